// The LP as callers hand it to the engine.  API mirrors
// include/sovereign/problem.hpp:29-178 of the reference:
//     optimize c'x + d  s.t.  row_lo <= A x <= row_hi,  col_lo <= x <= col_hi
// Additive: Problem::from_csr, a construction path that takes already
// compressed arrays (no names, no triplet assembly) for million-column
// instances.
#pragma once

#include <cstdint>
#include <limits>
#include <span>
#include <string>
#include <unordered_map>
#include <vector>

#include "sovereign/numeric.hpp"
#include "sovereign/sparse.hpp"
#include "sovereign/status.hpp"

namespace sov {

inline constexpr Real kInfinity = std::numeric_limits<Real>::infinity();

inline bool is_finite_bound(Real v) noexcept { return v > -kInfinity && v < kInfinity; }

enum class ObjSense { Minimize, Maximize };

constexpr const char* to_string(ObjSense s) noexcept {
    return s == ObjSense::Minimize ? "minimize" : "maximize";
}

enum class VarKind : std::uint8_t { Continuous, Integer };
enum class RowType : std::uint8_t { Free, LessEqual, GreaterEqual, Equality, Range };

class Problem {
public:
    Problem() = default;

    Idx num_rows() const noexcept { return matrix_.rows(); }
    Idx num_cols() const noexcept { return matrix_.cols(); }
    Idx num_nonzeros() const noexcept { return matrix_.nnz(); }
    const CsrMatrix& matrix() const noexcept { return matrix_; }

    ObjSense sense() const noexcept { return sense_; }
    std::span<const Real> objective() const noexcept { return objective_; }
    Real objective_constant() const noexcept { return objective_constant_; }
    Real evaluate_objective(std::span<const Real> x) const;

    std::span<const Real> row_lower() const noexcept { return row_lower_; }
    std::span<const Real> row_upper() const noexcept { return row_upper_; }
    std::span<const Real> col_lower() const noexcept { return col_lower_; }
    std::span<const Real> col_upper() const noexcept { return col_upper_; }
    RowType row_type(Idx i) const noexcept;

    std::span<const VarKind> col_kind() const noexcept { return col_kind_; }
    Idx num_integer_columns() const noexcept;
    bool is_mip() const noexcept { return num_integer_columns() > 0; }

    const std::string& name() const noexcept { return name_; }
    const std::string& objective_name() const noexcept { return objective_name_; }
    /// Names.  A Problem built by from_csr has none until asked: the first
    /// call materialises "R<i>" / "C<j>" (thread-unsafe on that first call).
    std::span<const std::string> row_names() const;
    std::span<const std::string> col_names() const;
    Idx num_free_rows() const noexcept;

    Status validate(std::string* why = nullptr) const;
    bool has_crossed_bounds() const noexcept;
    std::string summary() const;

    /// Additive fast path: adopt compressed arrays directly.  The matrix must
    /// already satisfy CsrMatrix::validate (checked; throws Error if not).
    static Problem from_csr(CsrMatrix matrix, std::vector<Real> objective,
                            std::vector<Real> col_lower, std::vector<Real> col_upper,
                            std::vector<Real> row_lower, std::vector<Real> row_upper,
                            ObjSense sense = ObjSense::Minimize, Real objective_constant = 0.0,
                            std::string name = {});

    class Builder;

private:
    friend class Builder;
    void ensure_names() const;

    std::string name_;
    std::string objective_name_;
    ObjSense sense_ = ObjSense::Minimize;
    CsrMatrix matrix_;
    std::vector<Real> objective_;
    Real objective_constant_ = 0.0;
    std::vector<Real> row_lower_, row_upper_;
    std::vector<Real> col_lower_, col_upper_;
    std::vector<VarKind> col_kind_;
    mutable std::vector<std::string> row_names_, col_names_;
    mutable bool names_ready_ = true;
};

class Problem::Builder {
public:
    void set_name(std::string n) { name_ = std::move(n); }
    void set_sense(ObjSense s) { sense_ = s; }
    void set_objective_name(std::string n) { objective_name_ = std::move(n); }
    void set_objective_constant(Real d) { objective_constant_ = d; }
    Idx add_row(std::string name, Real lower, Real upper);
    Idx add_column(std::string name);
    Idx find_row(const std::string& name) const;
    Idx find_column(const std::string& name) const;
    void set_row_bounds(Idx row, Real lower, Real upper);
    void set_column_bounds(Idx col, Real lower, Real upper);
    void set_column_kind(Idx col, VarKind kind);
    void set_objective_coefficient(Idx col, Real value);
    void add_coefficient(Idx row, Idx col, Real value);
    Real column_lower(Idx col) const { return col_lower_.at(static_cast<std::size_t>(col)); }
    Real column_upper(Idx col) const { return col_upper_.at(static_cast<std::size_t>(col)); }
    VarKind column_kind(Idx col) const { return col_kind_.at(static_cast<std::size_t>(col)); }
    Idx num_rows() const noexcept { return rows_.size(); }
    Idx num_columns() const noexcept { return cols_.size(); }
    Problem finish();

private:
    /// Names in insertion order plus their lookup; a repeated name throws.
    struct NameTable {
        std::vector<std::string> names;
        std::unordered_map<std::string, Idx> index;
        Idx size() const noexcept { return static_cast<Idx>(names.size()); }
        Idx insert(std::string name, const char* what);
        Idx find(const std::string& name) const;
    };
    std::string name_, objective_name_;
    ObjSense sense_ = ObjSense::Minimize;
    Real objective_constant_ = 0.0;
    NameTable rows_, cols_;
    std::vector<Real> row_lower_, row_upper_, col_lower_, col_upper_, objective_;
    std::vector<VarKind> col_kind_;
    std::vector<Triplet> entries_;
};

}  // namespace sov
