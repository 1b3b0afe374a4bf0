// Host CSR/CSC containers.  API mirrors include/sovereign/sparse.hpp:28-181 of
// the reference (the names callers use); the implementation is this
// project's own (sovereign-solver_amd/src/l0/sparse.cpp).  The host copy is
// the FP64 truth; FP32 images live only on the device.
#pragma once

#include <cstddef>
#include <span>
#include <string>
#include <vector>

#include "sovereign/numeric.hpp"
#include "sovereign/status.hpp"

namespace sov {

struct Triplet {
    Idx row = 0;
    Idx col = 0;
    Real value = 0.0;
};

class CscMatrix;

class CsrMatrix {
public:
    CsrMatrix() = default;
    CsrMatrix(Idx rows, Idx cols, std::vector<Idx> row_ptr, std::vector<Idx> col_idx,
              std::vector<Real> values);

    /// Sums duplicates, drops |v| <= tol::numeric_zero, sorts columns
    /// (reference sparse.cpp:30-70).  Linear-time bucket assembly here rather
    /// than an ordered map.
    static CsrMatrix from_triplets(Idx rows, Idx cols, std::span<const Triplet> entries);
    static CsrMatrix from_dense(Idx rows, Idx cols, std::span<const Real> row_major,
                                Real drop_tol = tol::numeric_zero);

    Idx rows() const noexcept { return rows_; }
    Idx cols() const noexcept { return cols_; }
    Idx nnz() const noexcept { return static_cast<Idx>(values_.size()); }
    std::span<const Idx> row_ptr() const noexcept { return row_ptr_; }
    std::span<const Idx> col_idx() const noexcept { return col_idx_; }
    std::span<const Real> values() const noexcept { return values_; }
    std::size_t byte_size() const noexcept {
        return row_ptr_.size() * sizeof(Idx) + col_idx_.size() * sizeof(Idx) +
               values_.size() * sizeof(Real);
    }

    Status validate(std::string* why = nullptr) const;
    void multiply(std::span<const Real> x, std::span<Real> y) const;
    void multiply_transpose(std::span<const Real> x, std::span<Real> y) const;
    CscMatrix to_csc() const;

    struct Extremes {
        Real max_abs = 0.0;
        Real min_abs_nonzero = 0.0;
    };
    Extremes coefficient_extremes() const noexcept;

private:
    Idx rows_ = 0;
    Idx cols_ = 0;
    std::vector<Idx> row_ptr_{0};
    std::vector<Idx> col_idx_;
    std::vector<Real> values_;
};

class CscMatrix {
public:
    CscMatrix() = default;
    CscMatrix(Idx rows, Idx cols, std::vector<Idx> col_ptr, std::vector<Idx> row_idx,
              std::vector<Real> values);
    Idx rows() const noexcept { return rows_; }
    Idx cols() const noexcept { return cols_; }
    Idx nnz() const noexcept { return static_cast<Idx>(values_.size()); }
    std::span<const Idx> col_ptr() const noexcept { return col_ptr_; }
    std::span<const Idx> row_idx() const noexcept { return row_idx_; }
    std::span<const Real> values() const noexcept { return values_; }
    Status validate(std::string* why = nullptr) const;
    void multiply(std::span<const Real> x, std::span<Real> y) const;
    CsrMatrix to_csr() const;

private:
    Idx rows_ = 0;
    Idx cols_ = 0;
    std::vector<Idx> col_ptr_{0};
    std::vector<Idx> row_idx_;
    std::vector<Real> values_;
};

/// Declared-but-empty blocked layouts (reference sparse.hpp:140-181); kept so
/// code naming them compiles.
struct VbcsrMatrix {
    Idx rows = 0, cols = 0;
    std::vector<Idx> block_row_ptr, block_col_idx, row_block_offset, col_block_offset;
    std::vector<Real> values;
    bool empty() const noexcept { return values.empty(); }
};
struct BsrMatrix {
    Idx block_rows = 0, block_cols = 0, block_dim = 0;
    std::vector<Idx> block_row_ptr, block_col_idx;
    std::vector<Real> values;
    bool empty() const noexcept { return values.empty(); }
};

}  // namespace sov
