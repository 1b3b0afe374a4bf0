// Problem model (API: include/sovereign/problem.hpp; the reference's
// include/sovereign/problem.hpp:29-178 and src/io/problem.cpp).
//
// Behaviour the callers rely on, kept from the reference: the objective
// value starts from the constant and adds c_j x_j in column order; the row
// type comes from which bounds are finite; a new Builder column is
// nonnegative and unbounded above (the MPS default); a repeated row or
// column name is an error; validate() reports size mismatches and crossed
// bounds as InvalidArgument and non-finite data as NumericalFailure, then
// defers to CsrMatrix::validate.  Additive: Problem::from_csr adopts already
// compressed arrays, and such a Problem names its rows and columns only when
// asked.
#include "sovereign/problem.hpp"

#include <cmath>
#include <sstream>

namespace sov {
namespace {

inline std::size_t at(Idx i) { return static_cast<std::size_t>(i); }

/// Formats a bound for messages ("inf" / "-inf" / the number).
std::string bound_text(Real v) {
    std::ostringstream os;
    os << v;
    return os.str();
}

}  // namespace

// ------------------------------------------------------------------ Problem

Real Problem::evaluate_objective(std::span<const Real> x) const {
    if (x.size() != objective_.size())
        throw Error(Status::DimensionMismatch, "evaluate_objective: x has " + std::to_string(x.size()) +
                                                   " entries for " + std::to_string(objective_.size()) + " columns");
    Real value = objective_constant_;
    const Real* c = objective_.data();
    for (std::size_t j = 0, n = objective_.size(); j < n; ++j) value += c[j] * x[j];
    return value;
}

RowType Problem::row_type(Idx i) const noexcept {
    const bool has_lo = is_finite_bound(row_lower_[at(i)]);
    const bool has_hi = is_finite_bound(row_upper_[at(i)]);
    switch ((has_lo ? 2 : 0) | (has_hi ? 1 : 0)) {
        case 0: return RowType::Free;
        case 1: return RowType::LessEqual;
        case 2: return RowType::GreaterEqual;
        default: return row_lower_[at(i)] == row_upper_[at(i)] ? RowType::Equality : RowType::Range;
    }
}

Idx Problem::num_integer_columns() const noexcept {
    Idx count = 0;
    for (const VarKind k : col_kind_)
        if (k == VarKind::Integer) ++count;
    return count;
}

Idx Problem::num_free_rows() const noexcept {
    Idx count = 0;
    for (Idx i = 0, m = num_rows(); i < m; ++i)
        if (row_type(i) == RowType::Free) ++count;
    return count;
}

void Problem::ensure_names() const {
    if (names_ready_) return;
    row_names_.clear();
    col_names_.clear();
    row_names_.reserve(at(num_rows()));
    col_names_.reserve(at(num_cols()));
    for (Idx i = 0; i < num_rows(); ++i) row_names_.push_back("R" + std::to_string(i));
    for (Idx j = 0; j < num_cols(); ++j) col_names_.push_back("C" + std::to_string(j));
    names_ready_ = true;
}

std::span<const std::string> Problem::row_names() const {
    ensure_names();
    return row_names_;
}

std::span<const std::string> Problem::col_names() const {
    ensure_names();
    return col_names_;
}

bool Problem::has_crossed_bounds() const noexcept {
    auto crossed = [](const std::vector<Real>& lo, const std::vector<Real>& hi) {
        for (std::size_t k = 0; k < lo.size() && k < hi.size(); ++k)
            if (lo[k] > hi[k]) return true;
        return false;
    };
    return crossed(row_lower_, row_upper_) || crossed(col_lower_, col_upper_);
}

Status Problem::validate(std::string* why) const {
    Status status = Status::Ok;
    auto flag = [&](Status s, std::string msg) {
        status = s;
        if (why != nullptr) *why = std::move(msg);
        return false;
    };
    const std::size_t m = at(num_rows()), n = at(num_cols());
    // 1. every per-row and per-column array has the matrix's length
    struct Length {
        const char* what;
        std::size_t have, want;
    };
    const Length lengths[] = {
        {"row lower bounds", row_lower_.size(), m},
        {"row upper bounds", row_upper_.size(), m},
        {"row names", names_ready_ ? row_names_.size() : m, m},
        {"column lower bounds", col_lower_.size(), n},
        {"column upper bounds", col_upper_.size(), n},
        {"column kinds", col_kind_.size(), n},
        {"column names", names_ready_ ? col_names_.size() : n, n},
        {"objective coefficients", objective_.size(), n},
    };
    for (const Length& l : lengths)
        if (l.have != l.want) {
            flag(Status::InvalidArgument, std::string(l.what) + ": " + std::to_string(l.have) + " entries, the " +
                                              std::to_string(m) + " x " + std::to_string(n) + " matrix needs " +
                                              std::to_string(l.want));
            return status;
        }
    auto row_label = [&](std::size_t i) { return "row '" + (names_ready_ ? row_names_[i] : "R" + std::to_string(i)) + "'"; };
    auto col_label = [&](std::size_t j) { return "column '" + (names_ready_ ? col_names_[j] : "C" + std::to_string(j)) + "'"; };
    // 2. bounds: no NaN, lower <= upper
    auto bounds_ok = [&](Real lo, Real hi, const std::string& label) {
        if (std::isnan(lo) || std::isnan(hi)) return flag(Status::NumericalFailure, label + ": a bound is NaN");
        if (lo > hi)
            return flag(Status::InvalidArgument,
                        label + ": lower bound " + bound_text(lo) + " exceeds upper bound " + bound_text(hi));
        return true;
    };
    if (!std::isfinite(objective_constant_)) {
        flag(Status::NumericalFailure, "the objective constant is not finite");
        return status;
    }
    for (std::size_t j = 0; j < n; ++j) {
        if (!std::isfinite(objective_[j])) {
            flag(Status::NumericalFailure, col_label(j) + ": objective coefficient is not finite");
            return status;
        }
        if (!bounds_ok(col_lower_[j], col_upper_[j], col_label(j))) return status;
    }
    for (std::size_t i = 0; i < m; ++i)
        if (!bounds_ok(row_lower_[i], row_upper_[i], row_label(i))) return status;
    // 3. integer columns: finite bounds must be whole numbers (a fractional
    //    one is legal but nearly always a reader mistake)
    for (std::size_t j = 0; j < n; ++j) {
        if (col_kind_[j] != VarKind::Integer) continue;
        for (const Real b : {col_lower_[j], col_upper_[j]})
            if (is_finite_bound(b) && std::abs(b - std::round(b)) > 1e-9) {
                flag(Status::InvalidArgument, col_label(j) + " is integer but has the fractional bound " + bound_text(b));
                return status;
            }
    }
    // 4. the matrix itself
    return matrix_.validate(why);
}

std::string Problem::summary() const {
    std::ostringstream os;
    os << (name_.empty() ? "<unnamed>" : name_) << ": " << num_rows() << " rows, " << num_cols() << " cols, "
       << num_nonzeros() << " nonzeros, " << to_string(sense_);
    if (const Idx k = num_integer_columns(); k > 0) os << ", " << k << " integer";
    if (const Idx k = num_free_rows(); k > 0) os << ", " << k << " free row(s)";
    if (objective_constant_ != 0.0) os << ", objective constant " << objective_constant_;
    return os.str();
}

Problem Problem::from_csr(CsrMatrix matrix, std::vector<Real> objective, std::vector<Real> col_lower,
                          std::vector<Real> col_upper, std::vector<Real> row_lower, std::vector<Real> row_upper,
                          ObjSense sense, Real objective_constant, std::string name) {
    std::string why;
    if (const Status s = matrix.validate(&why); s != Status::Ok) throw Error(s, "Problem::from_csr: " + why);
    Problem p;
    p.name_ = std::move(name);
    p.sense_ = sense;
    p.objective_constant_ = objective_constant;
    p.col_kind_.assign(at(matrix.cols()), VarKind::Continuous);
    p.matrix_ = std::move(matrix);
    p.objective_ = std::move(objective);
    p.col_lower_ = std::move(col_lower);
    p.col_upper_ = std::move(col_upper);
    p.row_lower_ = std::move(row_lower);
    p.row_upper_ = std::move(row_upper);
    p.names_ready_ = false;
    return p;
}

// --------------------------------------------------------- Problem::Builder

Idx Problem::Builder::NameTable::insert(std::string name, const char* what) {
    const Idx next = size();
    if (!index.try_emplace(name, next).second)
        throw Error(std::string(what) + " name '" + name + "' is used twice");
    names.push_back(std::move(name));
    return next;
}

Idx Problem::Builder::NameTable::find(const std::string& name) const {
    const auto hit = index.find(name);
    return hit == index.end() ? Idx{-1} : hit->second;
}

Idx Problem::Builder::add_row(std::string name, Real lower, Real upper) {
    const Idx i = rows_.insert(std::move(name), "row");
    row_lower_.push_back(lower);
    row_upper_.push_back(upper);
    return i;
}

Idx Problem::Builder::add_column(std::string name) {
    const Idx j = cols_.insert(std::move(name), "column");
    // unless told otherwise a column is x >= 0 (the MPS default)
    col_lower_.push_back(0.0);
    col_upper_.push_back(kInfinity);
    col_kind_.push_back(VarKind::Continuous);
    objective_.push_back(0.0);
    return j;
}

Idx Problem::Builder::find_row(const std::string& name) const { return rows_.find(name); }

Idx Problem::Builder::find_column(const std::string& name) const { return cols_.find(name); }

void Problem::Builder::set_row_bounds(Idx row, Real lower, Real upper) {
    row_lower_.at(at(row)) = lower;
    row_upper_.at(at(row)) = upper;
}

void Problem::Builder::set_column_bounds(Idx col, Real lower, Real upper) {
    col_lower_.at(at(col)) = lower;
    col_upper_.at(at(col)) = upper;
}

void Problem::Builder::set_column_kind(Idx col, VarKind kind) { col_kind_.at(at(col)) = kind; }

void Problem::Builder::set_objective_coefficient(Idx col, Real value) { objective_.at(at(col)) = value; }

void Problem::Builder::add_coefficient(Idx row, Idx col, Real value) { entries_.push_back({row, col, value}); }

Problem Problem::Builder::finish() {
    Problem p;
    p.matrix_ = CsrMatrix::from_triplets(rows_.size(), cols_.size(), entries_);
    p.name_ = std::move(name_);
    p.objective_name_ = std::move(objective_name_);
    p.sense_ = sense_;
    p.objective_constant_ = objective_constant_;
    p.objective_ = std::move(objective_);
    p.row_lower_ = std::move(row_lower_);
    p.row_upper_ = std::move(row_upper_);
    p.col_lower_ = std::move(col_lower_);
    p.col_upper_ = std::move(col_upper_);
    p.col_kind_ = std::move(col_kind_);
    p.row_names_ = std::move(rows_.names);
    p.col_names_ = std::move(cols_.names);
    *this = Builder{};
    return p;
}

}  // namespace sov
