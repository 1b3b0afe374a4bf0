// Problem model (API: include/sovereign/problem.hpp), following reference
// src/io/problem.cpp: evaluate_objective with the constant (:193-199),
// row_type (:201-208), crossed bounds (:222-228), validate (:230-280),
// Builder (:299-377).  Additive: from_csr, and names materialised lazily for
// problems built without them.
#include "sovereign/problem.hpp"

#include <cmath>
#include <sstream>

namespace sov {

Real Problem::evaluate_objective(std::span<const Real> x) const {
    if (static_cast<Idx>(x.size()) != num_cols()) throw Error(Status::DimensionMismatch, "evaluate_objective");
    Real acc = objective_constant_;
    for (std::size_t j = 0; j < objective_.size(); ++j) acc += objective_[j] * x[j];
    return acc;
}

RowType Problem::row_type(Idx i) const noexcept {
    const Real lo = row_lower_[static_cast<std::size_t>(i)], hi = row_upper_[static_cast<std::size_t>(i)];
    const bool lf = is_finite_bound(lo), hf = is_finite_bound(hi);
    if (!lf && !hf) return RowType::Free;
    if (lf && hf) return lo == hi ? RowType::Equality : RowType::Range;
    return hf ? RowType::LessEqual : RowType::GreaterEqual;
}

Idx Problem::num_integer_columns() const noexcept {
    Idx k = 0;
    for (VarKind v : col_kind_) k += v == VarKind::Integer ? 1 : 0;
    return k;
}

Idx Problem::num_free_rows() const noexcept {
    Idx k = 0;
    for (Idx i = 0; i < num_rows(); ++i) k += row_type(i) == RowType::Free ? 1 : 0;
    return k;
}

void Problem::ensure_names() const {
    if (names_ready_) return;
    row_names_.resize(static_cast<std::size_t>(num_rows()));
    col_names_.resize(static_cast<std::size_t>(num_cols()));
    for (std::size_t i = 0; i < row_names_.size(); ++i) row_names_[i] = "R" + std::to_string(i);
    for (std::size_t j = 0; j < col_names_.size(); ++j) col_names_[j] = "C" + std::to_string(j);
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
    for (std::size_t i = 0; i < row_lower_.size(); ++i)
        if (row_lower_[i] > row_upper_[i]) return true;
    for (std::size_t j = 0; j < col_lower_.size(); ++j)
        if (col_lower_[j] > col_upper_[j]) return true;
    return false;
}

Status Problem::validate(std::string* why) const {
    auto bad = [&](Status s, const std::string& msg) {
        if (why) *why = msg;
        return s;
    };
    const auto rows = static_cast<std::size_t>(num_rows());
    const auto cols = static_cast<std::size_t>(num_cols());
    if (row_lower_.size() != rows || row_upper_.size() != rows)
        return bad(Status::InvalidArgument, "row bound arrays do not match row count");
    if (names_ready_ && row_names_.size() != rows)
        return bad(Status::InvalidArgument, "row name array does not match row count");
    if (col_lower_.size() != cols || col_upper_.size() != cols)
        return bad(Status::InvalidArgument, "column bound arrays do not match column count");
    if (col_kind_.size() != cols || (names_ready_ && col_names_.size() != cols))
        return bad(Status::InvalidArgument, "column kind/name array does not match column count");
    if (objective_.size() != cols) return bad(Status::InvalidArgument, "objective length does not match column count");
    if (!std::isfinite(objective_constant_)) return bad(Status::NumericalFailure, "objective constant is not finite");
    auto cname = [&](std::size_t j) { return names_ready_ ? col_names_[j] : "C" + std::to_string(j); };
    auto rname = [&](std::size_t i) { return names_ready_ ? row_names_[i] : "R" + std::to_string(i); };
    for (std::size_t j = 0; j < cols; ++j) {
        if (!std::isfinite(objective_[j]))
            return bad(Status::NumericalFailure, "non-finite objective coefficient in column " + cname(j));
        if (std::isnan(col_lower_[j]) || std::isnan(col_upper_[j]))
            return bad(Status::NumericalFailure, "NaN bound on column " + cname(j));
        if (col_lower_[j] > col_upper_[j]) return bad(Status::InvalidArgument, "crossed bounds on column " + cname(j));
    }
    for (std::size_t i = 0; i < rows; ++i) {
        if (std::isnan(row_lower_[i]) || std::isnan(row_upper_[i]))
            return bad(Status::NumericalFailure, "NaN bound on row " + rname(i));
        if (row_lower_[i] > row_upper_[i]) return bad(Status::InvalidArgument, "crossed bounds on row " + rname(i));
    }
    for (std::size_t j = 0; j < cols; ++j) {
        if (col_kind_[j] != VarKind::Integer) continue;
        for (Real b : {col_lower_[j], col_upper_[j]})
            if (is_finite_bound(b) && std::abs(b - std::round(b)) > 1e-9)
                return bad(Status::InvalidArgument, "integer column " + cname(j) + " has a fractional bound");
    }
    return matrix_.validate(why);
}

std::string Problem::summary() const {
    std::ostringstream os;
    os << (name_.empty() ? "<unnamed>" : name_) << ": " << num_rows() << " rows, " << num_cols() << " cols, "
       << num_nonzeros() << " nonzeros, " << to_string(sense_);
    if (num_integer_columns() > 0) os << ", " << num_integer_columns() << " integer";
    if (num_free_rows() > 0) os << ", " << num_free_rows() << " free row(s)";
    if (objective_constant_ != 0.0) os << ", objective constant " << objective_constant_;
    return os.str();
}

Problem Problem::from_csr(CsrMatrix matrix, std::vector<Real> objective, std::vector<Real> col_lower,
                          std::vector<Real> col_upper, std::vector<Real> row_lower, std::vector<Real> row_upper,
                          ObjSense sense, Real objective_constant, std::string name) {
    std::string why;
    const Status s = matrix.validate(&why);
    if (s != Status::Ok) throw Error(s, "Problem::from_csr: " + why);
    Problem p;
    p.name_ = std::move(name);
    p.sense_ = sense;
    p.objective_constant_ = objective_constant;
    p.col_kind_.assign(static_cast<std::size_t>(matrix.cols()), VarKind::Continuous);
    p.matrix_ = std::move(matrix);
    p.objective_ = std::move(objective);
    p.col_lower_ = std::move(col_lower);
    p.col_upper_ = std::move(col_upper);
    p.row_lower_ = std::move(row_lower);
    p.row_upper_ = std::move(row_upper);
    p.names_ready_ = false;
    return p;
}

Idx Problem::Builder::add_row(std::string name, Real lower, Real upper) {
    const Idx index = static_cast<Idx>(row_names_.size());
    if (!row_index_.emplace(name, index).second) throw Error("duplicate row name '" + name + "'");
    row_names_.push_back(std::move(name));
    row_lower_.push_back(lower);
    row_upper_.push_back(upper);
    return index;
}

Idx Problem::Builder::add_column(std::string name) {
    const Idx index = static_cast<Idx>(col_names_.size());
    if (!col_index_.emplace(name, index).second) throw Error("duplicate column name '" + name + "'");
    col_names_.push_back(std::move(name));
    col_lower_.push_back(0.0);   // MPS default: nonnegative, unbounded above
    col_upper_.push_back(kInfinity);
    col_kind_.push_back(VarKind::Continuous);
    objective_.push_back(0.0);
    return index;
}

Idx Problem::Builder::find_row(const std::string& name) const {
    const auto it = row_index_.find(name);
    return it == row_index_.end() ? Idx{-1} : it->second;
}

Idx Problem::Builder::find_column(const std::string& name) const {
    const auto it = col_index_.find(name);
    return it == col_index_.end() ? Idx{-1} : it->second;
}

void Problem::Builder::set_row_bounds(Idx row, Real lower, Real upper) {
    row_lower_.at(static_cast<std::size_t>(row)) = lower;
    row_upper_.at(static_cast<std::size_t>(row)) = upper;
}

void Problem::Builder::set_column_bounds(Idx col, Real lower, Real upper) {
    col_lower_.at(static_cast<std::size_t>(col)) = lower;
    col_upper_.at(static_cast<std::size_t>(col)) = upper;
}

void Problem::Builder::set_column_kind(Idx col, VarKind kind) { col_kind_.at(static_cast<std::size_t>(col)) = kind; }

void Problem::Builder::set_objective_coefficient(Idx col, Real value) {
    objective_.at(static_cast<std::size_t>(col)) = value;
}

void Problem::Builder::add_coefficient(Idx row, Idx col, Real value) { entries_.push_back(Triplet{row, col, value}); }

Problem Problem::Builder::finish() {
    Problem p;
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
    p.row_names_ = std::move(row_names_);
    p.col_names_ = std::move(col_names_);
    p.matrix_ = CsrMatrix::from_triplets(static_cast<Idx>(p.row_names_.size()), static_cast<Idx>(p.col_names_.size()),
                                         entries_);
    entries_.clear();
    row_index_.clear();
    col_index_.clear();
    return p;
}

}  // namespace sov
