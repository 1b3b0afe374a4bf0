// Host CSR / CSC containers (API: include/sovereign/sparse.hpp, the
// reference's include/sovereign/sparse.hpp:28-181).
//
// Both orientations are the same object seen from two sides: a "compressed"
// index (a pointer array over the major dimension, minor indices and values).
// One checker, one transposer and one pair of products serve both classes:
//   * check_compressed   -- the structural and numerical checks behind
//                           CsrMatrix::validate / CscMatrix::validate;
//   * transpose          -- counting sort over the minor dimension (to_csc,
//                           to_csr), minor-major order kept stable;
//   * gather / scatter   -- y = M x as a per-major-line dot (CSR multiply),
//                           or as per-entry accumulation into y (CSR
//                           multiply_transpose, CSC multiply).
// Arithmetic order is part of the contract: the gather adds a row's entries
// in ascending position, the scatter visits entries in storage order into a
// zeroed y (reference src/l0/sparse.cpp:130-154), so host results equal the
// reference's bit for bit.  from_triplets sums duplicates in input order and
// drops |sum| <= tol::numeric_zero (reference :30-70) with a linear bucket
// pass instead of an ordered map.
#include "sovereign/sparse.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>

namespace sov {
namespace {

using std::size_t;

inline size_t at(Idx i) { return static_cast<size_t>(i); }

/// A read-only view of one compressed orientation.
struct Compressed {
    const char* kind;            // "CSR" or "CSC", for messages
    const char* major_name;      // "row" / "column"
    const char* minor_name;      // "column" / "row"
    Idx major = 0, minor = 0;
    std::span<const Idx> ptr;
    std::span<const Idx> idx;
    std::span<const Real> val;
};

/// The first defect of a compressed index, or Ok.  Structure defects are
/// InvalidArgument; a non-finite value is NumericalFailure.
Status check_compressed(const Compressed& c, std::string* why) {
    auto report = [&](Status s, std::string msg) {
        if (why != nullptr) *why = std::string(c.kind) + ": " + msg;
        return s;
    };
    if (c.ptr.size() != at(c.major) + 1)
        return report(Status::InvalidArgument, "pointer array holds " + std::to_string(c.ptr.size()) +
                                                   " entries for " + std::to_string(c.major) + " " +
                                                   c.major_name + "s (needs one more than the " + c.major_name +
                                                   " count)");
    if (c.idx.size() != c.val.size())
        return report(Status::InvalidArgument, std::to_string(c.idx.size()) + " indices but " +
                                                   std::to_string(c.val.size()) + " values");
    if (c.ptr.front() != 0)
        return report(Status::InvalidArgument, "pointer array starts at " + std::to_string(c.ptr.front()) +
                                                   ", not 0");
    if (at(c.ptr.back()) != c.val.size())
        return report(Status::InvalidArgument, "pointer array ends at " + std::to_string(c.ptr.back()) +
                                                   " but " + std::to_string(c.val.size()) + " entries are stored");
    for (Idx line = 0; line < c.major; ++line) {
        const Idx first = c.ptr[at(line)], last = c.ptr[at(line) + 1];
        if (last < first)
            return report(Status::InvalidArgument, std::string(c.major_name) + " " + std::to_string(line) +
                                                       " ends before it starts (pointer array decreases)");
        Idx prev = -1;
        for (Idx k = first; k < last; ++k) {
            const Idx j = c.idx[at(k)];
            if (j < 0 || j >= c.minor)
                return report(Status::InvalidArgument, std::string(c.minor_name) + " index " + std::to_string(j) +
                                                           " at entry " + std::to_string(k) + " is outside [0, " +
                                                           std::to_string(c.minor) + ")");
            // products and transposes rely on sorted, duplicate-free lines
            if (j <= prev)
                return report(Status::InvalidArgument, std::string(c.major_name) + " " + std::to_string(line) +
                                                           ": " + c.minor_name + " indices must increase strictly (" +
                                                           std::to_string(prev) + " then " + std::to_string(j) + ")");
            prev = j;
            if (!std::isfinite(c.val[at(k)]))
                return report(Status::NumericalFailure, "value at entry " + std::to_string(k) + " is not finite");
        }
    }
    if (why != nullptr) why->clear();
    return Status::Ok;
}

/// The same entries compressed over the other dimension: a counting sort on
/// the minor index, lines swept in order, so each new line lists its entries
/// by ascending old-major index.
struct Transposed {
    std::vector<Idx> ptr, idx;
    std::vector<Real> val;
};

Transposed transpose(const Compressed& c) {
    Transposed t;
    t.ptr.assign(at(c.minor) + 1, 0);
    for (const Idx j : c.idx) ++t.ptr[at(j) + 1];
    for (size_t j = 1; j < t.ptr.size(); ++j) t.ptr[j] += t.ptr[j - 1];
    t.idx.resize(c.idx.size());
    t.val.resize(c.val.size());
    std::vector<Idx> fill(t.ptr.begin(), t.ptr.end() - 1);
    for (Idx line = 0; line < c.major; ++line)
        for (Idx k = c.ptr[at(line)]; k < c.ptr[at(line) + 1]; ++k) {
            const size_t slot = at(fill[at(c.idx[at(k)])]++);
            t.idx[slot] = line;
            t.val[slot] = c.val[at(k)];
        }
    return t;
}

/// y[line] = sum over the line's entries, in storage order, of val * x[idx].
void gather(const Compressed& c, const Real* x, Real* y) {
    const Idx* ptr = c.ptr.data();
    const Idx* idx = c.idx.data();
    const Real* val = c.val.data();
    for (Idx line = 0; line < c.major; ++line) {
        Real dot = 0.0;
        for (Idx k = ptr[line]; k < ptr[line + 1]; ++k) dot += val[k] * x[idx[k]];
        y[line] = dot;
    }
}

/// y = 0, then y[idx] += val * x[line] over the entries in storage order.
void scatter(const Compressed& c, const Real* x, Real* y, size_t ylen) {
    std::fill(y, y + ylen, Real{0});
    const Idx* ptr = c.ptr.data();
    const Idx* idx = c.idx.data();
    const Real* val = c.val.data();
    for (Idx line = 0; line < c.major; ++line) {
        const Real xl = x[line];
        for (Idx k = ptr[line]; k < ptr[line + 1]; ++k) y[idx[k]] += val[k] * xl;
    }
}

void require_lengths(std::size_t got_x, Idx want_x, std::size_t got_y, Idx want_y, const char* what) {
    if (got_x != at(want_x) || got_y != at(want_y))
        throw Error(Status::DimensionMismatch, std::string(what) + ": x has " + std::to_string(got_x) +
                                                   " entries (needs " + std::to_string(want_x) + "), y has " +
                                                   std::to_string(got_y) + " (needs " + std::to_string(want_y) + ")");
}

}  // namespace

// ---------------------------------------------------------------- CsrMatrix

CsrMatrix::CsrMatrix(Idx rows, Idx cols, std::vector<Idx> row_ptr, std::vector<Idx> col_idx,
                     std::vector<Real> values)
    : rows_(rows), cols_(cols), row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)),
      values_(std::move(values)) {
    if (rows_ < 0 || cols_ < 0)
        throw Error("CsrMatrix(" + std::to_string(rows_) + " x " + std::to_string(cols_) + "): dimensions must be >= 0");
    if (row_ptr_.size() != at(rows_) + 1)
        throw Error("CsrMatrix: " + std::to_string(row_ptr_.size()) + " row pointers for " + std::to_string(rows_) +
                    " rows");
    if (col_idx_.size() != values_.size())
        throw Error("CsrMatrix: " + std::to_string(col_idx_.size()) + " column indices but " +
                    std::to_string(values_.size()) + " values");
}

CsrMatrix CsrMatrix::from_triplets(Idx rows, Idx cols, std::span<const Triplet> entries) {
    if (rows < 0 || cols < 0) throw Error("from_triplets: dimensions must be >= 0");
    // pass 1: validate and count per row
    std::vector<Idx> start(at(rows) + 1, 0);
    for (const Triplet& t : entries) {
        if (t.row < 0 || t.row >= rows || t.col < 0 || t.col >= cols)
            throw Error("from_triplets: (" + std::to_string(t.row) + ", " + std::to_string(t.col) +
                        ") lies outside the " + std::to_string(rows) + " x " + std::to_string(cols) + " matrix");
        if (!std::isfinite(t.value))
            throw Error("from_triplets: entry (" + std::to_string(t.row) + ", " + std::to_string(t.col) +
                        ") is not finite");
        ++start[at(t.row) + 1];
    }
    for (size_t r = 1; r < start.size(); ++r) start[r] += start[r - 1];
    // pass 2: bucket entry numbers by row (input order kept inside a row)
    std::vector<size_t> by_row(entries.size());
    {
        std::vector<Idx> next(start.begin(), start.end() - 1);
        for (size_t e = 0; e < entries.size(); ++e) by_row[at(next[at(entries[e].row)]++)] = e;
    }
    // pass 3: per row, a stable sort by column, then duplicates summed in
    // input order and numerical zeros dropped
    CsrMatrix out;
    out.rows_ = rows;
    out.cols_ = cols;
    out.row_ptr_.assign(at(rows) + 1, 0);
    out.col_idx_.reserve(entries.size());
    out.values_.reserve(entries.size());
    auto col_of = [&](size_t e) { return entries[e].col; };
    for (Idx r = 0; r < rows; ++r) {
        const auto first = by_row.begin() + start[at(r)], last = by_row.begin() + start[at(r) + 1];
        std::stable_sort(first, last, [&](size_t a, size_t b) { return col_of(a) < col_of(b); });
        for (auto run = first; run != last;) {
            const Idx col = col_of(*run);
            Real sum = 0.0;
            for (; run != last && col_of(*run) == col; ++run) sum += entries[*run].value;
            if (std::abs(sum) > tol::numeric_zero) {
                out.col_idx_.push_back(col);
                out.values_.push_back(sum);
            }
        }
        out.row_ptr_[at(r) + 1] = static_cast<Idx>(out.values_.size());
    }
    return out;
}

CsrMatrix CsrMatrix::from_dense(Idx rows, Idx cols, std::span<const Real> row_major, Real drop_tol) {
    if (rows < 0 || cols < 0 || at(rows) * at(cols) != row_major.size())
        throw Error("from_dense: " + std::to_string(row_major.size()) + " values for a " + std::to_string(rows) +
                    " x " + std::to_string(cols) + " matrix");
    CsrMatrix out;
    out.rows_ = rows;
    out.cols_ = cols;
    out.row_ptr_.assign(at(rows) + 1, 0);
    const Real* v = row_major.data();
    for (Idx r = 0; r < rows; ++r, v += cols) {
        for (Idx c = 0; c < cols; ++c)
            if (std::abs(v[c]) > drop_tol) {
                out.col_idx_.push_back(c);
                out.values_.push_back(v[c]);
            }
        out.row_ptr_[at(r) + 1] = static_cast<Idx>(out.values_.size());
    }
    return out;
}

Status CsrMatrix::validate(std::string* why) const {
    return check_compressed({"CSR", "row", "column", rows_, cols_, row_ptr_, col_idx_, values_}, why);
}

void CsrMatrix::multiply(std::span<const Real> x, std::span<Real> y) const {
    require_lengths(x.size(), cols_, y.size(), rows_, "CsrMatrix::multiply");
    gather({"CSR", "row", "column", rows_, cols_, row_ptr_, col_idx_, values_}, x.data(), y.data());
}

void CsrMatrix::multiply_transpose(std::span<const Real> x, std::span<Real> y) const {
    require_lengths(x.size(), rows_, y.size(), cols_, "CsrMatrix::multiply_transpose");
    scatter({"CSR", "row", "column", rows_, cols_, row_ptr_, col_idx_, values_}, x.data(), y.data(), y.size());
}

CscMatrix CsrMatrix::to_csc() const {
    Transposed t = transpose({"CSR", "row", "column", rows_, cols_, row_ptr_, col_idx_, values_});
    return CscMatrix(rows_, cols_, std::move(t.ptr), std::move(t.idx), std::move(t.val));
}

CsrMatrix::Extremes CsrMatrix::coefficient_extremes() const noexcept {
    Extremes e;
    Real smallest = std::numeric_limits<Real>::infinity();
    for (const Real v : values_) {
        const Real a = std::abs(v);
        e.max_abs = std::max(e.max_abs, a);
        if (a > tol::numeric_zero) smallest = std::min(smallest, a);
    }
    e.min_abs_nonzero = std::isinf(smallest) ? Real{0} : smallest;
    return e;
}

// ---------------------------------------------------------------- CscMatrix

CscMatrix::CscMatrix(Idx rows, Idx cols, std::vector<Idx> col_ptr, std::vector<Idx> row_idx,
                     std::vector<Real> values)
    : rows_(rows), cols_(cols), col_ptr_(std::move(col_ptr)), row_idx_(std::move(row_idx)),
      values_(std::move(values)) {
    if (rows_ < 0 || cols_ < 0)
        throw Error("CscMatrix(" + std::to_string(rows_) + " x " + std::to_string(cols_) + "): dimensions must be >= 0");
    if (col_ptr_.size() != at(cols_) + 1)
        throw Error("CscMatrix: " + std::to_string(col_ptr_.size()) + " column pointers for " +
                    std::to_string(cols_) + " columns");
    if (row_idx_.size() != values_.size())
        throw Error("CscMatrix: " + std::to_string(row_idx_.size()) + " row indices but " +
                    std::to_string(values_.size()) + " values");
}

Status CscMatrix::validate(std::string* why) const {
    return check_compressed({"CSC", "column", "row", cols_, rows_, col_ptr_, row_idx_, values_}, why);
}

void CscMatrix::multiply(std::span<const Real> x, std::span<Real> y) const {
    require_lengths(x.size(), cols_, y.size(), rows_, "CscMatrix::multiply");
    scatter({"CSC", "column", "row", cols_, rows_, col_ptr_, row_idx_, values_}, x.data(), y.data(), y.size());
}

CsrMatrix CscMatrix::to_csr() const {
    Transposed t = transpose({"CSC", "column", "row", cols_, rows_, col_ptr_, row_idx_, values_});
    return CsrMatrix(rows_, cols_, std::move(t.ptr), std::move(t.idx), std::move(t.val));
}

}  // namespace sov
