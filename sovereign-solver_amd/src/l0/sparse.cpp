// Host CSR / CSC containers (API: include/sovereign/sparse.hpp).
// Semantics follow reference src/l0/sparse.cpp: from_triplets sums duplicates
// in input order, drops |v| <= 1e-14 and sorts columns (:30-70); validate
// (:94-128); gather multiply (:130-141); scatter multiply_transpose
// (:143-154); counting-sort to_csc (:156-180).  Assembly here is a linear
// bucket pass plus a per-row stable sort instead of an ordered map.
#include "sovereign/sparse.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace sov {

CsrMatrix::CsrMatrix(Idx rows, Idx cols, std::vector<Idx> row_ptr, std::vector<Idx> col_idx,
                     std::vector<Real> values)
    : rows_(rows), cols_(cols), row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)),
      values_(std::move(values)) {
    if (rows_ < 0 || cols_ < 0) throw Error("CsrMatrix: negative dimensions");
    if (static_cast<Idx>(row_ptr_.size()) != rows_ + 1)
        throw Error("CsrMatrix: row_ptr must have rows+1 entries");
    if (col_idx_.size() != values_.size())
        throw Error("CsrMatrix: col_idx and values length mismatch");
}

CsrMatrix CsrMatrix::from_triplets(Idx rows, Idx cols, std::span<const Triplet> entries) {
    if (rows < 0 || cols < 0) throw Error("from_triplets: negative dimensions");
    const auto R = static_cast<std::size_t>(rows);
    std::vector<Idx> count(R + 1, 0);
    for (const Triplet& t : entries) {
        if (t.row < 0 || t.row >= rows || t.col < 0 || t.col >= cols)
            throw Error("from_triplets: entry out of range");
        if (!std::isfinite(t.value)) throw Error("from_triplets: non-finite coefficient");
        ++count[static_cast<std::size_t>(t.row) + 1];
    }
    for (std::size_t r = 0; r < R; ++r) count[r + 1] += count[r];
    // bucket by row, keeping input order inside each row
    std::vector<std::size_t> order(entries.size());
    {
        std::vector<Idx> cur(count.begin(), count.end() - 1);
        for (std::size_t e = 0; e < entries.size(); ++e)
            order[static_cast<std::size_t>(cur[static_cast<std::size_t>(entries[e].row)]++)] = e;
    }
    std::vector<Idx> row_ptr(R + 1, 0), col_idx;
    std::vector<Real> values;
    col_idx.reserve(entries.size());
    values.reserve(entries.size());
    for (std::size_t r = 0; r < R; ++r) {
        auto b = order.begin() + count[r], e = order.begin() + count[r + 1];
        std::stable_sort(b, e, [&](std::size_t p, std::size_t q) { return entries[p].col < entries[q].col; });
        for (auto it = b; it != e;) {
            const Idx c = entries[*it].col;
            Real acc = 0.0;
            while (it != e && entries[*it].col == c) acc += entries[*it++].value;
            if (std::abs(acc) <= tol::numeric_zero) continue;
            col_idx.push_back(c);
            values.push_back(acc);
        }
        row_ptr[r + 1] = static_cast<Idx>(values.size());
    }
    return CsrMatrix(rows, cols, std::move(row_ptr), std::move(col_idx), std::move(values));
}

CsrMatrix CsrMatrix::from_dense(Idx rows, Idx cols, std::span<const Real> a, Real drop_tol) {
    if (static_cast<std::size_t>(rows) * static_cast<std::size_t>(cols) != a.size())
        throw Error("from_dense: data size does not match dimensions");
    std::vector<Idx> rp(static_cast<std::size_t>(rows) + 1, 0), ci;
    std::vector<Real> v;
    for (Idx r = 0; r < rows; ++r) {
        for (Idx c = 0; c < cols; ++c) {
            const Real x = a[static_cast<std::size_t>(r) * static_cast<std::size_t>(cols) + static_cast<std::size_t>(c)];
            if (std::abs(x) > drop_tol) {
                ci.push_back(c);
                v.push_back(x);
            }
        }
        rp[static_cast<std::size_t>(r) + 1] = static_cast<Idx>(v.size());
    }
    return CsrMatrix(rows, cols, std::move(rp), std::move(ci), std::move(v));
}

Status CsrMatrix::validate(std::string* why) const {
    auto bad = [&](Status s, const char* msg) {
        if (why) *why = msg;
        return s;
    };
    if (static_cast<Idx>(row_ptr_.size()) != rows_ + 1) return bad(Status::InvalidArgument, "row_ptr size != rows+1");
    if (col_idx_.size() != values_.size()) return bad(Status::InvalidArgument, "col_idx/values length mismatch");
    if (row_ptr_.front() != 0) return bad(Status::InvalidArgument, "row_ptr does not start at 0");
    if (row_ptr_.back() != static_cast<Idx>(values_.size()))
        return bad(Status::InvalidArgument, "row_ptr does not end at nnz");
    for (Idx r = 0; r < rows_; ++r) {
        const Idx b = row_ptr_[static_cast<std::size_t>(r)], e = row_ptr_[static_cast<std::size_t>(r) + 1];
        if (b > e) return bad(Status::InvalidArgument, "row_ptr is not monotone");
        for (Idx k = b; k < e; ++k) {
            const Idx c = col_idx_[static_cast<std::size_t>(k)];
            if (c < 0 || c >= cols_) return bad(Status::InvalidArgument, "column index out of range");
            if (k > b && col_idx_[static_cast<std::size_t>(k) - 1] >= c)
                return bad(Status::InvalidArgument, "column indices not strictly ascending");
            if (!std::isfinite(values_[static_cast<std::size_t>(k)]))
                return bad(Status::NumericalFailure, "non-finite coefficient");
        }
    }
    if (why) why->clear();
    return Status::Ok;
}

void CsrMatrix::multiply(std::span<const Real> x, std::span<Real> y) const {
    if (static_cast<Idx>(x.size()) != cols_ || static_cast<Idx>(y.size()) != rows_)
        throw Error(Status::DimensionMismatch, "CsrMatrix::multiply");
    for (Idx r = 0; r < rows_; ++r) {
        Real s = 0.0;
        for (Idx k = row_ptr_[static_cast<std::size_t>(r)]; k < row_ptr_[static_cast<std::size_t>(r) + 1]; ++k)
            s += values_[static_cast<std::size_t>(k)] * x[static_cast<std::size_t>(col_idx_[static_cast<std::size_t>(k)])];
        y[static_cast<std::size_t>(r)] = s;
    }
}

void CsrMatrix::multiply_transpose(std::span<const Real> x, std::span<Real> y) const {
    if (static_cast<Idx>(x.size()) != rows_ || static_cast<Idx>(y.size()) != cols_)
        throw Error(Status::DimensionMismatch, "CsrMatrix::multiply_transpose");
    std::fill(y.begin(), y.end(), Real{0});
    for (Idx r = 0; r < rows_; ++r) {
        const Real xr = x[static_cast<std::size_t>(r)];
        for (Idx k = row_ptr_[static_cast<std::size_t>(r)]; k < row_ptr_[static_cast<std::size_t>(r) + 1]; ++k)
            y[static_cast<std::size_t>(col_idx_[static_cast<std::size_t>(k)])] += values_[static_cast<std::size_t>(k)] * xr;
    }
}

CscMatrix CsrMatrix::to_csc() const {
    const auto C = static_cast<std::size_t>(cols_);
    std::vector<Idx> cp(C + 1, 0);
    for (Idx c : col_idx_) ++cp[static_cast<std::size_t>(c) + 1];
    std::partial_sum(cp.begin(), cp.end(), cp.begin());
    std::vector<Idx> ri(col_idx_.size()), cur(cp.begin(), cp.end() - 1);
    std::vector<Real> v(values_.size());
    for (Idx r = 0; r < rows_; ++r)
        for (Idx k = row_ptr_[static_cast<std::size_t>(r)]; k < row_ptr_[static_cast<std::size_t>(r) + 1]; ++k) {
            const auto d = static_cast<std::size_t>(cur[static_cast<std::size_t>(col_idx_[static_cast<std::size_t>(k)])]++);
            ri[d] = r;
            v[d] = values_[static_cast<std::size_t>(k)];
        }
    return CscMatrix(rows_, cols_, std::move(cp), std::move(ri), std::move(v));
}

CsrMatrix::Extremes CsrMatrix::coefficient_extremes() const noexcept {
    Extremes e;
    bool any = false;
    for (Real v : values_) {
        const Real a = std::abs(v);
        e.max_abs = std::max(e.max_abs, a);
        if (a > tol::numeric_zero && (!any || a < e.min_abs_nonzero)) {
            e.min_abs_nonzero = a;
            any = true;
        }
    }
    return e;
}

CscMatrix::CscMatrix(Idx rows, Idx cols, std::vector<Idx> col_ptr, std::vector<Idx> row_idx,
                     std::vector<Real> values)
    : rows_(rows), cols_(cols), col_ptr_(std::move(col_ptr)), row_idx_(std::move(row_idx)),
      values_(std::move(values)) {
    if (static_cast<Idx>(col_ptr_.size()) != cols_ + 1) throw Error("CscMatrix: col_ptr must have cols+1 entries");
    if (row_idx_.size() != values_.size()) throw Error("CscMatrix: row_idx and values length mismatch");
}

Status CscMatrix::validate(std::string* why) const {
    auto bad = [&](const char* msg) {
        if (why) *why = msg;
        return Status::InvalidArgument;
    };
    if (static_cast<Idx>(col_ptr_.size()) != cols_ + 1) return bad("col_ptr size != cols+1");
    if (col_ptr_.back() != static_cast<Idx>(values_.size())) return bad("col_ptr does not end at nnz");
    for (Idx c = 0; c < cols_; ++c) {
        const Idx b = col_ptr_[static_cast<std::size_t>(c)], e = col_ptr_[static_cast<std::size_t>(c) + 1];
        if (b > e) return bad("col_ptr is not monotone");
        for (Idx k = b; k < e; ++k) {
            const Idx r = row_idx_[static_cast<std::size_t>(k)];
            if (r < 0 || r >= rows_) return bad("row index out of range");
            if (k > b && row_idx_[static_cast<std::size_t>(k) - 1] >= r) return bad("row indices not strictly ascending");
            if (!std::isfinite(values_[static_cast<std::size_t>(k)])) {
                if (why) *why = "non-finite coefficient";
                return Status::NumericalFailure;
            }
        }
    }
    if (why) why->clear();
    return Status::Ok;
}

void CscMatrix::multiply(std::span<const Real> x, std::span<Real> y) const {
    if (static_cast<Idx>(x.size()) != cols_ || static_cast<Idx>(y.size()) != rows_)
        throw Error(Status::DimensionMismatch, "CscMatrix::multiply");
    std::fill(y.begin(), y.end(), Real{0});
    for (Idx c = 0; c < cols_; ++c)
        for (Idx k = col_ptr_[static_cast<std::size_t>(c)]; k < col_ptr_[static_cast<std::size_t>(c) + 1]; ++k)
            y[static_cast<std::size_t>(row_idx_[static_cast<std::size_t>(k)])] +=
                values_[static_cast<std::size_t>(k)] * x[static_cast<std::size_t>(c)];
}

CsrMatrix CscMatrix::to_csr() const {
    const auto R = static_cast<std::size_t>(rows_);
    std::vector<Idx> rp(R + 1, 0);
    for (Idx r : row_idx_) ++rp[static_cast<std::size_t>(r) + 1];
    std::partial_sum(rp.begin(), rp.end(), rp.begin());
    std::vector<Idx> ci(row_idx_.size()), cur(rp.begin(), rp.end() - 1);
    std::vector<Real> v(values_.size());
    for (Idx c = 0; c < cols_; ++c)
        for (Idx k = col_ptr_[static_cast<std::size_t>(c)]; k < col_ptr_[static_cast<std::size_t>(c) + 1]; ++k) {
            const auto d = static_cast<std::size_t>(cur[static_cast<std::size_t>(row_idx_[static_cast<std::size_t>(k)])]++);
            ci[d] = c;
            v[d] = values_[static_cast<std::size_t>(k)];
        }
    return CsrMatrix(rows_, cols_, std::move(rp), std::move(ci), std::move(v));
}

}  // namespace sov
