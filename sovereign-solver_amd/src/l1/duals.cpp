// Dual checks (API: include/sovereign/duals.hpp).  Same semantics as
// reference src/l1/duals.cpp: term ranges (:105-109), Minkowski sums of the
// row side and the column side through d = A'y (:125-158), disjointness with
// slack (:160-168), complementary-slackness violation (:170-202).
// Additive: PrimalRay::certifies_unboundedness.
#include "sovereign/duals.hpp"

#include <algorithm>
#include <cmath>

namespace sov {
namespace {

FarkasCertificate::Range span_of(Real mult, Real lo, Real hi) {
    if (mult == 0.0) return {0.0, 0.0};
    if (mult > 0.0) return {mult * lo, mult * hi};
    return {mult * hi, mult * lo};
}

}  // namespace

FarkasCertificate::Range FarkasCertificate::row_side_range(const Problem& p) const {
    Range acc{0.0, 0.0};
    for (std::size_t i = 0; i < row_multipliers.size(); ++i) {
        const Range t = span_of(row_multipliers[i], p.row_lower()[i], p.row_upper()[i]);
        acc.lo += t.lo;
        acc.hi += t.hi;
    }
    return acc;
}

FarkasCertificate::Range FarkasCertificate::column_side_range(const Problem& p) const {
    const CsrMatrix& A = p.matrix();
    std::vector<Real> d(static_cast<std::size_t>(p.num_cols()), 0.0);
    for (Idx r = 0; r < A.rows(); ++r) {
        const Real y = row_multipliers[static_cast<std::size_t>(r)];
        if (y == 0.0) continue;
        for (Idx k = A.row_ptr()[static_cast<std::size_t>(r)]; k < A.row_ptr()[static_cast<std::size_t>(r) + 1]; ++k)
            d[static_cast<std::size_t>(A.col_idx()[static_cast<std::size_t>(k)])] += y * A.values()[static_cast<std::size_t>(k)];
    }
    Range acc{0.0, 0.0};
    for (std::size_t j = 0; j < d.size(); ++j) {
        const Range t = span_of(d[j], p.col_lower()[j], p.col_upper()[j]);
        acc.lo += t.lo;
        acc.hi += t.hi;
    }
    return acc;
}

bool FarkasCertificate::certifies_infeasibility(const Problem& p, Real slack) const {
    if (row_multipliers.size() != static_cast<std::size_t>(p.num_rows())) return false;
    const Range r = row_side_range(p), c = column_side_range(p);
    if (std::isnan(r.lo) || std::isnan(r.hi) || std::isnan(c.lo) || std::isnan(c.hi)) return false;
    return r.disjoint_from(c, slack);
}

bool PrimalRay::certifies_unboundedness(const Problem& p, Real tol) const {
    const auto n = static_cast<std::size_t>(p.num_cols());
    if (direction.size() != n || n == 0) return false;
    Real scale = 0.0;
    for (Real v : direction) {
        if (!std::isfinite(v)) return false;
        scale = std::max(scale, std::abs(v));
    }
    if (scale <= 0.0) return false;
    const Real sg = p.sense() == ObjSense::Maximize ? -1.0 : 1.0;
    Real obj = 0.0;
    for (std::size_t j = 0; j < n; ++j) {
        const Real d = direction[j] / scale;
        obj += sg * p.objective()[j] * d;
        if (is_finite_bound(p.col_lower()[j]) && d < -tol) return false;
        if (is_finite_bound(p.col_upper()[j]) && d > tol) return false;
    }
    if (!(obj < -tol)) return false;
    std::vector<Real> dn(n), ad(static_cast<std::size_t>(p.num_rows()));
    for (std::size_t j = 0; j < n; ++j) dn[j] = direction[j] / scale;
    p.matrix().multiply(dn, ad);
    for (std::size_t i = 0; i < ad.size(); ++i) {
        if (is_finite_bound(p.row_lower()[i]) && ad[i] < -tol) return false;
        if (is_finite_bound(p.row_upper()[i]) && ad[i] > tol) return false;
    }
    return true;
}

Real DualSolution::complementary_slackness_violation(const Problem& p, std::span<const Real> x, Real tol) const {
    Real worst = 0.0;
    if (!row_duals.empty()) {
        std::vector<Real> act(static_cast<std::size_t>(p.num_rows()), 0.0);
        p.matrix().multiply(x, act);
        for (std::size_t i = 0; i < row_duals.size(); ++i) {
            const Real lo = p.row_lower()[i], hi = p.row_upper()[i];
            const bool at_lo = is_finite_bound(lo) && std::abs(act[i] - lo) <= tol;
            const bool at_hi = is_finite_bound(hi) && std::abs(act[i] - hi) <= tol;
            if (!at_lo && !at_hi) worst = std::max(worst, std::abs(row_duals[i]));
        }
    }
    if (!reduced_costs.empty()) {
        for (std::size_t j = 0; j < reduced_costs.size(); ++j) {
            const Real lo = p.col_lower()[j], hi = p.col_upper()[j];
            const bool at_lo = is_finite_bound(lo) && std::abs(x[j] - lo) <= tol;
            const bool at_hi = is_finite_bound(hi) && std::abs(x[j] - hi) <= tol;
            if (!at_lo && !at_hi) worst = std::max(worst, std::abs(reduced_costs[j]));
        }
    }
    return worst;
}

}  // namespace sov
