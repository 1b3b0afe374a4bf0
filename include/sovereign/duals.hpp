// Dual solution checks and Farkas certificates.  API mirrors
// include/sovereign/duals.hpp:47-87 of the reference.  Additive:
// PrimalRay::certifies_unboundedness, the checker for the dual-infeasibility
// (unbounded LP) rays the PDHG engine emits.
#pragma once

#include <span>
#include <vector>

#include "sovereign/numeric.hpp"
#include "sovereign/problem.hpp"

namespace sov {

struct FarkasCertificate {
    std::vector<Real> row_multipliers;
    bool empty() const noexcept { return row_multipliers.empty(); }

    struct Range {
        Real lo = -kInfinity;
        Real hi = kInfinity;
        bool disjoint_from(const Range& other, Real slack) const noexcept {
            return hi < other.lo - slack || other.hi < lo - slack;
        }
    };
    Range row_side_range(const Problem& problem) const;
    Range column_side_range(const Problem& problem) const;
    bool certifies_infeasibility(const Problem& problem, Real slack = 1e-6) const;
};

/// Additive: a primal ray d proving the LP unbounded (feasible and improving
/// without limit).  Valid when, relative to ||d||_inf = 1: the objective
/// strictly improves (c'd < -tol in minimize sense), d stays in the
/// recession cone of the column box (d_j >= -tol where col_lo is finite,
/// d_j <= tol where col_hi is finite) and A d stays in the recession cone of
/// the row box.  Unboundedness also needs a feasible point, which this check
/// does not establish; callers pair it with a primal-feasibility test.
struct PrimalRay {
    std::vector<Real> direction;
    bool empty() const noexcept { return direction.empty(); }
    bool certifies_unboundedness(const Problem& problem, Real tol = 1e-6) const;
};

struct DualSolution {
    std::vector<Real> row_duals;
    std::vector<Real> reduced_costs;
    bool empty() const noexcept { return row_duals.empty() && reduced_costs.empty(); }
    Real complementary_slackness_violation(const Problem& problem, std::span<const Real> x,
                                           Real bound_tol = 1e-7) const;
};

}  // namespace sov
