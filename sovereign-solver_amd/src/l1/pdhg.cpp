// PDHG engine (API: include/sovereign/pdhg.hpp).
//
// solve() runs one of two loops:
//   solve_primitive -- reference src/l1/pdhg.cpp:218-502 written against the
//     virtual Backend, with every Status checked (the reference discards
//     them, e.g. :297, :316).  Any backend; the parity mode.
//   solve_fused     -- the MI355X loop on a Hip backend: device scaling
//     (sov_hip_scale_*), device power iteration, two fused launches per
//     iteration (sov_hip_pdhg_primal_step / _dual_step, replayed as HIP graphs
//     per check interval), FP64 device checks (sov_hip_pdhg_check) and device
//     restarts.  Same decisions as the reference: candidate = average when its
//     scaled residual is no worse (:354), converged when the candidate's
//     original-space residual < tolerance (:383), restart to the average when
//     its residual <= decrease * residual at the last restart (:391-393).
#include "sovereign/pdhg.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <system_error>
#include <thread>

#include <sched.h>
#include <utility>
#include <vector>

#include "sov_hip.h"
#include "sovereign/scaling.hpp"

namespace sov {

const char* to_string(PdhgStatus s) noexcept {
    switch (s) {
        case PdhgStatus::Optimal: return "optimal";
        case PdhgStatus::IterationLimit: return "iteration_limit";
        case PdhgStatus::TimeLimit: return "time_limit";
        case PdhgStatus::NumericalFailure: return "numerical_failure";
        case PdhgStatus::NotSolved: return "not_solved";
        case PdhgStatus::Cancelled: return "cancelled";
        case PdhgStatus::PrimalInfeasible: return "primal_infeasible";
        case PdhgStatus::DualInfeasible: return "dual_infeasible";
    }
    return "unknown";
}

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

/// inf over [lo, hi] of coeff * v (reference pdhg.cpp:36-40).
Real inf_over_box(Real coeff, Real lo, Real hi) {
    if (coeff == 0.0) return 0.0;
    if (coeff > 0.0) return is_finite_bound(lo) ? coeff * lo : -kInfinity;
    return is_finite_bound(hi) ? coeff * hi : -kInfinity;
}

Real dual_class(Real val, Real lo, Real hi, Real point, Real tol) {
    const bool at_lo = is_finite_bound(lo) && std::abs(point - lo) < tol;
    const bool at_hi = is_finite_bound(hi) && std::abs(point - hi) < tol;
    if (!at_lo && !at_hi) return std::abs(val);
    if (at_lo && !at_hi) return std::max(Real{0}, -val);
    if (at_hi && !at_lo) return std::max(Real{0}, val);
    return 0.0;
}

/// Scaled-space max-norm residual (reference pdhg.cpp:87-136).
Real kkt_residual(const Problem& s, std::span<const Real> c, std::span<const Real> x, std::span<const Real> y,
                  Real tol) {
    std::vector<Real> act(static_cast<std::size_t>(s.num_rows()));
    s.matrix().multiply(x, act);
    Real pr = 0.0, dr = 0.0;
    for (std::size_t i = 0; i < act.size(); ++i) {
        const Real lo = s.row_lower()[i], hi = s.row_upper()[i];
        if (is_finite_bound(lo)) pr = std::max(pr, lo - act[i]);
        if (is_finite_bound(hi)) pr = std::max(pr, act[i] - hi);
        dr = std::max(dr, dual_class(y[i], lo, hi, act[i], tol));
    }
    std::vector<Real> aty(static_cast<std::size_t>(s.num_cols()));
    s.matrix().multiply_transpose(y, aty);
    for (std::size_t j = 0; j < aty.size(); ++j)
        dr = std::max(dr, dual_class(c[j] - aty[j], s.col_lower()[j], s.col_upper()[j], x[j], tol));
    return std::max(pr, dr);
}

/// Original-space residual of an already-unscaled point (reference
/// pdhg.cpp:145-199 after its unscale at :148-152).
Real original_residual(const Problem& p, Real sign, std::span<const Real> x, std::span<const Real> y, Real tol) {
    std::vector<Real> act(static_cast<std::size_t>(p.num_rows()));
    p.matrix().multiply(x, act);
    Real pr = 0.0, dr = 0.0;
    for (std::size_t i = 0; i < act.size(); ++i) {
        const Real lo = p.row_lower()[i], hi = p.row_upper()[i];
        if (is_finite_bound(lo)) pr = std::max(pr, lo - act[i]);
        if (is_finite_bound(hi)) pr = std::max(pr, act[i] - hi);
        dr = std::max(dr, dual_class(sign * y[i], lo, hi, act[i], tol));
    }
    std::vector<Real> aty(static_cast<std::size_t>(p.num_cols()));
    p.matrix().multiply_transpose(y, aty);
    for (std::size_t j = 0; j < aty.size(); ++j)
        dr = std::max(dr, dual_class(sign * (p.objective()[j] - aty[j]), p.col_lower()[j], p.col_upper()[j], x[j], tol));
    return std::max(pr, dr);
}

struct RelKkt {
    Real rp = 0, rd = 0, gap = 0, pobj = 0, dobj = 0;
    Real worst() const { return std::max({rp, rd, gap}); }
};

/// Relative KKT of an original-space point (definitions in pdhg.hpp).
/// act = A x and aty = A^T y, as CsrMatrix::multiply / multiply_transpose
/// give them.
RelKkt relative_kkt(const Problem& p, std::span<const Real> x, std::span<const Real> y, std::span<const Real> act,
                    std::span<const Real> aty) {
    const Real sg = p.sense() == ObjSense::Maximize ? -1.0 : 1.0;
    Real pv2 = 0, bn2 = 0, dr2 = 0, cn2 = 0, pobj = 0, dobj = 0;
    for (std::size_t i = 0; i < act.size(); ++i) {
        const Real lo = p.row_lower()[i], hi = p.row_upper()[i];
        Real v = 0.0;
        if (is_finite_bound(lo) && act[i] < lo) v = lo - act[i];
        if (is_finite_bound(hi) && act[i] > hi) v = act[i] - hi;
        pv2 += v * v;
        if (is_finite_bound(lo)) bn2 += lo * lo;
        if (is_finite_bound(hi) && !(is_finite_bound(lo) && lo == hi)) bn2 += hi * hi;
        const Real yi = sg * y[i];
        Real yp = yi;
        if (!is_finite_bound(lo) && yp > 0.0) yp = 0.0;
        if (!is_finite_bound(hi) && yp < 0.0) yp = 0.0;
        dr2 += (yi - yp) * (yi - yp);
        if (yp > 0.0) dobj += yp * lo;
        else if (yp < 0.0) dobj += yp * hi;
    }
    for (std::size_t j = 0; j < aty.size(); ++j) {
        const Real ci = sg * p.objective()[j];
        const Real rc = ci - sg * aty[j];
        const Real lo = p.col_lower()[j], hi = p.col_upper()[j];
        Real lam = rc;
        if (!is_finite_bound(lo) && lam > 0.0) lam = 0.0;
        if (!is_finite_bound(hi) && lam < 0.0) lam = 0.0;
        dr2 += (rc - lam) * (rc - lam);
        cn2 += ci * ci;
        if (lam > 0.0) dobj += lam * lo;
        else if (lam < 0.0) dobj += lam * hi;
        pobj += ci * x[j];
    }
    RelKkt r;
    r.pobj = sg * pobj + p.objective_constant();
    r.dobj = sg * dobj + p.objective_constant();
    r.rp = std::sqrt(pv2) / (1.0 + std::sqrt(bn2));
    r.rd = std::sqrt(dr2) / (1.0 + std::sqrt(cn2));
    r.gap = std::abs(r.pobj - r.dobj) / (1.0 + std::abs(r.pobj) + std::abs(r.dobj));
    return r;
}

RelKkt relative_kkt(const Problem& p, std::span<const Real> x, std::span<const Real> y) {
    std::vector<Real> act(static_cast<std::size_t>(p.num_rows())), aty(static_cast<std::size_t>(p.num_cols()));
    p.matrix().multiply(x, act);
    p.matrix().multiply_transpose(y, aty);
    return relative_kkt(p, x, y, act, aty);
}

/// Unscale a scaled, internal-sense point (reference pdhg.cpp:205-214).
std::pair<std::vector<Real>, std::vector<Real>> unscale_point(Real sign, const Scaling& sc, std::span<const Real> xs,
                                                              std::span<const Real> ys) {
    std::vector<Real> x(xs.begin(), xs.end()), y(ys.begin(), ys.end());
    sc.unscale_primal(x);
    for (Real& v : y) v *= sign;
    sc.unscale_dual(y);
    return {std::move(x), std::move(y)};
}

/// Result population from the final original-space point (reference
/// pdhg.cpp:428-498) plus the additive relative-KKT fields.
void finish_result(const Problem& p, const PdhgOptions& o, std::vector<Real> x, std::vector<Real> y, PdhgResult& r) {
    r.primal = std::move(x);
    r.dual = std::move(y);
    r.objective = p.evaluate_objective(r.primal);
    const auto n = static_cast<std::size_t>(p.num_cols()), m = static_cast<std::size_t>(p.num_rows());
    std::vector<Real> aty(n);
    p.matrix().multiply_transpose(r.dual, aty);
    r.reduced_costs.resize(n);
    for (std::size_t j = 0; j < n; ++j) r.reduced_costs[j] = p.objective()[j] - aty[j];
    std::vector<Real> act(m);
    p.matrix().multiply(r.primal, act);
    r.primal_infeasibility = 0.0;
    for (std::size_t i = 0; i < m; ++i) {
        if (is_finite_bound(p.row_lower()[i])) r.primal_infeasibility = std::max(r.primal_infeasibility, p.row_lower()[i] - act[i]);
        if (is_finite_bound(p.row_upper()[i])) r.primal_infeasibility = std::max(r.primal_infeasibility, act[i] - p.row_upper()[i]);
    }
    DualSolution d;
    d.row_duals = r.dual;
    d.reduced_costs = r.reduced_costs;
    r.dual_infeasibility = d.complementary_slackness_violation(p, r.primal, std::max(o.tolerance, Real{1e-6}));
    Real dobj = 0.0;
    bool finite = true;
    for (std::size_t j = 0; j < n && finite; ++j) {
        const Real t = inf_over_box(r.reduced_costs[j], p.col_lower()[j], p.col_upper()[j]);
        if (!std::isfinite(t)) finite = false;
        else dobj += t;
    }
    for (std::size_t i = 0; i < m && finite; ++i) {
        const Real t = inf_over_box(r.dual[i], p.row_lower()[i], p.row_upper()[i]);
        if (!std::isfinite(t)) finite = false;
        else dobj += t;
    }
    r.duality_gap = finite ? std::abs(r.objective - dobj) / (1.0 + std::abs(r.objective) + std::abs(dobj)) : kInfinity;
    const RelKkt k = relative_kkt(p, r.primal, r.dual, act, aty);   // the products above, not again
    r.relative_primal_residual = k.rp;
    r.relative_dual_residual = k.rd;
    r.relative_gap = k.gap;
}

/// Additive infeasibility detection.  On an infeasible or unbounded LP the
/// PDHG iterates drift along the infimal displacement vector (Applegate,
/// Diaz, Lu, Lubin, "Infeasibility detection with primal-dual hybrid
/// gradient for large-scale linear programming", 2021): between two checks
/// with no restart in between, y moves along a Farkas ray and x along a
/// primal ray.  The drifts, in original units and scaled to max-norm 1, go to
/// the exact checkers (FarkasCertificate, PrimalRay); nothing is reported
/// that they do not verify.
struct Drift {
    std::vector<Real> x, y;   // original-space iterates at the previous check
    bool valid = false;

    void reset() { valid = false; }

    static bool unit(std::vector<Real>& v) {
        Real s = 0.0;
        for (Real t : v) {
            if (!std::isfinite(t)) return false;
            s = std::max(s, std::abs(t));
        }
        if (s <= 0.0) return false;
        for (Real& t : v) t /= s;
        return true;
    }

    /// Test the drift since the previous check; remember (ux, uy).  Sets the
    /// status and certificate of r and returns true when one verifies.
    bool test(const Problem& p, std::vector<Real> ux, std::vector<Real> uy, PdhgResult& r) {
        bool found = false;
        if (valid) {
            std::vector<Real> dy(uy.size()), dx(ux.size());
            for (std::size_t i = 0; i < dy.size(); ++i) dy[i] = uy[i] - y[i];
            for (std::size_t j = 0; j < dx.size(); ++j) dx[j] = ux[j] - x[j];
            FarkasCertificate fc;
            if (unit(dy)) {
                fc.row_multipliers = std::move(dy);
                if (fc.certifies_infeasibility(p)) {
                    r.status = PdhgStatus::PrimalInfeasible;
                    r.certificate = std::move(fc);
                    r.message = "primal infeasible: Farkas certificate from the dual drift";
                    found = true;
                }
            }
            PrimalRay ray;
            if (!found && unit(dx)) {
                ray.direction = std::move(dx);
                if (ray.certifies_unboundedness(p)) {
                    r.status = PdhgStatus::DualInfeasible;
                    r.ray = std::move(ray);
                    r.message = "dual infeasible: improving primal ray from the primal drift";
                    found = true;
                }
            }
        }
        x = std::move(ux);
        y = std::move(uy);
        valid = true;
        return found;
    }
};

bool prevalidate(const Problem& p, PdhgResult& r) {
    std::string why;
    if (p.validate(&why) != Status::Ok) {
        r.status = PdhgStatus::NumericalFailure;
        r.message = "invalid problem: " + why;
        return false;
    }
    if (p.has_crossed_bounds()) {
        r.status = PdhgStatus::NumericalFailure;
        r.message = "a bound pair is crossed (lower > upper)";
        return false;
    }
    return true;
}

Real next_primal_weight(Real omega, Real dx2, Real dy2, Real theta) {
    const Real dx = std::sqrt(dx2), dy = std::sqrt(dy2);
    if (dx > 1e-10 && dy > 1e-10) return std::exp(theta * std::log(dy / dx) + (1.0 - theta) * std::log(omega));
    return omega;
}

}  // namespace

// ===========================================================================
// dispatch
// ===========================================================================

PdhgResult Pdhg::solve(const Problem& problem, Backend& backend, const CancellationToken* cancel) const {
    const bool fused_ok = backend.kind() == BackendKind::Hip && backend.native_context() != nullptr;
    PdhgEngine e = options_.engine;
    if (e == PdhgEngine::Auto) e = fused_ok ? PdhgEngine::Fused : PdhgEngine::Primitive;
    if (e == PdhgEngine::Fused && !fused_ok) {
        PdhgResult r;
        r.status = PdhgStatus::NumericalFailure;
        r.message = "the fused engine needs a Hip backend";
        return r;
    }
    if (e == PdhgEngine::Primitive && options_.precision != Precision::F64) {
        PdhgResult r;
        r.status = PdhgStatus::NumericalFailure;
        r.message = "FP32 iterates need the fused engine (Hip backend)";
        return r;
    }
    return e == PdhgEngine::Fused ? solve_fused(problem, backend, cancel) : solve_primitive(problem, backend, cancel);
}

// ===========================================================================
// primitive loop (reference pdhg.cpp:218-502 over Backend primitives)
// ===========================================================================

PdhgResult Pdhg::solve_primitive(const Problem& problem, Backend& be, const CancellationToken* cancel) const {
    PdhgResult result;
    result.engine = "primitive";
    const auto t0 = Clock::now();
    if (!prevalidate(problem, result)) return result;
    const PdhgOptions& o = options_;
    try {
        auto [scaled, scaling] = Scaling::equilibrate(problem);
        const Idx n = scaled.num_cols(), m = scaled.num_rows();
        const auto nu = static_cast<std::size_t>(n), mu = static_cast<std::size_t>(m);
        DeviceCsr A = DeviceCsr::upload(be, scaled.matrix());
        const Real sign = scaled.sense() == ObjSense::Maximize ? -1.0 : 1.0;
        std::vector<Real> c_host(scaled.objective().begin(), scaled.objective().end());
        for (Real& v : c_host) v *= sign;
        DeviceBuffer c = DeviceBuffer::upload(be, c_host);
        DeviceBuffer col_lo = DeviceBuffer::upload(be, scaled.col_lower());
        DeviceBuffer col_hi = DeviceBuffer::upload(be, scaled.col_upper());
        DeviceBuffer row_lo = DeviceBuffer::upload(be, scaled.row_lower());
        DeviceBuffer row_hi = DeviceBuffer::upload(be, scaled.row_upper());

        auto ok = [](Status s, const char* what) { check(s, what); };

        // power iteration (reference pdhg.cpp:48-79)
        Real est = 0.0;
        {
            DeviceBuffer u = DeviceBuffer::upload(be, std::vector<Real>(nu, 1.0));
            DeviceBuffer v = DeviceBuffer::upload(be, std::vector<Real>(mu, 1.0));
            DeviceBuffer w(be, mu);
            for (int it = 0; it < o.power_iterations; ++it) {
                ok(be.spmv(A.view(), -1.0, u.data(), 0.0, w.data()), "power: spmv");
                ok(be.axpy(m, 1.0, v.data(), w.data()), "power: axpy");
                Real wn = 0.0;
                ok(be.norm2(m, w.data(), &wn), "power: norm2");
                if (wn < 1e-300) break;
                ok(be.scale(m, 1.0 / wn, w.data()), "power: scale");
                ok(be.spmv_transpose(A.view(), -1.0, w.data(), 0.0, u.data()), "power: spmv_transpose");
                ok(be.copy_device_to_device(w.data(), v.data(), mu * sizeof(Real)), "power: copy");
                Real un = 0.0, vn = 0.0;
                ok(be.norm2(n, u.data(), &un), "power: norm2");
                ok(be.norm2(m, v.data(), &vn), "power: norm2");
                est = std::sqrt(un * un + vn * vn);
                if (est < 1e-300) break;
                ok(be.scale(n, 1.0 / est, u.data()), "power: scale");
                ok(be.scale(m, 1.0 / est, v.data()), "power: scale");
            }
        }
        result.operator_norm_estimate = est;
        const Real eta = est > 1e-9 ? o.step_size_safety / est : o.step_size_safety;
        Real omega = o.use_primal_weight ? o.initial_primal_weight : 1.0;
        Real tau = eta / omega, sigma = eta * omega;

        DeviceBuffer x(be, nu), s(be, mu), y(be, mu);
        ok(be.project_box(n, col_lo.data(), col_hi.data(), x.data()), "project_box");
        ok(be.project_box(m, row_lo.data(), row_hi.data(), s.data()), "project_box");
        if (!o.warm_start_dual.empty()) {
            if (o.warm_start_dual.size() != mu) throw Error(Status::DimensionMismatch, "warm_start_dual");
            std::vector<Real> ys(mu);
            for (std::size_t i = 0; i < mu; ++i) ys[i] = sign * o.warm_start_dual[i] / scaling.row_scale()[i];
            y = DeviceBuffer::upload(be, ys);
        }
        if (!o.warm_start_primal.empty()) {
            // x = Proj_X(x0 / c^s), s = Proj_S(A_s x) (additive)
            if (o.warm_start_primal.size() != nu) throw Error(Status::DimensionMismatch, "warm_start_primal");
            std::vector<Real> xs(nu), act(mu, 0.0);
            for (std::size_t j = 0; j < nu; ++j) xs[j] = o.warm_start_primal[j] / scaling.col_scale()[j];
            x = DeviceBuffer::upload(be, xs);
            ok(be.project_box(n, col_lo.data(), col_hi.data(), x.data()), "project_box");
            const std::vector<Real> xp = x.to_host();
            scaled.matrix().multiply(xp, act);
            s = DeviceBuffer::upload(be, act);
            ok(be.project_box(m, row_lo.data(), row_hi.data(), s.data()), "project_box");
        }
        DeviceBuffer x_new(be, nu), s_new(be, mu), x_bar(be, nu), s_bar(be, mu), grad(be, nu), kb(be, mu);
        DeviceBuffer x_sum(be, nu), y_sum(be, mu);
        std::vector<Real> x_last = x.to_host(), y_last = y.to_host();
        long long since = 0;
        Real res_at_restart = kInfinity, prev_cand = kInfinity;
        const std::size_t nb = nu * sizeof(Real), mb = mu * sizeof(Real);
        std::vector<Real> best_x, best_y;
        Drift drift;
        bool converged = false;
        const Real otol = std::max(o.tolerance, Real{1e-6});
        const auto t_loop = Clock::now();
        result.setup_seconds = seconds_since(t0);

        for (result.iterations = 0; result.iterations < o.max_iterations; ++result.iterations) {
            if (cancel != nullptr && cancel->is_cancelled()) { result.status = PdhgStatus::Cancelled; break; }
            if (o.time_limit_seconds > 0.0 && seconds_since(t0) > o.time_limit_seconds) {
                result.status = PdhgStatus::TimeLimit;
                break;
            }
            ok(be.copy_device_to_device(c.data(), grad.data(), nb), "copy");
            ok(be.spmv_transpose(A.view(), -1.0, y.data(), 1.0, grad.data()), "spmv_transpose");
            ok(be.copy_device_to_device(x.data(), x_new.data(), nb), "copy");
            ok(be.axpy(n, -tau, grad.data(), x_new.data()), "axpy");
            ok(be.project_box(n, col_lo.data(), col_hi.data(), x_new.data()), "project_box");
            ok(be.copy_device_to_device(s.data(), s_new.data(), mb), "copy");
            ok(be.axpy(m, -tau, y.data(), s_new.data()), "axpy");
            ok(be.project_box(m, row_lo.data(), row_hi.data(), s_new.data()), "project_box");
            ok(be.copy_device_to_device(x_new.data(), x_bar.data(), nb), "copy");
            ok(be.scale(n, 2.0, x_bar.data()), "scale");
            ok(be.axpy(n, -1.0, x.data(), x_bar.data()), "axpy");
            ok(be.copy_device_to_device(s_new.data(), s_bar.data(), mb), "copy");
            ok(be.scale(m, 2.0, s_bar.data()), "scale");
            ok(be.axpy(m, -1.0, s.data(), s_bar.data()), "axpy");
            ok(be.spmv(A.view(), -1.0, x_bar.data(), 0.0, kb.data()), "spmv");
            ok(be.axpy(m, 1.0, s_bar.data(), kb.data()), "axpy");
            ok(be.axpy(m, sigma, kb.data(), y.data()), "axpy");
            ok(be.copy_device_to_device(x_new.data(), x.data(), nb), "copy");
            ok(be.copy_device_to_device(s_new.data(), s.data(), mb), "copy");
            if (since == 0) {
                ok(be.copy_device_to_device(x.data(), x_sum.data(), nb), "copy");
                ok(be.copy_device_to_device(y.data(), y_sum.data(), mb), "copy");
            } else {
                ok(be.axpy(n, 1.0, x.data(), x_sum.data()), "axpy");
                ok(be.axpy(m, 1.0, y.data(), y_sum.data()), "axpy");
            }
            ++since;
            if ((result.iterations + 1) % o.restart_check_period != 0) continue;

            const std::vector<Real> xh = x.to_host(), yh = y.to_host();
            const Real cur = kkt_residual(scaled, c_host, xh, yh, o.tolerance);
            std::vector<Real> xa = x_sum.to_host(), ya = y_sum.to_host();
            const Real inv = 1.0 / static_cast<Real>(since);
            for (Real& v : xa) v *= inv;
            for (Real& v : ya) v *= inv;
            const Real avg = kkt_residual(scaled, c_host, xa, ya, o.tolerance);
            if (!std::isfinite(cur) || !std::isfinite(avg)) {
                result.status = PdhgStatus::NumericalFailure;
                result.message = "non-finite KKT residual at iteration " + std::to_string(result.iterations);
                return result;
            }
            const bool adaptive = o.restart_scheme == PdhgRestartScheme::Adaptive;
            Real rel_c = 0.0, rel_a = 0.0;
            if (adaptive || o.termination == PdhgTermination::RelativeKkt) {
                auto [uxc, uyc] = unscale_point(sign, scaling, xh, yh);
                auto [uxa, uya] = unscale_point(sign, scaling, xa, ya);
                rel_c = relative_kkt(problem, uxc, uyc).worst();
                rel_a = relative_kkt(problem, uxa, uya).worst();
            }
            const bool avg_better = adaptive ? rel_a <= rel_c : avg <= cur;
            const std::vector<Real>& cx = avg_better ? xa : xh;
            const std::vector<Real>& cy = avg_better ? ya : yh;
            auto [ux, uy] = unscale_point(sign, scaling, cx, cy);
            const Real ores = original_residual(problem, sign, ux, uy, otol);
            if (!std::isfinite(ores)) {
                result.status = PdhgStatus::NumericalFailure;
                result.message = "non-finite KKT residual (original units) at iteration " + std::to_string(result.iterations);
                return result;
            }
            if (o.checkpoint) o.checkpoint(result.iterations, ux, uy, ores);
            const Real rel_cand = avg_better ? rel_a : rel_c;
            const bool stop = o.termination == PdhgTermination::RelativeKkt ? rel_cand <= o.tolerance : ores < o.tolerance;
            if (stop) {
                best_x = cx;
                best_y = cy;
                converged = true;
                ++result.iterations;
                break;
            }
            if (o.detect_infeasibility) {
                auto [dxo, dyo] = unscale_point(sign, scaling, xh, yh);
                if (drift.test(problem, std::move(dxo), std::move(dyo), result)) {
                    best_x = xh;
                    best_y = yh;
                    ++result.iterations;
                    break;
                }
            }
            bool do_restart = false;
            if (!adaptive) {
                if (!std::isfinite(res_at_restart)) res_at_restart = cur;
                do_restart = avg <= o.restart_sufficient_decrease * res_at_restart;
            } else {
                if (!std::isfinite(res_at_restart)) res_at_restart = rel_c;
                do_restart = rel_cand <= o.restart_beta_sufficient * res_at_restart ||
                             (rel_cand <= o.restart_beta_necessary * res_at_restart && rel_cand > prev_cand) ||
                             static_cast<Real>(since) >= o.restart_beta_artificial * static_cast<Real>(result.iterations + 1);
                prev_cand = rel_cand;
            }
            if (do_restart) {
                drift.reset();
                if (!adaptive || avg_better) {
                    x = DeviceBuffer::upload(be, xa);
                    y = DeviceBuffer::upload(be, ya);
                    ok(be.project_box(n, col_lo.data(), col_hi.data(), x.data()), "project_box");
                    std::vector<Real> act(mu, 0.0);
                    scaled.matrix().multiply(xa, act);
                    s = DeviceBuffer::upload(be, act);
                    ok(be.project_box(m, row_lo.data(), row_hi.data(), s.data()), "project_box");
                }
                if (o.use_primal_weight) {
                    const std::vector<Real> xn = x.to_host(), yn = y.to_host();
                    Real dx2 = 0.0, dy2 = 0.0;
                    for (std::size_t j = 0; j < nu; ++j) dx2 += (xn[j] - x_last[j]) * (xn[j] - x_last[j]);
                    for (std::size_t i = 0; i < mu; ++i) dy2 += (yn[i] - y_last[i]) * (yn[i] - y_last[i]);
                    omega = next_primal_weight(omega, dx2, dy2, o.primal_weight_smoothing);
                    tau = eta / omega;
                    sigma = eta * omega;
                    x_last = xn;
                    y_last = yn;
                }
                res_at_restart = adaptive ? rel_cand : avg;
                prev_cand = kInfinity;
                since = 0;
                ++result.restarts;
                if (o.verbose)
                    std::fprintf(stderr, "  [pdhg] restart at iter %lld, residual %.3g\n", result.iterations + 1,
                                 static_cast<double>(adaptive ? rel_cand : avg));
            }
        }
        result.loop_seconds = seconds_since(t_loop);
        if (best_x.empty()) {
            best_x = x.to_host();
            best_y = y.to_host();
        }
        if (converged) result.status = PdhgStatus::Optimal;
        else if (result.status == PdhgStatus::NotSolved) result.status = PdhgStatus::IterationLimit;
        auto [fx, fy] = unscale_point(sign, scaling, best_x, best_y);
        result.primal_weight = omega;
        finish_result(problem, o, std::move(fx), std::move(fy), result);
    } catch (const Error& e) {
        result.status = PdhgStatus::NumericalFailure;
        result.message = std::string("backend failure: ") + e.what();
    }
    result.seconds = seconds_since(t0);
    return result;
}

// ===========================================================================
// fused device loop
// ===========================================================================

namespace {

/// Device allocation owned by the fused loop.
struct DevMem {
    sov_hip_ctx* ctx = nullptr;
    void* p = nullptr;
    DevMem() = default;
    DevMem(sov_hip_ctx* c, std::size_t bytes) : ctx(c) {
        const int rc = sov_hip_malloc(c, bytes, &p);
        if (rc != SOV_OK) throw Error(static_cast<Status>(rc), std::string("device allocation: ") + sov_hip_last_error());
    }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : ctx(o.ctx), p(std::exchange(o.p, nullptr)) {}
    DevMem& operator=(DevMem&& o) noexcept {
        if (this != &o) {
            reset();
            ctx = o.ctx;
            p = std::exchange(o.p, nullptr);
        }
        return *this;
    }
    ~DevMem() { reset(); }
    void reset() {
        if (p != nullptr) (void)sov_hip_free(ctx, p);
        p = nullptr;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

void hip_ok(int rc, const char* what) {
    if (rc != SOV_OK) throw Error(static_cast<Status>(rc), std::string(what) + ": " + sov_hip_last_error());
}

struct CsrHandle {
    sov_hip_csr* h = nullptr;
    ~CsrHandle() { sov_hip_csr_destroy(h); }
};

/// Host threads this process may run on: the CPUs of its affinity mask (so a
/// run pinned to one core gets one thread), at most 16.
unsigned host_threads() {
    static const unsigned count = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        int c = 0;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) c = CPU_COUNT(&set);
        if (c <= 0) c = static_cast<int>(std::thread::hardware_concurrency());
        return static_cast<unsigned>(std::clamp(c, 1, 16));
    }();
    return count;
}

/// f(begin, end) over [0, n) in contiguous chunks on up to host_threads()
/// threads (one below 2^18 elements).  For element-wise work: every element
/// is computed as in a serial loop, so results do not depend on the split.
/// A thread that cannot be started leaves its chunk to the calling thread;
/// every started thread is joined before this returns (also when f throws).
template <class F>
void parallel_for(std::size_t n, F f) {
    const std::size_t parts = n < (std::size_t{1} << 18) ? 1 : host_threads();
    if (parts <= 1) {
        f(std::size_t{0}, n);
        return;
    }
    struct Joiner {
        std::vector<std::thread> pool;
        ~Joiner() {
            for (auto& t : pool)
                if (t.joinable()) t.join();
        }
    } j;
    j.pool.reserve(parts - 1);
    std::size_t k = 1;
    for (; k < parts; ++k) {
        try {
            j.pool.emplace_back(f, n * k / parts, n * (k + 1) / parts);
        } catch (const std::system_error&) {
            break;   // no more threads: the caller takes the rest
        }
    }
    f(std::size_t{0}, n / parts);
    for (; k < parts; ++k) f(n * k / parts, n * (k + 1) / parts);
}

/// Host copy of the handle's device order (sov_hip_csr_order): every vector
/// the fused kernels touch is stored with element p = natural index perm[p].
struct DeviceOrder {
    std::vector<int32_t> row, col;
    static std::vector<Real> to_dev(std::span<const Real> v, const std::vector<int32_t>& perm) {
        std::vector<Real> out(v.size());
        parallel_for(out.size(), [&](std::size_t b, std::size_t e) {
            for (std::size_t p = b; p < e; ++p) out[p] = v[static_cast<std::size_t>(perm[p])];
        });
        return out;
    }
    static std::vector<Real> to_nat(std::span<const Real> v, const std::vector<int32_t>& perm) {
        std::vector<Real> out(v.size());
        parallel_for(out.size(), [&](std::size_t b, std::size_t e) {
            for (std::size_t p = b; p < e; ++p) out[static_cast<std::size_t>(perm[p])] = v[p];
        });
        return out;
    }
};

struct GraphHandle {
    sov_hip_graph* g = nullptr;
    ~GraphHandle() { sov_hip_graph_destroy(g); }
    void reset() {
        sov_hip_graph_destroy(g);
        g = nullptr;
    }
};

/// The device-resident iterate set of one precision.
struct DeviceState {
    int prec = SOV_PREC_F64;
    std::size_t es = 8;
    DevMem c, clo, chi, rlo, rhi, x, xbar, xsum, s, y, ysum;
    sov_hip_pdhg_state st{};

    void allocate(sov_hip_ctx* ctx, int p, std::size_t m, std::size_t n) {
        prec = p;
        es = p == SOV_PREC_F64 ? 8 : 4;
        const std::size_t nb = std::max<std::size_t>(n, 1) * es, mb = std::max<std::size_t>(m, 1) * es;
        c = DevMem(ctx, nb); clo = DevMem(ctx, nb); chi = DevMem(ctx, nb);
        x = DevMem(ctx, nb); xbar = DevMem(ctx, nb); xsum = DevMem(ctx, nb);
        rlo = DevMem(ctx, mb); rhi = DevMem(ctx, mb);
        s = DevMem(ctx, mb); y = DevMem(ctx, mb); ysum = DevMem(ctx, mb);
        for (DevMem* d : {&x, &xbar, &xsum, &s, &y, &ysum}) hip_ok(sov_hip_fill_zero(ctx, d->p, d == &s || d == &y || d == &ysum ? mb : nb), "fill");
        st.prec = p;
        st.m = static_cast<int32_t>(m);
        st.n = static_cast<int32_t>(n);
        st.c = c.p; st.col_lo = clo.p; st.col_hi = chi.p; st.row_lo = rlo.p; st.row_hi = rhi.p;
        st.x = x.p; st.x_bar = xbar.p; st.x_sum = xsum.p; st.s = s.p; st.y = y.p; st.y_sum = ysum.p;
    }

    /// Upload an FP64 host vector into a device vector of this precision.
    void upload(sov_hip_ctx* ctx, const DevMem& dst, std::span<const Real> v) const {
        if (v.empty()) return;
        if (prec == SOV_PREC_F64) {
            hip_ok(sov_hip_memcpy_h2d(ctx, dst.p, v.data(), v.size() * 8), "upload");
        } else {
            std::vector<float> f(v.size());
            for (std::size_t i = 0; i < v.size(); ++i) f[i] = static_cast<float>(v[i]);
            hip_ok(sov_hip_memcpy_h2d(ctx, dst.p, f.data(), f.size() * 4), "upload");
        }
    }

    /// Download a device vector as FP64, optionally times `scale`.
    std::vector<Real> download(sov_hip_ctx* ctx, const DevMem& src, std::size_t len, Real scale = 1.0) const {
        std::vector<Real> out(len);
        if (len == 0) return out;
        if (prec == SOV_PREC_F64) {
            hip_ok(sov_hip_memcpy_d2h(ctx, out.data(), src.p, len * 8), "download");
        } else {
            std::vector<float> f(len);
            hip_ok(sov_hip_memcpy_d2h(ctx, f.data(), src.p, len * 4), "download");
            parallel_for(len, [&](std::size_t b, std::size_t e) {
                for (std::size_t i = b; i < e; ++i) out[i] = static_cast<Real>(f[i]);
            });
        }
        if (scale != 1.0)
            parallel_for(len, [&](std::size_t b, std::size_t e) {
                for (std::size_t i = b; i < e; ++i) out[i] *= scale;
            });
        return out;
    }
};

/// The fused engine's checkpoint hand-off (PdhgOptions::checkpoint, reference
/// pdhg.hpp:121-133, called at every check with the unscaled candidate).
/// post() snapshots the candidate on the device (ordered before every later
/// kernel) and queues its copy to pinned host memory on a copier stream, so
/// the next interval starts at once; drain(), called right after that
/// interval has been launched, waits for the copy, converts and unscales on
/// the host and calls the callback while the device computes.  Callbacks
/// come in check order, one interval late, with the values the synchronous
/// hand-off would give.
class CheckpointPipe {
public:
    CheckpointPipe(sov_hip_ctx* ctx, std::size_t m, std::size_t n) : ctx_(ctx), m_(m), n_(n) {
        hip_ok(sov_hip_copier_create(ctx, &copier_), "copier");
        const std::size_t bytes = std::max<std::size_t>(m + n, 1) * 8;
        staging_ = DevMem(ctx, bytes);
        hip_ok(sov_hip_host_alloc(ctx, bytes, &host_), "pinned checkpoint buffer");
    }
    ~CheckpointPipe() {
        sov_hip_copier_destroy(copier_);
        (void)sov_hip_host_free(ctx_, host_);
    }
    CheckpointPipe(const CheckpointPipe&) = delete;
    CheckpointPipe& operator=(const CheckpointPipe&) = delete;

    /// The candidate of this check: the average (x_sum, y_sum) * inv when
    /// avg, else (x, y), of the state S.
    void post(const DeviceState& S, bool avg, Real inv, long long iteration, Real residual) {
        const std::size_t es = S.es;
        char* dst = staging_.as<char>();
        if (n_) hip_ok(sov_hip_memcpy_d2d(ctx_, dst, (avg ? S.xsum : S.x).p, n_ * es), "checkpoint snapshot");
        if (m_) hip_ok(sov_hip_memcpy_d2d(ctx_, dst + n_ * es, (avg ? S.ysum : S.y).p, m_ * es), "checkpoint snapshot");
        hip_ok(sov_hip_copier_d2h(ctx_, copier_, host_, dst, (n_ + m_) * es), "checkpoint copy");
        pending_ = true;
        prec_ = S.prec;
        scale_ = avg ? inv : Real{1};
        iteration_ = iteration;
        residual_ = residual;
    }

    /// Calls f(iteration, x, y, residual) for the pending checkpoint, with x
    /// and y in device order and scaled space, as DeviceState::download
    /// would give them.
    template <class F>
    void drain(F&& f) {
        if (!pending_) return;
        pending_ = false;
        hip_ok(sov_hip_copier_wait(copier_), "checkpoint copy");
        std::vector<Real> x(n_), y(m_);
        auto widen = [&](std::vector<Real>& out, std::size_t offset) {
            parallel_for(out.size(), [&](std::size_t b, std::size_t e) {
                if (prec_ == SOV_PREC_F64) {
                    const double* src = static_cast<const double*>(host_) + offset;
                    for (std::size_t i = b; i < e; ++i) out[i] = src[i];
                } else {
                    const float* src = static_cast<const float*>(host_) + offset;
                    for (std::size_t i = b; i < e; ++i) out[i] = static_cast<Real>(src[i]);
                }
                if (scale_ != 1.0)
                    for (std::size_t i = b; i < e; ++i) out[i] *= scale_;
            });
        };
        widen(x, 0);
        widen(y, n_);
        f(iteration_, x, y, residual_);
    }

private:
    sov_hip_ctx* ctx_;
    std::size_t m_, n_;
    sov_hip_copier* copier_ = nullptr;
    DevMem staging_;
    void* host_ = nullptr;
    bool pending_ = false;
    int prec_ = SOV_PREC_F64;
    Real scale_ = 1.0;
    long long iteration_ = 0;
    Real residual_ = 0.0;
};

}  // namespace

PdhgResult Pdhg::solve_fused(const Problem& problem, Backend& backend, const CancellationToken* cancel) const {
    PdhgResult result;
    result.engine = "fused";
    const auto t0 = Clock::now();
    if (!prevalidate(problem, result)) return result;
    const PdhgOptions& o = options_;
    auto* ctx = static_cast<sov_hip_ctx*>(backend.native_context());
    try {
        const CsrMatrix& A0 = problem.matrix();
        const auto m = static_cast<std::size_t>(problem.num_rows()), n = static_cast<std::size_t>(problem.num_cols());
        const auto nnz = static_cast<std::size_t>(A0.nnz());
        const Real sign = problem.sense() == ObjSense::Maximize ? -1.0 : 1.0;
        // verbose: the setup's phases, each timed to the device's completion
        auto t_mark = Clock::now();
        auto mark = [&](const char* what) {
            if (!o.verbose) return;
            hip_ok(sov_hip_sync(ctx), "sync");
            std::fprintf(stderr, "  [pdhg] setup: %s %.1f ms\n", what, 1e3 * seconds_since(t_mark));
            t_mark = Clock::now();
        };

        // -- the matrix crosses the bus once (3 copies), then everything is device work
        DevMem d_rp(ctx, (m + 1) * sizeof(int32_t)), d_ci(ctx, std::max<std::size_t>(nnz, 1) * sizeof(int32_t)),
            d_v(ctx, std::max<std::size_t>(nnz, 1) * sizeof(double));
        hip_ok(sov_hip_memcpy_h2d(ctx, d_rp.p, A0.row_ptr().data(), (m + 1) * sizeof(int32_t)), "upload A");
        if (nnz > 0) {
            hip_ok(sov_hip_memcpy_h2d(ctx, d_ci.p, A0.col_idx().data(), nnz * sizeof(int32_t)), "upload A");
            hip_ok(sov_hip_memcpy_h2d(ctx, d_v.p, A0.values().data(), nnz * sizeof(double)), "upload A");
        }
        CsrHandle csr;
        hip_ok(sov_hip_csr_create(ctx, static_cast<int32_t>(m), static_cast<int32_t>(n), static_cast<int32_t>(nnz),
                                  d_rp.as<int32_t>(), d_ci.as<int32_t>(), d_v.as<double>(), SOV_CSR_OWN_VALUES, &csr.h),
               "csr_create");
        d_v.reset();   // the handle owns a copy of the values
        DeviceOrder ord;
        ord.row.resize(m);
        ord.col.resize(n);
        hip_ok(sov_hip_csr_order(ctx, csr.h, 0, ord.row.data()), "order");
        hip_ok(sov_hip_csr_order(ctx, csr.h, 1, ord.col.data()), "order");
        mark("upload, transpose and device order");

        // -- kernel (d): Ruiz + Pock-Chambolle on the device (reference scaling.cpp:19-104)
        DevMem d_rs(ctx, std::max<std::size_t>(m, 1) * 8), d_cs(ctx, std::max<std::size_t>(n, 1) * 8);
        {
            const std::vector<Real> ones_m(m, 1.0), ones_n(n, 1.0);
            if (m) hip_ok(sov_hip_memcpy_h2d(ctx, d_rs.p, ones_m.data(), m * 8), "upload");
            if (n) hip_ok(sov_hip_memcpy_h2d(ctx, d_cs.p, ones_n.data(), n * 8), "upload");
            const ScalingOptions so;
            if (so.apply_ruiz)
                for (int sweep = 0; sweep < so.ruiz_iterations; ++sweep) {
                    double worst = 0.0;
                    hip_ok(sov_hip_scale_ruiz_sweep(ctx, csr.h, d_rs.as<double>(), d_cs.as<double>(), &worst), "ruiz");
                    if (worst < so.ruiz_tolerance) break;
                }
            if (so.apply_pock_chambolle)
                hip_ok(sov_hip_scale_pock_chambolle(ctx, csr.h, so.pock_chambolle_alpha, d_rs.as<double>(),
                                                    d_cs.as<double>()),
                       "pock_chambolle");
        }
        std::vector<Real> rs(m), cs(n);
        if (m) hip_ok(sov_hip_memcpy_d2h(ctx, rs.data(), d_rs.p, m * 8), "download scaling");
        if (n) hip_ok(sov_hip_memcpy_d2h(ctx, cs.data(), d_cs.p, n * 8), "download scaling");
        const Scaling scaling = Scaling::from_factors(rs, cs);
        hip_ok(sov_hip_csr_scale(ctx, csr.h, d_rs.as<double>(), d_cs.as<double>()), "csr_scale");
        mark("scaling");

        // scaled vectors (reference scaling.cpp:172-192, pdhg.cpp:250-252)
        std::vector<Real> c_s(n), clo_s(n), chi_s(n), rlo_s(m), rhi_s(m);
        parallel_for(n, [&](std::size_t b, std::size_t e) {
            for (std::size_t j = b; j < e; ++j) {
                c_s[j] = (problem.objective()[j] * cs[j]) * sign;
                const Real lo = problem.col_lower()[j], hi = problem.col_upper()[j];
                clo_s[j] = is_finite_bound(lo) ? lo / cs[j] : lo;
                chi_s[j] = is_finite_bound(hi) ? hi / cs[j] : hi;
            }
        });
        for (std::size_t i = 0; i < m; ++i) {
            const Real lo = problem.row_lower()[i], hi = problem.row_upper()[i];
            rlo_s[i] = is_finite_bound(lo) ? lo * rs[i] : lo;
            rhi_s[i] = is_finite_bound(hi) ? hi * rs[i] : hi;
        }
        // everything below lives in device order
        c_s = DeviceOrder::to_dev(c_s, ord.col);
        clo_s = DeviceOrder::to_dev(clo_s, ord.col);
        chi_s = DeviceOrder::to_dev(chi_s, ord.col);
        rlo_s = DeviceOrder::to_dev(rlo_s, ord.row);
        rhi_s = DeviceOrder::to_dev(rhi_s, ord.row);
        // original-space arrays and factors for the device checks (FP64)
        DevMem d_olo_r(ctx, std::max<std::size_t>(m, 1) * 8), d_ohi_r(ctx, std::max<std::size_t>(m, 1) * 8),
            d_olo_c(ctx, std::max<std::size_t>(n, 1) * 8), d_ohi_c(ctx, std::max<std::size_t>(n, 1) * 8),
            d_oc(ctx, std::max<std::size_t>(n, 1) * 8), d_act(ctx, std::max<std::size_t>(m, 1) * 8),
            d_rs_dev(ctx, std::max<std::size_t>(m, 1) * 8), d_cs_dev(ctx, std::max<std::size_t>(n, 1) * 8);
        auto upload_dev = [&](const DevMem& d, std::span<const Real> v, const std::vector<int32_t>& perm) {
            if (v.empty()) return;
            const std::vector<Real> pv = DeviceOrder::to_dev(v, perm);
            hip_ok(sov_hip_memcpy_h2d(ctx, d.p, pv.data(), pv.size() * 8), "upload");
        };
        upload_dev(d_olo_r, problem.row_lower(), ord.row);
        upload_dev(d_ohi_r, problem.row_upper(), ord.row);
        upload_dev(d_rs_dev, rs, ord.row);
        upload_dev(d_olo_c, problem.col_lower(), ord.col);
        upload_dev(d_ohi_c, problem.col_upper(), ord.col);
        upload_dev(d_oc, problem.objective(), ord.col);
        upload_dev(d_cs_dev, cs, ord.col);
        mark("scaled vectors and check arrays");

        // -- step size: device power iteration in FP64 (reference pdhg.cpp:48-79)
        Real est = 0.0;
        {
            DevMem u(ctx, std::max<std::size_t>(n, 1) * 8), v(ctx, std::max<std::size_t>(m, 1) * 8),
                w(ctx, std::max<std::size_t>(m, 1) * 8);
            const std::vector<Real> ones_m(m, 1.0), ones_n(n, 1.0);
            if (n) hip_ok(sov_hip_memcpy_h2d(ctx, u.p, ones_n.data(), n * 8), "upload");
            if (m) hip_ok(sov_hip_memcpy_h2d(ctx, v.p, ones_m.data(), m * 8), "upload");
            const auto mm = static_cast<int64_t>(m), nn = static_cast<int64_t>(n);
            for (int it = 0; it < o.power_iterations; ++it) {
                hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, -1.0, u.p, 0.0, w.p), "power spmv");
                hip_ok(sov_hip_axpy_f64(ctx, mm, 1.0, v.as<double>(), w.as<double>()), "power axpy");
                double wn = 0.0;
                hip_ok(sov_hip_nrm2_f64(ctx, mm, w.as<double>(), &wn), "power nrm2");
                if (wn < 1e-300) break;
                hip_ok(sov_hip_scal_f64(ctx, mm, 1.0 / wn, w.as<double>()), "power scal");
                hip_ok(sov_hip_spmv_device(ctx, csr.h, 1, SOV_PREC_F64, -1.0, w.p, 0.0, u.p), "power spmv_t");
                hip_ok(sov_hip_memcpy_d2d(ctx, v.p, w.p, m * 8), "power copy");
                double un = 0.0, vn = 0.0;
                hip_ok(sov_hip_nrm2_f64(ctx, nn, u.as<double>(), &un), "power nrm2");
                hip_ok(sov_hip_nrm2_f64(ctx, mm, v.as<double>(), &vn), "power nrm2");
                est = std::sqrt(un * un + vn * vn);
                if (est < 1e-300) break;
                hip_ok(sov_hip_scal_f64(ctx, nn, 1.0 / est, u.as<double>()), "power scal");
                hip_ok(sov_hip_scal_f64(ctx, mm, 1.0 / est, v.as<double>()), "power scal");
            }
        }
        result.operator_norm_estimate = est;
        mark("power iteration");
        const Real eta = est > 1e-9 ? o.step_size_safety / est : o.step_size_safety;
        Real omega = o.use_primal_weight ? o.initial_primal_weight : 1.0;

        // -- iterate state
        int prec = o.precision == Precision::F32 ? SOV_PREC_F32 : SOV_PREC_F64;
        if (prec == SOV_PREC_F32) hip_ok(sov_hip_csr_make_f32(ctx, csr.h), "make_f32");
        DeviceState S;
        auto init_state = [&](int p) {
            S.allocate(ctx, p, m, n);
            S.upload(ctx, S.c, c_s);
            S.upload(ctx, S.clo, clo_s);
            S.upload(ctx, S.chi, chi_s);
            S.upload(ctx, S.rlo, rlo_s);
            S.upload(ctx, S.rhi, rhi_s);
        };
        init_state(prec);
        // x = Proj_X(0), s = Proj_S(0), y = 0 or the warm start (reference pdhg.cpp:268-270)
        {
            const std::vector<Real> zm(m, 0.0), zn(n, 0.0);
            DevMem zx(ctx, std::max<std::size_t>(n, 1) * 8), zs(ctx, std::max<std::size_t>(m, 1) * 8);
            if (n) hip_ok(sov_hip_memcpy_h2d(ctx, zx.p, zn.data(), n * 8), "upload");
            if (m) hip_ok(sov_hip_memcpy_h2d(ctx, zs.p, zm.data(), m * 8), "upload");
            DevMem wy;
            const double* ysrc = nullptr;
            if (!o.warm_start_dual.empty()) {
                if (o.warm_start_dual.size() != m) throw Error(Status::DimensionMismatch, "warm_start_dual length");
                std::vector<Real> ys(m);
                for (std::size_t i = 0; i < m; ++i) ys[i] = sign * o.warm_start_dual[i] / rs[i];
                ys = DeviceOrder::to_dev(ys, ord.row);
                wy = DevMem(ctx, std::max<std::size_t>(m, 1) * 8);
                if (m) hip_ok(sov_hip_memcpy_h2d(ctx, wy.p, ys.data(), m * 8), "upload warm start");
                ysrc = wy.as<double>();
            }
            if (!o.warm_start_primal.empty()) {
                // x = Proj_X(x0 / c^s), s = Proj_S(A_s x) (additive)
                if (o.warm_start_primal.size() != n) throw Error(Status::DimensionMismatch, "warm_start_primal length");
                std::vector<Real> xs(n);
                for (std::size_t j = 0; j < n; ++j) {
                    const Real lo = problem.col_lower()[j], hi = problem.col_upper()[j];
                    Real v = o.warm_start_primal[j] / cs[j];
                    if (is_finite_bound(lo)) v = std::max(v, lo / cs[j]);
                    if (is_finite_bound(hi)) v = std::min(v, hi / cs[j]);
                    xs[j] = v;
                }
                xs = DeviceOrder::to_dev(xs, ord.col);
                if (n) hip_ok(sov_hip_memcpy_h2d(ctx, zx.p, xs.data(), n * 8), "upload warm start");
                hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, 1.0, zx.p, 0.0, zs.p), "warm start activity");
            }
            hip_ok(sov_hip_pdhg_set_point(ctx, &S.st, zx.as<double>(), ysrc, zs.as<double>()), "set_point");
            hip_ok(sov_hip_sync(ctx), "sync");
        }
        // the autotuner (before the loop, or deferred to the first check
        // after autotune_after_seconds of iterations); forced tiles are set
        // again after it, since it also chooses tiles
        bool tuned = !(o.autotune && nnz > 0);
        auto set_forced_tiles = [&]() {
            if (o.tiles >= 0 && nnz > 0)
                for (int orient : {0, 1})
                    if ((orient == 0 ? m : n) > 0)
                        hip_ok(sov_hip_csr_set_tiles(ctx, csr.h, orient, o.tiles), "tiles");
        };
        if (!tuned && o.autotune_after_seconds <= 0.0) {
            hip_ok(sov_hip_csr_tune(ctx, csr.h, prec), "tune");
            tuned = true;
        }
        set_forced_tiles();
        mark("iterate state, tuning and layouts");

        // primal-weight bookkeeping (FP64 copies of the last restart point)
        DevMem d_xlast, d_ylast;
        if (o.use_primal_weight) {
            d_xlast = DevMem(ctx, std::max<std::size_t>(n, 1) * 8);
            d_ylast = DevMem(ctx, std::max<std::size_t>(m, 1) * 8);
            const std::vector<Real> xl = S.download(ctx, S.x, n), yl = S.download(ctx, S.y, m);
            if (n) hip_ok(sov_hip_memcpy_h2d(ctx, d_xlast.p, xl.data(), n * 8), "upload");
            if (m) hip_ok(sov_hip_memcpy_h2d(ctx, d_ylast.p, yl.data(), m * 8), "upload");
        }

        sov_hip_check_args ca{};
        ca.tol_scaled = o.tolerance;
        ca.tol_orig = std::max(o.tolerance, Real{1e-6});
        ca.obj_sign = sign;
        ca.obj_const = problem.objective_constant();
        ca.row_scale = d_rs_dev.as<double>();
        ca.col_scale = d_cs_dev.as<double>();
        ca.orig_row_lo = d_olo_r.as<double>();
        ca.orig_row_hi = d_ohi_r.as<double>();
        ca.orig_col_lo = d_olo_c.as<double>();
        ca.orig_col_hi = d_ohi_c.as<double>();
        ca.orig_c = d_oc.as<double>();
        ca.act_avg = d_act.as<double>();

        const long long period = std::max<long long>(o.restart_check_period, 1);
        GraphHandle g_first, g_cont;
        Real graph_tau = -1.0, graph_sigma = -1.0;
        int graph_prec = -1;
        auto run_steps = [&](long long count, bool first_is_fresh, Real tau, Real sigma) {
            for (long long k = 0; k < count; ++k) {
                const int first = (first_is_fresh && k == 0) ? 1 : 0;
                hip_ok(sov_hip_pdhg_primal_step(ctx, csr.h, &S.st, tau, first), "primal_step");
                hip_ok(sov_hip_pdhg_dual_step(ctx, csr.h, &S.st, tau, sigma, first), "dual_step");
            }
        };
        auto run_interval = [&](long long count, bool fresh, Real tau, Real sigma) {
            if (!o.use_graphs || count != period) {
                run_steps(count, fresh, tau, sigma);
                return;
            }
            if (graph_tau != tau || graph_sigma != sigma || graph_prec != S.prec) {
                g_first.reset();
                g_cont.reset();
                graph_tau = tau;
                graph_sigma = sigma;
                graph_prec = S.prec;
            }
            GraphHandle& g = fresh ? g_first : g_cont;
            if (g.g == nullptr) {
                hip_ok(sov_hip_graph_begin(ctx), "graph_begin");
                run_steps(count, fresh, tau, sigma);
                hip_ok(sov_hip_graph_end(ctx, &g.g), "graph_end");
            }
            hip_ok(sov_hip_graph_launch(ctx, g.g), "graph_launch");
        };

        long long since = 0;
        Real res_at_restart = kInfinity, prev_cand = kInfinity;
        bool converged = false;
        std::vector<Real> best_x, best_y;   // scaled, internal sense, FP64, device order
        Drift drift;
        const auto t_loop = Clock::now();
        result.setup_seconds = seconds_since(t0);
        long long polish_start = -1;
        Real best_rel = kInfinity;
        int stall = 0;
        bool timing = false;
        auto t_timed = Clock::now();
        // the checkpoint callback: unscale a device-order, scaled candidate
        auto deliver = [&](long long it, const std::vector<Real>& cx, const std::vector<Real>& cy, Real res) {
            auto [ux, uy] = unscale_point(sign, scaling, DeviceOrder::to_nat(cx, ord.col),
                                          DeviceOrder::to_nat(cy, ord.row));
            o.checkpoint(it, ux, uy, res);
        };
        std::unique_ptr<CheckpointPipe> pipe;
        if (o.checkpoint && o.checkpoint_async) pipe = std::make_unique<CheckpointPipe>(ctx, m, n);

        // iterative refinement: once on, S holds FP32 corrections (dx, ds, dy)
        // of the FP64 anchor (x0, A x0, y0) at scale delta, and R holds the
        // point they stand for (FP64, the original scaled data) for checks,
        // restarts, candidates and checkpoints
        const bool refine = o.fp32_refinement && prec == SOV_PREC_F32;
        bool refining = false;
        long long refine_from = -1;
        DeviceState R;
        DevMem d_x0, d_y0, d_s0, d_aty0, d_sgrad;
        Real delta = 1.0;
        auto combine = [&]() {
            hip_ok(sov_hip_pdhg_combine(ctx, &S.st, &R.st, d_x0.as<double>(), d_y0.as<double>(), since, 1.0 / delta),
                   "combine");
        };
        // R.x, R.y become the anchor; the correction restarts at zero
        auto anchor = [&](Real rel) {
            if (n) hip_ok(sov_hip_memcpy_d2d(ctx, d_x0.p, R.x.p, n * 8), "anchor");
            if (m) hip_ok(sov_hip_memcpy_d2d(ctx, d_y0.p, R.y.p, m * 8), "anchor");
            hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, 1.0, d_x0.p, 0.0, d_s0.p), "anchor activity");
            hip_ok(sov_hip_spmv_device(ctx, csr.h, 1, SOV_PREC_F64, 1.0, d_y0.p, 0.0, d_aty0.p), "anchor A^T y");
            delta = std::clamp(1.0 / std::max(rel, Real{1e-300}), Real{1}, std::max(o.refinement_max_scale, Real{1}));
            hip_ok(sov_hip_pdhg_anchor(ctx, &R.st, &S.st, d_x0.as<double>(), d_y0.as<double>(), d_s0.as<double>(),
                                       d_aty0.as<double>(), delta),
                   "anchor");
            ++result.refinement_anchors;
        };
        // the switch: R takes S's point in FP64 (a zero anchor at scale 1)
        auto begin_refinement = [&]() {
            R.allocate(ctx, SOV_PREC_F64, m, n);
            R.upload(ctx, R.c, c_s);
            R.upload(ctx, R.clo, clo_s);
            R.upload(ctx, R.chi, chi_s);
            R.upload(ctx, R.rlo, rlo_s);
            R.upload(ctx, R.rhi, rhi_s);
            d_x0 = DevMem(ctx, std::max<std::size_t>(n, 1) * 8);
            d_y0 = DevMem(ctx, std::max<std::size_t>(m, 1) * 8);
            d_s0 = DevMem(ctx, std::max<std::size_t>(m, 1) * 8);
            d_aty0 = DevMem(ctx, std::max<std::size_t>(n, 1) * 8);
            d_sgrad = DevMem(ctx, std::max<std::size_t>(m, 1) * S.es);
            hip_ok(sov_hip_fill_zero(ctx, d_x0.p, std::max<std::size_t>(n, 1) * 8), "fill");
            hip_ok(sov_hip_fill_zero(ctx, d_y0.p, std::max<std::size_t>(m, 1) * 8), "fill");
            delta = 1.0;
            combine();
            S.st.s_grad = d_sgrad.p;   // a new kernel argument: captured intervals are captured again
            g_first.reset();
            g_cont.reset();
            graph_prec = -1;
            refining = true;
            refine_from = result.iterations;
            if (o.verbose)
                std::fprintf(stderr, "  [pdhg] fp32 refinement from iter %lld\n", result.iterations);
        };

        result.iterations = 0;
        while (result.iterations < o.max_iterations) {
            if (cancel != nullptr && cancel->is_cancelled()) { result.status = PdhgStatus::Cancelled; break; }
            if (o.time_limit_seconds > 0.0 && seconds_since(t0) > o.time_limit_seconds) {
                result.status = PdhgStatus::TimeLimit;
                break;
            }
            const long long next_check = (result.iterations / period + 1) * period;
            const long long count = std::min(next_check, static_cast<long long>(o.max_iterations)) - result.iterations;
            const Real tau = eta / omega, sigma = eta * omega;
            const long long mark = o.timing_mark_iteration;
            if (mark >= 0 && mark == result.iterations && !timing) {
                hip_ok(sov_hip_sync(ctx), "sync");
                t_timed = Clock::now();
                timing = true;
            }
            if (mark > result.iterations && mark < result.iterations + count) {
                // the mark falls inside this interval: run it as two plain
                // launch sequences with the clock started between them
                const long long head = mark - result.iterations;
                run_steps(head, since == 0, tau, sigma);
                hip_ok(sov_hip_sync(ctx), "sync");
                t_timed = Clock::now();
                timing = true;
                run_steps(count - head, false, tau, sigma);
            } else {
                run_interval(count, since == 0, tau, sigma);
            }
            if (pipe) pipe->drain(deliver);   // the previous check's callback, while this interval runs
            since += count;
            result.iterations += count;
            if (result.iterations != next_check) break;   // iteration limit inside the interval
            if (!tuned && seconds_since(t_loop) >= o.autotune_after_seconds) {
                // deferred autotuning: new schedules and layouts, so the
                // captured intervals are captured again
                hip_ok(sov_hip_csr_tune(ctx, csr.h, S.prec), "tune");
                set_forced_tiles();
                tuned = true;
                g_first.reset();
                g_cont.reset();
                graph_prec = -1;
            }

            // refining: the check reads the corrections and adds the anchor's
            // terms; R (the combined point) is formed only where it is read
            ca.st = &S.st;
            ca.base = refining ? &R.st : nullptr;
            if (refining) {
                ca.base_x = d_x0.as<double>();
                ca.base_y = d_y0.as<double>();
                ca.base_act = d_s0.as<double>();
                ca.base_aty = d_aty0.as<double>();
                ca.inv_delta = 1.0 / delta;
            }
            bool r_current = false;
            auto point = [&]() -> DeviceState& {   // the point checked, for reading
                if (refining && !r_current) {
                    combine();
                    r_current = true;
                }
                return refining ? R : S;
            };
            ca.since_restart = since;
            sov_hip_check_out co{};
            hip_ok(sov_hip_pdhg_check(ctx, csr.h, &ca, &co), "check");
            const Real cur = co.kkt_cur, avg = co.kkt_avg;
            if (!std::isfinite(cur) || !std::isfinite(avg)) {
                result.status = PdhgStatus::NumericalFailure;
                result.message = "non-finite KKT residual at iteration " + std::to_string(result.iterations - 1);
                result.seconds = seconds_since(t0);
                return result;
            }
            const bool adaptive = o.restart_scheme == PdhgRestartScheme::Adaptive;
            const Real rel_c = std::max({co.rel_cur[0], co.rel_cur[1], co.rel_cur[2]});
            const Real rel_a = std::max({co.rel_avg[0], co.rel_avg[1], co.rel_avg[2]});
            const bool avg_better = adaptive ? rel_a <= rel_c : avg <= cur;
            const Real ores = avg_better ? co.orig_avg : co.orig_cur;
            const Real rel_worst = avg_better ? rel_a : rel_c;
            if (!std::isfinite(ores)) {
                result.status = PdhgStatus::NumericalFailure;
                result.message = "non-finite KKT residual (original units) at iteration " +
                                 std::to_string(result.iterations - 1);
                result.seconds = seconds_since(t0);
                return result;
            }
            const Real inv = 1.0 / static_cast<Real>(since);
            auto candidate = [&]() {
                DeviceState& V = point();
                std::pair<std::vector<Real>, std::vector<Real>> p;
                p.first = avg_better ? V.download(ctx, V.xsum, n, inv) : V.download(ctx, V.x, n);
                p.second = avg_better ? V.download(ctx, V.ysum, m, inv) : V.download(ctx, V.y, m);
                return p;
            };
            if (pipe) {
                pipe->post(point(), avg_better, inv, result.iterations - 1, ores);
            } else if (o.checkpoint) {
                auto [cx, cy] = candidate();
                deliver(result.iterations - 1, cx, cy, ores);
            }
            const bool stop = o.termination == PdhgTermination::Reference ? (ores < o.tolerance)
                                                                          : (rel_worst <= o.tolerance);
            if (stop) {
                auto [cx, cy] = candidate();
                best_x = std::move(cx);
                best_y = std::move(cy);
                converged = true;
                break;
            }
            if (o.detect_infeasibility) {
                DeviceState& V = point();
                std::vector<Real> xd = V.download(ctx, V.x, n), yd = V.download(ctx, V.y, m);
                auto [dxo, dyo] = unscale_point(sign, scaling, DeviceOrder::to_nat(xd, ord.col),
                                                DeviceOrder::to_nat(yd, ord.row));
                if (drift.test(problem, std::move(dxo), std::move(dyo), result)) {
                    best_x = std::move(xd);
                    best_y = std::move(yd);
                    break;
                }
            }
            // FP32 -> FP64 switch (additive): near tolerance or stalled.  With
            // refinement there is none: the corrections carry FP64 precision.
            const Real gauge = o.termination == PdhgTermination::Reference ? ores : rel_worst;
            if (S.prec == SOV_PREC_F32 && !refining && (o.fp64_polish || refine)) {
                if (rel_worst < best_rel * 0.99) {
                    best_rel = rel_worst;
                    stall = 0;
                } else {
                    ++stall;
                }
                if (!refine && o.fp64_polish &&
                    (gauge <= o.fp32_switch_factor * o.tolerance || stall >= o.fp32_stall_checks)) {
                    auto [cx, cy] = candidate();
                    init_state(SOV_PREC_F64);
                    DevMem hx(ctx, std::max<std::size_t>(n, 1) * 8), hy(ctx, std::max<std::size_t>(m, 1) * 8);
                    if (n) hip_ok(sov_hip_memcpy_h2d(ctx, hx.p, cx.data(), n * 8), "upload");
                    if (m) hip_ok(sov_hip_memcpy_h2d(ctx, hy.p, cy.data(), m * 8), "upload");
                    hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, 1.0, hx.p, 0.0, d_act.p),
                           "polish activity");
                    hip_ok(sov_hip_pdhg_set_point(ctx, &S.st, hx.as<double>(), hy.as<double>(), d_act.as<double>()),
                           "polish set_point");
                    hip_ok(sov_hip_sync(ctx), "sync");
                    since = 0;
                    res_at_restart = kInfinity;
                    prev_cand = kInfinity;
                    polish_start = result.iterations;
                    drift.reset();
                    if (o.verbose)
                        std::fprintf(stderr, "  [pdhg] fp64 polish from iter %lld (rel kkt %.3g)\n", result.iterations,
                                     static_cast<double>(rel_worst));
                    continue;
                }
            }
            bool do_restart = false;
            if (!adaptive) {
                if (!std::isfinite(res_at_restart)) res_at_restart = cur;
                do_restart = avg <= o.restart_sufficient_decrease * res_at_restart;
            } else {
                if (!std::isfinite(res_at_restart)) res_at_restart = rel_c;
                do_restart = rel_worst <= o.restart_beta_sufficient * res_at_restart ||
                             (rel_worst <= o.restart_beta_necessary * res_at_restart && rel_worst > prev_cand) ||
                             static_cast<Real>(since) >= o.restart_beta_artificial * static_cast<Real>(result.iterations);
                prev_cand = rel_worst;
            }
            // refinement begins at a restart near tolerance, or at once when
            // the FP32 iterates have stalled
            bool enter = false;
            if (refine && !refining && S.prec == SOV_PREC_F32) {
                if (stall >= o.fp32_stall_checks) {
                    do_restart = true;
                    enter = true;
                } else if (do_restart && gauge <= std::max(o.fp32_switch_factor * o.tolerance, o.refinement_start)) {
                    enter = true;
                }
            }
            if (do_restart) {
                drift.reset();
                if (enter) begin_refinement();
                DeviceState& Q = point();   // the state a restart moves
                double mv[2] = {0.0, 0.0};
                double* xl = o.use_primal_weight ? d_xlast.as<double>() : nullptr;
                double* yl = o.use_primal_weight ? d_ylast.as<double>() : nullptr;
                if (!adaptive || avg_better) {
                    hip_ok(sov_hip_pdhg_restart(ctx, &Q.st, since, d_act.as<double>(), xl, yl, mv), "restart");
                    if (refining) anchor(adaptive ? rel_a : std::max({co.rel_avg[0], co.rel_avg[1], co.rel_avg[2]}));
                } else {
                    if (o.use_primal_weight) hip_ok(sov_hip_pdhg_movement(ctx, &Q.st, xl, yl, mv), "movement");
                    if (enter) anchor(rel_c);
                }
                if (o.use_primal_weight) omega = next_primal_weight(omega, mv[0], mv[1], o.primal_weight_smoothing);
                res_at_restart = adaptive ? rel_worst : avg;
                prev_cand = kInfinity;
                since = 0;
                ++result.restarts;
                if (o.verbose)
                    std::fprintf(stderr, "  [pdhg] restart at iter %lld, residual %.3g, omega %.3g\n",
                                 result.iterations, static_cast<double>(adaptive ? rel_worst : avg),
                                 static_cast<double>(omega));
            }
        }
        if (pipe) pipe->drain(deliver);   // the last check's callback
        hip_ok(sov_hip_sync(ctx), "sync");
        result.loop_seconds = seconds_since(t_loop);
        if (timing) {
            result.timed_seconds = seconds_since(t_timed);
            result.timed_iterations = result.iterations - o.timing_mark_iteration;
        }
        if (refining) result.refinement_iterations = result.iterations - refine_from;
        if (best_x.empty()) {
            if (refining && since > 0) combine();
            DeviceState& V = refining ? R : S;
            best_x = V.download(ctx, V.x, n);
            best_y = V.download(ctx, V.y, m);
        }
        if (converged) result.status = PdhgStatus::Optimal;
        else if (result.status == PdhgStatus::NotSolved) result.status = PdhgStatus::IterationLimit;
        if (polish_start >= 0) result.fp64_polish_iterations = result.iterations - polish_start;
        result.primal_weight = omega;
        auto [fx, fy] = unscale_point(sign, scaling, DeviceOrder::to_nat(best_x, ord.col),
                                      DeviceOrder::to_nat(best_y, ord.row));
        finish_result(problem, o, std::move(fx), std::move(fy), result);
    } catch (const Error& e) {
        result.status = PdhgStatus::NumericalFailure;
        result.message = std::string("device failure: ") + e.what();
    }
    result.seconds = seconds_since(t0);
    return result;
}

// ===========================================================================
// batched engine and the live batch session (additive)
// ===========================================================================

namespace {

struct BatchHandle {
    sov_hip_batch* b = nullptr;
    ~BatchHandle() { sov_hip_batch_destroy(b); }
};

/// The same validation prevalidate() applies before a solve, for a session
/// update: throws with the reason instead of reporting a status.
void require_valid(const Problem& p, const char* what) {
    PdhgResult r;
    if (!prevalidate(p, r)) throw Error(Status::InvalidArgument, std::string(what) + ": " + r.message);
}

}  // namespace

struct PdhgBatchSession::Impl {
    PdhgOptions o;
    sov_hip_ctx* ctx = nullptr;
    std::vector<Problem> probs;                 // the session's copies, in input order
    std::vector<PdhgResult> rejected;           // prevalidate's verdict on instances not stacked
    std::vector<std::size_t> idx;               // stacked instance -> input index
    std::vector<std::ptrdiff_t> slot;           // input index -> stacked instance (-1: not stacked)
    std::size_t B = 0, m = 0, n = 0, nnz = 0;
    Real sign = 1.0;
    std::vector<int32_t> row_off{0}, col_off{0};
    DevMem a_rp, a_ci;                          // A's structure, borrowed by the handle (declared first: outlives it)
    CsrHandle csr;
    BatchHandle bh;
    DeviceOrder ord;
    std::vector<Real> rs, cs, oconst, est;
    std::vector<double> tau, sigma, weight;
    DevMem d_olo_r, d_ohi_r, d_olo_c, d_ohi_c, d_oc, d_act, d_act32, d_rs_dev, d_cs_dev;
    DeviceState S32, S64;
    bool f32 = false;
    std::vector<Real> x_start, y_start;         // next solve's start: scaled space, natural order, FP64
    std::vector<char> cold;                     // per stacked instance: start as the reference does (s = Proj_S(0))
    std::vector<Real> clo_s, chi_s;             // scaled column bounds, natural order
    std::vector<std::vector<Real>> last_x, last_y;   // last original-space solutions, input order
    bool data_dirty = true, start_dirty = true;
    double setup_seconds = 0.0;

    Impl(PdhgOptions options, const std::vector<const Problem*>& problems, Backend& backend,
         const std::vector<std::vector<Real>>* warm_dual, const std::vector<Real>* primal_weight);
    void stacked_vectors(std::vector<Real>& c, std::vector<Real>& clo, std::vector<Real>& chi,
                         std::vector<Real>& rlo, std::vector<Real>& rhi) const;
    void upload_data();
    void place_start();
    std::size_t stacked(std::size_t b, const char* what) const;
    void replace(std::size_t b, Problem p, const char* what);
    PdhgBatchResult run();
};

PdhgBatchSession::Impl::Impl(PdhgOptions options, const std::vector<const Problem*>& problems, Backend& backend,
                             const std::vector<std::vector<Real>>* warm_dual, const std::vector<Real>* primal_weight)
    : o(std::move(options)) {
    const auto t0 = Clock::now();
    ctx = static_cast<sov_hip_ctx*>(backend.native_context());
    if (backend.kind() != BackendKind::Hip || ctx == nullptr)
        throw Error(Status::Unsupported, "the batched engine needs a Hip backend");
    probs.reserve(problems.size());
    rejected.resize(problems.size());
    slot.assign(problems.size(), -1);
    last_x.resize(problems.size());
    last_y.resize(problems.size());
    for (std::size_t b = 0; b < problems.size(); ++b) {
        probs.push_back(problems[b] != nullptr ? *problems[b] : Problem{});
        rejected[b].engine = "batched";
        if (problems[b] == nullptr) {
            rejected[b].status = PdhgStatus::NumericalFailure;
            rejected[b].message = "no problem";
        } else if (prevalidate(probs[b], rejected[b])) {
            slot[b] = static_cast<std::ptrdiff_t>(idx.size());
            idx.push_back(b);
        }
    }
    B = idx.size();
    if (B == 0) return;
    const ObjSense sense = probs[idx[0]].sense();
    for (std::size_t k : idx)
        if (probs[k].sense() != sense) throw Error(Status::InvalidArgument, "PdhgBatch: instances differ in sense");
    sign = sense == ObjSense::Maximize ? -1.0 : 1.0;

    // -- stack the instances block-diagonally (natural order)
    for (std::size_t k : idx) {
        const Problem& p = probs[k];
        row_off.push_back(row_off.back() + p.num_rows());
        col_off.push_back(col_off.back() + p.num_cols());
        nnz += static_cast<std::size_t>(p.matrix().nnz());
    }
    m = static_cast<std::size_t>(row_off.back());
    n = static_cast<std::size_t>(col_off.back());
    std::vector<int32_t> rp(m + 1, 0), ci(std::max<std::size_t>(nnz, 1));
    std::vector<double> vals(std::max<std::size_t>(nnz, 1));
    oconst.resize(B);
    {
        std::size_t q = 0;
        for (std::size_t b = 0; b < B; ++b) {
            const Problem& p = probs[idx[b]];
            const CsrMatrix& A = p.matrix();
            const auto r0 = static_cast<std::size_t>(row_off[b]);
            for (Idx r = 0; r < A.rows(); ++r) {
                for (Idx k = A.row_ptr()[r]; k < A.row_ptr()[r + 1]; ++k, ++q) {
                    ci[q] = col_off[b] + A.col_idx()[k];
                    vals[q] = A.values()[k];
                }
                rp[r0 + static_cast<std::size_t>(r) + 1] = static_cast<int32_t>(q);
            }
            oconst[b] = p.objective_constant();
        }
    }
    {
        DevMem d_rp(ctx, (m + 1) * sizeof(int32_t)), d_ci(ctx, ci.size() * sizeof(int32_t)),
            d_v(ctx, vals.size() * sizeof(double));
        hip_ok(sov_hip_memcpy_h2d(ctx, d_rp.p, rp.data(), (m + 1) * sizeof(int32_t)), "upload");
        hip_ok(sov_hip_memcpy_h2d(ctx, d_ci.p, ci.data(), ci.size() * sizeof(int32_t)), "upload");
        hip_ok(sov_hip_memcpy_h2d(ctx, d_v.p, vals.data(), vals.size() * sizeof(double)), "upload");
        hip_ok(sov_hip_csr_create(ctx, static_cast<int32_t>(m), static_cast<int32_t>(n), static_cast<int32_t>(nnz),
                                  d_rp.as<int32_t>(), d_ci.as<int32_t>(), d_v.as<double>(), SOV_CSR_OWN_VALUES, &csr.h),
               "csr_create");
        // the handle copied the values and borrows the structure arrays
        a_rp = std::move(d_rp);
        a_ci = std::move(d_ci);
    }

    // -- Ruiz + Pock-Chambolle on the stack: row and column norms never
    //    cross instances, so these are each instance's own factors (the
    //    Ruiz early stop is taken over the whole batch)
    {
        DevMem d_rs(ctx, std::max<std::size_t>(m, 1) * 8), d_cs(ctx, std::max<std::size_t>(n, 1) * 8);
        const std::vector<Real> ones_m(m, 1.0), ones_n(n, 1.0);
        if (m) hip_ok(sov_hip_memcpy_h2d(ctx, d_rs.p, ones_m.data(), m * 8), "upload");
        if (n) hip_ok(sov_hip_memcpy_h2d(ctx, d_cs.p, ones_n.data(), n * 8), "upload");
        const ScalingOptions so;
        if (so.apply_ruiz)
            for (int sweep = 0; sweep < so.ruiz_iterations; ++sweep) {
                double worst = 0.0;
                hip_ok(sov_hip_scale_ruiz_sweep(ctx, csr.h, d_rs.as<double>(), d_cs.as<double>(), &worst), "ruiz");
                if (worst < so.ruiz_tolerance) break;
            }
        if (so.apply_pock_chambolle)
            hip_ok(sov_hip_scale_pock_chambolle(ctx, csr.h, so.pock_chambolle_alpha, d_rs.as<double>(),
                                                d_cs.as<double>()),
                   "pock_chambolle");
        rs.resize(m);
        cs.resize(n);
        if (m) hip_ok(sov_hip_memcpy_d2h(ctx, rs.data(), d_rs.p, m * 8), "download scaling");
        if (n) hip_ok(sov_hip_memcpy_d2h(ctx, cs.data(), d_cs.p, n * 8), "download scaling");
        hip_ok(sov_hip_csr_scale(ctx, csr.h, d_rs.as<double>(), d_cs.as<double>()), "csr_scale");
    }
    ord.row.resize(m);
    ord.col.resize(n);
    hip_ok(sov_hip_csr_order(ctx, csr.h, 0, ord.row.data()), "order");
    hip_ok(sov_hip_csr_order(ctx, csr.h, 1, ord.col.data()), "order");
    hip_ok(sov_hip_batch_create(ctx, csr.h, static_cast<int32_t>(B), row_off.data(), col_off.data(), &bh.b),
           "batch_create");
    if (o.batch_stage_bytes >= 0) hip_ok(sov_hip_batch_set_stage_bytes(bh.b, o.batch_stage_bytes), "stage_bytes");
    hip_ok(sov_hip_batch_set_tile_share(bh.b, o.batch_tile_share), "tile_share");
    hip_ok(sov_hip_batch_set_split(bh.b, o.batch_split ? 1 : 0), "split");
    const std::size_t mb = std::max<std::size_t>(m, 1) * 8, nb = std::max<std::size_t>(n, 1) * 8;
    d_olo_r = DevMem(ctx, mb);
    d_ohi_r = DevMem(ctx, mb);
    d_olo_c = DevMem(ctx, nb);
    d_ohi_c = DevMem(ctx, nb);
    d_oc = DevMem(ctx, nb);
    d_act = DevMem(ctx, mb);
    d_act32 = DevMem(ctx, mb);
    d_rs_dev = DevMem(ctx, mb);
    d_cs_dev = DevMem(ctx, nb);
    {
        const std::vector<Real> pr = DeviceOrder::to_dev(rs, ord.row), pc = DeviceOrder::to_dev(cs, ord.col);
        if (m) hip_ok(sov_hip_memcpy_h2d(ctx, d_rs_dev.p, pr.data(), m * 8), "upload");
        if (n) hip_ok(sov_hip_memcpy_h2d(ctx, d_cs_dev.p, pc.data(), n * 8), "upload");
    }

    // -- per-instance power iteration (reference pdhg.cpp:48-79 on each block)
    est.assign(B, 0.0);
    {
        DevMem u(ctx, nb), v(ctx, mb), w(ctx, mb);
        const std::vector<Real> ones_m(m, 1.0), ones_n(n, 1.0);
        if (n) hip_ok(sov_hip_memcpy_h2d(ctx, u.p, ones_n.data(), n * 8), "upload");
        if (m) hip_ok(sov_hip_memcpy_h2d(ctx, v.p, ones_m.data(), m * 8), "upload");
        std::vector<char> live(B, 1);
        std::vector<double> sq(B), sq2(B), fac(B);
        for (int it = 0; it < o.power_iterations; ++it) {
            hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, -1.0, u.p, 0.0, w.p), "power spmv");
            hip_ok(sov_hip_axpy_f64(ctx, static_cast<int64_t>(m), 1.0, v.as<double>(), w.as<double>()), "power axpy");
            hip_ok(sov_hip_batch_sumsq(ctx, bh.b, 0, w.as<double>(), sq.data()), "power norms");
            for (std::size_t b = 0; b < B; ++b) {
                const double wn = std::sqrt(sq[b]);
                if (live[b] && wn < 1e-300) live[b] = 0;
                fac[b] = live[b] ? 1.0 / wn : 1.0;
            }
            hip_ok(sov_hip_batch_scale(ctx, bh.b, 0, fac.data(), w.as<double>()), "power scale");
            hip_ok(sov_hip_spmv_device(ctx, csr.h, 1, SOV_PREC_F64, -1.0, w.p, 0.0, u.p), "power spmv_t");
            if (m) hip_ok(sov_hip_memcpy_d2d(ctx, v.p, w.p, m * 8), "power copy");
            hip_ok(sov_hip_batch_sumsq(ctx, bh.b, 1, u.as<double>(), sq.data()), "power norms");
            hip_ok(sov_hip_batch_sumsq(ctx, bh.b, 0, v.as<double>(), sq2.data()), "power norms");
            for (std::size_t b = 0; b < B; ++b) {
                if (live[b]) {
                    est[b] = std::sqrt(sq[b] + sq2[b]);
                    if (est[b] < 1e-300) live[b] = 0;
                }
                fac[b] = live[b] ? 1.0 / est[b] : 1.0;
            }
            hip_ok(sov_hip_batch_scale(ctx, bh.b, 1, fac.data(), u.as<double>()), "power scale");
            hip_ok(sov_hip_batch_scale(ctx, bh.b, 0, fac.data(), v.as<double>()), "power scale");
        }
    }
    tau.resize(B);
    sigma.resize(B);
    weight.resize(B);
    for (std::size_t b = 0; b < B; ++b) {
        const double eta = est[b] > 1e-9 ? o.step_size_safety / est[b] : o.step_size_safety;
        double w = 1.0;
        if (primal_weight != nullptr && idx[b] < primal_weight->size()) {
            w = (*primal_weight)[idx[b]];
            if (!(w > 0.0) || !std::isfinite(w)) w = 1.0;
        }
        tau[b] = eta / w;
        sigma[b] = eta * w;
        weight[b] = w;
    }

    // -- iterate states: FP64 always (the only one for an FP64 run, the
    //    polish state of an FP32 run); FP32 for an FP32 run.  Each instance
    //    lives in one of them (active = 1 or 2) and moves to FP64 on its own
    //    once it is near tolerance or stalled.
    f32 = o.precision == Precision::F32;
    if (f32) hip_ok(sov_hip_csr_make_f32(ctx, csr.h), "make_f32");
    if (f32) S32.allocate(ctx, SOV_PREC_F32, m, n);
    if (!f32 || o.fp64_polish) S64.allocate(ctx, SOV_PREC_F64, m, n);
    // starting point: x = 0 and y = 0 or the optional dual warm starts
    // (original units -> scaled)
    x_start.assign(n, 0.0);
    y_start.assign(m, 0.0);
    cold.assign(B, 1);
    if (warm_dual != nullptr)
        for (std::size_t b = 0; b < B; ++b) {
            if (idx[b] >= warm_dual->size() || (*warm_dual)[idx[b]].empty()) continue;
            const std::vector<Real>& w = (*warm_dual)[idx[b]];
            const auto r0 = static_cast<std::size_t>(row_off[b]);
            const auto rows = static_cast<std::size_t>(row_off[b + 1] - row_off[b]);
            if (w.size() != rows) throw Error(Status::DimensionMismatch, "PdhgBatch: warm_dual length");
            for (std::size_t i = 0; i < rows; ++i) y_start[r0 + i] = sign * w[i] / rs[r0 + i];
        }
    if (o.autotune && nnz > 0) hip_ok(sov_hip_csr_tune(ctx, csr.h, f32 ? SOV_PREC_F32 : SOV_PREC_F64), "tune");
    // the batched steps run on the tiled layouts while most entries iterate
    // in one precision (batch_tile_share), else on the CSR parts' block lists
    for (int orient : {0, 1}) {
        const bool tiled = o.batch_tile_share >= 0.0 && nnz > 0 && m > 0 && n > 0;
        hip_ok(sov_hip_csr_set_tiles(ctx, csr.h, orient, tiled ? 1 : 0), "tiles");
    }
    setup_seconds = seconds_since(t0);
}

/// The stacked natural-order objective and bounds of the instances as the
/// session holds them now.
void PdhgBatchSession::Impl::stacked_vectors(std::vector<Real>& c, std::vector<Real>& clo, std::vector<Real>& chi,
                                             std::vector<Real>& rlo, std::vector<Real>& rhi) const {
    c.resize(n);
    clo.resize(n);
    chi.resize(n);
    rlo.resize(m);
    rhi.resize(m);
    for (std::size_t b = 0; b < B; ++b) {
        const Problem& p = probs[idx[b]];
        const auto r0 = static_cast<std::size_t>(row_off[b]), c0 = static_cast<std::size_t>(col_off[b]);
        std::copy(p.objective().begin(), p.objective().end(), c.begin() + static_cast<std::ptrdiff_t>(c0));
        std::copy(p.col_lower().begin(), p.col_lower().end(), clo.begin() + static_cast<std::ptrdiff_t>(c0));
        std::copy(p.col_upper().begin(), p.col_upper().end(), chi.begin() + static_cast<std::ptrdiff_t>(c0));
        std::copy(p.row_lower().begin(), p.row_lower().end(), rlo.begin() + static_cast<std::ptrdiff_t>(r0));
        std::copy(p.row_upper().begin(), p.row_upper().end(), rhi.begin() + static_cast<std::ptrdiff_t>(r0));
    }
}

/// Objective and bounds into the iterate states (scaled, internal sense,
/// device order; reference scaling.cpp:172-192, pdhg.cpp:250-252) and the
/// original-space arrays of the checks.
void PdhgBatchSession::Impl::upload_data() {
    std::vector<Real> c_nat, clo, chi, rlo, rhi;
    stacked_vectors(c_nat, clo, chi, rlo, rhi);
    std::vector<Real> c_s(n), rlo_s(m), rhi_s(m);
    clo_s.resize(n);
    chi_s.resize(n);
    for (std::size_t j = 0; j < n; ++j) {
        c_s[j] = (c_nat[j] * cs[j]) * sign;
        clo_s[j] = is_finite_bound(clo[j]) ? clo[j] / cs[j] : clo[j];
        chi_s[j] = is_finite_bound(chi[j]) ? chi[j] / cs[j] : chi[j];
    }
    for (std::size_t i = 0; i < m; ++i) {
        rlo_s[i] = is_finite_bound(rlo[i]) ? rlo[i] * rs[i] : rlo[i];
        rhi_s[i] = is_finite_bound(rhi[i]) ? rhi[i] * rs[i] : rhi[i];
    }
    const std::vector<Real> c_d = DeviceOrder::to_dev(c_s, ord.col), clo_d = DeviceOrder::to_dev(clo_s, ord.col),
                            chi_d = DeviceOrder::to_dev(chi_s, ord.col), rlo_d = DeviceOrder::to_dev(rlo_s, ord.row),
                            rhi_d = DeviceOrder::to_dev(rhi_s, ord.row);
    for (DeviceState* S : {&S32, &S64}) {
        if (S->x.p == nullptr) continue;   // state not allocated
        S->upload(ctx, S->c, c_d);
        S->upload(ctx, S->clo, clo_d);
        S->upload(ctx, S->chi, chi_d);
        S->upload(ctx, S->rlo, rlo_d);
        S->upload(ctx, S->rhi, rhi_d);
    }
    auto upload_dev = [&](const DevMem& d, std::span<const Real> v, const std::vector<int32_t>& perm) {
        if (v.empty()) return;
        const std::vector<Real> pv = DeviceOrder::to_dev(v, perm);
        hip_ok(sov_hip_memcpy_h2d(ctx, d.p, pv.data(), pv.size() * 8), "upload");
    };
    upload_dev(d_olo_r, rlo, ord.row);
    upload_dev(d_ohi_r, rhi, ord.row);
    upload_dev(d_olo_c, clo, ord.col);
    upload_dev(d_ohi_c, chi, ord.col);
    upload_dev(d_oc, c_nat, ord.col);
}

/// Every state's point from (x_start, y_start): x projected onto the column
/// box and the slacks s = Proj_S(A x); a cold instance starts as the
/// reference does, x = Proj_X(0), s = Proj_S(0) (pdhg.cpp:268-270), with y
/// from the optional dual warm start.
void PdhgBatchSession::Impl::place_start() {
    const std::size_t mb = std::max<std::size_t>(m, 1) * 8, nb = std::max<std::size_t>(n, 1) * 8;
    std::vector<Real> x_act(n, 0.0);   // the point whose activity sets s (0 for cold instances)
    for (std::size_t b = 0; b < B; ++b) {
        if (cold[b]) continue;
        for (auto j = static_cast<std::size_t>(col_off[b]); j < static_cast<std::size_t>(col_off[b + 1]); ++j) {
            Real v = x_start[j];
            if (v < clo_s[j]) v = clo_s[j];
            if (v > chi_s[j]) v = chi_s[j];
            x_act[j] = v;
        }
    }
    DevMem dx(ctx, nb), dy(ctx, mb), da(ctx, nb);
    const std::vector<Real> xd = DeviceOrder::to_dev(x_start, ord.col), yd = DeviceOrder::to_dev(y_start, ord.row),
                            ad = DeviceOrder::to_dev(x_act, ord.col);
    if (n) hip_ok(sov_hip_memcpy_h2d(ctx, dx.p, xd.data(), n * 8), "upload start");
    if (m) hip_ok(sov_hip_memcpy_h2d(ctx, dy.p, yd.data(), m * 8), "upload start");
    if (n) hip_ok(sov_hip_memcpy_h2d(ctx, da.p, ad.data(), n * 8), "upload start");
    hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, 1.0, da.p, 0.0, d_act.p), "start activity");
    for (DeviceState* S : {&S32, &S64})
        if (S->x.p != nullptr)
            hip_ok(sov_hip_pdhg_set_point(ctx, &S->st, dx.as<double>(), dy.as<double>(), d_act.as<double>()),
                   "set_point");
    hip_ok(sov_hip_sync(ctx), "sync");
}

std::size_t PdhgBatchSession::Impl::stacked(std::size_t b, const char* what) const {
    if (b >= probs.size()) throw Error(Status::InvalidArgument, std::string(what) + ": no instance " + std::to_string(b));
    if (slot[b] < 0)
        throw Error(Status::InvalidArgument, std::string(what) + ": instance " + std::to_string(b) +
                                                 " failed validation and is not in the batch");
    return static_cast<std::size_t>(slot[b]);
}

void PdhgBatchSession::Impl::replace(std::size_t b, Problem p, const char* what) {
    require_valid(p, what);
    probs[b] = std::move(p);
    data_dirty = true;
}

PdhgBatchResult PdhgBatchSession::Impl::run() {
    PdhgBatchResult out;
    const auto t0 = Clock::now();
    out.results = rejected;
    if (B == 0) return out;
    try {
        if (data_dirty) upload_data();
        place_start();
        data_dirty = false;

        std::vector<int32_t> active(B, f32 ? 1 : 2), since(B, 0), flags32(B, 0), flags64(B, 0), since_stop(B, 0);
        std::vector<int32_t> stop_state(B, 0), promote(B, 0), use_avg(B, 0);
        std::vector<char> done(B, 0), cand_avg(B, 0), failed(B, 0);
        std::vector<long long> its(B, 0), promoted_at(B, -1), restarts(B, 0);
        std::vector<Real> res_at_restart(B, kInfinity), prev_cand(B, kInfinity), best_rel(B, kInfinity);
        std::vector<int> stall(B, 0);
        hip_ok(sov_hip_batch_set(ctx, bh.b, tau.data(), sigma.data(), active.data(), since.data()), "batch_set");
        std::vector<int32_t> active_sent = active;

        sov_hip_check_args ca{};
        ca.tol_scaled = o.tolerance;
        ca.tol_orig = std::max(o.tolerance, Real{1e-6});
        ca.obj_sign = sign;
        ca.obj_const = 0.0;
        ca.row_scale = d_rs_dev.as<double>();
        ca.col_scale = d_cs_dev.as<double>();
        ca.orig_row_lo = d_olo_r.as<double>();
        ca.orig_row_hi = d_ohi_r.as<double>();
        ca.orig_col_lo = d_olo_c.as<double>();
        ca.orig_col_hi = d_ohi_c.as<double>();
        ca.orig_c = d_oc.as<double>();
        std::vector<sov_hip_check_out> co32(B), co64(B);

        const long long period = std::max<long long>(o.restart_check_period, 1);
        GraphHandle graph;
        long long graph_key = -1, active_version = 0;
        bool has32 = f32, has64 = !f32;
        // an interval: the resident instances' whole interval in one launch
        // per precision and (batch_split) the FP64 step launches on the
        // batch's stream, beside the FP32 step launches on the context
        // stream, joined before the check
        auto run_steps = [&](long long count) {
            for (DeviceState* S : {&S32, &S64}) {
                if ((S == &S32 && !has32) || (S == &S64 && !has64)) continue;
                hip_ok(sov_hip_batch_resident_steps(ctx, csr.h, bh.b, &S->st, static_cast<int>(count)),
                       "batch resident steps");
            }
            for (long long k = 0; k < count; ++k)
                for (DeviceState* S : {&S32, &S64}) {
                    if ((S == &S32 && !has32) || (S == &S64 && !has64)) continue;
                    hip_ok(sov_hip_batch_primal_step(ctx, csr.h, bh.b, &S->st, static_cast<int>(k)), "batch primal_step");
                    hip_ok(sov_hip_batch_dual_step(ctx, csr.h, bh.b, &S->st, static_cast<int>(k)), "batch dual_step");
                }
            hip_ok(sov_hip_batch_resident_join(ctx, bh.b), "batch resident join");
        };
        const bool adaptive = o.restart_scheme == PdhgRestartScheme::Adaptive;
        std::size_t n_active = B;
        long long it = 0;
        const auto t_loop = Clock::now();
        out.setup_seconds = seconds_since(t0);
        bool timed_out = false;
        while (n_active > 0 && it < o.max_iterations) {
            if (o.time_limit_seconds > 0.0 && seconds_since(t0) > o.time_limit_seconds) {
                timed_out = true;
                break;
            }
            has32 = has64 = false;
            for (std::size_t b = 0; b < B; ++b) {
                has32 = has32 || active[b] == 1;
                has64 = has64 || active[b] == 2;
            }
            const long long count = std::min<long long>(period, o.max_iterations - it);
            if (o.use_graphs && count == period) {
                // the launches' grids follow the active blocks: re-capture when they change
                const long long key = active_version * 4 + (has32 ? 1 : 0) + (has64 ? 2 : 0);
                if (graph.g == nullptr || graph_key != key) {
                    graph.reset();
                    hip_ok(sov_hip_graph_begin(ctx), "graph_begin");
                    run_steps(count);
                    hip_ok(sov_hip_graph_end(ctx, &graph.g), "graph_end");
                    graph_key = key;
                }
                hip_ok(sov_hip_graph_launch(ctx, graph.g), "graph_launch");
            } else {
                run_steps(count);
            }
            it += count;
            for (std::size_t b = 0; b < B; ++b)
                if (active[b]) since[b] += static_cast<int32_t>(count);
            hip_ok(sov_hip_batch_set(ctx, bh.b, nullptr, nullptr, nullptr, since.data()), "batch_set");
            if (count != period) break;   // iteration limit inside the interval

            if (has32) {
                ca.st = &S32.st;
                ca.act_avg = d_act32.as<double>();
                hip_ok(sov_hip_batch_check(ctx, csr.h, bh.b, &ca, oconst.data(), co32.data()), "batch check");
            }
            if (has64) {
                ca.st = &S64.st;
                ca.act_avg = d_act.as<double>();
                hip_ok(sov_hip_batch_check(ctx, csr.h, bh.b, &ca, oconst.data(), co64.data()), "batch check");
            }
            bool any32 = false, any64 = false, any_promote = false;
            Real worst = 0.0;
            for (std::size_t b = 0; b < B; ++b) {
                if (!active[b]) continue;
                const sov_hip_check_out& q = active[b] == 1 ? co32[b] : co64[b];
                const Real rel_c = std::max({q.rel_cur[0], q.rel_cur[1], q.rel_cur[2]});
                const Real rel_a = std::max({q.rel_avg[0], q.rel_avg[1], q.rel_avg[2]});
                const bool ab = adaptive ? rel_a <= rel_c : q.kkt_avg <= q.kkt_cur;
                const Real ores = ab ? q.orig_avg : q.orig_cur;
                const Real rel_worst = ab ? rel_a : rel_c;
                worst = std::max(worst, rel_worst);
                if (!std::isfinite(q.kkt_cur) || !std::isfinite(q.kkt_avg) || !std::isfinite(ores)) {
                    failed[b] = 1;
                    stop_state[b] = active[b];
                    active[b] = 0;
                    its[b] = it;
                    --n_active;
                    continue;
                }
                const bool stop = o.termination == PdhgTermination::Reference ? (ores < o.tolerance)
                                                                              : (rel_worst <= o.tolerance);
                if (stop) {
                    done[b] = 1;
                    cand_avg[b] = ab;
                    since_stop[b] = since[b];
                    stop_state[b] = active[b];
                    its[b] = it;
                    active[b] = 0;
                    --n_active;
                    continue;
                }
                if (rel_worst < best_rel[b] * 0.99) {
                    best_rel[b] = rel_worst;
                    stall[b] = 0;
                } else {
                    ++stall[b];
                }
                if (active[b] == 1 && o.fp64_polish) {
                    const Real gauge = o.termination == PdhgTermination::Reference ? ores : rel_worst;
                    if (gauge <= o.fp32_switch_factor * o.tolerance || stall[b] >= o.fp32_stall_checks) {
                        promote[b] = 1;
                        use_avg[b] = ab;
                        any_promote = true;
                        continue;
                    }
                }
                bool do_restart;
                if (!adaptive) {
                    if (!std::isfinite(res_at_restart[b])) res_at_restart[b] = q.kkt_cur;
                    do_restart = q.kkt_avg <= o.restart_sufficient_decrease * res_at_restart[b];
                } else {
                    if (!std::isfinite(res_at_restart[b])) res_at_restart[b] = rel_c;
                    do_restart = rel_worst <= o.restart_beta_sufficient * res_at_restart[b] ||
                                 (rel_worst <= o.restart_beta_necessary * res_at_restart[b] && rel_worst > prev_cand[b]) ||
                                 static_cast<Real>(since[b]) >= o.restart_beta_artificial * static_cast<Real>(it);
                    prev_cand[b] = rel_worst;
                }
                if (do_restart) {
                    const int32_t f = (!adaptive || ab) ? 1 : 0;
                    (active[b] == 1 ? flags32 : flags64)[b] = f;
                    any32 = any32 || (f && active[b] == 1);
                    any64 = any64 || (f && active[b] == 2);
                    res_at_restart[b] = adaptive ? rel_worst : q.kkt_avg;
                    prev_cand[b] = kInfinity;
                    ++restarts[b];
                    since[b] = -1;   // reset once the restart kernels have read it
                }
            }
            if (any32) hip_ok(sov_hip_batch_restart(ctx, bh.b, &S32.st, flags32.data(), d_act32.as<double>()), "restart");
            if (any64) hip_ok(sov_hip_batch_restart(ctx, bh.b, &S64.st, flags64.data(), d_act.as<double>()), "restart");
            if (any_promote) {
                // candidate -> FP64 state; a restart with since = 1 places it, s from its FP64 activity
                hip_ok(sov_hip_batch_promote(ctx, bh.b, &S32.st, &S64.st, promote.data(), use_avg.data()), "promote");
                std::vector<int32_t> since1 = since;
                for (std::size_t b = 0; b < B; ++b)
                    if (promote[b]) since1[b] = 1;
                hip_ok(sov_hip_batch_set(ctx, bh.b, nullptr, nullptr, nullptr, since1.data()), "batch_set");
                hip_ok(sov_hip_spmv_device(ctx, csr.h, 0, SOV_PREC_F64, 1.0, S64.x.p, 0.0, d_act.p), "polish activity");
                hip_ok(sov_hip_batch_restart(ctx, bh.b, &S64.st, promote.data(), d_act.as<double>()), "polish place");
                for (std::size_t b = 0; b < B; ++b) {
                    if (!promote[b]) continue;
                    active[b] = 2;
                    since[b] = 0;
                    promoted_at[b] = it;
                    res_at_restart[b] = kInfinity;
                    prev_cand[b] = kInfinity;
                    promote[b] = 0;
                }
                if (o.verbose) std::fprintf(stderr, "  [pdhg batch] iter %lld: instances moved to FP64\n", it);
            }
            for (std::size_t b = 0; b < B; ++b) {
                flags32[b] = flags64[b] = 0;
                if (since[b] < 0) since[b] = 0;
            }
            // the block activity bits are recomputed only when `active` changed
            const bool changed = active != active_sent;
            hip_ok(sov_hip_batch_set(ctx, bh.b, nullptr, nullptr, changed ? active.data() : nullptr, since.data()),
                   "batch_set");
            if (changed) {
                active_sent = active;
                ++active_version;
            }
            if (o.verbose && (it / period) % 100 == 0)
                std::fprintf(stderr, "  [pdhg batch] iter %lld  active %zu/%zu  worst rel kkt %.3g  %.1fs\n", it,
                             n_active, B, static_cast<double>(worst), seconds_since(t0));
        }
        hip_ok(sov_hip_sync(ctx), "sync");
        out.loop_seconds = seconds_since(t_loop);
        out.iterations = it;

        // -- final points: each instance's candidate from the state it ended
        //    in, natural order, unscaled; the scaled one is the next start
        struct Host {
            std::vector<Real> x, y, xs, ys;
        };
        auto fetch = [&](DeviceState& S) {
            Host h;
            h.x = DeviceOrder::to_nat(S.download(ctx, S.x, n), ord.col);
            h.y = DeviceOrder::to_nat(S.download(ctx, S.y, m), ord.row);
            h.xs = DeviceOrder::to_nat(S.download(ctx, S.xsum, n), ord.col);
            h.ys = DeviceOrder::to_nat(S.download(ctx, S.ysum, m), ord.row);
            return h;
        };
        Host h32, h64;
        if (f32) h32 = fetch(S32);
        if (!f32 || o.fp64_polish) h64 = fetch(S64);
        for (std::size_t b = 0; b < B; ++b) {
            PdhgResult& r = out.results[idx[b]];
            const Problem& p = probs[idx[b]];
            const int where = (done[b] || failed[b]) ? stop_state[b] : active[b];
            const Host& hv = where == 1 ? h32 : h64;
            const auto r0 = static_cast<std::size_t>(row_off[b]), c0 = static_cast<std::size_t>(col_off[b]);
            const std::size_t mb = static_cast<std::size_t>(p.num_rows()), nb = static_cast<std::size_t>(p.num_cols());
            const bool ua = done[b] && cand_avg[b];
            const Real inv = 1.0 / static_cast<Real>(since_stop[b] > 0 ? since_stop[b] : 1);
            std::vector<Real> x(nb), y(mb);
            for (std::size_t j = 0; j < nb; ++j) x[j] = ua ? hv.xs[c0 + j] * inv : hv.x[c0 + j];
            for (std::size_t i = 0; i < mb; ++i) y[i] = ua ? hv.ys[r0 + i] * inv : hv.y[r0 + i];
            // the next solve starts from this candidate (a failed instance from zero)
            for (std::size_t j = 0; j < nb; ++j) x_start[c0 + j] = failed[b] ? 0.0 : x[j];
            for (std::size_t i = 0; i < mb; ++i) y_start[r0 + i] = failed[b] ? 0.0 : y[i];
            cold[b] = failed[b] ? 1 : 0;
            const Scaling sc = Scaling::from_factors(std::vector<Real>(rs.begin() + static_cast<std::ptrdiff_t>(r0),
                                                                       rs.begin() + static_cast<std::ptrdiff_t>(r0 + mb)),
                                                     std::vector<Real>(cs.begin() + static_cast<std::ptrdiff_t>(c0),
                                                                       cs.begin() + static_cast<std::ptrdiff_t>(c0 + nb)));
            auto [ux, uy] = unscale_point(sign, sc, x, y);
            r = PdhgResult{};
            r.engine = "batched";
            r.operator_norm_estimate = est[b];
            r.primal_weight = weight[b];
            r.iterations = done[b] || failed[b] ? its[b] : it;
            r.restarts = restarts[b];
            if (promoted_at[b] >= 0) r.fp64_polish_iterations = r.iterations - promoted_at[b];
            if (failed[b]) {
                r.status = PdhgStatus::NumericalFailure;
                r.message = "non-finite KKT residual";
            } else if (done[b]) {
                r.status = PdhgStatus::Optimal;
            } else {
                r.status = timed_out ? PdhgStatus::TimeLimit : PdhgStatus::IterationLimit;
            }
            finish_result(p, o, std::move(ux), std::move(uy), r);
            last_x[idx[b]] = r.primal;
            last_y[idx[b]] = r.dual;
            r.seconds = seconds_since(t0);
            r.setup_seconds = out.setup_seconds;
            r.loop_seconds = out.loop_seconds;
        }
    } catch (const Error& e) {
        out.message = std::string("device failure: ") + e.what();
        for (std::size_t k : idx) {
            out.results[k].status = PdhgStatus::NumericalFailure;
            out.results[k].message = out.message;
        }
    }
    out.seconds = seconds_since(t0);
    return out;
}

// ----------------------------------------------------------------- sessions

PdhgBatchSession::PdhgBatchSession(PdhgOptions options, const std::vector<const Problem*>& problems,
                                   Backend& backend, const std::vector<std::vector<Real>>* warm_dual,
                                   const std::vector<Real>* primal_weight)
    : impl_(std::make_unique<Impl>(std::move(options), problems, backend, warm_dual, primal_weight)) {}

PdhgBatchSession::~PdhgBatchSession() = default;

std::size_t PdhgBatchSession::size() const noexcept { return impl_->probs.size(); }

const Problem& PdhgBatchSession::problem(std::size_t b) const {
    if (b >= impl_->probs.size()) throw Error(Status::InvalidArgument, "PdhgBatchSession::problem: no instance " + std::to_string(b));
    return impl_->probs[b];
}

namespace {

/// The instance rebuilt from its matrix with the given vectors (its
/// sense and objective constant kept).
Problem rebuilt(const Problem& p, std::vector<Real> c, std::vector<Real> clo, std::vector<Real> chi,
                std::vector<Real> rlo, std::vector<Real> rhi) {
    return Problem::from_csr(p.matrix(), std::move(c), std::move(clo), std::move(chi), std::move(rlo), std::move(rhi),
                             p.sense(), p.objective_constant(), p.name());
}

std::vector<Real> vec(std::span<const Real> s) { return {s.begin(), s.end()}; }

void require_length(std::size_t got, Idx want, const char* what) {
    if (got != static_cast<std::size_t>(want))
        throw Error(Status::DimensionMismatch, std::string(what) + ": " + std::to_string(got) + " values for " +
                                                   std::to_string(want));
}

}  // namespace

void PdhgBatchSession::set_row_bounds(std::size_t b, std::span<const Real> lower, std::span<const Real> upper) {
    Impl& s = *impl_;
    s.stacked(b, "set_row_bounds");
    const Problem& p = s.probs[b];
    require_length(lower.size(), p.num_rows(), "set_row_bounds");
    require_length(upper.size(), p.num_rows(), "set_row_bounds");
    s.replace(b, rebuilt(p, vec(p.objective()), vec(p.col_lower()), vec(p.col_upper()), vec(lower), vec(upper)),
              "set_row_bounds");
}

void PdhgBatchSession::set_objective(std::size_t b, std::span<const Real> c) {
    Impl& s = *impl_;
    s.stacked(b, "set_objective");
    const Problem& p = s.probs[b];
    require_length(c.size(), p.num_cols(), "set_objective");
    s.replace(b, rebuilt(p, vec(c), vec(p.col_lower()), vec(p.col_upper()), vec(p.row_lower()), vec(p.row_upper())),
              "set_objective");
}

void PdhgBatchSession::set_column_bounds(std::size_t b, std::span<const Real> lower, std::span<const Real> upper) {
    Impl& s = *impl_;
    s.stacked(b, "set_column_bounds");
    const Problem& p = s.probs[b];
    require_length(lower.size(), p.num_cols(), "set_column_bounds");
    require_length(upper.size(), p.num_cols(), "set_column_bounds");
    s.replace(b, rebuilt(p, vec(p.objective()), vec(lower), vec(upper), vec(p.row_lower()), vec(p.row_upper())),
              "set_column_bounds");
}

void PdhgBatchSession::set_start(std::size_t b, std::span<const Real> x, std::span<const Real> y) {
    Impl& s = *impl_;
    const std::size_t k = s.stacked(b, "set_start");
    const Problem& p = s.probs[b];
    const auto r0 = static_cast<std::size_t>(s.row_off[k]), c0 = static_cast<std::size_t>(s.col_off[k]);
    if (!x.empty()) {
        require_length(x.size(), p.num_cols(), "set_start (x)");
        for (std::size_t j = 0; j < x.size(); ++j) s.x_start[c0 + j] = x[j] / s.cs[c0 + j];
    }
    if (!y.empty()) {
        require_length(y.size(), p.num_rows(), "set_start (y)");
        for (std::size_t i = 0; i < y.size(); ++i) s.y_start[r0 + i] = s.sign * y[i] / s.rs[r0 + i];
    }
    s.cold[k] = 0;
}

void PdhgBatchSession::warm_start_from(std::size_t b, std::size_t source) {
    Impl& s = *impl_;
    s.stacked(b, "warm_start_from");
    s.stacked(source, "warm_start_from");
    if (s.last_x[source].empty())
        throw Error(Status::InvalidArgument, "warm_start_from: instance " + std::to_string(source) + " has no solution yet");
    if (s.probs[b].num_rows() != s.probs[source].num_rows() || s.probs[b].num_cols() != s.probs[source].num_cols())
        throw Error(Status::DimensionMismatch, "warm_start_from: instances " + std::to_string(b) + " and " +
                                                   std::to_string(source) + " differ in size");
    const std::vector<Real> x = s.last_x[source], y = s.last_y[source];
    set_start(b, x, y);
}

PdhgBatchResult PdhgBatchSession::solve() {
    const double first_setup = impl_->setup_seconds;
    impl_->setup_seconds = 0.0;
    PdhgBatchResult out = impl_->run();
    // the first solve carries the session's construction
    out.setup_seconds += first_setup;
    out.seconds += first_setup;
    for (PdhgResult& r : out.results) {
        r.setup_seconds += first_setup;
        r.seconds += first_setup;
    }
    return out;
}

PdhgBatchResult PdhgBatch::solve(const std::vector<const Problem*>& problems, Backend& backend,
                                  const std::vector<std::vector<Real>>* warm_dual,
                                  const std::vector<Real>* primal_weight) const {
    try {
        PdhgBatchSession session(options_, problems, backend, warm_dual, primal_weight);
        return session.solve();
    } catch (const Error& e) {
        PdhgBatchResult out;
        out.message = e.what();
        out.results.resize(problems.size());
        for (PdhgResult& r : out.results) {
            r.engine = "batched";
            r.status = PdhgStatus::NumericalFailure;
            r.message = out.message;
        }
        return out;
    }
}

}  // namespace sov
