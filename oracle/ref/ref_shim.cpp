// C binding over the reference's own host PDHG, compiled from the reference
// sources in place (oracle/Makefile, target `ref`).  TEST INFRASTRUCTURE ONLY:
// used by tests/golden/make_golden.py to write fixtures, and by bench.py as
// the "reference" CPU baseline.  Nothing in the product links it.
//
// Exposes: Pdhg::solve on the Host backend (src/l1/pdhg.cpp:218), Scaling
// factors (src/l1/scaling.cpp:19), and the LP/MPS readers
// (src/io/lp_reader.cpp:498, src/io/mps_reader.cpp:566) so the hand-written
// test LPs of tests/test_pdhg.cpp can be turned into CSR fixtures.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sovereign/backend.hpp"
#include "sovereign/cancellation.hpp"
#include "sovereign/io.hpp"
#include "sovereign/pdhg.hpp"
#include "sovereign/problem.hpp"
#include "sovereign/scaling.hpp"

using namespace sov;

namespace {

Problem build(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
              const double* c, const double* clo, const double* chi, const double* rlo,
              const double* rhi, int32_t maximize, double obj_const) {
    Problem::Builder b;
    b.set_sense(maximize ? ObjSense::Maximize : ObjSense::Minimize);
    b.set_objective_constant(obj_const);
    for (int32_t j = 0; j < n; ++j) {
        const Idx col = b.add_column("c" + std::to_string(j));
        b.set_objective_coefficient(col, c[j]);
        b.set_column_bounds(col, clo[j], chi[j]);
    }
    for (int32_t i = 0; i < m; ++i) b.add_row("r" + std::to_string(i), rlo[i], rhi[i]);
    for (int32_t r = 0; r < m; ++r)
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) b.add_coefficient(r, ci[k], v[k]);
    return b.finish();
}

Problem g_parsed;

}  // namespace

#define REF_API __attribute__((visibility("default")))
extern "C" {

// stats: [objective, primal_inf, dual_inf, gap, iterations, restarts, seconds, op_norm]
REF_API int ref_pdhg_solve(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                   const double* c, const double* clo, const double* chi, const double* rlo,
                   const double* rhi, int32_t maximize, double obj_const, double tol,
                   int64_t max_it, int64_t period, double time_limit, double* x, double* y,
                   double* rc, double* stats, int64_t trace_cap, double* trace_res,
                   int64_t* trace_it, int64_t* trace_len) {
    try {
        const Problem p = build(m, n, rp, ci, v, c, clo, chi, rlo, rhi, maximize, obj_const);
        PdhgOptions opt;
        opt.tolerance = tol;
        opt.max_iterations = max_it;
        opt.restart_check_period = period;
        opt.time_limit_seconds = time_limit;
        int64_t nt = 0;
        if (trace_res != nullptr) {
            opt.checkpoint = [&](long long it, const std::vector<Real>&, const std::vector<Real>&,
                                 Real r) {
                if (nt < trace_cap) {
                    trace_res[nt] = r;
                    trace_it[nt] = it;
                }
                ++nt;
            };
        }
        auto be = make_backend(BackendKind::Host);
        const PdhgResult r = Pdhg(opt).solve(p, *be);
        if (trace_len) *trace_len = nt;
        if (!r.primal.empty()) std::memcpy(x, r.primal.data(), sizeof(double) * r.primal.size());
        if (!r.dual.empty()) std::memcpy(y, r.dual.data(), sizeof(double) * r.dual.size());
        if (!r.reduced_costs.empty())
            std::memcpy(rc, r.reduced_costs.data(), sizeof(double) * r.reduced_costs.size());
        stats[0] = r.objective;
        stats[1] = r.primal_infeasibility;
        stats[2] = r.dual_infeasibility;
        stats[3] = r.duality_gap;
        stats[4] = static_cast<double>(r.iterations);
        stats[5] = static_cast<double>(r.restarts);
        stats[6] = r.seconds;
        stats[7] = r.operator_norm_estimate;
        return static_cast<int>(r.status);
    } catch (...) {
        return -1;
    }
}

// Iteration rate of the reference's own loop on the Host backend (the CPU
// baseline): Pdhg::solve with the default options (check period 40), the
// wall clock stamped at each checkpoint (seconds since the solve started)
// with its iteration, and the solve cancelled through its CancellationToken
// once `checks` checkpoints have been stamped.  build_seconds receives the
// time Problem::Builder took; *stamped the checkpoints recorded.
REF_API int ref_pdhg_rate(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                          const double* c, const double* clo, const double* chi, const double* rlo,
                          const double* rhi, int32_t maximize, double obj_const, int64_t checks,
                          double* stamps, int64_t* iters, int64_t* stamped, double* build_seconds) {
    using clock = std::chrono::steady_clock;
    try {
        const auto tb = clock::now();
        const Problem p = build(m, n, rp, ci, v, c, clo, chi, rlo, rhi, maximize, obj_const);
        *build_seconds = std::chrono::duration<double>(clock::now() - tb).count();
        CancellationToken stop;
        PdhgOptions opt;
        int64_t k = 0;
        const auto t0 = clock::now();
        opt.checkpoint = [&](long long it, const std::vector<Real>&, const std::vector<Real>&, Real) {
            if (k < checks) {
                stamps[k] = std::chrono::duration<double>(clock::now() - t0).count();
                iters[k] = it;
                ++k;
            }
            if (k >= checks) stop.cancel();
        };
        auto be = make_backend(BackendKind::Host);
        const PdhgResult r = Pdhg(opt).solve(p, *be, &stop);
        *stamped = k;
        return static_cast<int>(r.status);
    } catch (...) {
        return -1;
    }
}

// The reference's checkpoint hand-off (pdhg.hpp:121-133, called at
// pdhg.cpp:373-381): the first `cap` checkpoints of a default-options solve,
// each its iteration, original-space residual and unscaled candidate (x: n,
// y: m values per checkpoint, row-major in xs / ys); *count = checkpoints
// the solve made.  Returns the status.
REF_API int ref_pdhg_checkpoints(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                                 const double* c, const double* clo, const double* chi, const double* rlo,
                                 const double* rhi, int32_t maximize, double obj_const, double tol,
                                 int64_t max_it, int64_t cap, int64_t* its, double* res, double* xs, double* ys,
                                 int64_t* count) {
    try {
        const Problem p = build(m, n, rp, ci, v, c, clo, chi, rlo, rhi, maximize, obj_const);
        PdhgOptions opt;
        opt.tolerance = tol;
        opt.max_iterations = max_it;
        int64_t k = 0;
        opt.checkpoint = [&](long long it, const std::vector<Real>& x, const std::vector<Real>& y, Real r) {
            if (k < cap) {
                its[k] = it;
                res[k] = r;
                std::memcpy(xs + k * n, x.data(), sizeof(double) * static_cast<std::size_t>(n));
                std::memcpy(ys + k * m, y.data(), sizeof(double) * static_cast<std::size_t>(m));
            }
            ++k;
        };
        auto be = make_backend(BackendKind::Host);
        const PdhgResult r = Pdhg(opt).solve(p, *be);
        *count = k;
        return static_cast<int>(r.status);
    } catch (...) {
        return -1;
    }
}

REF_API int ref_scaling(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                int32_t apply_ruiz, int32_t apply_pc, double* rs, double* cs) {
    try {
        std::vector<Idx> rpv(rp, rp + m + 1), civ(ci, ci + rp[m]);
        std::vector<Real> vv(v, v + rp[m]);
        const CsrMatrix A(m, n, std::move(rpv), std::move(civ), std::move(vv));
        ScalingOptions o;
        o.apply_ruiz = apply_ruiz != 0;
        o.apply_pock_chambolle = apply_pc != 0;
        const Scaling s = Scaling::equilibrate(A, o);
        std::memcpy(rs, s.row_scale().data(), sizeof(double) * static_cast<std::size_t>(m));
        std::memcpy(cs, s.col_scale().data(), sizeof(double) * static_cast<std::size_t>(n));
        return 0;
    } catch (...) {
        return -1;
    }
}

// Parse an LP (is_mps == 0) or MPS text; sizes out, arrays via ref_parsed_get.
REF_API int ref_parse(const char* text, int32_t is_mps, int32_t* m, int32_t* n, int32_t* nnz) {
    try {
        g_parsed = is_mps ? read_mps_string(text).problem : read_lp_string(text).problem;
        *m = g_parsed.num_rows();
        *n = g_parsed.num_cols();
        *nnz = g_parsed.num_nonzeros();
        return 0;
    } catch (...) {
        return -1;
    }
}

REF_API int ref_parsed_get(int32_t* rp, int32_t* ci, double* v, double* c, double* clo, double* chi,
                   double* rlo, double* rhi, int32_t* maximize, double* obj_const) {
    const Problem& p = g_parsed;
    const auto& A = p.matrix();
    std::memcpy(rp, A.row_ptr().data(), sizeof(int32_t) * A.row_ptr().size());
    std::memcpy(ci, A.col_idx().data(), sizeof(int32_t) * A.col_idx().size());
    std::memcpy(v, A.values().data(), sizeof(double) * A.values().size());
    std::memcpy(c, p.objective().data(), sizeof(double) * p.objective().size());
    std::memcpy(clo, p.col_lower().data(), sizeof(double) * p.col_lower().size());
    std::memcpy(chi, p.col_upper().data(), sizeof(double) * p.col_upper().size());
    std::memcpy(rlo, p.row_lower().data(), sizeof(double) * p.row_lower().size());
    std::memcpy(rhi, p.row_upper().data(), sizeof(double) * p.row_upper().size());
    *maximize = p.sense() == ObjSense::Maximize ? 1 : 0;
    *obj_const = p.objective_constant();
    return 0;
}

}  // extern "C"
