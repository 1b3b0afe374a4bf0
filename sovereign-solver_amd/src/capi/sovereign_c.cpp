// Flat C entry points over the C++ engine (API: include/sovereign_c.h).
#include "sovereign_c.h"

#include <cstring>
#include <memory>
#include <string>

#include "sovereign/backend.hpp"
#include "sovereign/generators.hpp"
#include "sovereign/io.hpp"
#include "sovereign/pdhg.hpp"
#include "sovereign/scaling.hpp"

struct sov_lp {
    sov::Problem p;
};

struct sov_family {
    sov::gen::TransportFamily f;
};

namespace {
thread_local std::string g_err;

int set_err(int code, const std::string& m) {
    g_err = m;
    return code;
}

template <class F>
sov_lp* make_lp(F&& f) {
    try {
        auto* lp = new sov_lp{f()};
        return lp;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
}  // namespace

extern "C" {

const char* sov_c_last_error(void) { return g_err.c_str(); }

sov_lp* sov_lp_from_csr(int32_t m, int32_t n, int32_t nnz, const int32_t* rp, const int32_t* ci, const double* v,
                        const double* c, const double* clo, const double* chi, const double* rlo, const double* rhi,
                        int32_t maximize, double obj_const) {
    return make_lp([&] {
        if (m < 0 || n < 0 || nnz < 0 || rp == nullptr) throw sov::Error("sov_lp_from_csr: bad arguments");
        sov::CsrMatrix A(m, n, std::vector<sov::Idx>(rp, rp + m + 1), std::vector<sov::Idx>(ci, ci + nnz),
                         std::vector<double>(v, v + nnz));
        return sov::Problem::from_csr(std::move(A), std::vector<double>(c, c + n), std::vector<double>(clo, clo + n),
                                      std::vector<double>(chi, chi + n), std::vector<double>(rlo, rlo + m),
                                      std::vector<double>(rhi, rhi + m),
                                      maximize ? sov::ObjSense::Maximize : sov::ObjSense::Minimize, obj_const);
    });
}

sov_lp* sov_lp_from_text(const char* text, int32_t fmt) {
    return make_lp([&] {
        if (text == nullptr) throw sov::Error("null text");
        return fmt == 1 ? sov::read_mps_string(text).problem : sov::read_lp_string(text).problem;
    });
}

sov_lp* sov_lp_read_file(const char* path) {
    return make_lp([&] {
        if (path == nullptr) throw sov::Error("null path");
        return sov::read_model(path).problem;
    });
}

void sov_lp_destroy(sov_lp* lp) { delete lp; }

int sov_lp_write_mps(const sov_lp* lp, const char* path) {
    if (lp == nullptr || path == nullptr) return set_err(3, "null argument");
    try {
        sov::write_mps(lp->p, path);
        return 0;
    } catch (const std::exception& e) {
        return set_err(6, e.what());
    }
}

int sov_lp_dims(const sov_lp* lp, int32_t* m, int32_t* n, int32_t* nnz) {
    if (lp == nullptr) return set_err(3, "null model");
    if (m) *m = lp->p.num_rows();
    if (n) *n = lp->p.num_cols();
    if (nnz) *nnz = lp->p.num_nonzeros();
    return 0;
}

int sov_lp_arrays(const sov_lp* lp, int32_t* rp, int32_t* ci, double* v, double* c, double* clo, double* chi,
                  double* rlo, double* rhi, int32_t* maximize, double* obj_const) {
    if (lp == nullptr) return set_err(3, "null model");
    const sov::Problem& p = lp->p;
    const auto& A = p.matrix();
    auto cp = [](auto* dst, auto span) {
        if (dst != nullptr && !span.empty()) std::memcpy(dst, span.data(), span.size_bytes());
    };
    cp(rp, A.row_ptr());
    cp(ci, A.col_idx());
    cp(v, A.values());
    cp(c, p.objective());
    cp(clo, p.col_lower());
    cp(chi, p.col_upper());
    cp(rlo, p.row_lower());
    cp(rhi, p.row_upper());
    if (maximize) *maximize = p.sense() == sov::ObjSense::Maximize ? 1 : 0;
    if (obj_const) *obj_const = p.objective_constant();
    return 0;
}

sov_lp* sov_gen_random_lp(int32_t m, int32_t n, int32_t k, uint64_t seed) {
    return make_lp([&] { return sov::gen::random_lp(m, n, k, seed).to_problem(); });
}

sov_lp* sov_gen_random_lp_zipf(int32_t m, int32_t n, double a, int32_t kmin, int32_t kmax, uint64_t seed) {
    return make_lp([&] { return sov::gen::random_lp_zipf(m, n, a, kmin, kmax, seed).to_problem(); });
}

sov_lp* sov_gen_transportation(int32_t s, int32_t d, uint64_t seed) {
    return make_lp([&] { return sov::gen::transportation(s, d, seed).to_problem(); });
}

sov_family* sov_family_make(int32_t sources, int32_t sinks, int32_t params, uint64_t seed) {
    if (sources <= 0 || sinks <= 0 || params <= 0) {
        g_err = "sov_family_make: bad sizes";
        return nullptr;
    }
    try {
        return new sov_family{sov::gen::TransportFamily::make(sources, sinks, params, seed)};
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void sov_family_destroy(sov_family* f) { delete f; }

sov_lp* sov_family_instance(const sov_family* f, uint64_t seed, uint64_t index, double* features) {
    if (f == nullptr) {
        g_err = "null family";
        return nullptr;
    }
    return make_lp([&] {
        std::vector<double> theta, z;
        f->f.draw(seed, index, theta, z);
        sov::gen::LpArrays a = f->f.instance(theta, z);
        if (features != nullptr) {
            std::copy(a.row_lo.begin(), a.row_lo.end(), features);
            std::copy(theta.begin(), theta.end(), features + a.row_lo.size());
        }
        return a.to_problem();
    });
}

int sov_gen_batch_sizes(int32_t count, int32_t lo, int32_t hi, uint64_t seed, int32_t* out) {
    if (count < 0 || out == nullptr || lo <= 0 || hi < lo) return set_err(3, "bad arguments");
    const auto v = sov::gen::batch_sizes(count, lo, hi, seed);
    std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    return 0;
}

int sov_scaling_factors(const sov_lp* lp, int32_t ruiz, int32_t pc, double* rs, double* cs) {
    if (lp == nullptr || rs == nullptr || cs == nullptr) return set_err(3, "null argument");
    sov::ScalingOptions o;
    o.apply_ruiz = ruiz != 0;
    o.apply_pock_chambolle = pc != 0;
    const sov::Scaling s = sov::Scaling::equilibrate(lp->p.matrix(), o);
    std::memcpy(rs, s.row_scale().data(), s.row_scale().size_bytes());
    std::memcpy(cs, s.col_scale().data(), s.col_scale().size_bytes());
    return 0;
}

void sov_pdhg_default_opts(sov_pdhg_opts* o) {
    if (o == nullptr) return;
    const sov::PdhgOptions d;
    std::memset(o, 0, sizeof(*o));
    o->tolerance = d.tolerance;
    o->max_iterations = d.max_iterations;
    o->time_limit_seconds = d.time_limit_seconds;
    o->restart_check_period = d.restart_check_period;
    o->restart_sufficient_decrease = d.restart_sufficient_decrease;
    o->step_size_safety = d.step_size_safety;
    o->power_iterations = d.power_iterations;
    o->verbose = 0;
    o->engine = SOV_ENGINE_AUTO;
    o->termination = SOV_TERM_REFERENCE;
    o->precision = 0;
    o->fp64_polish = d.fp64_polish ? 1 : 0;
    o->fp32_switch_factor = d.fp32_switch_factor;
    o->fp32_stall_checks = d.fp32_stall_checks;
    o->fp32_refinement = d.fp32_refinement ? 1 : 0;
    o->refinement_start = d.refinement_start;
    o->refinement_max_scale = d.refinement_max_scale;
    o->batch_stage_bytes = d.batch_stage_bytes;
    o->batch_tile_share = d.batch_tile_share;
    o->batch_split = d.batch_split ? 1 : 0;
    o->use_primal_weight = 0;
    o->initial_primal_weight = d.initial_primal_weight;
    o->primal_weight_smoothing = d.primal_weight_smoothing;
    o->detect_infeasibility = 0;
    o->use_graphs = d.use_graphs ? 1 : 0;
    o->restart_scheme = 0;
    o->autotune = d.autotune ? 1 : 0;
    o->tiles = d.tiles;
    o->autotune_after_seconds = d.autotune_after_seconds;
    o->timing_mark_iteration = d.timing_mark_iteration;
    o->checkpoint_async = d.checkpoint_async ? 1 : 0;
    o->restart_beta_sufficient = d.restart_beta_sufficient;
    o->restart_beta_necessary = d.restart_beta_necessary;
    o->restart_beta_artificial = d.restart_beta_artificial;
}

}  // extern "C"

namespace {

sov::PdhgOptions to_options(const sov_pdhg_opts* o) {
    sov::PdhgOptions p;
    p.tolerance = o->tolerance;
    p.max_iterations = o->max_iterations;
    p.time_limit_seconds = o->time_limit_seconds;
    p.restart_check_period = o->restart_check_period;
    p.restart_sufficient_decrease = o->restart_sufficient_decrease;
    p.step_size_safety = o->step_size_safety;
    p.power_iterations = o->power_iterations;
    p.verbose = o->verbose != 0;
    p.engine = static_cast<sov::PdhgEngine>(o->engine);
    p.termination = static_cast<sov::PdhgTermination>(o->termination);
    p.precision = o->precision == 1 ? sov::Precision::F32 : sov::Precision::F64;
    p.fp64_polish = o->fp64_polish != 0;
    p.fp32_switch_factor = o->fp32_switch_factor;
    p.fp32_stall_checks = o->fp32_stall_checks;
    p.fp32_refinement = o->fp32_refinement != 0;
    p.refinement_start = o->refinement_start;
    p.refinement_max_scale = o->refinement_max_scale;
    p.batch_stage_bytes = o->batch_stage_bytes;
    p.batch_tile_share = o->batch_tile_share;
    p.batch_split = o->batch_split != 0;
    p.use_primal_weight = o->use_primal_weight != 0;
    p.initial_primal_weight = o->initial_primal_weight;
    p.primal_weight_smoothing = o->primal_weight_smoothing;
    p.detect_infeasibility = o->detect_infeasibility != 0;
    p.use_graphs = o->use_graphs != 0;
    p.autotune = o->autotune != 0;
    p.tiles = o->tiles;
    p.autotune_after_seconds = o->autotune_after_seconds;
    p.timing_mark_iteration = o->timing_mark_iteration;
    p.checkpoint_async = o->checkpoint_async != 0;
    p.restart_scheme = o->restart_scheme == 1 ? sov::PdhgRestartScheme::Adaptive : sov::PdhgRestartScheme::Reference;
    p.restart_beta_sufficient = o->restart_beta_sufficient;
    p.restart_beta_necessary = o->restart_beta_necessary;
    p.restart_beta_artificial = o->restart_beta_artificial;
    return p;
}

void fill_result(const sov::PdhgResult& r, sov_pdhg_result* res) {
    std::memset(res, 0, sizeof(*res));
    res->status = static_cast<int32_t>(r.status);
    res->objective = r.objective;
    res->primal_infeasibility = r.primal_infeasibility;
    res->dual_infeasibility = r.dual_infeasibility;
    res->duality_gap = r.duality_gap;
    res->iterations = r.iterations;
    res->restarts = r.restarts;
    res->seconds = r.seconds;
    res->setup_seconds = r.setup_seconds;
    res->loop_seconds = r.loop_seconds;
    res->operator_norm_estimate = r.operator_norm_estimate;
    res->relative_primal = r.relative_primal_residual;
    res->relative_dual = r.relative_dual_residual;
    res->relative_gap = r.relative_gap;
    res->primal_weight = r.primal_weight;
    res->fp64_polish_iterations = r.fp64_polish_iterations;
    res->timed_seconds = r.timed_seconds;
    res->timed_iterations = r.timed_iterations;
    res->refinement_iterations = r.refinement_iterations;
    res->refinement_anchors = r.refinement_anchors;
}

}  // namespace

extern "C" {

int sov_pdhg_solve(const sov_lp* lp, int32_t backend, const sov_pdhg_opts* o, double* x, double* y, double* rc,
                   sov_pdhg_result* res, char* msg, int32_t msg_len) {
    if (lp == nullptr || o == nullptr || res == nullptr) return set_err(3, "null argument");
    try {
        sov::PdhgOptions p = to_options(o);
        if (o->warm_start_dual != nullptr)
            p.warm_start_dual.assign(o->warm_start_dual, o->warm_start_dual + lp->p.num_rows());
        if (o->warm_start_primal != nullptr)
            p.warm_start_primal.assign(o->warm_start_primal, o->warm_start_primal + lp->p.num_cols());
        int64_t nt = 0;
        const bool tracing = o->trace_res != nullptr && o->trace_cap > 0;
        if (tracing || o->checkpoint_fn != nullptr) {
            p.checkpoint = [&](long long it, const std::vector<double>& x, const std::vector<double>& y, double r) {
                if (tracing) {
                    if (nt < o->trace_cap) {
                        o->trace_res[nt] = r;
                        if (o->trace_it) o->trace_it[nt] = it;
                    }
                    ++nt;
                }
                if (o->checkpoint_fn != nullptr)
                    o->checkpoint_fn(o->checkpoint_user, it, x.data(), static_cast<int32_t>(x.size()), y.data(),
                                     static_cast<int32_t>(y.size()), r);
            };
        }
        const sov::BackendKind kind = backend == SOV_BACKEND_HIP ? sov::BackendKind::Hip : sov::BackendKind::Host;
        std::unique_ptr<sov::Backend> be = sov::make_backend(kind);
        const sov::PdhgResult r = sov::Pdhg(p).solve(lp->p, *be);
        fill_result(r, res);
        res->trace_len = nt;
        if (o->farkas_out && !r.certificate.row_multipliers.empty())
            std::memcpy(o->farkas_out, r.certificate.row_multipliers.data(), r.certificate.row_multipliers.size() * sizeof(double));
        if (o->ray_out && !r.ray.direction.empty())
            std::memcpy(o->ray_out, r.ray.direction.data(), r.ray.direction.size() * sizeof(double));
        if (x && !r.primal.empty()) std::memcpy(x, r.primal.data(), r.primal.size() * sizeof(double));
        if (y && !r.dual.empty()) std::memcpy(y, r.dual.data(), r.dual.size() * sizeof(double));
        if (rc && !r.reduced_costs.empty()) std::memcpy(rc, r.reduced_costs.data(), r.reduced_costs.size() * sizeof(double));
        if (msg != nullptr && msg_len > 0) {
            std::strncpy(msg, r.message.c_str(), static_cast<std::size_t>(msg_len - 1));
            msg[msg_len - 1] = '\0';
        }
        return 0;
    } catch (const sov::Error& e) {
        return set_err(static_cast<int>(e.status()), e.what());
    } catch (const std::exception& e) {
        return set_err(6, e.what());
    }
}

int sov_pdhg_solve_batch(const sov_lp* const* lps, int32_t count, const sov_pdhg_opts* o,
                         const double* const* warm_ys, const double* primal_weights, sov_pdhg_result* results,
                         double* const* xs, double* const* ys, double* timing) {
    if (lps == nullptr || o == nullptr || results == nullptr || count < 0) return set_err(3, "null argument");
    try {
        std::vector<const sov::Problem*> probs(static_cast<std::size_t>(count));
        for (int32_t b = 0; b < count; ++b) {
            if (lps[b] == nullptr) return set_err(3, "null model in batch");
            probs[static_cast<std::size_t>(b)] = &lps[b]->p;
        }
        std::vector<std::vector<double>> wy;
        if (warm_ys != nullptr) {
            wy.resize(static_cast<std::size_t>(count));
            for (int32_t b = 0; b < count; ++b)
                if (warm_ys[b] != nullptr)
                    wy[static_cast<std::size_t>(b)].assign(warm_ys[b], warm_ys[b] + lps[b]->p.num_rows());
        }
        std::vector<double> pw;
        if (primal_weights != nullptr) pw.assign(primal_weights, primal_weights + count);
        std::unique_ptr<sov::Backend> be = sov::make_backend(sov::BackendKind::Hip);
        const sov::PdhgBatchResult br = sov::PdhgBatch(to_options(o)).solve(probs, *be, warm_ys ? &wy : nullptr,
                                                                           primal_weights ? &pw : nullptr);
        for (int32_t b = 0; b < count; ++b) {
            const sov::PdhgResult& r = br.results[static_cast<std::size_t>(b)];
            fill_result(r, &results[b]);
            if (xs && xs[b] && !r.primal.empty()) std::memcpy(xs[b], r.primal.data(), r.primal.size() * sizeof(double));
            if (ys && ys[b] && !r.dual.empty()) std::memcpy(ys[b], r.dual.data(), r.dual.size() * sizeof(double));
        }
        if (timing) {
            timing[0] = br.seconds;
            timing[1] = br.setup_seconds;
            timing[2] = br.loop_seconds;
            timing[3] = static_cast<double>(br.iterations);
        }
        if (!br.message.empty()) g_err = br.message;
        return 0;
    } catch (const sov::Error& e) {
        return set_err(static_cast<int>(e.status()), e.what());
    } catch (const std::exception& e) {
        return set_err(6, e.what());
    }
}

}  // extern "C"

struct sov_batch_session {
    std::unique_ptr<sov::Backend> backend;   // declared first: outlives the session
    std::unique_ptr<sov::PdhgBatchSession> session;
};

namespace {

template <class F>
int session_call(sov_batch_session* s, F&& f) {
    if (s == nullptr) return set_err(3, "null session");
    try {
        f(*s->session);
        return 0;
    } catch (const sov::Error& e) {
        return set_err(static_cast<int>(e.status()), e.what());
    } catch (const std::exception& e) {
        return set_err(6, e.what());
    }
}

std::span<const double> span_of(const double* p, std::size_t n) {
    return p == nullptr ? std::span<const double>{} : std::span<const double>(p, n);
}

}  // namespace

extern "C" {

sov_batch_session* sov_batch_session_create(const sov_lp* const* lps, int32_t count, const sov_pdhg_opts* o,
                                            const double* const* warm_ys, const double* primal_weights) {
    if (lps == nullptr || o == nullptr || count < 0) {
        g_err = "null argument";
        return nullptr;
    }
    try {
        std::vector<const sov::Problem*> probs(static_cast<std::size_t>(count));
        for (int32_t b = 0; b < count; ++b) {
            if (lps[b] == nullptr) throw sov::Error(sov::Status::InvalidArgument, "null model in batch");
            probs[static_cast<std::size_t>(b)] = &lps[b]->p;
        }
        std::vector<std::vector<double>> wy;
        if (warm_ys != nullptr) {
            wy.resize(static_cast<std::size_t>(count));
            for (int32_t b = 0; b < count; ++b)
                if (warm_ys[b] != nullptr)
                    wy[static_cast<std::size_t>(b)].assign(warm_ys[b], warm_ys[b] + lps[b]->p.num_rows());
        }
        std::vector<double> pw;
        if (primal_weights != nullptr) pw.assign(primal_weights, primal_weights + count);
        auto* s = new sov_batch_session;
        try {
            s->backend = sov::make_backend(sov::BackendKind::Hip);
            s->session = std::make_unique<sov::PdhgBatchSession>(to_options(o), probs, *s->backend,
                                                                 warm_ys ? &wy : nullptr,
                                                                 primal_weights ? &pw : nullptr);
        } catch (...) {
            delete s;
            throw;
        }
        return s;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void sov_batch_session_destroy(sov_batch_session* s) {
    if (s == nullptr) return;
    s->session.reset();
    delete s;
}

int sov_batch_session_set_row_bounds(sov_batch_session* s, int32_t b, const double* lo, const double* hi) {
    return session_call(s, [&](sov::PdhgBatchSession& x) {
        const auto m = static_cast<std::size_t>(x.problem(static_cast<std::size_t>(b)).num_rows());
        x.set_row_bounds(static_cast<std::size_t>(b), span_of(lo, m), span_of(hi, m));
    });
}

int sov_batch_session_set_objective(sov_batch_session* s, int32_t b, const double* c) {
    return session_call(s, [&](sov::PdhgBatchSession& x) {
        const auto n = static_cast<std::size_t>(x.problem(static_cast<std::size_t>(b)).num_cols());
        x.set_objective(static_cast<std::size_t>(b), span_of(c, n));
    });
}

int sov_batch_session_set_col_bounds(sov_batch_session* s, int32_t b, const double* lo, const double* hi) {
    return session_call(s, [&](sov::PdhgBatchSession& x) {
        const auto n = static_cast<std::size_t>(x.problem(static_cast<std::size_t>(b)).num_cols());
        x.set_column_bounds(static_cast<std::size_t>(b), span_of(lo, n), span_of(hi, n));
    });
}

int sov_batch_session_set_start(sov_batch_session* s, int32_t b, const double* xv, const double* yv) {
    return session_call(s, [&](sov::PdhgBatchSession& x) {
        const sov::Problem& p = x.problem(static_cast<std::size_t>(b));
        x.set_start(static_cast<std::size_t>(b), span_of(xv, static_cast<std::size_t>(p.num_cols())),
                    span_of(yv, static_cast<std::size_t>(p.num_rows())));
    });
}

int sov_batch_session_warm_start_from(sov_batch_session* s, int32_t b, int32_t source) {
    return session_call(s, [&](sov::PdhgBatchSession& x) {
        if (b < 0 || source < 0) throw sov::Error(sov::Status::InvalidArgument, "negative instance");
        x.warm_start_from(static_cast<std::size_t>(b), static_cast<std::size_t>(source));
    });
}

int sov_batch_session_solve(sov_batch_session* s, sov_pdhg_result* results, double* const* xs, double* const* ys,
                            double* timing) {
    if (results == nullptr) return set_err(3, "null results");
    return session_call(s, [&](sov::PdhgBatchSession& x) {
        const sov::PdhgBatchResult br = x.solve();
        for (std::size_t b = 0; b < br.results.size(); ++b) {
            const sov::PdhgResult& r = br.results[b];
            fill_result(r, &results[b]);
            if (xs && xs[b] && !r.primal.empty()) std::memcpy(xs[b], r.primal.data(), r.primal.size() * sizeof(double));
            if (ys && ys[b] && !r.dual.empty()) std::memcpy(ys[b], r.dual.data(), r.dual.size() * sizeof(double));
        }
        if (timing) {
            timing[0] = br.seconds;
            timing[1] = br.setup_seconds;
            timing[2] = br.loop_seconds;
            timing[3] = static_cast<double>(br.iterations);
        }
        if (!br.message.empty()) g_err = br.message;
    });
}

int sov_backend_report(char* buf, int32_t len) {
    if (buf == nullptr || len <= 0) return set_err(3, "null buffer");
    const std::string s = sov::backend_report();
    std::strncpy(buf, s.c_str(), static_cast<std::size_t>(len - 1));
    buf[len - 1] = '\0';
    return 0;
}

}  // extern "C"
