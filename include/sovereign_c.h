/*
 * sovereign_c.h -- C entry points over the C++ engine (lib/libsovereign.so),
 * for FFI callers (Python ctypes in tests/ and bench.py, or any other
 * language).  The C++ API of include/sovereign/ headers is the primary
 * interface; this is a flat mirror of the pieces a caller of
 * `sovereign-cli solve --engine pdhg` needs: build or read a model, solve it
 * with PDHG on a chosen backend, read the result.  Nothing throws across it;
 * functions return sov::Status codes (0 = Ok) and sov_c_last_error() gives
 * the calling thread's message.
 */
#ifndef SOVEREIGN_C_H
#define SOVEREIGN_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sov_lp sov_lp;   /* an immutable sov::Problem */

const char* sov_c_last_error(void);

/* ---- models --------------------------------------------------------- */
sov_lp* sov_lp_from_csr(int32_t m, int32_t n, int32_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                        const double* val, const double* c, const double* col_lo, const double* col_hi,
                        const double* row_lo, const double* row_hi, int32_t maximize, double obj_const);
/* fmt: 0 = CPLEX LP text, 1 = MPS text */
sov_lp* sov_lp_from_text(const char* text, int32_t fmt);
sov_lp* sov_lp_read_file(const char* path);
void sov_lp_destroy(sov_lp* lp);
/* Write the model as MPS (sov::write_mps); reading it back gives the same arrays. */
int sov_lp_write_mps(const sov_lp* lp, const char* path);
int sov_lp_dims(const sov_lp* lp, int32_t* m, int32_t* n, int32_t* nnz);
/* Copy the model out; any pointer may be NULL. */
int sov_lp_arrays(const sov_lp* lp, int32_t* row_ptr, int32_t* col_idx, double* val, double* c, double* col_lo,
                  double* col_hi, double* row_lo, double* row_hi, int32_t* maximize, double* obj_const);

/* ---- seeded generators (include/sovereign/generators.hpp) -------------- */
sov_lp* sov_gen_random_lp(int32_t m, int32_t n, int32_t k, uint64_t seed);
sov_lp* sov_gen_random_lp_zipf(int32_t m, int32_t n, double a, int32_t kmin, int32_t kmax, uint64_t seed);
sov_lp* sov_gen_transportation(int32_t sources, int32_t sinks, uint64_t seed);
int sov_gen_batch_sizes(int32_t count, int32_t lo, int32_t hi, uint64_t seed, int32_t* m_out);

/* ---- host scaling (Scaling::equilibrate, for checks) ------------------ */
int sov_scaling_factors(const sov_lp* lp, int32_t apply_ruiz, int32_t apply_pc, double* row_scale,
                        double* col_scale);

/* ---- PDHG ---------------------------------------------------------------
 * Mirrors sov::PdhgOptions / PdhgResult (include/sovereign/pdhg.hpp). */
enum { SOV_BACKEND_HOST = 0, SOV_BACKEND_HIP = 2 };
enum { SOV_ENGINE_AUTO = 0, SOV_ENGINE_PRIMITIVE = 1, SOV_ENGINE_FUSED = 2 };
enum { SOV_TERM_REFERENCE = 0, SOV_TERM_RELATIVE_KKT = 1 };

typedef struct sov_pdhg_opts {
    double tolerance;
    int64_t max_iterations;
    double time_limit_seconds;
    int64_t restart_check_period;
    double restart_sufficient_decrease;
    double step_size_safety;
    int32_t power_iterations;
    int32_t verbose;
    int32_t engine;        /* SOV_ENGINE_* */
    int32_t termination;   /* SOV_TERM_* */
    int32_t precision;     /* 0 = f64, 1 = f32 */
    int32_t fp64_polish;
    double fp32_switch_factor;
    int32_t fp32_stall_checks;
    int32_t use_primal_weight;
    double initial_primal_weight;
    double primal_weight_smoothing;
    int32_t detect_infeasibility;
    int32_t use_graphs;
    const double* warm_start_dual;   /* NULL or num_rows values (original units) */
    int64_t trace_cap;               /* checkpoint trace: residual and iteration of each check */
    double* trace_res;
    int64_t* trace_it;
    int32_t restart_scheme;          /* 0 = reference rule, 1 = adaptive (PdhgRestartScheme) */
    int32_t autotune;                /* time lane counts and column segments before the loop */
    double restart_beta_sufficient;
    double restart_beta_necessary;
    double restart_beta_artificial;
    double* farkas_out;              /* NULL or num_rows: the Farkas certificate when status is primal_infeasible */
    double* ray_out;                 /* NULL or num_cols: the primal ray when status is dual_infeasible */
    const double* warm_start_primal; /* NULL or num_cols values (original units) */
    int32_t tiles;                   /* LDS-staged step kernels: -1 autotuner decides, 0 never, k >= 1 always (k launches) */
    double autotune_after_seconds;   /* the autotuner runs at the first check after this much loop time (0: before it) */
    int64_t timing_mark_iteration;   /* -1, or the iteration the timed part of the loop starts at (timed_seconds) */
    /* checkpoint callback (PdhgOptions::checkpoint): NULL, or called at every
     * check with the original-space candidate (x: n values, y: m values) and
     * its original-space residual; the arrays are valid during the call */
    void (*checkpoint_fn)(void* user, int64_t iteration, const double* x, int32_t n, const double* y, int32_t m,
                          double residual);
    void* checkpoint_user;
    int32_t checkpoint_async;        /* PdhgOptions::checkpoint_async (default 1) */
    int32_t fp32_refinement;         /* PdhgOptions::fp32_refinement (default 0) */
    double refinement_start;         /* PdhgOptions::refinement_start */
    double refinement_max_scale;     /* PdhgOptions::refinement_max_scale */
    int64_t batch_stage_bytes;       /* PdhgOptions::batch_stage_bytes (default -1) */
    double batch_tile_share;         /* PdhgOptions::batch_tile_share (default -1) */
    int32_t batch_split;             /* PdhgOptions::batch_split (default 1) */
} sov_pdhg_opts;

typedef struct sov_pdhg_result {
    int32_t status;        /* sov::PdhgStatus */
    int32_t reserved;
    double objective;
    double primal_infeasibility;
    double dual_infeasibility;
    double duality_gap;
    int64_t iterations;
    int64_t restarts;
    double seconds;
    double setup_seconds;
    double loop_seconds;
    double operator_norm_estimate;
    double relative_primal;
    double relative_dual;
    double relative_gap;
    double primal_weight;
    int64_t fp64_polish_iterations;
    int64_t trace_len;
    double timed_seconds;            /* PdhgResult::timed_seconds */
    int64_t timed_iterations;        /* PdhgResult::timed_iterations */
    int64_t refinement_iterations;   /* PdhgResult::refinement_iterations */
    int64_t refinement_anchors;      /* PdhgResult::refinement_anchors */
} sov_pdhg_result;

void sov_pdhg_default_opts(sov_pdhg_opts* o);
/* backend: SOV_BACKEND_*.  x[n], y[m], rc[n] may be NULL.  msg (optional)
 * receives PdhgResult::message.  Returns 0 when the solve ran (whatever its
 * PdhgStatus), else a Status code (e.g. the backend is unavailable). */
int sov_pdhg_solve(const sov_lp* lp, int32_t backend, const sov_pdhg_opts* o, double* x, double* y, double* rc,
                   sov_pdhg_result* res, char* msg, int32_t msg_len);

/* The parametric transportation family of the learned warm start
 * (sov::gen::TransportFamily): a fixed S x D transportation structure, base
 * supplies/demands and costs c0, and a fixed (S*D) x P matrix U.  Instance
 * `index` draws theta ~ N(0, I_P) and demand factors lognormal(sigma = 0.2);
 * its costs are c0 (1 + 0.2 tanh(U theta)).  features (NULL or S + D + P
 * doubles) receives (row right-hand sides, theta), the network's input. */
typedef struct sov_family sov_family;
sov_family* sov_family_make(int32_t sources, int32_t sinks, int32_t params, uint64_t seed);
void sov_family_destroy(sov_family* f);
sov_lp* sov_family_instance(const sov_family* f, uint64_t seed, uint64_t index, double* features);

/* Batched solve on the Hip backend (sov::PdhgBatch; no reference
 * counterpart): count models solved together, each with its own scaling,
 * step sizes, restarts and convergence.  warm_ys (NULL, or per model NULL or
 * num_rows duals) and primal_weights (NULL or count fixed weights) warm-start
 * the models.  results[b] per model; xs[b] / ys[b]
 * (NULL or num_cols / num_rows doubles) receive the points; timing (NULL or
 * 4 doubles) receives total, setup and loop seconds and the iterations the
 * batch ran.  All models must share one objective sense. */
int sov_pdhg_solve_batch(const sov_lp* const* lps, int32_t count, const sov_pdhg_opts* opts,
                         const double* const* warm_ys, const double* primal_weights, sov_pdhg_result* results,
                         double* const* xs, double* const* ys, double* timing);

/* A live batch for re-solving related LPs (sov::PdhgBatchSession; no
 * reference counterpart): created from count models like
 * sov_pdhg_solve_batch, it keeps them on the device (on its own Hip
 * backend); between solves an instance's row bounds, objective or column
 * bounds are overwritten in place, or its next start is set from a point
 * (original units; x or y may be NULL) or from another instance's last
 * solution.  Each solve starts every instance from its last candidate
 * unless told otherwise.  Functions return sov::Status codes; create
 * returns NULL on failure (sov_c_last_error). */
typedef struct sov_batch_session sov_batch_session;
sov_batch_session* sov_batch_session_create(const sov_lp* const* lps, int32_t count, const sov_pdhg_opts* opts,
                                            const double* const* warm_ys, const double* primal_weights);
void sov_batch_session_destroy(sov_batch_session* s);
int sov_batch_session_set_row_bounds(sov_batch_session* s, int32_t b, const double* lower, const double* upper);
int sov_batch_session_set_objective(sov_batch_session* s, int32_t b, const double* c);
int sov_batch_session_set_col_bounds(sov_batch_session* s, int32_t b, const double* lower, const double* upper);
int sov_batch_session_set_start(sov_batch_session* s, int32_t b, const double* x, const double* y);
int sov_batch_session_warm_start_from(sov_batch_session* s, int32_t b, int32_t source);
/* results[count]; xs / ys / timing as sov_pdhg_solve_batch. */
int sov_batch_session_solve(sov_batch_session* s, sov_pdhg_result* results, double* const* xs, double* const* ys,
                            double* timing);

/* Backend report line (available backends, default, index width). */
int sov_backend_report(char* buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
