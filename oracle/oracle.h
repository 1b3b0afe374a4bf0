/*
 * CPU oracle for the PDHG hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded, FP64 restatement of the reference host PDHG
 * (src/l1/pdhg.cpp), its scaling front-end (src/l1/scaling.cpp), the host
 * backend primitives it runs on (src/l0/backend_host.cpp), the CSR helpers
 * (src/l0/sparse.cpp) and the dual checkers (src/l1/duals.cpp).  Each function
 * cites the reference lines it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product never links it.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_spmv / orc_spmv_t / orc_dot     hand answers of test_l0_backend.cpp,
 *                                       dense brute force, oracle/_ref vectors
 *   orc_scaling                         oracle/_ref factors (bitwise), the
 *                                       invariants of test_scaling.cpp
 *   orc_operator_norm                   closed form sqrt(smax(A)^2+1) (numpy
 *                                       SVD), oracle/_ref estimate
 *   orc_kkt_residual / orc_original_... hand-computed tiny cases, zero at a
 *                                       known optimum, permutation invariance
 *   orc_pdhg_solve (reference mode)     oracle/_ref iteration/restart counts,
 *                                       objectives and residual trace
 *   orc_relative_kkt                    hand-computed cases, 0 at an optimum
 *   orc_cs_violation / orc_farkas_*     hand cases of test_duals.cpp
 *   adaptive restarts                   same decisions as the engine on the
 *                                       GPU (trajectory parity), textbook
 *                                       optima on the tiny LPs
 *   primal weight / relative-KKT stop   reduces to reference mode at omega=1
 *                                       (bitwise), textbook optima on the
 *                                       tiny LPs
 */
#ifndef SOV_ORACLE_H
#define SOV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A linear program in the reference's canonical form (problem.hpp:5-8):
 *   optimize c'x + obj_const  s.t. row_lo <= A x <= row_hi, col_lo <= x <= col_hi
 * A is CSR with strictly ascending column indices per row.  Infinite bounds
 * are +-INFINITY. */
typedef struct orc_lp {
    int32_t m, n, nnz;
    const int32_t* row_ptr;
    const int32_t* col_idx;
    const double* val;
    const double* c;
    const double* col_lo;
    const double* col_hi;
    const double* row_lo;
    const double* row_hi;
    int32_t maximize;
    double obj_const;
} orc_lp;

/* y := alpha*A*x + beta*y; beta == 0 overwrites (backend_host.cpp:133-149). */
void orc_spmv(int32_t m, const int32_t* rp, const int32_t* ci, const double* v,
              double alpha, const double* x, double beta, double* y);
/* y := alpha*A^T*x + beta*y by row scatter (backend_host.cpp:151-172). */
void orc_spmv_t(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                double alpha, const double* x, double beta, double* y);
/* Pairwise dot product, leaf block 128 (backend_host.cpp:204-215). */
double orc_dot(int64_t n, const double* x, const double* y);

/* Ruiz (<= ruiz_iters sweeps, early stop below ruiz_tol) then Pock-Chambolle
 * with exponent pc_alpha (scaling.cpp:19-104).  Writes row_scale[m],
 * col_scale[n]. */
void orc_scaling(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                 int32_t ruiz_iters, double ruiz_tol, int32_t apply_ruiz,
                 int32_t apply_pc, double pc_alpha, double* row_scale, double* col_scale);

/* Power iteration on K(u,v) = v - A u, start u=1, v=1 (pdhg.cpp:48-79). */
double orc_operator_norm(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                         const double* v, int32_t iterations);

/* Scaled-space max-norm KKT residual (pdhg.cpp:87-136).  c is the internal
 * (always-minimize) objective. */
double orc_kkt_residual(const orc_lp* scaled, const double* c, const double* x,
                        const double* y, double bound_tol);

/* Original-space residual of a scaled-space point (pdhg.cpp:145-199). */
double orc_original_space_residual(const orc_lp* orig, double obj_sign,
                                   const double* row_scale, const double* col_scale,
                                   const double* x_scaled, const double* y_scaled,
                                   double bound_tol);

/* Relative KKT of an ORIGINAL-space point (x, y) (y in the problem's own
 * sense).  out[0]=primal rel, out[1]=dual rel, out[2]=gap rel,
 * out[3]=primal objective (with constant, problem sense), out[4]=dual
 * objective (same).  Definitions:
 *   r_p = ||viol(Ax)||_2 / (1 + ||finite row bounds||_2)
 *   r_d = ||(rc - lambda, y - proj(y))||_2 / (1 + ||c||_2)
 *   gap = |p - d| / (1 + |p| + |d|)
 * with lambda the projection of rc = c - A^T y onto the sign cone the column
 * bounds allow and proj(y) the projection of y onto the sign cone the row
 * bounds allow, both in the internal minimize sense. */
void orc_relative_kkt(const orc_lp* orig, const double* x, const double* y, double out[5]);

/* An improving primal ray of the LP (no reference counterpart; pinned by
 * hand cases): d scaled to max-norm 1 has c'd < -tol (minimize sense), stays
 * in the recession cone of the column box and A d in that of the row box,
 * all within tol. */
int32_t orc_primal_ray_certifies(const orc_lp* lp, const double* d, double tol);

/* DualSolution::complementary_slackness_violation (duals.cpp:81-113). */
double orc_cs_violation(const orc_lp* lp, const double* x, const double* row_duals,
                        const double* reduced_costs, double bound_tol);

/* FarkasCertificate::certifies_infeasibility (duals.cpp:71-79, 125-168). */
int32_t orc_farkas_certifies(const orc_lp* lp, const double* y, double slack);

/* Status codes follow PdhgStatus (pdhg.hpp:74-81) plus two additive ones. */
enum {
    ORC_OPTIMAL = 0,
    ORC_ITERATION_LIMIT = 1,
    ORC_TIME_LIMIT = 2,
    ORC_NUMERICAL_FAILURE = 3,
    ORC_NOT_SOLVED = 4,
    ORC_CANCELLED = 5,
    ORC_PRIMAL_INFEASIBLE = 6,
    ORC_DUAL_INFEASIBLE = 7
};

/* Termination rules: 0 = the reference's original-space max-norm residual
 * (pdhg.cpp:383); 1 = relative KKT (max of r_p, r_d, gap) <= tolerance. */
typedef struct orc_options {
    double tolerance;              /* pdhg.hpp:89, default 1e-4 */
    int64_t max_iterations;        /* pdhg.hpp:96 */
    int64_t restart_check_period;  /* pdhg.hpp:103 */
    double restart_sufficient_decrease; /* pdhg.hpp:107 */
    double step_size_safety;       /* pdhg.hpp:113 */
    int32_t power_iterations;      /* pdhg.hpp:117 */
    int32_t termination;           /* 0 reference, 1 relative KKT */
    int32_t use_primal_weight;     /* 0: tau = sigma (reference); 1: PDLP weight */
    double primal_weight;          /* initial omega (1 = reference) */
    double primal_weight_smoothing;/* theta of the weight update */
    int32_t ruiz_iterations;       /* scaling.hpp:42 */
    int32_t apply_ruiz, apply_pc;  /* scaling.hpp:51-52 */
    const double* warm_y;          /* optional ORIGINAL-space dual start (NULL: y=0) */
    int32_t restart_scheme;        /* 0: reference rule (pdhg.cpp:391-408); 1: adaptive */
    double beta_sufficient;        /* adaptive: restart when K(cand) <= b_s K(last restart) */
    double beta_necessary;         /* ... or <= b_n K(last) and K(cand) grew since the last check */
    double beta_artificial;        /* ... or since-restart >= b_a * total iterations */
} orc_options;

void orc_default_options(orc_options* o);

typedef struct orc_result {
    int32_t status;
    double objective;
    double primal_infeasibility;
    double dual_infeasibility;
    double duality_gap;
    int64_t iterations;
    int64_t restarts;
    double operator_norm_estimate;
    double rel_kkt[3];           /* relative KKT of the reported point */
    double final_primal_weight;
} orc_result;

/* Pdhg::solve on the reference's Host backend (pdhg.cpp:218-502).  x[n],
 * y[m], rc[n] receive the reported ORIGINAL-space point.  If trace_res is not
 * NULL the original-space residual of each check is appended (up to
 * trace_cap entries) and *trace_len receives the count. */
void orc_pdhg_solve(const orc_lp* lp, const orc_options* opt, double* x, double* y,
                    double* rc, orc_result* res, double* trace_res, int64_t trace_cap,
                    int64_t* trace_len);

#ifdef __cplusplus
}
#endif
#endif
