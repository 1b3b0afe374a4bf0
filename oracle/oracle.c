/*
 * CPU oracle for the PDHG hot path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Plain C99, single thread, FP64.  Every function cites the reference lines
 * it restates.  Nothing here is linked into the product.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int finite_bound(double v) { return v > -INFINITY && v < INFINITY; }

/* std::max(a, b) as the reference uses it: (a < b) ? b : a.  A NaN in b is
 * therefore dropped and a NaN in a is kept -- restated as is. */
static double smax(double a, double b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------------- */
/* backend_host.cpp:133-149 */
void orc_spmv(int32_t m, const int32_t* rp, const int32_t* ci, const double* v,
              double alpha, const double* x, double beta, double* y) {
    for (int32_t r = 0; r < m; ++r) {
        double sum = 0.0;
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) sum += v[k] * x[ci[k]];
        y[r] = (beta == 0.0) ? alpha * sum : alpha * sum + beta * y[r];
    }
}

/* backend_host.cpp:151-172 */
void orc_spmv_t(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                double alpha, const double* x, double beta, double* y) {
    if (beta == 0.0) {
        for (int32_t j = 0; j < n; ++j) y[j] = 0.0;
    } else {
        for (int32_t j = 0; j < n; ++j) y[j] *= beta;
    }
    for (int32_t r = 0; r < m; ++r) {
        const double ax = alpha * x[r];
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) y[ci[k]] += v[k] * ax;
    }
}

/* backend_host.cpp:204-215: halve until a leaf of <= 128, serial inside. */
double orc_dot(int64_t n, const double* x, const double* y) {
    if (n <= 0) return 0.0;
    if (n <= 128) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
        return s;
    }
    const int64_t h = n / 2;
    return orc_dot(h, x, y) + orc_dot(n - h, x + h, y + h);
}

static double orc_nrm2(int64_t n, const double* x) { return sqrt(orc_dot(n, x, x)); }

/* ---------------------------------------------------------------------- */
/* scaling.cpp:11 */
static double safe_sqrt(double v) { return v > 0.0 ? sqrt(v) : 1.0; }

/* scaling.cpp:19-104 */
void orc_scaling(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                 int32_t ruiz_iters, double ruiz_tol, int32_t apply_ruiz,
                 int32_t apply_pc, double pc_alpha, double* rs, double* cs) {
    for (int32_t i = 0; i < m; ++i) rs[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) cs[j] = 1.0;
    double* rmax = (double*)calloc((size_t)m + 1, sizeof(double));
    double* cmax = (double*)calloc((size_t)n + 1, sizeof(double));
    if (apply_ruiz) {
        for (int32_t sweep = 0; sweep < ruiz_iters; ++sweep) {
            memset(rmax, 0, sizeof(double) * (size_t)m);
            memset(cmax, 0, sizeof(double) * (size_t)n);
            for (int32_t r = 0; r < m; ++r) {
                for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
                    const int32_t c = ci[k];
                    const double sc = fabs(v[k]) * rs[r] * cs[c];
                    rmax[r] = smax(rmax[r], sc);
                    cmax[c] = smax(cmax[c], sc);
                }
            }
            double worst = 0.0;
            for (int32_t r = 0; r < m; ++r)
                if (rmax[r] > 0.0) {
                    rs[r] /= safe_sqrt(rmax[r]);
                    worst = smax(worst, fabs(rmax[r] - 1.0));
                }
            for (int32_t c = 0; c < n; ++c)
                if (cmax[c] > 0.0) {
                    cs[c] /= safe_sqrt(cmax[c]);
                    worst = smax(worst, fabs(cmax[c] - 1.0));
                }
            if (worst < ruiz_tol) break;
        }
    }
    if (apply_pc) {
        /* Sums reuse the max buffers. */
        memset(rmax, 0, sizeof(double) * (size_t)m);
        memset(cmax, 0, sizeof(double) * (size_t)n);
        for (int32_t r = 0; r < m; ++r) {
            for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
                const int32_t c = ci[k];
                const double mg = fabs(v[k]) * rs[r] * cs[c];
                if (mg <= 0.0) continue;
                rmax[r] += pow(mg, 2.0 - pc_alpha);
                cmax[c] += pow(mg, pc_alpha);
            }
        }
        for (int32_t r = 0; r < m; ++r)
            if (rmax[r] > 0.0) rs[r] /= safe_sqrt(rmax[r]);
        for (int32_t c = 0; c < n; ++c)
            if (cmax[c] > 0.0) cs[c] /= safe_sqrt(cmax[c]);
    }
    free(rmax);
    free(cmax);
}

/* ---------------------------------------------------------------------- */
/* pdhg.cpp:48-79 */
double orc_operator_norm(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                         const double* v, int32_t iterations) {
    double* u = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    double* vv = (double*)malloc(sizeof(double) * ((size_t)m + 1));
    double* w = (double*)calloc((size_t)m + 1, sizeof(double));
    for (int32_t j = 0; j < n; ++j) u[j] = 1.0;
    for (int32_t i = 0; i < m; ++i) vv[i] = 1.0;
    double est = 0.0;
    for (int32_t it = 0; it < iterations; ++it) {
        orc_spmv(m, rp, ci, v, -1.0, u, 0.0, w);
        for (int32_t i = 0; i < m; ++i) w[i] += 1.0 * vv[i];
        const double wn = orc_nrm2(m, w);
        if (wn < 1e-300) break;
        for (int32_t i = 0; i < m; ++i) w[i] *= 1.0 / wn;
        orc_spmv_t(m, n, rp, ci, v, -1.0, w, 0.0, u);
        memcpy(vv, w, sizeof(double) * (size_t)m);
        const double un = orc_nrm2(n, u), vn = orc_nrm2(m, vv);
        est = sqrt(un * un + vn * vn);
        if (est < 1e-300) break;
        for (int32_t j = 0; j < n; ++j) u[j] *= 1.0 / est;
        for (int32_t i = 0; i < m; ++i) vv[i] *= 1.0 / est;
    }
    free(u);
    free(vv);
    free(w);
    return est;
}

/* ---------------------------------------------------------------------- */
/* Row-dual and column reduced-cost classification shared by the two residual
 * functions (pdhg.cpp:106-134 and 166-197). */
static double classify(double val, double lo, double hi, double point, double tol) {
    const int at_lo = finite_bound(lo) && fabs(point - lo) < tol;
    const int at_hi = finite_bound(hi) && fabs(point - hi) < tol;
    if (!at_lo && !at_hi) return fabs(val);
    if (at_lo && !at_hi) return smax(0.0, -val);
    if (at_hi && !at_lo) return smax(0.0, val);
    return 0.0;
}

/* pdhg.cpp:87-136 */
double orc_kkt_residual(const orc_lp* s, const double* c, const double* x, const double* y,
                        double bound_tol) {
    double* act = (double*)calloc((size_t)s->m + 1, sizeof(double));
    double* aty = (double*)calloc((size_t)s->n + 1, sizeof(double));
    orc_spmv(s->m, s->row_ptr, s->col_idx, s->val, 1.0, x, 0.0, act);
    double pr = 0.0, dr = 0.0;
    for (int32_t i = 0; i < s->m; ++i) {
        if (finite_bound(s->row_lo[i])) pr = smax(pr, s->row_lo[i] - act[i]);
        if (finite_bound(s->row_hi[i])) pr = smax(pr, act[i] - s->row_hi[i]);
        const int at_lo = finite_bound(s->row_lo[i]) && fabs(act[i] - s->row_lo[i]) < bound_tol;
        const int at_hi = finite_bound(s->row_hi[i]) && fabs(act[i] - s->row_hi[i]) < bound_tol;
        if (!at_lo && !at_hi) dr = smax(dr, fabs(y[i]));
        else if (at_lo && !at_hi) dr = smax(dr, smax(0.0, -y[i]));
        else if (at_hi && !at_lo) dr = smax(dr, smax(0.0, y[i]));
    }
    orc_spmv_t(s->m, s->n, s->row_ptr, s->col_idx, s->val, 1.0, y, 0.0, aty);
    for (int32_t j = 0; j < s->n; ++j) {
        const double r = c[j] - aty[j];
        dr = smax(dr, classify(r, s->col_lo[j], s->col_hi[j], x[j], bound_tol));
    }
    free(act);
    free(aty);
    return smax(pr, dr);
}

/* pdhg.cpp:145-199 */
double orc_original_space_residual(const orc_lp* p, double obj_sign, const double* rs,
                                   const double* cs, const double* xs, const double* ys,
                                   double bound_tol) {
    double* x = (double*)malloc(sizeof(double) * ((size_t)p->n + 1));
    double* y = (double*)malloc(sizeof(double) * ((size_t)p->m + 1));
    double* act = (double*)calloc((size_t)p->m + 1, sizeof(double));
    double* aty = (double*)calloc((size_t)p->n + 1, sizeof(double));
    for (int32_t j = 0; j < p->n; ++j) x[j] = xs[j] * cs[j];
    for (int32_t i = 0; i < p->m; ++i) {
        y[i] = ys[i] * obj_sign;
        y[i] *= rs[i];
    }
    orc_spmv(p->m, p->row_ptr, p->col_idx, p->val, 1.0, x, 0.0, act);
    double pr = 0.0, dr = 0.0;
    for (int32_t i = 0; i < p->m; ++i) {
        if (finite_bound(p->row_lo[i])) pr = smax(pr, p->row_lo[i] - act[i]);
        if (finite_bound(p->row_hi[i])) pr = smax(pr, act[i] - p->row_hi[i]);
        const double yi = obj_sign * y[i];
        const int at_lo = finite_bound(p->row_lo[i]) && fabs(act[i] - p->row_lo[i]) < bound_tol;
        const int at_hi = finite_bound(p->row_hi[i]) && fabs(act[i] - p->row_hi[i]) < bound_tol;
        if (!at_lo && !at_hi) dr = smax(dr, fabs(yi));
        else if (at_lo && !at_hi) dr = smax(dr, smax(0.0, -yi));
        else if (at_hi && !at_lo) dr = smax(dr, smax(0.0, yi));
    }
    orc_spmv_t(p->m, p->n, p->row_ptr, p->col_idx, p->val, 1.0, y, 0.0, aty);
    for (int32_t j = 0; j < p->n; ++j) {
        const double r = obj_sign * (p->c[j] - aty[j]);
        dr = smax(dr, classify(r, p->col_lo[j], p->col_hi[j], x[j], bound_tol));
    }
    free(x);
    free(y);
    free(act);
    free(aty);
    return smax(pr, dr);
}

/* ---------------------------------------------------------------------- */
/* Extension (no reference counterpart): relative KKT, see oracle.h. */
void orc_relative_kkt(const orc_lp* p, const double* x, const double* y, double out[5]) {
    const double sg = p->maximize ? -1.0 : 1.0;
    double* act = (double*)calloc((size_t)p->m + 1, sizeof(double));
    double* aty = (double*)calloc((size_t)p->n + 1, sizeof(double));
    orc_spmv(p->m, p->row_ptr, p->col_idx, p->val, 1.0, x, 0.0, act);
    orc_spmv_t(p->m, p->n, p->row_ptr, p->col_idx, p->val, 1.0, y, 0.0, aty);
    double pviol2 = 0.0, bnorm2 = 0.0, dres2 = 0.0, cnorm2 = 0.0;
    double pobj = 0.0, dobj = 0.0;
    for (int32_t i = 0; i < p->m; ++i) {
        const double lo = p->row_lo[i], hi = p->row_hi[i];
        double v = 0.0;
        if (finite_bound(lo) && act[i] < lo) v = lo - act[i];
        if (finite_bound(hi) && act[i] > hi) v = act[i] - hi;
        pviol2 += v * v;
        if (finite_bound(lo)) bnorm2 += lo * lo;
        if (finite_bound(hi) && !(finite_bound(lo) && lo == hi)) bnorm2 += hi * hi;
        /* internal-sense row dual: y_int >= 0 needs a finite lower bound,
         * y_int <= 0 a finite upper bound */
        const double yi = sg * y[i];
        double yp = yi;
        if (!finite_bound(lo) && yp > 0.0) yp = 0.0;
        if (!finite_bound(hi) && yp < 0.0) yp = 0.0;
        dres2 += (yi - yp) * (yi - yp);
        if (yp > 0.0) dobj += yp * lo;
        else if (yp < 0.0) dobj += yp * hi;
    }
    for (int32_t j = 0; j < p->n; ++j) {
        const double ci = sg * p->c[j];
        const double rc = ci - sg * aty[j];
        const double lo = p->col_lo[j], hi = p->col_hi[j];
        double lam = rc;
        if (!finite_bound(lo) && lam > 0.0) lam = 0.0;
        if (!finite_bound(hi) && lam < 0.0) lam = 0.0;
        dres2 += (rc - lam) * (rc - lam);
        cnorm2 += ci * ci;
        if (lam > 0.0) dobj += lam * lo;
        else if (lam < 0.0) dobj += lam * hi;
        pobj += ci * x[j];
    }
    free(act);
    free(aty);
    /* back to the problem's own sense, with the constant on both sides */
    pobj = sg * pobj + p->obj_const;
    dobj = sg * dobj + p->obj_const;
    out[0] = sqrt(pviol2) / (1.0 + sqrt(bnorm2));
    out[1] = sqrt(dres2) / (1.0 + sqrt(cnorm2));
    out[2] = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    out[3] = pobj;
    out[4] = dobj;
}

/* ---------------------------------------------------------------------- */
/* duals.cpp:81-113 */
double orc_cs_violation(const orc_lp* p, const double* x, const double* yd, const double* rc,
                        double tol) {
    double worst = 0.0;
    if (yd != NULL && p->m > 0) {
        double* act = (double*)calloc((size_t)p->m + 1, sizeof(double));
        orc_spmv(p->m, p->row_ptr, p->col_idx, p->val, 1.0, x, 0.0, act);
        for (int32_t i = 0; i < p->m; ++i) {
            const int at_lo = finite_bound(p->row_lo[i]) && fabs(act[i] - p->row_lo[i]) <= tol;
            const int at_hi = finite_bound(p->row_hi[i]) && fabs(act[i] - p->row_hi[i]) <= tol;
            if (!at_lo && !at_hi) worst = smax(worst, fabs(yd[i]));
        }
        free(act);
    }
    if (rc != NULL) {
        for (int32_t j = 0; j < p->n; ++j) {
            const int at_lo = finite_bound(p->col_lo[j]) && fabs(x[j] - p->col_lo[j]) <= tol;
            const int at_hi = finite_bound(p->col_hi[j]) && fabs(x[j] - p->col_hi[j]) <= tol;
            if (!at_lo && !at_hi) worst = smax(worst, fabs(rc[j]));
        }
    }
    return worst;
}

/* duals.cpp:105-121: range of mult*[lo,hi]; Minkowski accumulation. */
static void term_range(double mult, double lo, double hi, double* tl, double* th) {
    if (mult == 0.0) { *tl = 0.0; *th = 0.0; return; }
    if (mult > 0.0) { *tl = mult * lo; *th = mult * hi; return; }
    *tl = mult * hi;
    *th = mult * lo;
}

/* duals.cpp:125-168 */
int32_t orc_farkas_certifies(const orc_lp* p, const double* y, double slack) {
    double rlo = 0.0, rhi = 0.0, clo = 0.0, chi = 0.0, tl, th;
    for (int32_t i = 0; i < p->m; ++i) {
        term_range(y[i], p->row_lo[i], p->row_hi[i], &tl, &th);
        rlo += tl;
        rhi += th;
    }
    double* d = (double*)calloc((size_t)p->n + 1, sizeof(double));
    for (int32_t r = 0; r < p->m; ++r) {
        if (y[r] == 0.0) continue;
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k)
            d[p->col_idx[k]] += y[r] * p->val[k];
    }
    for (int32_t j = 0; j < p->n; ++j) {
        term_range(d[j], p->col_lo[j], p->col_hi[j], &tl, &th);
        clo += tl;
        chi += th;
    }
    free(d);
    if (isnan(rlo) || isnan(rhi) || isnan(clo) || isnan(chi)) return 0;
    return (rhi < clo - slack || chi < rlo - slack) ? 1 : 0;
}

/* A primal ray d of the LP (no reference counterpart; the definition of an
 * improving direction of unboundedness): with d scaled to max-norm 1,
 * c'd < -tol in minimize sense, d_j >= -tol where col_lo is finite, d_j <= tol
 * where col_hi is finite, and (Ad)_i >= -tol / <= tol where row_lo / row_hi
 * is finite.  Returns 0 for a zero or non-finite d. */
int32_t orc_primal_ray_certifies(const orc_lp* p, const double* d, double tol) {
    double s = 0.0, obj = 0.0;
    for (int32_t j = 0; j < p->n; ++j) {
        if (!isfinite(d[j])) return 0;
        if (fabs(d[j]) > s) s = fabs(d[j]);
    }
    if (p->n == 0 || s <= 0.0) return 0;
    const double sg = p->maximize ? -1.0 : 1.0;
    for (int32_t j = 0; j < p->n; ++j) {
        const double v = d[j] / s;
        obj += sg * p->c[j] * v;
        if (finite_bound(p->col_lo[j]) && v < -tol) return 0;
        if (finite_bound(p->col_hi[j]) && v > tol) return 0;
    }
    if (!(obj < -tol)) return 0;
    for (int32_t r = 0; r < p->m; ++r) {
        double a = 0.0;
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k) a += p->val[k] * (d[p->col_idx[k]] / s);
        if (finite_bound(p->row_lo[r]) && a < -tol) return 0;
        if (finite_bound(p->row_hi[r]) && a > tol) return 0;
    }
    return 1;
}

/* ---------------------------------------------------------------------- */
void orc_default_options(orc_options* o) {
    memset(o, 0, sizeof(*o));
    o->tolerance = 1e-4;
    o->max_iterations = 1000000;
    o->restart_check_period = 40;
    o->restart_sufficient_decrease = 0.5;
    o->step_size_safety = 0.9;
    o->power_iterations = 30;
    o->termination = 0;
    o->use_primal_weight = 0;
    o->primal_weight = 1.0;
    o->primal_weight_smoothing = 0.5;
    o->ruiz_iterations = 8;
    o->apply_ruiz = 1;
    o->apply_pc = 1;
    o->warm_y = NULL;
    o->restart_scheme = 0;
    o->beta_sufficient = 0.2;
    o->beta_necessary = 0.8;
    o->beta_artificial = 0.36;
}

/* Relative KKT (max of r_p, r_d, gap) of a scaled internal-sense point. */
static double rel_worst_scaled(const orc_lp* p, double sign, const double* rs, const double* cs,
                               const double* xs, const double* ys, double kk[5]) {
    double* ux = (double*)malloc(sizeof(double) * ((size_t)p->n + 1));
    double* uy = (double*)malloc(sizeof(double) * ((size_t)p->m + 1));
    for (int32_t j = 0; j < p->n; ++j) ux[j] = xs[j] * cs[j];
    for (int32_t i = 0; i < p->m; ++i) { uy[i] = ys[i] * sign; uy[i] *= rs[i]; }
    orc_relative_kkt(p, ux, uy, kk);
    free(ux);
    free(uy);
    double w = kk[0];
    if (kk[1] > w) w = kk[1];
    if (kk[2] > w) w = kk[2];
    return w;
}

static void clamp_vec(int32_t n, const double* lo, const double* hi, double* x) {
    /* backend_host.cpp:174-182 */
    for (int32_t i = 0; i < n; ++i) {
        if (x[i] < lo[i]) x[i] = lo[i];
        if (x[i] > hi[i]) x[i] = hi[i];
    }
}

/* pdhg.cpp:36-40 */
static double inf_over_box(double coeff, double lo, double hi) {
    if (coeff == 0.0) return 0.0;
    if (coeff > 0.0) return finite_bound(lo) ? coeff * lo : -INFINITY;
    return finite_bound(hi) ? coeff * hi : -INFINITY;
}

/* pdhg.cpp:218-502, plus the additive extensions declared in oracle.h. */
void orc_pdhg_solve(const orc_lp* p, const orc_options* o, double* xo, double* yo, double* rco,
                    orc_result* res, double* trace, int64_t trace_cap, int64_t* trace_len) {
    const int32_t m = p->m, n = p->n;
    const size_t mb = sizeof(double) * ((size_t)m + 1), nb = sizeof(double) * ((size_t)n + 1);
    memset(res, 0, sizeof(*res));
    res->status = ORC_NOT_SOLVED;
    if (trace_len) *trace_len = 0;
    const clock_t t0 = clock();

    /* pdhg.cpp:229-237 crossed bounds */
    for (int32_t i = 0; i < m; ++i)
        if (p->row_lo[i] > p->row_hi[i]) { res->status = ORC_NUMERICAL_FAILURE; return; }
    for (int32_t j = 0; j < n; ++j)
        if (p->col_lo[j] > p->col_hi[j]) { res->status = ORC_NUMERICAL_FAILURE; return; }

    /* pdhg.cpp:240, scaling.cpp:155-205 */
    double* rs = (double*)malloc(mb);
    double* cs = (double*)malloc(nb);
    orc_scaling(m, n, p->row_ptr, p->col_idx, p->val, o->ruiz_iterations, 1e-2, o->apply_ruiz,
                o->apply_pc, 1.0, rs, cs);
    int32_t* srp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)m + 1));
    int32_t* sci = (int32_t*)malloc(sizeof(int32_t) * ((size_t)p->nnz + 1));
    double* sv = (double*)malloc(sizeof(double) * ((size_t)p->nnz + 1));
    /* Scaling::apply (scaling.cpp:110-133) then from_triplets' zero drop
     * (sparse.cpp:238). */
    int32_t w = 0;
    srp[0] = 0;
    for (int32_t r = 0; r < m; ++r) {
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k) {
            const double val = p->val[k] * rs[r] * cs[p->col_idx[k]];
            if (fabs(val) <= 1e-14) continue;
            sci[w] = p->col_idx[k];
            sv[w] = val;
            ++w;
        }
        srp[r + 1] = w;
    }
    double* sc = (double*)malloc(nb);
    double* sclo = (double*)malloc(nb);
    double* schi = (double*)malloc(nb);
    double* srlo = (double*)malloc(mb);
    double* srhi = (double*)malloc(mb);
    const double obj_sign = p->maximize ? -1.0 : 1.0;
    for (int32_t j = 0; j < n; ++j) {
        const double csj = cs[j];
        sc[j] = obj_sign * (p->c[j] * csj);
        sclo[j] = finite_bound(p->col_lo[j]) ? p->col_lo[j] / csj : p->col_lo[j];
        schi[j] = finite_bound(p->col_hi[j]) ? p->col_hi[j] / csj : p->col_hi[j];
    }
    for (int32_t i = 0; i < m; ++i) {
        srlo[i] = finite_bound(p->row_lo[i]) ? p->row_lo[i] * rs[i] : p->row_lo[i];
        srhi[i] = finite_bound(p->row_hi[i]) ? p->row_hi[i] * rs[i] : p->row_hi[i];
    }
    orc_lp S = *p;
    S.nnz = w;
    S.row_ptr = srp;
    S.col_idx = sci;
    S.val = sv;
    S.c = sc;
    S.col_lo = sclo;
    S.col_hi = schi;
    S.row_lo = srlo;
    S.row_hi = srhi;

    /* pdhg.cpp:261-265 */
    const double opn = orc_operator_norm(m, n, srp, sci, sv, o->power_iterations);
    res->operator_norm_estimate = opn;
    const double eta = (opn > 1e-9) ? o->step_size_safety / opn : o->step_size_safety;
    double omega = o->use_primal_weight ? o->primal_weight : 1.0;
    double tau = eta / omega, sigma = eta * omega;

    /* pdhg.cpp:268-277 */
    double* x = (double*)calloc((size_t)n + 1, sizeof(double));
    double* s = (double*)calloc((size_t)m + 1, sizeof(double));
    double* y = (double*)calloc((size_t)m + 1, sizeof(double));
    clamp_vec(n, sclo, schi, x);
    clamp_vec(m, srlo, srhi, s);
    if (o->warm_y != NULL)
        for (int32_t i = 0; i < m; ++i) y[i] = obj_sign * o->warm_y[i] / rs[i];
    double* xn = (double*)calloc((size_t)n + 1, sizeof(double));
    double* sn = (double*)calloc((size_t)m + 1, sizeof(double));
    double* xb = (double*)calloc((size_t)n + 1, sizeof(double));
    double* sbar = (double*)calloc((size_t)m + 1, sizeof(double));
    double* grad = (double*)calloc((size_t)n + 1, sizeof(double));
    double* kb = (double*)calloc((size_t)m + 1, sizeof(double));
    double* xsum = (double*)calloc((size_t)n + 1, sizeof(double));
    double* ysum = (double*)calloc((size_t)m + 1, sizeof(double));
    double* xavg = (double*)malloc(nb);
    double* yavg = (double*)malloc(mb);
    /* primal-weight bookkeeping: the last restart point */
    double* xlast = (double*)malloc(nb);
    double* ylast = (double*)malloc(mb);
    memcpy(xlast, x, sizeof(double) * (size_t)n);
    memcpy(ylast, y, sizeof(double) * (size_t)m);
    int64_t since = 0;
    double res_at_restart = INFINITY;
    double prev_cand = INFINITY;
    const double* bx = NULL;
    const double* by = NULL;
    int converged = 0;
    int64_t ntrace = 0;
    const double otol = o->tolerance > 1e-6 ? o->tolerance : 1e-6;

    for (res->iterations = 0; res->iterations < o->max_iterations; ++res->iterations) {
        /* pdhg.cpp:296-300 */
        memcpy(grad, sc, sizeof(double) * (size_t)n);
        orc_spmv_t(m, n, srp, sci, sv, -1.0, y, 1.0, grad);
        for (int32_t j = 0; j < n; ++j) xn[j] = x[j] + (-tau) * grad[j];
        clamp_vec(n, sclo, schi, xn);
        /* pdhg.cpp:303-305 */
        for (int32_t i = 0; i < m; ++i) sn[i] = s[i] + (-tau) * y[i];
        clamp_vec(m, srlo, srhi, sn);
        /* pdhg.cpp:308-313 */
        for (int32_t j = 0; j < n; ++j) xb[j] = xn[j] * 2.0 + (-1.0) * x[j];
        for (int32_t i = 0; i < m; ++i) sbar[i] = sn[i] * 2.0 + (-1.0) * s[i];
        /* pdhg.cpp:316-318 */
        orc_spmv(m, srp, sci, sv, -1.0, xb, 0.0, kb);
        for (int32_t i = 0; i < m; ++i) kb[i] += 1.0 * sbar[i];
        for (int32_t i = 0; i < m; ++i) y[i] += sigma * kb[i];
        /* pdhg.cpp:320-331 */
        memcpy(x, xn, sizeof(double) * (size_t)n);
        memcpy(s, sn, sizeof(double) * (size_t)m);
        if (since == 0) {
            memcpy(xsum, x, sizeof(double) * (size_t)n);
            memcpy(ysum, y, sizeof(double) * (size_t)m);
        } else {
            for (int32_t j = 0; j < n; ++j) xsum[j] += 1.0 * x[j];
            for (int32_t i = 0; i < m; ++i) ysum[i] += 1.0 * y[i];
        }
        ++since;
        if ((res->iterations + 1) % o->restart_check_period != 0) continue;

        /* pdhg.cpp:336-352 */
        const double cur = orc_kkt_residual(&S, sc, x, y, o->tolerance);
        const double inv = 1.0 / (double)since;
        for (int32_t j = 0; j < n; ++j) xavg[j] = xsum[j] * inv;
        for (int32_t i = 0; i < m; ++i) yavg[i] = ysum[i] * inv;
        const double avg = orc_kkt_residual(&S, sc, xavg, yavg, o->tolerance);
        if (!isfinite(cur) || !isfinite(avg)) { res->status = ORC_NUMERICAL_FAILURE; goto done; }
        /* pdhg.cpp:354-371 (reference); adaptive: the candidate with the
         * smaller relative KKT */
        double kkc[5], kka[5];
        double rel_c = 0.0, rel_a = 0.0;
        if (o->restart_scheme == 1 || o->termination == 1) {
            rel_c = rel_worst_scaled(p, obj_sign, rs, cs, x, y, kkc);
            rel_a = rel_worst_scaled(p, obj_sign, rs, cs, xavg, yavg, kka);
        }
        const int avg_better = o->restart_scheme == 1 ? (rel_a <= rel_c) : (avg <= cur);
        const double* cx = avg_better ? xavg : x;
        const double* cy = avg_better ? yavg : y;
        const double ores = orc_original_space_residual(p, obj_sign, rs, cs, cx, cy, otol);
        if (!isfinite(ores)) { res->status = ORC_NUMERICAL_FAILURE; goto done; }
        if (trace != NULL && ntrace < trace_cap) trace[ntrace] = ores;
        ++ntrace;
        /* pdhg.cpp:383-389, or the relative-KKT extension */
        const double rel_cand = avg_better ? rel_a : rel_c;
        const int stop = (o->termination == 0) ? (ores < o->tolerance) : (rel_cand <= o->tolerance);
        if (stop) {
            if (avg_better) {
                memcpy(x, xavg, sizeof(double) * (size_t)n);
                memcpy(y, yavg, sizeof(double) * (size_t)m);
            }
            bx = x;
            by = y;
            converged = 1;
            ++res->iterations;
            break;
        }
        int do_restart = 0;
        if (o->restart_scheme == 0) {
            /* pdhg.cpp:391-393 */
            if (!isfinite(res_at_restart)) res_at_restart = cur;
            do_restart = avg <= o->restart_sufficient_decrease * res_at_restart;
        } else {
            if (!isfinite(res_at_restart)) res_at_restart = rel_c;
            if (rel_cand <= o->beta_sufficient * res_at_restart) do_restart = 1;
            else if (rel_cand <= o->beta_necessary * res_at_restart && rel_cand > prev_cand) do_restart = 1;
            else if ((double)since >= o->beta_artificial * (double)(res->iterations + 1)) do_restart = 1;
            prev_cand = rel_cand;
        }
        if (do_restart) {
            if (o->restart_scheme == 0 || avg_better) {
                /* pdhg.cpp:394-400: to the average, s from its activity */
                memcpy(x, xavg, sizeof(double) * (size_t)n);
                memcpy(y, yavg, sizeof(double) * (size_t)m);
                clamp_vec(n, sclo, schi, x);
                orc_spmv(m, srp, sci, sv, 1.0, xavg, 0.0, s);
                clamp_vec(m, srlo, srhi, s);
            }
            if (o->use_primal_weight) {
                /* PDLP primal-weight update (no reference counterpart); at
                 * omega == 1 with use_primal_weight == 0 it never runs. */
                double dx2 = 0.0, dy2 = 0.0;
                for (int32_t j = 0; j < n; ++j) dx2 += (x[j] - xlast[j]) * (x[j] - xlast[j]);
                for (int32_t i = 0; i < m; ++i) dy2 += (y[i] - ylast[i]) * (y[i] - ylast[i]);
                const double dx = sqrt(dx2), dy = sqrt(dy2);
                if (dx > 1e-10 && dy > 1e-10) {
                    const double th = o->primal_weight_smoothing;
                    omega = exp(th * log(dy / dx) + (1.0 - th) * log(omega));
                    tau = eta / omega;
                    sigma = eta * omega;
                }
                memcpy(xlast, x, sizeof(double) * (size_t)n);
                memcpy(ylast, y, sizeof(double) * (size_t)m);
            }
            res_at_restart = o->restart_scheme == 0 ? avg : rel_cand;
            prev_cand = INFINITY;
            since = 0;
            ++res->restarts;
        }
    }
done:
    if (trace_len) *trace_len = ntrace;
    if (res->status == ORC_NUMERICAL_FAILURE) {
        res->final_primal_weight = omega;
    } else {
        /* pdhg.cpp:420-425 */
        if (bx == NULL) { bx = x; by = y; }
        if (converged) res->status = ORC_OPTIMAL;
        else if (res->status == ORC_NOT_SOLVED) res->status = ORC_ITERATION_LIMIT;
        /* pdhg.cpp:205-214, 428-432 */
        for (int32_t j = 0; j < n; ++j) xo[j] = bx[j] * cs[j];
        for (int32_t i = 0; i < m; ++i) { yo[i] = by[i] * obj_sign; yo[i] *= rs[i]; }
        double obj = p->obj_const;
        for (int32_t j = 0; j < n; ++j) obj += p->c[j] * xo[j];
        res->objective = obj;
        /* pdhg.cpp:434-442 */
        double* aty = (double*)calloc((size_t)n + 1, sizeof(double));
        orc_spmv_t(m, n, p->row_ptr, p->col_idx, p->val, 1.0, yo, 0.0, aty);
        for (int32_t j = 0; j < n; ++j) rco[j] = p->c[j] - aty[j];
        free(aty);
        /* pdhg.cpp:444-457 */
        double* act = (double*)calloc((size_t)m + 1, sizeof(double));
        orc_spmv(m, p->row_ptr, p->col_idx, p->val, 1.0, xo, 0.0, act);
        double pinf = 0.0;
        for (int32_t i = 0; i < m; ++i) {
            if (finite_bound(p->row_lo[i])) pinf = smax(pinf, p->row_lo[i] - act[i]);
            if (finite_bound(p->row_hi[i])) pinf = smax(pinf, act[i] - p->row_hi[i]);
        }
        free(act);
        res->primal_infeasibility = pinf;
        /* pdhg.cpp:460-472 */
        res->dual_infeasibility = orc_cs_violation(p, xo, yo, rco, otol);
        /* pdhg.cpp:474-498 */
        double dobj = 0.0;
        int fin = 1;
        for (int32_t j = 0; j < n && fin; ++j) {
            const double t = inf_over_box(rco[j], p->col_lo[j], p->col_hi[j]);
            if (!isfinite(t)) { fin = 0; break; }
            dobj += t;
        }
        for (int32_t i = 0; i < m && fin; ++i) {
            const double t = inf_over_box(yo[i], p->row_lo[i], p->row_hi[i]);
            if (!isfinite(t)) { fin = 0; break; }
            dobj += t;
        }
        res->duality_gap = fin ? fabs(res->objective - dobj) / (1.0 + fabs(res->objective) + fabs(dobj))
                               : INFINITY;
        double kk[5];
        orc_relative_kkt(p, xo, yo, kk);
        res->rel_kkt[0] = kk[0];
        res->rel_kkt[1] = kk[1];
        res->rel_kkt[2] = kk[2];
        res->final_primal_weight = omega;
    }
    (void)t0;
    free(rs); free(cs); free(srp); free(sci); free(sv); free(sc); free(sclo); free(schi);
    free(srlo); free(srhi); free(x); free(s); free(y); free(xn); free(sn); free(xb); free(sbar);
    free(grad); free(kb); free(xsum); free(ysum); free(xavg); free(yavg); free(xlast); free(ylast);
}
