// Fused PDHG kernels: the primal and dual steps (kernels a+b), the KKT /
// restart reductions (kernel c), restart and point setting.
//
// Reference counterparts: the ~21 backend calls per iteration of
// src/l1/pdhg.cpp:295-331, the host checks kkt_residual (:87-136) and
// original_space_residual (:145-199) with their D2H copies (:336-345), and
// the host restart with its H2D uploads (:393-404).
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "sov_hip_internal.hpp"

namespace sovhip {

struct DeviceGuard3 {
    int prev = -1;
    explicit DeviceGuard3(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard3() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// ---------------------------------------------------------------------------
// step kernels
// ---------------------------------------------------------------------------

// Per column j (a row of A^T): g = c_j - (A^T y)_j; x+ = clamp(x - tau g);
// x_bar = 2 x+ - x; x = x+; x_sum = first ? x+ : x_sum + x+.
template <typename T>
struct PrimalOp {
    const int32_t* ptr;
    const int32_t* idx;
    const T* val;
    const T* y;
    const T* c;
    const T* lo;
    const T* hi;
    T* x;
    T* xbar;
    T* xsum;
    T* acc;
    T tau;
    int first;
    int mode;
    Sched s;
    template <int G>
    __device__ void run(const ClassView& v) {
        int lane;
        const int row = row_of<G>(v, lane);
        T aty = row_dot<G, T>(ptr, idx, val, y, row, lane);
        if (row >= 0 && lane == 0) {
            if (mode == kFirst) { acc[row] = aty; return; }
            if (mode == kMiddle) { acc[row] = acc[row] + aty; return; }
            if (mode == kLast) aty = acc[row] + aty;
            const T g = c[row] - aty;
            const T xo = x[row];
            T xn = xo + (-tau) * g;
            const T l = lo[row], h = hi[row];
            if (xn < l) xn = l;
            if (xn > h) xn = h;
            xbar[row] = xn * T(2) - xo;
            x[row] = xn;
            xsum[row] = first ? xn : xsum[row] + xn;
        }
    }
};

// Per row i: s+ = clamp(s - tau y); s_bar = 2 s+ - s;
// y+ = y + sigma ((-(A x_bar)_i) + s_bar); y_sum = first ? y+ : y_sum + y+.
template <typename T>
struct DualOp {
    const int32_t* ptr;
    const int32_t* idx;
    const T* val;
    const T* xbar;
    const T* lo;
    const T* hi;
    T* s_;
    T* y;
    T* ysum;
    T* acc;
    T tau, sigma;
    int first;
    int mode;
    Sched s;
    template <int G>
    __device__ void run(const ClassView& v) {
        int lane;
        const int row = row_of<G>(v, lane);
        T ax = row_dot<G, T>(ptr, idx, val, xbar, row, lane);
        if (row >= 0 && lane == 0) {
            if (mode == kFirst) { acc[row] = ax; return; }
            if (mode == kMiddle) { acc[row] = acc[row] + ax; return; }
            if (mode == kLast) ax = acc[row] + ax;
            const T so = s_[row], yo = y[row];
            T sn = so + (-tau) * yo;
            const T l = lo[row], h = hi[row];
            if (sn < l) sn = l;
            if (sn > h) sn = h;
            const T sb = sn * T(2) - so;
            const T yn = yo + sigma * ((-ax) + sb);
            s_[row] = sn;
            y[row] = yn;
            ysum[row] = first ? yn : ysum[row] + yn;
        }
    }
};

template <class Op>
__global__ void __launch_bounds__(kBlock) k_step(Op op) {
    const ClassView v = class_view(op.s);
    dispatch_G(v, op);
}

// ---------------------------------------------------------------------------
// check kernels (FP64 arithmetic regardless of storage type)
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool finite_b(double v) { return v > -INFINITY && v < INFINITY; }

// max(primal violation, dual classification) of one row (pdhg.cpp:106-120)
// or, with the activity replaced by x_j and no primal part, one column
// (pdhg.cpp:124-134).
__device__ __forceinline__ double dual_class(double val, double lo, double hi, double point, double tol) {
    const bool at_lo = finite_b(lo) && fabs(point - lo) < tol;
    const bool at_hi = finite_b(hi) && fabs(point - hi) < tol;
    if (!at_lo && !at_hi) return fabs(val);
    if (at_lo && !at_hi) return nmax(0.0, -val);
    if (at_hi && !at_lo) return nmax(0.0, val);
    return 0.0;
}

__device__ __forceinline__ double row_res(double act, double lo, double hi, double yv, double tol) {
    double pr = 0.0;
    if (finite_b(lo)) pr = nmax(pr, lo - act);
    if (finite_b(hi)) pr = nmax(pr, act - hi);
    return nmax(pr, dual_class(yv, lo, hi, act, tol));
}

__device__ __forceinline__ double proj_sign(double v, double lo, double hi) {
    if (!finite_b(lo) && v > 0.0) v = 0.0;
    if (!finite_b(hi) && v < 0.0) v = 0.0;
    return v;
}

__device__ __forceinline__ double box_inf(double v, double lo, double hi) {
    return v > 0.0 ? v * lo : (v < 0.0 ? v * hi : 0.0);
}

// Fold one part's two dots of a row into its running pair; true when the row
// is complete (unsegmented, or the last part) and its epilogue should run.
__device__ __forceinline__ bool combine_parts(int mode, double* pair, double& d1, double& d2) {
    if (mode == kOnly) return true;
    if (mode == kFirst) {
        pair[0] = d1;
        pair[1] = d2;
        return false;
    }
    if (mode == kMiddle) {
        pair[0] += d1;
        pair[1] += d2;
        return false;
    }
    d1 += pair[0];
    d2 += pair[1];
    return true;
}

template <typename T>
struct CheckRowOp {
    const int32_t* ptr;
    const int32_t* idx;
    const T* val;
    const T* x;
    const T* xsum;
    const T* y;
    const T* ysum;
    const T* lo_s;
    const T* hi_s;
    const double* lo_o;
    const double* hi_o;
    const double* rs;
    double* act_avg;
    double* partial;
    double* acc;      // 2 per row: running (current, average) dots across parts
    double inv, tol_s, tol_o;
    int mode;
    Sched s;
    // sums: pv2_c, pv2_a, bn2, yres2_c, yres2_a, dobj_c, dobj_a; maxes: kkt_c, kkt_a, orig_c, orig_a
    template <int G>
    __device__ void run(const ClassView& v, double (&sm)[7], double (&mx)[4]) const {
        int lane;
        const int i = row_of<G>(v, lane);
        double ac, aa;
        row_dot2<G>(ptr, idx, val, x, xsum, inv, i, lane, ac, aa);
        if (i >= 0 && lane == 0) {
            if (!combine_parts(mode, acc + 2 * (int64_t)i, ac, aa)) return;
            if (act_avg != nullptr) act_avg[i] = aa;
            const double yc = double(y[i]), ya = double(ysum[i]) * inv;
            mx[0] = nmax(mx[0], row_res(ac, double(lo_s[i]), double(hi_s[i]), yc, tol_s));
            mx[1] = nmax(mx[1], row_res(aa, double(lo_s[i]), double(hi_s[i]), ya, tol_s));
            const double r = rs[i], lo = lo_o[i], hi = hi_o[i];
            const double aoc = ac / r, aoa = aa / r;
            const double yic = yc * r, yia = ya * r;
            mx[2] = nmax(mx[2], row_res(aoc, lo, hi, yic, tol_o));
            mx[3] = nmax(mx[3], row_res(aoa, lo, hi, yia, tol_o));
            double vc = 0.0, va = 0.0;
            if (finite_b(lo) && aoc < lo) vc = lo - aoc;
            if (finite_b(hi) && aoc > hi) vc = aoc - hi;
            if (finite_b(lo) && aoa < lo) va = lo - aoa;
            if (finite_b(hi) && aoa > hi) va = aoa - hi;
            sm[0] += vc * vc;
            sm[1] += va * va;
            if (finite_b(lo)) sm[2] += lo * lo;
            if (finite_b(hi) && !(finite_b(lo) && lo == hi)) sm[2] += hi * hi;
            const double pc = proj_sign(yic, lo, hi), pa = proj_sign(yia, lo, hi);
            sm[3] += (yic - pc) * (yic - pc);
            sm[4] += (yia - pa) * (yia - pa);
            sm[5] += box_inf(pc, lo, hi);
            sm[6] += box_inf(pa, lo, hi);
        }
    }
};

template <typename T>
struct CheckColOp {
    const int32_t* ptr;
    const int32_t* idx;
    const T* val;
    const T* y;
    const T* ysum;
    const T* x;
    const T* xsum;
    const T* c_s;
    const T* lo_s;
    const T* hi_s;
    const double* lo_o;
    const double* hi_o;
    const double* c_o;
    const double* cs;
    double* partial;
    double* acc;
    double inv, tol_s, tol_o, sign;
    int mode;
    Sched s;
    // sums: dres2_c, dres2_a, cn2, pobj_c, pobj_a, dobj_c, dobj_a; maxes: kkt_c, kkt_a, orig_c, orig_a
    template <int G>
    __device__ void run(const ClassView& v, double (&sm)[7], double (&mx)[4]) const {
        int lane;
        const int j = row_of<G>(v, lane);
        double tc, ta;
        row_dot2<G>(ptr, idx, val, y, ysum, inv, j, lane, tc, ta);
        if (j >= 0 && lane == 0) {
            if (!combine_parts(mode, acc + 2 * (int64_t)j, tc, ta)) return;
            const double xc = double(x[j]), xa = double(xsum[j]) * inv;
            const double cj = double(c_s[j]);
            mx[0] = nmax(mx[0], dual_class(cj - tc, double(lo_s[j]), double(hi_s[j]), xc, tol_s));
            mx[1] = nmax(mx[1], dual_class(cj - ta, double(lo_s[j]), double(hi_s[j]), xa, tol_s));
            const double sc = cs[j], lo = lo_o[j], hi = hi_o[j];
            const double ci = sign * c_o[j];
            const double xoc = xc * sc, xoa = xa * sc;
            const double rc = ci - tc / sc, ra = ci - ta / sc;
            mx[2] = nmax(mx[2], dual_class(rc, lo, hi, xoc, tol_o));
            mx[3] = nmax(mx[3], dual_class(ra, lo, hi, xoa, tol_o));
            const double lc = proj_sign(rc, lo, hi), la = proj_sign(ra, lo, hi);
            sm[0] += (rc - lc) * (rc - lc);
            sm[1] += (ra - la) * (ra - la);
            sm[2] += ci * ci;
            sm[3] += ci * xoc;
            sm[4] += ci * xoa;
            sm[5] += box_inf(lc, lo, hi);
            sm[6] += box_inf(la, lo, hi);
        }
    }
};

template <class Op>
struct CheckRunner {
    const Op& op;
    double (&sm)[7];
    double (&mx)[4];
    template <int G>
    __device__ void run(const ClassView& v) {
        op.template run<G>(v, sm, mx);
    }
};

template <class Op>
__global__ void __launch_bounds__(kBlock) k_check(Op op) {
    double sm[7], mx[4];
#pragma unroll
    for (int i = 0; i < 7; ++i) sm[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) mx[i] = 0.0;
    const ClassView v = class_view(op.s);
    CheckRunner<Op> r{op, sm, mx};
    dispatch_G(v, r);
    // mode is uniform over the launch: only a launch that completes rows reduces
    if (op.mode == kOnly || op.mode == kLast)
        block_reduce_write<7, 4>(sm, mx, op.partial + (int64_t)blockIdx.x * kCheckSlots);
}

// Block b reduces slot (b % 11) of region (b / 11): slots 0..6 sums, 7..10 maxima.
__global__ void k_check_final(const double* __restrict__ rowpart, int64_t rblocks,
                              const double* __restrict__ colpart, int64_t cblocks, double* __restrict__ out) {
    const int region = blockIdx.x / 11, slot = blockIdx.x % 11;
    const double* p = region == 0 ? rowpart : colpart;
    const int64_t nb = region == 0 ? rblocks : cblocks;
    const bool is_max = slot >= 7;
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) {
        const double v = p[b * kCheckSlots + slot];
        acc = is_max ? nmax(acc, v) : acc + v;
    }
    double sm[1] = {acc}, mx[1] = {acc};
    if (is_max) block_reduce_write<0, 1>(sm, mx, out + blockIdx.x);
    else block_reduce_write<1, 0>(sm, mx, out + blockIdx.x);
}

// ---------------------------------------------------------------------------
// restart / set point
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_restart_x(int32_t n, const T* __restrict__ xsum, double inv, const T* __restrict__ lo,
                            const T* __restrict__ hi, T* __restrict__ x, double* __restrict__ xlast,
                            double* __restrict__ partial) {
    double acc = 0.0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        double v = double(xsum[j]) * inv;
        const double l = double(lo[j]), h = double(hi[j]);
        if (v < l) v = l;
        if (v > h) v = h;
        const T t = T(v);
        x[j] = t;
        if (xlast != nullptr) {
            const double d = double(t) - xlast[j];
            acc += d * d;
            xlast[j] = double(t);
        }
    }
    double sm[1] = {acc}, mx[1] = {0.0};
    block_reduce_write<1, 0>(sm, mx, partial + blockIdx.x);
}

template <typename T>
__global__ void k_restart_y(int32_t m, const T* __restrict__ ysum, double inv, const double* __restrict__ act,
                            const T* __restrict__ lo, const T* __restrict__ hi, T* __restrict__ y,
                            T* __restrict__ s, double* __restrict__ ylast, double* __restrict__ partial) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const T t = T(double(ysum[i]) * inv);
        y[i] = t;
        double a = act[i];
        const double l = double(lo[i]), h = double(hi[i]);
        if (a < l) a = l;
        if (a > h) a = h;
        s[i] = T(a);
        if (ylast != nullptr) {
            const double d = double(t) - ylast[i];
            acc += d * d;
            ylast[i] = double(t);
        }
    }
    double sm[1] = {acc}, mx[1] = {0.0};
    block_reduce_write<1, 0>(sm, mx, partial + blockIdx.x);
}

__global__ void k_sum_partials2(const double* __restrict__ p0, int64_t n0, const double* __restrict__ p1,
                                int64_t n1, double* __restrict__ out) {
    const double* p = blockIdx.x == 0 ? p0 : p1;
    const int64_t nb = blockIdx.x == 0 ? n0 : n1;
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) acc += p[b];
    double sm[1] = {acc}, mx[1] = {0.0};
    block_reduce_write<1, 0>(sm, mx, out + blockIdx.x);
}

template <typename T>
__global__ void k_set_clamped(int32_t n, const double* __restrict__ src, const T* __restrict__ lo,
                              const T* __restrict__ hi, T* __restrict__ dst) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        double v = src[j];
        if (lo != nullptr) {
            const double l = double(lo[j]);
            if (v < l) v = l;
        }
        if (hi != nullptr) {
            const double h = double(hi[j]);
            if (v > h) v = h;
        }
        dst[j] = T(v);
    }
}

template <typename T>
__global__ void k_movement(int32_t len, const T* __restrict__ v, double* __restrict__ last,
                           double* __restrict__ partial) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        const double t = double(v[i]);
        const double d = t - last[i];
        acc += d * d;
        last[i] = t;
    }
    double sm[1] = {acc}, mx[1] = {0.0};
    block_reduce_write<1, 0>(sm, mx, partial + blockIdx.x);
}

constexpr int kRestartBlocks = 512;

static inline double nmax_host(double a, double b) { return (a > b || a != a) ? a : b; }

template <typename T>
static const T* vals_of(const Part& p) {
    return std::is_same<T, double>::value ? reinterpret_cast<const T*>(p.val64) : reinterpret_cast<const T*>(p.val32);
}

template <typename T>
static int check_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_check_args* a, sov_hip_check_out* out) {
    const sov_hip_pdhg_state* st = a->st;
    const double inv = 1.0 / static_cast<double>(a->since_restart > 0 ? a->since_restart : 1);
    const Orient& OA = h->a;
    const Orient& OT = h->t;
    const int npa = OA.nparts(), npt = OT.nparts();
    // the launch that completes rows (the last part) owns the block partials
    const int rb = OA.part(npa - 1).blocks(), cb = OT.part(npt - 1).blocks();
    for (const Orient* o : {&OA, &OT})
        for (const Part& p : o->parts)
            if (p.nnz > 0 && vals_of<T>(p) == nullptr) return fail(SOV_INVALID_ARGUMENT, "check: value image missing");
    double* rowpart = h->partials;
    double* colpart = h->partials + (int64_t)(rb + 1) * kCheckSlots;
    if ((int64_t)(rb + cb + 2) > h->partial_blocks) return fail(SOV_INVALID_ARGUMENT, "check: partials too small");
    int rc = ctx->ensure_red(64);
    if (rc != SOV_OK) return rc;
    for (int k = 0; k < npa; ++k) {
        const Part& p = OA.part(k);
        const Sched s = p.sched();
        const int blocks = s.block_begin[kClasses];
        if (blocks == 0) continue;
        CheckRowOp<T> op{};
        op.ptr = p.ptr; op.idx = p.idx; op.val = vals_of<T>(p);
        op.x = static_cast<const T*>(st->x); op.xsum = static_cast<const T*>(st->x_sum);
        op.y = static_cast<const T*>(st->y); op.ysum = static_cast<const T*>(st->y_sum);
        op.lo_s = static_cast<const T*>(st->row_lo); op.hi_s = static_cast<const T*>(st->row_hi);
        op.lo_o = a->orig_row_lo; op.hi_o = a->orig_row_hi; op.rs = a->row_scale;
        op.act_avg = a->act_avg; op.partial = rowpart; op.acc = OA.chk_acc;
        op.inv = inv; op.tol_s = a->tol_scaled; op.tol_o = a->tol_orig; op.mode = seg_mode(k, npa); op.s = s;
        k_check<CheckRowOp<T>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    for (int k = 0; k < npt; ++k) {
        const Part& p = OT.part(k);
        const Sched s = p.sched();
        const int blocks = s.block_begin[kClasses];
        if (blocks == 0) continue;
        CheckColOp<T> op{};
        op.ptr = p.ptr; op.idx = p.idx; op.val = vals_of<T>(p);
        op.y = static_cast<const T*>(st->y); op.ysum = static_cast<const T*>(st->y_sum);
        op.x = static_cast<const T*>(st->x); op.xsum = static_cast<const T*>(st->x_sum);
        op.c_s = static_cast<const T*>(st->c);
        op.lo_s = static_cast<const T*>(st->col_lo); op.hi_s = static_cast<const T*>(st->col_hi);
        op.lo_o = a->orig_col_lo; op.hi_o = a->orig_col_hi; op.c_o = a->orig_c; op.cs = a->col_scale;
        op.partial = colpart; op.acc = OT.chk_acc;
        op.inv = inv; op.tol_s = a->tol_scaled; op.tol_o = a->tol_orig; op.sign = a->obj_sign;
        op.mode = seg_mode(k, npt); op.s = s;
        k_check<CheckColOp<T>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    k_check_final<<<22, kBlock, 0, ctx->stream>>>(rowpart, rb, colpart, cb, ctx->d_red);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, ctx->d_red, 22 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double* r = ctx->h_pinned;        // [0..10] rows, [11..21] columns
    const double* q = ctx->h_pinned + 11;
    out->kkt_cur = nmax_host(r[7], q[7]);
    out->kkt_avg = nmax_host(r[8], q[8]);
    out->orig_cur = nmax_host(r[9], q[9]);
    out->orig_avg = nmax_host(r[10], q[10]);
    const double bn = std::sqrt(r[2]), cn = std::sqrt(q[2]);
    const double sg = a->obj_sign;
    for (int w = 0; w < 2; ++w) {
        const double pv2 = r[0 + w], yres2 = r[3 + w], dobj_r = r[5 + w];
        const double dres2 = q[0 + w], pobj = q[3 + w], dobj_c = q[5 + w];
        const double p = sg * pobj + a->obj_const;
        const double d = sg * (dobj_r + dobj_c) + a->obj_const;
        double* rel = w == 0 ? out->rel_cur : out->rel_avg;
        rel[0] = std::sqrt(pv2) / (1.0 + bn);
        rel[1] = std::sqrt(yres2 + dres2) / (1.0 + cn);
        rel[2] = std::fabs(p - d) / (1.0 + std::fabs(p) + std::fabs(d));
        if (w == 0) { out->pobj_cur = p; out->dobj_cur = d; }
        else { out->pobj_avg = p; out->dobj_avg = d; }
    }
    return SOV_OK;
}

}  // namespace sovhip

using namespace sovhip;

namespace {

template <typename T>
T* orient_acc(const Orient& o) {
    return std::is_same<T, double>::value ? reinterpret_cast<T*>(o.acc64) : reinterpret_cast<T*>(o.acc32);
}

template <typename T>
int primal_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_pdhg_state* st, double tau, int first) {
    const Orient& o = h->t;
    const int np = o.nparts();
    for (int k = 0; k < np; ++k) {
        const Part& p = o.part(k);
        const Sched s = p.sched();
        const int blocks = s.block_begin[kClasses];
        if (blocks == 0) continue;
        const T* val = vals_of<T>(p);
        if (val == nullptr && p.nnz > 0) return fail(SOV_INVALID_ARGUMENT, "primal_step: value image missing");
        PrimalOp<T> op{p.ptr, p.idx, val, static_cast<const T*>(st->y), static_cast<const T*>(st->c),
                       static_cast<const T*>(st->col_lo), static_cast<const T*>(st->col_hi), static_cast<T*>(st->x),
                       static_cast<T*>(st->x_bar), static_cast<T*>(st->x_sum), orient_acc<T>(o), T(tau), first,
                       seg_mode(k, np), s};
        k_step<PrimalOp<T>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

template <typename T>
int dual_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_pdhg_state* st, double tau, double sigma,
              int first) {
    const Orient& o = h->a;
    const int np = o.nparts();
    for (int k = 0; k < np; ++k) {
        const Part& p = o.part(k);
        const Sched s = p.sched();
        const int blocks = s.block_begin[kClasses];
        if (blocks == 0) continue;
        const T* val = vals_of<T>(p);
        if (val == nullptr && p.nnz > 0) return fail(SOV_INVALID_ARGUMENT, "dual_step: value image missing");
        DualOp<T> op{p.ptr, p.idx, val, static_cast<const T*>(st->x_bar), static_cast<const T*>(st->row_lo),
                     static_cast<const T*>(st->row_hi), static_cast<T*>(st->s), static_cast<T*>(st->y),
                     static_cast<T*>(st->y_sum), orient_acc<T>(o), T(tau), T(sigma), first, seg_mode(k, np), s};
        k_step<DualOp<T>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int state_ok(const sov_hip_csr* h, const sov_hip_pdhg_state* st) {
    if (st == nullptr) return fail(SOV_INVALID_ARGUMENT, "state is null");
    if (st->prec != SOV_PREC_F64 && st->prec != SOV_PREC_F32) return fail(SOV_INVALID_ARGUMENT, "state precision");
    if (h != nullptr && (st->m != h->m || st->n != h->n)) return fail(SOV_DIMENSION_MISMATCH, "state dims vs csr");
    if ((st->n > 0 && (!st->c || !st->col_lo || !st->col_hi || !st->x || !st->x_bar || !st->x_sum)) ||
        (st->m > 0 && (!st->row_lo || !st->row_hi || !st->s || !st->y || !st->y_sum)))
        return fail(SOV_INVALID_ARGUMENT, "state has a null vector");
    return SOV_OK;
}

template <typename T>
int restart_impl(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, int64_t since, const double* act, double* xl,
                 double* yl, double movement[2]) {
    const double inv = 1.0 / static_cast<double>(since > 0 ? since : 1);
    int rc = ctx->ensure_red(2 * kRestartBlocks + 8);
    if (rc != SOV_OK) return rc;
    double* px = ctx->d_red;
    double* py = px + kRestartBlocks;
    double* res = py + kRestartBlocks;
    const int bx = std::min(launch_blocks(st->n), kRestartBlocks);
    const int by = std::min(launch_blocks(st->m), kRestartBlocks);
    k_restart_x<T><<<bx, kBlock, 0, ctx->stream>>>(st->n, static_cast<const T*>(st->x_sum), inv,
                                                   static_cast<const T*>(st->col_lo), static_cast<const T*>(st->col_hi),
                                                   static_cast<T*>(st->x), xl, px);
    k_restart_y<T><<<by, kBlock, 0, ctx->stream>>>(st->m, static_cast<const T*>(st->y_sum), inv, act,
                                                   static_cast<const T*>(st->row_lo), static_cast<const T*>(st->row_hi),
                                                   static_cast<T*>(st->y), static_cast<T*>(st->s), yl, py);
    k_sum_partials2<<<2, kBlock, 0, ctx->stream>>>(px, bx, py, by, res);
    SOV_HIP_TRY(check_launch());
    if (movement != nullptr) {
        SOV_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, res, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
        movement[0] = ctx->h_pinned[0];
        movement[1] = ctx->h_pinned[1];
    }
    return SOV_OK;
}

template <typename T>
int movement_impl(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, double* xl, double* yl, double movement[2]) {
    int rc = ctx->ensure_red(2 * kRestartBlocks + 8);
    if (rc != SOV_OK) return rc;
    double* px = ctx->d_red;
    double* py = px + kRestartBlocks;
    double* res = py + kRestartBlocks;
    const int bx = std::min(launch_blocks(st->n), kRestartBlocks);
    const int by = std::min(launch_blocks(st->m), kRestartBlocks);
    k_movement<T><<<bx, kBlock, 0, ctx->stream>>>(st->n, static_cast<const T*>(st->x), xl, px);
    k_movement<T><<<by, kBlock, 0, ctx->stream>>>(st->m, static_cast<const T*>(st->y), yl, py);
    k_sum_partials2<<<2, kBlock, 0, ctx->stream>>>(px, bx, py, by, res);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, res, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    movement[0] = ctx->h_pinned[0];
    movement[1] = ctx->h_pinned[1];
    return SOV_OK;
}

template <typename T>
int set_point_impl(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, const double* xs, const double* ys,
                   const double* act) {
    if (xs != nullptr && st->n > 0)
        k_set_clamped<T><<<launch_blocks(st->n), kBlock, 0, ctx->stream>>>(
            st->n, xs, static_cast<const T*>(st->col_lo), static_cast<const T*>(st->col_hi), static_cast<T*>(st->x));
    if (ys != nullptr && st->m > 0)
        k_set_clamped<T><<<launch_blocks(st->m), kBlock, 0, ctx->stream>>>(st->m, ys, nullptr, nullptr,
                                                                           static_cast<T*>(st->y));
    if (act != nullptr && st->m > 0)
        k_set_clamped<T><<<launch_blocks(st->m), kBlock, 0, ctx->stream>>>(
            st->m, act, static_cast<const T*>(st->row_lo), static_cast<const T*>(st->row_hi), static_cast<T*>(st->s));
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

}  // namespace

extern "C" {

int sov_hip_pdhg_primal_step(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_pdhg_state* st, double tau,
                             int first) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "primal_step args");
    const int rc = state_ok(h, st);
    if (rc != SOV_OK) return rc;
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? primal_impl<double>(ctx, h, st, tau, first)
                                    : primal_impl<float>(ctx, h, st, tau, first);
}

int sov_hip_pdhg_dual_step(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_pdhg_state* st, double tau,
                           double sigma, int first) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "dual_step args");
    const int rc = state_ok(h, st);
    if (rc != SOV_OK) return rc;
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? dual_impl<double>(ctx, h, st, tau, sigma, first)
                                    : dual_impl<float>(ctx, h, st, tau, sigma, first);
}

int sov_hip_pdhg_check(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_check_args* a, sov_hip_check_out* out) {
    if (ctx == nullptr || h == nullptr || a == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "check args");
    int rc = state_ok(h, a->st);
    if (rc != SOV_OK) return rc;
    if ((h->m > 0 && (!a->row_scale || !a->orig_row_lo || !a->orig_row_hi)) ||
        (h->n > 0 && (!a->col_scale || !a->orig_col_lo || !a->orig_col_hi || !a->orig_c)))
        return fail(SOV_INVALID_ARGUMENT, "check: null original-space array");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "check during capture");
    DeviceGuard3 g(ctx->device);
    return a->st->prec == SOV_PREC_F64 ? check_impl<double>(ctx, h, a, out) : check_impl<float>(ctx, h, a, out);
}

int sov_hip_pdhg_restart(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, int64_t since, const double* act,
                         double* xl, double* yl, double movement[2]) {
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "restart args");
    int rc = state_ok(nullptr, st);
    if (rc != SOV_OK) return rc;
    if (st->m > 0 && act == nullptr) return fail(SOV_INVALID_ARGUMENT, "restart: act is null");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "restart during capture");
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? restart_impl<double>(ctx, st, since, act, xl, yl, movement)
                                    : restart_impl<float>(ctx, st, since, act, xl, yl, movement);
}

int sov_hip_pdhg_movement(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, double* xl, double* yl,
                          double movement[2]) {
    if (ctx == nullptr || movement == nullptr) return fail(SOV_INVALID_ARGUMENT, "movement args");
    int rc = state_ok(nullptr, st);
    if (rc != SOV_OK) return rc;
    if ((st->n > 0 && xl == nullptr) || (st->m > 0 && yl == nullptr)) return fail(SOV_INVALID_ARGUMENT, "movement: null last point");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "movement during capture");
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? movement_impl<double>(ctx, st, xl, yl, movement)
                                    : movement_impl<float>(ctx, st, xl, yl, movement);
}

int sov_hip_pdhg_set_point(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, const double* xs, const double* ys,
                           const double* act) {
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "set_point args");
    int rc = state_ok(nullptr, st);
    if (rc != SOV_OK) return rc;
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? set_point_impl<double>(ctx, st, xs, ys, act)
                                    : set_point_impl<float>(ctx, st, xs, ys, act);
}

}  // extern "C"
