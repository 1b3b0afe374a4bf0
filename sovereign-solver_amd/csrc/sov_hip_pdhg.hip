// Fused PDHG kernels: the primal and dual steps (kernels a+b), the KKT /
// restart reductions (kernel c), restart and point setting.
//
// Reference counterparts: the ~21 backend calls per iteration of
// src/l1/pdhg.cpp:295-331, the host checks kkt_residual (:87-136) and
// original_space_residual (:145-199) with their D2H copies (:336-345), and
// the host restart with its H2D uploads (:393-404).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>

#include "sov_hip_internal.hpp"
#include "sov_hip_tiles.hpp"

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

// Per-instance parameters of a block-diagonal batch of LPs, read by the
// batched (B = true) variants of the kernels: the instance of each
// device-order row, its step sizes, where it iterates (0 = stopped, 1 = in
// the FP32 state, 2 = in the FP64 state: a step launch on a state of one
// precision skips the rows of instances living in the other), and how many
// iterations its Cesaro sums hold.  k is the iteration's index inside the
// captured interval (a launch argument, fixed per graph node).
struct BatchView {
    const int32_t* inst = nullptr;
    const double* tau = nullptr;
    const double* sigma = nullptr;
    const int32_t* active = nullptr;
    const int32_t* since = nullptr;
    int k = 0;
    int want = 0;   // the active value this launch's state serves
    const int32_t* list = nullptr;  // the blocks this launch runs (compacted), or null for all
    const int32_t* skip = nullptr;  // per instance: 1 = resident (its rows run in the resident launch instead)
};

// The per-row updates of the two steps, shared by every kernel that makes
// them (so that they make them bit for bit alike).  x, x_bar, x_sum (s, y,
// y_sum) in and out; x_sum's (y_sum's) input is unused when first.
template <typename T>
__device__ __forceinline__ void primal_update(T c, T aty, T l, T h, T tau, int first, T& x, T& xbar, T& xsum) {
    const T g = c - aty;
    const T xo = x;
    T xn = xo + (-tau) * g;
    if (xn < l) xn = l;
    if (xn > h) xn = h;
    xbar = xn * T(2) - xo;
    x = xn;
    xsum = first ? xn : xsum + xn;
}
template <typename T>
__device__ __forceinline__ void dual_update(T ax, T l, T h, T sgv, bool has_sg, T tau, T sigma, int first, T& s,
                                            T& y, T& ysum) {
    const T so = s, yo = y;
    const T gs = has_sg ? yo + sgv : yo;   // the slacks' gradient
    T sn = so + (-tau) * gs;
    if (sn < l) sn = l;
    if (sn > h) sn = h;
    const T sb = sn * T(2) - so;
    const T yn = yo + sigma * ((-ax) + sb);
    s = sn;
    y = yn;
    ysum = first ? yn : ysum + yn;
}

// Per column j (a row of A^T): g = c_j - (A^T y)_j; x+ = clamp(x - tau g);
// x_bar = 2 x+ - x; x = x+; x_sum = first ? x+ : x_sum + x+.
template <typename T, bool B = false>
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
    BatchView bv;
    // a batch launch runs the listed blocks in XCD-aware order: an XCD's
    // share of the list is a run of consecutive blocks (fewer instances)
    __device__ int block() const {
        return (B && bv.list != nullptr) ? bv.list[xcd_slot((int)blockIdx.x, (int)gridDim.x)] : (int)blockIdx.x;
    }
    // The row's own operand loads (bounds, iterates, a segment's partial sum
    // and, in a batch, its instance) are issued before the product's, so that
    // the row waits for one chain of memory round trips, not for one after
    // the other; a batch row whose instance does not iterate in this launch
    // computes its product and drops it.  Same arithmetic as apply().
    template <int G>
    __device__ void run(const ClassView& v) {
        int lane;
        const int row = row_of<G>(v, lane);
        const bool mine = row >= 0 && lane == 0;
        T cr = T(0), lr = T(0), hr = T(0), xv = T(0), xs = T(0), ar = T(0);
        int b = 0;
        if (mine) {
            if (mode == kMiddle || mode == kLast) ar = acc[row];
            if (mode == kOnly || mode == kLast) {
                cr = c[row];
                lr = lo[row];
                hr = hi[row];
                xv = x[row];
                xs = xsum[row];
            }
            if constexpr (B) b = bv.inst[row];
        }
        T aty = row_dot<G, T>(ptr, idx, val, y, row, lane);
        if (!mine) return;
        T tau_r = tau;
        int first_r = first;
        if constexpr (B) {
            if (bv.active[b] != bv.want || (bv.skip != nullptr && bv.skip[b])) return;
            tau_r = T(bv.tau[b]);
            first_r = (bv.k == 0 && bv.since[b] == 0) ? 1 : 0;
        }
        if (mode == kFirst) { acc[row] = aty; return; }
        if (mode == kMiddle) { acc[row] = ar + aty; return; }
        if (mode == kLast) aty = ar + aty;
        T xb;
        if (first_r) xs = T(0);
        primal_update(cr, aty, lr, hr, tau_r, first_r, xv, xb, xs);
        xbar[row] = xb;
        x[row] = xv;
        xsum[row] = xs;
    }
    // a row of batch instance b, with b's step size and Cesaro state
    __device__ __forceinline__ void apply_batch(int row, T aty, int b) {
        apply(row, aty, T(bv.tau[b]), (bv.k == 0 && bv.since[b] == 0) ? 1 : 0);
    }
    // the row's full product (A^T y)_row, or a segment's part of it
    __device__ __forceinline__ void apply(int row, T aty, T tau_r, int first_r) {
        if (mode == kFirst) { acc[row] = aty; return; }
        if (mode == kMiddle) { acc[row] = acc[row] + aty; return; }
        if (mode == kLast) aty = acc[row] + aty;
        T xv = x[row], xb, xs = first_r ? T(0) : xsum[row];
        primal_update(c[row], aty, lo[row], hi[row], tau_r, first_r, xv, xb, xs);
        xbar[row] = xb;
        x[row] = xv;
        xsum[row] = xs;
    }
    // a row's sum from the tiled kernels; in a batch, only the rows of
    // instances iterating in this launch's precision (and not resident)
    __device__ __forceinline__ void finish(int row, T aty) {
        if constexpr (B) {
            const int b = bv.inst[row];
            if (bv.active[b] == bv.want && (bv.skip == nullptr || !bv.skip[b])) apply_batch(row, aty, b);
        } else {
            apply(row, aty, tau, first);
        }
    }
};

// Per row i: s+ = clamp(s - tau y); s_bar = 2 s+ - s;
// y+ = y + sigma ((-(A x_bar)_i) + s_bar); y_sum = first ? y+ : y_sum + y+.
template <typename T, bool B = false>
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
    BatchView bv;
    const T* sg = nullptr;   // optional objective of the slacks (refinement corrections)
    __device__ int block() const {
        return (B && bv.list != nullptr) ? bv.list[xcd_slot((int)blockIdx.x, (int)gridDim.x)] : (int)blockIdx.x;
    }
    // as PrimalOp::run
    template <int G>
    __device__ void run(const ClassView& v) {
        int lane;
        const int row = row_of<G>(v, lane);
        const bool mine = row >= 0 && lane == 0;
        T lr = T(0), hr = T(0), sv = T(0), yv = T(0), ys = T(0), sgv = T(0), ar = T(0);
        int b = 0;
        if (mine) {
            if (mode == kMiddle || mode == kLast) ar = acc[row];
            if (mode == kOnly || mode == kLast) {
                lr = lo[row];
                hr = hi[row];
                sv = s_[row];
                yv = y[row];
                ys = ysum[row];
                if (sg != nullptr) sgv = sg[row];
            }
            if constexpr (B) b = bv.inst[row];
        }
        T ax = row_dot<G, T>(ptr, idx, val, xbar, row, lane);
        if (!mine) return;
        T tau_r = tau, sigma_r = sigma;
        int first_r = first;
        if constexpr (B) {
            if (bv.active[b] != bv.want || (bv.skip != nullptr && bv.skip[b])) return;
            tau_r = T(bv.tau[b]);
            sigma_r = T(bv.sigma[b]);
            first_r = (bv.k == 0 && bv.since[b] == 0) ? 1 : 0;
        }
        if (mode == kFirst) { acc[row] = ax; return; }
        if (mode == kMiddle) { acc[row] = ar + ax; return; }
        if (mode == kLast) ax = ar + ax;
        if (first_r) ys = T(0);
        dual_update(ax, lr, hr, sgv, sg != nullptr, tau_r, sigma_r, first_r, sv, yv, ys);
        s_[row] = sv;
        y[row] = yv;
        ysum[row] = ys;
    }
    // a row of batch instance b, with b's step sizes and Cesaro state
    __device__ __forceinline__ void apply_batch(int row, T ax, int b) {
        apply(row, ax, T(bv.tau[b]), T(bv.sigma[b]), (bv.k == 0 && bv.since[b] == 0) ? 1 : 0);
    }
    // the row's full product (A x_bar)_row, or a segment's part of it
    __device__ __forceinline__ void apply(int row, T ax, T tau_r, T sigma_r, int first_r) {
        if (mode == kFirst) { acc[row] = ax; return; }
        if (mode == kMiddle) { acc[row] = acc[row] + ax; return; }
        if (mode == kLast) ax = acc[row] + ax;
        T sv = s_[row], yv = y[row], ys = first_r ? T(0) : ysum[row];
        dual_update(ax, lo[row], hi[row], sg != nullptr ? sg[row] : T(0), sg != nullptr, tau_r, sigma_r, first_r,
                    sv, yv, ys);
        s_[row] = sv;
        y[row] = yv;
        ysum[row] = ys;
    }
    // as PrimalOp::finish
    __device__ __forceinline__ void finish(int row, T ax) {
        if constexpr (B) {
            const int b = bv.inst[row];
            if (bv.active[b] == bv.want && (bv.skip == nullptr || !bv.skip[b])) apply_batch(row, ax, b);
        } else {
            apply(row, ax, tau, sigma, first);
        }
    }
};

template <class Op>
__global__ void __launch_bounds__(kBlock) k_step(Op op) {
    const ClassView v = class_view(op.s, op.block());
    dispatch_G(v, op);
}

// The same steps over the tiled layout: x is the gathered vector (y for the
// primal step, x_bar for the dual step).
template <typename T, class Op>
__global__ void __launch_bounds__(kTileThreads) k_tile_step(Op op, TileArgs ta, const T* x) {
    tile_rows<T>(ta, x, op);
}

// (start, length) of instance b's rows of class c in dimension dim.
__device__ __forceinline__ int2 batch_range(const int32_t* seg, int b, int dim, int c) {
    const int32_t* p = seg + (((int64_t)b * 2 + dim) * kClasses + c) * 2;
    return make_int2(p[0], p[1]);
}

// Resident batch intervals.  A small instance of a block-diagonal batch runs
// a whole interval of iterations in one workgroup: its iterates (x, x_bar,
// x_sum over its columns, s, y, y_sum over its rows) and its rows' bounds
// and objective are copied into LDS, its entries of both orientations into
// registers, and each iteration makes the primal step over its rows of A^T,
// a barrier, the dual step over its rows of A, a barrier; the iterates are
// written back at the end.  An instance's columns (rows) in device order are
// the runs seg(b, 1, c) (seg(b, 0, c)) of the 8 length classes; LDS holds
// them concatenated in class order, and an entry's lidx is its column's place
// in that concatenation.
//
// The work of a pass: the rows of class c take G_c lanes each (the default
// lanes of class c, kDefaultLanes), so the class is ceil(len_c G_c / 64)
// wave-sized chunks;
// the chunks of all classes are numbered in class order and wave w holds
// chunks w, w + 16, ... (at most kResidentSlotsP / kResidentSlotsD of them,
// with at most kResidentEntriesP / kResidentEntriesD entries per lane: the
// host admits only instances that fit).  Each lane adds its entries in
// ascending position, the G lanes of a row combine by group_sum's xor
// butterfly, and every row is updated by primal_update / dual_update, so
// where the CSR schedule has its default lanes (no autotuning) an interval
// ends where `count` CSR step launches would, bit for bit; with tuned lanes
// the sums differ from theirs in rounding only.  Which instances are
// resident depends on each instance alone, never on the tuning.
template <typename T>
struct ResidentArgs {
    const int32_t* a_ptr;     // A (rows: instance rows), one part, device order
    const uint16_t* a_lidx;
    const T* a_val;
    const int32_t* t_ptr;     // A^T (rows: instance columns)
    const uint16_t* t_lidx;
    const T* t_val;
    const int32_t* seg;
    const int32_t* list;      // the instances this launch runs
    T* x;
    T* xbar;
    T* xsum;
    const T* c;
    const T* clo;
    const T* chi;
    T* y;
    T* s;
    T* ysum;
    const T* rlo;
    const T* rhi;
    BatchView bv;
    int count;                // iterations
    int ga[kClasses], gt[kClasses];   // lanes per class of A's and A^T's rows (kDefaultLanes)
};

constexpr int kResidentThreads = 1024;
constexpr int kResidentWaves = kResidentThreads / kWave;
constexpr int kResidentSlotsP = 5;     // chunks per wave of the primal pass
constexpr int kResidentEntriesP = 2;   // entries per lane of the primal pass
constexpr int kResidentSlotsD = 3;     // chunks per wave of the dual pass
constexpr int kResidentEntriesD = 4;   // entries per lane of the dual pass

// One pass's rows of one instance: per class, the first device row, the
// count, the first local row, the lanes per row, and the first chunk
// (pre[kClasses] = the pass's chunk count).
struct ResidentPass {
    int r0[kClasses];
    int len[kClasses];
    int lb[kClasses];
    int g[kClasses];
    int pre[kClasses + 1];
};

// group_sum with a runtime (power-of-two) G: the same additions in the same order
template <typename T>
__device__ __forceinline__ T group_sum_rt(T v, int G) {
    for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// A wave's chunks of one pass, in registers for the whole interval: per
// chunk, the lanes per row, the local row this lane updates (-1: none),
// and the lane's entries (value, local index of the gathered element).
template <typename T, int S, int E>
struct ResidentChunks {
    int G[S];
    int upd[S];
    int cnt[S];
    T w[S][E];
    int li[S][E];
};

template <typename T, int S, int E>
__device__ __forceinline__ void resident_load(const ResidentPass& P, const int32_t* __restrict__ ptr,
                                              const T* __restrict__ val, const uint16_t* __restrict__ lidx,
                                              ResidentChunks<T, S, E>& ch) {
    // the chunk numbers and lanes per row are wave-uniform: scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    const int C = P.pre[kClasses];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int j = wave + u * kResidentWaves;   // wave-uniform
        int c = 0;
#pragma unroll
        for (int cc = 1; cc < kClasses; ++cc)
            if (j >= P.pre[cc]) c = cc;
        const int G = P.g[c];
        const int rr = ((j - P.pre[c]) * kWave + lane) / G;
        const int q = lane & (G - 1);
        const bool ok = j < C && rr < P.len[c];
        ch.G[u] = j < C ? G : 1;
        ch.upd[u] = (ok && q == 0) ? P.lb[c] + rr : -1;
        int e0 = 0, e1 = 0;
        if (ok) {
            e0 = ptr[P.r0[c] + rr] + q;
            e1 = ptr[P.r0[c] + rr + 1];
        }
        ch.cnt[u] = e1 > e0 ? min((e1 - e0 + G - 1) / G, E) : 0;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            ch.w[u][k] = T(0);
            ch.li[u][k] = 0;
            if (k < ch.cnt[u]) {
                ch.w[u][k] = val[e0 + k * G];
                ch.li[u][k] = lidx[e0 + k * G];
            }
        }
    }
}

// One pass (primal: Pr, over A^T's rows, gathering y; dual: over A's rows,
// gathering x_bar) of one iteration, in LDS and registers.  k0/k1/k2: the
// rows' objective (primal only) and bounds.
template <bool Pr, typename T, int S, int E>
__device__ __forceinline__ void resident_pass(const ResidentChunks<T, S, E>& ch, const T* __restrict__ vin,
                                              const T* __restrict__ k0, const T* __restrict__ k1,
                                              const T* __restrict__ k2, T* __restrict__ u0, T* __restrict__ u1,
                                              T* __restrict__ u2, T tau, T sigma, int first) {
#pragma unroll
    for (int u = 0; u < S; ++u) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < E; ++k)
            if (k < ch.cnt[u]) acc += ch.w[u][k] * vin[ch.li[u][k]];
        acc = group_sum_rt(acc, ch.G[u]);
        const int l = ch.upd[u];
        if (l >= 0) {
            if constexpr (Pr) primal_update(k0[l], acc, k1[l], k2[l], tau, first, u0[l], u1[l], u2[l]);
            else dual_update(acc, k1[l], k2[l], T(0), false, tau, sigma, first, u0[l], u1[l], u2[l]);
        }
    }
}

// copy an instance's runs of a dimension between a device vector and its
// concatenation in LDS
template <bool In, typename D, typename T>
__device__ __forceinline__ void resident_copy(const ResidentPass& P, D* dev, T* lds) {
    for (int c = 0; c < kClasses; ++c)
        for (int i = threadIdx.x; i < P.len[c]; i += kResidentThreads) {
            if constexpr (In) lds[P.lb[c] + i] = dev[P.r0[c] + i];
            else dev[P.r0[c] + i] = lds[P.lb[c] + i];
        }
}

// LDS bytes of a resident instance with n columns and m rows: x, x_bar,
// x_sum, c, col_lo, col_hi, then s, y, y_sum, row_lo, row_hi.
__host__ __device__ constexpr int64_t resident_lds_bytes(int64_t n, int64_t m, int64_t tsize) {
    return (6 * n + 5 * m) * tsize;
}

template <typename T>
__global__ void __launch_bounds__(kResidentThreads) k_batch_resident(ResidentArgs<T> a) {
    __shared__ ResidentPass P[2];   // 0: the primal pass (A^T), 1: the dual pass (A)
    const int b = a.list[blockIdx.x];
    if (a.bv.active[b] != a.bv.want) return;   // uniform: the whole instance
    if (threadIdx.x < 2) {
        const int p = static_cast<int>(threadIdx.x);
        int lb = 0, pre = 0;
        for (int c = 0; c < kClasses; ++c) {
            const int2 rr = batch_range(a.seg, b, p == 0 ? 1 : 0, c);
            const int g = p == 0 ? a.gt[c] : a.ga[c];
            P[p].r0[c] = rr.x;
            P[p].len[c] = rr.y;
            P[p].lb[c] = lb;
            P[p].g[c] = g;
            P[p].pre[c] = pre;
            lb += rr.y;
            pre += (rr.y * g + kWave - 1) / kWave;
        }
        P[p].pre[kClasses] = pre;
    }
    __syncthreads();
    const int n = P[0].lb[kClasses - 1] + P[0].len[kClasses - 1];
    const int m = P[1].lb[kClasses - 1] + P[1].len[kClasses - 1];
    extern __shared__ __align__(16) unsigned char res_lds[];
    T* xl = reinterpret_cast<T*>(res_lds);
    T* xbl = xl + n;
    T* xsl = xbl + n;
    T* cl = xsl + n;
    T* clol = cl + n;
    T* chil = clol + n;
    T* sl = chil + n;
    T* yl = sl + m;
    T* ysl = yl + m;
    T* rlol = ysl + m;
    T* rhil = rlol + m;
    ResidentChunks<T, kResidentSlotsP, kResidentEntriesP> cp;
    ResidentChunks<T, kResidentSlotsD, kResidentEntriesD> cd;
    resident_load(P[0], a.t_ptr, a.t_val, a.t_lidx, cp);
    resident_load(P[1], a.a_ptr, a.a_val, a.a_lidx, cd);
    const T tau = T(a.bv.tau[b]), sigma = T(a.bv.sigma[b]);
    const int since = a.bv.since[b];
    resident_copy<true>(P[0], a.x, xl);
    resident_copy<true>(P[0], a.xbar, xbl);
    resident_copy<true>(P[0], a.xsum, xsl);
    resident_copy<true>(P[0], a.c, cl);
    resident_copy<true>(P[0], a.clo, clol);
    resident_copy<true>(P[0], a.chi, chil);
    resident_copy<true>(P[1], a.s, sl);
    resident_copy<true>(P[1], a.y, yl);
    resident_copy<true>(P[1], a.ysum, ysl);
    resident_copy<true>(P[1], a.rlo, rlol);
    resident_copy<true>(P[1], a.rhi, rhil);
    __syncthreads();
    for (int k = 0; k < a.count; ++k) {
        const int first = (k == 0 && since == 0) ? 1 : 0;
        resident_pass<true>(cp, yl, cl, clol, chil, xl, xbl, xsl, tau, sigma, first);
        __syncthreads();
        resident_pass<false>(cd, xbl, static_cast<const T*>(nullptr), rlol, rhil, sl, yl, ysl, tau, sigma, first);
        __syncthreads();
    }
    resident_copy<false>(P[0], a.x, xl);
    resident_copy<false>(P[0], a.xbar, xbl);
    resident_copy<false>(P[0], a.xsum, xsl);
    resident_copy<false>(P[1], a.s, sl);
    resident_copy<false>(P[1], a.y, yl);
    resident_copy<false>(P[1], a.ysum, ysl);
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

// The Cesaro divisor of a row: 1/since of the run (single LP) or of the
// row's instance (batch).
template <bool B>
__device__ __forceinline__ double row_inv(const BatchView& bv, double inv, int row) {
    if constexpr (B) {
        if (row >= 0) {
            const int32_t k = bv.since[bv.inst[row]];
            return 1.0 / double(k > 0 ? k : 1);
        }
    }
    return inv;
}

// A row's 7 summed and 4 maximised check terms go either into the thread's
// running accumulators (single LP: block partials follow) or, for a batch,
// to per-row arrays (slot-major, stride `stride`) that a segmented reduction
// sums per instance.
__device__ __forceinline__ void emit_terms(const double (&t)[11], double* rowout, int64_t stride, int64_t row,
                                           double (&sm)[7], double (&mx)[4]) {
    if (rowout != nullptr) {
#pragma unroll
        for (int k = 0; k < 11; ++k) rowout[k * stride + row] = t[k];
        return;
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) sm[k] += t[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) mx[k] = nmax(mx[k], t[7 + k]);
}

template <typename T, bool B = false>
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
    double* rowout;   // batch: 11 x rows per-row terms
    double inv, tol_s, tol_o;
    int mode;
    Sched s;
    BatchView bv;
    // a refinement correction (b_act != null): the point checked is the
    // anchor plus idelta times the state's vectors, with A x0 = b_act,
    // y0 = b_y and the LP's scaled row box b_lo, b_hi (FP64)
    const double* b_act = nullptr;
    const double* b_y = nullptr;
    const double* b_lo = nullptr;
    const double* b_hi = nullptr;
    double idelta = 1.0;
    // (x, x_sum) interleaved (k_pack_pairs), or null: gather them separately
    const T* pair = nullptr;
    int64_t pstride = 1;   // slot stride of the block partials (slot-major)
    __device__ int block() const { return (B && bv.list != nullptr) ? bv.list[blockIdx.x] : (int)blockIdx.x; }
    // sums: pv2_c, pv2_a, bn2, yres2_c, yres2_a, dobj_c, dobj_a; maxes: kkt_c, kkt_a, orig_c, orig_a
    template <int G>
    __device__ void run(const ClassView& v, double (&sm)[7], double (&mx)[4]) const {
        int lane;
        const int i = row_of<G>(v, lane);
        const double inv_r = row_inv<B>(bv, inv, i);
        double ac, aa;
        if (pair != nullptr) row_dot2<G, true>(ptr, idx, val, pair, pair, inv_r, i, lane, ac, aa);
        else row_dot2<G>(ptr, idx, val, x, xsum, inv_r, i, lane, ac, aa);
        if (i >= 0 && lane == 0) {
            if (!combine_parts(mode, acc + 2 * (int64_t)i, ac, aa)) return;
            double yc = double(y[i]), ya = double(ysum[i]) * inv_r;
            double ls, hs;
            if (b_act != nullptr) {
                const double a0 = b_act[i], y0 = b_y[i];
                ac = a0 + ac * idelta;
                aa = a0 + aa * idelta;
                yc = y0 + yc * idelta;
                ya = y0 + ya * idelta;
                ls = b_lo[i];
                hs = b_hi[i];
            } else {
                ls = double(lo_s[i]);
                hs = double(hi_s[i]);
            }
            if (act_avg != nullptr) act_avg[i] = aa;
            double t[11];
            t[7] = row_res(ac, ls, hs, yc, tol_s);
            t[8] = row_res(aa, ls, hs, ya, tol_s);
            const double r = rs[i], lo = lo_o[i], hi = hi_o[i];
            const double aoc = ac / r, aoa = aa / r;
            const double yic = yc * r, yia = ya * r;
            t[9] = row_res(aoc, lo, hi, yic, tol_o);
            t[10] = row_res(aoa, lo, hi, yia, tol_o);
            double vc = 0.0, va = 0.0;
            if (finite_b(lo) && aoc < lo) vc = lo - aoc;
            if (finite_b(hi) && aoc > hi) vc = aoc - hi;
            if (finite_b(lo) && aoa < lo) va = lo - aoa;
            if (finite_b(hi) && aoa > hi) va = aoa - hi;
            t[0] = vc * vc;
            t[1] = va * va;
            t[2] = 0.0;
            if (finite_b(lo)) t[2] += lo * lo;
            if (finite_b(hi) && !(finite_b(lo) && lo == hi)) t[2] += hi * hi;
            const double pc = proj_sign(yic, lo, hi), pa = proj_sign(yia, lo, hi);
            t[3] = (yic - pc) * (yic - pc);
            t[4] = (yia - pa) * (yia - pa);
            t[5] = box_inf(pc, lo, hi);
            t[6] = box_inf(pa, lo, hi);
            emit_terms(t, rowout, stride, i, sm, mx);
        }
    }
    int64_t stride;   // rows of the orientation (batch per-row arrays)
};

template <typename T, bool B = false>
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
    double* rowout;
    double inv, tol_s, tol_o, sign;
    int mode;
    Sched s;
    BatchView bv;
    // a refinement correction (b_aty != null): the point checked is the
    // anchor plus idelta times the state's vectors, with A^T y0 = b_aty,
    // x0 = b_x and the LP's scaled objective b_c and column box b_lo, b_hi
    const double* b_aty = nullptr;
    const double* b_x = nullptr;
    const double* b_c = nullptr;
    const double* b_lo = nullptr;
    const double* b_hi = nullptr;
    double idelta = 1.0;
    // (y, y_sum) interleaved (k_pack_pairs), or null: gather them separately
    const T* pair = nullptr;
    int64_t pstride = 1;   // slot stride of the block partials (slot-major)
    __device__ int block() const { return (B && bv.list != nullptr) ? bv.list[blockIdx.x] : (int)blockIdx.x; }
    // sums: dres2_c, dres2_a, cn2, pobj_c, pobj_a, dobj_c, dobj_a; maxes: kkt_c, kkt_a, orig_c, orig_a
    template <int G>
    __device__ void run(const ClassView& v, double (&sm)[7], double (&mx)[4]) const {
        int lane;
        const int j = row_of<G>(v, lane);
        const double inv_r = row_inv<B>(bv, inv, j);
        double tc, ta;
        if (pair != nullptr) row_dot2<G, true>(ptr, idx, val, pair, pair, inv_r, j, lane, tc, ta);
        else row_dot2<G>(ptr, idx, val, y, ysum, inv_r, j, lane, tc, ta);
        if (j >= 0 && lane == 0) {
            if (!combine_parts(mode, acc + 2 * (int64_t)j, tc, ta)) return;
            double t[11];
            double xc = double(x[j]), xa = double(xsum[j]) * inv_r;
            double cj, ls, hs;
            if (b_aty != nullptr) {
                const double t0 = b_aty[j], x0 = b_x[j];
                tc = t0 + tc * idelta;
                ta = t0 + ta * idelta;
                xc = x0 + xc * idelta;
                xa = x0 + xa * idelta;
                cj = b_c[j];
                ls = b_lo[j];
                hs = b_hi[j];
            } else {
                cj = double(c_s[j]);
                ls = double(lo_s[j]);
                hs = double(hi_s[j]);
            }
            t[7] = dual_class(cj - tc, ls, hs, xc, tol_s);
            t[8] = dual_class(cj - ta, ls, hs, xa, tol_s);
            const double sc = cs[j], lo = lo_o[j], hi = hi_o[j];
            const double ci = sign * c_o[j];
            const double xoc = xc * sc, xoa = xa * sc;
            const double rc = ci - tc / sc, ra = ci - ta / sc;
            t[9] = dual_class(rc, lo, hi, xoc, tol_o);
            t[10] = dual_class(ra, lo, hi, xoa, tol_o);
            const double lc = proj_sign(rc, lo, hi), la = proj_sign(ra, lo, hi);
            t[0] = (rc - lc) * (rc - lc);
            t[1] = (ra - la) * (ra - la);
            t[2] = ci * ci;
            t[3] = ci * xoc;
            t[4] = ci * xoa;
            t[5] = box_inf(lc, lo, hi);
            t[6] = box_inf(la, lo, hi);
            emit_terms(t, rowout, stride, j, sm, mx);
        }
    }
    int64_t stride;
};

template <class Op>
struct CheckRunner {
    const Op& op;
    double (&sm)[7];
    double (&mx)[4];
    template <int G>
    __device__ __forceinline__ void run(const ClassView& v) {
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
    const ClassView v = class_view(op.s, op.block());
    CheckRunner<Op> r{op, sm, mx};
    dispatch_G(v, r);
    // mode is uniform over the launch: only a launch that completes rows reduces
    if ((op.mode == kOnly || op.mode == kLast) && op.rowout == nullptr)
        block_reduce_write<7, 4>(sm, mx, op.partial + blockIdx.x, op.pstride);
}

// Block b reduces slot (b % 11) of region (b / 11): slots 0..6 sums, 7..10
// maxima.  A region holds slot k of block partial j at k * (blocks + 1) + j
// (slot-major, so a slot's partials are contiguous).
__global__ void k_check_final(const double* __restrict__ rowpart, int64_t rblocks,
                              const double* __restrict__ colpart, int64_t cblocks, double* __restrict__ out) {
    const int region = blockIdx.x / 11, slot = blockIdx.x % 11;
    const int64_t nb = region == 0 ? rblocks : cblocks;
    const double* p = (region == 0 ? rowpart : colpart) + slot * (nb + 1);
    const bool is_max = slot >= 7;
    double acc = 0.0;
    // thread t folds partials t, t + blockDim, t + 2 blockDim, ... in that
    // order, with the loads of kAhead of them in flight at a time
    constexpr int kAhead = 8;
    for (int64_t b0 = threadIdx.x; b0 < nb; b0 += (int64_t)kAhead * blockDim.x) {
        double v[kAhead];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const int64_t b = b0 + (int64_t)u * blockDim.x;
            v[u] = b < nb ? p[b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u)
            if (b0 + (int64_t)u * blockDim.x < nb) acc = is_max ? nmax(acc, v[u]) : acc + v[u];
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

// dst = base + d * inv (current) and dst_sum = since * base + d_sum * inv
// (running sum) for one vector, in FP64 from a correction of type T.
template <typename T>
__global__ void k_combine(int32_t len, const double* __restrict__ base, const T* __restrict__ d,
                          const T* __restrict__ d_sum, double since, double inv, double* __restrict__ dst,
                          double* __restrict__ dst_sum) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        const double b = base[i];
        dst[i] = b + double(d[i]) * inv;
        dst_sum[i] = since * b + double(d_sum[i]) * inv;
    }
}

// The state's data fields are read-only to the step kernels; a correction
// problem is the one place that writes them.
template <typename T>
static T* writable(const void* p) { return static_cast<T*>(const_cast<void*>(p)); }

// One side of a correction problem around the anchor point p0:
//   box' = delta (box - p0), v = Proj_box'(0),
//   grad' = delta (g - g_sub)  (columns: c - A^T y0; rows: y0 with g_sub = null)
// and, for rows, the dual correction starts at zero (zero != null).
template <typename T>
__global__ void k_anchor(int32_t len, const double* __restrict__ lo, const double* __restrict__ hi,
                         const double* __restrict__ p0, const double* __restrict__ g,
                         const double* __restrict__ g_sub, double delta, T* __restrict__ lo_c, T* __restrict__ hi_c,
                         T* __restrict__ v, T* __restrict__ grad, T* __restrict__ zero) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        const double p = p0[i];
        const T l = T(delta * (lo[i] - p)), h = T(delta * (hi[i] - p));
        lo_c[i] = l;
        hi_c[i] = h;
        T z = T(0);
        if (z < l) z = l;
        if (z > h) z = h;
        v[i] = z;
        grad[i] = T(delta * (g_sub != nullptr ? g[i] - g_sub[i] : g[i]));
        if (zero != nullptr) zero[i] = T(0);
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

// The check results from the 22 reduced values: [0..10] over rows (7 sums,
// 4 maxima), [11..21] over columns.
static void finish_check(const double* red, double sign, double obj_const, sov_hip_check_out* out) {
    const double* r = red;
    const double* q = red + 11;
    out->kkt_cur = nmax_host(r[7], q[7]);
    out->kkt_avg = nmax_host(r[8], q[8]);
    out->orig_cur = nmax_host(r[9], q[9]);
    out->orig_avg = nmax_host(r[10], q[10]);
    const double bn = std::sqrt(r[2]), cn = std::sqrt(q[2]);
    for (int w = 0; w < 2; ++w) {
        const double pv2 = r[0 + w], yres2 = r[3 + w], dobj_r = r[5 + w];
        const double dres2 = q[0 + w], pobj = q[3 + w], dobj_c = q[5 + w];
        const double p = sign * pobj + obj_const;
        const double d = sign * (dobj_r + dobj_c) + obj_const;
        double* rel = w == 0 ? out->rel_cur : out->rel_avg;
        rel[0] = std::sqrt(pv2) / (1.0 + bn);
        rel[1] = std::sqrt(yres2 + dres2) / (1.0 + cn);
        rel[2] = std::fabs(p - d) / (1.0 + std::fabs(p) + std::fabs(d));
        if (w == 0) { out->pobj_cur = p; out->dobj_cur = d; }
        else { out->pobj_avg = p; out->dobj_avg = d; }
    }
}

template <typename T>
static const T* vals_of(const Part& p) {
    return std::is_same<T, double>::value ? reinterpret_cast<const T*>(p.val64) : reinterpret_cast<const T*>(p.val32);
}

// Block partials the check writes: one set per block of the launch that
// completes the rows of each orientation (its last part), plus one spare each.
static int64_t check_blocks(const sov_hip_csr* h) {
    return (int64_t)h->a.part(h->a.nparts() - 1).blocks() + (int64_t)h->t.part(h->t.nparts() - 1).blocks() + 2;
}

// The allocation-free check (sov_hip_pdhg_check_static) keeps its block
// partials and its 22 reduced values in this file-scope device array, one
// copy per GPU, held until the process ends: kStaticCheckBlocks blocks of
// kCheckSlots doubles (8 MiB) and 64 doubles after them.  Host code must not
// pass `scratch` itself to a kernel (on the host it names a stand-in, not
// device memory): it uses the address hipGetSymbolAddress gives for the
// current device, taken once per device by static_scratch_address() when a
// context is created on it.
static __device__ __align__(256) double scratch[kStaticCheckBlocks * kCheckSlots + 64];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxScratchDevices] = {};
}  // namespace

int static_scratch_address(int device, double** addr) {
    *addr = nullptr;
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(SOV_UNSUPPORTED, "static scratch: device ordinal " + std::to_string(device) + " out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device] == nullptr) {
        // the caller has made `device` current
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_addr[device] = static_cast<double*>(p);
    }
    *addr = g_scratch_addr[device];
    return SOV_OK;
}

// out[2 j] = a[j], out[2 j + 1] = b[j]: a vector and its running sum as
// pairs, so a check's gather of both is one load.
template <typename T>
__global__ void k_pack_pairs(int32_t n, const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out) {
    typedef T T2 __attribute__((ext_vector_type(2)));
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        T2 q;
        q.x = a[j];
        q.y = b[j];
        reinterpret_cast<T2*>(out)[j] = q;
    }
}

// partials: the block partials of both orientations, (rb + cb + 2) *
// kCheckSlots doubles, checked by the caller; red: 22 doubles; pairs:
// 2 (n + m) elements of T for the interleaved vectors (null: the check
// gathers each vector separately, with the same sums).
template <typename T>
static int check_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_check_args* a, sov_hip_check_out* out,
                      double* partials, double* red, void* pairs) {
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
    double* rowpart = partials;
    double* colpart = partials + (int64_t)(rb + 1) * kCheckSlots;
    T* xpair = nullptr;
    T* ypair = nullptr;
    if (pairs != nullptr) {
        xpair = static_cast<T*>(pairs);
        ypair = xpair + 2 * (int64_t)h->n;
        if (h->n > 0)
            k_pack_pairs<T><<<launch_blocks(h->n), kBlock, 0, ctx->stream>>>(
                h->n, static_cast<const T*>(st->x), static_cast<const T*>(st->x_sum), xpair);
        if (h->m > 0)
            k_pack_pairs<T><<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(
                h->m, static_cast<const T*>(st->y), static_cast<const T*>(st->y_sum), ypair);
    }
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
        op.pair = xpair;
        op.pstride = rb + 1;
        if (a->base != nullptr) {
            op.b_act = a->base_act; op.b_y = a->base_y; op.idelta = a->inv_delta;
            op.b_lo = static_cast<const double*>(a->base->row_lo); op.b_hi = static_cast<const double*>(a->base->row_hi);
        }
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
        op.pair = ypair;
        op.pstride = cb + 1;
        if (a->base != nullptr) {
            op.b_aty = a->base_aty; op.b_x = a->base_x; op.idelta = a->inv_delta;
            op.b_c = static_cast<const double*>(a->base->c);
            op.b_lo = static_cast<const double*>(a->base->col_lo); op.b_hi = static_cast<const double*>(a->base->col_hi);
        }
        k_check<CheckColOp<T>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    k_check_final<<<22, kBlock, 0, ctx->stream>>>(rowpart, rb, colpart, cb, red);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, red, 22 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    finish_check(ctx->h_pinned, a->obj_sign, a->obj_const, out);
    return SOV_OK;
}

// ---------------------------------------------------------------------------
// batches of LPs (block-diagonal in one CSR)
// ---------------------------------------------------------------------------


// Reduces each instance's 11 per-row terms of both dimensions: slots 0..6
// summed, 7..10 maximised; out[b * 22 + dim * 11 + slot].  Instances not
// iterating in the checked state (active[b] != want) are skipped.  One
// workgroup per chunk of the chunk table (batch_reduce_chunks):
// * a whole chunk (c < 0) is all of instance b's entries of one dimension:
//   one sweep, each term keeping its own accumulation order (classes, then
//   entries, ascending) and reduction tree, written straight to out;
// * an instance dimension with more than kReduceChunk entries is split
//   into chunks of at most kReduceChunk entries of one class, so that a
//   large instance is read by many CUs at once.  Each chunk's block result
//   goes to its slot of part with device-scope stores; the chunk that
//   arrives last at the instance dimension's counter resets it for the next
//   launch and adds the slots in table order (classes, then entries,
//   ascending) with device-scope loads, so the sums do not depend on the
//   arrival order.
constexpr int kReduceChunk = 8192;
constexpr int kReduceRec = 8;   // b, dim, c, i0, i1, slot, chunks of (b, dim), first slot

__global__ void __launch_bounds__(kBlock) k_batch_reduce(const int32_t* __restrict__ seg,
                                                         const double* __restrict__ rows_t, int64_t m,
                                                         const double* __restrict__ cols_t, int64_t n,
                                                         const int32_t* __restrict__ active, int want,
                                                         const int32_t* __restrict__ chunks,
                                                         double* __restrict__ part, int32_t* __restrict__ cnt,
                                                         double* __restrict__ out) {
    const int32_t* rec = chunks + (int64_t)blockIdx.x * kReduceRec;
    const int b = rec[0];
    if (active[b] != want) return;   // uniform per block
    const int dim = rec[1];
    const double* t = dim == 0 ? rows_t : cols_t;
    const int64_t stride = dim == 0 ? m : n;
    double sm[7], mx[4];
#pragma unroll
    for (int k = 0; k < 7; ++k) sm[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) mx[k] = 0.0;
    auto add = [&](int64_t at) {
        double v[11];
#pragma unroll
        for (int k = 0; k < 11; ++k) v[k] = t[k * stride + at];
#pragma unroll
        for (int k = 0; k < 7; ++k) sm[k] = sm[k] + v[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) mx[k] = nmax(mx[k], v[7 + k]);
    };
    if (rec[2] < 0) {
        for (int c = 0; c < kClasses; ++c) {
            const int2 r = batch_range(seg, b, dim, c);
            for (int i = threadIdx.x; i < r.y; i += kBlock) add(r.x + (int64_t)i);
        }
        block_reduce_write<7, 4>(sm, mx, out + (int64_t)b * 22 + dim * 11);
        return;
    }
    const int2 r = batch_range(seg, b, dim, rec[2]);
    for (int i = rec[3] + threadIdx.x; i < rec[4]; i += kBlock) add(r.x + (int64_t)i);
    __shared__ double res[11];
    __shared__ int arrived;
    block_reduce_write<7, 4>(sm, mx, res);
    __syncthreads();
    if (threadIdx.x < 11)
        __hip_atomic_store(part + (int64_t)rec[5] * 11 + threadIdx.x, res[threadIdx.x], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);   // the stores have completed before the block counts itself in
    __syncthreads();
    int32_t* ctr = cnt + (int64_t)b * 2 + dim;
    if (threadIdx.x == 0) arrived = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (arrived != rec[6] - 1) return;
    if (threadIdx.x == 0) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 11) {
        const int k = threadIdx.x;
        double acc = 0.0;
        for (int j = 0; j < rec[6]; ++j) {
            const double v = __hip_atomic_load(part + ((int64_t)rec[7] + j) * 11 + k, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
            acc = k < 7 ? acc + v : nmax(acc, v);
        }
        out[(int64_t)b * 22 + dim * 11 + k] = acc;
    }
}

// out[b] = sum of v^2 over instance b's entries of dimension dim.
__global__ void __launch_bounds__(kBlock) k_batch_sumsq(const int32_t* __restrict__ seg, int dim,
                                                        const double* __restrict__ v, double* __restrict__ out) {
    const int b = blockIdx.x;
    double acc = 0.0;
    for (int c = 0; c < kClasses; ++c) {
        const int2 r = batch_range(seg, b, dim, c);
        for (int i = threadIdx.x; i < r.y; i += kBlock) {
            const double t = v[r.x + i];
            acc += t * t;
        }
    }
    double sm[1] = {acc}, mx[1] = {0.0};
    block_reduce_write<1, 0>(sm, mx, out + b);
}

__global__ void k_batch_scale(int32_t len, const int32_t* __restrict__ inst, const double* __restrict__ fac,
                              double* __restrict__ v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        v[i] *= fac[inst[i]];
}

template <typename T>
__global__ void k_batch_restart_x(int32_t n, const int32_t* __restrict__ inst, const int32_t* __restrict__ flags,
                                  const int32_t* __restrict__ since, const T* __restrict__ xsum,
                                  const T* __restrict__ lo, const T* __restrict__ hi, T* __restrict__ x) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const int b = inst[j];
        if (!flags[b]) continue;
        const int32_t k = since[b];
        double v = double(xsum[j]) * (1.0 / double(k > 0 ? k : 1));
        const double l = double(lo[j]), h = double(hi[j]);
        if (v < l) v = l;
        if (v > h) v = h;
        x[j] = T(v);
    }
}

template <typename T>
__global__ void k_batch_restart_y(int32_t m, const int32_t* __restrict__ inst, const int32_t* __restrict__ flags,
                                  const int32_t* __restrict__ since, const T* __restrict__ ysum,
                                  const double* __restrict__ act, const T* __restrict__ lo,
                                  const T* __restrict__ hi, T* __restrict__ y, T* __restrict__ s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = inst[i];
        if (!flags[b]) continue;
        const int32_t k = since[b];
        y[i] = T(double(ysum[i]) * (1.0 / double(k > 0 ? k : 1)));
        double a = act[i];
        const double l = double(lo[i]), h = double(hi[i]);
        if (a < l) a = l;
        if (a > h) a = h;
        s[i] = T(a);
    }
}

// Activity bits of every block of one part's step launch (see sov_hip_batch::blk).
__global__ void k_block_flags(Sched s, int32_t blocks, const int32_t* __restrict__ inst,
                              const int32_t* __restrict__ active, const int32_t* __restrict__ m32,
                              const int32_t* __restrict__ m64, int32_t* __restrict__ flags) {
    for (int64_t blk = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; blk < blocks;
         blk += (int64_t)gridDim.x * blockDim.x) {
        int c = 0;
#pragma unroll
        for (int k = 1; k < kClasses; ++k)
            if (blk >= pick(s.block_begin, k)) c = k;
        const int G = pick(s.G, c), rpb = kBlock / G;
        const int64_t local0 = (blk - pick(s.block_begin, c)) * (int64_t)rpb;
        const int64_t cnt = min((int64_t)rpb, (int64_t)pick(s.count, c) - local0);
        const int64_t r0 = pick(s.offset, c) + local0;
        int f = 0;
        for (int64_t r = 0; r < cnt; ++r) {
            const int b = inst[r0 + r];
            const int a = active[b];
            // bits 0 / 1: a row the CSR step launch of FP32 / FP64 runs (not
            // one of a resident instance); bits 2 / 3: a row iterating in
            // FP32 / FP64 (the checks)
            if (a == 1 && (m32 == nullptr || !m32[b])) f |= 1;
            if (a == 2 && (m64 == nullptr || !m64[b])) f |= 2;
            if (a == 1) f |= 4;
            if (a == 2) f |= 8;
        }
        flags[blk] = f;
    }
}

// keep[i] = 1 when bit `bit` of flags[i] is set.
__global__ void k_flag_bit(int32_t n, const int32_t* __restrict__ flags, int bit, int32_t* __restrict__ keep) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        keep[i + 1] = (flags[i] >> bit) & 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) keep[0] = 0;
}

// list[pos[i]] = i for the kept blocks (pos = exclusive prefix sum, scan[i]).
__global__ void k_compact(int32_t n, const int32_t* __restrict__ scan, int32_t* __restrict__ list) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (scan[i + 1] != scan[i]) list[scan[i]] = (int32_t)i;
}

// (Re)build the resident instances when A's or A^T's single part is not the
// one their indices were built from (or stage_bytes changed: built_from
// reset).  Host side, from the parts' device-order CSRs.
static int batch_resident_build(sov_hip_ctx* ctx, sov_hip_batch* bt) {
    const sov_hip_csr* h = bt->csr;
    auto& R = bt->res;
    const bool one = h->a.nparts() == 1 && h->t.nparts() == 1;
    if (one && R.built_from[0] == h->a.parts_version && R.built_from[1] == h->t.parts_version) return SOV_OK;
    for (int q = 0; q < 2; ++q) {
        for (void** p : {(void**)&R.lidx[q], (void**)&R.mask[q], (void**)&R.list[q]}) {
            (void)hipFree(*p);
            *p = nullptr;
        }
        R.h_inst[q].clear();
        R.nlist[q] = 0;
        R.lds[q] = 0;
    }
    R.built_from[0] = R.built_from[1] = -1;
    if (!one) return SOV_OK;
    const int count = bt->count;
    auto run = [&](int b, int dim, int c, int k) {
        return bt->h_seg[((static_cast<size_t>(b) * 2 + dim) * kClasses + c) * 2 + k];
    };
    // the device-order CSRs of A (o = 0) and A^T (o = 1) on the host
    std::vector<int32_t> hptr[2], hidx[2];
    for (int o = 0; o < 2; ++o) {
        const Part& part = (o == 0 ? h->a : h->t).part(0);
        const int32_t rows = (o == 0 ? h->a : h->t).nat.rows, nnz = part.nnz;
        hptr[o].assign(static_cast<size_t>(rows) + 1, 0);
        hidx[o].assign(static_cast<size_t>(std::max(nnz, 1)), 0);
        SOV_HIP_TRY(hipMemcpyAsync(hptr[o].data(), part.ptr, sizeof(int32_t) * hptr[o].size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        if (nnz > 0)
            SOV_HIP_TRY(hipMemcpyAsync(hidx[o].data(), part.idx, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost,
                                       ctx->stream));
    }
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    // does instance b's pass over orientation o fit the kernel's registers:
    // at most `slots` chunks per wave, `entries` entries per lane
    auto fits = [&](int b, int o, int slots, int entries) {
        const int rdim = o == 0 ? 0 : 1;
        int64_t chunks = 0;
        for (int c = 0; c < kClasses; ++c) {
            const int32_t r0 = run(b, rdim, c, 0), rl = run(b, rdim, c, 1), G = kDefaultLanes[c];
            chunks += (static_cast<int64_t>(rl) * G + kWave - 1) / kWave;
            for (int32_t r = r0; r < r0 + rl; ++r) {
                const int32_t len = hptr[o][static_cast<size_t>(r) + 1] - hptr[o][static_cast<size_t>(r)];
                if ((len + G - 1) / G > entries) return false;
            }
        }
        return chunks <= static_cast<int64_t>(kResidentWaves) * slots;
    };
    std::vector<int32_t> mask[2];
    for (int q = 0; q < 2; ++q) mask[q].assign(static_cast<size_t>(count), 0);
    for (int b = 0; b < count; ++b) {
        int64_t nb = 0, mb = 0;
        for (int c = 0; c < kClasses; ++c) {
            nb += run(b, 1, c, 1);
            mb += run(b, 0, c, 1);
        }
        if (nb <= 0 || mb <= 0 || nb > 65535 || mb > 65535) continue;
        if (!fits(b, 1, kResidentSlotsP, kResidentEntriesP) || !fits(b, 0, kResidentSlotsD, kResidentEntriesD)) continue;
        for (int q = 0; q < 2; ++q) {
            // the iterates, objective and bounds, beside the static pass tables
            const int64_t need = resident_lds_bytes(nb, mb, q == 0 ? 4 : 8);
            if (need > std::min<int64_t>(bt->stage_bytes, bt->lds_max - 2 * (int64_t)sizeof(ResidentPass))) continue;
            mask[q][static_cast<size_t>(b)] = 1;
            R.h_inst[q].push_back(b);
            R.lds[q] = std::max(R.lds[q], need);
        }
    }
    if (!R.h_inst[0].empty()) {
        // local indices of both orientations (FP32 residency covers FP64's)
        for (int o = 0; o < 2; ++o) {
            const Part& part = (o == 0 ? h->a : h->t).part(0);
            const int32_t nnz = part.nnz;
            const int rdim = o == 0 ? 0 : 1, gdim = 1 - rdim;
            const std::vector<int32_t>& ptr = hptr[o];
            const std::vector<int32_t>& idx = hidx[o];
            std::vector<uint16_t> lidx(static_cast<size_t>(std::max(nnz, 1)), 0);
            for (int b : R.h_inst[0]) {
                int32_t gs[kClasses], gl[kClasses], go[kClasses];
                int32_t off = 0;
                for (int c = 0; c < kClasses; ++c) {
                    gs[c] = run(b, gdim, c, 0);
                    gl[c] = run(b, gdim, c, 1);
                    go[c] = off;
                    off += gl[c];
                }
                for (int c = 0; c < kClasses; ++c) {
                    const int32_t r0 = run(b, rdim, c, 0), rl = run(b, rdim, c, 1);
                    for (int32_t r = r0; r < r0 + rl; ++r)
                        for (int32_t q = ptr[static_cast<size_t>(r)]; q < ptr[static_cast<size_t>(r) + 1]; ++q) {
                            const int32_t j = idx[static_cast<size_t>(q)];
                            int k = 0;
                            while (k < kClasses && !(gl[k] > 0 && j >= gs[k] && j < gs[k] + gl[k])) ++k;
                            if (k == kClasses)
                                return fail(SOV_INVALID_ARGUMENT, "batch: an entry outside its instance's block");
                            lidx[static_cast<size_t>(q)] = static_cast<uint16_t>(go[k] + (j - gs[k]));
                        }
                }
            }
            SOV_HIP_TRY(hipMalloc(&R.lidx[o], sizeof(uint16_t) * lidx.size()));
            SOV_HIP_TRY(hipMemcpyAsync(R.lidx[o], lidx.data(), sizeof(uint16_t) * lidx.size(), hipMemcpyHostToDevice,
                                       ctx->stream));
            SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));   // lidx is freed next
        }
        for (int q = 0; q < 2; ++q) {
            SOV_HIP_TRY(hipMalloc(&R.mask[q], sizeof(int32_t) * count));
            SOV_HIP_TRY(hipMemsetAsync(R.mask[q], 0, sizeof(int32_t) * count, ctx->stream));
            SOV_HIP_TRY(hipMalloc(&R.list[q], sizeof(int32_t) * std::max<size_t>(R.h_inst[q].size(), 1)));
        }
    }
    for (int q = 0; q < 2; ++q) {
        R.h_mask[q] = std::move(mask[q]);
        R.on[q] = 0;   // the device masks start all zero
    }
    R.built_from[0] = h->a.parts_version;
    R.built_from[1] = h->t.parts_version;
    return SOV_OK;
}

// The resident instances iterating in each precision (h_active), and the
// device masks the step launches skip.
static int batch_resident_lists(sov_hip_ctx* ctx, sov_hip_batch* bt) {
    auto& R = bt->res;
    for (int q = 0; q < 2; ++q) {
        // resident only while every instance iterating in precision q fits:
        // then its intervals make no step launch at all (a resident launch
        // beside step launches of the same precision measured slower)
        std::vector<int32_t>& keep = R.h_list[q];   // outlives the copy (the caller synchronises)
        keep.clear();
        bool all = R.mask[q] != nullptr;
        for (size_t b = 0; b < bt->h_active.size() && all; ++b)
            if (bt->h_active[b] == q + 1) {
                if (R.h_mask[q][b]) keep.push_back(static_cast<int32_t>(b));
                else all = false;
            }
        if (!all) keep.clear();
        const int on = keep.empty() ? 0 : 1;
        // ordered on the context stream, before the block flags that read it
        if (R.mask[q] != nullptr && on != R.on[q]) {
            if (on) {
                SOV_HIP_TRY(hipMemcpyAsync(R.mask[q], R.h_mask[q].data(), sizeof(int32_t) * R.h_mask[q].size(),
                                           hipMemcpyHostToDevice, ctx->stream));
            } else {
                SOV_HIP_TRY(hipMemsetAsync(R.mask[q], 0, sizeof(int32_t) * R.h_mask[q].size(), ctx->stream));
            }
            R.on[q] = on;
        }
        R.nlist[q] = static_cast<int32_t>(keep.size());
        if (!keep.empty())
            SOV_HIP_TRY(hipMemcpyAsync(R.list[q], keep.data(), sizeof(int32_t) * keep.size(), hipMemcpyHostToDevice,
                                       ctx->stream));
    }
    return SOV_OK;
}

// Per precision: tiled step launches while its iterating instances hold at
// least tile_frac of the entries (h_active).
static void batch_tile_choice(sov_hip_batch* bt) {
    int64_t act[2] = {0, 0};
    for (size_t b = 0; b < bt->h_nnz.size(); ++b) {
        const int a = bt->h_active[b];
        if (a == 1 || a == 2) act[a - 1] += bt->h_nnz[b];
    }
    for (int q = 0; q < 2; ++q)
        bt->tiles_on[q] = bt->tile_frac >= 0.0 && bt->nnz_total > 0 && act[q] > 0 &&
                          static_cast<double>(act[q]) >= bt->tile_frac * static_cast<double>(bt->nnz_total);
}

// (Re)compute the block activity bits of every part of both orientations.
static int batch_block_flags(sov_hip_ctx* ctx, sov_hip_batch* bt) {
    const sov_hip_csr* h = bt->csr;
    {
        int rc = batch_resident_build(ctx, bt);
        if (rc == SOV_OK) rc = batch_resident_lists(ctx, bt);
        if (rc != SOV_OK) return rc;
    }
    batch_tile_choice(bt);
    for (int o = 0; o < 2; ++o) {
        const Orient& orient = o == 0 ? h->a : h->t;
        const int np = orient.nparts();
        if ((int)bt->blk[o].size() < np) {
            bt->blk[o].resize(np, nullptr);
            bt->list[o].resize(4 * np, nullptr);
            bt->nlist[o].resize(4 * np, 0);
            bt->blk_cap[o].resize(np, 0);
        }
        for (int q = 0; q < np; ++q) {
            const Sched sch = orient.part(q).sched();
            const int32_t blocks = sch.block_begin[kClasses];
            if (blocks > bt->blk_cap[o][q]) {
                for (int32_t** p : {&bt->blk[o][q], &bt->list[o][4 * q], &bt->list[o][4 * q + 1],
                                    &bt->list[o][4 * q + 2], &bt->list[o][4 * q + 3]}) {
                    (void)hipFree(*p);
                    *p = nullptr;
                }
                // room for every lane choice: at most one block per 4 rows plus one per class
                const int64_t cap = std::max<int64_t>(blocks, (int64_t)orient.nat.rows / 4 + kClasses + 1);
                SOV_HIP_TRY(hipMalloc(&bt->blk[o][q], sizeof(int32_t) * cap));
                for (int bit = 0; bit < 4; ++bit)
                    SOV_HIP_TRY(hipMalloc(&bt->list[o][4 * q + bit], sizeof(int32_t) * cap));
                bt->blk_cap[o][q] = cap;
            }
            if (blocks + 1 > bt->scan_cap) {
                (void)hipFree(bt->scan);
                bt->scan = nullptr;
                bt->scan_cap = 0;
                SOV_HIP_TRY(hipMalloc(&bt->scan, sizeof(int32_t) * (blocks + 1)));
                bt->scan_cap = blocks + 1;
            }
            for (int bit = 0; bit < 4; ++bit) bt->nlist[o][4 * q + bit] = 0;
            if (blocks == 0) continue;
            k_block_flags<<<launch_blocks(blocks), kBlock, 0, ctx->stream>>>(
                sch, blocks, o == 0 ? bt->row_inst : bt->col_inst, bt->active, bt->res.mask[0], bt->res.mask[1],
                bt->blk[o][q]);
            for (int bit = 0; bit < 4; ++bit) {
                k_flag_bit<<<launch_blocks(blocks), kBlock, 0, ctx->stream>>>(blocks, bt->blk[o][q], bit, bt->scan);
                SOV_HIP_TRY(check_launch());
                const int rc = inclusive_scan(ctx, bt->scan, (int64_t)blocks + 1);
                if (rc != SOV_OK) return rc;
                k_compact<<<launch_blocks(blocks), kBlock, 0, ctx->stream>>>(blocks, bt->scan, bt->list[o][4 * q + bit]);
                SOV_HIP_TRY(check_launch());
                int32_t cnt = 0;
                SOV_HIP_TRY(hipMemcpyAsync(&cnt, bt->scan + blocks, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
                SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
                bt->nlist[o][4 * q + bit] = cnt;
            }
        }
    }
    SOV_HIP_TRY(check_launch());
    ++bt->version;
    return SOV_OK;
}

// check: the lists of every block with a row iterating in T (the checks
// cover every instance); else those the CSR step launches run (resident
// instances left out).
template <typename T>
static BatchView batch_view(const sov_hip_batch* bt, bool rows, int k, int part = -1, bool check = false) {
    BatchView v;
    v.inst = rows ? bt->row_inst : bt->col_inst;
    v.tau = bt->tau;
    v.sigma = bt->sigma;
    v.active = bt->active;
    v.since = bt->since;
    v.k = k;
    v.want = std::is_same<T, float>::value ? 1 : 2;
    if (part >= 0) v.list = bt->list[rows ? 0 : 1][4 * part + (check ? 2 : 0) + (v.want - 1)];
    return v;
}

template <typename T>
static int batch_blocks(const sov_hip_batch* bt, bool rows, int part, bool check = false) {
    return bt->nlist[rows ? 0 : 1][4 * part + (check ? 2 : 0) + (std::is_same<T, float>::value ? 0 : 1)];
}

// Move flagged instances from an FP32 state to the FP64 one: their candidate
// (the current point, or the Cesaro average when use_avg) becomes both the
// FP64 point and its running sum (a restart with since = 1 then lands on it).
template <bool Rows>
__global__ void k_batch_promote(int32_t len, const int32_t* __restrict__ inst, const int32_t* __restrict__ flags,
                                const int32_t* __restrict__ use_avg, const int32_t* __restrict__ since,
                                const float* __restrict__ cur32, const float* __restrict__ sum32,
                                double* __restrict__ cur64, double* __restrict__ sum64) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = inst[i];
        if (!flags[b]) continue;
        const int32_t k = since[b];
        const double v = use_avg[b] ? double(sum32[i]) * (1.0 / double(k > 0 ? k : 1)) : double(cur32[i]);
        cur64[i] = v;
        sum64[i] = v;
    }
}

template <typename T>
static int batch_check_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt,
                            const sov_hip_check_args* a, const double* obj_const, sov_hip_check_out* out) {
    const sov_hip_pdhg_state* st = a->st;
    const Orient& OA = h->a;
    const Orient& OT = h->t;
    const int npa = OA.nparts(), npt = OT.nparts();
    for (const Orient* o : {&OA, &OT})
        for (const Part& p : o->parts)
            if (p.nnz > 0 && vals_of<T>(p) == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch check: value image missing");
    double* rows_t = bt->terms;
    double* cols_t = bt->terms + (int64_t)11 * bt->m;
    for (int k = 0; k < npa; ++k) {
        const Part& p = OA.part(k);
        const Sched s = p.sched();
        const int blocks = batch_blocks<T>(bt, true, k, true);
        if (blocks == 0) continue;
        CheckRowOp<T, true> op{};
        op.ptr = p.ptr; op.idx = p.idx; op.val = vals_of<T>(p);
        op.x = static_cast<const T*>(st->x); op.xsum = static_cast<const T*>(st->x_sum);
        op.y = static_cast<const T*>(st->y); op.ysum = static_cast<const T*>(st->y_sum);
        op.lo_s = static_cast<const T*>(st->row_lo); op.hi_s = static_cast<const T*>(st->row_hi);
        op.lo_o = a->orig_row_lo; op.hi_o = a->orig_row_hi; op.rs = a->row_scale;
        op.act_avg = a->act_avg; op.acc = OA.chk_acc; op.rowout = rows_t; op.stride = bt->m;
        op.inv = 1.0; op.tol_s = a->tol_scaled; op.tol_o = a->tol_orig; op.mode = seg_mode(k, npa); op.s = s;
        op.bv = batch_view<T>(bt, true, 0, k, true);
        k_check<CheckRowOp<T, true>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    for (int k = 0; k < npt; ++k) {
        const Part& p = OT.part(k);
        const Sched s = p.sched();
        const int blocks = batch_blocks<T>(bt, false, k, true);
        if (blocks == 0) continue;
        CheckColOp<T, true> op{};
        op.ptr = p.ptr; op.idx = p.idx; op.val = vals_of<T>(p);
        op.y = static_cast<const T*>(st->y); op.ysum = static_cast<const T*>(st->y_sum);
        op.x = static_cast<const T*>(st->x); op.xsum = static_cast<const T*>(st->x_sum);
        op.c_s = static_cast<const T*>(st->c);
        op.lo_s = static_cast<const T*>(st->col_lo); op.hi_s = static_cast<const T*>(st->col_hi);
        op.lo_o = a->orig_col_lo; op.hi_o = a->orig_col_hi; op.c_o = a->orig_c; op.cs = a->col_scale;
        op.acc = OT.chk_acc; op.rowout = cols_t; op.stride = bt->n;
        op.inv = 1.0; op.tol_s = a->tol_scaled; op.tol_o = a->tol_orig; op.sign = a->obj_sign;
        op.mode = seg_mode(k, npt); op.s = s;
        op.bv = batch_view<T>(bt, false, 0, k, true);
        k_check<CheckColOp<T, true>><<<blocks, kBlock, 0, ctx->stream>>>(op);
    }
    k_batch_reduce<<<bt->n_chunks, kBlock, 0, ctx->stream>>>(bt->seg, rows_t, bt->m, cols_t, bt->n, bt->active,
                                                             std::is_same<T, float>::value ? 1 : 2, bt->chunks,
                                                             bt->chunk_part, bt->chunk_cnt, bt->red);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(bt->h_red, bt->red, sizeof(double) * 22 * bt->count, hipMemcpyDeviceToHost,
                               ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < bt->count; ++b)
        finish_check(bt->h_red + (int64_t)b * 22, a->obj_sign, obj_const ? obj_const[b] : 0.0, out + b);
    return SOV_OK;
}

}  // namespace sovhip

using namespace sovhip;

namespace {

template <typename T>
T* orient_acc(const Orient& o) {
    return std::is_same<T, double>::value ? reinterpret_cast<T*>(o.acc64) : reinterpret_cast<T*>(o.acc32);
}

template <typename T, class Op>
int launch_tile_step(sov_hip_ctx* ctx, const Orient& o, Op op, const T* x) {
    const Tiles& t = o.tiles_of(sizeof(T) == 8);
    SOV_HIP_TRY(tile_lds_attr(ctx, reinterpret_cast<const void*>(&k_tile_step<T, Op>), sizeof(T)));
    const size_t lds = tile_lds_bytes(t.W, t.rmax, sizeof(T));
    const int np = sizeof(T) == 8 ? t.parts64 : t.parts;
    for (int k = 0; k < np; ++k) {
        op.mode = seg_mode(k, np);
        const TileArgs ta = tile_args(t, o.nat.rows, o.nat.cols, sizeof(T) == 8, k, np);
        k_tile_step<T, Op><<<t.nb * t.Q, kTileThreads, lds, ctx->stream>>>(op, ta, x);
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

template <typename T>
int primal_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_pdhg_state* st, double tau, int first) {
    const Orient& o = h->t;
    if (tiles_usable(ctx, o.tiles_of(sizeof(T) == 8), sizeof(T) == 8, st->y)) {
        PrimalOp<T> op{nullptr, nullptr, nullptr, static_cast<const T*>(st->y), static_cast<const T*>(st->c),
                       static_cast<const T*>(st->col_lo), static_cast<const T*>(st->col_hi), static_cast<T*>(st->x),
                       static_cast<T*>(st->x_bar), static_cast<T*>(st->x_sum), orient_acc<T>(o), T(tau), first,
                       kOnly, Sched{}};
        return launch_tile_step<T>(ctx, o, op, static_cast<const T*>(st->y));
    }
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
    if (tiles_usable(ctx, o.tiles_of(sizeof(T) == 8), sizeof(T) == 8, st->x_bar)) {
        DualOp<T> op{nullptr, nullptr, nullptr, static_cast<const T*>(st->x_bar), static_cast<const T*>(st->row_lo),
                     static_cast<const T*>(st->row_hi), static_cast<T*>(st->s), static_cast<T*>(st->y),
                     static_cast<T*>(st->y_sum), orient_acc<T>(o), T(tau), T(sigma), first, kOnly, Sched{}};
        op.sg = static_cast<const T*>(st->s_grad);
        return launch_tile_step<T>(ctx, o, op, static_cast<const T*>(st->x_bar));
    }
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
        op.sg = static_cast<const T*>(st->s_grad);
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

// The batch's own stream, forked from the context stream at its first use
// since the last join (sov_hip_batch_resident_join joins it back).
static hipStream_t batch_side(sov_hip_ctx* ctx, const sov_hip_batch* bt, hipError_t* err) {
    *err = hipSuccess;
    if (!bt->forked) {
        *err = hipEventRecord(bt->ev_fork, ctx->stream);
        if (*err == hipSuccess) *err = hipStreamWaitEvent(bt->side, bt->ev_fork, 0);
        if (*err != hipSuccess) return nullptr;
        bt->forked = true;
    }
    return bt->side;
}

// The context stream waits for what the batch's own stream holds (a no-op
// when nothing was forked since the last join).
static int batch_join(sov_hip_ctx* ctx, const sov_hip_batch* bt) {
    if (!bt->forked) return SOV_OK;
    SOV_HIP_TRY(hipEventRecord(bt->ev_join, bt->side));
    SOV_HIP_TRY(hipStreamWaitEvent(ctx->stream, bt->ev_join, 0));
    bt->forked = false;
    return SOV_OK;
}

// The stream of a batch step launch: the batch's own stream for the FP64
// state when the batch splits its precisions, else the context stream.
template <typename T>
static hipStream_t batch_step_stream(sov_hip_ctx* ctx, const sov_hip_batch* bt, hipError_t* err) {
    *err = hipSuccess;
    if (std::is_same<T, double>::value && bt->split) return batch_side(ctx, bt, err);
    return ctx->stream;
}

template <typename T>
int batch_primal_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt, const sov_hip_pdhg_state* st,
                      int k) {
    const Orient& o = h->t;
    constexpr bool f64 = std::is_same<T, double>::value;
    if (bt->tiles_on[f64 ? 1 : 0] && tiles_usable(ctx, o.tiles_of(f64), f64, st->y)) {
        PrimalOp<T, true> op{nullptr, nullptr, nullptr, static_cast<const T*>(st->y), static_cast<const T*>(st->c),
                             static_cast<const T*>(st->col_lo), static_cast<const T*>(st->col_hi),
                             static_cast<T*>(st->x), static_cast<T*>(st->x_bar), static_cast<T*>(st->x_sum),
                             orient_acc<T>(o), T(0), 0, kOnly, Sched{}, batch_view<T>(bt, false, k)};
        op.bv.skip = bt->res.mask[f64 ? 1 : 0];
        return launch_tile_step<T>(ctx, o, op, static_cast<const T*>(st->y));
    }
    const int np = o.nparts();
    for (int q = 0; q < np; ++q) {
        const Part& p = o.part(q);
        const Sched s = p.sched();
        const int blocks = batch_blocks<T>(bt, false, q);
        if (blocks == 0) continue;
        const T* val = vals_of<T>(p);
        if (val == nullptr && p.nnz > 0) return fail(SOV_INVALID_ARGUMENT, "batch primal_step: value image missing");
        PrimalOp<T, true> op{p.ptr, p.idx, val, static_cast<const T*>(st->y), static_cast<const T*>(st->c),
                             static_cast<const T*>(st->col_lo), static_cast<const T*>(st->col_hi),
                             static_cast<T*>(st->x), static_cast<T*>(st->x_bar), static_cast<T*>(st->x_sum),
                             orient_acc<T>(o), T(0), 0, seg_mode(q, np), s, batch_view<T>(bt, false, k, q)};
        op.bv.skip = bt->res.mask[std::is_same<T, float>::value ? 0 : 1];
        hipError_t se;
        const hipStream_t stream = batch_step_stream<T>(ctx, bt, &se);
        SOV_HIP_TRY(se);
        k_step<PrimalOp<T, true>><<<blocks, kBlock, 0, stream>>>(op);
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

template <typename T>
int batch_dual_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt, const sov_hip_pdhg_state* st,
                    int k) {
    const Orient& o = h->a;
    constexpr bool f64 = std::is_same<T, double>::value;
    if (bt->tiles_on[f64 ? 1 : 0] && tiles_usable(ctx, o.tiles_of(f64), f64, st->x_bar)) {
        DualOp<T, true> op{nullptr, nullptr, nullptr, static_cast<const T*>(st->x_bar),
                           static_cast<const T*>(st->row_lo), static_cast<const T*>(st->row_hi),
                           static_cast<T*>(st->s), static_cast<T*>(st->y), static_cast<T*>(st->y_sum),
                           orient_acc<T>(o), T(0), T(0), 0, kOnly, Sched{}, batch_view<T>(bt, true, k)};
        op.bv.skip = bt->res.mask[f64 ? 1 : 0];
        return launch_tile_step<T>(ctx, o, op, static_cast<const T*>(st->x_bar));
    }
    const int np = o.nparts();
    for (int q = 0; q < np; ++q) {
        const Part& p = o.part(q);
        const Sched s = p.sched();
        const int blocks = batch_blocks<T>(bt, true, q);
        if (blocks == 0) continue;
        const T* val = vals_of<T>(p);
        if (val == nullptr && p.nnz > 0) return fail(SOV_INVALID_ARGUMENT, "batch dual_step: value image missing");
        DualOp<T, true> op{p.ptr, p.idx, val, static_cast<const T*>(st->x_bar), static_cast<const T*>(st->row_lo),
                           static_cast<const T*>(st->row_hi), static_cast<T*>(st->s), static_cast<T*>(st->y),
                           static_cast<T*>(st->y_sum), orient_acc<T>(o), T(0), T(0), 0, seg_mode(q, np), s,
                           batch_view<T>(bt, true, k, q)};
        op.bv.skip = bt->res.mask[std::is_same<T, float>::value ? 0 : 1];
        hipError_t se;
        const hipStream_t stream = batch_step_stream<T>(ctx, bt, &se);
        SOV_HIP_TRY(se);
        k_step<DualOp<T, true>><<<blocks, kBlock, 0, stream>>>(op);
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

template <typename T>
int batch_resident_impl(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt, const sov_hip_pdhg_state* st,
                        int count) {
    const int q = std::is_same<T, float>::value ? 0 : 1;
    const auto& R = bt->res;
    if (R.nlist[q] == 0 || count <= 0) return SOV_OK;
    const Part& pa = h->a.part(0);
    const Part& pt = h->t.part(0);
    ResidentArgs<T> a{};
    a.a_ptr = pa.ptr;
    a.a_lidx = R.lidx[0];
    a.a_val = vals_of<T>(pa);
    a.t_ptr = pt.ptr;
    a.t_lidx = R.lidx[1];
    a.t_val = vals_of<T>(pt);
    if (a.a_val == nullptr || a.t_val == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch resident: value image missing");
    a.seg = bt->seg;
    a.list = R.list[q];
    a.x = static_cast<T*>(st->x);
    a.xbar = static_cast<T*>(st->x_bar);
    a.xsum = static_cast<T*>(st->x_sum);
    a.c = static_cast<const T*>(st->c);
    a.clo = static_cast<const T*>(st->col_lo);
    a.chi = static_cast<const T*>(st->col_hi);
    a.y = static_cast<T*>(st->y);
    a.s = static_cast<T*>(st->s);
    a.ysum = static_cast<T*>(st->y_sum);
    a.rlo = static_cast<const T*>(st->row_lo);
    a.rhi = static_cast<const T*>(st->row_hi);
    a.bv = batch_view<T>(bt, true, 0);
    a.count = count;
    for (int c = 0; c < kClasses; ++c) a.ga[c] = a.gt[c] = kDefaultLanes[c];
    SOV_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_resident<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(R.lds[q])));
    hipError_t se;
    const hipStream_t side = batch_side(ctx, bt, &se);   // after everything queued so far
    SOV_HIP_TRY(se);
    k_batch_resident<T><<<R.nlist[q], kResidentThreads, static_cast<size_t>(R.lds[q]), side>>>(a);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int batch_ok(const sov_hip_csr* h, const sov_hip_batch* bt, const sov_hip_pdhg_state* st) {
    if (bt == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch is null");
    if (h != nullptr && h != bt->csr) return fail(SOV_INVALID_ARGUMENT, "batch belongs to another csr handle");
    if (h != nullptr)
        for (int o = 0; o < 2; ++o) {
            const Orient& orient = o == 0 ? h->a : h->t;
            if ((int)bt->blk[o].size() < orient.nparts())
                return fail(SOV_INVALID_ARGUMENT, "batch: segments changed since the last sov_hip_batch_set(active)");
            for (int q = 0; q < orient.nparts(); ++q)
                if (orient.part(q).blocks() > bt->blk_cap[o][q])
                    return fail(SOV_INVALID_ARGUMENT, "batch: lanes changed since the last sov_hip_batch_set(active)");
            // the resident instances index the parts they were built from
            if (bt->res.lidx[o] != nullptr && bt->res.built_from[o] != orient.parts_version)
                return fail(SOV_INVALID_ARGUMENT, "batch: segments changed since the last sov_hip_batch_set(active)");
        }
    if (h != nullptr && (bt->m != h->m || bt->n != h->n)) return fail(SOV_DIMENSION_MISMATCH, "batch dims vs csr");
    if (st != nullptr && (st->m != bt->m || st->n != bt->n)) return fail(SOV_DIMENSION_MISMATCH, "batch dims vs state");
    return SOV_OK;
}

// The check reduction's chunk table (k_batch_reduce): per instance and
// dimension, one whole record when it has at most kReduceChunk entries,
// else one record per kReduceChunk entries of each class, in class order.
// Returns the split chunks' slot count.
int32_t batch_reduce_chunks(const std::vector<int32_t>& seg, int count, std::vector<int32_t>& table) {
    table.clear();
    int32_t slots = 0;
    for (int b = 0; b < count; ++b)
        for (int dim = 0; dim < 2; ++dim) {
            auto len = [&](int c) { return seg[((static_cast<size_t>(b) * 2 + dim) * kClasses + c) * 2 + 1]; };
            int64_t total = 0;
            for (int c = 0; c < kClasses; ++c) total += len(c);
            if (total <= kReduceChunk) {
                table.insert(table.end(), {b, dim, -1, 0, 0, 0, 1, 0});
                continue;
            }
            const int32_t first = slots;
            int32_t nch = 0;
            for (int c = 0; c < kClasses; ++c) nch += (len(c) + kReduceChunk - 1) / kReduceChunk;
            for (int c = 0; c < kClasses; ++c)
                for (int32_t i0 = 0; i0 < len(c); i0 += kReduceChunk)
                    table.insert(table.end(), {b, dim, c, i0, std::min(len(c), i0 + kReduceChunk), slots++, nch, first});
        }
    return slots;
}

void free_batch(sov_hip_batch* bt) {
    for (void* p : {(void*)bt->row_inst, (void*)bt->col_inst, (void*)bt->seg, (void*)bt->tau, (void*)bt->sigma,
                    (void*)bt->active, (void*)bt->since, (void*)bt->flags, (void*)bt->use_avg, (void*)bt->fac,
                    (void*)bt->terms, (void*)bt->red, (void*)bt->chunks, (void*)bt->chunk_part,
                    (void*)bt->chunk_cnt})
        (void)hipFree(p);
    if (bt->h_red) (void)hipHostFree(bt->h_red);
    if (bt->side) {
        (void)hipStreamSynchronize(bt->side);
        (void)hipStreamDestroy(bt->side);
    }
    if (bt->ev_fork) (void)hipEventDestroy(bt->ev_fork);
    if (bt->ev_join) (void)hipEventDestroy(bt->ev_join);
    for (int o = 0; o < 2; ++o) {
        for (int32_t* p : bt->blk[o]) (void)hipFree(p);
        for (int32_t* p : bt->list[o]) (void)hipFree(p);
    }
    (void)hipFree(bt->scan);
    for (int q = 0; q < 2; ++q) {
        (void)hipFree(bt->res.lidx[q]);
        (void)hipFree(bt->res.mask[q]);
        (void)hipFree(bt->res.list[q]);
    }
    delete bt;
}

// Instance of each device-order row of one dimension, and the per-class runs.
int batch_layout(const std::vector<int32_t>& perm, const std::vector<int32_t>& off, const Orient& o, int dim,
                 std::vector<int32_t>& inst_dev, std::vector<int32_t>& seg, int count) {
    const int32_t rows = static_cast<int32_t>(perm.size());
    std::vector<int32_t> inst_nat(static_cast<size_t>(rows));
    for (int b = 0; b < count; ++b)
        for (int32_t r = off[b]; r < off[b + 1]; ++r) inst_nat[static_cast<size_t>(r)] = b;
    inst_dev.resize(static_cast<size_t>(rows));
    for (int32_t p = 0; p < rows; ++p) inst_dev[static_cast<size_t>(p)] = inst_nat[static_cast<size_t>(perm[p])];
    for (int c = 0; c < kClasses; ++c) {
        const int32_t lo = o.class_offset[c], hi = lo + o.class_count[c];
        int32_t prev = -1;
        for (int32_t p = lo; p < hi; ++p) {
            const int32_t b = inst_dev[static_cast<size_t>(p)];
            if (b < prev) return fail(SOV_INVALID_ARGUMENT, "batch: instance rows are not contiguous per class");
            int32_t* e = &seg[((static_cast<size_t>(b) * 2 + dim) * kClasses + c) * 2];
            if (b != prev) e[0] = p;
            ++e[1];
            prev = b;
        }
    }
    return SOV_OK;
}

}  // namespace

extern "C" {

int sov_hip_batch_create(sov_hip_ctx* ctx, const sov_hip_csr* h, int32_t count, const int32_t* row_off,
                         const int32_t* col_off, sov_hip_batch** out) {
    if (ctx == nullptr || h == nullptr || out == nullptr || count < 1 || row_off == nullptr || col_off == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "batch_create args");
    *out = nullptr;
    if (row_off[0] != 0 || col_off[0] != 0 || row_off[count] != h->m || col_off[count] != h->n)
        return fail(SOV_DIMENSION_MISMATCH, "batch_create: offsets do not cover the matrix");
    for (int b = 0; b < count; ++b)
        if (row_off[b + 1] < row_off[b] || col_off[b + 1] < col_off[b])
            return fail(SOV_INVALID_ARGUMENT, "batch_create: offsets must be non-decreasing");
    DeviceGuard3 g(ctx->device);
    std::vector<int32_t> pr(static_cast<size_t>(h->m)), pc(static_cast<size_t>(h->n));
    int rc = sov_hip_csr_order(ctx, h, 0, pr.data());
    if (rc == SOV_OK) rc = sov_hip_csr_order(ctx, h, 1, pc.data());
    if (rc != SOV_OK) return rc;
    std::vector<int32_t> ro(row_off, row_off + count + 1), co(col_off, col_off + count + 1);
    std::vector<int32_t> rinst, cinst, seg(static_cast<size_t>(count) * 2 * kClasses * 2, 0);
    rc = batch_layout(pr, ro, h->a, 0, rinst, seg, count);
    if (rc == SOV_OK) rc = batch_layout(pc, co, h->t, 1, cinst, seg, count);
    if (rc != SOV_OK) return rc;
    auto* bt = new (std::nothrow) sov_hip_batch();
    if (bt == nullptr) return fail(SOV_OUT_OF_MEMORY, "batch");
    bt->count = count;
    bt->m = h->m;
    bt->n = h->n;
    bt->device = ctx->device;
    bt->stream = ctx->stream;
    bt->csr = h;
    bt->h_seg = seg;
    bt->h_active.assign(static_cast<size_t>(count), 2);
    bt->lds_max = ctx->lds_max;
    bt->stage_bytes = ctx->lds_max;
    {
        // entries of each instance: its row runs' spans of A's device-order CSR
        std::vector<int32_t> ptr(static_cast<size_t>(h->m) + 1, 0);
        hipError_t e = hipSuccess;
        if (h->m > 0) {
            e = hipMemcpyAsync(ptr.data(), h->a.d_ptr, sizeof(int32_t) * ptr.size(), hipMemcpyDeviceToHost,
                               ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        }
        if (e != hipSuccess) {
            delete bt;
            return hip_fail(e, "batch_create ptr");
        }
        bt->h_nnz.assign(static_cast<size_t>(count), 0);
        for (int b = 0; b < count; ++b)
            for (int c = 0; c < kClasses; ++c) {
                const int32_t* r = &seg[((static_cast<size_t>(b) * 2) * kClasses + c) * 2];
                if (r[1] > 0) bt->h_nnz[static_cast<size_t>(b)] += ptr[static_cast<size_t>(r[0] + r[1])] - ptr[static_cast<size_t>(r[0])];
            }
        for (int64_t v : bt->h_nnz) bt->nnz_total += v;
    }
    const size_t m1 = std::max<int32_t>(h->m, 1), n1 = std::max<int32_t>(h->n, 1);
    hipError_t e = hipMalloc(&bt->row_inst, sizeof(int32_t) * m1);
    if (e == hipSuccess) e = hipMalloc(&bt->col_inst, sizeof(int32_t) * n1);
    if (e == hipSuccess) e = hipMalloc(&bt->seg, sizeof(int32_t) * seg.size());
    if (e == hipSuccess) e = hipMalloc(&bt->tau, sizeof(double) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->sigma, sizeof(double) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->active, sizeof(int32_t) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->since, sizeof(int32_t) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->flags, sizeof(int32_t) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->use_avg, sizeof(int32_t) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->fac, sizeof(double) * count);
    if (e == hipSuccess) e = hipMalloc(&bt->terms, sizeof(double) * 11 * (m1 + n1));
    if (e == hipSuccess) e = hipMalloc(&bt->red, sizeof(double) * 22 * count);
    if (e == hipSuccess) e = hipHostMalloc(&bt->h_red, sizeof(double) * 22 * count, hipHostMallocDefault);
    std::vector<int32_t> table;
    const int32_t slots = batch_reduce_chunks(seg, count, table);
    bt->n_chunks = static_cast<int32_t>(table.size() / kReduceRec);
    if (e == hipSuccess) e = hipMalloc(&bt->chunks, sizeof(int32_t) * table.size());
    if (e == hipSuccess) e = hipMalloc(&bt->chunk_part, sizeof(double) * 11 * std::max<int32_t>(slots, 1));
    if (e == hipSuccess) e = hipMalloc(&bt->chunk_cnt, sizeof(int32_t) * 2 * count);
    if (e == hipSuccess)
        e = hipMemcpy(bt->chunks, table.data(), sizeof(int32_t) * table.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(bt->chunk_cnt, 0, sizeof(int32_t) * 2 * count, ctx->stream);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&bt->side, hipStreamNonBlocking);   // default priority
    if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->ev_join, hipEventDisableTiming);
    if (e == hipSuccess && h->m > 0)
        e = hipMemcpy(bt->row_inst, rinst.data(), sizeof(int32_t) * h->m, hipMemcpyHostToDevice);
    if (e == hipSuccess && h->n > 0)
        e = hipMemcpy(bt->col_inst, cinst.data(), sizeof(int32_t) * h->n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->seg, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice);
    std::vector<int32_t> zero(static_cast<size_t>(count), 0), one(static_cast<size_t>(count), 2);
    std::vector<double> unit(static_cast<size_t>(count), 1.0);
    if (e == hipSuccess) e = hipMemcpy(bt->tau, unit.data(), sizeof(double) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->sigma, unit.data(), sizeof(double) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->active, one.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->since, zero.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->flags, zero.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_batch(bt);
        return hip_fail(e, "batch_create");
    }
    rc = batch_block_flags(ctx, bt);
    if (rc == SOV_OK) {
        const hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e2 != hipSuccess) rc = hip_fail(e2, "batch_create flags");
    }
    if (rc != SOV_OK) {
        free_batch(bt);
        return rc;
    }
    *out = bt;
    return SOV_OK;
}

void sov_hip_batch_destroy(sov_hip_batch* bt) {
    if (bt == nullptr) return;
    DeviceGuard3 g(bt->device);
    // Only the creating context's stream can still be using these buffers.
    if (bt->stream != nullptr) (void)hipStreamSynchronize(bt->stream);
    free_batch(bt);
}

int sov_hip_batch_set_stage_bytes(sov_hip_batch* bt, int64_t bytes) {
    if (bt == nullptr || bytes < 0) return fail(SOV_INVALID_ARGUMENT, "batch_set_stage_bytes args");
    bt->stage_bytes = std::min<int64_t>(bytes, bt->lds_max);
    bt->res.built_from[0] = bt->res.built_from[1] = -1;   // rebuilt at the next batch_set(active)
    return SOV_OK;
}

int sov_hip_batch_set_split(sov_hip_batch* bt, int on) {
    if (bt == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch_set_split args");
    if (bt->forked) return fail(SOV_INVALID_ARGUMENT, "batch_set_split between a fork and its join");
    bt->split = on != 0;
    return SOV_OK;
}

int sov_hip_batch_set_tile_share(sov_hip_batch* bt, double share) {
    if (bt == nullptr || share != share) return fail(SOV_INVALID_ARGUMENT, "batch_set_tile_share args");
    bt->tile_frac = share;
    return SOV_OK;
}

int sov_hip_batch_set(sov_hip_ctx* ctx, sov_hip_batch* bt, const double* tau, const double* sigma,
                      const int32_t* active, const int32_t* since) {
    if (ctx == nullptr || bt == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch_set args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "batch_set during capture");
    DeviceGuard3 g(ctx->device);
    const size_t c = static_cast<size_t>(bt->count);
    if (tau) SOV_HIP_TRY(hipMemcpyAsync(bt->tau, tau, sizeof(double) * c, hipMemcpyHostToDevice, ctx->stream));
    if (sigma) SOV_HIP_TRY(hipMemcpyAsync(bt->sigma, sigma, sizeof(double) * c, hipMemcpyHostToDevice, ctx->stream));
    if (active) {
        SOV_HIP_TRY(hipMemcpyAsync(bt->active, active, sizeof(int32_t) * c, hipMemcpyHostToDevice, ctx->stream));
        bt->h_active.assign(active, active + c);
        const int rc = batch_block_flags(ctx, bt);
        if (rc != SOV_OK) return rc;
    }
    if (since) SOV_HIP_TRY(hipMemcpyAsync(bt->since, since, sizeof(int32_t) * c, hipMemcpyHostToDevice, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));   // the host arrays may be reused at once
    return SOV_OK;
}

int sov_hip_batch_primal_step(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt,
                              const sov_hip_pdhg_state* st, int k) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch primal_step args");
    int rc = state_ok(h, st);
    if (rc == SOV_OK) rc = batch_ok(h, bt, st);
    if (rc != SOV_OK) return rc;
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? batch_primal_impl<double>(ctx, h, bt, st, k)
                                    : batch_primal_impl<float>(ctx, h, bt, st, k);
}

int sov_hip_batch_dual_step(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt,
                            const sov_hip_pdhg_state* st, int k) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch dual_step args");
    int rc = state_ok(h, st);
    if (rc == SOV_OK) rc = batch_ok(h, bt, st);
    if (rc != SOV_OK) return rc;
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? batch_dual_impl<double>(ctx, h, bt, st, k)
                                    : batch_dual_impl<float>(ctx, h, bt, st, k);
}

int sov_hip_batch_resident_join(sov_hip_ctx* ctx, const sov_hip_batch* bt) {
    if (ctx == nullptr || bt == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch resident_join args");
    DeviceGuard3 g(ctx->device);
    return batch_join(ctx, bt);
}

int sov_hip_batch_resident_steps(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt,
                                 const sov_hip_pdhg_state* st, int count) {
    if (ctx == nullptr || h == nullptr || count < 0) return fail(SOV_INVALID_ARGUMENT, "batch resident_steps args");
    int rc = state_ok(h, st);
    if (rc == SOV_OK) rc = batch_ok(h, bt, st);
    if (rc != SOV_OK) return rc;
    DeviceGuard3 g(ctx->device);
    return st->prec == SOV_PREC_F64 ? batch_resident_impl<double>(ctx, h, bt, st, count)
                                    : batch_resident_impl<float>(ctx, h, bt, st, count);
}

int sov_hip_batch_check(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_batch* bt,
                        const sov_hip_check_args* a, const double* obj_const, sov_hip_check_out* out) {
    if (ctx == nullptr || h == nullptr || a == nullptr || out == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "batch check args");
    int rc = state_ok(h, a->st);
    if (rc == SOV_OK) rc = batch_ok(h, bt, a->st);
    if (rc != SOV_OK) return rc;
    if ((h->m > 0 && (!a->row_scale || !a->orig_row_lo || !a->orig_row_hi)) ||
        (h->n > 0 && (!a->col_scale || !a->orig_col_lo || !a->orig_col_hi || !a->orig_c)))
        return fail(SOV_INVALID_ARGUMENT, "batch check: null original-space array");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "batch check during capture");
    DeviceGuard3 g(ctx->device);
    rc = batch_join(ctx, bt);   // the iterates may still be in flight on the batch's stream
    if (rc != SOV_OK) return rc;
    return a->st->prec == SOV_PREC_F64 ? batch_check_impl<double>(ctx, h, bt, a, obj_const, out)
                                       : batch_check_impl<float>(ctx, h, bt, a, obj_const, out);
}

int sov_hip_batch_restart(sov_hip_ctx* ctx, const sov_hip_batch* bt, const sov_hip_pdhg_state* st,
                          const int32_t* flags, const double* act) {
    if (ctx == nullptr || flags == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch restart args");
    int rc = state_ok(nullptr, st);
    if (rc == SOV_OK) rc = batch_ok(nullptr, bt, st);
    if (rc != SOV_OK) return rc;
    if (st->m > 0 && act == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch restart: act is null");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "batch restart during capture");
    DeviceGuard3 g(ctx->device);
    rc = batch_join(ctx, bt);
    if (rc != SOV_OK) return rc;
    SOV_HIP_TRY(hipMemcpyAsync(bt->flags, flags, sizeof(int32_t) * bt->count, hipMemcpyHostToDevice, ctx->stream));
    if (st->prec == SOV_PREC_F64) {
        if (st->n > 0)
            k_batch_restart_x<double><<<launch_blocks(st->n), kBlock, 0, ctx->stream>>>(
                st->n, bt->col_inst, bt->flags, bt->since, static_cast<const double*>(st->x_sum),
                static_cast<const double*>(st->col_lo), static_cast<const double*>(st->col_hi),
                static_cast<double*>(st->x));
        if (st->m > 0)
            k_batch_restart_y<double><<<launch_blocks(st->m), kBlock, 0, ctx->stream>>>(
                st->m, bt->row_inst, bt->flags, bt->since, static_cast<const double*>(st->y_sum), act,
                static_cast<const double*>(st->row_lo), static_cast<const double*>(st->row_hi),
                static_cast<double*>(st->y), static_cast<double*>(st->s));
    } else {
        if (st->n > 0)
            k_batch_restart_x<float><<<launch_blocks(st->n), kBlock, 0, ctx->stream>>>(
                st->n, bt->col_inst, bt->flags, bt->since, static_cast<const float*>(st->x_sum),
                static_cast<const float*>(st->col_lo), static_cast<const float*>(st->col_hi),
                static_cast<float*>(st->x));
        if (st->m > 0)
            k_batch_restart_y<float><<<launch_blocks(st->m), kBlock, 0, ctx->stream>>>(
                st->m, bt->row_inst, bt->flags, bt->since, static_cast<const float*>(st->y_sum), act,
                static_cast<const float*>(st->row_lo), static_cast<const float*>(st->row_hi),
                static_cast<float*>(st->y), static_cast<float*>(st->s));
    }
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));   // flags' host array may be reused at once
    return SOV_OK;
}

int sov_hip_batch_sumsq(sov_hip_ctx* ctx, const sov_hip_batch* bt, int which, const double* v, double* out) {
    if (ctx == nullptr || bt == nullptr || out == nullptr || (which != 0 && which != 1) ||
        ((which == 0 ? bt->m : bt->n) > 0 && v == nullptr))
        return fail(SOV_INVALID_ARGUMENT, "batch sumsq args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "batch sumsq during capture");
    DeviceGuard3 g(ctx->device);
    k_batch_sumsq<<<bt->count, kBlock, 0, ctx->stream>>>(bt->seg, which, v, bt->red);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(bt->h_red, bt->red, sizeof(double) * bt->count, hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::copy(bt->h_red, bt->h_red + bt->count, out);
    return SOV_OK;
}

int sov_hip_batch_promote(sov_hip_ctx* ctx, const sov_hip_batch* bt, const sov_hip_pdhg_state* st32,
                          const sov_hip_pdhg_state* st64, const int32_t* flags, const int32_t* use_avg) {
    if (ctx == nullptr || flags == nullptr || use_avg == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch promote args");
    int rc = state_ok(nullptr, st32);
    if (rc == SOV_OK) rc = state_ok(nullptr, st64);
    if (rc == SOV_OK) rc = batch_ok(nullptr, bt, st32);
    if (rc == SOV_OK) rc = batch_ok(nullptr, bt, st64);
    if (rc != SOV_OK) return rc;
    if (st32->prec != SOV_PREC_F32 || st64->prec != SOV_PREC_F64)
        return fail(SOV_INVALID_ARGUMENT, "batch promote: needs an FP32 and an FP64 state");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "batch promote during capture");
    DeviceGuard3 g(ctx->device);
    rc = batch_join(ctx, bt);
    if (rc != SOV_OK) return rc;
    const size_t c = static_cast<size_t>(bt->count);
    SOV_HIP_TRY(hipMemcpyAsync(bt->flags, flags, sizeof(int32_t) * c, hipMemcpyHostToDevice, ctx->stream));
    SOV_HIP_TRY(hipMemcpyAsync(bt->use_avg, use_avg, sizeof(int32_t) * c, hipMemcpyHostToDevice, ctx->stream));
    if (st32->n > 0)
        k_batch_promote<false><<<launch_blocks(st32->n), kBlock, 0, ctx->stream>>>(
            st32->n, bt->col_inst, bt->flags, bt->use_avg, bt->since, static_cast<const float*>(st32->x),
            static_cast<const float*>(st32->x_sum), static_cast<double*>(st64->x), static_cast<double*>(st64->x_sum));
    if (st32->m > 0)
        k_batch_promote<true><<<launch_blocks(st32->m), kBlock, 0, ctx->stream>>>(
            st32->m, bt->row_inst, bt->flags, bt->use_avg, bt->since, static_cast<const float*>(st32->y),
            static_cast<const float*>(st32->y_sum), static_cast<double*>(st64->y), static_cast<double*>(st64->y_sum));
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_batch_scale(sov_hip_ctx* ctx, const sov_hip_batch* bt, int which, const double* factors, double* v) {
    if (ctx == nullptr || bt == nullptr || factors == nullptr || (which != 0 && which != 1))
        return fail(SOV_INVALID_ARGUMENT, "batch scale args");
    const int32_t len = which == 0 ? bt->m : bt->n;
    if (len > 0 && v == nullptr) return fail(SOV_INVALID_ARGUMENT, "batch scale: v is null");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "batch scale during capture");
    DeviceGuard3 g(ctx->device);
    SOV_HIP_TRY(hipMemcpyAsync(bt->fac, factors, sizeof(double) * bt->count, hipMemcpyHostToDevice, ctx->stream));
    if (len > 0)
        k_batch_scale<<<launch_blocks(len), kBlock, 0, ctx->stream>>>(len, which == 0 ? bt->row_inst : bt->col_inst,
                                                                      bt->fac, v);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}


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

// The refinement fields of a check: all set (with an FP64 base of the
// state's dimensions) or base null.
static int check_base_ok(const sov_hip_csr* h, const sov_hip_check_args* a) {
    if (a->base == nullptr) return SOV_OK;
    const int rc = state_ok(h, a->base);
    if (rc != SOV_OK) return rc;
    if (a->base->prec != SOV_PREC_F64) return fail(SOV_INVALID_ARGUMENT, "check: the base state must be FP64");
    if ((h->n > 0 && (!a->base_x || !a->base_aty)) || (h->m > 0 && (!a->base_y || !a->base_act)))
        return fail(SOV_INVALID_ARGUMENT, "check: null anchor array");
    if (!(a->inv_delta > 0.0) || !(a->inv_delta < INFINITY))
        return fail(SOV_INVALID_ARGUMENT, "check: inv_delta must be positive");
    return SOV_OK;
}

int sov_hip_pdhg_check(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_check_args* a, sov_hip_check_out* out) {
    if (ctx == nullptr || h == nullptr || a == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "check args");
    int rc = state_ok(h, a->st);
    if (rc != SOV_OK) return rc;
    if ((h->m > 0 && (!a->row_scale || !a->orig_row_lo || !a->orig_row_hi)) ||
        (h->n > 0 && (!a->col_scale || !a->orig_col_lo || !a->orig_col_hi || !a->orig_c)))
        return fail(SOV_INVALID_ARGUMENT, "check: null original-space array");
    rc = check_base_ok(h, a);
    if (rc != SOV_OK) return rc;
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "check during capture");
    DeviceGuard3 g(ctx->device);
    if (check_blocks(h) > h->partial_blocks) return fail(SOV_INVALID_ARGUMENT, "check: partials too small");
    rc = ctx->ensure_red(64);
    if (rc != SOV_OK) return rc;
    return a->st->prec == SOV_PREC_F64 ? check_impl<double>(ctx, h, a, out, h->partials, ctx->d_red, h->pairs)
                                       : check_impl<float>(ctx, h, a, out, h->partials, ctx->d_red, h->pairs);
}

int sov_hip_pdhg_check_static(sov_hip_ctx* ctx, const sov_hip_csr* h, const sov_hip_check_args* a,
                              sov_hip_check_out* out) {
    if (ctx == nullptr || h == nullptr || a == nullptr || out == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "check_static args");
    int rc = state_ok(h, a->st);
    if (rc != SOV_OK) return rc;
    if ((h->m > 0 && (!a->row_scale || !a->orig_row_lo || !a->orig_row_hi)) ||
        (h->n > 0 && (!a->col_scale || !a->orig_col_lo || !a->orig_col_hi || !a->orig_c)))
        return fail(SOV_INVALID_ARGUMENT, "check_static: null original-space array");
    rc = check_base_ok(h, a);
    if (rc != SOV_OK) return rc;
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "check_static during capture");
    if (ctx->check_scratch == nullptr)
        return fail(SOV_BACKEND_ERROR, "check_static: the context has no static scratch address");
    const int64_t need = check_blocks(h);
    if (need > kStaticCheckBlocks)
        return fail(SOV_INVALID_ARGUMENT, "check_static: the check needs " + std::to_string(need) +
                                              " blocks of partials; the static scratch holds " +
                                              std::to_string(kStaticCheckBlocks) + " (use sov_hip_pdhg_check)");
    DeviceGuard3 g(ctx->device);
    double* partials = ctx->check_scratch;
    double* red = ctx->check_scratch + kStaticCheckBlocks * kCheckSlots;
    return a->st->prec == SOV_PREC_F64 ? check_impl<double>(ctx, h, a, out, partials, red, nullptr)
                                       : check_impl<float>(ctx, h, a, out, partials, red, nullptr);
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

int sov_hip_pdhg_combine(sov_hip_ctx* ctx, const sov_hip_pdhg_state* corr, const sov_hip_pdhg_state* dst,
                         const double* x0, const double* y0, int64_t since, double inv_delta) {
    if (ctx == nullptr || x0 == nullptr || y0 == nullptr) return fail(SOV_INVALID_ARGUMENT, "combine args");
    int rc = state_ok(nullptr, corr);
    if (rc == SOV_OK) rc = state_ok(nullptr, dst);
    if (rc != SOV_OK) return rc;
    if (dst->prec != SOV_PREC_F64) return fail(SOV_INVALID_ARGUMENT, "combine: the target state must be FP64");
    if (corr->m != dst->m || corr->n != dst->n) return fail(SOV_DIMENSION_MISMATCH, "combine: state dims");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "combine during capture");
    DeviceGuard3 g(ctx->device);
    const double sn = static_cast<double>(since);
    auto launch = [&](auto tag) {
        using T = decltype(tag);
        if (corr->n > 0)
            k_combine<T><<<launch_blocks(corr->n), kBlock, 0, ctx->stream>>>(
                corr->n, x0, static_cast<const T*>(corr->x), static_cast<const T*>(corr->x_sum), sn, inv_delta,
                static_cast<double*>(dst->x), static_cast<double*>(dst->x_sum));
        if (corr->m > 0)
            k_combine<T><<<launch_blocks(corr->m), kBlock, 0, ctx->stream>>>(
                corr->m, y0, static_cast<const T*>(corr->y), static_cast<const T*>(corr->y_sum), sn, inv_delta,
                static_cast<double*>(dst->y), static_cast<double*>(dst->y_sum));
    };
    if (corr->prec == SOV_PREC_F64) launch(double{});
    else launch(float{});
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_pdhg_anchor(sov_hip_ctx* ctx, const sov_hip_pdhg_state* base, const sov_hip_pdhg_state* corr,
                        const double* x0, const double* y0, const double* s0, const double* aty0, double delta) {
    if (ctx == nullptr || x0 == nullptr || y0 == nullptr || s0 == nullptr || aty0 == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "anchor args");
    int rc = state_ok(nullptr, base);
    if (rc == SOV_OK) rc = state_ok(nullptr, corr);
    if (rc != SOV_OK) return rc;
    if (base->prec != SOV_PREC_F64) return fail(SOV_INVALID_ARGUMENT, "anchor: the base state must be FP64");
    if (corr->m != base->m || corr->n != base->n) return fail(SOV_DIMENSION_MISMATCH, "anchor: state dims");
    if (corr->m > 0 && corr->s_grad == nullptr) return fail(SOV_INVALID_ARGUMENT, "anchor: s_grad is not set");
    if (!(delta > 0.0) || !(delta < INFINITY)) return fail(SOV_INVALID_ARGUMENT, "anchor: delta must be positive");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "anchor during capture");
    DeviceGuard3 g(ctx->device);
    auto launch = [&](auto tag) {
        using T = decltype(tag);
        if (corr->n > 0)
            k_anchor<T><<<launch_blocks(corr->n), kBlock, 0, ctx->stream>>>(
                corr->n, static_cast<const double*>(base->col_lo), static_cast<const double*>(base->col_hi), x0,
                static_cast<const double*>(base->c), aty0, delta, writable<T>(corr->col_lo),
                writable<T>(corr->col_hi), static_cast<T*>(corr->x), writable<T>(corr->c), nullptr);
        if (corr->m > 0)
            k_anchor<T><<<launch_blocks(corr->m), kBlock, 0, ctx->stream>>>(
                corr->m, static_cast<const double*>(base->row_lo), static_cast<const double*>(base->row_hi), s0,
                y0, nullptr, delta, writable<T>(corr->row_lo), writable<T>(corr->row_hi),
                static_cast<T*>(corr->s), writable<T>(corr->s_grad), static_cast<T*>(corr->y));
    };
    if (corr->prec == SOV_PREC_F64) launch(double{});
    else launch(float{});
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
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
