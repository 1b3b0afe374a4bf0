// CSR handle, device-built transpose, row-length schedules, SpMV, the lane
// autotuner and the Ruiz / Pock-Chambolle scaling kernels.
//
// Reference counterparts: cuSPARSE SpMV on one CSR in both orientations
// (src/l0/backend_cuda.cu:220-284), the host scatter SpMV^T
// (src/l0/backend_host.cpp:151-172), and host scaling (src/l1/scaling.cpp:19-104).
// Here A^T is materialised once, on the device, as its own CSR, so both
// products are gathers with one writer per output element.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sov_hip_internal.hpp"

namespace sovhip {

struct DeviceGuard2 {
    int prev = -1;
    explicit DeviceGuard2(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard2() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

Sched Orient::sched() const {
    Sched s{};
    int b = 0;
    for (int c = 0; c < kClasses; ++c) {
        const int G = lanes[c] > 0 ? lanes[c] : kDefaultLanes[c];
        s.G[c] = G;
        s.count[c] = class_count[c];
        s.offset[c] = class_offset[c];
        s.block_begin[c] = b;
        const int rpb = kBlock / G;
        b += (class_count[c] + rpb - 1) / rpb;
    }
    s.block_begin[kClasses] = b;
    s.list = list;
    return s;
}

int Orient::blocks() const { return sched().block_begin[kClasses]; }

// ---------------------------------------------------------------------------
// transpose construction
// ---------------------------------------------------------------------------

__global__ void k_count_cols(int32_t nnz, const int32_t* __restrict__ ci, int32_t* __restrict__ cnt1) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt1[ci[k] + 1], 1);
}

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

// Inclusive scan of each 2048-element tile in place; tile totals to sums[].
__global__ void k_scan_tiles(int64_t len, int32_t* __restrict__ a, int32_t* __restrict__ sums) {
    __shared__ int32_t lds[kBlock];
    const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * (int64_t)kScanItems;
    int32_t v[kScanItems];
    int32_t run = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t p = base + i;
        run += (p < len) ? a[p] : 0;
        v[i] = run;
    }
    lds[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int32_t t = (threadIdx.x >= (unsigned)off) ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    const int32_t excl = lds[threadIdx.x] - run;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t p = base + i;
        if (p < len) a[p] = v[i] + excl;
    }
    if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = lds[kBlock - 1];
}

// Exclusive scan of the tile totals (one thread: a few thousand entries).
__global__ void k_scan_sums(int64_t count, int32_t* __restrict__ sums) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t run = 0;
    for (int64_t i = 0; i < count; ++i) {
        const int32_t t = sums[i];
        sums[i] = run;
        run += t;
    }
}

__global__ void k_scan_add(int64_t len, int32_t* __restrict__ a, const int32_t* __restrict__ sums) {
    const int64_t tile = blockIdx.x;
    if (tile == 0) return;
    const int32_t add = sums[tile];
    for (int i = threadIdx.x; i < kScanTile; i += kBlock) {
        const int64_t p = tile * kScanTile + i;
        if (p < len) a[p] += add;
    }
}

static int inclusive_scan(sov_hip_ctx* ctx, int32_t* a, int64_t len) {
    const int64_t tiles = (len + kScanTile - 1) / kScanTile;
    int32_t* sums = nullptr;
    SOV_HIP_TRY(hipMalloc(&sums, sizeof(int32_t) * (tiles + 1)));
    k_scan_tiles<<<(int)tiles, kBlock, 0, ctx->stream>>>(len, a, sums);
    k_scan_sums<<<1, 64, 0, ctx->stream>>>(tiles, sums);
    k_scan_add<<<(int)tiles, kBlock, 0, ctx->stream>>>(len, a, sums);
    hipError_t e = check_launch();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(sums);
    if (e != hipSuccess) return hip_fail(e, "inclusive_scan");
    return SOV_OK;
}

// Largest r with ptr[r] <= k (the row holding nonzero k), ptr of rows+1.
__device__ __forceinline__ int32_t find_row(const int32_t* __restrict__ ptr, int32_t rows, int32_t k) {
    int32_t lo = 0, hi = rows;   // invariant: ptr[lo] <= k < ptr[hi]
    while (hi - lo > 1) {
        const int32_t mid = lo + (hi - lo) / 2;
        if (ptr[mid] <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_fill_transpose(int32_t m, int32_t nnz, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci, int32_t* __restrict__ cursor,
                                 int32_t* __restrict__ trow, int32_t* __restrict__ tsrc) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = find_row(rp, m, (int32_t)k);
        const int32_t pos = atomicAdd(&cursor[ci[k]], 1);
        trow[pos] = r;
        tsrc[pos] = (int32_t)k;
    }
}

// Deterministic placement: the entries of one A^T row have distinct A-row
// indices, so each entry's rank among them is its final slot.
__global__ void k_rank_transpose(int32_t n, int32_t nnz, const int32_t* __restrict__ tptr,
                                 const int32_t* __restrict__ trow_u, const int32_t* __restrict__ tsrc_u,
                                 int32_t* __restrict__ tidx, int32_t* __restrict__ tsrc) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = find_row(tptr, n, (int32_t)p);
        const int32_t b = tptr[j], e = tptr[j + 1];
        const int32_t me = trow_u[p];
        int32_t rank = 0;
        for (int32_t q = b; q < e; ++q) rank += (trow_u[q] < me) ? 1 : 0;
        tidx[b + rank] = me;
        tsrc[b + rank] = tsrc_u[p];
    }
}

__global__ void k_gather_vals(int32_t nnz, const int32_t* __restrict__ src, const double* __restrict__ v,
                              double* __restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * blockDim.x)
        out[p] = v[src[p]];
}

__global__ void k_copy_i32(int64_t n, const int32_t* __restrict__ s, int32_t* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__global__ void k_zero_i32(int64_t n, int32_t* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = 0;
}

__global__ void k_copy_f64(int64_t n, const double* __restrict__ s, double* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ s, float* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}

// Host: group rows by length class (stable), upload the list.
int build_schedule(sov_hip_ctx* ctx, Orient& o) {
    std::vector<int32_t> ptr(static_cast<size_t>(o.rows) + 1);
    SOV_HIP_TRY(hipMemcpyAsync(ptr.data(), o.ptr, sizeof(int32_t) * ptr.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> cls(static_cast<size_t>(o.rows));
    int32_t cnt[kClasses] = {};
    o.max_len = 0;
    for (int32_t r = 0; r < o.rows; ++r) {
        const int32_t len = ptr[r + 1] - ptr[r];
        o.max_len = std::max<int64_t>(o.max_len, len);
        int c = 0;
        while (c < kClasses - 1 && len > kClassMax[c]) ++c;
        cls[r] = c;
        ++cnt[c];
    }
    int32_t off = 0;
    for (int c = 0; c < kClasses; ++c) {
        o.class_count[c] = cnt[c];
        o.class_offset[c] = off;
        off += cnt[c];
        if (o.lanes[c] == 0) o.lanes[c] = kDefaultLanes[c];
    }
    std::vector<int32_t> list(static_cast<size_t>(o.rows));
    int32_t cur[kClasses];
    for (int c = 0; c < kClasses; ++c) cur[c] = o.class_offset[c];
    for (int32_t r = 0; r < o.rows; ++r) list[cur[cls[r]]++] = r;
    if (o.rows > 0) {
        SOV_HIP_TRY(hipMalloc(&o.list, sizeof(int32_t) * o.rows));
        SOV_HIP_TRY(hipMemcpyAsync(o.list, list.data(), sizeof(int32_t) * o.rows, hipMemcpyHostToDevice,
                                   ctx->stream));
        SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return SOV_OK;
}

int ensure_partials(sov_hip_csr* csr) {
    const int64_t need = (int64_t)csr->a.blocks() + (int64_t)csr->t.blocks() + 64;
    if (need <= csr->partial_blocks) return SOV_OK;
    if (csr->partials) (void)hipFree(csr->partials);
    csr->partials = nullptr;
    csr->partial_blocks = 0;
    SOV_HIP_TRY(hipMalloc(&csr->partials, sizeof(double) * kCheckSlots * need));
    csr->partial_blocks = need;
    return SOV_OK;
}

// ---------------------------------------------------------------------------
// SpMV over a schedule
// ---------------------------------------------------------------------------

template <typename V, typename T>
struct SpmvOp {
    const int32_t* ptr;
    const int32_t* idx;
    const V* val;
    const T* x;
    T* y;
    T alpha, beta;
    Sched s;
    int cls;
    template <int G>
    __device__ void run() {
        constexpr int RPB = kBlock / G;
        const int local = (blockIdx.x - s.block_begin[cls]) * RPB + threadIdx.x / G;
        const int lane = threadIdx.x % G;
        const int row = (local < s.count[cls]) ? s.list[s.offset[cls] + local] : -1;
        const T sum = row_dot<G, T>(ptr, idx, val, x, row, lane);
        if (row >= 0 && lane == 0) y[row] = (beta == T(0)) ? alpha * sum : alpha * sum + beta * y[row];
    }
};

template <typename V, typename T>
__global__ void __launch_bounds__(kBlock) k_spmv(SpmvOp<V, T> op) {
    op.cls = block_class(op.s);
    dispatch_G(op.s.G[op.cls], op);
}

template <typename V, typename T>
static int launch_spmv(sov_hip_ctx* ctx, const Orient& o, const V* val, double alpha, const void* x,
                       double beta, void* y, const Sched& s) {
    const int blocks = s.block_begin[kClasses];
    if (blocks == 0) return SOV_OK;
    SpmvOp<V, T> op{o.ptr, o.idx, val, static_cast<const T*>(x), static_cast<T*>(y), T(alpha), T(beta), s, 0};
    k_spmv<V, T><<<blocks, kBlock, 0, ctx->stream>>>(op);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

static int spmv_dispatch(sov_hip_ctx* ctx, const Orient& o, int prec, double alpha, const void* x,
                         double beta, void* y, const Sched& s) {
    if (prec == SOV_PREC_F64) return launch_spmv<double, double>(ctx, o, o.val64, alpha, x, beta, y, s);
    if (o.val32 == nullptr) return fail(SOV_INVALID_ARGUMENT, "FP32 SpMV needs sov_hip_csr_make_f32 first");
    return launch_spmv<float, float>(ctx, o, o.val32, alpha, x, beta, y, s);
}

// ---------------------------------------------------------------------------
// scaling kernels (thread per row, serial over the row in index order)
// ---------------------------------------------------------------------------

// Ruiz: out[r] = max_k (|v_k| * own[r]) * other[idx_k]  (same product order as
// the reference, scaling.cpp:48-50; for the A^T pass own = col factor, and
// the product is (|v| * r_i) * c_j, so the operands are swapped accordingly)
__global__ void k_ruiz_rowmax(int32_t rows, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                              const double* __restrict__ val, const double* __restrict__ rs,
                              const double* __restrict__ cs, double* __restrict__ out) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        double mx = 0.0;
        const double rr = rs[r];
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const double sc = fabs(val[k]) * rr * cs[idx[k]];
            mx = (mx < sc) ? sc : mx;
        }
        out[r] = mx;
    }
}

__global__ void k_ruiz_colmax(int32_t cols, const int32_t* __restrict__ tptr, const int32_t* __restrict__ tidx,
                              const double* __restrict__ tval, const double* __restrict__ rs,
                              const double* __restrict__ cs, double* __restrict__ out) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < cols;
         j += (int64_t)gridDim.x * blockDim.x) {
        double mx = 0.0;
        const double cc = cs[j];
        for (int32_t k = tptr[j]; k < tptr[j + 1]; ++k) {
            const double sc = fabs(tval[k]) * rs[tidx[k]] * cc;
            mx = (mx < sc) ? sc : mx;
        }
        out[j] = mx;
    }
}

// Factor update of one Ruiz sweep; per-block max |norm - 1| to partial.
__global__ void k_ruiz_update(int32_t m, int32_t n, const double* __restrict__ rmax,
                              const double* __restrict__ cmax, double* __restrict__ rs,
                              double* __restrict__ cs, double* __restrict__ partial) {
    double worst = 0.0;
    const int64_t total = (int64_t)m + n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < m) {
            const double v = rmax[i];
            if (v > 0.0) {
                rs[i] /= sqrt(v);
                const double d = fabs(v - 1.0);
                worst = (worst < d) ? d : worst;
            }
        } else {
            const double v = cmax[i - m];
            if (v > 0.0) {
                cs[i - m] /= sqrt(v);
                const double d = fabs(v - 1.0);
                worst = (worst < d) ? d : worst;
            }
        }
    }
    double s0[1] = {0.0}, mx[1] = {worst};
    block_reduce_write<0, 1>(s0, mx, partial + blockIdx.x);
}

__global__ void k_max_partials(int64_t count, const double* __restrict__ partial, double* __restrict__ out) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += blockDim.x) acc = nmax(acc, partial[i]);
    double s0[1] = {0.0}, mx[1] = {acc};
    block_reduce_write<0, 1>(s0, mx, out);
}

// Pock-Chambolle sums: m^e over positive m = (|v| r) c, serial in index order.
__global__ void k_pc_rowsum(int32_t rows, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                            const double* __restrict__ val, const double* __restrict__ rs,
                            const double* __restrict__ cs, double expo, double* __restrict__ out) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        const double rr = rs[r];
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const double mg = fabs(val[k]) * rr * cs[idx[k]];
            if (mg <= 0.0) continue;
            acc += (expo == 1.0) ? mg : pow(mg, expo);
        }
        out[r] = acc;
    }
}

__global__ void k_pc_colsum(int32_t cols, const int32_t* __restrict__ tptr, const int32_t* __restrict__ tidx,
                            const double* __restrict__ tval, const double* __restrict__ rs,
                            const double* __restrict__ cs, double expo, double* __restrict__ out) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < cols;
         j += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        const double cc = cs[j];
        for (int32_t k = tptr[j]; k < tptr[j + 1]; ++k) {
            const double mg = fabs(tval[k]) * rs[tidx[k]] * cc;
            if (mg <= 0.0) continue;
            acc += (expo == 1.0) ? mg : pow(mg, expo);
        }
        out[j] = acc;
    }
}

__global__ void k_pc_update(int32_t m, int32_t n, const double* __restrict__ rsum,
                            const double* __restrict__ csum, double* __restrict__ rs, double* __restrict__ cs) {
    const int64_t total = (int64_t)m + n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < m) {
            if (rsum[i] > 0.0) rs[i] /= sqrt(rsum[i]);
        } else {
            if (csum[i - m] > 0.0) cs[i - m] /= sqrt(csum[i - m]);
        }
    }
}

// A <- diag(r) A diag(c) on the owned A values; magnitudes <= 1e-14 -> 0.
__global__ void k_scale_vals(int32_t m, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                             double* __restrict__ v, const double* __restrict__ rs,
                             const double* __restrict__ cs) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        const double rr = rs[r];
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
            const double s = v[k] * rr * cs[ci[k]];
            v[k] = (fabs(s) <= 1e-14) ? 0.0 : s;
        }
    }
}

}  // namespace sovhip

using namespace sovhip;

extern "C" {

int sov_hip_csr_create(sov_hip_ctx* ctx, int32_t m, int32_t n, int32_t nnz, const int32_t* d_rp,
                       const int32_t* d_ci, const double* d_val, int flags, sov_hip_csr** out) {
    if (ctx == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "csr_create args");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0) return fail(SOV_INVALID_ARGUMENT, "csr_create: negative size");
    if (d_rp == nullptr || (nnz > 0 && (d_ci == nullptr || d_val == nullptr)))
        return fail(SOV_INVALID_ARGUMENT, "csr_create: null array");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "csr_create during capture");
    DeviceGuard2 g(ctx->device);
    auto* h = new (std::nothrow) sov_hip_csr();
    if (h == nullptr) return fail(SOV_OUT_OF_MEMORY, "csr");
    h->m = m;
    h->n = n;
    h->nnz = nnz;
    h->device = ctx->device;
    auto cleanup = [&](int rc) {
        sov_hip_csr_destroy(h);
        return rc;
    };
    const int64_t nz = std::max<int32_t>(nnz, 1);
    // A
    h->a.rows = m;
    h->a.cols = n;
    h->a.ptr = d_rp;
    h->a.idx = d_ci;
    if (flags & SOV_CSR_OWN_VALUES) {
        hipError_t e = hipMalloc(&h->a.val64, sizeof(double) * nz);
        if (e != hipSuccess) return cleanup(hip_fail(e, "csr values"));
        h->a.owns_val64 = true;
        if (nnz > 0) k_copy_f64<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, d_val, h->a.val64);
    } else {
        h->a.val64 = const_cast<double*>(d_val);
    }
    // A^T structure
    int32_t *cursor = nullptr, *trow_u = nullptr, *tsrc_u = nullptr;
    hipError_t e = hipMalloc(&h->t_ptr, sizeof(int32_t) * ((int64_t)n + 1));
    if (e == hipSuccess) e = hipMalloc(&h->t_idx, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&h->t_src, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&h->t.val64, sizeof(double) * nz);
    if (e == hipSuccess) e = hipMalloc(&cursor, sizeof(int32_t) * ((int64_t)n + 1));
    if (e == hipSuccess) e = hipMalloc(&trow_u, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&tsrc_u, sizeof(int32_t) * nz);
    if (e != hipSuccess) {
        (void)hipFree(cursor); (void)hipFree(trow_u); (void)hipFree(tsrc_u);
        return cleanup(hip_fail(e, "csr transpose buffers"));
    }
    h->t.owns_val64 = true;
    k_zero_i32<<<launch_blocks((int64_t)n + 1), kBlock, 0, ctx->stream>>>((int64_t)n + 1, h->t_ptr);
    if (nnz > 0) k_count_cols<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, d_ci, h->t_ptr);
    int rc = inclusive_scan(ctx, h->t_ptr, (int64_t)n + 1);
    if (rc == SOV_OK && nnz > 0) {
        k_copy_i32<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, h->t_ptr, cursor);
        k_fill_transpose<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(m, nnz, d_rp, d_ci, cursor, trow_u, tsrc_u);
        k_rank_transpose<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(n, nnz, h->t_ptr, trow_u, tsrc_u,
                                                                          h->t_idx, h->t_src);
        k_gather_vals<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, h->t_src, h->a.val64, h->t.val64);
        e = check_launch();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "csr transpose");
    }
    (void)hipFree(cursor);
    (void)hipFree(trow_u);
    (void)hipFree(tsrc_u);
    if (rc != SOV_OK) return cleanup(rc);
    h->t.rows = n;
    h->t.cols = m;
    h->t.ptr = h->t_ptr;
    h->t.idx = h->t_idx;
    rc = build_schedule(ctx, h->a);
    if (rc == SOV_OK) rc = build_schedule(ctx, h->t);
    if (rc == SOV_OK) rc = ensure_partials(h);
    if (rc != SOV_OK) return cleanup(rc);
    *out = h;
    return SOV_OK;
}

void sov_hip_csr_destroy(sov_hip_csr* h) {
    if (h == nullptr) return;
    DeviceGuard2 g(h->device);
    (void)hipDeviceSynchronize();
    if (h->a.owns_val64) (void)hipFree(h->a.val64);
    (void)hipFree(h->a.val32);
    (void)hipFree(h->a.list);
    (void)hipFree(h->t.val64);
    (void)hipFree(h->t.val32);
    (void)hipFree(h->t.list);
    (void)hipFree(h->t_ptr);
    (void)hipFree(h->t_idx);
    (void)hipFree(h->t_src);
    (void)hipFree(h->partials);
    delete h;
}

int sov_hip_csr_dims(const sov_hip_csr* h, int32_t* m, int32_t* n, int32_t* nnz) {
    if (h == nullptr) return fail(SOV_INVALID_ARGUMENT, "csr is null");
    if (m) *m = h->m;
    if (n) *n = h->n;
    if (nnz) *nnz = h->nnz;
    return SOV_OK;
}

int sov_hip_csr_make_f32(sov_hip_ctx* ctx, sov_hip_csr* h) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "make_f32 args");
    DeviceGuard2 g(ctx->device);
    const int64_t nz = std::max<int32_t>(h->nnz, 1);
    if (h->a.val32 == nullptr) SOV_HIP_TRY(hipMalloc(&h->a.val32, sizeof(float) * nz));
    if (h->t.val32 == nullptr) SOV_HIP_TRY(hipMalloc(&h->t.val32, sizeof(float) * nz));
    if (h->nnz > 0) {
        k_to_f32<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, h->a.val64, h->a.val32);
        k_to_f32<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, h->t.val64, h->t.val32);
        SOV_HIP_TRY(check_launch());
    }
    return SOV_OK;
}

int sov_hip_csr_scale(sov_hip_ctx* ctx, sov_hip_csr* h, const double* rs, const double* cs) {
    if (ctx == nullptr || h == nullptr || rs == nullptr || cs == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "csr_scale args");
    if (!h->a.owns_val64) return fail(SOV_INVALID_ARGUMENT, "csr_scale needs SOV_CSR_OWN_VALUES");
    DeviceGuard2 g(ctx->device);
    if (h->nnz == 0) return SOV_OK;
    k_scale_vals<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, h->a.ptr, h->a.idx, h->a.val64, rs, cs);
    k_gather_vals<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, h->t_src, h->a.val64, h->t.val64);
    SOV_HIP_TRY(check_launch());
    if (h->a.val32 != nullptr) return sov_hip_csr_make_f32(ctx, h);
    return SOV_OK;
}

int sov_hip_csr_transpose_arrays(sov_hip_ctx* ctx, const sov_hip_csr* h, int32_t* d_tptr, int32_t* d_tidx,
                                 double* d_tval) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "transpose_arrays args");
    DeviceGuard2 g(ctx->device);
    if (d_tptr) SOV_HIP_TRY(hipMemcpyAsync(d_tptr, h->t_ptr, sizeof(int32_t) * ((int64_t)h->n + 1),
                                           hipMemcpyDeviceToDevice, ctx->stream));
    if (d_tidx && h->nnz > 0)
        SOV_HIP_TRY(hipMemcpyAsync(d_tidx, h->t_idx, sizeof(int32_t) * h->nnz, hipMemcpyDeviceToDevice, ctx->stream));
    if (d_tval && h->nnz > 0)
        SOV_HIP_TRY(hipMemcpyAsync(d_tval, h->t.val64, sizeof(double) * h->nnz, hipMemcpyDeviceToDevice, ctx->stream));
    return SOV_OK;
}

int sov_hip_csr_set_lanes(sov_hip_csr* h, int orient, const int32_t lanes[8]) {
    if (h == nullptr || lanes == nullptr || (orient != 0 && orient != 1))
        return fail(SOV_INVALID_ARGUMENT, "set_lanes args");
    Orient& o = orient == 0 ? h->a : h->t;
    for (int c = 0; c < kClasses; ++c) {
        const int G = lanes[c];
        if (G == 0) { o.lanes[c] = kDefaultLanes[c]; continue; }
        if (G < 1 || G > 64 || (G & (G - 1)) != 0) return fail(SOV_INVALID_ARGUMENT, "lanes must be a power of two <= 64");
        o.lanes[c] = G;
    }
    return ensure_partials(h);
}

int sov_hip_csr_get_lanes(const sov_hip_csr* h, int orient, int32_t lanes[8]) {
    if (h == nullptr || lanes == nullptr || (orient != 0 && orient != 1))
        return fail(SOV_INVALID_ARGUMENT, "get_lanes args");
    const Orient& o = orient == 0 ? h->a : h->t;
    for (int c = 0; c < kClasses; ++c) lanes[c] = o.lanes[c];
    return SOV_OK;
}

int sov_hip_spmv(sov_hip_ctx* ctx, const sov_hip_csr* h, int transpose, int prec, double alpha, const void* x,
                 double beta, void* y) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv args");
    if (prec != SOV_PREC_F64 && prec != SOV_PREC_F32) return fail(SOV_INVALID_ARGUMENT, "spmv: precision");
    const Orient& o = transpose ? h->t : h->a;
    if (o.rows > 0 && y == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv: y is null");
    if (h->nnz > 0 && x == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv: x is null");
    DeviceGuard2 g(ctx->device);
    return spmv_dispatch(ctx, o, prec, alpha, x, beta, y, o.sched());
}

int sov_hip_csr_tune(sov_hip_ctx* ctx, sov_hip_csr* h, int prec) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "tune args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "tune during capture");
    if (prec == SOV_PREC_F32 && h->a.val32 == nullptr) return fail(SOV_INVALID_ARGUMENT, "tune f32 needs make_f32");
    DeviceGuard2 g(ctx->device);
    const size_t es = prec == SOV_PREC_F64 ? sizeof(double) : sizeof(float);
    void *xv = nullptr, *yv = nullptr;
    const int64_t len = std::max<int64_t>(std::max(h->m, h->n), 1);
    SOV_HIP_TRY(hipMalloc(&xv, es * len));
    hipError_t e = hipMalloc(&yv, es * len);
    if (e != hipSuccess) { (void)hipFree(xv); return hip_fail(e, "tune buffers"); }
    (void)hipMemsetAsync(xv, 0, es * len, ctx->stream);
    (void)hipMemsetAsync(yv, 0, es * len, ctx->stream);
    static const int cand[7] = {1, 2, 4, 8, 16, 32, 64};
    int rc = SOV_OK;
    for (int orient = 0; orient < 2 && rc == SOV_OK; ++orient) {
        Orient& o = orient == 0 ? h->a : h->t;
        for (int c = 0; c < kClasses && rc == SOV_OK; ++c) {
            if (o.class_count[c] == 0) continue;
            int best = o.lanes[c];
            float best_ms = 1e30f;
            for (int gi = 0; gi < 7 && rc == SOV_OK; ++gi) {
                const int G = cand[gi];
                // a schedule holding only class c, with G lanes per row
                Orient tmp = o;
                for (int k = 0; k < kClasses; ++k) tmp.class_count[k] = (k == c) ? o.class_count[c] : 0;
                tmp.lanes[c] = G;
                const Sched s = tmp.sched();
                float tot = 1e30f;
                for (int rep = 0; rep < 4 && rc == SOV_OK; ++rep) {
                    (void)hipEventRecord(ctx->ev0, ctx->stream);
                    rc = spmv_dispatch(ctx, o, prec, 1.0, xv, 0.0, yv, s);
                    (void)hipEventRecord(ctx->ev1, ctx->stream);
                    (void)hipEventSynchronize(ctx->ev1);
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
                    if (rep > 0) tot = std::min(tot, ms);   // first run warms caches
                }
                if (tot < best_ms) { best_ms = tot; best = G; }
            }
            o.lanes[c] = best;
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(xv);
    (void)hipFree(yv);
    if (rc != SOV_OK) return rc;
    return ensure_partials(h);
}

int sov_hip_scale_ruiz_sweep(sov_hip_ctx* ctx, const sov_hip_csr* h, double* rs, double* cs, double* worst) {
    if (ctx == nullptr || h == nullptr || rs == nullptr || cs == nullptr || worst == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "ruiz args");
    DeviceGuard2 g(ctx->device);
    const int64_t need = (int64_t)h->m + h->n + 1024 + 8;
    int rc = ctx->ensure_red(need);
    if (rc != SOV_OK) return rc;
    double* rmax = ctx->d_red;
    double* cmax = rmax + h->m;
    double* part = cmax + h->n;
    if (h->m > 0)
        k_ruiz_rowmax<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, h->a.ptr, h->a.idx, h->a.val64, rs, cs, rmax);
    if (h->n > 0)
        k_ruiz_colmax<<<launch_blocks(h->n), kBlock, 0, ctx->stream>>>(h->n, h->t.ptr, h->t.idx, h->t.val64, rs, cs, cmax);
    int blocks = launch_blocks((int64_t)h->m + h->n);
    if (blocks > 1024) blocks = 1024;
    k_ruiz_update<<<blocks, kBlock, 0, ctx->stream>>>(h->m, h->n, rmax, cmax, rs, cs, part);
    k_max_partials<<<1, kBlock, 0, ctx->stream>>>(blocks, part, part + 1024);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, part + 1024, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *worst = ctx->h_pinned[0];
    return SOV_OK;
}

int sov_hip_scale_pock_chambolle(sov_hip_ctx* ctx, const sov_hip_csr* h, double alpha, double* rs, double* cs) {
    if (ctx == nullptr || h == nullptr || rs == nullptr || cs == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "pock_chambolle args");
    DeviceGuard2 g(ctx->device);
    int rc = ctx->ensure_red((int64_t)h->m + h->n + 8);
    if (rc != SOV_OK) return rc;
    double* rsum = ctx->d_red;
    double* csum = rsum + h->m;
    if (h->m > 0)
        k_pc_rowsum<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, h->a.ptr, h->a.idx, h->a.val64, rs, cs,
                                                                      2.0 - alpha, rsum);
    if (h->n > 0)
        k_pc_colsum<<<launch_blocks(h->n), kBlock, 0, ctx->stream>>>(h->n, h->t.ptr, h->t.idx, h->t.val64, rs, cs,
                                                                      alpha, csum);
    if ((int64_t)h->m + h->n > 0)
        k_pc_update<<<launch_blocks((int64_t)h->m + h->n), kBlock, 0, ctx->stream>>>(h->m, h->n, rsum, csum, rs, cs);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

}  // extern "C"
