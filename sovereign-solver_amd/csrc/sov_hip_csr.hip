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

Sched Part::sched() const {
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

int Part::blocks() const { return sched().block_begin[kClasses]; }

int Orient::max_blocks() const {
    int b = whole.blocks();
    for (const Part& p : seg) b = std::max(b, p.blocks());
    return b;
}

void Orient::free_segments() {
    for (Part& p : seg) (void)hipFree(p.list);
    seg.clear();
    seg_width = 0;
    (void)hipFree(seg_ptr);
    (void)hipFree(seg_idx);
    (void)hipFree(seg_src);
    (void)hipFree(seg_val64);
    (void)hipFree(seg_val32);
    (void)hipFree(acc64);
    (void)hipFree(acc32);
    seg_ptr = seg_idx = seg_src = nullptr;
    seg_val64 = nullptr;
    seg_val32 = nullptr;
    acc64 = nullptr;
    acc32 = nullptr;
}

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
int build_schedule(sov_hip_ctx* ctx, Part& o) {
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
    (void)hipFree(o.list);
    o.list = nullptr;
    if (o.rows > 0) {
        SOV_HIP_TRY(hipMalloc(&o.list, sizeof(int32_t) * o.rows));
        SOV_HIP_TRY(hipMemcpyAsync(o.list, list.data(), sizeof(int32_t) * o.rows, hipMemcpyHostToDevice,
                                   ctx->stream));
        SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return SOV_OK;
}

int ensure_partials(sov_hip_csr* csr) {
    const int64_t need = (int64_t)csr->a.max_blocks() + (int64_t)csr->t.max_blocks() + 64;
    if (need <= csr->partial_blocks) return SOV_OK;
    if (csr->partials) (void)hipFree(csr->partials);
    csr->partials = nullptr;
    csr->partial_blocks = 0;
    SOV_HIP_TRY(hipMalloc(&csr->partials, sizeof(double) * kCheckSlots * need));
    csr->partial_blocks = need;
    return SOV_OK;
}

// ---------------------------------------------------------------------------
// column segments
// ---------------------------------------------------------------------------

// bounds[k * rows + r] = first position of row r whose column is >= k * W,
// for k = 0..K (k = 0 is ptr[r], k = K is ptr[r + 1]).
__global__ void k_seg_bounds(int32_t rows, int32_t K, int32_t W, const int32_t* __restrict__ ptr,
                             const int32_t* __restrict__ idx, int32_t* __restrict__ bounds) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = ptr[r], e = ptr[r + 1];
        bounds[r] = b;
        for (int32_t k = 1; k < K; ++k) {
            const int64_t key = (int64_t)k * W;
            int32_t lo = b, hi = e;   // first position with idx >= key
            while (lo < hi) {
                const int32_t mid = lo + (hi - lo) / 2;
                if ((int64_t)idx[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            bounds[(int64_t)k * rows + r] = lo;
        }
        bounds[(int64_t)K * rows + r] = e;
    }
}

// Row lengths of segment k into sptr[k * (rows+1) + r + 1] (sptr[.. + 0] = 0).
__global__ void k_seg_counts(int32_t rows, int32_t K, const int32_t* __restrict__ bounds, int32_t* __restrict__ sptr) {
    const int64_t total = (int64_t)K * rows;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = t / rows, r = t % rows;
        sptr[k * (rows + 1) + r + 1] = bounds[(k + 1) * rows + r] - bounds[k * rows + r];
        if (r == 0) sptr[k * (rows + 1)] = 0;
    }
}

// Copy the entries of every (row, segment) run into segment-major order.
__global__ void k_seg_fill(int32_t rows, int32_t K, const int32_t* __restrict__ bounds,
                           const int32_t* __restrict__ sptr, const int32_t* __restrict__ base,
                           const int32_t* __restrict__ idx, const int32_t* __restrict__ src_in,
                           int32_t* __restrict__ sidx, int32_t* __restrict__ ssrc) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        for (int32_t k = 0; k < K; ++k) {
            const int32_t b = bounds[(int64_t)k * rows + r], e = bounds[(int64_t)(k + 1) * rows + r];
            int64_t d = (int64_t)base[k] + sptr[(int64_t)k * (rows + 1) + r];
            for (int32_t p = b; p < e; ++p, ++d) {
                sidx[d] = idx[p];
                ssrc[d] = src_in ? src_in[p] : p;
            }
        }
    }
}

__global__ void k_gather_f64_to_f32(int32_t nnz, const int32_t* __restrict__ src, const double* __restrict__ v,
                                    float* __restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x)
        out[p] = (float)v[src[p]];
}

// Build (or rebuild) the K-segment copy of an orientation.  K <= 1 removes it.
static int build_segments(sov_hip_ctx* ctx, sov_hip_csr* h, Orient& o, int K, bool want_f32) {
    o.free_segments();
    const Part& w = o.whole;
    if (K <= 1 || w.rows == 0 || w.nnz == 0) return SOV_OK;
    const int32_t rows = w.rows, W = (w.cols + K - 1) / K;
    if (W <= 0) return SOV_OK;
    int32_t* bounds = nullptr;
    int32_t* base_d = nullptr;
    auto fail_free = [&](hipError_t e, const char* what) {
        (void)hipFree(bounds);
        (void)hipFree(base_d);
        o.free_segments();
        return hip_fail(e, what);
    };
    hipError_t e = hipMalloc(&bounds, sizeof(int32_t) * (int64_t)(K + 1) * rows);
    if (e == hipSuccess) e = hipMalloc(&o.seg_ptr, sizeof(int32_t) * (int64_t)K * (rows + 1));
    if (e == hipSuccess) e = hipMalloc(&o.seg_idx, sizeof(int32_t) * w.nnz);
    if (e == hipSuccess) e = hipMalloc(&o.seg_src, sizeof(int32_t) * w.nnz);
    if (e == hipSuccess) e = hipMalloc(&o.seg_val64, sizeof(double) * w.nnz);
    if (e == hipSuccess) e = hipMalloc(&o.acc64, sizeof(double) * rows);
    if (e == hipSuccess) e = hipMalloc(&o.acc32, sizeof(float) * rows);
    if (e == hipSuccess) e = hipMalloc(&base_d, sizeof(int32_t) * (K + 1));
    if (e != hipSuccess) return fail_free(e, "segment buffers");
    k_seg_bounds<<<launch_blocks(rows), kBlock, 0, ctx->stream>>>(rows, K, W, w.ptr, w.idx, bounds);
    k_seg_counts<<<launch_blocks((int64_t)K * rows), kBlock, 0, ctx->stream>>>(rows, K, bounds, o.seg_ptr);
    e = check_launch();
    if (e != hipSuccess) return fail_free(e, "segment counts");
    std::vector<int32_t> base(static_cast<size_t>(K) + 1, 0);
    for (int k = 0; k < K; ++k) {
        const int rc = inclusive_scan(ctx, o.seg_ptr + (int64_t)k * (rows + 1), rows + 1);
        if (rc != SOV_OK) {
            (void)hipFree(bounds);
            (void)hipFree(base_d);
            o.free_segments();
            return rc;
        }
        int32_t nk = 0;
        e = hipMemcpy(&nk, o.seg_ptr + (int64_t)k * (rows + 1) + rows, sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail_free(e, "segment sizes");
        base[static_cast<size_t>(k) + 1] = base[static_cast<size_t>(k)] + nk;
    }
    e = hipMemcpyAsync(base_d, base.data(), sizeof(int32_t) * (K + 1), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail_free(e, "segment bases");
    k_seg_fill<<<launch_blocks(rows), kBlock, 0, ctx->stream>>>(rows, K, bounds, o.seg_ptr, base_d, w.idx, w.src,
                                                                 o.seg_idx, o.seg_src);
    k_gather_vals<<<launch_blocks(w.nnz), kBlock, 0, ctx->stream>>>(w.nnz, o.seg_src, h->a.whole.val64, o.seg_val64);
    e = check_launch();
    if (e == hipSuccess && want_f32) {
        e = hipMalloc(&o.seg_val32, sizeof(float) * w.nnz);
        if (e == hipSuccess)
            k_gather_f64_to_f32<<<launch_blocks(w.nnz), kBlock, 0, ctx->stream>>>(w.nnz, o.seg_src, h->a.whole.val64,
                                                                                o.seg_val32);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(bounds);
    bounds = nullptr;
    (void)hipFree(base_d);
    base_d = nullptr;
    if (e != hipSuccess) return fail_free(e, "segment fill");
    o.seg_width = W;
    o.seg.resize(static_cast<size_t>(K));
    for (int k = 0; k < K; ++k) {
        Part& p = o.seg[static_cast<size_t>(k)];
        p.rows = rows;
        p.cols = w.cols;
        p.nnz = base[static_cast<size_t>(k) + 1] - base[static_cast<size_t>(k)];
        p.ptr = o.seg_ptr + (int64_t)k * (rows + 1);
        p.idx = o.seg_idx + base[static_cast<size_t>(k)];
        p.src = o.seg_src + base[static_cast<size_t>(k)];
        p.val64 = o.seg_val64 + base[static_cast<size_t>(k)];
        p.val32 = o.seg_val32 ? o.seg_val32 + base[static_cast<size_t>(k)] : nullptr;
        const int rc = build_schedule(ctx, p);
        if (rc != SOV_OK) {
            o.free_segments();
            return rc;
        }
    }
    return SOV_OK;
}

// Re-derive every value image of both orientations from A's FP64 values.
static int refresh_values(sov_hip_ctx* ctx, sov_hip_csr* h) {
    if (h->nnz == 0) return SOV_OK;
    const double* av = h->a.whole.val64;
    k_gather_vals<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, h->t_src, av, h->t.whole.val64);
    for (Orient* o : {&h->a, &h->t}) {
        if (o->whole.val32 != nullptr && o != &h->a)
            k_gather_f64_to_f32<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, h->t_src, av, o->whole.val32);
        if (o->seg_val64 != nullptr)
            k_gather_vals<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, o->seg_src, av, o->seg_val64);
        if (o->seg_val32 != nullptr)
            k_gather_f64_to_f32<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, o->seg_src, av, o->seg_val32);
    }
    if (h->a.whole.val32 != nullptr)
        k_to_f32<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, av, h->a.whole.val32);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

// ---------------------------------------------------------------------------
// SpMV over the parts of an orientation
// ---------------------------------------------------------------------------

template <typename V, typename T>
struct SpmvOp {
    const int32_t* ptr;
    const int32_t* idx;
    const V* val;
    const T* x;
    T* y;
    T* acc;
    T alpha, beta;
    int mode;
    Sched s;
    template <int G>
    __device__ void run(const ClassView& v) {
        int lane;
        const int row = row_of<G>(s, v, lane);
        T sum = row_dot<G, T>(ptr, idx, val, x, row, lane);
        if (row >= 0 && lane == 0) {
            if (mode == kFirst) { acc[row] = sum; return; }
            if (mode == kMiddle) { acc[row] = acc[row] + sum; return; }
            if (mode == kLast) sum = acc[row] + sum;
            y[row] = (beta == T(0)) ? alpha * sum : alpha * sum + beta * y[row];
        }
    }
};

template <typename V, typename T>
__global__ void __launch_bounds__(kBlock) k_spmv(SpmvOp<V, T> op) {
    const ClassView v = class_view(op.s);
    dispatch_G(v, op);
}

template <typename V, typename T>
static int launch_spmv_part(sov_hip_ctx* ctx, const Part& p, const V* val, int mode, double alpha, const void* x,
                            double beta, void* y, T* acc, const Sched& s) {
    const int blocks = s.block_begin[kClasses];
    if (blocks == 0) return SOV_OK;
    SpmvOp<V, T> op{p.ptr, p.idx, val, static_cast<const T*>(x), static_cast<T*>(y), acc, T(alpha), T(beta), mode, s};
    k_spmv<V, T><<<blocks, kBlock, 0, ctx->stream>>>(op);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

static int spmv_orient(sov_hip_ctx* ctx, const Orient& o, int prec, double alpha, const void* x, double beta,
                       void* y) {
    const int np = o.nparts();
    for (int k = 0; k < np; ++k) {
        const Part& p = o.part(k);
        const int mode = seg_mode(k, np);
        int rc;
        if (prec == SOV_PREC_F64) {
            rc = launch_spmv_part<double, double>(ctx, p, p.val64, mode, alpha, x, beta, y, o.acc64, p.sched());
        } else {
            if (p.val32 == nullptr) return fail(SOV_INVALID_ARGUMENT, "FP32 SpMV needs sov_hip_csr_make_f32 first");
            rc = launch_spmv_part<float, float>(ctx, p, p.val32, mode, alpha, x, beta, y, o.acc32, p.sched());
        }
        if (rc != SOV_OK) return rc;
    }
    return SOV_OK;
}

// ---------------------------------------------------------------------------
// scaling kernels (thread per row, serial over the row in index order)
// ---------------------------------------------------------------------------

// Ruiz: out[r] = max_k (|v_k| * r_r) * c_{idx_k}  (the product order of the
// reference, scaling.cpp:48-50; the A^T pass forms (|v| * r_i) * c_j too)
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

// Time one SpMV of an orientation (min over reps after a warm run).
static float time_spmv(sov_hip_ctx* ctx, const Orient& o, int prec, const void* x, void* y, int reps) {
    float best = 1e30f;
    for (int rep = 0; rep <= reps; ++rep) {
        (void)hipEventRecord(ctx->ev0, ctx->stream);
        if (spmv_orient(ctx, o, prec, 1.0, x, 0.0, y) != SOV_OK) return 1e30f;
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        (void)hipEventSynchronize(ctx->ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        if (rep > 0) best = std::min(best, ms);
    }
    return best;
}

// Tune the lanes of every class of one part against the orientation's SpMV.
static int tune_part_lanes(sov_hip_ctx* ctx, Orient& o, Part& p, int prec, const void* x, void* y) {
    static const int cand[7] = {1, 2, 4, 8, 16, 32, 64};
    for (int c = 0; c < kClasses; ++c) {
        if (p.class_count[c] == 0) continue;
        int best = p.lanes[c];
        float best_ms = 1e30f;
        for (int G : cand) {
            // a schedule holding only class c of this part
            Part tmp = p;
            for (int k = 0; k < kClasses; ++k) tmp.class_count[k] = (k == c) ? p.class_count[c] : 0;
            tmp.lanes[c] = G;
            const Sched s = tmp.sched();
            float ms_best = 1e30f;
            for (int rep = 0; rep < 4; ++rep) {
                (void)hipEventRecord(ctx->ev0, ctx->stream);
                int rc;
                if (prec == SOV_PREC_F64)
                    rc = launch_spmv_part<double, double>(ctx, p, p.val64, kOnly, 1.0, x, 0.0, y, o.acc64, s);
                else
                    rc = launch_spmv_part<float, float>(ctx, p, p.val32, kOnly, 1.0, x, 0.0, y, o.acc32, s);
                if (rc != SOV_OK) return rc;
                (void)hipEventRecord(ctx->ev1, ctx->stream);
                (void)hipEventSynchronize(ctx->ev1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
                if (rep > 0) ms_best = std::min(ms_best, ms);
            }
            if (ms_best < best_ms) {
                best_ms = ms_best;
                best = G;
            }
        }
        p.lanes[c] = best;
    }
    return SOV_OK;
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
    Part& A = h->a.whole;
    A.rows = m;
    A.cols = n;
    A.nnz = nnz;
    A.ptr = d_rp;
    A.idx = d_ci;
    if (flags & SOV_CSR_OWN_VALUES) {
        hipError_t e = hipMalloc(&A.val64, sizeof(double) * nz);
        if (e != hipSuccess) return cleanup(hip_fail(e, "csr values"));
        h->owns_a_val64 = true;
        if (nnz > 0) k_copy_f64<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, d_val, A.val64);
    } else {
        A.val64 = const_cast<double*>(d_val);
    }
    // A^T structure, built on the device
    Part& T = h->t.whole;
    int32_t *cursor = nullptr, *trow_u = nullptr, *tsrc_u = nullptr;
    hipError_t e = hipMalloc(&h->t_ptr, sizeof(int32_t) * ((int64_t)n + 1));
    if (e == hipSuccess) e = hipMalloc(&h->t_idx, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&h->t_src, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&T.val64, sizeof(double) * nz);
    if (e == hipSuccess) e = hipMalloc(&cursor, sizeof(int32_t) * ((int64_t)n + 1));
    if (e == hipSuccess) e = hipMalloc(&trow_u, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&tsrc_u, sizeof(int32_t) * nz);
    if (e != hipSuccess) {
        (void)hipFree(cursor); (void)hipFree(trow_u); (void)hipFree(tsrc_u);
        return cleanup(hip_fail(e, "csr transpose buffers"));
    }
    k_zero_i32<<<launch_blocks((int64_t)n + 1), kBlock, 0, ctx->stream>>>((int64_t)n + 1, h->t_ptr);
    if (nnz > 0) k_count_cols<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, d_ci, h->t_ptr);
    int rc = inclusive_scan(ctx, h->t_ptr, (int64_t)n + 1);
    if (rc == SOV_OK && nnz > 0) {
        k_copy_i32<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, h->t_ptr, cursor);
        k_fill_transpose<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(m, nnz, d_rp, d_ci, cursor, trow_u, tsrc_u);
        k_rank_transpose<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(n, nnz, h->t_ptr, trow_u, tsrc_u,
                                                                          h->t_idx, h->t_src);
        k_gather_vals<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, h->t_src, A.val64, T.val64);
        e = check_launch();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "csr transpose");
    }
    (void)hipFree(cursor);
    (void)hipFree(trow_u);
    (void)hipFree(tsrc_u);
    if (rc != SOV_OK) return cleanup(rc);
    T.rows = n;
    T.cols = m;
    T.nnz = nnz;
    T.ptr = h->t_ptr;
    T.idx = h->t_idx;
    T.src = h->t_src;
    rc = build_schedule(ctx, A);
    if (rc == SOV_OK) rc = build_schedule(ctx, T);
    if (rc == SOV_OK) rc = ensure_partials(h);
    if (rc != SOV_OK) return cleanup(rc);
    *out = h;
    return SOV_OK;
}

void sov_hip_csr_destroy(sov_hip_csr* h) {
    if (h == nullptr) return;
    DeviceGuard2 g(h->device);
    (void)hipDeviceSynchronize();
    h->a.free_segments();
    h->t.free_segments();
    if (h->owns_a_val64) (void)hipFree(h->a.whole.val64);
    (void)hipFree(h->a.whole.val32);
    (void)hipFree(h->a.whole.list);
    (void)hipFree(h->t.whole.val64);
    (void)hipFree(h->t.whole.val32);
    (void)hipFree(h->t.whole.list);
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
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "make_f32 during capture");
    DeviceGuard2 g(ctx->device);
    const int64_t nz = std::max<int32_t>(h->nnz, 1);
    if (h->a.whole.val32 == nullptr) SOV_HIP_TRY(hipMalloc(&h->a.whole.val32, sizeof(float) * nz));
    if (h->t.whole.val32 == nullptr) SOV_HIP_TRY(hipMalloc(&h->t.whole.val32, sizeof(float) * nz));
    for (Orient* o : {&h->a, &h->t}) {
        if (o->seg_val64 != nullptr && o->seg_val32 == nullptr) {
            SOV_HIP_TRY(hipMalloc(&o->seg_val32, sizeof(float) * nz));
            int64_t off = 0;
            for (Part& p : o->seg) {
                p.val32 = o->seg_val32 + off;
                off += p.nnz;
            }
        }
    }
    return refresh_values(ctx, h);
}

int sov_hip_csr_scale(sov_hip_ctx* ctx, sov_hip_csr* h, const double* rs, const double* cs) {
    if (ctx == nullptr || h == nullptr || rs == nullptr || cs == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "csr_scale args");
    if (!h->owns_a_val64) return fail(SOV_INVALID_ARGUMENT, "csr_scale needs SOV_CSR_OWN_VALUES");
    DeviceGuard2 g(ctx->device);
    if (h->nnz == 0) return SOV_OK;
    k_scale_vals<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, h->a.whole.ptr, h->a.whole.idx,
                                                                   h->a.whole.val64, rs, cs);
    SOV_HIP_TRY(check_launch());
    return refresh_values(ctx, h);
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
        SOV_HIP_TRY(hipMemcpyAsync(d_tval, h->t.whole.val64, sizeof(double) * h->nnz, hipMemcpyDeviceToDevice,
                                   ctx->stream));
    return SOV_OK;
}

int sov_hip_csr_set_lanes(sov_hip_csr* h, int orient, const int32_t lanes[8]) {
    if (h == nullptr || lanes == nullptr || (orient != 0 && orient != 1))
        return fail(SOV_INVALID_ARGUMENT, "set_lanes args");
    Orient& o = orient == 0 ? h->a : h->t;
    for (int c = 0; c < kClasses; ++c) {
        const int G = lanes[c];
        if (G != 0 && (G < 1 || G > 64 || (G & (G - 1)) != 0))
            return fail(SOV_INVALID_ARGUMENT, "lanes must be a power of two <= 64");
    }
    for (int k = -1; k < (int)o.seg.size(); ++k) {
        Part& p = k < 0 ? o.whole : o.seg[static_cast<size_t>(k)];
        for (int c = 0; c < kClasses; ++c) p.lanes[c] = lanes[c] == 0 ? kDefaultLanes[c] : lanes[c];
    }
    return ensure_partials(h);
}

int sov_hip_csr_get_lanes(const sov_hip_csr* h, int orient, int32_t lanes[8]) {
    if (h == nullptr || lanes == nullptr || (orient != 0 && orient != 1))
        return fail(SOV_INVALID_ARGUMENT, "get_lanes args");
    const Orient& o = orient == 0 ? h->a : h->t;
    for (int c = 0; c < kClasses; ++c) lanes[c] = o.part(0).lanes[c];
    return SOV_OK;
}

int sov_hip_csr_set_segments(sov_hip_ctx* ctx, sov_hip_csr* h, int orient, int count) {
    if (ctx == nullptr || h == nullptr || (orient != 0 && orient != 1) || count < 1 || count > 64)
        return fail(SOV_INVALID_ARGUMENT, "set_segments args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "set_segments during capture");
    DeviceGuard2 g(ctx->device);
    Orient& o = orient == 0 ? h->a : h->t;
    const int rc = build_segments(ctx, h, o, count, h->a.whole.val32 != nullptr);
    if (rc != SOV_OK) return rc;
    return ensure_partials(h);
}

int sov_hip_csr_get_segments(const sov_hip_csr* h, int orient, int* count) {
    if (h == nullptr || count == nullptr || (orient != 0 && orient != 1))
        return fail(SOV_INVALID_ARGUMENT, "get_segments args");
    *count = (orient == 0 ? h->a : h->t).nparts();
    return SOV_OK;
}

int sov_hip_spmv(sov_hip_ctx* ctx, const sov_hip_csr* h, int transpose, int prec, double alpha, const void* x,
                 double beta, void* y) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv args");
    if (prec != SOV_PREC_F64 && prec != SOV_PREC_F32) return fail(SOV_INVALID_ARGUMENT, "spmv: precision");
    const Orient& o = transpose ? h->t : h->a;
    if (o.whole.rows > 0 && y == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv: y is null");
    if (h->nnz > 0 && x == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv: x is null");
    DeviceGuard2 g(ctx->device);
    return spmv_orient(ctx, o, prec, alpha, x, beta, y);
}

int sov_hip_csr_tune(sov_hip_ctx* ctx, sov_hip_csr* h, int prec) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "tune args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "tune during capture");
    if (prec == SOV_PREC_F32 && h->a.whole.val32 == nullptr) return fail(SOV_INVALID_ARGUMENT, "tune f32 needs make_f32");
    DeviceGuard2 g(ctx->device);
    const size_t es = prec == SOV_PREC_F64 ? sizeof(double) : sizeof(float);
    void *xv = nullptr, *yv = nullptr;
    const int64_t len = std::max<int64_t>(std::max(h->m, h->n), 1);
    SOV_HIP_TRY(hipMalloc(&xv, es * len));
    hipError_t e = hipMalloc(&yv, es * len);
    if (e != hipSuccess) { (void)hipFree(xv); return hip_fail(e, "tune buffers"); }
    (void)hipMemsetAsync(xv, 0, es * len, ctx->stream);
    (void)hipMemsetAsync(yv, 0, es * len, ctx->stream);
    int rc = SOV_OK;
    const bool f32 = h->a.whole.val32 != nullptr;
    for (int orient = 0; orient < 2 && rc == SOV_OK; ++orient) {
        Orient& o = orient == 0 ? h->a : h->t;
        // segment counts worth trying: keep each gathered slice >= 64k elements
        const int64_t cols = o.whole.cols;
        int best_K = 1;
        float best_ms = 1e30f;
        for (int K : {1, 2, 4, 8}) {
            if (K > 1 && cols / K < 65536) break;
            rc = build_segments(ctx, h, o, K, f32);
            if (rc != SOV_OK) break;
            for (int k = 0; k < o.nparts() && rc == SOV_OK; ++k) rc = tune_part_lanes(ctx, o, o.part(k), prec, xv, yv);
            if (rc != SOV_OK) break;
            const float ms = time_spmv(ctx, o, prec, xv, yv, 4);
            if (ms < best_ms) {
                best_ms = ms;
                best_K = K;
            }
        }
        if (rc == SOV_OK && best_K != o.nparts()) {
            rc = build_segments(ctx, h, o, best_K, f32);
            for (int k = 0; k < o.nparts() && rc == SOV_OK; ++k) rc = tune_part_lanes(ctx, o, o.part(k), prec, xv, yv);
        }
        // the whole part (checks, primitive SpMV fallback) keeps tuned lanes too
        if (rc == SOV_OK && !o.seg.empty()) {
            Orient whole_only;
            whole_only.whole = o.whole;
            rc = tune_part_lanes(ctx, whole_only, o.whole, prec, xv, yv);
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
    const Part& A = h->a.whole;
    const Part& T = h->t.whole;
    if (h->m > 0)
        k_ruiz_rowmax<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, A.ptr, A.idx, A.val64, rs, cs, rmax);
    if (h->n > 0)
        k_ruiz_colmax<<<launch_blocks(h->n), kBlock, 0, ctx->stream>>>(h->n, T.ptr, T.idx, T.val64, rs, cs, cmax);
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
    const Part& A = h->a.whole;
    const Part& T = h->t.whole;
    if (h->m > 0)
        k_pc_rowsum<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, A.ptr, A.idx, A.val64, rs, cs,
                                                                      2.0 - alpha, rsum);
    if (h->n > 0)
        k_pc_colsum<<<launch_blocks(h->n), kBlock, 0, ctx->stream>>>(h->n, T.ptr, T.idx, T.val64, rs, cs,
                                                                      alpha, csum);
    if ((int64_t)h->m + h->n > 0)
        k_pc_update<<<launch_blocks((int64_t)h->m + h->n), kBlock, 0, ctx->stream>>>(h->m, h->n, rsum, csum, rs, cs);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

}  // extern "C"
