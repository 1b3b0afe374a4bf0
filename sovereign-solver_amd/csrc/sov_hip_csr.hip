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

#include <rocprim/device/device_radix_sort.hpp>

#include "sov_hip_internal.hpp"
#include "sov_hip_tiles.hpp"

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
    return s;
}

int Part::blocks() const { return sched().block_begin[kClasses]; }

int Orient::max_blocks() const {
    int b = 0;
    for (const Part& p : parts) b = std::max(b, p.blocks());
    return b;
}

void Orient::free_parts() {
    parts.clear();
    seg_width = 0;
    (void)hipFree(p_ptr);
    (void)hipFree(p_idx);
    (void)hipFree(p_src);
    (void)hipFree(p_val64);
    (void)hipFree(p_val32);
    (void)hipFree(acc64);
    (void)hipFree(acc32);
    (void)hipFree(chk_acc);
    p_ptr = p_idx = p_src = nullptr;
    p_val64 = nullptr;
    p_val32 = nullptr;
    acc64 = nullptr;
    acc32 = nullptr;
    chk_acc = nullptr;
}

void Orient::free_order() {
    (void)hipFree(perm);
    (void)hipFree(inv);
    (void)hipFree(d_ptr);
    (void)hipFree(d_idx);
    (void)hipFree(d_src);
    perm = inv = d_ptr = d_idx = d_src = nullptr;
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

int inclusive_scan(sov_hip_ctx* ctx, int32_t* a, int64_t len) {
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

// The transpose as one sort: entry k of A gets the key (column << rbits) |
// row, unique because a row's columns are distinct, so sorting the keys
// lists A^T's rows in order with each one's entries by ascending A row --
// the reference's scatter order -- whatever the column lengths (the sort is
// linear in nnz).  The value carried along is k, the entry's position in A.
__global__ void k_transpose_keys(int32_t m, int32_t nnz, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci, int rbits, uint64_t* __restrict__ keys,
                                 int32_t* __restrict__ pos) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = find_row(rp, m, (int32_t)k);
        keys[k] = ((uint64_t)(uint32_t)ci[k] << rbits) | (uint64_t)(uint32_t)r;
        pos[k] = (int32_t)k;
    }
}

// Sorted keys -> A^T column indices (the A rows).
__global__ void k_transpose_rows(int32_t nnz, int rbits, const uint64_t* __restrict__ keys,
                                 int32_t* __restrict__ tidx) {
    const uint64_t mask = ((uint64_t)1 << rbits) - 1;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * blockDim.x)
        tidx[p] = (int32_t)(keys[p] & mask);
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

template <typename T>
__global__ void k_fill_value(int64_t n, T v, T* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = v;
}

int ensure_partials(sov_hip_csr* csr) {
    if (csr->pairs == nullptr)
        SOV_HIP_TRY(hipMalloc(&csr->pairs, sizeof(double) * 2 * ((size_t)csr->m + (size_t)csr->n + 1)));
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
// device order and parts
// ---------------------------------------------------------------------------

static int row_class(int32_t len) {
    int c = 0;
    while (c < kClasses - 1 && len > kClassMax[c]) ++c;
    return c;
}

// Device-order CSR of one orientation: row p is natural row perm[p], its
// column indices mapped to device positions through other_inv.
__global__ void k_device_fill(int32_t rows, const int32_t* __restrict__ perm, const int32_t* __restrict__ dptr,
                              const int32_t* __restrict__ nptr, const int32_t* __restrict__ nidx,
                              const int32_t* __restrict__ nsrc, const int32_t* __restrict__ other_inv,
                              int32_t* __restrict__ didx, int32_t* __restrict__ dsrc) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < rows; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = perm[p];
        int32_t d = dptr[p];
        for (int32_t q = nptr[r]; q < nptr[r + 1]; ++q, ++d) {
            didx[d] = other_inv[nidx[q]];
            dsrc[d] = nsrc ? nsrc[q] : q;
        }
    }
}

// Host: the stable class order of an orientation's rows from its natural
// row pointers; uploads perm / inv / device row pointers.
static int build_order_host(const std::vector<int32_t>& nptr, Orient& o) {
    const int32_t rows = o.nat.rows;
    std::vector<int32_t> perm(static_cast<size_t>(rows)), inv(static_cast<size_t>(rows));
    std::vector<int32_t> dptr(static_cast<size_t>(rows) + 1, 0);
    int32_t cnt[kClasses] = {};
    for (int32_t r = 0; r < rows; ++r) ++cnt[row_class(nptr[r + 1] - nptr[r])];
    int32_t cur[kClasses], off = 0;
    for (int c = 0; c < kClasses; ++c) {
        o.class_count[c] = cnt[c];
        o.class_offset[c] = off;
        cur[c] = off;
        off += cnt[c];
    }
    for (int32_t r = 0; r < rows; ++r) {
        const int32_t p = cur[row_class(nptr[r + 1] - nptr[r])]++;
        perm[p] = r;
        inv[r] = p;
    }
    for (int32_t p = 0; p < rows; ++p) dptr[p + 1] = dptr[p] + (nptr[perm[p] + 1] - nptr[perm[p]]);
    const int64_t r1 = std::max<int32_t>(rows, 1);
    SOV_HIP_TRY(hipMalloc(&o.perm, sizeof(int32_t) * r1));
    SOV_HIP_TRY(hipMalloc(&o.inv, sizeof(int32_t) * r1));
    SOV_HIP_TRY(hipMalloc(&o.d_ptr, sizeof(int32_t) * (rows + 1)));
    SOV_HIP_TRY(hipMalloc(&o.d_idx, sizeof(int32_t) * std::max<int32_t>(o.nat.nnz, 1)));
    SOV_HIP_TRY(hipMalloc(&o.d_src, sizeof(int32_t) * std::max<int32_t>(o.nat.nnz, 1)));
    if (rows > 0) {
        SOV_HIP_TRY(hipMemcpy(o.perm, perm.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
        SOV_HIP_TRY(hipMemcpy(o.inv, inv.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
    }
    SOV_HIP_TRY(hipMemcpy(o.d_ptr, dptr.data(), sizeof(int32_t) * (rows + 1), hipMemcpyHostToDevice));
    return SOV_OK;
}

// Device order of both orientations and their device-order CSRs.
static int build_order(sov_hip_ctx* ctx, sov_hip_csr* h) {
    std::vector<int32_t> aptr(static_cast<size_t>(h->m) + 1), tptr(static_cast<size_t>(h->n) + 1);
    SOV_HIP_TRY(hipMemcpyAsync(aptr.data(), h->a.nat.ptr, sizeof(int32_t) * aptr.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    SOV_HIP_TRY(hipMemcpyAsync(tptr.data(), h->t.nat.ptr, sizeof(int32_t) * tptr.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    int rc = build_order_host(aptr, h->a);
    if (rc == SOV_OK) rc = build_order_host(tptr, h->t);
    if (rc != SOV_OK) return rc;
    if (h->nnz > 0) {
        if (h->m > 0)
            k_device_fill<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(
                h->m, h->a.perm, h->a.d_ptr, h->a.nat.ptr, h->a.nat.idx, nullptr, h->t.inv, h->a.d_idx, h->a.d_src);
        if (h->n > 0)
            k_device_fill<<<launch_blocks(h->n), kBlock, 0, ctx->stream>>>(
                h->n, h->t.perm, h->t.d_ptr, h->t.nat.ptr, h->t.nat.idx, h->t.nat.src, h->a.inv, h->t.d_idx,
                h->t.d_src);
        SOV_HIP_TRY(check_launch());
    }
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

// lens[k * (rows + 1) + r + 1] = entries of device row r in column segment k
// (device column / W == k); lens[k * (rows + 1)] = 0.
__global__ void k_seg_lens(int32_t rows, int32_t K, int32_t W, const int32_t* __restrict__ ptr,
                           const int32_t* __restrict__ idx, int32_t* __restrict__ lens) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = ptr[r], e = ptr[r + 1];
        for (int32_t k = 0; k < K; ++k) {
            int32_t c = 0;
            for (int32_t q = b; q < e; ++q) c += (idx[q] / W == k) ? 1 : 0;
            lens[(int64_t)k * (rows + 1) + r + 1] = c;
            if (r == 0) lens[(int64_t)k * (rows + 1)] = 0;
        }
    }
}

// Copy the entries of segment k of every device row, in row order.
__global__ void k_seg_fill(int32_t rows, int32_t k, int32_t W, const int32_t* __restrict__ ptr,
                           const int32_t* __restrict__ idx, const int32_t* __restrict__ src,
                           const int32_t* __restrict__ pptr, int32_t* __restrict__ pidx,
                           int32_t* __restrict__ psrc) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        int32_t d = pptr[r];
        for (int32_t q = ptr[r]; q < ptr[r + 1]; ++q) {
            if (idx[q] / W != k) continue;
            pidx[d] = idx[q];
            psrc[d] = src[q];
            ++d;
        }
    }
}

__global__ void k_gather_f64_to_f32(int32_t nnz, const int32_t* __restrict__ src, const double* __restrict__ v,
                                    float* __restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x)
        out[p] = (float)v[src[p]];
}

// (Re)build the K column-segment parts of an orientation from its
// device-order CSR and A's current FP64 values.
int build_parts(sov_hip_ctx* ctx, sov_hip_csr* h, Orient& o, int K, bool want_f32) {
    o.free_parts();
    ++o.parts_version;
    const int32_t rows = o.nat.rows, cols = o.nat.cols, nnz = o.nat.nnz;
    if (K < 1 || cols < 2 * K || nnz == 0) K = 1;
    const int32_t W = (K == 1) ? std::max<int32_t>(cols, 1) : (cols + K - 1) / K;
    const int64_t r1 = std::max<int32_t>(rows, 1), nz = std::max<int32_t>(nnz, 1);
    auto fail_free = [&](hipError_t e, const char* what) {
        o.free_parts();
        return hip_fail(e, what);
    };
    hipError_t e = hipMalloc(&o.p_ptr, sizeof(int32_t) * (int64_t)K * (rows + 1));
    if (e == hipSuccess) e = hipMalloc(&o.p_idx, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&o.p_src, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&o.p_val64, sizeof(double) * nz);
    if (e == hipSuccess && want_f32) e = hipMalloc(&o.p_val32, sizeof(float) * nz);
    if (e == hipSuccess) e = hipMalloc(&o.acc64, sizeof(double) * r1);
    if (e == hipSuccess) e = hipMalloc(&o.acc32, sizeof(float) * r1);
    if (e == hipSuccess) e = hipMalloc(&o.chk_acc, sizeof(double) * 2 * r1);
    if (e != hipSuccess) return fail_free(e, "part buffers");
    std::vector<int64_t> base(static_cast<size_t>(K) + 1, 0);
    if (rows > 0) {
        k_seg_lens<<<launch_blocks(rows), kBlock, 0, ctx->stream>>>(rows, K, W, o.d_ptr, o.d_idx, o.p_ptr);
        e = check_launch();
        if (e != hipSuccess) return fail_free(e, "part lengths");
        for (int k = 0; k < K; ++k) {
            const int rc = inclusive_scan(ctx, o.p_ptr + (int64_t)k * (rows + 1), rows + 1);
            if (rc != SOV_OK) {
                o.free_parts();
                return rc;
            }
            int32_t nk = 0;
            e = hipMemcpy(&nk, o.p_ptr + (int64_t)k * (rows + 1) + rows, sizeof(int32_t), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return fail_free(e, "part sizes");
            base[static_cast<size_t>(k) + 1] = base[static_cast<size_t>(k)] + nk;
        }
    } else {
        e = hipMemsetAsync(o.p_ptr, 0, sizeof(int32_t) * K, ctx->stream);
        if (e != hipSuccess) return fail_free(e, "part pointers");
    }
    if (base[static_cast<size_t>(K)] != nnz) return fail_free(hipErrorUnknown, "part fill: entry count mismatch");
    o.parts.assign(static_cast<size_t>(K), Part{});
    for (int k = 0; k < K; ++k) {
        Part& p = o.parts[static_cast<size_t>(k)];
        const int64_t b = base[static_cast<size_t>(k)];
        p.rows = rows;
        p.cols = cols;
        p.nnz = static_cast<int32_t>(base[static_cast<size_t>(k) + 1] - b);
        p.ptr = o.p_ptr + (int64_t)k * (rows + 1);
        p.idx = o.p_idx + b;
        p.src = o.p_src + b;
        p.val64 = o.p_val64 + b;
        p.val32 = o.p_val32 ? o.p_val32 + b : nullptr;
        for (int c = 0; c < kClasses; ++c) {
            p.class_count[c] = o.class_count[c];
            p.class_offset[c] = o.class_offset[c];
            p.lanes[c] = kDefaultLanes[c];
        }
        if (rows > 0 && p.nnz > 0)
            k_seg_fill<<<launch_blocks(rows), kBlock, 0, ctx->stream>>>(rows, k, W, o.d_ptr, o.d_idx, o.d_src, p.ptr,
                                                                        o.p_idx + b, o.p_src + b);
    }
    if (nnz > 0) {
        k_gather_vals<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, o.p_src, h->a.nat.val64, o.p_val64);
        if (o.p_val32)
            k_gather_f64_to_f32<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, o.p_src, h->a.nat.val64,
                                                                                o.p_val32);
    }
    e = check_launch();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail_free(e, "part fill");
    o.seg_width = K > 1 ? W : 0;
    return SOV_OK;
}

// dst[p] = src[perm[p]] (natural -> device) or dst[perm[p]] = src[p].
template <typename T>
__global__ void k_permute(int32_t n, const int32_t* __restrict__ perm, int to_natural, const T* __restrict__ src,
                          T* __restrict__ dst) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (to_natural) dst[perm[p]] = src[p];
        else dst[p] = src[perm[p]];
    }
}

static int permute(sov_hip_ctx* ctx, const Orient& o, int to_natural, int prec, const void* src, void* dst) {
    const int32_t n = o.nat.rows;
    if (n == 0) return SOV_OK;
    if (prec == SOV_PREC_F64)
        k_permute<double><<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, o.perm, to_natural,
                                                                        static_cast<const double*>(src),
                                                                        static_cast<double*>(dst));
    else
        k_permute<float><<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, o.perm, to_natural,
                                                                       static_cast<const float*>(src),
                                                                       static_cast<float*>(dst));
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

// Re-derive every value image of both orientations from A's FP64 values.
static int refresh_values(sov_hip_ctx* ctx, sov_hip_csr* h) {
    if (h->nnz == 0) return SOV_OK;
    const double* av = h->a.nat.val64;
    k_gather_vals<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, h->t_src, av, h->t_val64);
    for (Orient* o : {&h->a, &h->t}) {
        if (o->p_val64 != nullptr)
            k_gather_vals<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, o->p_src, av, o->p_val64);
        if (o->p_val32 != nullptr)
            k_gather_f64_to_f32<<<launch_blocks(h->nnz), kBlock, 0, ctx->stream>>>(h->nnz, o->p_src, av, o->p_val32);
    }
    SOV_HIP_TRY(check_launch());
    for (Orient* o : {&h->a, &h->t}) {
        const int rc = refresh_tile_values(ctx, h, *o);
        if (rc != SOV_OK) return rc;
    }
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
        const int row = row_of<G>(v, lane);
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
    if (tiles_usable(ctx, o.tiles_of(prec == SOV_PREC_F64), prec == SOV_PREC_F64, x))
        return spmv_tiles(ctx, o, prec, alpha, x, beta, y);
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

// SpMV on natural-order vectors: stage x (and y when beta != 0) in device
// order, multiply, scatter y back.
static int spmv_natural(sov_hip_ctx* ctx, sov_hip_csr* h, int transpose, int prec, double alpha, const void* x,
                        double beta, void* y) {
    const Orient& o = transpose ? h->t : h->a;       // rows = outputs
    const Orient& g = transpose ? h->a : h->t;       // rows = the gathered dimension
    int rc = SOV_OK;
    if (g.nat.rows > 0) rc = permute(ctx, g, 0, prec, x, h->vx);
    if (rc == SOV_OK && beta != 0.0 && o.nat.rows > 0) rc = permute(ctx, o, 0, prec, y, h->vy);
    if (rc == SOV_OK) rc = spmv_orient(ctx, o, prec, alpha, h->vx, beta, h->vy);
    if (rc == SOV_OK && o.nat.rows > 0) rc = permute(ctx, o, 1, prec, h->vy, y);
    return rc;
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

// A^T's column indices (tidx, the A rows) and each entry's position in A
// (tsrc), in A^T row order with ascending A rows inside a row: a stable
// radix sort of the (column << rbits | row) keys over the bits they use.
// Temporary device memory: 2 x 12 bytes per entry plus the sort's storage.
static int sort_transpose(sov_hip_ctx* ctx, int32_t m, int32_t n, int32_t nnz, const int32_t* d_rp,
                          const int32_t* d_ci, int32_t* tidx, int32_t* tsrc) {
    int rbits = 1, cbits = 1;
    while (rbits < 31 && (1LL << rbits) < m) ++rbits;
    while (cbits < 31 && (1LL << cbits) < n) ++cbits;
    uint64_t *keys_in = nullptr, *keys_out = nullptr;
    int32_t* pos_in = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t e = hipMalloc(&keys_in, sizeof(uint64_t) * nnz);
    if (e == hipSuccess) e = hipMalloc(&keys_out, sizeof(uint64_t) * nnz);
    if (e == hipSuccess) e = hipMalloc(&pos_in, sizeof(int32_t) * nnz);
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in, keys_out, pos_in, tsrc, (size_t)nnz, 0,
                                      (unsigned)(rbits + cbits), ctx->stream);
    if (e == hipSuccess) e = hipMalloc(&temp, std::max<size_t>(temp_bytes, 16));
    if (e == hipSuccess) {
        k_transpose_keys<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(m, nnz, d_rp, d_ci, rbits, keys_in, pos_in);
        e = check_launch();
    }
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, pos_in, tsrc, (size_t)nnz, 0,
                                      (unsigned)(rbits + cbits), ctx->stream);
    if (e == hipSuccess) {
        k_transpose_rows<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, rbits, keys_out, tidx);
        e = check_launch();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(keys_in);
    (void)hipFree(keys_out);
    (void)hipFree(pos_in);
    (void)hipFree(temp);
    return e == hipSuccess ? SOV_OK : hip_fail(e, "csr transpose sort");
}

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
    h->stream = ctx->stream;
    auto cleanup = [&](int rc) {
        sov_hip_csr_destroy(h);
        return rc;
    };
    const int64_t nz = std::max<int32_t>(nnz, 1);
    Natural& A = h->a.nat;
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
    // natural A^T, built on the device: row pointers from a column count and
    // a scan, entries from one radix sort of (column, row) keys
    hipError_t e = hipMalloc(&h->t_ptr, sizeof(int32_t) * ((int64_t)n + 1));
    if (e == hipSuccess) e = hipMalloc(&h->t_idx, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&h->t_src, sizeof(int32_t) * nz);
    if (e == hipSuccess) e = hipMalloc(&h->t_val64, sizeof(double) * nz);
    if (e != hipSuccess) return cleanup(hip_fail(e, "csr transpose buffers"));
    k_zero_i32<<<launch_blocks((int64_t)n + 1), kBlock, 0, ctx->stream>>>((int64_t)n + 1, h->t_ptr);
    if (nnz > 0) k_count_cols<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, d_ci, h->t_ptr);
    int rc = inclusive_scan(ctx, h->t_ptr, (int64_t)n + 1);
    if (rc == SOV_OK && nnz > 0) rc = sort_transpose(ctx, m, n, nnz, d_rp, d_ci, h->t_idx, h->t_src);
    if (rc == SOV_OK && nnz > 0) {
        k_gather_vals<<<launch_blocks(nnz), kBlock, 0, ctx->stream>>>(nnz, h->t_src, A.val64, h->t_val64);
        e = check_launch();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "csr transpose");
    }
    if (rc != SOV_OK) return cleanup(rc);
    Natural& T = h->t.nat;
    T.rows = n;
    T.cols = m;
    T.nnz = nnz;
    T.ptr = h->t_ptr;
    T.idx = h->t_idx;
    T.val64 = h->t_val64;
    T.src = h->t_src;
    {
        const int64_t len = std::max<int64_t>(std::max<int64_t>(m, n), 1);
        hipError_t e2 = hipMalloc(&h->vx, sizeof(double) * len);
        if (e2 == hipSuccess) e2 = hipMalloc(&h->vy, sizeof(double) * len);
        if (e2 != hipSuccess) return cleanup(hip_fail(e2, "csr staging"));
    }
    rc = build_order(ctx, h);
    if (rc == SOV_OK) rc = build_parts(ctx, h, h->a, 1, false);
    if (rc == SOV_OK) rc = build_parts(ctx, h, h->t, 1, false);
    if (rc == SOV_OK) rc = ensure_partials(h);
    if (rc != SOV_OK) return cleanup(rc);
    *out = h;
    return SOV_OK;
}

void sov_hip_csr_destroy(sov_hip_csr* h) {
    if (h == nullptr) return;
    DeviceGuard2 g(h->device);
    // Only the creating context's stream can still be using these buffers.
    if (h->stream != nullptr) (void)hipStreamSynchronize(h->stream);
    for (Orient* o : {&h->a, &h->t}) {
        o->tiles.free();
        o->tiles64.free();
    }
    h->a.free_parts();
    h->t.free_parts();
    h->a.free_order();
    h->t.free_order();
    (void)hipFree(h->vx);
    (void)hipFree(h->vy);
    if (h->owns_a_val64) (void)hipFree(h->a.nat.val64);
    (void)hipFree(h->t_ptr);
    (void)hipFree(h->t_idx);
    (void)hipFree(h->t_src);
    (void)hipFree(h->t_val64);
    (void)hipFree(h->partials);
    (void)hipFree(h->pairs);
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
    for (Orient* o : {&h->a, &h->t}) {
        if (o->p_val32 != nullptr) continue;
        SOV_HIP_TRY(hipMalloc(&o->p_val32, sizeof(float) * nz));
        int64_t off = 0;
        for (Part& p : o->parts) {
            p.val32 = o->p_val32 + off;
            off += p.nnz;
        }
    }
    h->f32 = true;
    return refresh_values(ctx, h);
}

int sov_hip_csr_scale(sov_hip_ctx* ctx, sov_hip_csr* h, const double* rs, const double* cs) {
    if (ctx == nullptr || h == nullptr || rs == nullptr || cs == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "csr_scale args");
    if (!h->owns_a_val64) return fail(SOV_INVALID_ARGUMENT, "csr_scale needs SOV_CSR_OWN_VALUES");
    DeviceGuard2 g(ctx->device);
    if (h->nnz == 0) return SOV_OK;
    k_scale_vals<<<launch_blocks(h->m), kBlock, 0, ctx->stream>>>(h->m, h->a.nat.ptr, h->a.nat.idx,
                                                                   h->a.nat.val64, rs, cs);
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
        SOV_HIP_TRY(hipMemcpyAsync(d_tval, h->t_val64, sizeof(double) * h->nnz, hipMemcpyDeviceToDevice,
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
    for (Part& p : o.parts)
        for (int c = 0; c < kClasses; ++c) p.lanes[c] = lanes[c] == 0 ? kDefaultLanes[c] : lanes[c];
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
    const int rc = build_parts(ctx, h, o, count, h->f32);
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
    if (o.nat.rows > 0 && y == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv: y is null");
    if (h->nnz > 0 && x == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv: x is null");
    if (prec == SOV_PREC_F32 && !h->f32 && h->nnz > 0)
        return fail(SOV_INVALID_ARGUMENT, "FP32 SpMV needs sov_hip_csr_make_f32 first");
    DeviceGuard2 g(ctx->device);
    return spmv_natural(ctx, const_cast<sov_hip_csr*>(h), transpose, prec, alpha, x, beta, y);
}

int sov_hip_spmv_device(sov_hip_ctx* ctx, const sov_hip_csr* h, int transpose, int prec, double alpha,
                        const void* x, double beta, void* y) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv_device args");
    if (prec != SOV_PREC_F64 && prec != SOV_PREC_F32) return fail(SOV_INVALID_ARGUMENT, "spmv_device: precision");
    const Orient& o = transpose ? h->t : h->a;
    if (o.nat.rows > 0 && y == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv_device: y is null");
    if (h->nnz > 0 && x == nullptr) return fail(SOV_INVALID_ARGUMENT, "spmv_device: x is null");
    DeviceGuard2 g(ctx->device);
    return spmv_orient(ctx, o, prec, alpha, x, beta, y);
}

int sov_hip_csr_order(sov_hip_ctx* ctx, const sov_hip_csr* h, int which, int32_t* perm) {
    if (ctx == nullptr || h == nullptr || perm == nullptr || (which != 0 && which != 1))
        return fail(SOV_INVALID_ARGUMENT, "csr_order args");
    DeviceGuard2 g(ctx->device);
    const Orient& o = which == 0 ? h->a : h->t;
    if (o.nat.rows == 0) return SOV_OK;
    SOV_HIP_TRY(hipMemcpyAsync(perm, o.perm, sizeof(int32_t) * o.nat.rows, hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_csr_permute(sov_hip_ctx* ctx, const sov_hip_csr* h, int which, int to_natural, int prec,
                        const void* src, void* dst) {
    if (ctx == nullptr || h == nullptr || (which != 0 && which != 1) || (prec != SOV_PREC_F64 && prec != SOV_PREC_F32))
        return fail(SOV_INVALID_ARGUMENT, "csr_permute args");
    const Orient& o = which == 0 ? h->a : h->t;
    if (o.nat.rows > 0 && (src == nullptr || dst == nullptr || src == dst))
        return fail(SOV_INVALID_ARGUMENT, "csr_permute: null or aliased vectors");
    DeviceGuard2 g(ctx->device);
    return permute(ctx, o, to_natural, prec, src, dst);
}

int sov_hip_csr_tune(sov_hip_ctx* ctx, sov_hip_csr* h, int prec) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "tune args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "tune during capture");
    if (prec == SOV_PREC_F32 && !h->f32) return fail(SOV_INVALID_ARGUMENT, "tune f32 needs make_f32");
    DeviceGuard2 g(ctx->device);
    const int64_t len = std::max<int64_t>(std::max(h->m, h->n), 1);
    void *xv = nullptr, *yv = nullptr;   // room for either precision
    SOV_HIP_TRY(hipMalloc(&xv, sizeof(double) * len));
    hipError_t e = hipMalloc(&yv, sizeof(double) * len);
    if (e != hipSuccess) { (void)hipFree(xv); return hip_fail(e, "tune buffers"); }
    int rc = SOV_OK;
    // x = 1 in the element type of precision q
    auto fill = [&](int q) {
        if (q == SOV_PREC_F64)
            k_fill_value<double><<<launch_blocks(len), kBlock, 0, ctx->stream>>>(len, 1.0, static_cast<double*>(xv));
        else
            k_fill_value<float><<<launch_blocks(len), kBlock, 0, ctx->stream>>>(len, 1.0f, static_cast<float*>(xv));
        const hipError_t fe = check_launch();
        return fe == hipSuccess ? SOV_OK : hip_fail(fe, "tune fill");
    };
    rc = fill(prec);
    for (int orient = 0; orient < 2 && rc == SOV_OK; ++orient) {
        Orient& o = orient == 0 ? h->a : h->t;
        o.tiles.parts = o.tiles64.parts64 = 0;   // the CSR timings below must not run on the tiles
        // segment counts worth trying: keep each gathered slice >= 64k elements
        const int64_t cols = o.nat.cols;
        int best_K = 1;
        float best_ms = 1e30f;
        std::vector<std::vector<int32_t>> best_lanes;
        for (int K : {1, 2, 4, 8}) {
            if (K > 1 && cols / K < 65536) break;
            rc = build_parts(ctx, h, o, K, h->f32);
            for (int k = 0; k < o.nparts() && rc == SOV_OK; ++k) rc = tune_part_lanes(ctx, o, o.part(k), prec, xv, yv);
            if (rc != SOV_OK) break;
            const float ms = time_spmv(ctx, o, prec, xv, yv, 4);
            if (ms < best_ms) {
                best_ms = ms;
                best_K = K;
                best_lanes.clear();
                for (const Part& p : o.parts) best_lanes.emplace_back(p.lanes, p.lanes + kClasses);
            }
        }
        if (rc == SOV_OK && best_K != o.nparts()) {
            rc = build_parts(ctx, h, o, best_K, h->f32);
            if (rc == SOV_OK && static_cast<int>(best_lanes.size()) == o.nparts())
                for (int k = 0; k < o.nparts(); ++k)
                    std::copy(best_lanes[static_cast<size_t>(k)].begin(), best_lanes[static_cast<size_t>(k)].end(),
                              o.part(k).lanes);
        }
        // the LDS-staged tiles, for products large enough to amortise their
        // per-group staging: timed in both precisions (FP32 iterates end in
        // an FP64 polish), each on its own layout, each kept where its SpMV
        // is within kTileTuneMargin of the CSR parts' (the fused steps add a
        // per-row epilogue, which the tiles run on coalesced 64-row chunks
        // from LDS and the CSR parts per row: on the 1M-column Zipf LP in
        // FP64 the A^T step takes 145 us on the tiles and 194 us on the CSR
        // parts, yet an SpMV-only comparison kept the CSR parts)
        if (rc == SOV_OK && o.nat.nnz >= kTileTuneMinNnz && o.nat.rows >= 64) {
            int choice[2] = {0, 0};   // by SOV_PREC_*
            const int other = prec == SOV_PREC_F64 ? static_cast<int>(SOV_PREC_F32) : static_cast<int>(SOV_PREC_F64);
            for (int q : {prec, other}) {
                if (rc != SOV_OK || (q == SOV_PREC_F32 && !h->f32)) continue;
                const bool f64 = q == SOV_PREC_F64;
                Tiles& t = o.tiles_of(f64);
                if (!t.built()) rc = build_tiles(ctx, h, o, f64);
                if (rc == SOV_OK && q != prec) rc = fill(q);
                if (rc != SOV_OK) break;
                (f64 ? t.parts64 : t.parts) = 0;
                float bms = (q == prec ? best_ms : time_spmv(ctx, o, q, xv, yv, 4)) * kTileTuneMargin;
                for (int P : {1, 2, 4}) {
                    if (P > std::max(t.S, 1)) break;
                    (f64 ? t.parts64 : t.parts) = P;
                    if (!tiles_usable(ctx, t, f64, xv)) break;
                    const float ms = time_spmv(ctx, o, q, xv, yv, 4);
                    if (ms < bms) {
                        bms = ms;
                        choice[q] = P;
                    }
                }
            }
            if (rc == SOV_OK) {
                o.tiles.parts = choice[SOV_PREC_F32];
                o.tiles64.parts64 = choice[SOV_PREC_F64];
                if (o.tiles.parts == 0) o.tiles.free();
                if (o.tiles64.parts64 == 0) o.tiles64.free();
            }
            if (rc == SOV_OK) rc = fill(prec);   // the next orientation times precision prec again
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
    const Natural& A = h->a.nat;
    const Natural& T = h->t.nat;
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
    const Natural& A = h->a.nat;
    const Natural& T = h->t.nat;
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
