// Tiled layout (struct Tiles) of an orientation: host build, value images,
// the tiled SpMV, and the C-ABI that switches an orientation's products to
// the tiles.
//
// Why: a row product gathers one vector element per entry.  Gathered from
// L2, each lane is its own L2 request, and on the 1M-column Zipf LP the CSR
// kernels run at about two thirds of the random-gather rate the GPU can
// sustain, far below HBM bandwidth.  Staging a slice of the vector into LDS
// with 16-byte coalesced loads moves 16 (FP32) values per L1 request
// instead of one; the entries then gather from LDS.  The price is that every
// row group reads the whole vector once per product (L2-resident: the
// groups walk the slices in the same order, so the hot slices are shared).
//
// Reference counterpart: the same products as the CSR parts
// (src/l1/pdhg.cpp:295-331 via Backend::spmv / spmv_transpose,
// include/sovereign/backend.hpp:124-129).
#include <algorithm>
#include <thread>
#include <vector>

#include "sov_hip_tiles.hpp"

namespace sovhip {

struct DeviceGuardT {
    int prev = -1;
    explicit DeviceGuardT(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuardT() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

void Tiles::free() {
    (void)hipFree(keys);
    (void)hipFree(src);
    (void)hipFree(pidx);
    (void)hipFree(pdesc);
    (void)hipFree(units);
    (void)hipFree(val32);
    (void)hipFree(val64);
    (void)hipFree(part);
    (void)hipFree(cnt);
    part = nullptr;
    cnt = nullptr;
    h_cost.clear();
    Q = 1;
    keys = nullptr;
    src = nullptr;
    pidx = nullptr;
    pdesc = nullptr;
    passes = 0;
    units = nullptr;
    val32 = nullptr;
    val64 = nullptr;
    h_units.clear();
    nb = S = rmax = U = cols = 0;
    E = 0;
    parts = 0;
    parts64 = 0;
}

// Launch k of `parts` covers the units whose first slice lies in
// [k S / parts, (k + 1) S / parts), split into t.Q halves of contiguous
// units with about equal estimated time (h_cost): the split that makes the
// slower half fastest.
TileArgs tile_args(const Tiles& t, int32_t rows, int32_t cols, bool f64, int k, int parts) {
    TileArgs a{};
    a.keys = t.keys;
    a.vals = f64 ? static_cast<const void*>(t.val64) : static_cast<const void*>(t.val32);
    a.pidx = t.pidx;
    a.pdesc = t.pdesc;
    a.sent = static_cast<int32_t>(t.E);
    a.units = t.units;
    a.part = t.part;
    a.cnt = t.cnt;
    a.nb = t.nb;
    a.U = t.U;
    a.W = t.W;
    a.rmax = t.rmax;
    a.rows = rows;
    a.cols = cols;
    a.Q = t.Q;
    const int64_t lo = (int64_t)k * t.S / parts, hi = (int64_t)(k + 1) * t.S / parts;
    int u0 = 0;
    while (u0 < t.U && t.h_units[4 * static_cast<size_t>(u0)] < lo) ++u0;
    int u1 = u0;
    while (u1 < t.U && t.h_units[4 * static_cast<size_t>(u1)] < hi) ++u1;
    a.hu[0] = u0;
    a.hu[t.Q] = u1;
    if (t.Q == 2) {
        double total = 0.0;
        for (int u = u0; u < u1; ++u) total += t.h_cost[static_cast<size_t>(u)];
        double left = 0.0, best = total;
        int cut = u0;
        for (int c = u0; c <= u1; ++c) {
            const double worse = std::max(left, total - left);
            if (worse < best) {
                best = worse;
                cut = c;
            }
            if (c < u1) left += t.h_cost[static_cast<size_t>(c)];
        }
        a.hu[1] = cut;
    }
    for (int h = 0; h < t.Q; ++h) {
        int v = a.hu[h];
        while (v < a.hu[h + 1] && !t.h_units[4 * static_cast<size_t>(v) + 1]) ++v;
        a.hs[h] = v;
    }
    return a;
}

bool tiles_usable(const sov_hip_ctx* ctx, const Tiles& t, bool f64, const void* x) {
    if ((f64 ? t.parts64 : t.parts) <= 0 || t.nb <= 0 || !t.built()) return false;
    if (f64 ? (t.val64 == nullptr && t.E > 0) : (t.val32 == nullptr && t.E > 0)) return false;
    if ((reinterpret_cast<uintptr_t>(x) & 15u) != 0) return false;
    // the staging waves load the gathered vector through a buffer resource,
    // whose byte offsets are 32-bit
    if ((int64_t)t.cols * (f64 ? 8 : 4) > INT32_MAX) return false;
    return (int64_t)tile_lds_bytes(t.W, t.rmax, f64 ? 8 : 4) <= ctx->lds_max;
}

hipError_t tile_lds_attr(const sov_hip_ctx* ctx, const void* kernel, int tsize) {
    (void)tsize;   // every layout fits kTileLdsBytes
    const int64_t want = std::min<int64_t>(ctx->lds_max, (int64_t)kTileLdsBytes);
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(want));
}

namespace {

// One row group's entries, in storage order.
struct GroupOut {
    std::vector<uint32_t> keys;
    std::vector<int32_t> src;
    std::vector<int32_t> wrel;   // U * kTileWaves starts, relative to the group
};

void build_group(int b, const Tiles& t, int lanew, int32_t rows, const std::vector<int32_t>& ptr,
                 const std::vector<int32_t>& idx, const std::vector<int32_t>& srcv,
                 const std::vector<int32_t>& unit_of_slice, GroupOut& g) {
    const int U = t.U, W = t.W;
    std::vector<int64_t> cnt(static_cast<size_t>(U) + 1, 0);
    int nloc = 0;
    for (int l = 0; l < t.rmax; ++l) {
        const int p = tile_row(b, t.nb, l);
        if (p >= rows) break;
        nloc = l + 1;
        for (int32_t q = ptr[p]; q < ptr[p + 1]; ++q)
            ++cnt[static_cast<size_t>(unit_of_slice[static_cast<size_t>(idx[q] / W)]) + 1];
    }
    for (int u = 0; u < U; ++u) cnt[static_cast<size_t>(u) + 1] += cnt[static_cast<size_t>(u)];
    // bucket the entries by unit, local rows ascending
    std::vector<uint32_t> bk(static_cast<size_t>(cnt[static_cast<size_t>(U)]));
    std::vector<int32_t> bs(bk.size());
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int l = 0; l < nloc; ++l) {
        const int p = tile_row(b, t.nb, l);
        if (p >= rows) continue;
        for (int32_t q = ptr[p]; q < ptr[p + 1]; ++q) {
            const int u = unit_of_slice[static_cast<size_t>(idx[q] / W)];
            const int64_t base = static_cast<int64_t>(t.h_units[4 * static_cast<size_t>(u)]) * W;
            const int64_t at = fill[static_cast<size_t>(u)]++;
            bk[static_cast<size_t>(at)] = (static_cast<uint32_t>(l) << kTileColBits) |
                                          static_cast<uint32_t>(idx[q] - base);
            bs[static_cast<size_t>(at)] = srcv[q];
        }
    }
    // split each tile at row boundaries into kTileWaves sub-ranges, pad each
    // to a multiple of lanew (the entries per lane per pass)
    g.wrel.assign(static_cast<size_t>(U) * kTileWaves, 0);
    g.keys.clear();
    g.src.clear();
    g.keys.reserve(bk.size() + static_cast<size_t>(U) * kTileWaves * (lanew - 1));
    g.src.reserve(g.keys.capacity());
    for (int u = 0; u < U; ++u) {
        const int64_t b0 = cnt[static_cast<size_t>(u)], n = cnt[static_cast<size_t>(u) + 1] - b0;
        int64_t cut[kTileWaves + 1];
        cut[0] = 0;
        cut[kTileWaves] = n;
        for (int w = 1; w < kTileWaves; ++w) {
            int64_t at = std::max<int64_t>(n * w / kTileWaves, cut[w - 1]);
            while (at > 0 && at < n &&
                   (bk[static_cast<size_t>(b0 + at)] >> kTileColBits) ==
                       (bk[static_cast<size_t>(b0 + at - 1)] >> kTileColBits))
                ++at;
            cut[w] = at;
        }
        for (int w = 0; w < kTileWaves; ++w) {
            g.wrel[static_cast<size_t>(u) * kTileWaves + w] = static_cast<int32_t>(g.keys.size());
            for (int64_t i = cut[w]; i < cut[w + 1]; ++i) {
                g.keys.push_back(bk[static_cast<size_t>(b0 + i)]);
                g.src.push_back(bs[static_cast<size_t>(b0 + i)]);
            }
            while (g.keys.size() % lanew != 0) {
                g.keys.push_back(kTileSentinel);
                g.src.push_back(-1);
            }
        }
    }
}

template <typename T>
__global__ void k_tile_vals(int64_t E, const int32_t* __restrict__ src, const double* __restrict__ av,
                            T* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < E; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = src[i];
        out[i] = s >= 0 ? static_cast<T>(av[s]) : T(0);
    }
}

template <typename T>
int gather_tile_values(sov_hip_ctx* ctx, const sov_hip_csr* h, const Tiles& t, T* out) {
    const int64_t n = t.E + kTileLaneMax;   // the arrays' full length (padding and the sentinel block read 0)
    k_tile_vals<T><<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, t.src, h->a.nat.val64, out);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

}  // namespace

// The pass list of a layout (what the kernels read instead of wstart): for
// each (group, wave), unit by unit, the passes of its sub-range, each up to
// kWave * lanew entries.  Host only.
static int pass_tables(const Tiles& t, const std::vector<int32_t>& wstart, int lanew, bool pairs,
                       std::vector<int32_t>& pidx, std::vector<int32_t>& pdesc) {
    const int U = t.U;
    const int kPass = kWave * lanew;
    pidx.assign(static_cast<size_t>(t.nb) * kTileWaves * (U + 1), 0);
    pdesc.clear();
    for (int b = 0; b < t.nb; ++b)
        for (int w = 0; w < kTileWaves; ++w) {
            const size_t base = (static_cast<size_t>(b) * kTileWaves + w) * (U + 1);
            for (int u = 0; u < U; ++u) {
                pidx[base + u] = static_cast<int32_t>(pdesc.size() / 2);
                const size_t at = (static_cast<size_t>(b) * U + u) * kTileWaves + w;
                // passes of kPass entries, the last one short; then, for
                // kernels that add passes in pairs, an empty pass (its lanes
                // load only the sentinel block) when the count is odd.  So
                // only the last real pass of a sub-range is short: a pass
                // with a successor fills all 64 lanes, and its last row is
                // lane 63's, which the carry takes to the successor.
                const int64_t e0 = wstart[at], e1 = wstart[at + 1];
                int64_t np = 0;
                for (int64_t e = e0; e < e1; e += kPass, ++np) {
                    pdesc.push_back(static_cast<int32_t>(e));
                    pdesc.push_back(static_cast<int32_t>(std::min<int64_t>(kPass, e1 - e)));
                }
                if (pairs && (np & 1)) {
                    pdesc.push_back(static_cast<int32_t>(e1));
                    pdesc.push_back(0);
                }
            }
            if (pdesc.size() / 2 >= static_cast<size_t>(INT32_MAX))
                return fail(SOV_INVALID_ARGUMENT, "tiles: too many passes");
            pidx[base + U] = static_cast<int32_t>(pdesc.size() / 2);
        }
    return SOV_OK;
}

// Estimated time of a unit per group, in entry-equivalents: entries
// gathered from memory cost more than those gathered from LDS, and every
// unit pays its barrier, table reads and the fill and drain of the entry
// waves' passes.  Measured on the 1M-column Zipf LP's FP32 dual step (two
// gather units of about 8,000 entries per group, then staged units of
// 1,500 to 21,000): gather weight 1 / 2 / 4 / 8 with unit cost 1,000 gave
// 70.6 / 65.3 / 65.1 / 63.2 us; 4 with 3,000, 60.6-61.9 us; 2 with 3,500
// (a fit to per-unit timestamps of an instrumented copy), 64.9-65.2 us;
// the primal step (31 equal staged units) does not move.
constexpr double kGatherEntryCost = 4.0;
constexpr double kUnitCost = 3000.0;

// The tiled layout of a device-order CSR (rows x cols, nnz entries with
// source positions src) for products of element size tsize, for a device
// with `cus` compute units: fills the shape fields, unit table and unit
// costs of t and the host arrays.  Host only.
static int layout_host(int32_t rows, int32_t cols, const std::vector<int32_t>& ptr, const std::vector<int32_t>& idx,
                       const std::vector<int32_t>& srcv, int cus, int tsize, Tiles& t, std::vector<uint32_t>& keys,
                       std::vector<int32_t>& src, std::vector<int32_t>& wstart) {
    if (rows <= 0) return fail(SOV_INVALID_ARGUMENT, "tiles: no rows");
    const int64_t nchunks = (rows + 63) / 64;
    const int64_t cap = kTileRowCap / 64;
    t.Q = kTileHalves;
    int64_t nb = std::min<int64_t>(std::max(cus / t.Q, 1), nchunks);
    nb = std::max<int64_t>(nb, (nchunks + cap - 1) / cap);
    t.nb = static_cast<int32_t>(nb);
    t.rmax = static_cast<int32_t>((nchunks + nb - 1) / nb * 64);
    t.W = tile_slice_bytes(t.rmax, tsize) / tsize;
    if (t.W <= 0) return fail(SOV_INVALID_ARGUMENT, "tiles: the row sums leave no room for a slice");
    t.cols = cols;
    t.S = static_cast<int32_t>((static_cast<int64_t>(cols) + t.W - 1) / t.W);
    // units: a slice with at least kTileStageMin entries per group on
    // average is staged; runs of the others become gather units
    std::vector<int64_t> per_slice(static_cast<size_t>(t.S), 0);
    for (int32_t q = 0; q < ptr[rows]; ++q) ++per_slice[static_cast<size_t>(idx[q] / t.W)];
    std::vector<int32_t> unit_of_slice(static_cast<size_t>(t.S));
    t.h_units.clear();
    const int gmax = tile_gather_slices(t.W);
    for (int s = 0; s < t.S;) {
        const bool staged = per_slice[static_cast<size_t>(s)] >= static_cast<int64_t>(kTileStageMin) * nb;
        int e = s + 1;
        if (!staged)
            while (e < t.S && e - s < gmax &&
                   per_slice[static_cast<size_t>(e)] < static_cast<int64_t>(kTileStageMin) * nb)
                ++e;
        const int u = static_cast<int>(t.h_units.size() / 4);
        t.h_units.insert(t.h_units.end(), {s, staged ? 1 : 0, 0, 0});
        for (int k = s; k < e; ++k) unit_of_slice[static_cast<size_t>(k)] = u;
        s = e;
    }
    t.U = static_cast<int32_t>(t.h_units.size() / 4);
    int ordinal = 0, next = t.U;
    for (int u = 0; u < t.U; ++u)
        if (t.h_units[4 * static_cast<size_t>(u) + 1])
            t.h_units[4 * static_cast<size_t>(u) + 2] = (ordinal++) % kTileBuffers;
    for (int u = t.U - 1; u >= 0; --u) {
        t.h_units[4 * static_cast<size_t>(u) + 3] = next;
        if (t.h_units[4 * static_cast<size_t>(u) + 1]) next = u;
    }
    t.h_units.push_back(t.S);
    std::vector<GroupOut> groups(static_cast<size_t>(nb));
    const int nth = static_cast<int>(std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency())));
    {
        std::vector<std::thread> pool;
        for (int th = 0; th < nth; ++th)
            pool.emplace_back([&, th] {
                for (int64_t b = th; b < nb; b += nth)
                    build_group(static_cast<int>(b), t, tile_lane(tsize), rows, ptr, idx, srcv, unit_of_slice,
                                groups[static_cast<size_t>(b)]);
            });
        for (auto& th : pool) th.join();
    }
    int64_t E = 0;
    for (const GroupOut& g : groups) E += static_cast<int64_t>(g.keys.size());
    if (E + kTileLaneMax >= INT32_MAX) return fail(SOV_INVALID_ARGUMENT, "tiles: too many entries for 32-bit offsets");
    wstart.assign(static_cast<size_t>(nb) * t.U * kTileWaves + 1, 0);
    keys.assign(static_cast<size_t>(E + kTileLaneMax), kTileSentinel);   // entries, then the sentinel block
    src.assign(keys.size(), -1);
    int64_t off = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const GroupOut& g = groups[static_cast<size_t>(b)];
        for (size_t i = 0; i < g.wrel.size(); ++i)
            wstart[static_cast<size_t>(b) * t.U * kTileWaves + i] = static_cast<int32_t>(off + g.wrel[i]);
        std::copy(g.keys.begin(), g.keys.end(), keys.begin() + off);
        std::copy(g.src.begin(), g.src.end(), src.begin() + off);
        off += static_cast<int64_t>(g.keys.size());
    }
    wstart.back() = static_cast<int32_t>(E);
    t.E = E;
    t.h_cost.assign(static_cast<size_t>(t.U), 0.0);
    for (int u = 0; u < t.U; ++u) {
        int64_t eu = 0;
        for (int64_t b = 0; b < nb; ++b) {
            const size_t at = (static_cast<size_t>(b) * t.U + u) * kTileWaves;
            eu += wstart[at + kTileWaves] - wstart[at];
        }
        const double per_group = static_cast<double>(eu) / static_cast<double>(nb);
        t.h_cost[static_cast<size_t>(u)] =
            per_group * (t.h_units[4 * static_cast<size_t>(u) + 1] ? 1.0 : kGatherEntryCost) + kUnitCost;
    }
    return SOV_OK;
}

int build_tiles(sov_hip_ctx* ctx, const sov_hip_csr* h, Orient& o, bool f64) {
    Tiles& t = o.tiles_of(f64);
    const int parts = t.parts, parts64 = t.parts64;
    t.free();
    const int32_t rows = o.nat.rows, nnz = o.nat.nnz;
    if (rows == 0) return fail(SOV_INVALID_ARGUMENT, "tiles: no rows");
    // the device-order CSR on the host
    std::vector<int32_t> ptr(static_cast<size_t>(rows) + 1), idx(static_cast<size_t>(std::max(nnz, 1))),
        srcv(idx.size());
    SOV_HIP_TRY(hipMemcpyAsync(ptr.data(), o.d_ptr, sizeof(int32_t) * ptr.size(), hipMemcpyDeviceToHost, ctx->stream));
    if (nnz > 0) {
        SOV_HIP_TRY(hipMemcpyAsync(idx.data(), o.d_idx, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, ctx->stream));
        SOV_HIP_TRY(hipMemcpyAsync(srcv.data(), o.d_src, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, ctx->stream));
    }
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<uint32_t> keys;
    std::vector<int32_t> src, wstart;
    int rc = layout_host(rows, o.nat.cols, ptr, idx, srcv, ctx->cus, f64 ? 8 : 4, t, keys, src, wstart);
    std::vector<int32_t> pidx, pdesc;
    if (rc == SOV_OK) rc = pass_tables(t, wstart, tile_lane(f64 ? 8 : 4), tile_pairs(f64 ? 8 : 4), pidx, pdesc);
    if (rc != SOV_OK) {
        t.free();
        return rc;
    }
    t.passes = static_cast<int64_t>(pdesc.size() / 2);
    if (pdesc.empty()) pdesc.assign(2, 0);
    const int64_t E = t.E;
    const int64_t e1 = E + kTileLaneMax;   // entries and the sentinel block
    hipError_t e = hipMalloc(&t.keys, sizeof(uint32_t) * e1);
    if (e == hipSuccess) e = hipMalloc(&t.src, sizeof(int32_t) * e1);
    if (e == hipSuccess) e = hipMalloc(&t.pidx, sizeof(int32_t) * pidx.size());
    if (e == hipSuccess) e = hipMalloc(&t.pdesc, sizeof(int32_t) * pdesc.size());
    if (e == hipSuccess) e = hipMalloc(&t.units, sizeof(int32_t) * t.h_units.size());
    if (e == hipSuccess)
        e = hipMemcpy(t.units, t.h_units.data(), sizeof(int32_t) * t.h_units.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && t.Q > 1) e = hipMalloc(&t.part, (size_t)(f64 ? 8 : 4) * t.Q * t.nb * t.rmax);
    if (e == hipSuccess && t.Q > 1) e = hipMalloc(&t.cnt, sizeof(int32_t) * t.nb);
    if (e == hipSuccess && t.Q > 1) e = hipMemsetAsync(t.cnt, 0, sizeof(int32_t) * t.nb, ctx->stream);
    if (e == hipSuccess && f64) e = hipMalloc(&t.val64, sizeof(double) * e1);
    if (e == hipSuccess && !f64) e = hipMalloc(&t.val32, sizeof(float) * e1);
    if (e == hipSuccess) e = hipMemcpy(t.keys, keys.data(), sizeof(uint32_t) * e1, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.src, src.data(), sizeof(int32_t) * e1, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.pidx, pidx.data(), sizeof(int32_t) * pidx.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.pdesc, pdesc.data(), sizeof(int32_t) * pdesc.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        t.free();
        return hip_fail(e, "tiles upload");
    }
    t.parts = parts;
    t.parts64 = parts64;
    rc = refresh_tile_values(ctx, h, o);
    if (rc != SOV_OK) {
        t.free();
        return rc;
    }
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int refresh_tile_values(sov_hip_ctx* ctx, const sov_hip_csr* h, Orient& o) {
    int rc = SOV_OK;
    for (const Tiles* t : {&o.tiles, &o.tiles64}) {
        if (!t->built()) continue;
        if (rc == SOV_OK && t->val64) rc = gather_tile_values<double>(ctx, h, *t, t->val64);
        if (rc == SOV_OK && t->val32) rc = gather_tile_values<float>(ctx, h, *t, t->val32);
    }
    return rc;
}

// ---------------------------------------------------------------------------
// tiled SpMV: y = alpha (M x) + beta y, device order
// ---------------------------------------------------------------------------

template <typename T>
struct TileSpmvOp {
    T* y;
    T* acc;
    T alpha, beta;
    int mode;
    __device__ __forceinline__ void finish(int row, T sum) {
        if (mode == kFirst) { acc[row] = sum; return; }
        if (mode == kMiddle) { acc[row] = acc[row] + sum; return; }
        if (mode == kLast) sum = acc[row] + sum;
        y[row] = (beta == T(0)) ? alpha * sum : alpha * sum + beta * y[row];
    }
};

template <typename T>
__global__ void __launch_bounds__(kTileThreads) k_tile_spmv(TileSpmvOp<T> op, TileArgs ta, const T* x) {
    tile_rows<T>(ta, x, op);
}

template <typename T>
static int spmv_tiles_t(sov_hip_ctx* ctx, const Orient& o, double alpha, const void* x, double beta, void* y) {
    const Tiles& t = o.tiles_of(sizeof(T) == 8);
    const size_t lds = tile_lds_bytes(t.W, t.rmax, sizeof(T));
    SOV_HIP_TRY(tile_lds_attr(ctx, reinterpret_cast<const void*>(&k_tile_spmv<T>), sizeof(T)));
    const int np = sizeof(T) == 8 ? t.parts64 : t.parts;
    for (int k = 0; k < np; ++k) {
        const TileArgs ta = tile_args(t, o.nat.rows, o.nat.cols, sizeof(T) == 8, k, np);
        TileSpmvOp<T> op{static_cast<T*>(y),
                         std::is_same<T, double>::value ? reinterpret_cast<T*>(o.acc64) : reinterpret_cast<T*>(o.acc32),
                         T(alpha), T(beta), seg_mode(k, np)};
        k_tile_spmv<T><<<t.nb * t.Q, kTileThreads, lds, ctx->stream>>>(op, ta, static_cast<const T*>(x));
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int spmv_tiles(sov_hip_ctx* ctx, const Orient& o, int prec, double alpha, const void* x, double beta, void* y) {
    return prec == SOV_PREC_F64 ? spmv_tiles_t<double>(ctx, o, alpha, x, beta, y)
                                : spmv_tiles_t<float>(ctx, o, alpha, x, beta, y);
}

}  // namespace sovhip

using namespace sovhip;

extern "C" {

int sov_hip_csr_set_tiles(sov_hip_ctx* ctx, sov_hip_csr* h, int orient, int parts) {
    if (ctx == nullptr || h == nullptr || (orient != 0 && orient != 1) || parts < 0 || parts > 64)
        return fail(SOV_INVALID_ARGUMENT, "set_tiles args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "set_tiles during capture");
    DeviceGuardT g(ctx->device);
    Orient& o = orient == 0 ? h->a : h->t;
    if (parts == 0) {
        o.tiles.free();
        o.tiles64.free();
        return SOV_OK;
    }
    if (o.nat.rows == 0) return fail(SOV_INVALID_ARGUMENT, "set_tiles: orientation has no rows");
    // FP64 products always (FP64 runs and the FP64 polish); FP32 ones when
    // the handle has its FP32 image
    for (const bool f64 : {true, false}) {
        if (!f64 && !h->f32) continue;
        Tiles& t = o.tiles_of(f64);
        if (!t.built()) {
            const int rc = build_tiles(ctx, h, o, f64);
            if (rc != SOV_OK) return rc;
        }
        (f64 ? t.parts64 : t.parts) = std::max(1, std::min(parts, std::max(t.S, 1)));
    }
    return SOV_OK;
}

int sov_hip_tile_layout(int32_t rows, int32_t cols, int32_t nnz, const int32_t* ptr, const int32_t* idx,
                        const int32_t* src, int32_t cus, int32_t tsize, int32_t shape[8], int64_t* entries,
                        uint32_t* keys, int32_t* src_out, int32_t* wstart, int32_t* units) {
    if (rows <= 0 || cols < 0 || nnz < 0 || ptr == nullptr || shape == nullptr || entries == nullptr ||
        (nnz > 0 && (idx == nullptr || src == nullptr)) || (tsize != 4 && tsize != 8))
        return fail(SOV_INVALID_ARGUMENT, "tile_layout args");
    if (ptr[0] != 0 || ptr[rows] != nnz) return fail(SOV_INVALID_ARGUMENT, "tile_layout: row pointers");
    for (int32_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r]) return fail(SOV_INVALID_ARGUMENT, "tile_layout: row pointers");
    for (int32_t q = 0; q < nnz; ++q)
        if (idx[q] < 0 || idx[q] >= cols) return fail(SOV_INVALID_ARGUMENT, "tile_layout: column index");
    std::vector<int32_t> p(ptr, ptr + rows + 1), ix(idx ? idx : ptr, idx ? idx + nnz : ptr),
        sv(src ? src : ptr, src ? src + nnz : ptr);
    Tiles t;
    std::vector<uint32_t> k;
    std::vector<int32_t> so, ws;
    const int rc = layout_host(rows, cols, p, ix, sv, cus, tsize, t, k, so, ws);
    if (rc != SOV_OK) return rc;
    shape[0] = t.nb;
    shape[1] = t.S;
    shape[2] = t.W;
    shape[3] = t.rmax;
    shape[4] = t.U;
    shape[5] = kTileWaves;
    shape[6] = t.Q;
    shape[7] = kTileBuffers;
    *entries = t.E;
    if (keys) std::copy(k.begin(), k.begin() + t.E, keys);
    if (src_out) std::copy(so.begin(), so.begin() + t.E, src_out);
    if (wstart) std::copy(ws.begin(), ws.end(), wstart);
    if (units) std::copy(t.h_units.begin(), t.h_units.end(), units);
    return SOV_OK;
}

int sov_hip_tile_passes(int32_t rows, int32_t cols, int32_t nnz, const int32_t* ptr, const int32_t* idx,
                        const int32_t* src, int32_t cus, int32_t tsize, int64_t* passes, int32_t* pidx,
                        int32_t* pdesc) {
    if (rows <= 0 || cols < 0 || nnz < 0 || ptr == nullptr || passes == nullptr ||
        (nnz > 0 && (idx == nullptr || src == nullptr)) || (tsize != 4 && tsize != 8))
        return fail(SOV_INVALID_ARGUMENT, "tile_passes args");
    if (ptr[0] != 0 || ptr[rows] != nnz) return fail(SOV_INVALID_ARGUMENT, "tile_passes: row pointers");
    for (int32_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r]) return fail(SOV_INVALID_ARGUMENT, "tile_passes: row pointers");
    for (int32_t q = 0; q < nnz; ++q)
        if (idx[q] < 0 || idx[q] >= cols) return fail(SOV_INVALID_ARGUMENT, "tile_passes: column index");
    std::vector<int32_t> p(ptr, ptr + rows + 1), ix(idx ? idx : ptr, idx ? idx + nnz : ptr),
        sv(src ? src : ptr, src ? src + nnz : ptr);
    Tiles t;
    std::vector<uint32_t> k;
    std::vector<int32_t> so, ws, pi, pd;
    int rc = layout_host(rows, cols, p, ix, sv, cus, tsize, t, k, so, ws);
    if (rc == SOV_OK) rc = pass_tables(t, ws, tile_lane(tsize), tile_pairs(tsize), pi, pd);
    if (rc != SOV_OK) return rc;
    *passes = static_cast<int64_t>(pd.size() / 2);
    if (pidx) std::copy(pi.begin(), pi.end(), pidx);
    if (pdesc) std::copy(pd.begin(), pd.end(), pdesc);
    return SOV_OK;
}

int sov_hip_csr_tiles_usable(const sov_hip_ctx* ctx, const sov_hip_csr* h, int orient, int prec, int* usable) {
    if (ctx == nullptr || h == nullptr || usable == nullptr || (orient != 0 && orient != 1) ||
        (prec != SOV_PREC_F64 && prec != SOV_PREC_F32))
        return fail(SOV_INVALID_ARGUMENT, "tiles_usable args");
    const Orient& o = orient == 0 ? h->a : h->t;
    // a 256-byte aligned stand-in for a device vector (hipMalloc alignment)
    const bool f64 = prec == SOV_PREC_F64;
    *usable = tiles_usable(ctx, o.tiles_of(f64), f64, reinterpret_cast<const void*>(uintptr_t{256})) ? 1 : 0;
    return SOV_OK;
}

int sov_hip_csr_get_tiles(const sov_hip_csr* h, int orient, int* parts) {
    if (h == nullptr || parts == nullptr || (orient != 0 && orient != 1))
        return fail(SOV_INVALID_ARGUMENT, "get_tiles args");
    *parts = (orient == 0 ? h->a : h->t).tiles.parts;
    return SOV_OK;
}

int sov_hip_csr_get_tiles_prec(const sov_hip_csr* h, int orient, int prec, int* parts) {
    if (h == nullptr || parts == nullptr || (orient != 0 && orient != 1) ||
        (prec != SOV_PREC_F64 && prec != SOV_PREC_F32))
        return fail(SOV_INVALID_ARGUMENT, "get_tiles_prec args");
    const Orient& o = orient == 0 ? h->a : h->t;
    *parts = prec == SOV_PREC_F64 ? o.tiles64.parts64 : o.tiles.parts;
    return SOV_OK;
}

}  // extern "C"
