// Internal types and device helpers shared by the sov_hip translation units.
// gfx950 (CDNA4): 64-wide wavefronts; all reductions are written for wave64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sov_hip.h"

namespace sovhip {

// ---------------------------------------------------------------------------
// error plumbing: nothing throws across the C-ABI
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define SOV_HIP_TRY(expr)                                            \
    do {                                                             \
        hipError_t _e = (expr);                                      \
        if (_e != hipSuccess) return ::sovhip::hip_fail(_e, #expr);  \
    } while (0)

constexpr int kBlock = 256;          // threads per workgroup (4 waves)
constexpr int kWave = 64;
constexpr int kXcds = 8;             // gfx950: workgroup p of a launch runs on XCD p % 8

// XCD-aware order of a launch's n work items: workgroup p (on XCD p % kXcds)
// takes item xcd_slot(p, n), so XCD k runs the k-th contiguous share of the
// items (q + 1 items for k < n % kXcds, else q = n / kXcds) and its L2 holds
// the operands of that share only.  A bijection of [0, n) (checked below).
__host__ __device__ constexpr int xcd_slot(int p, int n) {
    const int q = n / kXcds, r = n % kXcds, k = p % kXcds;
    return k * q + (k < r ? k : r) + p / kXcds;
}
constexpr bool xcd_slot_bijective(int nmax) {
    for (int n = 0; n <= nmax; ++n) {
        bool seen[200] = {};
        for (int p = 0; p < n; ++p) {
            const int s = xcd_slot(p, n);
            if (s < 0 || s >= n || seen[s]) return false;
            seen[s] = true;
        }
    }
    return true;
}
static_assert(xcd_slot_bijective(200), "xcd_slot must permute [0, n)");
constexpr int kClasses = 8;          // row-length classes of a schedule
constexpr int kCheckSlots = 16;      // FP64 partials per block in the check kernels
// Blocks of check partials the allocation-free check's static scratch holds.
constexpr int64_t kStaticCheckBlocks = 65536;

// Row-length class boundaries: class c holds rows with length <= kClassMax[c].
constexpr int kClassMax[kClasses] = {2, 4, 8, 16, 32, 64, 256, 0x7fffffff};
// Default lanes per row for each class.
constexpr int kDefaultLanes[kClasses] = {1, 2, 4, 4, 8, 16, 32, 64};

// Device order.  Every vector the PDHG kernels touch is stored permuted:
// position p of the row dimension holds natural row perm_r[p], position q of
// the column dimension natural column perm_c[q], where each permutation
// groups the rows of its orientation (rows of A for perm_r, rows of A^T =
// columns of A for perm_c) by length class, stably.  Row p of an orientation
// therefore writes element p of its output, the rows of one class are one
// contiguous range, and every per-row access of a kernel is coalesced.  The
// stored matrix indices are device positions of the gathered dimension, so
// the gathers read the permuted vectors directly.  (For skewed column counts
// the heavy columns also end up adjacent, which keeps their gathers in L2.)

// Schedule passed by value to the row kernels.  Rows of class c are
// [offset[c], offset[c] + count[c]); they are processed by G[c] lanes per
// row, kBlock / G[c] rows per workgroup, in workgroups
// [block_begin[c], block_begin[c+1]).  Kernels read the per-class values with
// pick() (unrolled selects on kernel-argument scalars), never by a runtime
// array index, so nothing spills to scratch.
struct Sched {
    int32_t G[kClasses];
    int32_t count[kClasses];
    int32_t offset[kClasses];
    int32_t block_begin[kClasses + 1];
};

// One CSR "part" of an orientation, in device order: all its columns, or one
// column segment (entries whose device column lies in [k W, (k+1) W)).  Every
// part holds every row (possibly empty), so all parts share the class ranges.
struct Part {
    int32_t rows = 0, cols = 0, nnz = 0;
    const int32_t* ptr = nullptr;   // rows + 1
    const int32_t* idx = nullptr;   // nnz, device positions in the gathered dimension
    double* val64 = nullptr;        // nnz
    float* val32 = nullptr;         // nnz, optional
    const int32_t* src = nullptr;   // nnz: position of each entry in A's natural value array
    int32_t class_count[kClasses] = {};
    int32_t class_offset[kClasses] = {};
    int32_t lanes[kClasses] = {};
    Sched sched() const;
    int blocks() const;
};

// The natural (row-ordered, caller-indexed) CSR of an orientation: what the
// scaling kernels walk row by row and what device order is derived from.
struct Natural {
    int32_t rows = 0, cols = 0, nnz = 0;
    const int32_t* ptr = nullptr;
    const int32_t* idx = nullptr;
    double* val64 = nullptr;
    const int32_t* src = nullptr;   // null for A itself
};

// Tiled layout of an orientation for the LDS-staged step kernels.
//
// Rows (device order) are cut into 64-row chunks; chunk c belongs to group
// c % nb, where it holds local rows (c / nb) * 64 .. + 63, so every group
// mixes all length classes and the per-row work is coalesced per chunk.
// A group is run by Q workgroups ("halves"), one per CU: each takes a
// contiguous range of the units (below) and keeps the group's row sums over
// its units in LDS; the last half to finish adds the halves' sums in half
// order and runs the per-row epilogue.  So a CU stages only its half of the
// gathered vector, and each of its units holds the entries of twice as many
// rows as one workgroup per group would give it.
//
// The gathered dimension is cut into slices of W = tile_slice_bytes() /
// (element size) columns, and consecutive slices form units.  A
// staged unit is one slice whose entries are dense enough per group (on
// average at least kTileStageMin per group) to repay copying the vector
// slice into one of the kTileBuffers LDS buffers; its entries then gather
// from LDS.  A gather unit is a run of up to tile_gather_slices(W) sparser
// slices whose entries gather straight from the vector in memory (staging
// them would move more bytes than it saves).
//
// Entries are stored group-major, then unit, then wave: each (group, unit)
// tile is split at row boundaries into kTileWaves contiguous sub-ranges of
// about equal size, one per entry wave, each padded to a multiple of
// tile_lane() entries with key kTileSentinel (value 0); kTileLaneMax more
// sentinels follow the last entry (what a lane with no entries in a pass loads).  key = (local
// row << kTileColBits) | (column - first column of the unit).
// wstart[(b U + u) kTileWaves + w] is the first entry of wave w of tile
// (b, u); the array ends with the total count E.  units[4 u .. 4 u + 3] =
// {first slice, staged, LDS buffer of a staged unit (its ordinal among the
// staged units modulo kTileBuffers), next staged unit (U if none)},
// units[4 U] = S.
//
// What the kernels read instead of wstart: the pass list.  A pass is up to
// kWave * tile_lane() consecutive entries of one sub-range; only a sub-range's
// last pass is short, and for kernels that add passes in pairs
// (tile_pairs) an empty pass (count 0) evens each (wave, unit) count.  The
// passes of wave w of group b are listed unit by unit: pidx[(b kTileWaves +
// w) (U + 1) + u] is the index of its first pass in unit u (entry U: one
// past its last), and pdesc[2 i], pdesc[2 i + 1] = first entry and entry
// count of pass i.
constexpr int kTileThreads = 1024;                    // workgroup of a half
constexpr int kTileStagers = 4;                       // waves that copy vector slices into LDS (one per SIMD)
constexpr int kTileWaves = kTileThreads / kWave - kTileStagers;   // 12 waves that add the entries
// Entries per lane per pass.  Measured on the 1M-column Zipf LP: 8 FP32
// entries per lane (one segmented scan and one row-sum batch per 512
// entries) made the FP32 steps slower, 64.5 / 68.4 us against 59.1 /
// 61.8 us with 4 and pairs of passes (with pairs of 8-entry passes: 69.0 /
// 70.4 us, 109 VGPRs): the pass's latency chain, not its instruction
// count, sets the time.
__host__ __device__ constexpr int tile_lane(int) { return 4; }
constexpr int kTileLaneMax = 8;                       // sentinel entries after the last one
constexpr int kTileHalves = 2;                        // workgroups (CUs) per row group
constexpr int kTileLdsBytes = 163840;                 // LDS of one gfx950 workgroup
constexpr int kTileColBits = 19;                      // key bits of the column within a unit
constexpr uint32_t kTileColMask = (1u << kTileColBits) - 1u;
constexpr uint32_t kTileSentinel = 0xFFF80000u;       // row 8191, column 0: never a real key (rows < kTileRowCap)
// Two LDS slice buffers: one is consumed while the next slice lands in the
// other.  A layout's slice is 64 KiB for FP32 and 48 KiB for FP64, or the
// largest smaller one (48, 32 KiB) whose two buffers fit beside the row
// sums (tile_slice_bytes).  Measured on the 1M-column Zipf LP (7,872 /
// 3,968 rows per group in the primal / dual step): FP64 with three 32 KiB
// buffers 108.7 / 93.4 us, two 48 KiB 92.5-94.4 / 84.2-85.3 us, two 64 KiB
// (dual only) 88.5 us; FP32 with two 64 KiB 58.1 / 60.1 us, two 48 KiB
// 62.4 / 65.4 us.
constexpr int kTileBuffers = 2;
// local rows per group: 13 key bits, and the LDS of two 32 KiB buffers
constexpr int kTileRowCap = 8128;
static_assert(kTileRowCap < 8191, "sentinel row");
__host__ __device__ constexpr size_t tile_lds_bytes(int W, int rmax, int tsize);
// slices per gather unit: its columns must fit the key's column bits
__host__ __device__ constexpr int tile_gather_slices(int W) {
    return ((1 << kTileColBits) / W - 1) < 63 ? ((1 << kTileColBits) / W - 1) : 63;
}
// The slice bytes of a layout with rmax local rows per group (0: none fits).
__host__ __device__ constexpr int tile_slice_bytes(int rmax, int tsize) {
    for (int sb = tsize == 4 ? 65536 : 49152; sb >= 32768; sb -= 16384)
        if ((size_t)kTileBuffers * sb + (size_t)(rmax + kWave) * tsize <= (size_t)kTileLdsBytes) return sb;
    return 0;
}
// Entries per group that make a slice staged.  Measured on the 1M-column
// Zipf LP (FP32 dual step, where slices of 400-1100 entries per group lie
// between the short columns and the dense ones): 384 stages them, 69.6 us;
// 800 gathers the sparser ones from memory in one unit, 65.6 us (fewer
// barrier-ended units); 1200, 65.7 us; the FP64 steps and the FP32 primal
// step do not move.  Again with the two-buffer 64 KiB slices: 500 / 800 /
// 1400 give an FP32 dual step of 61.1 / 59.3 / 59.3 us and an FP64 one of
// 87.5 / 85.1 / 90.4 us.
constexpr int kTileStageMin = 800;
// Whether the tiled kernels of element size tsize add passes in pairs (and
// the pass list evens every (wave, unit) count for them).  Measured on the
// 1M-column Zipf LP: pairs cut the FP32 primal step from 66 to 60 us; the
// FP64 tiles hold about one pass per wave and unit, where the padding only
// halves the load queue's lookahead.
constexpr bool tile_pairs(int tsize) { return tsize == 4; }
// Passes in an entry wave's load queue (register slots).  Measured on the
// 1M-column Zipf LP with 12 entry waves: 8 FP32 / 6 FP64 slots (95 / 92
// VGPRs) made the FP32 primal step slower (58.5 -> 61.3 us) and left the
// others within noise, so the entry data is not what the waves wait for.
constexpr int tile_ring(int) { return 4; }

struct Tiles {
    int32_t nb = 0, S = 0, W = 0, rmax = 0;           // groups, slices, slice width, local rows per group
    int32_t Q = 1;                                    // halves (workgroups) per group
    int32_t U = 0;                                    // units
    int32_t cols = 0;                                 // length of the gathered vector
    int64_t E = 0;                                    // stored entries (with padding)
    uint32_t* keys = nullptr;                         // E
    int32_t* src = nullptr;                           // E: A natural position, -1 for padding
    int32_t* pidx = nullptr;                          // nb * kTileWaves * (U + 1)
    int32_t* pdesc = nullptr;                         // 2 * passes
    int64_t passes = 0;
    int32_t* units = nullptr;                         // 4 U + 1
    std::vector<int32_t> h_units;                     // host copy of units
    std::vector<double> h_cost;                       // U: each unit's estimated time (splits a launch into halves)
    void* part = nullptr;                             // Q * nb * rmax elements: the halves' row sums
    int32_t* cnt = nullptr;                           // nb: halves finished (0 between launches)
    float* val32 = nullptr;                           // E + kTileLaneMax: the FP32 layout's values
    double* val64 = nullptr;                          // E + kTileLaneMax: the FP64 layout's values
    int parts = 0;                                    // FP32 layout: launches per product (0 = on the CSR parts)
    int parts64 = 0;                                  // FP64 layout: the same
    bool built() const { return pidx != nullptr; }
    void free();
};

// An orientation (A or A^T): its natural CSR, its device order, the
// device-order CSR the parts are cut from, and the parts (one, or one per
// column segment when the gathered vector is larger than a comfortable share
// of an XCD's L2).
struct Orient {
    Tiles tiles;                    // FP32 products (W = 16384)
    Tiles tiles64;                  // FP64 products (W = 8192)
    Natural nat;
    int32_t* perm = nullptr;        // rows: natural row at each device position
    int32_t* inv = nullptr;         // rows: device position of each natural row
    int32_t* d_ptr = nullptr;       // rows + 1: device-order CSR
    int32_t* d_idx = nullptr;       // nnz (device positions of the gathered dimension)
    int32_t* d_src = nullptr;       // nnz
    int32_t class_count[kClasses] = {};
    int32_t class_offset[kClasses] = {};
    std::vector<Part> parts;
    int64_t parts_version = 0;      // bumped whenever the parts are rebuilt
    int32_t seg_width = 0;          // 0 when unsegmented
    // owned storage of the parts
    int32_t* p_ptr = nullptr;       // nparts * (rows + 1)
    int32_t* p_idx = nullptr;       // nnz
    int32_t* p_src = nullptr;       // nnz
    double* p_val64 = nullptr;      // nnz
    float* p_val32 = nullptr;       // nnz
    double* acc64 = nullptr;        // rows: running partial sums across segments
    float* acc32 = nullptr;         // rows
    double* chk_acc = nullptr;      // 2 * rows: the checks' two running dots
    Tiles& tiles_of(bool f64) { return f64 ? tiles64 : tiles; }
    const Tiles& tiles_of(bool f64) const { return f64 ? tiles64 : tiles; }
    int nparts() const { return static_cast<int>(parts.size()); }
    const Part& part(int k) const { return parts[static_cast<size_t>(k)]; }
    Part& part(int k) { return parts[static_cast<size_t>(k)]; }
    int max_blocks() const;
    void free_parts();
    void free_order();
};

// How a part's launch combines with the others of its orientation.
enum SegMode : int {
    kOnly = 0,    // unsegmented: sum is the full row product
    kFirst = 1,   // acc = sum
    kMiddle = 2,  // acc += sum
    kLast = 3,    // sum += acc, then the epilogue
};

__host__ __device__ constexpr int seg_mode(int k, int nparts) {
    return nparts == 1 ? kOnly : (k == 0 ? kFirst : (k == nparts - 1 ? kLast : kMiddle));
}

}  // namespace sovhip

struct sov_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    int cus = 0;
    int64_t l2_bytes = 0;
    int64_t mall_bytes = 0;
    int64_t lds_max = 65536;   // LDS bytes one workgroup may allocate
    // FP64 scratch for two-stage reductions (device) and pinned host staging
    double* d_red = nullptr;
    size_t d_red_count = 0;
    double* h_pinned = nullptr;   // 64 doubles
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool capturing = false;
    // device address of the allocation-free check's static scratch on this
    // context's device (sovhip::static_scratch_address)
    double* check_scratch = nullptr;
    int ensure_red(size_t count);
};

struct sov_hip_csr {
    int32_t m = 0, n = 0, nnz = 0;
    sovhip::Orient a;     // A   : m rows, gathers n-vectors
    sovhip::Orient t;     // A^T : n rows, gathers m-vectors
    int32_t* t_ptr = nullptr;   // owned natural A^T structure
    int32_t* t_idx = nullptr;
    int32_t* t_src = nullptr;   // owned: A position of each A^T entry
    double* t_val64 = nullptr;  // owned natural A^T values
    bool owns_a_val64 = false;
    bool f32 = false;           // parts carry FP32 value images (sov_hip_csr_make_f32)
    double* partials = nullptr; // owned: kCheckSlots * (largest part blocks of both orientations + 64) FP64
    void* pairs = nullptr;      // owned: 2 (n + m) FP64, a check's (current, sum) pairs of x then y
    void* vx = nullptr;         // owned: max(m, n) FP64, device-order staging of natural-order SpMV operands
    void* vy = nullptr;
    int64_t partial_blocks = 0;
    int device = 0;
    hipStream_t stream = nullptr;   // the creating context's stream (destroy waits for it only)
};

// A block-diagonal batch of LPs stacked in one sov_hip_csr: instance b owns
// natural rows [row_off[b], row_off[b+1]) and columns [col_off[b],
// col_off[b+1]).  Device order groups rows by length class stably, so inside
// each class range the rows of one instance stay contiguous: seg holds, per
// instance, dimension (0 rows, 1 columns) and class, the (start, length) of
// that run, which makes every per-instance reduction a sum over 8 contiguous
// ranges.
struct sov_hip_batch {
    int32_t count = 0, m = 0, n = 0;
    int device = 0;
    hipStream_t stream = nullptr;   // the creating context's stream (destroy waits for it only)
    int32_t* row_inst = nullptr;    // m, device order
    int32_t* col_inst = nullptr;    // n, device order
    int32_t* seg = nullptr;         // count * 2 * kClasses * 2
    double* tau = nullptr;          // count
    double* sigma = nullptr;        // count
    int32_t* active = nullptr;      // count
    int32_t* since = nullptr;       // count
    int32_t* flags = nullptr;       // count
    int32_t* use_avg = nullptr;     // count
    double* fac = nullptr;          // count
    double* terms = nullptr;        // 11 * (m + n): per-row check terms
    // per step-kernel block of each part of A (blk[0]) and A^T (blk[1]):
    // bit 0 / 1 = holds a row the CSR step launch of the FP32 / FP64 state
    // runs (an instance iterating in it and not resident now), bit 2 / 3 =
    // holds a row of an instance iterating in FP32 /
    // FP64; and per part the compacted list of the blocks with bit j set
    // (list[.][4q + j]): bits 0-1 for the step launches, 2-3 for the checks.
    // Recomputed whenever `active` is set.
    std::vector<int32_t*> blk[2];
    std::vector<int32_t*> list[2];
    std::vector<int32_t> nlist[2];
    std::vector<int64_t> blk_cap[2];
    int32_t* scan = nullptr;        // staging for the compaction
    int64_t scan_cap = 0;
    int64_t version = 0;            // bumped when the block lists change
    const sov_hip_csr* csr = nullptr;
    double* red = nullptr;          // count * 22
    int32_t* chunks = nullptr;      // the check reduction's chunk table, n_chunks records of 8
    int32_t n_chunks = 0;
    double* chunk_part = nullptr;   // 11 per split chunk: their block results
    int32_t* chunk_cnt = nullptr;   // count * 2: arrivals per split instance dimension (0 between launches)
    double* h_red = nullptr;        // pinned host copy, count * 22
    // Resident instances: small instances whose iterates, objective and
    // bounds fit stage_bytes of LDS and whose rows fit the resident kernel's
    // registers run whole intervals in one workgroup each
    // (sov_hip_batch_resident_steps) while every instance iterating in
    // their precision fits; the CSR step launches then skip them (mask).
    // Built from single-part A and A^T (none when either has several
    // parts), again whenever a part is rebuilt.
    struct Resident {
        int64_t built_from[2] = {-1, -1};          // parts_version of A and A^T the indices follow
        uint16_t* lidx[2] = {nullptr, nullptr};    // per orientation: per entry, its column's place in its instance
        int32_t* mask[2] = {nullptr, nullptr};     // per precision (0 FP32, 1 FP64), per instance: 1 = resident now
        std::vector<int32_t> h_mask[2];            // per precision, per instance: 1 = fits
        std::vector<int32_t> h_list[2];            // host copy of list (the source of its upload)
        std::vector<int32_t> h_inst[2];            // per precision: the resident instances
        int32_t* list[2] = {nullptr, nullptr};     // per precision: the resident instances iterating in it
        int32_t nlist[2] = {0, 0};
        int64_t lds[2] = {0, 0};                   // per precision: LDS bytes of the largest resident instance
        int on[2] = {0, 0};                        // per precision: mask holds h_mask (1) or zeros (0)
    } res;
    // The resident launches (and with `split` the FP64 step launches) run on
    // the side stream, beside the context stream's launches: the first such
    // launch after a join forks it from the context stream (ev_fork) and
    // sov_hip_batch_resident_join makes the context stream wait for it
    // (ev_join).
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    mutable bool forked = false;    // work is queued on `side` since the last join
    bool split = false;             // FP64 step launches on `side` too (sov_hip_batch_set_split)
    int64_t stage_bytes = 0;        // LDS a resident instance may take (0: none resident)
    int64_t lds_max = 0;            // the device's LDS per workgroup
    // Tiled steps: a precision's step launches run on the tiled layouts of
    // the whole stacked matrix (the rows of instances not iterating in it
    // are summed and dropped) while its iterating instances hold at least
    // tile_frac of the entries, else on the CSR block lists.  Decided
    // whenever `active` is set; negative tile_frac: never.
    std::vector<int64_t> h_nnz;     // entries of each instance
    int64_t nnz_total = 0;
    double tile_frac = -1.0;
    bool tiles_on[2] = {false, false};
    std::vector<int32_t> h_seg;     // host copy of seg
    std::vector<int32_t> h_active;  // host copy of active
};

struct sov_hip_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace sovhip {

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// max that keeps a NaN from either side (residual maxima must not lose one)
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off, kWave));
    return v;
}

// Block-wide reduction of NS sums and NM maxima; thread 0 writes them to
// out[0..NS) then out[NS..NS+NM), element i at out[i * stride].  All
// threads of the block must call it.
template <int NS, int NM>
__device__ __forceinline__ void block_reduce_write(const double* s, const double* mx,
                                                   double* out, int64_t stride = 1) {
    constexpr int W = kBlock / kWave;
    __shared__ double lds[(NS + NM) * W];
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const double r = wave_sum(s[i]);
        if (lane == 0) lds[i * W + wv] = r;
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const double r = wave_max(mx[i]);
        if (lane == 0) lds[(NS + i) * W + wv] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            double acc = 0.0;
            for (int w = 0; w < W; ++w) acc += lds[i * W + w];
            out[i * stride] = acc;
        }
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            double acc = 0.0;
            for (int w = 0; w < W; ++w) acc = nmax(acc, lds[(NS + i) * W + w]);
            out[(NS + i) * stride] = acc;
        }
    }
}

// The class a schedule block belongs to (uniform per workgroup).
__device__ __forceinline__ int block_class(const Sched& s, int blk) {
    int c = 0;
#pragma unroll
    for (int k = 1; k < kClasses; ++k)
        if (blk >= s.block_begin[k]) c = k;
    return c;
}

// a[c] for a runtime c without a dynamic index (keeps the kernel-argument
// arrays in scalar registers).
template <int N>
__device__ __forceinline__ int pick(const int32_t (&a)[N], int c) {
    int v = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k) v = (c == k) ? a[k] : v;
    return v;
}

// Where a workgroup's rows come from: resolved once per block.  blk is the
// schedule block the workgroup runs: blockIdx.x, or an entry of a compacted
// block list (batches skip the blocks of stopped instances).
struct ClassView {
    int cls, G, begin, count, offset, blk;
};

__device__ __forceinline__ ClassView class_view(const Sched& s, int blk = -1) {
    ClassView v;
    v.blk = blk >= 0 ? blk : (int)blockIdx.x;
    v.cls = block_class(s, v.blk);
    v.G = pick(s.G, v.cls);
    v.begin = pick(s.block_begin, v.cls);
    v.count = pick(s.count, v.cls);
    v.offset = pick(s.offset, v.cls);
    return v;
}

// The device-order row of this thread's group of G lanes (-1 past its
// class) and the lane within the group.
template <int G>
__device__ __forceinline__ int row_of(const ClassView& v, int& lane) {
    constexpr int RPB = kBlock / G;
    const int local = (v.blk - v.begin) * RPB + (int)threadIdx.x / G;
    lane = (int)threadIdx.x % G;
    return (local < v.count) ? v.offset + local : -1;
}

// Entries per lane that one chunk of a row product loads before it gathers:
// all index/value loads of a chunk are issued together, then all gathers,
// so a lane waits for two memory round trips per chunk of kChunk entries.
constexpr int kChunk = 8;

// Dot product of row r with a gathered vector, G lanes per row.  Every lane
// of the group returns the full sum.  Accumulation type Acc.  Each lane adds
// its entries in ascending position (chunking does not change the order).
template <int G, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc row_dot(const int32_t* __restrict__ ptr,
                                       const int32_t* __restrict__ idx,
                                       const V* __restrict__ val, const X* __restrict__ x,
                                       int r, int lane) {
    Acc acc = Acc(0);
    if (r >= 0) {
        const int b = ptr[r], e = ptr[r + 1];
        for (int base = b + lane; base < e; base += kChunk * G) {
            int j[kChunk];
            V w[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int k = base + u * G;
                j[u] = -1;
                w[u] = V(0);
                if (k < e) {
                    j[u] = idx[k];
                    w[u] = val[k];
                }
            }
            X xv[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                xv[u] = X(0);
                if (j[u] >= 0) xv[u] = x[j[u]];
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                if (j[u] >= 0) acc += Acc(w[u]) * Acc(xv[u]);
        }
    }
    return group_sum<G>(acc);
}

// Two dots of the same row (current and second vector scaled by s2), FP64.
// P: x1 holds both vectors interleaved (x1[2 j], x1[2 j + 1] = the first's
// and the second's element j) and x2 is unused, so each entry makes one
// gather of a pair instead of two; the sums are the same.
template <int G, bool P = false, typename V, typename X>
__device__ __forceinline__ void row_dot2(const int32_t* __restrict__ ptr,
                                         const int32_t* __restrict__ idx,
                                         const V* __restrict__ val, const X* __restrict__ x1,
                                         const X* __restrict__ x2, double s2, int r, int lane,
                                         double& d1, double& d2) {
    typedef X X2 __attribute__((ext_vector_type(2)));
    double a1 = 0.0, a2 = 0.0;
    if (r >= 0) {
        const int b = ptr[r], e = ptr[r + 1];
        for (int base = b + lane; base < e; base += kChunk * G) {
            int j[kChunk];
            V w[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int k = base + u * G;
                j[u] = -1;
                w[u] = V(0);
                if (k < e) {
                    j[u] = idx[k];
                    w[u] = val[k];
                }
            }
            X v1[kChunk], v2[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                v1[u] = X(0);
                v2[u] = X(0);
                if (j[u] >= 0) {
                    if constexpr (P) {
                        const X2 q = reinterpret_cast<const X2*>(x1)[j[u]];
                        v1[u] = q.x;
                        v2[u] = q.y;
                    } else {
                        v1[u] = x1[j[u]];
                        v2[u] = x2[j[u]];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                if (j[u] >= 0) {
                    const double v = double(w[u]);
                    a1 += v * double(v1[u]);
                    a2 += v * (double(v2[u]) * s2);
                }
        }
    }
    d1 = group_sum<G>(a1);
    d2 = group_sum<G>(a2);
}

// Dispatch a functor templated on G with the block's lane count.
template <class F>
__device__ __forceinline__ void dispatch_G(const ClassView& v, F& f) {
    switch (v.G) {
        case 1: f.template run<1>(v); break;
        case 2: f.template run<2>(v); break;
        case 4: f.template run<4>(v); break;
        case 8: f.template run<8>(v); break;
        case 16: f.template run<16>(v); break;
        case 32: f.template run<32>(v); break;
        default: f.template run<64>(v); break;
    }
}

// Host helpers
int launch_blocks(int64_t n);      // blocks for an elementwise kernel over n
hipError_t check_launch();         // hipGetLastError wrapper
int build_parts(sov_hip_ctx* ctx, sov_hip_csr* h, Orient& o, int K, bool want_f32);
int ensure_partials(sov_hip_csr* csr);
int inclusive_scan(sov_hip_ctx* ctx, int32_t* a, int64_t len);   // in place, synchronous
// Device address of the allocation-free check's static scratch on `device`
// (which must be current); looked up once per device.
int static_scratch_address(int device, double** addr);

}  // namespace sovhip
