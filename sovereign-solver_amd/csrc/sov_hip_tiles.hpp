// LDS-staged row products over the tiled layout (struct Tiles): the device
// routine shared by the tiled SpMV and the tiled PDHG step kernels.
//
// One workgroup of kTileThreads runs one row group.  For every unit u of
// the launch's range, each wave adds the products of its row-aligned
// sub-range of tile (b, u) into the group's LDS row sums acc[local row],
// gathering the vector from LDS (staged unit: the slice was copied there
// beforehand, two buffers when they fit) or from memory (gather unit).
// After the last unit the epilogue functor sees every row of the group
// once, with its full sum.
//
// Summation order (fixed, so results are run-to-run identical): inside a
// tile a row's entries are summed per lane in storage order, lanes combine
// by a segmented Hillis-Steele scan over the wave; the tile sums of a row
// are added to acc in ascending unit order.
#pragma once

#include "sov_hip_internal.hpp"

namespace sovhip {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Read-only tables addressed through the constant address space: the
// compiler treats them as invariant, so uniform reads become scalar loads
// (which never wait behind a wave's vector loads) wherever they occur.
typedef const int32_t __attribute__((address_space(4))) cint32;
__device__ __forceinline__ const cint32* const_table(const int32_t* p) {
    return (const cint32*)(p);
}

// What one tiled launch reads: units [hu[0], hu[Q]) of every group's
// tiles, half h of a group taking units [hu[h], hu[h + 1]).
struct TileArgs {
    const uint32_t* keys;
    const void* vals;         // T
    const int32_t* pidx;      // nb * kTileWaves * (U + 1): first pass of (group, wave) in unit u
    const int32_t* pdesc;     // 2 per pass: first entry, entries in the pass
    const int32_t* units;     // 4 U + 1
    void* part;               // Q * nb * rmax elements of T: the halves' row sums
    int32_t* cnt;             // nb: halves of the group finished in this launch
    int32_t nb, U, W, rmax, rows, cols;
    int32_t Q;
    int32_t hu[kTileHalves + 1];
    int32_t hs[kTileHalves];  // first staged unit of each half (its end if none)
    int32_t sent;             // first entry of the sentinel block
};

// Device row of local row l of group b.
__host__ __device__ __forceinline__ int tile_row(int b, int nb, int l) { return ((l >> 6) * nb + b) * 64 + (l & 63); }

// LDS of a tiled launch: kTileBuffers slice buffers (one is consumed while
// the next slice lands in the other), the row sums and one slot per lane
// that takes the writes a lane makes no real write with.
__host__ __device__ constexpr size_t tile_lds_bytes(int W, int rmax, int tsize) {
    return (size_t)kTileBuffers * W * tsize + (size_t)(rmax + kWave) * tsize;
}

// One lane's L entries of a pass as loaded (sentinels where it has none):
// keys and values, each read with 16-byte loads.
template <typename T, int L>
struct PassData {
    uint32_t k[L];
    T v[L];
    int32_t n;   // entries in the pass (wave-uniform; 0: an empty pass)
};

// The lane's L entries of the pass whose first entry is `start` and which
// holds `cnt` entries; a lane at or past cnt loads the sentinel block (keys
// kTileSentinel, values 0), so what a lane holds needs no mask.  The loads
// are issued on every path and nothing reads the loaded registers here, so
// every refill issues the same loads, the compiler can count exactly how
// many loads follow each queue slot, and the refill never waits for its own
// loads.
template <typename T, int L>
__device__ __forceinline__ PassData<T, L> load_pass(const TileArgs& ta, int32_t start, int32_t cnt, int lane) {
    static_assert(L % 4 == 0, "16-byte key loads");
    PassData<T, L> d;
    const uint32_t at = static_cast<uint32_t>(lane * L < cnt ? start + lane * L : ta.sent);
#pragma unroll
    for (int j = 0; j < L / 4; ++j) {
        const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(ta.keys + at) + j);
        d.k[4 * j] = q.x;
        d.k[4 * j + 1] = q.y;
        d.k[4 * j + 2] = q.z;
        d.k[4 * j + 3] = q.w;
    }
    const T* vals = static_cast<const T*>(ta.vals) + at;
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int j = 0; j < L / 4; ++j) {
            const f32x4 q = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(vals) + j);
            d.v[4 * j] = q.x;
            d.v[4 * j + 1] = q.y;
            d.v[4 * j + 2] = q.z;
            d.v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < L / 2; ++j) {
            const f64x2 q = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(vals) + j);
            d.v[2 * j] = q.x;
            d.v[2 * j + 1] = q.y;
        }
    }
    d.n = cnt;
    return d;
}

// Cross-lane moves on the VALU (DPP, GFX9 controls).  update_dpp with
// old = 0 and bound_ctrl off: a lane whose source lies outside its row (or
// whose row is masked off) reads 0.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;

template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, dpp<CTRL, ROWS>(__builtin_bit_cast(int, v)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = dpp<CTRL, ROWS>(static_cast<int>(b));
    const int hi = dpp<CTRL, ROWS>(static_cast<int>(b >> 32));
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double readlane(double v, int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float readlane(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// One step of the segmented scan: (S, F) of the lane CTRL reaches combined
// in front of the lane's own: S = F ? S : S_src + S, F = F | F_src.
template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ void seg_step(T& S, int& F) {
    const T Su = dpp<CTRL, ROWS>(S);
    const int Fu = dpp<CTRL, ROWS>(F);
    if (!F) S = Su + S;
    F = F | Fu;
}

// Inclusive segmented scan over the 64 lanes: within rows of 16 by shifts
// of 1, 2, 4, 8, then row 0's total into row 1 (and 2's into 3), then the
// first half's into the second half.  A fixed tree: run-to-run identical.
template <typename T>
__device__ __forceinline__ void seg_scan(T& S, int& F) {
    seg_step<kDppRowShr1, 0xF>(S, F);
    seg_step<kDppRowShr2, 0xF>(S, F);
    seg_step<kDppRowShr4, 0xF>(S, F);
    seg_step<kDppRowShr8, 0xF>(S, F);
    seg_step<kDppRowBcast15, 0xA>(S, F);
    seg_step<kDppRowBcast31, 0xC>(S, F);
}

// What a pass computes from its own entries alone: rows, validity, the
// lanes' running sums and the row exchanges with the neighbouring lanes.
template <typename T, int L>
struct PassLocal {
    int r[L];
    T s[L];
    bool ok[L];
    bool endin[L - 1];   // a run ends inside the lane at u < L - 1 (entry u + 1 has another row)
    bool single;         // the lane's entries are one run
    // rows are encoded +1 in the lane exchanges so that 0 (what a lane with
    // no source reads, or a lane whose entry is a sentinel) means "no row"
    int fr1, tl1, prev_tl1, next_fr1, lane0_fr1;
};

template <typename T, int L>
__device__ __forceinline__ PassLocal<T, L> pass_local(const PassData<T, L>& d, const T (&xv)[L]) {
    PassLocal<T, L> p;
#pragma unroll
    for (int u = 0; u < L; ++u) {
        p.ok[u] = d.k[u] < kTileSentinel;
        p.r[u] = static_cast<int>(d.k[u] >> kTileColBits);
    }
    p.s[0] = d.v[0] * xv[0];
    bool any = false;
#pragma unroll
    for (int u = 1; u < L; ++u) {
        const T q = d.v[u] * xv[u];
        p.endin[u - 1] = p.r[u] != p.r[u - 1];
        any = any | p.endin[u - 1];
        p.s[u] = p.endin[u - 1] ? q : p.s[u - 1] + q;
    }
    p.single = !any;
    p.fr1 = p.ok[0] ? p.r[0] + 1 : 0;
    p.tl1 = p.ok[L - 1] ? p.r[L - 1] + 1 : 0;
    p.prev_tl1 = dpp<kDppWaveShr1, 0xF>(p.tl1);   // lane - 1's tail row (+1), 0 for lane 0
    p.next_fr1 = dpp<kDppWaveShl1, 0xF>(p.fr1);   // lane + 1's first row (+1), 0 for lane 63
    p.lane0_fr1 = __builtin_amdgcn_readfirstlane(p.fr1);
    return p;
}

// The segmented inclusive scan of a pass's tails, before the carry is
// added.  carried: lane 0's first run continues the incoming carry row.
template <typename T, int L>
__device__ __forceinline__ void pass_scan(const PassLocal<T, L>& p, bool carried, int lane, T& S, int& F, bool& cont) {
    cont = (p.fr1 != 0) & (lane > 0 ? p.prev_tl1 == p.fr1 : carried);
    S = p.ok[L - 1] ? p.s[L - 1] : T(0);
    F = (p.single & cont) ? 0 : 1;
    seg_scan(S, F);
}

// The pass's writes once its scan has the carry in (cin: what the lane's
// first run takes from earlier lanes): runs ending at u = 0 .. L - 2, and
// the tail when its row ends in this lane (lane 63's goes to the carry
// instead); wm / wv: L entries each.
template <typename T, int L>
__device__ __forceinline__ void pass_writes(const PassLocal<T, L>& p, T S, bool cont, bool carried, T carry_val,
                                            int lane, bool* wm, T* wv) {
    T Sprev = dpp<kDppWaveShr1, 0xF>(S);
    if (lane == 0) Sprev = carried ? carry_val : T(0);
    const T cin = cont ? Sprev : T(0);
#pragma unroll
    for (int u = 0; u < L - 1; ++u) {
        wm[u] = p.endin[u] & p.ok[u];
        wv[u] = p.s[u] + (p.r[u] == p.r[0] ? cin : T(0));
    }
    wm[L - 1] = (p.tl1 != 0) & (lane < kWave - 1) & (p.next_fr1 != p.tl1);
    wv[L - 1] = S;
}

// Add one pass (64 lanes x L entries, rows ascending, sentinels last; xv =
// the gathered vector values of the lane's entries) into acc.  carry_row /
// carry_val: the last row of the previous pass of this sub-range, not yet
// written (-1: none); wave-uniform.
//
// Each lane sums its runs of equal rows in storage order.  A run that ends
// inside the lane (the next entry has another row, a sentinel included) is
// complete, except that the lane's first run may continue a row of earlier
// lanes, whose part arrives as cin.  The lane's tail run (the one holding
// its last entry, if real) may continue into later lanes: the tails are
// joined by a segmented scan over the wave and written by the lane where
// the row ends (lane 63's goes to the carry).  Sentinels have row 8191 and
// value 0, so they never join a real run.
//
// The writes are one batch: all reads of acc first, then all writes.  The
// rows written in one pass are distinct (a lane's runs are distinct rows,
// a run crossing lanes is written once), so the batch reads no value it
// writes; a write a lane does not make goes to its own slot acc[dummy]
// (dummy = rmax + lane), so the batch needs no branches.
template <typename T, int L>
__device__ __forceinline__ void tile_pass(const PassData<T, L>& d, const T (&xv)[L], T* __restrict__ acc, int dummy,
                                          int lane, int& carry_row, T& carry_val) {
    const PassLocal<T, L> p = pass_local<T, L>(d, xv);
    const bool carried = (carry_row >= 0) & (p.lane0_fr1 == carry_row + 1);
    T S;
    int F;
    bool cont;
    pass_scan(p, carried, lane, S, F, cont);
    if (carried & !F) S = carry_val + S;
    bool wm[L];
    T wv[L];
    pass_writes(p, S, cont, carried, carry_val, lane, wm, wv);
    // the carried row, when this pass does not continue it (wave-uniform)
    if ((carry_row >= 0) & !carried) {
        if (lane == 0) acc[carry_row] = acc[carry_row] + carry_val;
    }
    T old[L];
#pragma unroll
    for (int u = 0; u < L; ++u) old[u] = acc[wm[u] ? p.r[u] : 0];
#pragma unroll
    for (int u = 0; u < L; ++u) acc[wm[u] ? p.r[u] : dummy] = old[u] + wv[u];
    const int tl63 = __builtin_amdgcn_readlane(p.tl1, kWave - 1);
    const T s63 = readlane(S, kWave - 1);
    carry_row = tl63 - 1;
    carry_val = tl63 ? s63 : T(0);
}

// Two consecutive passes A, B of one sub-range, with the result of
// tile_pass(A) then tile_pass(B), bit for bit.  B's dependence on A is its
// incoming carry: the carry row is A's lane-63 tail row, known from A's keys
// alone, so both scans run side by side and only the carry value (A's
// lane-63 sum) joins them afterwards.  The two batches are one: A's rows,
// the flushed carries and B's rows are all distinct (rows ascend through
// the sub-range, and A's tail row is written by B or flushed, never both).
template <typename T, int L>
__device__ __forceinline__ void tile_pass2(const PassData<T, L>& da, const T (&xa)[L], const PassData<T, L>& db,
                                           const T (&xb)[L], T* __restrict__ acc, int dummy, int lane, int& carry_row,
                                           T& carry_val) {
    const PassLocal<T, L> a = pass_local<T, L>(da, xa);
    const PassLocal<T, L> b = pass_local<T, L>(db, xb);
    const bool carriedA = (carry_row >= 0) & (a.lane0_fr1 == carry_row + 1);
    const int tlA63 = __builtin_amdgcn_readlane(a.tl1, kWave - 1);
    const int rowA = tlA63 - 1;   // A's outgoing carry row
    const bool carriedB = (rowA >= 0) & (b.lane0_fr1 == rowA + 1);
    T SA, SB;
    int FA, FB;
    bool contA, contB;
    pass_scan(a, carriedA, lane, SA, FA, contA);
    pass_scan(b, carriedB, lane, SB, FB, contB);
    if (carriedA & !FA) SA = carry_val + SA;
    const T sA63 = readlane(SA, kWave - 1);
    const T valA = tlA63 ? sA63 : T(0);   // A's outgoing carry value
    if (carriedB & !FB) SB = valA + SB;
    bool wm[2 * L];
    T wv[2 * L];
    pass_writes(a, SA, contA, carriedA, carry_val, lane, wm, wv);
    pass_writes(b, SB, contB, carriedB, valA, lane, wm + L, wv + L);
    if ((carry_row >= 0) & !carriedA) {
        if (lane == 0) acc[carry_row] = acc[carry_row] + carry_val;
    }
    if ((rowA >= 0) & !carriedB) {
        if (lane == 0) acc[rowA] = acc[rowA] + valA;
    }
    int rw[2 * L];
#pragma unroll
    for (int u = 0; u < L; ++u) {
        rw[u] = a.r[u];
        rw[L + u] = b.r[u];
    }
    T old[2 * L];
#pragma unroll
    for (int u = 0; u < 2 * L; ++u) old[u] = acc[wm[u] ? rw[u] : 0];
#pragma unroll
    for (int u = 0; u < 2 * L; ++u) acc[wm[u] ? rw[u] : dummy] = old[u] + wv[u];
    const int tl63 = __builtin_amdgcn_readlane(b.tl1, kWave - 1);
    const T s63 = readlane(SB, kWave - 1);
    carry_row = tl63 - 1;
    carry_val = tl63 ? s63 : T(0);
}

// x itself, through a DPP move (identity lane pattern).  A gathered value
// passed through it is consumed where it is loaded, so the wait for the
// gather stays on the gather path instead of following the value to where
// the paths join (where it would also wait for the entry queue).
__device__ __forceinline__ float lane_identity(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xE4, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_identity(double x) {
    const long long v = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0xE4, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(v >> 32), 0xE4, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Workgroup barrier for LDS traffic only.  The kernels below share nothing
// between waves through global memory, so the barrier fences the LDS
// address space alone and leaves a wave's global loads in flight
// (a full __syncthreads would wait for all of them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The gathered vector as a buffer resource of cols elements (raw: no
// stride, no format conversion; 32-bit data format as the raw loads need).
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vector_rsrc(const T* x, int cols) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(x), static_cast<short>(0), cols * static_cast<int>(sizeof(T)),
                                             0x00020000);
}

// The group's row sums over this half's units into acc (LDS, rmax
// entries), then op.finish(device row, sum) for each row of the group,
// with the halves' sums added in half order.  x: the gathered vector
// (device order, 16-byte aligned).  All threads call it.
//
// Waves kTileStagers .. 15 add the entries.  Each walks its stream of
// passes (its sub-ranges of the units, in order) with the loads of the
// next tile_ring() passes in flight: a queue of that many register slots,
// each consumed in place and refilled with the pass that far on, so passes
// are used in the order their loads were issued (a wave's vector memory
// counter completes in that order) and no slot is copied while its load is
// pending.  The first kTileStagers waves (one per SIMD, the first to start:
// a workgroup's waves are launched one after another) copy vector slices
// into LDS with direct buffer-to-LDS loads (16 bytes per lane, no
// registers), one staged unit ahead of the entry waves.
// Staging runs in waves of its own so its loads never sit in an entry
// wave's queue.  Every unit ends in a barrier of all waves: the next
// unit's sub-ranges give rows to other waves.

// s_waitcnt immediate that waits until at most n vector memory operations
// of the wave are in flight (expcnt and lgkmcnt left alone).
__host__ __device__ constexpr int vmcnt_imm(int n) { return 0x0F70 | ((n >> 4) << 14) | (n & 15); }

template <typename T, class Op>
__device__ __forceinline__ void tile_rows(const TileArgs& ta, const T* __restrict__ x, Op& op) {
    constexpr int NB = kTileBuffers;
    const int SB = ta.W * static_cast<int>(sizeof(T));   // slice bytes: 64, 48 or 32 KiB
    constexpr int L = tile_lane(sizeof(T));
    extern __shared__ __align__(16) unsigned char tile_lds[];
    T* xs = reinterpret_cast<T*>(tile_lds);
    T* acc = xs + (size_t)NB * ta.W;
    // group g, half h: consecutive workgroups are the halves of one group
    // (workgroups are dealt to the XCDs round-robin, so with two halves the
    // even XCDs stage one half of the vector and the odd ones the other)
    const int g = blockIdx.x / ta.Q, h = blockIdx.x - g * ta.Q;
    const int u0 = ta.hu[h], u1 = ta.hu[h + 1];
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    // wave-uniform, so the unit table and sub-range bounds are scalar loads
    // (their own counter: they never wait for the vector loads in flight)
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const cint32* un = const_table(ta.units);
    for (int l = tid; l < ta.rmax; l += kTileThreads) acc[l] = T(0);
    if (wave < kTileStagers) {
        // ---- staging waves.  With s_0, s_1, ... the staged units of the
        // range: s_0 .. s_{NB-1} are issued before the first barrier (every
        // buffer is free), and s_{j+NB-1} (j >= 1) right after the barrier
        // that starts s_j, into the buffer s_{j-1} used (the unit table
        // gives staged unit j buffer j mod NB).  Before the barrier that
        // starts s_j the wave waits for its own part of s_j, leaving the
        // younger slices' loads in flight.  A slice is kStageChunks chunks
        // of kWave x 16 bytes per staging wave (16, 12 or 8), chunk k of
        // wave w at (k kTileStagers + w) kWave x 16 bytes into the slice.
        const int sw = wave;
        constexpr int NV = 16 / static_cast<int>(sizeof(T));
        constexpr int kChunkBytes = kWave * 16;
        const int kStageChunks = SB / (kChunkBytes * kTileStagers);
        static_assert(NB == 2, "one younger slice at most");
        // The resource ends at the last whole 16-byte vector, so a lane's
        // load is either inside it or wholly past it (the hardware returns
        // zeros there).  A gathered length that is not a whole number of
        // 16-byte vectors leaves rem elements: the lane whose load covers
        // them writes them after its load has landed, read with scalar
        // loads (x does not change during the launch).
        const int rem = ta.cols % NV;
        const __amdgpu_buffer_rsrc_t rs = vector_rsrc(x, ta.cols - rem);
        const int last = (ta.cols - 1) / ta.W;                        // the slice holding them
        const int lo = ta.cols - rem - last * ta.W;                   // their first element there
        const int lo_chunk = lo * static_cast<int>(sizeof(T)) / kChunkBytes;
        const bool owner = rem != 0 && sw == lo_chunk % kTileStagers &&
                           lane == (lo * static_cast<int>(sizeof(T)) % kChunkBytes) / 16;
        auto next_staged = [&](int v) { return v < u1 ? un[4 * v + 3] : u1; };
        auto buffer = [&](int v) { return xs + (size_t)un[4 * v + 2] * ta.W; };
        int issued = 0;   // slices of the range issued so far
        int upcoming = ta.hs[h];   // the next to issue (the half's first staged unit: no table load)
        auto issue_next = [&]() {
            const int base = un[4 * upcoming] * SB;
            unsigned char* buf = reinterpret_cast<unsigned char*>(buffer(upcoming));
#pragma unroll 4
            for (int k = 0; k < kStageChunks; ++k) {
                const int c = (k * kTileStagers + sw) * kChunkBytes;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (__attribute__((address_space(3))) void*)(buf + c), 16, lane * 16, base + c, 0, 0);
            }
            ++issued;
            upcoming = next_staged(upcoming);
        };
        int j = 0;   // ordinal of the next staged unit to start
        // before the barrier that starts staged unit v (ordinal j): the
        // wave's part of v has landed, and the rem tail is written over it
        auto land = [&](int v) {
            const int younger = issued - j - 1;
            if (younger < 1) __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
            else if (kStageChunks == 16) __builtin_amdgcn_s_waitcnt(vmcnt_imm(16));
            else if (kStageChunks == 12) __builtin_amdgcn_s_waitcnt(vmcnt_imm(12));
            else __builtin_amdgcn_s_waitcnt(vmcnt_imm(8));
            if (owner && un[4 * v] == last) {
                typedef const T __attribute__((address_space(4))) ctail;
                const ctail* xt = (const ctail*)(x + (ta.cols - rem));
                T* buf = buffer(v);
                for (int i = 0; i < NV - 1; ++i)
                    if (i < rem) buf[lo + i] = xt[i];
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): the tail is in LDS
            }
        };
        // The staging waves' barrier.  The entry waves' lds_barrier is a
        // release of the LDS address space, which the compiler implements
        // here by waiting for every direct-to-LDS load in flight
        // (vmcnt(0)), the younger slices' included.  What must be in LDS
        // before this barrier, land() has already waited for, so it is a
        // bare s_barrier (the memory clobber keeps the compiler from moving
        // LDS accesses across it).
        auto stage_barrier = []() { asm volatile("s_barrier" ::: "memory"); };
        for (int k = 0; k < NB && upcoming < u1; ++k) issue_next();
        if (u0 < u1 && un[4 * u0 + 1]) land(u0);
        stage_barrier();
        for (int u = u0; u < u1; ++u) {
            if (un[4 * u + 1]) {   // staged unit j has started
                if (j >= 1 && upcoming < u1) issue_next();
                ++j;
            }
            if (u + 1 < u1 && un[4 * (u + 1) + 1]) land(u + 1);
            stage_barrier();
        }
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));   // nothing is left in flight
    } else {
        // ---- entry waves
        // The wave's passes of units [u0, u1) are [i, iend) of the pass
        // list.  Lane l holds the descriptor of pass dbase + l (reloaded
        // every kWave passes); a refill reads its pass's descriptor with
        // readlane, so it waits for no memory.
        const int ew = wave - kTileStagers;   // entry wave
        const cint32* pi = const_table(ta.pidx) + (int64_t)(g * kTileWaves + ew) * (ta.U + 1);
        int i = pi[u0];
        const int iend = pi[u1];
        int ld = i, dbase = i;
        auto load_desc = [&]() {
            const int at = dbase + lane;
            return at < iend ? *reinterpret_cast<const int2*>(ta.pdesc + 2 * (int64_t)at) : make_int2(0, 0);
        };
        int2 dl = load_desc();
        auto fetch = [&]() -> PassData<T, L> {
            if (ld - dbase == kWave) {
                dbase = ld;
                dl = load_desc();
                // consume the reload here, so the wait for it stays on this
                // (rare) path instead of joining every refill
                asm volatile("" : : "v"(dl.x), "v"(dl.y));
            }
            const int j = ld - dbase;
            const int32_t s0 = __builtin_amdgcn_readlane(dl.x, j);
            const int32_t c0 = __builtin_amdgcn_readlane(dl.y, j);
            ++ld;
            return load_pass<T, L>(ta, s0, c0, lane);
        };
        constexpr int R = tile_ring(sizeof(T));
        static_assert(R == 4 || R == 6 || R == 8, "ring");
        PassData<T, L> q0 = fetch(), q1 = fetch(), q2 = fetch(), q3 = fetch();
        PassData<T, L> q4, q5, q6, q7;
        if constexpr (R >= 6) {
            q4 = fetch();
            q5 = fetch();
        }
        if constexpr (R >= 8) {
            q6 = fetch();
            q7 = fetch();
        }
        lds_barrier();
        // consumer state: unit cu, its end in the pass list
        int cu = u0, nxt = i, staged = 0, carry_row = -1;
        T carry_val = T(0);
        const T* xl = xs;
        const T* xg = x;
        const int dummy = ta.rmax + lane;
        auto open_unit = [&](int u) {
            staged = un[4 * u + 1];
            xl = xs + (size_t)un[4 * u + 2] * ta.W;
            xg = x + (int64_t)un[4 * u] * ta.W;
            nxt = pi[u + 1];
            carry_row = -1;
            carry_val = T(0);
        };
        auto close_unit = [&](int u) {
            if (carry_row >= 0 && lane == 0) acc[carry_row] = acc[carry_row] + carry_val;
            lds_barrier();
        };
        auto gather = [&](const PassData<T, L>& slot, T (&xv)[L]) {
            uint32_t q[L];
#pragma unroll
            for (int k = 0; k < L; ++k) q[k] = slot.k[k] & kTileColMask;
            if (staged) {
#pragma unroll
                for (int k = 0; k < L; ++k) xv[k] = xl[q[k]];
            } else {
#pragma unroll
                for (int k = 0; k < L; ++k) xv[k] = lane_identity(xg[q[k]]);
            }
        };
        // the gathers of two passes, all issued before any is waited for
        auto gather2 = [&](const PassData<T, L>& sa, const PassData<T, L>& sb, T (&xa)[L], T (&xb)[L]) {
            uint32_t qa[L], qb[L];
#pragma unroll
            for (int k = 0; k < L; ++k) {
                qa[k] = sa.k[k] & kTileColMask;
                qb[k] = sb.k[k] & kTileColMask;
            }
            if (staged) {
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    xa[k] = xl[qa[k]];
                    xb[k] = xl[qb[k]];
                }
            } else {
                T ga[L], gb[L];
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    ga[k] = xg[qa[k]];
                    gb[k] = xg[qb[k]];
                }
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    xa[k] = lane_identity(ga[k]);
                    xb[k] = lane_identity(gb[k]);
                }
            }
        };
        auto consume_pass = [&](const PassData<T, L>& slot) {
            T xv[L];
            gather(slot, xv);
            tile_pass<T, L>(slot, xv, acc, dummy, lane, carry_row, carry_val);
        };
        // close the units this wave has finished (and those it has no
        // passes in) before pass i
        auto close_done = [&]() {
            while (cu < u1 && i == nxt) {
                close_unit(cu);
                ++cu;
                if (cu < u1) open_unit(cu);
            }
        };
        // two queue slots, passes i and i + 1: when both lie in the current
        // unit (always, as the pass list gives every (wave, unit) an even
        // count) they are added together (tile_pass2: their gathers, scans
        // and row-sum batches overlap), or pass i alone when i + 1 is the
        // empty pass that evens the count; else one at a time with the unit
        // change between them.  Then both slots are refilled (loads are
        // issued on every path, so each slot is always followed by the same
        // number of loads, and the loop has its only exit at the top).
        auto step2 = [&](PassData<T, L>& sa, PassData<T, L>& sb) {
            close_done();
            if (cu < u1) {
                if (i + 1 < nxt && sb.n == 0) {
                    consume_pass(sa);
                    i += 2;
                } else if (i + 1 < nxt) {
                    T xa[L], xb[L];
                    gather2(sa, sb, xa, xb);
                    tile_pass2<T, L>(sa, xa, sb, xb, acc, dummy, lane, carry_row, carry_val);
                    i += 2;
                } else {
                    consume_pass(sa);
                    ++i;
                    close_done();
                    if (cu < u1) {
                        consume_pass(sb);
                        ++i;
                    }
                }
            }
            sa = fetch();
            sb = fetch();
        };
        // one queue slot, pass i (the pass list of unpaired kernels has no
        // empty passes)
        auto step = [&](PassData<T, L>& slot) {
            close_done();
            if (cu < u1) {
                consume_pass(slot);
                ++i;
            }
            slot = fetch();
        };
        if (cu < u1) open_unit(cu);
        if constexpr (tile_pairs(sizeof(T))) {
            while (cu < u1) {
                step2(q0, q1);
                step2(q2, q3);
                if constexpr (R >= 6) step2(q4, q5);
                if constexpr (R >= 8) step2(q6, q7);
            }
        } else {
            while (cu < u1) {
                step(q0);
                step(q1);
                step(q2);
                step(q3);
                if constexpr (R >= 6) {
                    step(q4);
                    step(q5);
                }
                if constexpr (R >= 8) {
                    step(q6);
                    step(q7);
                }
            }
        }
    }
    lds_barrier();
    if (ta.Q == 1) {
        for (int l = tid; l < ta.rmax; l += kTileThreads) {
            const int p = tile_row(g, ta.nb, l);
            if (p < ta.rows) op.finish(p, acc[l]);
        }
        return;
    }
    // Halves: each publishes its row sums with device-scope stores (sc1:
    // written through to where every XCD reads them, and counted as done
    // only once they are), then counts itself in at the group's counter
    // (a device-scope atomic).  The last half to arrive resets the counter
    // for the next launch and finishes the rows, reading the other halves'
    // sums with device-scope loads (past any stale cache line).
    T* part = static_cast<T*>(ta.part);
    T* mine = part + ((size_t)h * ta.nb + g) * ta.rmax;
    for (int l = tid; l < ta.rmax; l += kTileThreads)
        __hip_atomic_store(mine + l, acc[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    // the arrival order goes through lane 0's spare slot (no pass writes
    // it any more), so the kernel needs no LDS beyond tile_lds_bytes
    volatile int* arrived = reinterpret_cast<volatile int*>(acc + ta.rmax);
    __syncthreads();
    if (tid == 0) *arrived = __hip_atomic_fetch_add(ta.cnt + g, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (*arrived != ta.Q - 1) return;
    if (tid == 0) __hip_atomic_store(ta.cnt + g, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int l = tid; l < ta.rmax; l += kTileThreads) {
        const int p = tile_row(g, ta.nb, l);
        if (p >= ta.rows) continue;
        T sum = T(0);
        for (int k = 0; k < ta.Q; ++k) {
            const T v = k == h ? acc[l]
                               : __hip_atomic_load(part + ((size_t)k * ta.nb + g) * ta.rmax + l, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            sum = k == 0 ? v : sum + v;
        }
        op.finish(p, sum);
    }
}

// Products with fewer entries than this are not offered the tiles by the
// autotuner (their vectors stay in L2 and the CSR kernels win).
constexpr int64_t kTileTuneMinNnz = 1 << 20;
// The autotuner keeps the tiles when their SpMV takes at most this factor
// of the CSR parts' time.
constexpr float kTileTuneMargin = 1.1f;

// Host: the launch arguments of slice range k of `parts` for one precision.
TileArgs tile_args(const Tiles& t, int32_t rows, int32_t cols, bool f64, int k, int parts);

// Host: build the tiled layout of an orientation's products of one
// precision (o.tiles64 when f64, else o.tiles) from its device-order CSR,
// with that precision's image of A's natural FP64 values.
int build_tiles(sov_hip_ctx* ctx, const sov_hip_csr* h, Orient& o, bool f64);

// Host: re-derive both layouts' value images from A's natural FP64 values.
int refresh_tile_values(sov_hip_ctx* ctx, const sov_hip_csr* h, Orient& o);

// Host: whether a tiled launch of precision f64 is possible on this device
// for these tiles (LDS) and these vectors (alignment).
bool tiles_usable(const sov_hip_ctx* ctx, const Tiles& t, bool f64, const void* x);

// Host: let `kernel` (a tiled kernel of element size tsize) allocate the LDS
// its largest launch needs.
hipError_t tile_lds_attr(const sov_hip_ctx* ctx, const void* kernel, int tsize);

// Host: y = alpha (M x) + beta y over the tiles of an orientation.
int spmv_tiles(sov_hip_ctx* ctx, const Orient& o, int prec, double alpha, const void* x, double beta, void* y);

}  // namespace sovhip
