/*
 * sov_hip.h -- the C-ABI over this project's hand-written gfx950 kernels.
 *
 * This is the drop-in boundary under the reference's L0 Backend seam
 * (reference include/sovereign/backend.hpp:82-165).  The C++ wrapper
 * sovereign-solver_amd/src/l0/backend_hip.cpp implements BackendKind::Hip on
 * top of it, and src/l1/pdhg.cpp drives the fused PDHG kernels through it.
 * Plain pointers and sizes only; nothing throws across it.
 *
 * Return codes equal sov::Status (reference include/sovereign/status.hpp:25-34):
 * 0 Ok, 1 Unsupported, 2 NotImplemented, 3 InvalidArgument, 4 DimensionMismatch,
 * 5 OutOfMemory, 6 BackendError, 7 NumericalFailure.  sov_hip_last_error()
 * returns the calling thread's last message.
 *
 * Threading: a context owns one HIP stream and its own scratch; use one
 * context per host thread.  Every call is ordered on that stream; calls that
 * return host scalars synchronise that stream only (never the device).
 * A handle (csr, batch, mlp) remembers the stream of the context that
 * created it; its destroy waits for that stream only, so it must run before
 * that context is destroyed.  (The runtime's own hipFree may still wait for
 * other streams' queued work: that delays the freeing thread, never theirs.)
 *
 * Device pointers are those returned by sov_hip_malloc (or any allocation of
 * the same device).  "Vector" arguments whose element type depends on a
 * precision argument are void*: SOV_PREC_F64 -> double, SOV_PREC_F32 -> float.
 */
#ifndef SOV_HIP_H
#define SOV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SOV_OK = 0,
    SOV_UNSUPPORTED = 1,
    SOV_NOT_IMPLEMENTED = 2,
    SOV_INVALID_ARGUMENT = 3,
    SOV_DIMENSION_MISMATCH = 4,
    SOV_OUT_OF_MEMORY = 5,
    SOV_BACKEND_ERROR = 6,
    SOV_NUMERICAL_FAILURE = 7
};

enum { SOV_PREC_F64 = 0, SOV_PREC_F32 = 1 };

typedef struct sov_hip_ctx sov_hip_ctx;
typedef struct sov_hip_csr sov_hip_csr;
typedef struct sov_hip_graph sov_hip_graph;

/* ---- runtime / context ------------------------------------------------
 * Replaces the reference HIP backend's constructor and device query
 * (src/l0/backend_hip.cpp:56-90). */
int sov_hip_device_count(int* count);
int sov_hip_ctx_create(int device, sov_hip_ctx** out);
void sov_hip_ctx_destroy(sov_hip_ctx* ctx);
const char* sov_hip_ctx_device_name(const sov_hip_ctx* ctx);
/* Number of compute units and the device's L2/MALL sizes, for schedules. */
int sov_hip_ctx_props(const sov_hip_ctx* ctx, int* cus, int64_t* l2_bytes, int64_t* mall_bytes);
const char* sov_hip_last_error(void);

/* ---- memory -----------------------------------------------------------
 * Backend::allocate / deallocate / copy_to_device / copy_to_host /
 * copy_device_to_device / fill_zero / synchronize (backend.hpp:100-108;
 * reference HIP versions backend_hip.cpp:92-124).  Copies are ordered on the
 * context stream; h2d/d2h return when the copy is complete.  sov_hip_sync
 * synchronises the context's stream, not the device. */
int sov_hip_malloc(sov_hip_ctx* ctx, size_t bytes, void** out);
int sov_hip_free(sov_hip_ctx* ctx, void* ptr);
int sov_hip_memcpy_h2d(sov_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int sov_hip_memcpy_d2h(sov_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int sov_hip_memcpy_d2d(sov_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int sov_hip_fill_zero(sov_hip_ctx* ctx, void* dst, size_t bytes);
int sov_hip_sync(sov_hip_ctx* ctx);

/* ---- asynchronous device -> host copies ---------------------------------
 * No reference counterpart (the reference's checkpoint hand-off copies
 * synchronously, pdhg.cpp:336-345 / :373-381).  A copier owns a second
 * stream of default priority: sov_hip_copier_d2h queues a copy into host
 * memory that starts once everything queued so far on the context stream
 * is done and returns at once, so the context stream goes on with later
 * work (which must not write `src` until the copy is done: copy from a
 * snapshot); sov_hip_copier_wait blocks until the copier's copies are done.
 * dst should be pinned host memory (sov_hip_host_alloc) for the copy to run
 * asynchronously.  A copier is used by one thread. */
typedef struct sov_hip_copier sov_hip_copier;
int sov_hip_copier_create(sov_hip_ctx* ctx, sov_hip_copier** out);
void sov_hip_copier_destroy(sov_hip_copier* c);
int sov_hip_copier_d2h(sov_hip_ctx* ctx, sov_hip_copier* c, void* dst, const void* src, size_t bytes);
int sov_hip_copier_wait(sov_hip_copier* c);
int sov_hip_host_alloc(sov_hip_ctx* ctx, size_t bytes, void** out);
int sov_hip_host_free(sov_hip_ctx* ctx, void* ptr);

/* ---- level 1, FP64 ----------------------------------------------------
 * Backend::axpy/scale/dot/norm2/max_abs (backend.hpp:111-115; reference
 * rocBLAS calls backend_hip.cpp:126-146, Unsupported max_abs :169-175).
 * dot/nrm2/amax write a HOST scalar; reductions are two-stage with a fixed
 * order, so results are run-to-run reproducible.  n == 0 is a no-op (dot and
 * nrm2 write 0). */
int sov_hip_axpy_f64(sov_hip_ctx* ctx, int64_t n, double alpha, const double* x, double* y);
int sov_hip_scal_f64(sov_hip_ctx* ctx, int64_t n, double alpha, double* x);
int sov_hip_dot_f64(sov_hip_ctx* ctx, int64_t n, const double* x, const double* y, double* result);
int sov_hip_nrm2_f64(sov_hip_ctx* ctx, int64_t n, const double* x, double* result);
int sov_hip_amax_f64(sov_hip_ctx* ctx, int64_t n, const double* x, double* result);
/* Backend::project_box (backend.hpp:135; reference kernel
 * backend_hip.cpp:46-53).  lo or hi may be NULL (unbounded side). */
int sov_hip_project_box_f64(sov_hip_ctx* ctx, int64_t n, const double* lo, const double* hi,
                            double* x);
/* Precision images. */
int sov_hip_convert_f64_f32(sov_hip_ctx* ctx, int64_t n, const double* src, float* dst);
int sov_hip_convert_f32_f64(sov_hip_ctx* ctx, int64_t n, const float* src, double* dst);

/* ---- CSR handle -------------------------------------------------------
 * Replaces DeviceCsr's device half and the per-call cuSPARSE descriptors of
 * the reference (backend_cuda.cu:242-297).  The handle takes A (m x n, nnz,
 * int32 row_ptr/col_idx with strictly ascending columns, FP64 values) already
 * on the device and builds A^T as its own CSR ON THE DEVICE (rows of A^T in
 * ascending A-row order), so the transposed product is a gather with no
 * atomics and no extra host->device traffic.  It also builds the row-length
 * schedules both orientations use.
 *   flags & SOV_CSR_OWN_VALUES: copy the values so sov_hip_csr_scale may
 *     rescale them in place (otherwise A's value array is borrowed and must
 *     outlive the handle; the structure arrays are always borrowed).
 * Device order: the handle fixes a permutation of the rows (which = 0) and
 * of the columns (which = 1): the rows of A, resp. of A^T, grouped by
 * length class (<= 2, 4, 8, 16, 32, 64, 256, more entries), stably.  The
 * PDHG kernels and sov_hip_spmv_device read and write every vector in this
 * order (row p of an orientation writes element p; all per-row accesses are
 * coalesced); sov_hip_spmv and the scaling entry points take natural order.
 * Row-length schedules: each class is processed by G lanes per row,
 * G in {1,..,64}; the lane counts come from sov_hip_csr_tune (an autotuner
 * that times candidates) or from a length heuristic. */
enum { SOV_CSR_OWN_VALUES = 1 };
int sov_hip_csr_create(sov_hip_ctx* ctx, int32_t m, int32_t n, int32_t nnz,
                       const int32_t* d_row_ptr, const int32_t* d_col_idx, const double* d_values,
                       int flags, sov_hip_csr** out);
void sov_hip_csr_destroy(sov_hip_csr* csr);
int sov_hip_csr_dims(const sov_hip_csr* csr, int32_t* m, int32_t* n, int32_t* nnz);
/* Build the FP32 value images of both orientations from the current FP64 values. */
int sov_hip_csr_make_f32(sov_hip_ctx* ctx, sov_hip_csr* csr);
/* A <- diag(row_scale) A diag(col_scale) on the owned values (both
 * orientations), value = (v * r_i) * c_j; entries whose magnitude falls to
 * <= 1e-14 become exact zeros (the reference drops them when it rebuilds
 * the scaled problem, scaling.cpp:194-204 / sparse.cpp:238). */
int sov_hip_csr_scale(sov_hip_ctx* ctx, sov_hip_csr* csr, const double* d_row_scale,
                      const double* d_col_scale);
/* Copy the current FP64 values of A^T into dst (nnz) / its row_ptr (n+1) and
 * col_idx (nnz) -- for tests. */
int sov_hip_csr_transpose_arrays(sov_hip_ctx* ctx, const sov_hip_csr* csr, int32_t* d_tptr,
                                 int32_t* d_tidx, double* d_tval);
/* Set the lanes-per-row of each length class: lanes[c] for class c in
 * [0, 8), both orientations (orient 0 = A, 1 = A^T).  0 keeps the heuristic. */
int sov_hip_csr_set_lanes(sov_hip_csr* csr, int orient, const int32_t lanes[8]);
int sov_hip_csr_get_lanes(const sov_hip_csr* csr, int orient, int32_t lanes[8]);
/* Column segments: split an orientation into `count` ranges of equal width
 * of the gathered dimension (in device order), stored segment-major, so each
 * pass gathers from a slice of the vector that stays resident in an XCD's
 * L2.  The step, check and SpMV kernels run the segments in order and carry
 * per-row partial sums between them (a fixed summation order: results are
 * reproducible).  count = 1 removes the split.  The scaling kernels use the
 * natural CSR. */
int sov_hip_csr_set_segments(sov_hip_ctx* ctx, sov_hip_csr* csr, int orient, int count);
int sov_hip_csr_get_segments(const sov_hip_csr* csr, int orient, int* count);
/* LDS-staged tiles (no reference counterpart: a layout of the same
 * products).  The rows of an orientation are dealt in 64-row chunks to
 * row groups (one workgroup each, about one per CU); the gathered vector is
 * cut into 64 KiB slices (16384 FP32 or 8192 FP64 elements).  Slices dense
 * enough in entries per group are copied into LDS by each workgroup with
 * 16-byte loads issued two staged slices ahead, and their entries gather
 * from LDS; runs of sparser slices gather from memory, as the CSR kernels
 * do.  Each precision has its own layout.  A product runs on the tiles
 * only when its gathered vector is 16-byte aligned and shorter than 2 GiB
 * (else on the CSR parts).  parts >= 1
 * builds the FP64 layout (and, after make_f32, the FP32 one) and runs the
 * orientation's SpMV and PDHG steps on it, in `parts` launches over ranges
 * of slices joined by per-row partial sums; parts = 0 frees both and
 * returns to the CSR parts (the autotuner chooses per precision).  Launches
 * need 160 KiB of LDS per workgroup and otherwise use the CSR parts.  The
 * summation order is fixed (results are reproducible run to run) but
 * differs from the CSR kernels'.  get_tiles reports the FP32 launch count
 * (0: FP32 products on the CSR). */
int sov_hip_csr_set_tiles(sov_hip_ctx* ctx, sov_hip_csr* csr, int orient, int parts);
int sov_hip_csr_get_tiles(const sov_hip_csr* csr, int orient, int* parts);
/* The launches per product of the tiles of precision prec (0: that
 * precision's products run on the CSR parts).  sov_hip_csr_get_tiles gives
 * the FP32 value.  No reference counterpart. */
int sov_hip_csr_get_tiles_prec(const sov_hip_csr* csr, int orient, int prec, int* parts);
/* *usable = 1 when the orientation's products of precision prec run on the
 * tiles (tiles set, value image present, LDS of the device large enough,
 * gathered vector shorter than 2 GiB). */
int sov_hip_csr_tiles_usable(const sov_hip_ctx* ctx, const sov_hip_csr* csr, int orient, int prec, int* usable);
/* The tiled layout of a host CSR (row pointers, column indices in
 * [0, cols), and a source position per entry), as sov_hip_csr_set_tiles
 * builds it for a device with `cus` compute units and the products of
 * element size tsize (4: FP32, slices of 16384 columns; 8: FP64, slices of
 * 6144; either fewer when a group's row sums leave less LDS).  shape = {groups, slices, slice width, local rows per group,
 * units, sub-ranges per tile, workgroups per group, LDS slice buffers};
 * *entries = stored entries including padding.  keys / src_out (entries each), wstart (groups * units *
 * sub-ranges + 1) and units (4 * units + 1: first slice, staged, buffer,
 * next staged unit; then the slice count) may be null (first call for the
 * sizes).  Host only: no device is touched. */
int sov_hip_tile_layout(int32_t rows, int32_t cols, int32_t nnz, const int32_t* row_ptr,
                        const int32_t* col_idx, const int32_t* src, int32_t cus, int32_t tsize,
                        int32_t shape[8], int64_t* entries, uint32_t* keys, int32_t* src_out,
                        int32_t* wstart, int32_t* units);
/* The pass list the tiled kernels read for the same layout: *passes = pass
 * count; pidx (groups * sub-ranges per tile * (units + 1)) = index of the
 * first pass of (group, wave) in each unit, then one past its last; pdesc
 * (2 * passes) = first entry and entry count of each pass (at most 256
 * entries, inside one sub-range, only the sub-range's last one short; for
 * tsize 4, whose kernels add passes in pairs, an empty pass evens each
 * (group, wave, unit) count).  pidx / pdesc may be null (first call for
 * the size).  Host only.  No reference counterpart (layout of the tiled
 * products that replace Backend::spmv / spmv_transpose). */
int sov_hip_tile_passes(int32_t rows, int32_t cols, int32_t nnz, const int32_t* row_ptr,
                        const int32_t* col_idx, const int32_t* src, int32_t cus, int32_t tsize,
                        int64_t* passes, int32_t* pidx, int32_t* pdesc);
/* Autotuner: for each orientation, try 1, 2, 4 and 8 column segments (slices
 * of at least 64k elements); for each, time every G in {1,2,4,8,16,32,64} per
 * length class of every segment on the SpMV of precision prec; keep the
 * fastest segment count with its lanes.  Orientations with at least 2^20
 * entries then also time the tiles (1, 2 and 4 launches), in precision prec
 * and in the other one (FP32 iterates end in an FP64 polish), and keep them
 * for each precision where they beat the CSR parts. */
int sov_hip_csr_tune(sov_hip_ctx* ctx, sov_hip_csr* csr, int prec);

/* ---- sparse level 2 ---------------------------------------------------
 * Backend::spmv / spmv_transpose (backend.hpp:124-129; reference cuSPARSE
 * calls backend_cuda.cu:220-227).  y := alpha*op(A)*x + beta*y, op = A
 * (transpose == 0) or A^T (transpose == 1).  beta == 0 overwrites y (NaN-safe).
 * prec selects the value image and the element type of x and y. */
int sov_hip_spmv(sov_hip_ctx* ctx, const sov_hip_csr* csr, int transpose, int prec,
                 double alpha, const void* x, double beta, void* y);
/* The same product with x and y in device order (no staging copies).
 * No reference counterpart: the fused engine's power iteration. */
int sov_hip_spmv_device(sov_hip_ctx* ctx, const sov_hip_csr* csr, int transpose, int prec,
                        double alpha, const void* x, double beta, void* y);
/* The device-order permutation of rows (which = 0, m entries) or columns
 * (which = 1, n entries) into host memory: perm[p] = natural index at
 * device position p.  No reference counterpart. */
int sov_hip_csr_order(sov_hip_ctx* ctx, const sov_hip_csr* csr, int which, int32_t* perm);
/* Reorder a device vector of the row (which = 0) or column (which = 1)
 * dimension: natural -> device (to_natural = 0) or back (1).  src and dst
 * must not alias.  No reference counterpart. */
int sov_hip_csr_permute(sov_hip_ctx* ctx, const sov_hip_csr* csr, int which, int to_natural, int prec,
                        const void* src, void* dst);

/* ---- scaling (kernel d) -----------------------------------------------
 * Ruiz and Pock-Chambolle factors (reference scaling.cpp:36-101) from the
 * handle's CURRENT values.  One Ruiz sweep: row/column inf-norms of
 * |a_ij| r_i c_j, r_i /= sqrt(row norm), c_j /= sqrt(col norm) where the norm
 * is positive; *worst receives max |norm - 1| (host scalar).  Pock-Chambolle:
 * r_i /= sqrt(sum_j m^(2-alpha)), c_j /= sqrt(sum_i m^alpha), m = |a| r c.
 * Column reductions run over the A^T CSR (no atomics); sums are serial per
 * row in ascending index order, so factors match the host restatement. */
int sov_hip_scale_ruiz_sweep(sov_hip_ctx* ctx, const sov_hip_csr* csr, double* d_row_scale,
                             double* d_col_scale, double* worst);
int sov_hip_scale_pock_chambolle(sov_hip_ctx* ctx, const sov_hip_csr* csr, double alpha,
                                 double* d_row_scale, double* d_col_scale);

/* ---- fused PDHG step kernels (kernels a+b) ----------------------------
 * One iteration of reference pdhg.cpp:295-331 is two launches:
 *   primal: per column j, g = c_j - (A^T y)_j; x+ = clamp(x - tau g, lo, hi);
 *           x_bar = 2 x+ - x; x = x+; x_sum = (first ? x+ : x_sum + x+)
 *   dual:   per row i, s+ = clamp(s - tau y, lo, hi); s_bar = 2 s+ - s;
 *           y+ = y + sigma (s_bar - (A x_bar)_i); s = s+; y = y+;
 *           y_sum = (first ? y+ : y_sum + y+)
 * All vectors have the element type of prec and are in the handle's device
 * order (sov_hip_csr_order); bounds may hold +-inf. */
typedef struct sov_hip_pdhg_state {
    int32_t prec;
    int32_t m, n;
    const void* c;       /* n: internal-sense (minimize) scaled objective */
    const void* col_lo;  /* n */
    const void* col_hi;  /* n */
    const void* row_lo;  /* m */
    const void* row_hi;  /* m */
    void* x;             /* n */
    void* x_bar;         /* n */
    void* x_sum;         /* n */
    void* s;             /* m */
    void* y;             /* m */
    void* y_sum;         /* m */
    /* optional (NULL = none), m: an objective on the slacks, so that the
     * dual step's slack update is s+ = clamp(s - tau (y + s_grad)) -- the
     * correction problems of iterative refinement need it; batch steps
     * ignore it */
    const void* s_grad;
} sov_hip_pdhg_state;

int sov_hip_pdhg_primal_step(sov_hip_ctx* ctx, const sov_hip_csr* csr,
                             const sov_hip_pdhg_state* st, double tau, int first);
int sov_hip_pdhg_dual_step(sov_hip_ctx* ctx, const sov_hip_csr* csr,
                           const sov_hip_pdhg_state* st, double tau, double sigma, int first);

/* ---- KKT / restart reductions (kernel c) ------------------------------
 * Evaluates, in FP64 on the device, for the current point (x, y) and the
 * Cesaro average (x_sum, y_sum) / since_restart:
 *   kkt_*   the scaled-space max-norm residual of pdhg.cpp:87-136
 *           (bound_tol = tol_scaled)
 *   orig_*  the original-space residual of pdhg.cpp:145-199
 *           (bound_tol = tol_orig), via x = c_s x_s, y = sign r_s y_s
 *   rel_*   relative KKT [r_p, r_d, gap] and the primal/dual objectives
 *           (problem sense, with obj_const) of the original-space point.
 * Max reductions propagate NaN (a NaN anywhere reads back as NaN).  If
 * act_avg is not NULL it receives A_s (x_sum / since_restart) (m, FP64),
 * the activity a restart needs.  Every array is in device order. */
typedef struct sov_hip_check_args {
    const sov_hip_pdhg_state* st;
    int64_t since_restart;
    double tol_scaled;
    double tol_orig;
    double obj_sign;
    double obj_const;
    const double* row_scale;    /* m */
    const double* col_scale;    /* n */
    const double* orig_row_lo;  /* m */
    const double* orig_row_hi;  /* m */
    const double* orig_col_lo;  /* n */
    const double* orig_col_hi;  /* n */
    const double* orig_c;       /* n, problem sense */
    double* act_avg;            /* m, optional */
    /* Iterative refinement (no reference counterpart; all NULL / ignored
     * otherwise): with base set, st is a correction (sov_hip_pdhg_anchor)
     * and the check evaluates the point anchor + inv_delta * correction,
     * the current one and the average, against the LP whose FP64 scaled
     * data (c, col_lo, col_hi, row_lo, row_hi) base holds.  base_x = x0,
     * base_y = y0, base_act = A x0, base_aty = A^T y0 (FP64, device
     * order).  The products run on the correction, in its precision; the
     * anchor's terms are added in FP64, so an FP32 correction is checked to
     * FP32 rounding of the distance from the anchor, not of the point.
     * act_avg then receives the combined average's activity. */
    const sov_hip_pdhg_state* base;
    const double* base_x;       /* n */
    const double* base_y;       /* m */
    const double* base_act;     /* m */
    const double* base_aty;     /* n */
    double inv_delta;
} sov_hip_check_args;

typedef struct sov_hip_check_out {
    double kkt_cur, kkt_avg;
    double orig_cur, orig_avg;
    double rel_cur[3], rel_avg[3];
    double pobj_cur, dobj_cur, pobj_avg, dobj_avg;
} sov_hip_check_out;

int sov_hip_pdhg_check(sov_hip_ctx* ctx, const sov_hip_csr* csr, const sov_hip_check_args* a,
                       sov_hip_check_out* out);

/* The same check without any allocation.  No reference counterpart: it is
 * for callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels, summation order and
 * results as sov_hip_pdhg_check; its block partials live in a static device
 * array of the library (8 MiB per GPU, kept until the process ends) instead
 * of the handle's and context's buffers.  A check that needs more than 65536
 * blocks of partials returns SOV_INVALID_ARGUMENT (use sov_hip_pdhg_check).
 * All calls on one GPU share that array: calls that overlap in time on one
 * GPU (from two contexts of that device) race and must be serialised by the
 * caller. */
int sov_hip_pdhg_check_static(sov_hip_ctx* ctx, const sov_hip_csr* csr, const sov_hip_check_args* a,
                              sov_hip_check_out* out);

/* Restart to the average (pdhg.cpp:393-404): x = clamp(x_sum/k, col box),
 * y = y_sum/k, s = clamp(act_avg, row box).  If x_last/y_last are not NULL
 * (FP64, n and m), movement[0] = ||x_new - x_last||_2^2 and
 * movement[1] = ||y_new - y_last||_2^2 are returned (host) and x_last/y_last
 * are overwritten with the new point (primal-weight bookkeeping). */
int sov_hip_pdhg_restart(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, int64_t since_restart,
                         const double* act_avg, double* x_last, double* y_last,
                         double movement[2]);

/* Restart in place at the current point (adaptive restarts keep it when it
 * beats the average): movement[0] = ||x - x_last||_2^2, movement[1] =
 * ||y - y_last||_2^2 (host), then x_last = x, y_last = y (FP64 copies). */
int sov_hip_pdhg_movement(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, double* x_last,
                          double* y_last, double movement[2]);

/* Set (x, s, y) from FP64 sources: x = clamp(x_src, col box), y = y_src,
 * s = clamp(act, row box); NULL sources leave the vector as is.  Used for
 * warm starts and for the FP32 -> FP64 switch. */
int sov_hip_pdhg_set_point(sov_hip_ctx* ctx, const sov_hip_pdhg_state* st, const double* x_src,
                           const double* y_src, const double* act);

/* Iterative refinement (no reference counterpart).  With an FP64 anchor
 * (x0, y0, s0 = A x0) and a scale delta > 0, the substitution
 *   x = x0 + dx / delta,  s = s0 + ds / delta,  y = y0 + dy / delta
 * turns the scaled saddle problem c x + y (s - A x) into one of the same
 * form in (dx, ds, dy): objective delta (c - A^T y0), column box
 * delta (col - x0), row box delta (row - s0) and slack objective delta y0.
 * PDHG on it, with the same step sizes, takes the same steps as PDHG on
 * the original from the anchor, so FP32 storage of the corrections keeps
 * FP64-level precision in the combined point.
 *
 * sov_hip_pdhg_anchor writes that problem into `correction` (c, col_lo,
 * col_hi, row_lo, row_hi and the m values at s_grad, which must be set)
 * from the FP64 scaled data of `base` (c, col_lo, col_hi, row_lo, row_hi)
 * and sets its start: x = Proj(0), s = Proj(0), y = 0.  aty0 = A^T y0.
 * All vectors in device order. */
int sov_hip_pdhg_anchor(sov_hip_ctx* ctx, const sov_hip_pdhg_state* base, const sov_hip_pdhg_state* correction,
                        const double* x0, const double* y0, const double* s0, const double* aty0, double delta);

/* The point an anchor and a correction stand for, into the FP64 state dst:
 *   x = x0 + inv_delta x_c,  x_sum = since x0 + inv_delta x_sum_c,
 *   y = y0 + inv_delta y_c,  y_sum = since y0 + inv_delta y_sum_c
 * (the correction of either precision), so sov_hip_pdhg_check on dst
 * evaluates the combined current point and its average. */
int sov_hip_pdhg_combine(sov_hip_ctx* ctx, const sov_hip_pdhg_state* correction, const sov_hip_pdhg_state* dst,
                         const double* x0, const double* y0, int64_t since, double inv_delta);

/* ---- batches of LPs ---------------------------------------------------
 * No reference counterpart (many LPs solved in one launch sequence).  A batch is a
 * block-diagonal stack of independent LPs in ONE csr handle: instance b owns
 * natural rows [row_off[b], row_off[b+1]) and columns [col_off[b],
 * col_off[b+1]) (count + 1 offsets each, covering the matrix).  All state
 * vectors are the stacked ones, in the handle's device order.  Each
 * instance has its own step sizes (tau, sigma), an active flag (inactive
 * instances are skipped by the step kernels) and `since` = the iterations its
 * Cesaro sums hold.  One step launch covers every instance; k is the
 * iteration's index inside a captured interval: an instance starts a new
 * Cesaro sum at k == 0 when its `since` is 0.  active: 0 = stopped, 1 =
 * iterating in an FP32 state, 2 = in an FP64 state; a step launch on a state
 * of one precision skips the rows of the other instances, so a batch can
 * keep an FP32 and an FP64 state side by side and move instances between
 * them one at a time (sov_hip_batch_promote). */
typedef struct sov_hip_batch sov_hip_batch;
int sov_hip_batch_create(sov_hip_ctx* ctx, const sov_hip_csr* csr, int32_t count, const int32_t* row_off,
                         const int32_t* col_off, sov_hip_batch** out);
void sov_hip_batch_destroy(sov_hip_batch* batch);
/* Resident instances: an instance whose iterates, objective and bounds
 * ((6 n_b + 5 m_b) values of the state's precision) fit `bytes` of LDS, and
 * whose rows fit the resident kernel's registers (with the default lanes per
 * length class: at most 80 wave-sized chunks of A^T's rows with 2 entries
 * per lane, 48 of A's rows with 4), runs whole intervals in one workgroup
 * (sov_hip_batch_resident_steps) while every instance iterating in its
 * precision fits (then that precision makes no step launch; otherwise all
 * of its instances stay on the step launches), and the batched step
 * launches skip it.  Default: the device's LDS per workgroup (160 KiB);
 * 0 = none resident.
 * Takes effect at the next sov_hip_batch_set with `active`.  No reference
 * counterpart. */
int sov_hip_batch_set_stage_bytes(sov_hip_batch* batch, int64_t bytes);
/* Tiled batch steps: while the instances iterating in a precision hold at
 * least `share` of the batch's entries, that precision's step launches run
 * on the csr handle's tiled layouts of the whole stacked matrix (built by
 * sov_hip_csr_set_tiles; rows of other instances are summed and dropped),
 * else on the CSR parts' block lists.  Negative = never (the default).
 * Takes effect at the next sov_hip_batch_set with `active`.  No reference
 * counterpart. */
int sov_hip_batch_set_tile_share(sov_hip_batch* batch, double share);
/* Split precisions: when on, the FP64 state's batched step launches run on
 * the batch's own stream (forked from the context stream at the first such
 * launch, like sov_hip_batch_resident_steps), beside the FP32 state's on the
 * context stream; sov_hip_batch_resident_join joins it back and must come
 * before anything reads the FP64 state.  Default off.  Not between a fork
 * and its join.  No reference counterpart. */
int sov_hip_batch_set_split(sov_hip_batch* batch, int on);
/* `count` iterations (k = 0 .. count-1 of an interval) of the resident
 * instances iterating in st's precision, in one launch: each instance's
 * workgroup keeps its iterates in LDS and its entries in registers for the
 * whole interval.  With the csr handle's default lanes (no autotuning) they
 * are the iterations `count` pairs of sov_hip_batch_primal_step / _dual_step
 * calls would make, bit for bit; with tuned lanes they differ in rounding.
 * The launch runs on the batch's own stream, after everything queued so far
 * on the context stream and beside what is queued there next (the step
 * launches, which skip the resident instances); sov_hip_batch_resident_join
 * makes the context stream wait for it.  sov_hip_batch_check, _restart and
 * _promote join first by themselves; anything else that reads the iterates
 * (a copy, a product) must come after a join.  No reference counterpart
 * (a persistent batched launch). */
int sov_hip_batch_resident_steps(sov_hip_ctx* ctx, const sov_hip_csr* csr, const sov_hip_batch* batch,
                                 const sov_hip_pdhg_state* st, int count);
int sov_hip_batch_resident_join(sov_hip_ctx* ctx, const sov_hip_batch* batch);
/* Upload per-instance parameters from host arrays of `count` (NULL leaves
 * that parameter as it is).  Defaults: tau = sigma = 1, active = 2, since = 0. */
int sov_hip_batch_set(sov_hip_ctx* ctx, sov_hip_batch* batch, const double* tau, const double* sigma,
                      const int32_t* active, const int32_t* since);
int sov_hip_batch_primal_step(sov_hip_ctx* ctx, const sov_hip_csr* csr, const sov_hip_batch* batch,
                              const sov_hip_pdhg_state* st, int k);
int sov_hip_batch_dual_step(sov_hip_ctx* ctx, const sov_hip_csr* csr, const sov_hip_batch* batch,
                            const sov_hip_pdhg_state* st, int k);
/* The check of sov_hip_pdhg_check for every instance at once (averages over
 * each instance's own `since`; a->since_restart is ignored); obj_const is a
 * host array of `count` (NULL = 0); out receives `count` results. */
int sov_hip_batch_check(sov_hip_ctx* ctx, const sov_hip_csr* csr, const sov_hip_batch* batch,
                        const sov_hip_check_args* a, const double* obj_const, sov_hip_check_out* out);
/* Restart the instances whose host flag is 1 to their average (as
 * sov_hip_pdhg_restart, with each instance's `since`). */
int sov_hip_batch_restart(sov_hip_ctx* ctx, const sov_hip_batch* batch, const sov_hip_pdhg_state* st,
                          const int32_t* flags, const double* act_avg);
/* For the instances whose host flag is 1: write their candidate from the
 * FP32 state (the current point, or the average x_sum/since when use_avg[b])
 * into the FP64 state's x and x_sum (y and y_sum likewise).  A following
 * sov_hip_batch_restart of the FP64 state with since = 1 then places them
 * there, with s from the caller's FP64 activity. */
int sov_hip_batch_promote(sov_hip_ctx* ctx, const sov_hip_batch* batch, const sov_hip_pdhg_state* st32,
                          const sov_hip_pdhg_state* st64, const int32_t* flags, const int32_t* use_avg);
/* Per-instance sum of squares of an FP64 vector of the row (which = 0) or
 * column (1) dimension into host out[count]; and v *= factors[instance]. */
int sov_hip_batch_sumsq(sov_hip_ctx* ctx, const sov_hip_batch* batch, int which, const double* v, double* out);
int sov_hip_batch_scale(sov_hip_ctx* ctx, const sov_hip_batch* batch, int which, const double* factors,
                        double* v);

/* ---- the warm-start network (BF16 MFMA MLP) -----------------------------
 * No reference counterpart.  A fully connected network with `layers` layers
 * of widths widths[0] (input) .. widths[layers] (output), GELU after every
 * layer but the last, BF16 weights and activations on the matrix cores with
 * FP32 accumulation, FP32 master parameters trained with Adam.  Parameters
 * are one flat FP32 array: for each layer W (out x in, row-major) then b.
 * Batches are row-major [batch x features] FP32 device arrays. */
typedef struct sov_hip_mlp sov_hip_mlp;
int sov_hip_mlp_create(sov_hip_ctx* ctx, int32_t layers, const int32_t* widths, int32_t max_batch,
                       sov_hip_mlp** out);
void sov_hip_mlp_destroy(sov_hip_mlp* mlp);
int64_t sov_hip_mlp_num_params(const sov_hip_mlp* mlp);
/* Host FP32 parameters in (and the Adam moments reset) / out; grads (host,
 * optional) receives the last backward pass's gradients. */
int sov_hip_mlp_set_params(sov_hip_ctx* ctx, sov_hip_mlp* mlp, const float* params);
int sov_hip_mlp_get_params(sov_hip_ctx* ctx, const sov_hip_mlp* mlp, float* params, float* grads);
/* y = net(x) for `batch` rows (y may be NULL: the activations are kept for
 * the backward pass either way). */
int sov_hip_mlp_forward(sov_hip_ctx* ctx, sov_hip_mlp* mlp, int32_t batch, const float* d_x, float* d_y);
/* Gradients of a loss with dL/dy = d_dy (device, batch x out) for the last
 * forward pass, into the network's gradient array. */
int sov_hip_mlp_backward(sov_hip_ctx* ctx, sov_hip_mlp* mlp, const float* d_dy);
/* One Adam step with the current gradients (step counts from 1). */
int sov_hip_mlp_adam(sov_hip_ctx* ctx, sov_hip_mlp* mlp, float lr, float beta1, float beta2, float eps,
                     int32_t step);
/* Weighted mean squared error L = sum_ij w_j (y_ij - t_ij)^2 / (batch out)
 * (d_weight NULL = all ones): dL/dy into d_dy, L into *loss (host, optional). */
int sov_hip_mlp_mse(sov_hip_ctx* ctx, sov_hip_mlp* mlp, int32_t batch, const float* d_y, const float* d_target,
                    const float* d_weight, float* d_dy, double* loss);

/* ---- launch graphs ----------------------------------------------------
 * Capture the context stream's launches between begin and end into a graph
 * and replay it.  Only the step kernels above may be captured (no host
 * scalars, no allocation inside a capture). */
int sov_hip_graph_begin(sov_hip_ctx* ctx);
int sov_hip_graph_end(sov_hip_ctx* ctx, sov_hip_graph** out);
int sov_hip_graph_launch(sov_hip_ctx* ctx, sov_hip_graph* g);
void sov_hip_graph_destroy(sov_hip_graph* g);

/* ---- timing -------------------------------------------------------------
 * Event pair on the context stream, for in-process kernel timing. */
int sov_hip_timer_start(sov_hip_ctx* ctx);
int sov_hip_timer_stop(sov_hip_ctx* ctx, double* milliseconds);
/* An array of `count` events; record event i on the context stream; after a
 * sov_hip_sync, elapsed time between events i and j. */
typedef struct sov_hip_events sov_hip_events;
int sov_hip_events_create(sov_hip_ctx* ctx, int count, sov_hip_events** out);
int sov_hip_events_record(sov_hip_ctx* ctx, sov_hip_events* ev, int i);
int sov_hip_events_elapsed(sov_hip_events* ev, int i, int j, double* milliseconds);
void sov_hip_events_destroy(sov_hip_events* ev);

#ifdef __cplusplus
}
#endif
#endif /* SOV_HIP_H */
