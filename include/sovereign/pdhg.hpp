// PDHG / PDLP engine.  API mirrors include/sovereign/pdhg.hpp:74-172 of the
// reference: the same PdhgStatus/PdhgOptions/PdhgResult/Pdhg names, fields and
// defaults, so `sovereign-cli solve --engine pdhg`, race_solve and
// Crossover::solve take this engine unchanged.
//
// The iteration (reference pdhg.hpp:9-31) is Chambolle-Pock on the
// slack-augmented saddle point  min_{x in X, s in S} max_y c'x + y'(s - Ax):
//     x+ = Proj_X(x - tau (c - A'y))
//     s+ = Proj_S(s - tau y)
//     y+ = y + sigma ((2 s+ - s) - A (2 x+ - x))
// with Cesaro-average restarts checked every restart_check_period iterations.
//
// Two loops sit behind one solve():
//   * Primitive -- the reference's loop written against the virtual Backend
//     primitives; runs on any backend and is the parity mode.
//   * Fused     -- selected automatically on a Hip backend: per iteration one
//     kernel fuses A'y with the primal update/projection/extrapolation/Cesaro
//     sum and one fuses A x_bar with the slack and dual updates; scaling and
//     all KKT/restart reductions run on the GPU in FP64, and only a few
//     scalars cross the bus per check.
//
// Everything below the "additive" markers has no reference counterpart and
// defaults to the reference's behaviour.
#pragma once

#include <functional>
#include <memory>
#include <span>
#include <string>
#include <vector>

#include "sovereign/backend.hpp"
#include "sovereign/cancellation.hpp"
#include "sovereign/duals.hpp"
#include "sovereign/numeric.hpp"
#include "sovereign/problem.hpp"

namespace sov {

enum class PdhgStatus {
    Optimal,
    IterationLimit,
    TimeLimit,
    NumericalFailure,
    NotSolved,
    Cancelled,
    // additive
    PrimalInfeasible,   ///< a Farkas certificate in PdhgResult::certificate verifies
    DualInfeasible,     ///< a primal ray in PdhgResult::ray verifies (LP unbounded)
};

const char* to_string(PdhgStatus s) noexcept;

/// Additive: which loop solve() runs.
enum class PdhgEngine { Auto, Primitive, Fused };

/// Additive: the convergence test.  Reference = the original-space max-norm
/// residual of pdhg.cpp:145-199 (< tolerance).  RelativeKkt = max(r_p, r_d,
/// gap) <= tolerance with
///   r_p = ||row violation of Ax||_2 / (1 + ||finite row bounds||_2)
///   r_d = ||(c - A'y - lambda, y - proj(y))||_2 / (1 + ||c||_2)
///   gap = |p - d| / (1 + |p| + |d|)
/// all in original units, lambda/proj(y) being the projections onto the sign
/// cones the column/row bounds allow.
enum class PdhgTermination { Reference, RelativeKkt };

/// Additive: the restart rule.  Reference = pdhg.cpp:391-408 (restart to the
/// average when its scaled residual <= restart_sufficient_decrease x the
/// residual at the last restart).  Adaptive = the published PDLP rule on the
/// relative KKT K of the candidate (the better of average and current):
/// restart when K <= beta_sufficient K_last, or K <= beta_necessary K_last
/// and K grew since the previous check, or the run since the last restart is
/// >= beta_artificial of all iterations; the restart moves to the candidate.
enum class PdhgRestartScheme { Reference, Adaptive };

struct PdhgOptions {
    Real tolerance = 1e-4;
    long long max_iterations = 1000000;
    double time_limit_seconds = 0.0;
    long long restart_check_period = 40;
    Real restart_sufficient_decrease = 0.5;
    Real step_size_safety = 0.9;
    int power_iterations = 30;
    bool verbose = false;
    std::function<void(long long iteration, const std::vector<Real>& x_original,
                       const std::vector<Real>& y_original, Real original_residual)>
        checkpoint;

    // ---- additive --------------------------------------------------------
    PdhgEngine engine = PdhgEngine::Auto;
    PdhgTermination termination = PdhgTermination::Reference;
    /// F64: everything FP64.  F32: FP32 matrix values and iterates for the
    /// SpMVs and updates, FP64 for every reduction, restart and weight
    /// decision, then an FP64 polish (below).  Fused engine only.
    Precision precision = Precision::F64;
    /// FP32 runs switch to FP64 iterates (same matrix, FP64 values) once the
    /// candidate's relative KKT is below fp32_switch_factor * tolerance or
    /// its best value has not improved by 1% for fp32_stall_checks checks.
    /// (On the 1M-column Zipf LP, FP32 iterates stall near 1e-5 relative KKT;
    /// PDHG's own non-monotone phases last thousands of iterations, so the
    /// stall window is long.)
    bool fp64_polish = true;
    Real fp32_switch_factor = 20.0;
    int fp32_stall_checks = 250;
    /// FP32 runs with iterative refinement (fused engine, single LP): at the
    /// first restart whose candidate is below
    /// max(fp32_switch_factor * tolerance, refinement_start) relative KKT (or
    /// when FP32 iterates stall), the FP32 state turns into a correction
    /// around an FP64 anchor, re-anchored at every restart to the average;
    /// checks, restarts and the reported point are the combined FP64 point.
    /// PDHG on a correction takes the steps PDHG on the LP would take from
    /// the anchor, so no FP64 polish follows (fp64_polish applies to runs
    /// without refinement and to batches).
    bool fp32_refinement = true;
    Real refinement_start = 1e-3;
    Real refinement_max_scale = 1e8;   ///< the largest scale 1 / residual given to a correction
    /// PDLP primal weight omega: tau = eta/omega, sigma = eta*omega, updated at
    /// restarts from the restart-to-restart movement.  Off = the reference
    /// (omega = 1 throughout).
    bool use_primal_weight = false;
    Real initial_primal_weight = 1.0;
    Real primal_weight_smoothing = 0.5;
    PdhgRestartScheme restart_scheme = PdhgRestartScheme::Reference;
    Real restart_beta_sufficient = 0.2;
    Real restart_beta_necessary = 0.8;
    Real restart_beta_artificial = 0.36;
    /// Certificate checks at every check period (fused engine).
    bool detect_infeasibility = false;
    /// Optional warm start: original-space row duals (length num_rows).
    std::vector<Real> warm_start_dual;
    /// Optional warm start: original-space primal point (length num_cols);
    /// projected onto the column box, with the slack set from its activity.
    std::vector<Real> warm_start_primal;
    /// Capture each check interval as a HIP graph and replay it.
    bool use_graphs = true;
    /// Time SpMV lane counts per row-length class and column-segment counts
    /// before the loop and keep the fastest (fused engine).  The choice
    /// changes summation order, so runs with it on are not bit-reproducible
    /// across processes; off = the fixed default schedule.
    bool autotune = true;
    /// LDS-staged tiles for the step kernels (sov_hip_csr_set_tiles): -1 =
    /// the autotuner decides (off without autotune), 0 = never, k >= 1 =
    /// always, in k launches per product.
    int tiles = -1;
    /// When the autotuner runs (fused engine): at the first check after the
    /// iterations have taken this many seconds; 0 = before the first
    /// iteration.  Tuning costs about as much as a few thousand iterations
    /// on a large LP, so a short solve finishes on the default schedule
    /// instead of paying for it.
    double autotune_after_seconds = 0.5;
    /// Device ordinal for a fused solve that creates its own backend context.
    int device = 0;
    /// Measurement mark (fused engine): when >= 0, the loop waits for the
    /// device once this many iterations are done (splitting the interval that
    /// holds the mark into two plain launch sequences) and
    /// PdhgResult::timed_seconds / timed_iterations then cover the rest of
    /// the loop, checks and restarts included, up to the device's completion.
    long long timing_mark_iteration = -1;
    /// Fused engine: hand checkpoints over without stalling the device.  The
    /// candidate is snapshot on the device at its check and copied to pinned
    /// host memory while the next interval runs; the callback is called
    /// after that interval has been launched (one interval late, in check
    /// order, with the same arguments).  false = call it at the check.
    bool checkpoint_async = true;
    /// Batches (PdhgBatch, PdhgBatchSession): instances whose iterates fit
    /// this many bytes of LDS run every check interval in one workgroup
    /// launch, iterates resident in LDS (the same iterations, bit for bit);
    /// -1 = the device's LDS per workgroup, 0 = none.
    long long batch_stage_bytes = -1;
    /// Batches: a precision's steps run on the tiled layouts of the whole
    /// stacked matrix while the instances iterating in it hold at least this
    /// share of the entries, else on the CSR block lists; negative = never
    /// (the default: the stacked device order spreads an instance's rows over
    /// the length classes, so a row group's column slices hold few entries
    /// each and the tiles run slower than the block lists).
    double batch_tile_share = -1.0;
    /// Batches: the FP64 instances' step launches run on a stream of their
    /// own, beside the FP32 instances' (joined before each check).
    bool batch_split = true;
};

struct PdhgResult {
    PdhgStatus status = PdhgStatus::NotSolved;
    Real objective = 0.0;
    std::vector<Real> primal;
    std::vector<Real> dual;
    std::vector<Real> reduced_costs;
    Real primal_infeasibility = 0.0;
    Real dual_infeasibility = 0.0;
    Real duality_gap = 0.0;
    long long iterations = 0;
    long long restarts = 0;
    double seconds = 0.0;
    Real operator_norm_estimate = 0.0;
    std::string message;

    // ---- additive --------------------------------------------------------
    Real relative_primal_residual = 0.0;   ///< r_p of the reported point
    Real relative_dual_residual = 0.0;     ///< r_d
    Real relative_gap = 0.0;               ///< gap
    Real primal_weight = 1.0;              ///< final omega
    long long fp64_polish_iterations = 0;  ///< iterations run after an FP32->FP64 switch
    long long refinement_iterations = 0;   ///< iterations run as FP32 corrections (fp32_refinement)
    long long refinement_anchors = 0;      ///< anchors the corrections were taken around
    std::string engine;                    ///< "primitive" or "fused"
    FarkasCertificate certificate;         ///< set when PrimalInfeasible
    PrimalRay ray;                         ///< set when DualInfeasible
    double setup_seconds = 0.0;            ///< scaling, upload, operator norm
    double loop_seconds = 0.0;             ///< the iteration loop alone
    double timed_seconds = 0.0;            ///< from PdhgOptions::timing_mark_iteration to the loop's end
    long long timed_iterations = 0;        ///< iterations inside timed_seconds
};

class Pdhg {
public:
    explicit Pdhg(PdhgOptions options = {}) : options_(std::move(options)) {}

    PdhgResult solve(const Problem& problem, Backend& backend,
                     const CancellationToken* cancel = nullptr) const;

    const PdhgOptions& options() const noexcept { return options_; }

private:
    PdhgResult solve_primitive(const Problem& problem, Backend& backend,
                               const CancellationToken* cancel) const;
    PdhgResult solve_fused(const Problem& problem, Backend& backend,
                           const CancellationToken* cancel) const;
    PdhgOptions options_;
};

/// Additive (no reference counterpart): many independent LPs solved together
/// on a Hip backend.  The instances are stacked block-diagonally into one
/// device matrix, so each PDHG step is one launch for the whole batch; every
/// instance keeps its own scaling, step size, Cesaro average, restarts and
/// convergence test, and stops iterating (its rows are skipped) once it
/// converges.  Check intervals are captured as HIP graphs.  Uses from
/// PdhgOptions: tolerance, max_iterations, time_limit_seconds,
/// restart_check_period, step_size_safety, power_iterations, termination,
/// precision, fp64_polish, fp32_switch_factor, fp32_stall_checks,
/// restart_scheme and its betas, use_graphs, autotune.  The primal weight is
/// fixed per instance (1 unless given); all instances must share one
/// objective sense.
struct PdhgBatchResult {
    std::vector<PdhgResult> results;   ///< one per instance, in input order
    long long iterations = 0;          ///< iterations the batch ran
    double seconds = 0.0;
    double setup_seconds = 0.0;
    double loop_seconds = 0.0;
    std::string message;
};

class PdhgBatch {
public:
    explicit PdhgBatch(PdhgOptions options = {}) : options_(std::move(options)) {}
    /// warm_dual (optional): per instance, original-space row duals (empty =
    /// cold); primal_weight (optional): per instance, a fixed primal weight
    /// omega (tau = eta / omega, sigma = eta omega; default 1).
    PdhgBatchResult solve(const std::vector<const Problem*>& problems, Backend& backend,
                          const std::vector<std::vector<Real>>* warm_dual = nullptr,
                          const std::vector<Real>* primal_weight = nullptr) const;
    const PdhgOptions& options() const noexcept { return options_; }

private:
    PdhgOptions options_;
};

/// Additive (no reference counterpart): a live batch for re-solving related
/// LPs -- bounding the nodes of a branch-and-bound tree, parametric sweeps,
/// rolling horizons.  The instances are stacked once (matrix, scaling,
/// layouts and schedules, operator norms, iterate states all stay on the
/// device); between solves the caller overwrites an instance's row bounds
/// (right-hand sides), objective or column bounds in place, or starts an
/// instance from a given point or from another instance's (a neighbour's)
/// last solution, and solve() runs the batched engine again from the points
/// the instances hold.  Scaling depends on A only, so no update rescales;
/// the matrix itself cannot change.  By default an instance restarts from
/// its own last candidate.  Updates are validated when they are made (an
/// update that would make the instance invalid throws Error and changes
/// nothing).  The backend must outlive the session; one thread uses it.
class PdhgBatchSession {
public:
    /// Same arguments as PdhgBatch::solve; throws Error when the backend is
    /// not a Hip backend or the device work fails.  Instances that fail
    /// validation are not stacked: every solve reports them with the
    /// validation status, and updating them throws.
    PdhgBatchSession(PdhgOptions options, const std::vector<const Problem*>& problems, Backend& backend,
                     const std::vector<std::vector<Real>>* warm_dual = nullptr,
                     const std::vector<Real>* primal_weight = nullptr);
    ~PdhgBatchSession();
    PdhgBatchSession(const PdhgBatchSession&) = delete;
    PdhgBatchSession& operator=(const PdhgBatchSession&) = delete;

    std::size_t size() const noexcept;
    /// The instance as the session holds it (with every update applied).
    const Problem& problem(std::size_t b) const;
    void set_row_bounds(std::size_t b, std::span<const Real> lower, std::span<const Real> upper);
    void set_objective(std::size_t b, std::span<const Real> c);
    void set_column_bounds(std::size_t b, std::span<const Real> lower, std::span<const Real> upper);
    /// Start instance b's next solve from an original-space point (either
    /// vector may be empty: keep that part); x is projected onto the column
    /// bounds and the slacks follow from A x.
    void set_start(std::size_t b, std::span<const Real> x, std::span<const Real> y);
    /// Start instance b from instance `source`'s last solution (same
    /// dimensions; source must have been solved).
    void warm_start_from(std::size_t b, std::size_t source);
    /// Run the batched engine on every instance; results in input order.
    PdhgBatchResult solve();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace sov
