"""Python binding of the MI355X PDHG engine (ctypes over the in-tree builds).

Loads ``lib/libsovereign.so`` (C++ engine, include/sovereign_c.h) and
``lib/libsov_hip.so`` (HIP kernels, include/sov_hip.h) from this package's
directory.  Nothing here computes: every solve and every kernel runs in the
native libraries, and a missing library raises instead of falling back.

Mirrors the reference's solve entry (apps/sovereign_cli.cpp:196-236 /
src/l1/pdhg.cpp:218): ``solve(lp, backend="hip", **options)`` returns the
PdhgResult fields plus the primal/dual/reduced-cost vectors.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")

STATUS = ["optimal", "iteration_limit", "time_limit", "numerical_failure", "not_solved",
          "cancelled", "primal_infeasible", "dual_infeasible"]

CHECKPOINT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.c_int32,
                            C.POINTER(C.c_double), C.c_int32, C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class PdhgOpts(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double), ("max_iterations", C.c_int64), ("time_limit_seconds", C.c_double),
        ("restart_check_period", C.c_int64), ("restart_sufficient_decrease", C.c_double),
        ("step_size_safety", C.c_double), ("power_iterations", C.c_int32), ("verbose", C.c_int32),
        ("engine", C.c_int32), ("termination", C.c_int32), ("precision", C.c_int32), ("fp64_polish", C.c_int32),
        ("fp32_switch_factor", C.c_double), ("fp32_stall_checks", C.c_int32), ("use_primal_weight", C.c_int32),
        ("initial_primal_weight", C.c_double), ("primal_weight_smoothing", C.c_double),
        ("detect_infeasibility", C.c_int32), ("use_graphs", C.c_int32), ("warm_start_dual", _f64p),
        ("trace_cap", C.c_int64), ("trace_res", _f64p), ("trace_it", _i64p),
        ("restart_scheme", C.c_int32), ("autotune", C.c_int32), ("restart_beta_sufficient", C.c_double),
        ("restart_beta_necessary", C.c_double), ("restart_beta_artificial", C.c_double),
        ("farkas_out", _f64p), ("ray_out", _f64p), ("warm_start_primal", _f64p), ("tiles", C.c_int32),
        ("autotune_after_seconds", C.c_double), ("timing_mark_iteration", C.c_int64),
        ("checkpoint_fn", CHECKPOINT_FN), ("checkpoint_user", C.c_void_p), ("checkpoint_async", C.c_int32),
        ("fp32_refinement", C.c_int32), ("refinement_start", C.c_double), ("refinement_max_scale", C.c_double),
        ("batch_stage_bytes", C.c_int64),
        ("batch_tile_share", C.c_double),
        ("batch_split", C.c_int32),
    ]


class PdhgRes(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("reserved", C.c_int32), ("objective", C.c_double),
        ("primal_infeasibility", C.c_double), ("dual_infeasibility", C.c_double), ("duality_gap", C.c_double),
        ("iterations", C.c_int64), ("restarts", C.c_int64), ("seconds", C.c_double), ("setup_seconds", C.c_double),
        ("loop_seconds", C.c_double), ("operator_norm_estimate", C.c_double), ("relative_primal", C.c_double),
        ("relative_dual", C.c_double), ("relative_gap", C.c_double), ("primal_weight", C.c_double),
        ("fp64_polish_iterations", C.c_int64), ("trace_len", C.c_int64), ("timed_seconds", C.c_double),
        ("timed_iterations", C.c_int64), ("refinement_iterations", C.c_int64), ("refinement_anchors", C.c_int64),
    ]


_engine = None
_hip = None


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def engine() -> C.CDLL:
    """libsovereign.so (raises if it was not built)."""
    global _engine
    if _engine is None:
        p = lib_path("libsovereign.so")
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: run __graft_entry__.build() first")
        _engine = C.CDLL(p)
        L = _engine
        L.sov_c_last_error.restype = C.c_char_p
        L.sov_lp_from_csr.restype = C.c_void_p
        L.sov_lp_from_csr.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                      _f64p, _f64p, C.c_int32, C.c_double]
        L.sov_lp_from_text.restype = C.c_void_p
        L.sov_lp_from_text.argtypes = [C.c_char_p, C.c_int32]
        L.sov_lp_read_file.restype = C.c_void_p
        L.sov_lp_read_file.argtypes = [C.c_char_p]
        L.sov_lp_destroy.argtypes = [C.c_void_p]
        L.sov_lp_dims.argtypes = [C.c_void_p, _i32p, _i32p, _i32p]
        L.sov_lp_arrays.argtypes = [C.c_void_p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p,
                                    _f64p]
        L.sov_gen_random_lp.restype = C.c_void_p
        L.sov_gen_random_lp.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint64]
        L.sov_gen_random_lp_zipf.restype = C.c_void_p
        L.sov_gen_random_lp_zipf.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_uint64]
        L.sov_gen_transportation.restype = C.c_void_p
        L.sov_gen_transportation.argtypes = [C.c_int32, C.c_int32, C.c_uint64]
        L.sov_gen_batch_sizes.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _i32p]
        L.sov_scaling_factors.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, _f64p]
        L.sov_pdhg_default_opts.argtypes = [C.POINTER(PdhgOpts)]
        L.sov_pdhg_solve.argtypes = [C.c_void_p, C.c_int32, C.POINTER(PdhgOpts), _f64p, _f64p, _f64p,
                                     C.POINTER(PdhgRes), C.c_char_p, C.c_int32]
        L.sov_backend_report.argtypes = [C.c_char_p, C.c_int32]
    return _engine


def hip() -> C.CDLL:
    """libsov_hip.so, the C-ABI of include/sov_hip.h (raises if not built)."""
    global _hip
    if _hip is None:
        p = lib_path("libsov_hip.so")
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: run __graft_entry__.build() first")
        _hip = C.CDLL(p, mode=C.RTLD_GLOBAL)
        _hip.sov_hip_last_error.restype = C.c_char_p
        _hip.sov_hip_ctx_device_name.restype = C.c_char_p
    return _hip


def ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def last_error() -> str:
    return engine().sov_c_last_error().decode()


def backend_report() -> str:
    buf = C.create_string_buffer(512)
    engine().sov_backend_report(buf, 512)
    return buf.value.decode()


def hip_device_count() -> int:
    n = C.c_int(0)
    hip().sov_hip_device_count(C.byref(n))
    return n.value


class LP:
    """An immutable model held by the native library (sov::Problem)."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("model construction failed: " + last_error())
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and engine() is not None:
            engine().sov_lp_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_csr(cls, m, n, row_ptr, col_idx, val, c, col_lo, col_hi, row_lo, row_hi, maximize=False,
                 obj_const=0.0):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col_idx, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (c, col_lo, col_hi, row_lo, row_hi)]
        h = engine().sov_lp_from_csr(m, n, len(v), ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double),
                                     *[ptr(a, C.c_double) for a in arrs], int(bool(maximize)), float(obj_const))
        return cls(h)

    @classmethod
    def from_text(cls, text: str, mps: bool = False):
        return cls(engine().sov_lp_from_text(text.encode(), 1 if mps else 0))

    @classmethod
    def read(cls, path: str):
        return cls(engine().sov_lp_read_file(path.encode()))

    @classmethod
    def random(cls, m, n, k, seed):
        return cls(engine().sov_gen_random_lp(m, n, k, seed))

    @classmethod
    def random_zipf(cls, m, n, seed, a=1.5, kmin=2, kmax=200):
        return cls(engine().sov_gen_random_lp_zipf(m, n, a, kmin, kmax, seed))

    @classmethod
    def transportation(cls, sources, sinks, seed):
        return cls(engine().sov_gen_transportation(sources, sinks, seed))

    def write_mps(self, path: str):
        L = engine()
        L.sov_lp_write_mps.argtypes = [C.c_void_p, C.c_char_p]
        if L.sov_lp_write_mps(self._h, path.encode()) != 0:
            raise RuntimeError("write_mps failed: " + last_error())

    def dims(self):
        m, n, z = C.c_int32(), C.c_int32(), C.c_int32()
        engine().sov_lp_dims(self._h, C.byref(m), C.byref(n), C.byref(z))
        return m.value, n.value, z.value

    def arrays(self) -> dict:
        m, n, nnz = self.dims()
        out = dict(m=m, n=n, nnz=nnz, row_ptr=np.zeros(m + 1, np.int32), col_idx=np.zeros(max(nnz, 1), np.int32),
                   val=np.zeros(max(nnz, 1)), c=np.zeros(max(n, 1)), col_lo=np.zeros(max(n, 1)),
                   col_hi=np.zeros(max(n, 1)), row_lo=np.zeros(max(m, 1)), row_hi=np.zeros(max(m, 1)))
        mx, k = C.c_int32(), C.c_double()
        engine().sov_lp_arrays(self._h, ptr(out["row_ptr"], C.c_int32), ptr(out["col_idx"], C.c_int32),
                               ptr(out["val"], C.c_double), ptr(out["c"], C.c_double), ptr(out["col_lo"], C.c_double),
                               ptr(out["col_hi"], C.c_double), ptr(out["row_lo"], C.c_double),
                               ptr(out["row_hi"], C.c_double), C.byref(mx), C.byref(k))
        for key, ln in (("col_idx", nnz), ("val", nnz), ("c", n), ("col_lo", n), ("col_hi", n), ("row_lo", m),
                        ("row_hi", m)):
            out[key] = out[key][:ln]
        out["maximize"] = bool(mx.value)
        out["obj_const"] = k.value
        return out

    def scaling(self, ruiz=True, pc=True):
        m, n, _ = self.dims()
        rs, cs = np.zeros(max(m, 1)), np.zeros(max(n, 1))
        engine().sov_scaling_factors(self._h, int(ruiz), int(pc), ptr(rs, C.c_double), ptr(cs, C.c_double))
        return rs[:m], cs[:n]


class Family:
    """The parametric transportation family (sov_family_* of sovereign_c.h)."""

    def __init__(self, sources=100, sinks=200, params=16, seed=0):
        L = engine()
        L.sov_family_make.restype = C.c_void_p
        L.sov_family_make.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint64]
        L.sov_family_destroy.argtypes = [C.c_void_p]
        L.sov_family_instance.restype = C.c_void_p
        L.sov_family_instance.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, _f64p]
        self.S, self.D, self.P = sources, sinks, params
        self.h = L.sov_family_make(sources, sinks, params, seed)
        if not self.h:
            raise RuntimeError("family: " + last_error())

    @property
    def features(self):
        return self.S + self.D + self.P

    def instance(self, seed, index):
        """(LP, features) of instance `index` of draw stream `seed`."""
        f = np.zeros(self.features)
        h = engine().sov_family_instance(self.h, seed, index, ptr(f, C.c_double))
        if not h:
            raise RuntimeError("family instance: " + last_error())
        return LP(h), f

    def __del__(self):
        try:
            engine().sov_family_destroy(self.h)
        except Exception:
            pass


def batch_sizes(count, lo, hi, seed):
    out = np.zeros(count, np.int32)
    engine().sov_gen_batch_sizes(count, lo, hi, seed, ptr(out, C.c_int32))
    return out


@dataclass
class Result:
    status: str
    objective: float
    primal_infeasibility: float
    dual_infeasibility: float
    duality_gap: float
    iterations: int
    restarts: int
    seconds: float
    setup_seconds: float
    loop_seconds: float
    operator_norm_estimate: float
    relative_kkt: tuple
    primal_weight: float
    fp64_polish_iterations: int
    x: np.ndarray
    y: np.ndarray
    rc: np.ndarray
    message: str = ""
    trace_res: np.ndarray = field(default_factory=lambda: np.zeros(0))
    trace_it: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    farkas: np.ndarray = None     # row multipliers when status == "primal_infeasible"
    ray: np.ndarray = None        # primal ray when status == "dual_infeasible"
    timed_seconds: float = 0.0    # from the timing mark to the end of the loop
    timed_iterations: int = 0
    refinement_iterations: int = 0   # iterations run as FP32 corrections around an FP64 anchor
    refinement_anchors: int = 0


_ENGINES = {"auto": 0, "primitive": 1, "fused": 2}


def solve(lp: LP, backend: str = "hip", tolerance=1e-4, max_iterations=1000000, time_limit=0.0,
          restart_check_period=40, engine_kind="auto", termination="reference", precision="f64",
          fp64_polish=True, primal_weight=False, initial_primal_weight=1.0, warm_start_dual=None,
          use_graphs=True, trace_cap=0, verbose=False, fp32_switch_factor=None, fp32_stall_checks=None,
          restarts="reference", autotune=True, detect_infeasibility=False, primal_weight_smoothing=None,
          warm_start_primal=None, tiles=None, autotune_after=None, timing_mark=None, checkpoint=None,
          checkpoint_async=True, refinement=None, refinement_start=None) -> Result:
    """Pdhg(options).solve(problem, backend) through the C entry point.
    tiles: None = the autotuner decides, 0 = CSR step kernels only, k >= 1 =
    the LDS-staged tiled step kernels in k launches per product.
    autotune_after: seconds of iterations before the autotuner runs (None:
    the engine's default; 0: before the first iteration).
    timing_mark: the iteration from which Result.timed_seconds measures the
    loop (None: no timed part).
    checkpoint: None, or f(iteration, x, y, residual) called at every check
    with copies of the original-space candidate (PdhgOptions::checkpoint);
    checkpoint_async: hand checkpoints over without stalling the device.
    refinement: FP32 iterates become corrections around an FP64 anchor
    (PdhgOptions::fp32_refinement; None: the engine's default, on);
    refinement_start: the relative KKT at which they do."""
    L = engine()
    o = PdhgOpts()
    L.sov_pdhg_default_opts(C.byref(o))
    o.tolerance = tolerance
    o.max_iterations = max_iterations
    o.time_limit_seconds = time_limit
    o.restart_check_period = restart_check_period
    o.engine = _ENGINES[engine_kind]
    o.termination = 1 if termination == "relative" else 0
    o.precision = 1 if precision == "f32" else 0
    o.fp64_polish = int(fp64_polish)
    o.use_primal_weight = int(primal_weight)
    o.initial_primal_weight = initial_primal_weight
    o.use_graphs = int(use_graphs)
    o.verbose = int(verbose)
    o.restart_scheme = 1 if restarts == "adaptive" else 0
    o.autotune = int(autotune)
    if tiles is not None:
        o.tiles = int(tiles)
    if autotune_after is not None:
        o.autotune_after_seconds = float(autotune_after)
    if timing_mark is not None:
        o.timing_mark_iteration = int(timing_mark)
    o.checkpoint_async = int(bool(checkpoint_async))
    callback = None   # kept alive for the whole native call
    if checkpoint is not None:
        def _fn(user, it, xp, nx, yp, ny, res):
            checkpoint(int(it), np.ctypeslib.as_array(xp, (nx,)).copy() if nx else np.zeros(0),
                       np.ctypeslib.as_array(yp, (ny,)).copy() if ny else np.zeros(0), float(res))
        callback = CHECKPOINT_FN(_fn)
        o.checkpoint_fn = callback
    o.detect_infeasibility = int(detect_infeasibility)
    if primal_weight_smoothing is not None:
        o.primal_weight_smoothing = primal_weight_smoothing
    if fp32_switch_factor is not None:
        o.fp32_switch_factor = fp32_switch_factor
    if fp32_stall_checks is not None:
        o.fp32_stall_checks = fp32_stall_checks
    if refinement is not None:
        o.fp32_refinement = int(bool(refinement))
    if refinement_start is not None:
        o.refinement_start = float(refinement_start)
    m, n, _ = lp.dims()
    keep = []
    if warm_start_dual is not None:
        w = np.ascontiguousarray(warm_start_dual, np.float64)
        assert w.shape == (m,)
        keep.append(w)
        o.warm_start_dual = ptr(w, C.c_double)
    if warm_start_primal is not None:
        wp = np.ascontiguousarray(warm_start_primal, np.float64)
        assert wp.shape == (n,)
        keep.append(wp)
        o.warm_start_primal = ptr(wp, C.c_double)
    tr = np.zeros(max(trace_cap, 1))
    ti = np.zeros(max(trace_cap, 1), np.int64)
    if trace_cap > 0:
        o.trace_cap = trace_cap
        o.trace_res = ptr(tr, C.c_double)
        o.trace_it = ptr(ti, C.c_int64)
    x, y, rc = np.zeros(max(n, 1)), np.zeros(max(m, 1)), np.zeros(max(n, 1))
    farkas, ray = np.zeros(max(m, 1)), np.zeros(max(n, 1))
    o.farkas_out, o.ray_out = ptr(farkas, C.c_double), ptr(ray, C.c_double)
    r = PdhgRes()
    msg = C.create_string_buffer(1024)
    rc_code = L.sov_pdhg_solve(lp.handle, 2 if backend == "hip" else 0, C.byref(o), ptr(x, C.c_double),
                               ptr(y, C.c_double), ptr(rc, C.c_double), C.byref(r), msg, 1024)
    if rc_code != 0:
        raise RuntimeError(f"sov_pdhg_solve failed ({rc_code}): {last_error()}")
    nt = min(int(r.trace_len), trace_cap)
    return Result(STATUS[r.status], r.objective, r.primal_infeasibility, r.dual_infeasibility, r.duality_gap,
                  int(r.iterations), int(r.restarts), r.seconds, r.setup_seconds, r.loop_seconds,
                  r.operator_norm_estimate, (r.relative_primal, r.relative_dual, r.relative_gap), r.primal_weight,
                  int(r.fp64_polish_iterations), x[:n], y[:m], rc[:n], msg.value.decode(), tr[:nt], ti[:nt],
                  farkas[:m] if STATUS[r.status] == "primal_infeasible" else None,
                  ray[:n] if STATUS[r.status] == "dual_infeasible" else None,
                  r.timed_seconds, int(r.timed_iterations), int(r.refinement_iterations),
                  int(r.refinement_anchors))


def solve_batch(lps, tolerance=1e-6, max_iterations=1000000, time_limit=0.0, restart_check_period=40,
                termination="relative", precision="f32", fp64_polish=True, use_graphs=True, autotune=True,
                restarts="adaptive", fp32_switch_factor=None, fp32_stall_checks=None, verbose=False,
                warm_duals=None, primal_weights=None, stage_bytes=None, tile_share=None, split=None):
    """PdhgBatch(options).solve(problems, hip backend) through
    sov_pdhg_solve_batch: every model solved in one stacked device batch.
    Returns (list of Result, dict of batch timings).

    stage_bytes: LDS a resident instance may take (None: the device's 160
    KiB, 0: none; resident instances run whole check intervals in one
    workgroup while every instance of their precision fits).  tile_share:
    the share of the entries from which a precision's steps run on the tiled
    layouts (None: never).  split: the FP64 instances' steps on a stream of
    their own (None: on)."""
    L = engine()
    o = PdhgOpts()
    L.sov_pdhg_default_opts(C.byref(o))
    o.tolerance = tolerance
    o.max_iterations = max_iterations
    o.time_limit_seconds = time_limit
    o.restart_check_period = restart_check_period
    o.termination = 1 if termination == "relative" else 0
    o.precision = 1 if precision == "f32" else 0
    o.fp64_polish = int(fp64_polish)
    o.use_graphs = int(use_graphs)
    o.autotune = int(autotune)
    o.restart_scheme = 1 if restarts == "adaptive" else 0
    o.verbose = int(verbose)
    if fp32_switch_factor is not None:
        o.fp32_switch_factor = fp32_switch_factor
    if fp32_stall_checks is not None:
        o.fp32_stall_checks = fp32_stall_checks
    if stage_bytes is not None:   # LDS a resident instance may take (0: every instance on the step launches)
        o.batch_stage_bytes = int(stage_bytes)
    if tile_share is not None:    # share of entries from which a precision's steps run tiled (negative: never)
        o.batch_tile_share = float(tile_share)
    if split is not None:         # FP64 step launches on a stream of their own
        o.batch_split = int(bool(split))
    count = len(lps)
    handles = (C.c_void_p * count)(*[lp.handle for lp in lps])
    res = (PdhgRes * count)()
    dims = [lp.dims() for lp in lps]
    xs = [np.zeros(max(d[1], 1)) for d in dims]
    ys = [np.zeros(max(d[0], 1)) for d in dims]
    xp = (_f64p * count)(*[ptr(x, C.c_double) for x in xs])
    yp = (_f64p * count)(*[ptr(y, C.c_double) for y in ys])
    timing = np.zeros(4)
    keep = []
    wyp = None
    if warm_duals is not None:
        arrs = [None if w is None else np.ascontiguousarray(w, np.float64) for w in warm_duals]
        keep += [a for a in arrs if a is not None]
        wyp = (_f64p * count)(*[ptr(a, C.c_double) if a is not None else _f64p() for a in arrs])
    pwp = None
    if primal_weights is not None:
        pwa = np.ascontiguousarray(primal_weights, np.float64)
        keep.append(pwa)
        pwp = ptr(pwa, C.c_double)
    L.sov_pdhg_solve_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(PdhgOpts), C.POINTER(_f64p),
                                       _f64p, C.POINTER(PdhgRes), C.POINTER(_f64p), C.POINTER(_f64p), _f64p]
    rc_code = L.sov_pdhg_solve_batch(handles, count, C.byref(o), wyp, pwp, res, xp, yp, ptr(timing, C.c_double))
    if rc_code != 0:
        raise RuntimeError(f"sov_pdhg_solve_batch failed ({rc_code}): {last_error()}")
    out = []
    for b in range(count):
        r = res[b]
        m, n, _ = dims[b]
        out.append(Result(STATUS[r.status], r.objective, r.primal_infeasibility, r.dual_infeasibility,
                          r.duality_gap, int(r.iterations), int(r.restarts), r.seconds, r.setup_seconds,
                          r.loop_seconds, r.operator_norm_estimate, (r.relative_primal, r.relative_dual,
                                                                     r.relative_gap), r.primal_weight,
                          int(r.fp64_polish_iterations), xs[b][:n], ys[b][:m], np.zeros(0)))
    return out, {"seconds": timing[0], "setup_seconds": timing[1], "loop_seconds": timing[2],
                 "iterations": int(timing[3])}


def _batch_opts(tolerance, max_iterations, time_limit, restart_check_period, termination, precision, fp64_polish,
                use_graphs, autotune, restarts, verbose):
    L = engine()
    o = PdhgOpts()
    L.sov_pdhg_default_opts(C.byref(o))
    o.tolerance = tolerance
    o.max_iterations = max_iterations
    o.time_limit_seconds = time_limit
    o.restart_check_period = restart_check_period
    o.termination = 1 if termination == "relative" else 0
    o.precision = 1 if precision == "f32" else 0
    o.fp64_polish = int(fp64_polish)
    o.use_graphs = int(use_graphs)
    o.autotune = int(autotune)
    o.restart_scheme = 1 if restarts == "adaptive" else 0
    o.verbose = int(verbose)
    return o


class BatchSession:
    """A live batch for re-solving related LPs (sov::PdhgBatchSession through
    sov_batch_session_* of sovereign_c.h): the models stay on the device;
    between solve() calls an instance's row bounds, objective or column
    bounds are overwritten in place, or its next start is set from a point
    or from another instance's last solution.  Every solve starts each
    instance from its last candidate unless told otherwise."""

    def __init__(self, lps, tolerance=1e-6, max_iterations=1000000, time_limit=0.0, restart_check_period=40,
                 termination="relative", precision="f32", fp64_polish=True, use_graphs=True, autotune=True,
                 restarts="adaptive", verbose=False, warm_duals=None, primal_weights=None):
        L = engine()
        L.sov_batch_session_create.restype = C.c_void_p
        L.sov_batch_session_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(PdhgOpts),
                                               C.POINTER(_f64p), _f64p]
        L.sov_batch_session_destroy.argtypes = [C.c_void_p]
        for f in ("sov_batch_session_set_row_bounds", "sov_batch_session_set_col_bounds",
                  "sov_batch_session_set_start"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p]
        L.sov_batch_session_set_objective.argtypes = [C.c_void_p, C.c_int32, _f64p]
        L.sov_batch_session_warm_start_from.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.sov_batch_session_solve.argtypes = [C.c_void_p, C.POINTER(PdhgRes), C.POINTER(_f64p), C.POINTER(_f64p),
                                              _f64p]
        self.lps = list(lps)
        self.dims = [lp.dims() for lp in self.lps]
        o = _batch_opts(tolerance, max_iterations, time_limit, restart_check_period, termination, precision,
                        fp64_polish, use_graphs, autotune, restarts, verbose)
        count = len(self.lps)
        handles = (C.c_void_p * count)(*[lp.handle for lp in self.lps])
        keep, wyp, pwp = [], None, None
        if warm_duals is not None:
            arrs = [None if w is None else np.ascontiguousarray(w, np.float64) for w in warm_duals]
            keep += [a for a in arrs if a is not None]
            wyp = (_f64p * count)(*[ptr(a, C.c_double) if a is not None else _f64p() for a in arrs])
        if primal_weights is not None:
            pwa = np.ascontiguousarray(primal_weights, np.float64)
            keep.append(pwa)
            pwp = ptr(pwa, C.c_double)
        h = L.sov_batch_session_create(handles, count, C.byref(o), wyp, pwp)
        if not h:
            raise RuntimeError("sov_batch_session_create failed: " + last_error())
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            engine().sov_batch_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        rc = getattr(engine(), name)(self.h, *args)
        if rc != 0:
            raise RuntimeError(f"{name} failed ({rc}): {last_error()}")

    @staticmethod
    def _p(a):
        return None if a is None else ptr(np.ascontiguousarray(a, np.float64), C.c_double)

    def set_row_bounds(self, b, lower, upper):
        lo, hi = np.ascontiguousarray(lower, np.float64), np.ascontiguousarray(upper, np.float64)
        assert lo.shape == hi.shape == (self.dims[b][0],)
        self._call("sov_batch_session_set_row_bounds", b, ptr(lo, C.c_double), ptr(hi, C.c_double))

    def set_col_bounds(self, b, lower, upper):
        lo, hi = np.ascontiguousarray(lower, np.float64), np.ascontiguousarray(upper, np.float64)
        assert lo.shape == hi.shape == (self.dims[b][1],)
        self._call("sov_batch_session_set_col_bounds", b, ptr(lo, C.c_double), ptr(hi, C.c_double))

    def set_objective(self, b, c):
        cc = np.ascontiguousarray(c, np.float64)
        assert cc.shape == (self.dims[b][1],)
        self._call("sov_batch_session_set_objective", b, ptr(cc, C.c_double))

    def set_start(self, b, x=None, y=None):
        xa = None if x is None else np.ascontiguousarray(x, np.float64)
        ya = None if y is None else np.ascontiguousarray(y, np.float64)
        self._call("sov_batch_session_set_start", b, None if xa is None else ptr(xa, C.c_double),
                   None if ya is None else ptr(ya, C.c_double))

    def warm_start_from(self, b, source):
        self._call("sov_batch_session_warm_start_from", b, source)

    def solve(self):
        """(list of Result, dict of batch timings), as solve_batch."""
        count = len(self.lps)
        res = (PdhgRes * count)()
        xs = [np.zeros(max(d[1], 1)) for d in self.dims]
        ys = [np.zeros(max(d[0], 1)) for d in self.dims]
        xp = (_f64p * count)(*[ptr(x, C.c_double) for x in xs])
        yp = (_f64p * count)(*[ptr(y, C.c_double) for y in ys])
        timing = np.zeros(4)
        self._call("sov_batch_session_solve", res, xp, yp, ptr(timing, C.c_double))
        out = []
        for b in range(count):
            r = res[b]
            m, n, _ = self.dims[b]
            out.append(Result(STATUS[r.status], r.objective, r.primal_infeasibility, r.dual_infeasibility,
                              r.duality_gap, int(r.iterations), int(r.restarts), r.seconds, r.setup_seconds,
                              r.loop_seconds, r.operator_norm_estimate, (r.relative_primal, r.relative_dual,
                                                                         r.relative_gap), r.primal_weight,
                              int(r.fp64_polish_iterations), xs[b][:n], ys[b][:m], np.zeros(0)))
        return out, {"seconds": timing[0], "setup_seconds": timing[1], "loop_seconds": timing[2],
                     "iterations": int(timing[3])}
