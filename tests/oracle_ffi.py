"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, when built, the
reference shim (oracle/_ref/libref.so).  Test infrastructure: used only as
the checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class OrcLp(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int32), ("row_ptr", _i32p), ("col_idx", _i32p),
                ("val", _f64p), ("c", _f64p), ("col_lo", _f64p), ("col_hi", _f64p), ("row_lo", _f64p),
                ("row_hi", _f64p), ("maximize", C.c_int32), ("obj_const", C.c_double)]


class OrcOptions(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int64), ("restart_check_period", C.c_int64),
                ("restart_sufficient_decrease", C.c_double), ("step_size_safety", C.c_double),
                ("power_iterations", C.c_int32), ("termination", C.c_int32), ("use_primal_weight", C.c_int32),
                ("primal_weight", C.c_double), ("primal_weight_smoothing", C.c_double),
                ("ruiz_iterations", C.c_int32), ("apply_ruiz", C.c_int32), ("apply_pc", C.c_int32),
                ("warm_y", _f64p), ("restart_scheme", C.c_int32), ("beta_sufficient", C.c_double),
                ("beta_necessary", C.c_double), ("beta_artificial", C.c_double)]


class OrcResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("objective", C.c_double), ("primal_infeasibility", C.c_double),
                ("dual_infeasibility", C.c_double), ("duality_gap", C.c_double), ("iterations", C.c_int64),
                ("restarts", C.c_int64), ("operator_norm_estimate", C.c_double), ("rel_kkt", C.c_double * 3),
                ("final_primal_weight", C.c_double)]


STATUS = ["optimal", "iteration_limit", "time_limit", "numerical_failure", "not_solved", "cancelled",
          "primal_infeasible", "dual_infeasible"]

_lib = None
_ref = None


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def oracle() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        _lib = C.CDLL(ORACLE_SO)
        _lib.orc_dot.restype = C.c_double
        _lib.orc_operator_norm.restype = C.c_double
        _lib.orc_kkt_residual.restype = C.c_double
        _lib.orc_original_space_residual.restype = C.c_double
        _lib.orc_cs_violation.restype = C.c_double
        _lib.orc_farkas_certifies.restype = C.c_int32
        _lib.orc_primal_ray_certifies.restype = C.c_int32
        lp = C.POINTER(OrcLp)
        _lib.orc_spmv.argtypes = [C.c_int32, _i32p, _i32p, _f64p, C.c_double, _f64p, C.c_double, _f64p]
        _lib.orc_spmv_t.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_double, _f64p, C.c_double, _f64p]
        _lib.orc_dot.argtypes = [C.c_int64, _f64p, _f64p]
        _lib.orc_scaling.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_int32, C.c_double, C.c_int32,
                                     C.c_int32, C.c_double, _f64p, _f64p]
        _lib.orc_operator_norm.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_int32]
        _lib.orc_kkt_residual.argtypes = [lp, _f64p, _f64p, _f64p, C.c_double]
        _lib.orc_original_space_residual.argtypes = [lp, C.c_double, _f64p, _f64p, _f64p, _f64p, C.c_double]
        _lib.orc_relative_kkt.argtypes = [lp, _f64p, _f64p, _f64p]
        _lib.orc_cs_violation.argtypes = [lp, _f64p, _f64p, _f64p, C.c_double]
        _lib.orc_farkas_certifies.argtypes = [lp, _f64p, C.c_double]
        _lib.orc_primal_ray_certifies.argtypes = [lp, _f64p, C.c_double]
        _lib.orc_default_options.argtypes = [C.POINTER(OrcOptions)]
        _lib.orc_pdhg_solve.argtypes = [lp, C.POINTER(OrcOptions), _f64p, _f64p, _f64p, C.POINTER(OrcResult),
                                        _f64p, C.c_int64, _i64p]
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_pdhg_solve.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                        C.c_int32, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_double, _f64p,
                                        _f64p, _f64p, _f64p, C.c_int64, _f64p, _i64p, _i64p]
        _ref.ref_scaling.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_int32, C.c_int32, _f64p, _f64p]
        _ref.ref_pdhg_checkpoints.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                              _f64p, C.c_int32, C.c_double, C.c_double, C.c_int64, C.c_int64, _i64p,
                                              _f64p, _f64p, _f64p, _i64p]
        _ref.ref_pdhg_rate.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                       C.c_int32, C.c_double, C.c_int64, _f64p, _i64p, _i64p, _f64p]
        _ref.ref_parse.argtypes = [C.c_char_p, C.c_int32, _i32p, _i32p, _i32p]
        _ref.ref_parsed_get.argtypes = [_i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p, _f64p]
    return _ref


class Model:
    """Plain arrays of an LP, kept alive for the structs that point at them."""

    def __init__(self, d: dict):
        self.m, self.n = int(d["m"]), int(d["n"])
        self.row_ptr = np.ascontiguousarray(d["row_ptr"], np.int32)
        self.col_idx = np.ascontiguousarray(d["col_idx"], np.int32)
        self.val = np.ascontiguousarray(d["val"], np.float64)
        self.nnz = int(self.row_ptr[-1])
        self.col_idx = self.col_idx[: max(self.nnz, 1)] if self.nnz else np.zeros(1, np.int32)
        self.val = self.val[: max(self.nnz, 1)] if self.nnz else np.zeros(1)
        for k in ("c", "col_lo", "col_hi", "row_lo", "row_hi"):
            setattr(self, k, np.ascontiguousarray(d[k], np.float64).copy())
        self.maximize = bool(d.get("maximize", False))
        self.obj_const = float(d.get("obj_const", 0.0))

    def lp(self) -> OrcLp:
        s = OrcLp()
        s.m, s.n, s.nnz = self.m, self.n, self.nnz
        s.row_ptr, s.col_idx, s.val = _p(self.row_ptr, C.c_int32), _p(self.col_idx, C.c_int32), _p(self.val, C.c_double)
        s.c, s.col_lo, s.col_hi = _p(self.c, C.c_double), _p(self.col_lo, C.c_double), _p(self.col_hi, C.c_double)
        s.row_lo, s.row_hi = _p(self.row_lo, C.c_double), _p(self.row_hi, C.c_double)
        s.maximize, s.obj_const = int(self.maximize), self.obj_const
        return s

    def dense(self) -> np.ndarray:
        A = np.zeros((self.m, self.n))
        for r in range(self.m):
            for k in range(self.row_ptr[r], self.row_ptr[r + 1]):
                A[r, self.col_idx[k]] += self.val[k]
        return A


def spmv(M: Model, x, alpha=1.0, beta=0.0, y=None):
    y = np.zeros(M.m) if y is None else np.array(y, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    oracle().orc_spmv(M.m, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32), _p(M.val, C.c_double),
                      C.c_double(alpha), _p(x, C.c_double), C.c_double(beta), _p(y, C.c_double))
    return y


def spmv_t(M: Model, x, alpha=1.0, beta=0.0, y=None):
    y = np.zeros(M.n) if y is None else np.array(y, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    oracle().orc_spmv_t(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32), _p(M.val, C.c_double),
                        C.c_double(alpha), _p(x, C.c_double), C.c_double(beta), _p(y, C.c_double))
    return y


def dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return oracle().orc_dot(len(x), _p(x, C.c_double), _p(y, C.c_double))


def scaling(M: Model, ruiz=True, pc=True, ruiz_iters=8):
    rs, cs = np.zeros(max(M.m, 1)), np.zeros(max(M.n, 1))
    oracle().orc_scaling(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32), _p(M.val, C.c_double),
                         ruiz_iters, C.c_double(1e-2), int(ruiz), int(pc), C.c_double(1.0), _p(rs, C.c_double),
                         _p(cs, C.c_double))
    return rs[: M.m], cs[: M.n]


def operator_norm(M: Model, iterations=30):
    return oracle().orc_operator_norm(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32),
                                      _p(M.val, C.c_double), iterations)


def kkt_residual(M: Model, c_internal, x, y, tol):
    s = M.lp()
    c_internal = np.ascontiguousarray(c_internal, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return oracle().orc_kkt_residual(C.byref(s), _p(c_internal, C.c_double), _p(x, C.c_double), _p(y, C.c_double),
                                     C.c_double(tol))


def original_space_residual(M: Model, sign, rs, cs, xs, ys, tol):
    s = M.lp()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (rs, cs, xs, ys)]
    return oracle().orc_original_space_residual(C.byref(s), C.c_double(sign), *[_p(a, C.c_double) for a in arrs],
                                                C.c_double(tol))


def relative_kkt(M: Model, x, y):
    s = M.lp()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    out = np.zeros(5)
    oracle().orc_relative_kkt(C.byref(s), _p(x, C.c_double), _p(y, C.c_double), _p(out, C.c_double))
    return out


def cs_violation(M: Model, x, y, rc, tol):
    s = M.lp()
    x, y, rc = (np.ascontiguousarray(a, np.float64) for a in (x, y, rc))
    return oracle().orc_cs_violation(C.byref(s), _p(x, C.c_double), _p(y, C.c_double), _p(rc, C.c_double),
                                     C.c_double(tol))


def farkas_certifies(M: Model, y, slack=1e-6):
    s = M.lp()
    y = np.ascontiguousarray(y, np.float64)
    return bool(oracle().orc_farkas_certifies(C.byref(s), _p(y, C.c_double), C.c_double(slack)))


def primal_ray_certifies(M: Model, d, tol=1e-6):
    s = M.lp()
    d = np.ascontiguousarray(d, np.float64)
    return bool(oracle().orc_primal_ray_certifies(C.byref(s), _p(d, C.c_double), C.c_double(tol)))


def pdhg(M: Model, tolerance=1e-4, max_iterations=1000000, period=40, termination=0, primal_weight=False,
         omega=1.0, warm_y=None, trace_cap=0, restart_scheme=0):
    o = OrcOptions()
    oracle().orc_default_options(C.byref(o))
    o.tolerance, o.max_iterations, o.restart_check_period = tolerance, max_iterations, period
    o.termination, o.use_primal_weight, o.primal_weight = termination, int(primal_weight), omega
    o.restart_scheme = restart_scheme
    keep = None
    if warm_y is not None:
        keep = np.ascontiguousarray(warm_y, np.float64)
        o.warm_y = _p(keep, C.c_double)
    x, y, rc = np.zeros(max(M.n, 1)), np.zeros(max(M.m, 1)), np.zeros(max(M.n, 1))
    r = OrcResult()
    tr = np.zeros(max(trace_cap, 1))
    tl = C.c_int64(0)
    s = M.lp()
    oracle().orc_pdhg_solve(C.byref(s), C.byref(o), _p(x, C.c_double), _p(y, C.c_double), _p(rc, C.c_double),
                            C.byref(r), _p(tr, C.c_double) if trace_cap else None, trace_cap, C.byref(tl))
    del keep
    return dict(status=STATUS[r.status], objective=r.objective, primal_infeasibility=r.primal_infeasibility,
                dual_infeasibility=r.dual_infeasibility, duality_gap=r.duality_gap, iterations=int(r.iterations),
                restarts=int(r.restarts), operator_norm=r.operator_norm_estimate, rel_kkt=tuple(r.rel_kkt),
                primal_weight=r.final_primal_weight, x=x[: M.n], y=y[: M.m], rc=rc[: M.n],
                trace=tr[: min(int(tl.value), trace_cap)])


# ---- reference shim ------------------------------------------------------

def ref_solve(M: Model, tolerance=1e-4, max_iterations=1000000, period=40, time_limit=0.0, trace_cap=0):
    L = ref()
    x, y, rc = np.zeros(max(M.n, 1)), np.zeros(max(M.m, 1)), np.zeros(max(M.n, 1))
    stats = np.zeros(8)
    tr = np.zeros(max(trace_cap, 1))
    ti = np.zeros(max(trace_cap, 1), np.int64)
    tl = C.c_int64(0)
    st = L.ref_pdhg_solve(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32), _p(M.val, C.c_double),
                          _p(M.c, C.c_double), _p(M.col_lo, C.c_double), _p(M.col_hi, C.c_double),
                          _p(M.row_lo, C.c_double), _p(M.row_hi, C.c_double), int(M.maximize),
                          C.c_double(M.obj_const), C.c_double(tolerance), C.c_int64(max_iterations),
                          C.c_int64(period), C.c_double(time_limit), _p(x, C.c_double), _p(y, C.c_double),
                          _p(rc, C.c_double), _p(stats, C.c_double), C.c_int64(trace_cap),
                          _p(tr, C.c_double) if trace_cap else None, _p(ti, C.c_int64) if trace_cap else None,
                          C.byref(tl))
    if st < 0:
        raise RuntimeError("reference solve threw")
    nt = min(int(tl.value), trace_cap)
    return dict(status=STATUS[st], objective=stats[0], primal_infeasibility=stats[1], dual_infeasibility=stats[2],
                duality_gap=stats[3], iterations=int(stats[4]), restarts=int(stats[5]), seconds=stats[6],
                operator_norm=stats[7], x=x[: M.n], y=y[: M.m], rc=rc[: M.n], trace=tr[:nt], trace_it=ti[:nt])


def ref_checkpoints(M: Model, tolerance=1e-4, max_iterations=1000000, cap=16):
    """The first `cap` checkpoints of the reference's own solve: (status,
    total checkpoints, iterations, residuals, xs [k x n], ys [k x m])."""
    its, res = np.zeros(cap, np.int64), np.zeros(cap)
    xs, ys = np.zeros(cap * max(M.n, 1)), np.zeros(cap * max(M.m, 1))
    k = C.c_int64(0)
    st = ref().ref_pdhg_checkpoints(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32),
                                    _p(M.val, C.c_double), _p(M.c, C.c_double), _p(M.col_lo, C.c_double),
                                    _p(M.col_hi, C.c_double), _p(M.row_lo, C.c_double), _p(M.row_hi, C.c_double),
                                    int(M.maximize), C.c_double(M.obj_const), C.c_double(tolerance),
                                    C.c_int64(max_iterations), C.c_int64(cap), _p(its, C.c_int64),
                                    _p(res, C.c_double), _p(xs, C.c_double), _p(ys, C.c_double), C.byref(k))
    if st < 0:
        raise RuntimeError("reference solve threw")
    got = min(int(k.value), cap)
    return (STATUS[st], int(k.value), its[:got], res[:got], xs[: got * M.n].reshape(got, M.n),
            ys[: got * M.m].reshape(got, M.m))


def ref_rate(M: Model, checks=2):
    """The reference's own PDHG loop on the Host backend, cancelled after
    `checks` checkpoints: (checkpoint wall-clock stamps since the solve
    started, their iterations, Problem::Builder seconds, status)."""
    L = ref()
    stamps, its = np.zeros(checks), np.zeros(checks, np.int64)
    k, tb = C.c_int64(0), C.c_double(0.0)
    st = L.ref_pdhg_rate(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32), _p(M.val, C.c_double),
                         _p(M.c, C.c_double), _p(M.col_lo, C.c_double), _p(M.col_hi, C.c_double),
                         _p(M.row_lo, C.c_double), _p(M.row_hi, C.c_double), int(M.maximize),
                         C.c_double(M.obj_const), C.c_int64(checks), _p(stamps, C.c_double), _p(its, C.c_int64),
                         C.byref(k), C.byref(tb))
    if st < 0:
        raise RuntimeError("reference solve threw")
    n = int(k.value)
    return stamps[:n], its[:n], tb.value, STATUS[st]


def ref_scaling(M: Model, ruiz=True, pc=True):
    rs, cs = np.zeros(max(M.m, 1)), np.zeros(max(M.n, 1))
    ref().ref_scaling(M.m, M.n, _p(M.row_ptr, C.c_int32), _p(M.col_idx, C.c_int32), _p(M.val, C.c_double),
                      int(ruiz), int(pc), _p(rs, C.c_double), _p(cs, C.c_double))
    return rs[: M.m], cs[: M.n]


def ref_parse(text: str, mps=False) -> dict:
    L = ref()
    m, n, z = C.c_int32(), C.c_int32(), C.c_int32()
    if L.ref_parse(text.encode(), int(mps), C.byref(m), C.byref(n), C.byref(z)) != 0:
        raise RuntimeError("reference parser rejected the model")
    m, n, z = m.value, n.value, z.value
    d = dict(m=m, n=n, row_ptr=np.zeros(m + 1, np.int32), col_idx=np.zeros(max(z, 1), np.int32),
             val=np.zeros(max(z, 1)), c=np.zeros(max(n, 1)), col_lo=np.zeros(max(n, 1)), col_hi=np.zeros(max(n, 1)),
             row_lo=np.zeros(max(m, 1)), row_hi=np.zeros(max(m, 1)))
    mx, k = C.c_int32(), C.c_double()
    L.ref_parsed_get(_p(d["row_ptr"], C.c_int32), _p(d["col_idx"], C.c_int32), _p(d["val"], C.c_double),
                     _p(d["c"], C.c_double), _p(d["col_lo"], C.c_double), _p(d["col_hi"], C.c_double),
                     _p(d["row_lo"], C.c_double), _p(d["row_hi"], C.c_double), C.byref(mx), C.byref(k))
    for key, ln in (("col_idx", z), ("val", z), ("c", n), ("col_lo", n), ("col_hi", n), ("row_lo", m), ("row_hi", m)):
        d[key] = d[key][:ln]
    d["maximize"] = bool(mx.value)
    d["obj_const"] = k.value
    return d
