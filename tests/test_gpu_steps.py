"""Kernel-level parity of the fused PDHG kernels through the C-ABI: one primal
step, one dual step and one KKT/restart check against the oracle's SpMVs and
residual functions, unsegmented and with the column-segmented parts the
autotuner can choose (the running row sums across parts must not change a
result beyond summation-order rounding).

All state vectors cross the C-ABI in the handle's device order
(sov_hip_csr_order); the tests permute on the host.

Oracle functions: orc_spmv / orc_spmv_t (backend_host.cpp:133-172),
orc_kkt_residual (pdhg.cpp:87-136), orc_original_space_residual
(pdhg.cpp:145-199), orc_relative_kkt (relative KKT of the reported point)."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import Csr, Ctx, PdhgState, CheckArgs, CheckOut, check

pytestmark = pytest.mark.gpu

RT = 1e-12   # FP64: the device dots sum in another order than the oracle's


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    yield c
    c.close()


def scaled_problem(seed, m=3000, n=7000):
    """Original model, scale factors and the scaled model A_s = R A C."""
    lp = S.LP.random_zipf(m, n, seed)
    a = lp.arrays()
    M = O.Model(a)
    rs, cs = O.scaling(M)
    rows = np.repeat(np.arange(M.m), np.diff(M.row_ptr))
    vs = rs[rows] * M.val[: M.nnz] * cs[M.col_idx[: M.nnz]]
    d = dict(m=M.m, n=M.n, row_ptr=M.row_ptr, col_idx=M.col_idx[: M.nnz], val=vs, c=M.c * cs,
             col_lo=M.col_lo / cs, col_hi=M.col_hi / cs, row_lo=M.row_lo * rs, row_hi=M.row_hi * rs)
    return M, O.Model(d), rs, cs


def abs_model(Ms):
    """|A_s|, for the rounding bound of a dot product."""
    return O.Model(dict(m=Ms.m, n=Ms.n, row_ptr=Ms.row_ptr, col_idx=Ms.col_idx[: Ms.nnz], val=np.abs(Ms.val[: Ms.nnz]),
                        c=Ms.c, col_lo=Ms.col_lo, col_hi=Ms.col_hi, row_lo=Ms.row_lo, row_hi=Ms.row_hi))


def random_point(rng, Ms, since):
    """A point with some coordinates exactly on their bounds (exercises the
    bound classification), its running sums and a slack vector."""
    n, m = Ms.n, Ms.m
    lo, hi = Ms.col_lo, Ms.col_hi
    x = lo + rng.uniform(0, 1, n) * np.where(np.isfinite(hi - lo), hi - lo, 1.0)
    at = rng.uniform(size=n) < 0.3
    x[at] = lo[at]
    y = rng.standard_normal(m)
    xs = (x + 0.1 * rng.standard_normal(n)) * since
    ys = (y + 0.1 * rng.standard_normal(m)) * since
    s = Ms.row_lo.copy()
    return x, y, xs, ys, s


COLS = ("x", "x_bar", "x_sum", "c", "col_lo", "col_hi")


class Device:
    """The state of sov_hip_pdhg_state on the device, in the handle's device
    order; put/get take and return natural-order host vectors."""

    def __init__(self, ctx, Ms, K, prec=0):
        self.ctx, self.Ms, self.prec = ctx, Ms, prec
        self.dt = np.float64 if prec == 0 else np.float32
        self.A = Csr(ctx, Ms.m, Ms.n, Ms.row_ptr, Ms.col_idx, Ms.val)
        self.pr, self.pc = self.A.order(0), self.A.order(1)
        assert sorted(self.pr) == list(range(Ms.m)) and sorted(self.pc) == list(range(Ms.n))
        if prec == 1:
            check(ctx.L.sov_hip_csr_make_f32(ctx.h, self.A.h))
        for orient in (0, 1):
            check(ctx.L.sov_hip_csr_set_segments(ctx.h, self.A.h, orient, K))
        st = PdhgState()
        st.prec, st.m, st.n = prec, Ms.m, Ms.n
        for f, k in (("x", Ms.n), ("x_bar", Ms.n), ("x_sum", Ms.n), ("s", Ms.m), ("y", Ms.m), ("y_sum", Ms.m),
                     ("c", Ms.n), ("col_lo", Ms.n), ("col_hi", Ms.n), ("row_lo", Ms.m), ("row_hi", Ms.m)):
            setattr(st, f, ctx.upload(np.zeros(k, self.dt)))
        self.st = st
        self.put(c=Ms.c, col_lo=Ms.col_lo, col_hi=Ms.col_hi, row_lo=Ms.row_lo, row_hi=Ms.row_hi)

    def perm(self, f):
        return self.pc if f in COLS else self.pr

    def put(self, **kw):
        for f, v in kw.items():
            v = np.ascontiguousarray(np.asarray(v, self.dt)[self.perm(f)])
            check(self.ctx.L.sov_hip_memcpy_h2d(self.ctx.h, getattr(self.st, f), v.ctypes.data, v.nbytes))

    def get(self, f, count):
        d = self.ctx.download(getattr(self.st, f), count, self.dt).astype(np.float64)
        out = np.empty_like(d)
        out[self.perm(f)] = d
        return out


@pytest.mark.parametrize("K", [1, 3])
def test_primal_and_dual_steps_match_oracle(ctx, K):
    rng = np.random.default_rng(40 + K)
    M, Ms, rs, cs = scaled_problem(5)
    D = Device(ctx, Ms, K)
    Ma = abs_model(Ms)
    x, y, xs, ys, s = random_point(rng, Ms, 3)
    tau, sigma = 0.37, 0.81
    for first in (1, 0):
        D.put(x=x, y=y, x_sum=xs, y_sum=ys, s=s)
        check(ctx.L.sov_hip_pdhg_primal_step(ctx.h, D.A.h, C.byref(D.st), tau, first))
        g = Ms.c - O.spmv_t(Ms, y)
        xn = np.clip(x - tau * g, Ms.col_lo, Ms.col_hi)
        gx = D.get("x", Ms.n)
        assert np.all(np.abs(gx - xn) <= RT * (np.abs(x) + tau * (np.abs(Ms.c) + O.spmv_t(Ma, np.abs(y)))))
        np.testing.assert_array_equal(D.get("x_bar", Ms.n), 2 * gx - x)
        np.testing.assert_array_equal(D.get("x_sum", Ms.n), gx if first else xs + gx)
        # dual step on the new x_bar
        xbar = D.get("x_bar", Ms.n)
        check(ctx.L.sov_hip_pdhg_dual_step(ctx.h, D.A.h, C.byref(D.st), tau, sigma, first))
        sn = np.clip(s - tau * y, Ms.row_lo, Ms.row_hi)
        yn = y + sigma * (-O.spmv(Ms, xbar) + (2 * sn - s))
        gy = D.get("y", Ms.m)
        assert np.all(np.abs(gy - yn) <= RT * (np.abs(y) + sigma * (O.spmv(Ma, np.abs(xbar)) + 3 * np.abs(s))))
        np.testing.assert_array_equal(D.get("s", Ms.m), sn)
        np.testing.assert_array_equal(D.get("y_sum", Ms.m), gy if first else ys + gy)
    D.A.destroy()


@pytest.mark.parametrize("K,prec", [(1, 0), (3, 0), (1, 1), (4, 1)])
def test_check_matches_oracle_residuals(ctx, K, prec):
    rng = np.random.default_rng(7 + K + 10 * prec)
    M, Ms, rs, cs = scaled_problem(11)
    D = Device(ctx, Ms, K, prec)
    since = 7
    x, y, xs, ys, s = random_point(rng, Ms, since)
    if prec == 1:   # the oracle sees exactly the FP32 values the device holds
        x, y, xs, ys = (v.astype(np.float32).astype(np.float64) for v in (x, y, xs, ys))
        Ms32 = O.Model(dict(m=Ms.m, n=Ms.n, row_ptr=Ms.row_ptr, col_idx=Ms.col_idx[: Ms.nnz],
                            val=Ms.val[: Ms.nnz].astype(np.float32).astype(np.float64),
                            c=Ms.c.astype(np.float32).astype(np.float64),
                            col_lo=Ms.col_lo.astype(np.float32).astype(np.float64),
                            col_hi=Ms.col_hi.astype(np.float32).astype(np.float64),
                            row_lo=Ms.row_lo.astype(np.float32).astype(np.float64),
                            row_hi=Ms.row_hi.astype(np.float32).astype(np.float64)))
    else:
        Ms32 = Ms
    D.put(x=x, y=y, x_sum=xs, y_sum=ys, s=s)
    a = CheckArgs()
    a.st = C.pointer(D.st)
    a.since_restart = since
    a.tol_scaled, a.tol_orig = 1e-4, 1e-4
    a.obj_sign, a.obj_const = 1.0, M.obj_const
    pr, pc = D.pr, D.pc   # check arrays are in device order too
    a.row_scale, a.col_scale = ctx.upload(rs[pr]), ctx.upload(cs[pc])
    a.orig_row_lo, a.orig_row_hi = ctx.upload(M.row_lo[pr]), ctx.upload(M.row_hi[pr])
    a.orig_col_lo, a.orig_col_hi, a.orig_c = ctx.upload(M.col_lo[pc]), ctx.upload(M.col_hi[pc]), ctx.upload(M.c[pc])
    act = ctx.upload(np.zeros(M.m))
    a.act_avg = act
    out = CheckOut()
    check(ctx.L.sov_hip_pdhg_check(ctx.h, D.A.h, C.byref(a), C.byref(out)))
    inv = 1.0 / since
    xa, ya = xs * inv, ys * inv
    rt = RT if prec == 0 else 1e-9   # FP32 storage, FP64 arithmetic on the same values
    want = dict(kkt_cur=O.kkt_residual(Ms32, Ms32.c, x, y, 1e-4), kkt_avg=O.kkt_residual(Ms32, Ms32.c, xa, ya, 1e-4))
    for k, v in want.items():
        assert abs(getattr(out, k) - v) <= rt * max(1.0, abs(v)), (k, getattr(out, k), v)
    if prec == 0:
        oc = O.original_space_residual(M, 1.0, rs, cs, x, y, 1e-4)
        oa = O.original_space_residual(M, 1.0, rs, cs, xa, ya, 1e-4)
        assert abs(out.orig_cur - oc) <= RT * max(1.0, oc)
        assert abs(out.orig_avg - oa) <= RT * max(1.0, oa)
        for pt, rel, p, d in (((x, y), out.rel_cur, out.pobj_cur, out.dobj_cur),
                              ((xa, ya), out.rel_avg, out.pobj_avg, out.dobj_avg)):
            w = O.relative_kkt(M, pt[0] * cs, pt[1] * rs)
            np.testing.assert_allclose(list(rel), w[:3], rtol=1e-10, atol=1e-14)
            np.testing.assert_allclose([p, d], w[3:5], rtol=1e-11)
        act_nat = np.empty(M.m)
        act_nat[pr] = ctx.download(act, M.m)
        np.testing.assert_allclose(act_nat, O.spmv(Ms, xa), rtol=1e-11, atol=1e-11)
    D.A.destroy()
