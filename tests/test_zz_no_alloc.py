"""The allocation-free check (sov_hip_pdhg_check_static, include/sov_hip.h)
against the allocating one (sov_hip_pdhg_check) through the C-ABI: same
state, same arguments, results identical to the last bit (same kernels, same
summation order), on two calls in a row of which the second has a smaller
input (the static scratch keeps the first call's contents, which the second
must not see), in FP64 and FP32, unsegmented and with column segments."""
import ctypes as C

import numpy as np
import pytest

from sovereign_amd.device import CheckArgs, CheckOut, Ctx, check
from test_gpu_steps import Device, random_point, scaled_problem

pytestmark = pytest.mark.gpu

FIELDS = ("kkt_cur", "kkt_avg", "orig_cur", "orig_avg", "pobj_cur", "dobj_cur", "pobj_avg", "dobj_avg")


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    yield c
    c.close()


def prepared(ctx, seed, m, n, K, prec):
    """A scaled LP on the device with a random point and its check arguments."""
    M, Ms, rs, cs = scaled_problem(seed, m=m, n=n)
    D = Device(ctx, Ms, K, prec)
    rng = np.random.default_rng(seed)
    since = 5
    x, y, xs, ys, s = random_point(rng, Ms, since)
    D.put(x=x, y=y, x_sum=xs, y_sum=ys, s=s)
    a = CheckArgs()
    a.st = C.pointer(D.st)
    a.since_restart = since
    a.tol_scaled, a.tol_orig = 1e-4, 1e-4
    a.obj_sign, a.obj_const = 1.0, M.obj_const
    pr, pc = D.pr, D.pc
    a.row_scale, a.col_scale = ctx.upload(rs[pr]), ctx.upload(cs[pc])
    a.orig_row_lo, a.orig_row_hi = ctx.upload(M.row_lo[pr]), ctx.upload(M.row_hi[pr])
    a.orig_col_lo, a.orig_col_hi, a.orig_c = ctx.upload(M.col_lo[pc]), ctx.upload(M.col_hi[pc]), ctx.upload(M.c[pc])
    a.act_avg = ctx.upload(np.zeros(M.m))
    return D, a, M.m


def both(ctx, D, a, m):
    """(allocating result, its activity, static result, its activity)."""
    out = CheckOut()
    check(ctx.L.sov_hip_pdhg_check(ctx.h, D.A.h, C.byref(a), C.byref(out)))
    act = ctx.download(a.act_avg, m)
    check(ctx.L.sov_hip_memcpy_h2d(ctx.h, a.act_avg, np.full(m, np.nan).ctypes.data, 8 * m))
    out_s = CheckOut()
    check(ctx.L.sov_hip_pdhg_check_static(ctx.h, D.A.h, C.byref(a), C.byref(out_s)))
    return out, act, out_s, ctx.download(a.act_avg, m)


def assert_same(out, act, out_s, act_s):
    for f in FIELDS:
        assert getattr(out_s, f) == getattr(out, f), (f, getattr(out_s, f), getattr(out, f))
    assert list(out_s.rel_cur) == list(out.rel_cur) and list(out_s.rel_avg) == list(out.rel_avg)
    np.testing.assert_array_equal(act_s, act)


@pytest.mark.parametrize("K,prec", [(1, 0), (3, 0), (1, 1), (2, 1)])
def test_static_check_equals_the_allocating_check(ctx, K, prec):
    big = prepared(ctx, 21, 3000, 7000, K, prec)
    small = prepared(ctx, 22, 200, 500, K, prec)
    print(f"K={K} prec={prec}: large then small", flush=True)
    r_big = both(ctx, *big)
    assert_same(*r_big)
    r_small = both(ctx, *small)
    assert_same(*r_small)
    assert np.isfinite(r_small[2].kkt_cur)
    for D, _, _ in (big, small):
        D.A.destroy()


def test_static_check_rejects_bad_arguments(ctx):
    D, a, m = prepared(ctx, 23, 200, 500, 1, 0)
    out = CheckOut()
    assert ctx.L.sov_hip_pdhg_check_static(ctx.h, D.A.h, None, C.byref(out)) == 3
    assert ctx.L.sov_hip_pdhg_check_static(None, D.A.h, C.byref(a), C.byref(out)) == 3
    D.A.destroy()
