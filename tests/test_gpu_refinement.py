"""FP32 iterative refinement (PdhgOptions::fp32_refinement): FP32 iterates
become corrections (dx, ds, dy) of an FP64 anchor (x0, s0 = A x0, y0) at a
scale delta,

    x = x0 + dx / delta,  s = s0 + ds / delta,  y = y0 + dy / delta,

and the correction problem has objective delta (c - A^T y0), column box
delta (X - x0), row box delta (S - s0) and slack objective delta y0.  The
reference has no counterpart (its PDHG is FP64 throughout,
src/l1/pdhg.cpp:218-502); what fixes the method is the algebra itself:

  * kernel level, through the C-ABI: sov_hip_pdhg_anchor writes exactly that
    problem and start (bit for bit against numpy), and k steps of PDHG on the
    correction, combined by sov_hip_pdhg_combine, are k steps of PDHG on the
    scaled LP from the anchor (FP64 corrections: to rounding; FP32
    corrections: to FP32 rounding of the distance moved), on the CSR and on
    the tiled step kernels.  A flipped sign in the slack objective, a box
    taken around the wrong point or a dropped 1/delta fails these;
  * the check of a correction with its anchor (the products on the FP32
    correction, the anchor's terms in FP64) against the oracle's residuals
    of the combined point;
  * solve level: FP32 + refinement to 1e-6 relative KKT on LPs with exact
    optima (HiGHS through scipy, tests/golden/exact_optima.json), with the
    objective within 1e-6, the host oracle's relative KKT of the returned
    point within the tolerance, and every iteration in FP32 (no FP64
    polish).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import CheckArgs, CheckOut, Ctx, PdhgState, check
from test_gpu_steps import Device, scaled_problem
from test_gpu_tiles import set_tiles

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def exact():
    with open(os.path.join(HERE, "golden", "exact_optima.json")) as f:
        return json.load(f)["instances"]


def dev_vec(ctx, D, v, field):
    """An FP64 device vector in D's device order for a natural-order host
    vector `v` (the order of state field `field`)."""
    return ctx.upload(np.ascontiguousarray(np.asarray(v, np.float64)[D.perm(field)]))


def host_vec(ctx, D, p, field, count, dtype=np.float64):
    d = ctx.download(p, count, dtype).astype(np.float64)
    out = np.empty_like(d)
    out[D.perm(field)] = d
    return out


@pytest.mark.parametrize("layout,prec", [("csr", 0), ("tiles", 0), ("csr", 1), ("tiles", 1)])
def test_correction_steps_are_the_original_steps(ctx, layout, prec):
    rng = np.random.default_rng(17 + prec)
    M, Ms, rs, cs = scaled_problem(5)
    m, n = Ms.m, Ms.n
    base = Device(ctx, Ms, 1)                 # the scaled LP, FP64
    corr = Device(ctx, Ms, 1, prec)           # its correction problem
    assert np.array_equal(base.pr, corr.pr) and np.array_equal(base.pc, corr.pc)   # one device order
    if layout == "tiles":
        set_tiles(ctx, base.A, 1, 0)
        set_tiles(ctx, corr.A, 1, prec)
    sg = ctx.upload(np.zeros(m, corr.dt))
    corr.st.s_grad = sg
    # an anchor inside the box, some coordinates on their bounds
    lo, hi = Ms.col_lo, Ms.col_hi
    x0 = lo + rng.uniform(0, 1, n) * np.where(np.isfinite(hi - lo), hi - lo, 1.0)
    on = rng.uniform(size=n) < 0.3
    x0[on] = lo[on]
    y0 = rng.standard_normal(m)
    s0, aty0 = O.spmv(Ms, x0), O.spmv_t(Ms, y0)
    delta = 1024.0 + 0.5
    d_x0, d_y0 = dev_vec(ctx, base, x0, "x"), dev_vec(ctx, base, y0, "y")
    d_s0, d_aty0 = dev_vec(ctx, base, s0, "s"), dev_vec(ctx, base, aty0, "x")
    check(ctx.L.sov_hip_pdhg_anchor(ctx.h, C.byref(base.st), C.byref(corr.st), d_x0, d_y0, d_s0, d_aty0, delta))

    # the correction problem and its start, element for element
    dt = corr.dt
    want = dict(c=(delta * (Ms.c - aty0)).astype(dt), col_lo=(delta * (lo - x0)).astype(dt),
                col_hi=(delta * (hi - x0)).astype(dt), row_lo=(delta * (Ms.row_lo - s0)).astype(dt),
                row_hi=(delta * (Ms.row_hi - s0)).astype(dt))
    for f, v in want.items():
        np.testing.assert_array_equal(corr.get(f, n if f.startswith(("c", "col")) else m), v, err_msg=f)
    np.testing.assert_array_equal(host_vec(ctx, corr, sg, "s", m, dt), (delta * y0).astype(dt))
    np.testing.assert_array_equal(corr.get("x", n), np.clip(0, want["col_lo"], want["col_hi"]))
    np.testing.assert_array_equal(corr.get("s", m), np.clip(0, want["row_lo"], want["row_hi"]))
    np.testing.assert_array_equal(corr.get("y", m), np.zeros(m))

    # the scaled LP from the point the correction's start stands for
    base.put(x=x0, y=y0, s=np.clip(s0, Ms.row_lo, Ms.row_hi))
    tau, sigma, k = 0.31, 0.77, 6
    for it in range(k):
        for D in (base, corr):
            check(ctx.L.sov_hip_pdhg_primal_step(ctx.h, D.A.h, C.byref(D.st), tau, int(it == 0)))
            check(ctx.L.sov_hip_pdhg_dual_step(ctx.h, D.A.h, C.byref(D.st), tau, sigma, int(it == 0)))
    # combine into a third FP64 state and compare with the direct iterates
    out = PdhgState()
    out.prec, out.m, out.n = 0, m, n
    for f, cnt in (("x", n), ("x_sum", n), ("y", m), ("y_sum", m), ("x_bar", n), ("s", m), ("c", n),
                   ("col_lo", n), ("col_hi", n), ("row_lo", m), ("row_hi", m)):
        setattr(out, f, ctx.upload(np.zeros(cnt)))
    check(ctx.L.sov_hip_pdhg_combine(ctx.h, C.byref(corr.st), C.byref(out), d_x0, d_y0, k, 1.0 / delta))
    u = 2.0 ** -24 if prec == 1 else 2.0 ** -52
    for f, cnt, anchor, scale in (("x", n, x0, 1), ("x_sum", n, x0, k), ("y", m, y0, 1), ("y_sum", m, y0, k)):
        got = host_vec(ctx, base, getattr(out, f), f, cnt)
        ref = base.get(f, cnt)
        moved = np.abs(ref - scale * anchor)
        # FP64 corrections: rounding of both runs; FP32: the moved distance
        # carries FP32 rounding, the anchor none
        bound = 64 * u * (moved + np.max(moved)) + 1e-13 * (1 + np.abs(ref))
        err = np.abs(got - ref)
        print(f"{layout} prec={prec} {f}: max err {err.max():.3g}, max moved {moved.max():.3g}", flush=True)
        assert np.all(err <= bound), (f, float(np.max(err / bound)))
        assert moved.max() > 1e-3   # the steps did move the point
    # the slack: s = s0 + ds / delta
    s_dir = base.get("s", m)
    s_cor = s0 + corr.get("s", m) / delta
    s_moved = np.abs(s_dir - s0)
    assert np.all(np.abs(s_cor - s_dir) <= 64 * u * (s_moved + s_moved.max()) + 1e-12 * (1 + np.abs(s_dir)))
    for D in (base, corr):
        D.A.destroy()


def test_anchored_check_is_the_check_of_the_combined_point(ctx):
    """sov_hip_pdhg_check on an FP32 correction with its anchor (the
    refinement fields of sov_hip_check_args) against the oracle's residuals
    of the combined point x0 + x_c / delta, y0 + y_c / delta and of its
    average, with the activity of the average written to act_avg."""
    rng = np.random.default_rng(23)
    M, Ms, rs, cs = scaled_problem(11)
    m, n = Ms.m, Ms.n
    base = Device(ctx, Ms, 1)
    corr = Device(ctx, Ms, 2, 1)
    pr, pc = base.pr, base.pc
    lo, hi = Ms.col_lo, Ms.col_hi
    x0 = lo + rng.uniform(0, 1, n) * np.where(np.isfinite(hi - lo), hi - lo, 1.0)
    y0 = rng.standard_normal(m)
    s0, aty0 = O.spmv(Ms, x0), O.spmv_t(Ms, y0)
    delta, since = 4096.0, 5
    # a correction of FP32 values (what the device holds), and its sums
    xc = (rng.standard_normal(n) * 50).astype(np.float32).astype(np.float64)
    yc = (rng.standard_normal(m) * 50).astype(np.float32).astype(np.float64)
    xs = ((xc + rng.standard_normal(n)) * since).astype(np.float32).astype(np.float64)
    ys = ((yc + rng.standard_normal(m)) * since).astype(np.float32).astype(np.float64)
    corr.put(x=xc, y=yc, x_sum=xs, y_sum=ys)
    a = CheckArgs()
    a.st = C.pointer(corr.st)
    a.since_restart = since
    a.tol_scaled, a.tol_orig = 1e-4, 1e-4
    a.obj_sign, a.obj_const = 1.0, M.obj_const
    a.row_scale, a.col_scale = ctx.upload(rs[pr]), ctx.upload(cs[pc])
    a.orig_row_lo, a.orig_row_hi = ctx.upload(M.row_lo[pr]), ctx.upload(M.row_hi[pr])
    a.orig_col_lo, a.orig_col_hi, a.orig_c = ctx.upload(M.col_lo[pc]), ctx.upload(M.col_hi[pc]), ctx.upload(M.c[pc])
    act = ctx.upload(np.zeros(m))
    a.act_avg = act
    a.base = C.pointer(base.st)
    a.base_x, a.base_y = dev_vec(ctx, base, x0, "x"), dev_vec(ctx, base, y0, "y")
    a.base_act, a.base_aty = dev_vec(ctx, base, s0, "s"), dev_vec(ctx, base, aty0, "x")
    a.inv_delta = 1.0 / delta
    out = CheckOut()
    check(ctx.L.sov_hip_pdhg_check(ctx.h, corr.A.h, C.byref(a), C.byref(out)))
    # the combined points on the host
    x, y = x0 + xc / delta, y0 + yc / delta
    xa, ya = x0 + xs / since / delta, y0 + ys / since / delta
    # The products ran on the FP32 image of A_s times the correction; the
    # anchor's products are exact FP64.  So the residuals match the oracle's
    # of the combined point up to FP32 rounding of |A_s| |correction| / delta
    # (the bound below) plus summation order.
    Ma = O.Model(dict(m=m, n=n, row_ptr=Ms.row_ptr, col_idx=Ms.col_idx[: Ms.nnz], val=np.abs(Ms.val[: Ms.nnz]),
                      c=Ms.c, col_lo=Ms.col_lo, col_hi=Ms.col_hi, row_lo=Ms.row_lo, row_hi=Ms.row_hi))
    for pt, cr, kkt, rel_, p, d in (((x, y), (xc, yc), out.kkt_cur, out.rel_cur, out.pobj_cur, out.dobj_cur),
                                     ((xa, ya), (xs / since, ys / since), out.kkt_avg, out.rel_avg, out.pobj_avg,
                                      out.dobj_avg)):
        want = O.kkt_residual(Ms, Ms.c, pt[0], pt[1], 1e-4)
        err = 4 * 2.0 ** -24 / delta * max(np.max(O.spmv(Ma, np.abs(cr[0]))), np.max(O.spmv_t(Ma, np.abs(cr[1]))))
        assert abs(kkt - want) <= 1e-12 * max(1.0, abs(want)) + err, (kkt, want, err)
        w = O.relative_kkt(M, pt[0] * cs, pt[1] * rs)
        np.testing.assert_allclose(list(rel_), w[:3], rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose([p, d], w[3:5], rtol=1e-7)
    act_nat = np.empty(m)
    act_nat[pr] = ctx.download(act, m)
    bound = 4 * 2.0 ** -24 / delta * O.spmv(Ma, np.abs(xs / since)) + 1e-12 * (1 + np.abs(O.spmv(Ms, xa)))
    assert np.all(np.abs(act_nat - O.spmv(Ms, xa)) <= bound)
    # a base state in FP32 is refused
    bad = Device(ctx, Ms, 1, 1)
    a.base = C.pointer(bad.st)
    assert ctx.L.sov_hip_pdhg_check(ctx.h, corr.A.h, C.byref(a), C.byref(out)) != 0
    for D in (base, corr, bad):
        D.A.destroy()


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


CASES = [
    ("lp_2000x5000_s0", lambda: S.LP.random(2000, 5000, 5, 0)),
    ("transport_200x200_s0", lambda: S.LP.transportation(200, 200, 0)),
    ("zipf_3000x7000_s2", lambda: S.LP.random_zipf(3000, 7000, 2)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_fp32_refinement_reaches_the_exact_optimum(gpu, exact, name, make):
    lp = make()
    e = exact[name]
    kw = dict(backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative", restarts="adaptive",
              max_iterations=2000000)
    r = S.solve(lp, precision="f32", **kw)   # refinement is the FP32 default
    r64 = S.solve(lp, precision="f64", **kw)
    host = O.relative_kkt(O.Model(lp.arrays()), r.x, r.y)[:3]
    print(f"{name}: refined {r.status} it={r.iterations} (corrections {r.refinement_iterations}, anchors "
          f"{r.refinement_anchors}, fp64 polish {r.fp64_polish_iterations}) obj {r.objective:.12g} exact "
          f"{e['objective']:.12g} host rel {host} {r.seconds:.2f}s | fp64 it={r64.iterations} {r64.seconds:.2f}s",
          flush=True)
    assert r.status == "optimal", r.message
    assert max(r.relative_kkt) <= 1e-6 * 1.0001 and max(host) <= 1e-6 * 1.001
    assert rel(r.objective, e["objective"]) < 1e-6
    # every iteration in FP32: the plain phase, then corrections; no FP64 polish
    assert r.refinement_iterations > 0 and r.refinement_anchors >= 1 and r.fp64_polish_iterations == 0
