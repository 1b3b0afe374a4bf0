"""The batched engine (sov::PdhgBatch): many LPs stacked block-diagonally in
one device matrix.  Kernel-level parity of the per-instance check (the
segmented reductions and the per-instance Cesaro divisors) against the
oracle's residual functions, instance by instance; end-to-end, every
instance of a batch reaches the oracle's objective (FP64 oracle solve, 1e-6
relative) and its own relative KKT tolerance, in FP64 and in FP32 + polish,
and an instance's answer does not depend on its batch mates.

Oracle functions: orc_kkt_residual (pdhg.cpp:87-136), orc_relative_kkt,
orc_pdhg_solve (pdhg.cpp:218-502, relative-KKT stop, adaptive restarts)."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import Csr, Ctx, PdhgState, CheckArgs, CheckOut, check

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-6


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


def stack(models):
    """Block-diagonal stack of oracle models: (arrays, row_off, col_off)."""
    rp, ci, v = [0], [], []
    row_off, col_off = [0], [0]
    cat = {k: [] for k in ("c", "col_lo", "col_hi", "row_lo", "row_hi")}
    for M in models:
        base = rp[-1]
        rp += [base + int(t) for t in M.row_ptr[1:]]
        ci += [int(j) + col_off[-1] for j in M.col_idx[: M.nnz]]
        v += list(M.val[: M.nnz])
        for k in cat:
            cat[k] += list(getattr(M, k))
        row_off.append(row_off[-1] + M.m)
        col_off.append(col_off[-1] + M.n)
    d = dict(m=row_off[-1], n=col_off[-1], row_ptr=np.array(rp, np.int32), col_idx=np.array(ci, np.int32),
             val=np.array(v), **{k: np.array(x) for k, x in cat.items()})
    return d, np.array(row_off, np.int32), np.array(col_off, np.int32)


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    L = c.L
    vp = C.c_void_p
    L.sov_hip_batch_create.argtypes = [vp, vp, C.c_int32, vp, vp, C.POINTER(vp)]
    L.sov_hip_batch_destroy.argtypes = [vp]
    L.sov_hip_batch_set.argtypes = [vp, vp, vp, vp, vp, vp]
    L.sov_hip_batch_check.argtypes = [vp, vp, vp, vp, vp, vp]
    L.sov_hip_batch_sumsq.argtypes = [vp, vp, C.c_int, vp, vp]
    L.sov_hip_batch_scale.argtypes = [vp, vp, C.c_int, vp, vp]
    yield c
    c.close()


@pytest.mark.parametrize("K", [1, 3])
def test_batch_check_and_norms_match_each_instance(ctx, K):
    rng = np.random.default_rng(17 + K)
    models = [O.Model(S.LP.random_zipf(m, n, seed, kmax=min(m, 200)).arrays()) for m, n, seed in
              ((300, 700, 1), (40, 90, 2), (900, 2100, 3), (120, 260, 4), (9000, 20000, 5))]
    # the last instance's rows and columns exceed the check reduction's
    # 8192-entry chunk, so its sums are split over chunks and finished by
    # the last chunk to arrive
    d, ro, co = stack(models)
    B = len(models)
    A = Csr(ctx, d["m"], d["n"], d["row_ptr"], d["col_idx"], d["val"])
    for orient in (0, 1):
        check(ctx.L.sov_hip_csr_set_segments(ctx.h, A.h, orient, K))
    bt = C.c_void_p()
    check(ctx.L.sov_hip_batch_create(ctx.h, A.h, B, ro.ctypes.data, co.ctypes.data, C.byref(bt)))
    pr, pc = A.order(0), A.order(1)
    since = np.array([3, 5, 1, 7, 4], np.int32)
    check(ctx.L.sov_hip_batch_set(ctx.h, bt, None, None, None, since.ctypes.data))
    # a point with some columns on their bounds, per-instance running sums
    x = d["col_lo"] + rng.uniform(0, 1, d["n"]) * (d["col_hi"] - d["col_lo"])
    at = rng.uniform(size=d["n"]) < 0.3
    x[at] = d["col_lo"][at]
    y = rng.standard_normal(d["m"])
    cinst = np.repeat(np.arange(B), np.diff(co))
    rinst = np.repeat(np.arange(B), np.diff(ro))
    xa, ya = x + 0.1 * rng.standard_normal(d["n"]), y + 0.1 * rng.standard_normal(d["m"])
    xs, ys = xa * since[cinst], ya * since[rinst]
    st = PdhgState()
    st.prec, st.m, st.n = 0, d["m"], d["n"]
    up_c = lambda v: ctx.upload(np.ascontiguousarray(np.asarray(v, np.float64)[pc]))
    up_r = lambda v: ctx.upload(np.ascontiguousarray(np.asarray(v, np.float64)[pr]))
    st.c, st.col_lo, st.col_hi, st.x, st.x_bar, st.x_sum = (up_c(v) for v in (d["c"], d["col_lo"], d["col_hi"], x,
                                                                               x, xs))
    st.row_lo, st.row_hi, st.s, st.y, st.y_sum = (up_r(v) for v in (d["row_lo"], d["row_hi"], d["row_lo"], y, ys))
    a = CheckArgs()
    a.st = C.pointer(st)
    a.tol_scaled, a.tol_orig, a.obj_sign = 1e-4, 1e-4, 1.0
    a.row_scale, a.col_scale = up_r(np.ones(d["m"])), up_c(np.ones(d["n"]))
    a.orig_row_lo, a.orig_row_hi = up_r(d["row_lo"]), up_r(d["row_hi"])
    a.orig_col_lo, a.orig_col_hi, a.orig_c = up_c(d["col_lo"]), up_c(d["col_hi"]), up_c(d["c"])
    a.act_avg = up_r(np.zeros(d["m"]))
    out = (CheckOut * B)()
    oc = np.zeros(B)
    check(ctx.L.sov_hip_batch_check(ctx.h, A.h, bt, C.byref(a), oc.ctypes.data, out))
    for b, M in enumerate(models):
        xb, yb = x[co[b]:co[b + 1]], y[ro[b]:ro[b + 1]]
        xab, yab = xs[co[b]:co[b + 1]] * (1.0 / since[b]), ys[ro[b]:ro[b + 1]] * (1.0 / since[b])
        for got, want in ((out[b].kkt_cur, O.kkt_residual(M, M.c, xb, yb, 1e-4)),
                          (out[b].kkt_avg, O.kkt_residual(M, M.c, xab, yab, 1e-4))):
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (b, got, want)
        for pt, rl in (((xb, yb), out[b].rel_cur), ((xab, yab), out[b].rel_avg)):
            np.testing.assert_allclose(list(rl), O.relative_kkt(M, *pt)[:3], rtol=1e-10, atol=1e-14)
    # per-instance sums of squares and scaling
    v = rng.standard_normal(d["m"])
    dv = up_r(v)
    sq = np.zeros(B)
    check(ctx.L.sov_hip_batch_sumsq(ctx.h, bt, 0, dv, sq.ctypes.data))
    np.testing.assert_allclose(sq, [np.sum(v[ro[b]:ro[b + 1]] ** 2) for b in range(B)], rtol=1e-13)
    fac = np.array([2.0, -1.0, 0.5, 3.0, 1.5])
    check(ctx.L.sov_hip_batch_scale(ctx.h, bt, 0, fac.ctypes.data, dv))
    got = np.empty(d["m"])
    got[pr] = ctx.download(dv, d["m"])
    np.testing.assert_array_equal(got, v * fac[rinst])
    ctx.L.sov_hip_batch_destroy(bt)
    A.destroy()


def c0_models(sizes, seed0=10):
    return [S.LP.random(m, int(2.5 * m), 5, seed0 + k) for k, m in enumerate(sizes)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_solves_every_instance_to_the_oracles_objective(gpu, prec):
    lps = c0_models([60, 400, 150, 1000, 90, 230])
    res, timing = S.solve_batch(lps, tolerance=1e-6, precision=prec, max_iterations=400000)
    print(prec, timing, [(r.status, r.iterations) for r in res], flush=True)
    for lp, r in zip(lps, res):
        o = O.pdhg(O.Model(lp.arrays()), tolerance=1e-6, termination=1, restart_scheme=1, max_iterations=400000)
        assert o["status"] == "optimal"
        assert r.status == "optimal", r
        assert rel(r.objective, o["objective"]) < OBJ_RTOL, (r.objective, o["objective"])
        # the relative KKT of the reported point, recomputed on the host from x and y
        assert max(r.relative_kkt) <= 1e-6 * 1.0001
    assert timing["iterations"] == max(r.iterations for r in res)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_resident_intervals_equal_the_step_launches(gpu, prec):
    """Instances whose iterates fit the LDS budget run each interval resident
    in one workgroup; the others stay on the CSR step launches.  Both add a
    row's entries in the same order with the same lanes and update it with
    the same arithmetic, so a batch solved with every instance on the CSR
    launches (budget 0), with a 32 KiB budget (the smaller instances
    resident) and with the full budget takes the same iterations to the
    same points, bit for bit (tiled steps off: they add in another order)."""
    lps = c0_models([60, 400, 150, 1000, 90, 230])
    runs = {sb: S.solve_batch(lps, tolerance=1e-6, precision=prec, autotune=False, max_iterations=400000,
                              stage_bytes=sb, tile_share=-1.0)[0] for sb in (0, 32768, None)}
    for sb in (32768, None):
        for b, (r0, r) in enumerate(zip(runs[0], runs[sb])):
            assert r.status == r0.status == "optimal", (sb, b)
            assert r.iterations == r0.iterations and r.objective == r0.objective, (sb, b, r.iterations, r0.iterations)
            np.testing.assert_array_equal(r.x, r0.x)
            np.testing.assert_array_equal(r.y, r0.y)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_whole_solves_in_resident_intervals(gpu, prec):
    """Every instance small enough to be resident from the first interval:
    each solve runs entirely in resident launches and ends where the step
    launches end, bit for bit, at the oracle's objective."""
    lps = c0_models([60, 150, 90, 230, 300, 120], seed0=90)
    kw = dict(tolerance=1e-6, precision=prec, autotune=False, max_iterations=400000, tile_share=-1.0)
    steps = S.solve_batch(lps, stage_bytes=0, **kw)[0]
    res = S.solve_batch(lps, **kw)[0]
    for b, (lp, r0, r) in enumerate(zip(lps, steps, res)):
        assert r.status == r0.status == "optimal", b
        assert r.iterations == r0.iterations and r.objective == r0.objective, (b, r.iterations, r0.iterations)
        np.testing.assert_array_equal(r.x, r0.x)
        np.testing.assert_array_equal(r.y, r0.y)
        o = O.pdhg(O.Model(lp.arrays()), tolerance=1e-6, termination=1, restart_scheme=1, max_iterations=400000)
        tol = 2e-6 * (1.0 + abs(r.objective) + abs(o["objective"]))
        assert abs(r.objective - o["objective"]) <= tol, (b, r.objective, o["objective"])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_streams_and_graphs_do_not_change_the_iterations(gpu, prec):
    """The FP64 instances' steps and the resident intervals on the batch's
    own stream (split) or all on the context stream, in captured interval
    graphs or as plain launches: the same kernels make the same sums, so
    every instance takes the same iterations to the same point, bit for
    bit."""
    lps = c0_models([60, 400, 150, 1000, 90, 230], seed0=70)
    kw = dict(tolerance=1e-6, precision=prec, autotune=False, max_iterations=400000)
    base = S.solve_batch(lps, split=True, use_graphs=True, **kw)[0]
    for split, graphs in ((False, True), (True, False)):
        res = S.solve_batch(lps, split=split, use_graphs=graphs, **kw)[0]
        for b, (r0, r) in enumerate(zip(base, res)):
            assert r.status == r0.status == "optimal", (split, graphs, b)
            assert r.iterations == r0.iterations and r.objective == r0.objective, (split, graphs, b)
            np.testing.assert_array_equal(r.x, r0.x)
            np.testing.assert_array_equal(r.y, r0.y)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tiled_batch_steps_reach_the_oracles_optima(gpu, prec):
    """The batch's steps on the tiled layouts of the whole stacked matrix
    (share 0: tiled while any instance iterates in the precision) and on the
    CSR block lists only (share -1) solve every instance to the oracle's
    objective; the tiled sums differ from the CSR ones in rounding only."""
    lps = c0_models([60, 400, 150, 1000, 90, 230], seed0=40)
    runs = {sh: S.solve_batch(lps, tolerance=1e-6, precision=prec, max_iterations=400000, stage_bytes=0,
                              tile_share=sh, autotune=False)[0] for sh in (0.0, -1.0)}
    for b, lp in enumerate(lps):
        o = O.pdhg(O.Model(lp.arrays()), tolerance=1e-6, termination=1, restart_scheme=1, max_iterations=400000)
        for sh, res in runs.items():
            r = res[b]
            assert r.status == "optimal", (sh, b, r)
            assert max(r.relative_kkt) <= 1e-6 * 1.0001, (sh, b)
            # both points end with a relative duality gap <= 1e-6 of
            # 1 + |primal objective| + |dual objective|, so their objectives
            # agree to twice that
            tol = 2e-6 * (1.0 + abs(r.objective) + abs(o["objective"]))
            assert abs(r.objective - o["objective"]) <= tol, (sh, b, r.objective, o["objective"])


def test_an_instance_does_not_depend_on_its_batch(gpu):
    lps = c0_models([200, 500, 80])
    alone, _ = S.solve_batch(lps[:1], tolerance=1e-6, precision="f64", autotune=False)
    mixed, _ = S.solve_batch(lps, tolerance=1e-6, precision="f64", autotune=False)
    single = S.solve(lps[0], backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative",
                     restarts="adaptive", autotune=False)
    for r in (alone[0], mixed[0]):
        assert r.status == "optimal" and rel(r.objective, single.objective) < OBJ_RTOL


def test_invalid_instance_fails_alone(gpu):
    good = c0_models([100, 120])
    bad = S.LP.from_csr(1, 1, [0, 1], [0], [1.0], [1.0], [2.0], [1.0], [0.0], [5.0])   # crossed column bounds
    res, _ = S.solve_batch([good[0], bad, good[1]], tolerance=1e-6, precision="f64")
    assert [r.status for r in res] == ["optimal", "numerical_failure", "optimal"]
