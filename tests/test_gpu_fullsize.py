"""Parity at the sizes the benchmark runs.

  * the 2,000 x 5,000 LP (5 nonzeros per column, seed 0) in FP64 on the fused
    engine against the reference's own host PDHG run on it (the committed
    golden: 20,000 iterations, status, restarts, every check's residual), and
    to 1e-6 relative KKT in FP64, FP32 with iterative refinement and FP32 +
    FP64 polish against its exact optimum;
  * the 2,000 x 2,000 transportation LP (4M columns, 8M nonzeros) with FP32
    iterates + iterative refinement to 1e-6 relative KKT, on the default schedule and
    on the LDS-staged tiles, against its exact optimum;
  * the 1M-column Zipf LP (m = 500,000, about 21.5M nonzeros): one FP32 and
    one FP64 primal step and dual step on the tiles against the oracle's
    products under the dot-product rounding bound, and the FP32 solve with
    iterative refinement to 1e-6 relative KKT.

Exact optima come from tests/golden/exact_optima.json (HiGHS, the solver the
reference's own PDHG-vs-oracle check uses, via scipy; written by
tests/golden/make_exact_optima.py).  Every solution is also re-checked on the
host: the oracle's relative KKT (orc_relative_kkt, definitions of
include/sovereign/pdhg.hpp) of the returned original-space (x, y) must be
within the tolerance, independently of what the engine reports.  Where no
exact optimum exists (the Zipf LP), the gap term of that host check bounds
the objective error.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import Ctx, check
from test_gpu_steps import Device, abs_model, random_point
from test_gpu_tiles import lens, set_tiles
from golden.make_golden import digest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
OBJ_RTOL = 1e-6
U = {0: 2.0 ** -53, 1: 2.0 ** -24}


@pytest.fixture(scope="module")
def exact():
    with open(os.path.join(HERE, "golden", "exact_optima.json")) as f:
        return json.load(f)["instances"]


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def arrays_digest(a):
    return digest(a["row_ptr"], a["col_idx"], a["val"], a["c"], a["row_lo"])


def host_relative_kkt(lp, r):
    """The oracle's relative KKT [r_p, r_d, gap] of the returned point."""
    return O.relative_kkt(O.Model(lp.arrays()), r.x, r.y)[:3]


def assert_solved(lp, r, tol, optimum=None, label=""):
    rk = host_relative_kkt(lp, r)
    print(f"{label}: {r.status} obj {r.objective:.12g} (exact {optimum}) it={r.iterations} "
          f"polish={r.fp64_polish_iterations} engine rel={r.relative_kkt} host rel={rk} {r.seconds:.2f}s",
          flush=True)
    assert r.status == "optimal", r.message
    assert max(r.relative_kkt) <= tol * 1.0001
    assert max(rk) <= tol * 1.001
    if optimum is not None:
        assert rel(r.objective, optimum) < OBJ_RTOL


def test_lp_2000x5000_fp64_follows_the_reference_golden(gpu, golden):
    case = golden["random"]["c0_2000x5000_s0_tol0.0001"]
    g = case["result"]
    lp = S.LP.random(2000, 5000, 5, 0)
    assert arrays_digest(lp.arrays()) == case["digest"]
    r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=case["tolerance"],
                max_iterations=case["max_iterations"], autotune=False, trace_cap=100000)
    print(f"gpu {r.status} it={r.iterations} rs={r.restarts} obj {r.objective:.15g} | ref {g['status']} "
          f"it={g['iterations']} rs={g['restarts']} obj {g['objective']:.15g}", flush=True)
    assert r.status == g["status"]
    assert r.iterations == g["iterations"] and r.restarts == g["restarts"]
    assert abs(r.operator_norm_estimate - g["operator_norm"]) < 1e-12 * g["operator_norm"]
    # every check of the 20,000 iterations: the reference's original-space
    # residual, to rounding of the device's summation order
    assert len(r.trace_res) == len(g["trace"])
    np.testing.assert_allclose(r.trace_res, g["trace"], rtol=1e-8)
    assert rel(r.objective, g["objective"]) < 1e-10


@pytest.mark.parametrize("precision", ["f64", "f32", "f32-polish"])
def test_lp_2000x5000_to_1e6_reaches_the_exact_optimum(gpu, exact, precision):
    e = exact["lp_2000x5000_s0"]
    lp = S.LP.random(2000, 5000, 5, 0)
    assert arrays_digest(lp.arrays()) == e["digest"]
    # "f32": FP32 iterates with iterative refinement (the default);
    # "f32-polish": FP32 iterates, then FP64 iterates (refinement off)
    r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative",
                restarts="adaptive", precision=precision[:3], refinement=precision == "f32",
                max_iterations=2000000)
    assert_solved(lp, r, 1e-6, e["objective"], f"lp_2000x5000 {precision}")
    if precision == "f32":
        assert r.refinement_iterations > 0 and r.fp64_polish_iterations == 0
    if precision == "f32-polish":
        assert r.refinement_iterations == 0 and r.fp64_polish_iterations > 0


@pytest.fixture(scope="module")
def transport_2000():
    lp = S.LP.transportation(2000, 2000, 0)
    return lp


@pytest.mark.parametrize("tiles", [None, 1])
def test_transport_2000x2000_fp32_reaches_the_exact_optimum(gpu, exact, transport_2000, tiles):
    e = exact["transport_2000x2000_s0"]
    lp = transport_2000
    assert arrays_digest(lp.arrays()) == e["digest"]
    kw = {} if tiles is None else dict(tiles=tiles, autotune_after=0.0)
    r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative",
                restarts="adaptive", precision="f32", max_iterations=2000000, **kw)
    assert_solved(lp, r, 1e-6, e["objective"], f"transport_2000x2000 f32 tiles={tiles}")


@pytest.fixture(scope="module")
def zipf_scaled():
    """The 1M-column Zipf LP, its scale factors and the scaled model."""
    lp = S.LP.random_zipf(500_000, 1_000_000, 0)
    a = lp.arrays()
    M = O.Model(a)
    rs, cs = O.scaling(M)
    rows = np.repeat(np.arange(M.m), np.diff(M.row_ptr))
    vs = rs[rows] * M.val[: M.nnz] * cs[M.col_idx[: M.nnz]]
    d = dict(m=M.m, n=M.n, row_ptr=M.row_ptr, col_idx=M.col_idx[: M.nnz], val=vs, c=M.c * cs,
             col_lo=M.col_lo / cs, col_hi=M.col_hi / cs, row_lo=M.row_lo * rs, row_hi=M.row_hi * rs)
    return lp, O.Model(d)


@pytest.mark.parametrize("prec", [1, 0])
def test_zipf_1m_tiled_steps_match_oracle(gpu, zipf_scaled, prec):
    """One primal step and one dual step of the full 1M-column Zipf LP on the
    tiles (one launch per product, the bench's layout) against the oracle's
    FP64 step on the same FP32- (or FP64-) rounded inputs, under the rounding
    bound of a dot product in the kernel's arithmetic."""
    _, Ms = zipf_scaled
    ctx = Ctx(0)
    try:
        D = Device(ctx, Ms, 1, prec)
        set_tiles(ctx, D.A, 1, prec)
        Ma = abs_model(Ms)
        u = U[prec]
        dt = D.dt
        rng = np.random.default_rng(11 + prec)
        x, y, xs, ys, s = random_point(rng, Ms, 3)
        x, y, xs, ys, s = (v.astype(dt).astype(np.float64) for v in (x, y, xs, ys, s))
        # the kernels see the values rounded to their precision
        Mr = O.Model(dict(m=Ms.m, n=Ms.n, row_ptr=Ms.row_ptr, col_idx=Ms.col_idx[: Ms.nnz],
                          val=Ms.val[: Ms.nnz].astype(dt).astype(np.float64), c=Ms.c.astype(dt).astype(np.float64),
                          col_lo=Ms.col_lo.astype(dt).astype(np.float64),
                          col_hi=Ms.col_hi.astype(dt).astype(np.float64),
                          row_lo=Ms.row_lo.astype(dt).astype(np.float64),
                          row_hi=Ms.row_hi.astype(dt).astype(np.float64)))
        tau, sigma = 0.37, 0.81
        lt, la = lens(Ms, True), lens(Ms, False)
        D.put(x=x, y=y, x_sum=xs, y_sum=ys, s=s)
        check(ctx.L.sov_hip_pdhg_primal_step(ctx.h, D.A.h, C.byref(D.st), tau, 0))
        g = Mr.c - O.spmv_t(Mr, y)
        xn = np.clip(x - tau * g, Mr.col_lo, Mr.col_hi)
        gx = D.get("x", Ms.n)
        tol = (lt + 8) * u * 4 * (np.abs(x) + tau * (np.abs(Mr.c) + O.spmv_t(Ma, np.abs(y))))
        print(f"prec {prec} primal: max err / bound {float(np.max(np.abs(gx - xn) / np.maximum(tol, 1e-300))):.3g}",
              flush=True)
        assert np.all(np.abs(gx - xn) <= tol)
        xbar = D.get("x_bar", Ms.n)
        np.testing.assert_allclose(xbar, 2 * gx - x, rtol=4 * u, atol=4 * u * np.max(np.abs(x)))
        check(ctx.L.sov_hip_pdhg_dual_step(ctx.h, D.A.h, C.byref(D.st), tau, sigma, 0))
        sn = np.clip(s - tau * y, Mr.row_lo, Mr.row_hi)
        yn = y + sigma * (-O.spmv(Mr, xbar) + (2 * sn - s))
        gy = D.get("y", Ms.m)
        tol = (la + 8) * u * 4 * (np.abs(y) + sigma * (O.spmv(Ma, np.abs(xbar)) + 3 * np.abs(s)))
        print(f"prec {prec} dual: max err / bound {float(np.max(np.abs(gy - yn) / np.maximum(tol, 1e-300))):.3g}",
              flush=True)
        assert np.all(np.abs(gy - yn) <= tol)
        D.A.destroy()
    finally:
        ctx.close()


def test_zipf_1m_fp32_refined_to_1e6(gpu, zipf_scaled):
    """The headline solve: the 1M-column Zipf LP with FP32 iterates and
    iterative refinement to 1e-6 relative KKT, every iteration in FP32.  No
    exact optimum is available at this size; the host re-check of the
    relative KKT (gap term <= 1e-6) bounds the objective error."""
    lp, _ = zipf_scaled
    r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative",
                restarts="adaptive", precision="f32", max_iterations=2000000, time_limit=110)
    assert_solved(lp, r, 1e-6, None, "zipf_500000x1000000 f32")
    assert r.refinement_iterations > 0 and r.fp64_polish_iterations == 0
