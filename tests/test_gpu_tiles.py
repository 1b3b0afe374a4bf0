"""The LDS-staged tiled products (sov_hip_csr_set_tiles) against the oracle:
SpMV in both orientations, the fused primal and dual steps, and whole
solves, in FP64 and FP32, in one launch and in several launches joined by
per-row partial sums.  The tiled kernels sum in another order than the
oracle (and than the CSR kernels), so products are held to the rounding
bound of a dot product: |error_i| <= (len_i + c) u (|A| |x|)_i with u the
unit roundoff of the arithmetic.  Results must also be run-to-run
identical (fixed summation order).

Oracle functions: orc_spmv / orc_spmv_t (backend_host.cpp:133-172),
orc_pdhg_solve (pdhg.cpp:218-502, relative-KKT stop, adaptive restarts)."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import Csr, Ctx, check
from test_gpu_steps import Device, abs_model, random_point, scaled_problem
from test_tile_layout import ragged_csr

pytestmark = pytest.mark.gpu

U = {0: 2.0 ** -53, 1: 2.0 ** -24}


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    yield c
    c.close()


def model(name):
    rng = np.random.default_rng(3)
    if name == "zipf":
        return O.Model(S.LP.random_zipf(3000, 7000, 2).arrays())
    if name == "wide":
        # 3 slices of A's columns (the last one ragged), a row with an entry in
        # every column (many wave passes per slice), empty rows
        m, n = 300, 40000
        lens = rng.integers(0, 30, m)
        lens[7] = n
        lens[50:60] = 0
        rp = np.zeros(m + 1, np.int32)
        rp[1:] = np.cumsum(lens)
        ci = np.concatenate([np.sort(rng.choice(n, int(k), replace=False)) for k in lens]).astype(np.int32)
        v = rng.uniform(0.5, 2.0, rp[-1]) * rng.choice([-1.0, 1.0], rp[-1])
        return O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=np.ones(n), col_lo=np.zeros(n),
                            col_hi=np.full(n, np.inf), row_lo=np.zeros(m), row_hi=np.ones(m)))
    if name == "tall":
        # more 64-row chunks than compute units in both orientations
        return O.Model(S.LP.random(30000, 75000, 6, 9).arrays())
    raise KeyError(name)


def lens(M, transpose):
    return np.bincount(M.col_idx[: M.nnz], minlength=M.n) if transpose else np.diff(M.row_ptr)


def set_tiles(ctx, A, parts, prec):
    if prec == 1:
        check(ctx.L.sov_hip_csr_make_f32(ctx.h, A.h))
    for orient in (0, 1):
        check(ctx.L.sov_hip_csr_set_tiles(ctx.h, A.h, orient, parts))
        got, ok = C.c_int(), C.c_int()
        check(ctx.L.sov_hip_csr_get_tiles_prec(A.h, orient, prec, C.byref(got)))
        check(ctx.L.sov_hip_csr_tiles_usable(ctx.h, A.h, orient, prec, C.byref(ok)))
        assert got.value >= 1 and ok.value == 1, (orient, got.value, ok.value)


@pytest.mark.parametrize("name", ["zipf", "wide", "tall"])
@pytest.mark.parametrize("prec,parts", [(0, 1), (0, 3), (1, 1), (1, 2)])
def test_tiled_spmv_matches_oracle(ctx, name, prec, parts):
    M = model(name)
    Ma = O.Model(dict(m=M.m, n=M.n, row_ptr=M.row_ptr, col_idx=M.col_idx[: M.nnz], val=np.abs(M.val[: M.nnz]),
                      c=M.c, col_lo=M.col_lo, col_hi=M.col_hi, row_lo=M.row_lo, row_hi=M.row_hi))
    A = Csr(ctx, M.m, M.n, M.row_ptr, M.col_idx[: M.nnz], M.val[: M.nnz])
    set_tiles(ctx, A, parts, prec)
    dt = np.float64 if prec == 0 else np.float32
    rng = np.random.default_rng(parts + 10 * prec)
    vals = M.val[: M.nnz].astype(dt).astype(np.float64)
    Mr = O.Model(dict(m=M.m, n=M.n, row_ptr=M.row_ptr, col_idx=M.col_idx[: M.nnz], val=vals, c=M.c,
                      col_lo=M.col_lo, col_hi=M.col_hi, row_lo=M.row_lo, row_hi=M.row_hi))
    for transpose in (False, True):
        x = rng.standard_normal(M.m if transpose else M.n).astype(dt).astype(np.float64)
        got = A.spmv(x, transpose=transpose, prec=prec).astype(np.float64)
        again = A.spmv(x, transpose=transpose, prec=prec).astype(np.float64)
        np.testing.assert_array_equal(got, again)
        want = O.spmv_t(Mr, x) if transpose else O.spmv(Mr, x)
        bound = (lens(M, transpose) + 4) * U[prec] * 2 * (O.spmv_t(Ma, np.abs(x)) if transpose
                                                         else O.spmv(Ma, np.abs(x)))
        bad = np.abs(got - want) > bound
        print(name, prec, parts, transpose, "max err / bound", float(np.max(np.abs(got - want) / np.maximum(bound, 1e-300))),
              flush=True)
        assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], want[bad][:5])
    A.destroy()


@pytest.mark.parametrize("name", ["zipf", "tall"])
@pytest.mark.parametrize("prec,parts", [(0, 1), (1, 1), (1, 2)])
def test_tiled_products_in_sequence(ctx, name, prec, parts):
    """Two workgroups share each row group, and the last to finish adds the
    other's published row sums: products of different vectors back to back
    (both orientations interleaved) must each be the product of their own
    vector, which a stale or missing half would break far beyond rounding."""
    M = model(name)
    A = Csr(ctx, M.m, M.n, M.row_ptr, M.col_idx[: M.nnz], M.val[: M.nnz])
    set_tiles(ctx, A, parts, prec)
    dt = np.float64 if prec == 0 else np.float32
    vals = M.val[: M.nnz].astype(dt).astype(np.float64)
    Mr = O.Model(dict(m=M.m, n=M.n, row_ptr=M.row_ptr, col_idx=M.col_idx[: M.nnz], val=vals, c=M.c,
                      col_lo=M.col_lo, col_hi=M.col_hi, row_lo=M.row_lo, row_hi=M.row_hi))
    Ma = O.Model(dict(m=M.m, n=M.n, row_ptr=M.row_ptr, col_idx=M.col_idx[: M.nnz], val=np.abs(vals), c=M.c,
                      col_lo=M.col_lo, col_hi=M.col_hi, row_lo=M.row_lo, row_hi=M.row_hi))
    rng = np.random.default_rng(40 + parts + 10 * prec)
    for k in range(6):
        transpose = k % 2 == 1
        x = rng.standard_normal(M.m if transpose else M.n).astype(dt).astype(np.float64) * (k + 1)
        got = A.spmv(x, transpose=transpose, prec=prec).astype(np.float64)
        want = O.spmv_t(Mr, x) if transpose else O.spmv(Mr, x)
        bound = (lens(M, transpose) + 4) * U[prec] * 2 * (O.spmv_t(Ma, np.abs(x)) if transpose
                                                         else O.spmv(Ma, np.abs(x)))
        bad = np.abs(got - want) > bound
        assert not bad.any(), (k, np.flatnonzero(bad)[:5], got[bad][:5], want[bad][:5])
    A.destroy()


@pytest.mark.parametrize("prec,parts", [(0, 1), (0, 2), (1, 1), (1, 2)])
def test_tiled_spmv_with_ragged_lengths(ctx, prec, parts):
    """Gathered lengths past the last whole 16-byte vector (n = 40003,
    m = 301), with that partial vector in a staged unit in both orientations
    (pinned on the CPU by test_tile_layout.py): the staging waves load whole
    vectors only and write the partial one's elements separately."""
    m, n, ptr, idx = ragged_csr()
    rng = np.random.default_rng(5 + parts + 10 * prec)
    val = rng.uniform(0.5, 2.0, len(idx)) * rng.choice([-1.0, 1.0], len(idx))
    dt = np.float64 if prec == 0 else np.float32
    vals = val.astype(dt).astype(np.float64)
    ones = dict(c=np.ones(n), col_lo=np.zeros(n), col_hi=np.ones(n), row_lo=np.zeros(m), row_hi=np.ones(m))
    Mr = O.Model(dict(m=m, n=n, row_ptr=ptr, col_idx=idx, val=vals, **ones))
    Ma = O.Model(dict(m=m, n=n, row_ptr=ptr, col_idx=idx, val=np.abs(vals), **ones))
    A = Csr(ctx, m, n, ptr, idx, val)
    set_tiles(ctx, A, parts, prec)
    for transpose in (False, True):
        x = rng.standard_normal(m if transpose else n).astype(dt).astype(np.float64)
        got = A.spmv(x, transpose=transpose, prec=prec).astype(np.float64)
        want = O.spmv_t(Mr, x) if transpose else O.spmv(Mr, x)
        ln = np.bincount(idx, minlength=n) if transpose else np.diff(ptr)
        bound = (ln + 4) * U[prec] * 2 * (O.spmv_t(Ma, np.abs(x)) if transpose else O.spmv(Ma, np.abs(x)))
        print("ragged", prec, parts, transpose, "max err / bound",
              float(np.max(np.abs(got - want) / np.maximum(bound, 1e-300))), flush=True)
        assert np.all(np.abs(got - want) <= bound)
    A.destroy()


@pytest.mark.parametrize("prec,parts", [(0, 1), (0, 2), (1, 1), (1, 3)])
def test_tiled_steps_match_oracle(ctx, prec, parts):
    rng = np.random.default_rng(70 + parts + 10 * prec)
    M, Ms, rs, cs = scaled_problem(5, m=3000, n=40000)
    D = Device(ctx, Ms, 1, prec)
    set_tiles(ctx, D.A, parts, prec)
    Ma = abs_model(Ms)
    u = U[prec]
    dt = D.dt
    x, y, xs, ys, s = random_point(rng, Ms, 3)
    x, y, xs, ys, s = (v.astype(dt).astype(np.float64) for v in (x, y, xs, ys, s))
    tau, sigma = 0.37, 0.81
    lt, la = lens(Ms, True), lens(Ms, False)
    for first in (1, 0):
        D.put(x=x, y=y, x_sum=xs, y_sum=ys, s=s)
        check(ctx.L.sov_hip_pdhg_primal_step(ctx.h, D.A.h, C.byref(D.st), tau, first))
        g = Ms.c - O.spmv_t(Ms, y)
        xn = np.clip(x - tau * g, Ms.col_lo, Ms.col_hi)
        gx = D.get("x", Ms.n)
        tol = (lt + 8) * u * 4 * (np.abs(x) + tau * (np.abs(Ms.c) + O.spmv_t(Ma, np.abs(y))))
        assert np.all(np.abs(gx - xn) <= tol)
        if prec == 0:
            np.testing.assert_array_equal(D.get("x_bar", Ms.n), 2 * gx - x)
            np.testing.assert_array_equal(D.get("x_sum", Ms.n), gx if first else xs + gx)
        xbar = D.get("x_bar", Ms.n)
        check(ctx.L.sov_hip_pdhg_dual_step(ctx.h, D.A.h, C.byref(D.st), tau, sigma, first))
        sn = np.clip(s - tau * y, Ms.row_lo, Ms.row_hi)
        yn = y + sigma * (-O.spmv(Ms, xbar) + (2 * sn - s))
        gy = D.get("y", Ms.m)
        tol = (la + 8) * u * 4 * (np.abs(y) + sigma * (O.spmv(Ma, np.abs(xbar)) + 3 * np.abs(s)))
        assert np.all(np.abs(gy - yn) <= tol)
    D.A.destroy()


@pytest.fixture(scope="module")
def solve_case():
    """An LP whose A spans two slices (the second ragged), with its oracle solve."""
    lp = S.LP.random(8000, 20000, 3, 5)
    o = O.pdhg(O.Model(lp.arrays()), tolerance=1e-6, termination=1, restart_scheme=1, max_iterations=400000)
    assert o["status"] == "optimal"
    return lp, o


@pytest.mark.parametrize("precision,tiles,after", [("f64", 1, None), ("f32", 2, None), ("f32", 1, 1e-9)])
def test_solve_on_tiles_reaches_the_oracles_objective(gpu, solve_case, precision, tiles, after):
    # after = 1e-9: the autotuner runs at the first check, mid-solve, and the
    # forced tiles are set again after it
    lp, o = solve_case
    kw = dict(backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative", restarts="adaptive",
              precision=precision, tiles=tiles, max_iterations=400000, autotune_after=after)
    r = S.solve(lp, **kw)
    print(precision, tiles, r.status, r.iterations, r.objective, o["objective"], flush=True)
    assert r.status == "optimal"
    assert abs(r.objective - o["objective"]) <= 1e-6 * max(1.0, abs(o["objective"]))
    assert max(r.relative_kkt) <= 1e-6 * 1.0001
    # the same solve again: bit-identical (fixed summation order, autotuner off)
    r2 = S.solve(lp, autotune=False, **kw)
    r3 = S.solve(lp, autotune=False, **kw)
    assert r2.iterations == r3.iterations and r2.objective == r3.objective
