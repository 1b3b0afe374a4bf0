"""GPU parity of the HIP kernels through the C-ABI (include/sov_hip.h) against
the CPU oracle and the reference's hand answers (tests/test_l0_backend.cpp)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import Csr, Ctx, check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    print("device:", c.L.sov_hip_ctx_device_name(c.h).decode(), flush=True)
    yield c
    c.close()


def ref4x3(ctx):
    return Csr(ctx, 4, 3, [0, 2, 4, 5, 6], [0, 2, 1, 2, 0, 2], [2.0, -1.0, 3.0, 4.0, 1.0, 5.0])


def test_level1_hand_answers(ctx):
    L = ctx.L
    x = ctx.upload(np.array([1.0, 2.0, 3.0, 4.0]))
    y = ctx.upload(np.array([10.0, 20.0, 30.0, 40.0]))
    check(L.sov_hip_axpy_f64(ctx.h, 4, 2.0, x, y))
    out = ctx.download(y, 4)
    assert out[0] == 12.0 and out[3] == 48.0
    check(L.sov_hip_scal_f64(ctx.h, 4, 0.5, y))
    assert ctx.download(y, 4)[0] == 6.0
    d = C.c_double()
    check(L.sov_hip_dot_f64(ctx.h, 4, x, x, C.byref(d)))
    assert d.value == 30.0
    check(L.sov_hip_nrm2_f64(ctx.h, 4, x, C.byref(d)))
    assert abs(d.value - math.sqrt(30.0)) < 1e-12
    z = ctx.upload(np.array([1.0, -7.5, 3.0, 2.0]))
    check(L.sov_hip_amax_f64(ctx.h, 4, z, C.byref(d)))
    assert d.value == 7.5
    d.value = 123.0
    check(L.sov_hip_dot_f64(ctx.h, 0, None, None, C.byref(d)))
    assert d.value == 0.0
    check(L.sov_hip_axpy_f64(ctx.h, 0, 1.0, None, None))


def test_reductions_on_large_vectors(ctx):
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal(3_000_001), rng.standard_normal(3_000_001)
    da, db = ctx.upload(a), ctx.upload(b)
    d = C.c_double()
    check(ctx.L.sov_hip_dot_f64(ctx.h, len(a), da, db, C.byref(d)))
    assert abs(d.value - O.dot(a, b)) < 1e-9 * np.sqrt(len(a))
    d2 = C.c_double()
    check(ctx.L.sov_hip_dot_f64(ctx.h, len(a), da, db, C.byref(d2)))
    assert d.value == d2.value   # fixed-order reduction: reproducible
    check(ctx.L.sov_hip_amax_f64(ctx.h, len(a), da, C.byref(d)))
    assert d.value == np.max(np.abs(a))


def test_project_box_with_null_bounds(ctx):
    L = ctx.L
    lo = ctx.upload(np.array([0.0, 0.0, -1.0, -5.0]))
    hi = ctx.upload(np.array([1.0, 10.0, 1.0, 5.0]))
    x = ctx.upload(np.array([-3.0, 4.0, 7.0, 0.5]))
    check(L.sov_hip_project_box_f64(ctx.h, 4, lo, hi, x))
    np.testing.assert_array_equal(ctx.download(x, 4), [0.0, 4.0, 1.0, 0.5])
    x2 = ctx.upload(np.array([-3.0, 1e9]))
    check(L.sov_hip_project_box_f64(ctx.h, 2, lo, None, x2))
    np.testing.assert_array_equal(ctx.download(x2, 2), [0.0, 1e9])


def test_spmv_hand_answers(ctx):
    A = ref4x3(ctx)
    np.testing.assert_array_equal(A.spmv([1, 2, 3]), [-1, 18, 1, 15])
    np.testing.assert_array_equal(A.spmv([1, 2, 3], alpha=2.0, beta=3.0, y=[1, 1, 1, 1]), [1, 39, 5, 33])
    np.testing.assert_array_equal(A.spmv([1, 2, 3, 4], transpose=True), [5, 6, 27])
    out = A.spmv([1, 2, 3], beta=0.0, y=[np.nan] * 4)
    assert not np.isnan(out).any() and out[1] == 18
    check(ctx.L.sov_hip_csr_make_f32(ctx.h, A.h))
    np.testing.assert_array_equal(A.spmv([1, 2, 3], prec=1), np.array([-1, 18, 1, 15], np.float32))
    A.destroy()


def random_csr(rng, m, n, density=None, zipf=False):
    rows = []
    for i in range(m):
        k = int(min(n, max(0, rng.zipf(1.5) if zipf else rng.poisson(density * n))))
        cols = np.sort(rng.choice(n, size=k, replace=False))
        rows.append(cols)
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    return rp, ci, rng.standard_normal(int(rp[-1]))


@pytest.mark.parametrize("shape,zipf", [((300, 500), False), ((2000, 700), True), ((50, 4000), True),
                                        ((1, 5000), False), ((5000, 1), False)])
def test_spmv_matches_oracle(ctx, shape, zipf):
    rng = np.random.default_rng(shape[0] + shape[1])
    m, n = shape
    rp, ci, v = random_csr(rng, m, n, density=0.02 if not zipf else None, zipf=zipf)
    if shape == (1, 5000):
        rp, ci = np.array([0, 5000], np.int32), np.arange(5000, dtype=np.int32)
        v = rng.standard_normal(5000)
    M = O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=np.zeros(n), col_lo=np.zeros(n),
                     col_hi=np.zeros(n), row_lo=np.zeros(m), row_hi=np.zeros(m)))
    A = Csr(ctx, m, n, rp, ci, v)
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    got, want = A.spmv(x), O.spmv(M, x)
    scale = np.abs(M.dense()) @ np.abs(x) + 1e-300
    assert np.max(np.abs(got - want) / scale) < 1e-14
    got, want = A.spmv(y, transpose=True), O.spmv_t(M, y)
    scale = np.abs(M.dense()).T @ np.abs(y) + 1e-300
    assert np.max(np.abs(got - want) / scale) < 1e-14
    # FP32 image: error bound k * eps32 relative to |A||x|
    check(ctx.L.sov_hip_csr_make_f32(ctx.h, A.h))
    got32 = A.spmv(x, prec=1).astype(np.float64)
    kmax = max(int(np.max(np.diff(rp))), 1)
    scale = np.abs(M.dense()) @ np.abs(x) + 1e-30
    assert np.max(np.abs(got32 - O.spmv(M, x)) / scale) < (kmax + 2) * 1.2e-7
    A.destroy()


def test_device_transpose_is_the_csc_of_a(ctx):
    rng = np.random.default_rng(7)
    m, n = 700, 900
    rp, ci, v = random_csr(rng, m, n, zipf=True)
    A = Csr(ctx, m, n, rp, ci, v)
    nnz = int(rp[-1])
    tp, ti, tv = ctx.alloc(4 * (n + 1)), ctx.alloc(4 * nnz), ctx.alloc(8 * nnz)
    check(ctx.L.sov_hip_csr_transpose_arrays(ctx.h, A.h, tp, ti, tv))
    tp, ti, tv = ctx.download(tp, n + 1, np.int32), ctx.download(ti, nnz, np.int32), ctx.download(tv, nnz)
    # host counting-sort transpose (ascending rows within each column)
    cnt = np.bincount(ci, minlength=n)
    ptr = np.concatenate([[0], np.cumsum(cnt)])
    rows = np.repeat(np.arange(m), np.diff(rp))
    order = np.lexsort((rows, ci))
    np.testing.assert_array_equal(tp, ptr)
    np.testing.assert_array_equal(ti, rows[order])
    np.testing.assert_array_equal(tv, v[order])
    A.destroy()


def test_device_transpose_with_a_dense_column(ctx):
    """One column touching every row (m = 400,000) next to short ones: the
    transpose is a sort, linear in nnz, so a dense column costs no more than
    its entries; the result is still the CSC of A with ascending rows."""
    import time
    rng = np.random.default_rng(8)
    m, n = 400_000, 64
    extra = [np.sort(rng.choice(np.arange(1, n), int(k), replace=False)) for k in rng.integers(0, 3, m)]
    rows_cols = [np.concatenate([[0], e]) for e in extra]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows_cols])]).astype(np.int32)
    ci = np.concatenate(rows_cols).astype(np.int32)
    v = rng.standard_normal(len(ci))
    t0 = time.perf_counter()
    A = Csr(ctx, m, n, rp, ci, v)
    check(ctx.L.sov_hip_sync(ctx.h))
    built = time.perf_counter() - t0
    nnz = int(rp[-1])
    tp, ti, tv = ctx.alloc(4 * (n + 1)), ctx.alloc(4 * nnz), ctx.alloc(8 * nnz)
    check(ctx.L.sov_hip_csr_transpose_arrays(ctx.h, A.h, tp, ti, tv))
    tp, ti, tv = ctx.download(tp, n + 1, np.int32), ctx.download(ti, nnz, np.int32), ctx.download(tv, nnz)
    rows = np.repeat(np.arange(m), np.diff(rp))
    order = np.lexsort((rows, ci))
    print(f"dense column of {m} entries, nnz {nnz}: handle built in {built:.2f} s", flush=True)
    np.testing.assert_array_equal(tp, np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n))]))
    np.testing.assert_array_equal(ti, rows[order])
    np.testing.assert_array_equal(tv, v[order])
    x = rng.standard_normal(m)
    zero = dict(c=np.zeros(n), col_lo=np.zeros(n), col_hi=np.zeros(n), row_lo=np.zeros(m), row_hi=np.zeros(m))
    want = O.spmv_t(O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, **zero)), x)
    mag = O.spmv_t(O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=np.abs(v), **zero)), np.abs(x))
    bound = (np.bincount(ci, minlength=n) + 4) * 2.0 ** -53 * 2 * mag
    assert np.all(np.abs(A.spmv(x, transpose=True) - want) <= bound)
    assert built < 10.0
    A.destroy()


def test_empty_matrix_and_empty_rows(ctx):
    A = Csr(ctx, 3, 4, [0, 0, 0, 0], [], [])
    np.testing.assert_array_equal(A.spmv([1, 2, 3, 4], beta=0.0, y=[np.nan] * 3), [0, 0, 0])
    np.testing.assert_array_equal(A.spmv([1, 2, 3], transpose=True), [0, 0, 0, 0])
    A.destroy()


def test_lane_autotuner_keeps_results(ctx):
    rng = np.random.default_rng(9)
    m, n = 3000, 6000
    rp, ci, v = random_csr(rng, m, n, zipf=True)
    A = Csr(ctx, m, n, rp, ci, v)
    x = rng.standard_normal(n)
    before = A.spmv(x)
    check(ctx.L.sov_hip_csr_tune(ctx.h, A.h, 0))
    lanes = (C.c_int32 * 8)()
    check(ctx.L.sov_hip_csr_get_lanes(A.h, 0, lanes))
    assert all(l in (1, 2, 4, 8, 16, 32, 64) for l in lanes)
    after = A.spmv(x)
    scale = np.abs(before) + 1e-12
    assert np.max(np.abs(after - before) / scale) < 1e-12
    for G in (1, 64):
        check(ctx.L.sov_hip_csr_set_lanes(A.h, 1, (C.c_int32 * 8)(*([G] * 8))))
        t = A.spmv(rng.standard_normal(m), transpose=True)
        assert np.all(np.isfinite(t))
    A.destroy()


@pytest.mark.parametrize("seed", [0, 1])
def test_device_scaling_matches_oracle(ctx, seed):
    lp = S.LP.random_zipf(2000, 4000, seed) if seed else S.LP.random(200, 500, 5, 0)
    a = lp.arrays()
    M = O.Model(a)
    A = Csr(ctx, a["m"], a["n"], a["row_ptr"], a["col_idx"], a["val"])
    rs, cs = ctx.upload(np.ones(a["m"])), ctx.upload(np.ones(a["n"]))
    w = C.c_double()
    for _ in range(8):
        check(ctx.L.sov_hip_scale_ruiz_sweep(ctx.h, A.h, rs, cs, C.byref(w)))
        if w.value < 1e-2:
            break
    check(ctx.L.sov_hip_scale_pock_chambolle(ctx.h, A.h, 1.0, rs, cs))
    grs, gcs = ctx.download(rs, a["m"]), ctx.download(cs, a["n"])
    ors, ocs = O.scaling(M)
    # same operations in the same order: equal up to the device sqrt's rounding
    assert np.max(np.abs(grs - ors) / ors) < 4e-16
    assert np.max(np.abs(gcs - ocs) / ocs) < 4e-16
    print("bitwise equal factors:", np.mean(grs == ors), np.mean(gcs == ocs))
    A.destroy()


@pytest.mark.parametrize("K", [2, 3, 8])
def test_column_segments_keep_results(ctx, K):
    rng = np.random.default_rng(21 + K)
    m, n = 4000, 9000
    rp, ci, v = random_csr(rng, m, n, zipf=True)
    M = O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=np.zeros(n), col_lo=np.zeros(n),
                     col_hi=np.zeros(n), row_lo=np.zeros(m), row_hi=np.zeros(m)))
    A = Csr(ctx, m, n, rp, ci, v)
    check(ctx.L.sov_hip_csr_set_segments(ctx.h, A.h, 0, K))
    check(ctx.L.sov_hip_csr_set_segments(ctx.h, A.h, 1, K))
    cnt = C.c_int()
    check(ctx.L.sov_hip_csr_get_segments(A.h, 0, C.byref(cnt)))
    assert cnt.value == K
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    Ad = np.abs(M.dense())
    for got, want, scale in ((A.spmv(x, alpha=2.0, beta=-1.0, y=np.ones(m)), 2 * O.spmv(M, x) - 1,
                              2 * Ad @ np.abs(x) + 1),
                             (A.spmv(y, transpose=True), O.spmv_t(M, y), Ad.T @ np.abs(y) + 1e-300)):
        assert np.max(np.abs(got - want) / scale) < 1e-14
    check(ctx.L.sov_hip_csr_make_f32(ctx.h, A.h))
    got32 = A.spmv(x, prec=1).astype(np.float64)
    kmax = int(np.max(np.diff(rp)))
    assert np.max(np.abs(got32 - O.spmv(M, x)) / (Ad @ np.abs(x) + 1e-30)) < (kmax + 2) * 1.2e-7
    A.destroy()


def test_device_order_and_permute(ctx):
    """The handle's device order groups the rows of A (and of A^T) by length
    class, stably; device-order SpMV equals the natural one permuted, and the
    permute entry point is exact both ways."""
    rng = np.random.default_rng(5)
    m, n = 3000, 7000
    rp, ci, v = random_csr(rng, m, n, zipf=True)
    M = O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=np.zeros(n), col_lo=np.zeros(n),
                     col_hi=np.zeros(n), row_lo=np.zeros(m), row_hi=np.zeros(m)))
    A = Csr(ctx, m, n, rp, ci, v)
    pr, pc = A.order(0), A.order(1)
    maxes = [2, 4, 8, 16, 32, 64, 256]
    cls = lambda lens: np.searchsorted(maxes, lens, side="left")
    np.testing.assert_array_equal(pr, np.argsort(cls(np.diff(rp)), kind="stable"))
    np.testing.assert_array_equal(pc, np.argsort(cls(np.bincount(ci[: rp[-1]], minlength=n)), kind="stable"))
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    Ad = np.abs(M.dense())
    for transpose, vin, pin, pout, want, scale in ((0, x, pc, pr, O.spmv(M, x), Ad @ np.abs(x)),
                                                   (1, y, pr, pc, O.spmv_t(M, y), Ad.T @ np.abs(y))):
        dx = ctx.upload(np.ascontiguousarray(vin[pin]))
        dy = ctx.upload(np.zeros(len(pout)))
        check(ctx.L.sov_hip_spmv_device(ctx.h, A.h, transpose, 0, 1.0, dx, 0.0, dy))
        got = np.empty(len(pout))
        got[pout] = ctx.download(dy, len(pout))
        assert np.max(np.abs(got - want) / (scale + 1e-300)) < 1e-14
        ctx.free(dx)
        ctx.free(dy)
    for which, vec, perm in ((0, y, pr), (1, x, pc)):
        for prec, dt in ((0, np.float64), (1, np.float32)):
            src = ctx.upload(vec.astype(dt))
            dev, back = ctx.upload(np.zeros(len(vec), dt)), ctx.upload(np.zeros(len(vec), dt))
            check(ctx.L.sov_hip_csr_permute(ctx.h, A.h, which, 0, prec, src, dev))
            check(ctx.L.sov_hip_csr_permute(ctx.h, A.h, which, 1, prec, dev, back))
            np.testing.assert_array_equal(ctx.download(dev, len(vec), dt), vec.astype(dt)[perm])
            np.testing.assert_array_equal(ctx.download(back, len(vec), dt), vec.astype(dt))
    A.destroy()
