"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

Checks: the reference's own hand answers (tests/test_l0_backend.cpp,
tests/test_l0_sparse.cpp, tests/test_scaling.cpp, tests/test_duals.cpp
semantics), dense brute force, closed forms, and bitwise agreement with the
reference host PDHG recorded in tests/golden/reference_pdhg.json.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O

INF = math.inf


def model(d):
    d = dict(d)
    for k in ("col_lo", "col_hi", "row_lo", "row_hi"):
        d[k] = [float(v) for v in d[k]]
    return O.Model(d)


def ref_matrix():
    # reference tests/test_l0_backend.cpp:13-21: the 4x3 matrix
    return O.Model(dict(m=4, n=3, row_ptr=[0, 2, 4, 5, 6], col_idx=[0, 2, 1, 2, 0, 2],
                        val=[2.0, -1.0, 3.0, 4.0, 1.0, 5.0], c=[0] * 3, col_lo=[0] * 3, col_hi=[0] * 3,
                        row_lo=[0] * 4, row_hi=[0] * 4))


def test_spmv_hand_answers():
    M = ref_matrix()
    np.testing.assert_array_equal(O.spmv(M, [1, 2, 3]), [-1, 18, 1, 15])
    np.testing.assert_array_equal(O.spmv_t(M, [1, 2, 3, 4]), [5, 6, 27])
    np.testing.assert_array_equal(O.spmv(M, [1, 2, 3], alpha=2.0, beta=3.0, y=[1, 1, 1, 1]), [1, 39, 5, 33])
    out = O.spmv(M, [1, 2, 3], beta=0.0, y=[np.nan] * 4)
    assert not np.isnan(out).any() and out[1] == 18


def test_dot_hand_and_pairwise():
    assert O.dot([1, 2, 3, 4], [1, 2, 3, 4]) == 30.0
    rng = np.random.default_rng(0)
    x = rng.standard_normal(1000)
    # pairwise with 128-leaves: split 500/500, 250/250, 125/125
    leaf = [x[i:i + 125] @ x[i:i + 125] for i in range(0, 1000, 125)]
    s = 0.0
    tree = [sum_serial(x[i:i + 125]) for i in range(0, 1000, 125)]
    s = ((tree[0] + tree[1]) + (tree[2] + tree[3])) + ((tree[4] + tree[5]) + (tree[6] + tree[7]))
    assert O.dot(x, x) == s
    assert abs(O.dot(x, x) - sum(leaf)) < 1e-9


def sum_serial(v):
    s = 0.0
    for e in v:
        s += e * e
    return s


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmv_matches_dense_brute_force(seed):
    rng = np.random.default_rng(seed)
    m, n = 17, 23
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.3)
    rp, ci, v = [0], [], []
    for r in range(m):
        for c in range(n):
            if A[r, c] != 0:
                ci.append(c)
                v.append(A[r, c])
        rp.append(len(v))
    M = O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=[0] * n, col_lo=[0] * n, col_hi=[0] * n,
                     row_lo=[0] * m, row_hi=[0] * m))
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    np.testing.assert_allclose(O.spmv(M, x), A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.spmv_t(M, y), A.T @ y, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("name", ["ill_scaled_4x4", "c0_200x500_s0"])
@pytest.mark.parametrize("flags", [(1, 1), (1, 0), (0, 1)])
def test_scaling_bitwise_vs_reference(golden, name, flags):
    sc = golden["scaling"]
    if name == "ill_scaled_4x4":
        M = model(sc["ill_scaled_4x4_model"])
    else:
        import sovereign_amd as S
        M = O.Model(S.LP.random(200, 500, 5, 0).arrays())
    rs, cs = O.scaling(M, ruiz=flags[0], pc=flags[1])
    g = sc[f"{name}_r{flags[0]}p{flags[1]}"]
    np.testing.assert_array_equal(rs, g["row_scale"])
    np.testing.assert_array_equal(cs, g["col_scale"])


def test_scaling_invariants_of_reference_tests(golden):
    # reference tests/test_scaling.cpp:35-86
    M = model(golden["scaling"]["ill_scaled_4x4_model"])
    A = M.dense()
    rs, cs = O.scaling(M, ruiz=True, pc=False)
    S = np.abs(rs[:, None] * A * cs[None, :])
    rn = S.max(axis=1)
    cn = S.max(axis=0)
    assert rn.max() / rn.min() < 1.5 and cn.max() / cn.min() < 1.5
    nz = np.abs(A[A != 0])
    nzs = S[A != 0]
    assert nzs.max() / nzs.min() < (nz.max() / nz.min()) / 100
    empty = O.Model(dict(m=3, n=3, row_ptr=[0, 1, 1, 2], col_idx=[0, 2], val=[5.0, 7.0], c=[0] * 3, col_lo=[0] * 3,
                         col_hi=[0] * 3, row_lo=[0] * 3, row_hi=[0] * 3))
    rs, cs = O.scaling(empty)
    assert rs[1] == 1.0 and cs[1] == 1.0 and np.all(rs > 0) and np.all(cs > 0)
    rs, cs = O.scaling(M, ruiz=False, pc=False)
    assert np.all(rs == 1.0) and np.all(cs == 1.0)


def test_pock_chambolle_closed_form():
    # alpha = 1: after Ruiz-off, the factors are 1/sqrt(row / column L1 norms)
    rng = np.random.default_rng(3)
    m, n = 6, 9
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    A[0, 0] = 1.5
    rp, ci, v = [0], [], []
    for r in range(m):
        for c in range(n):
            if A[r, c] != 0:
                ci.append(c)
                v.append(A[r, c])
        rp.append(len(v))
    M = O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=[0] * n, col_lo=[0] * n, col_hi=[0] * n,
                     row_lo=[0] * m, row_hi=[0] * m))
    rs, cs = O.scaling(M, ruiz=False, pc=True)
    r1, c1 = np.abs(A).sum(axis=1), np.abs(A).sum(axis=0)
    np.testing.assert_allclose(rs, np.where(r1 > 0, 1 / np.sqrt(np.where(r1 > 0, r1, 1)), 1), rtol=1e-15)
    np.testing.assert_allclose(cs, np.where(c1 > 0, 1 / np.sqrt(np.where(c1 > 0, c1, 1)), 1), rtol=1e-15)


def test_operator_norm_closed_form():
    rng = np.random.default_rng(5)
    m, n = 8, 12
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    rp, ci, v = [0], [], []
    for r in range(m):
        for c in range(n):
            if A[r, c] != 0:
                ci.append(c)
                v.append(A[r, c])
        rp.append(len(v))
    M = O.Model(dict(m=m, n=n, row_ptr=rp, col_idx=ci, val=v, c=[0] * n, col_lo=[0] * n, col_hi=[0] * n,
                     row_lo=[0] * m, row_hi=[0] * m))
    K = np.hstack([-A, np.eye(m)])
    exact = np.linalg.svd(K, compute_uv=False)[0]
    assert abs(exact - math.sqrt(np.linalg.svd(A, compute_uv=False)[0] ** 2 + 1)) < 1e-12
    est = O.operator_norm(M, 400)
    assert abs(est - exact) / exact < 1e-6


def tiny_optimum():
    # max x + y s.t. x + 2y <= 4, 4x + 2y <= 12: optimum (8/3, 2/3), duals (1/3, 1/6)
    M = O.Model(dict(m=2, n=2, row_ptr=[0, 2, 4], col_idx=[0, 1, 0, 1], val=[1.0, 2.0, 4.0, 2.0], c=[1.0, 1.0],
                     col_lo=[0.0, 0.0], col_hi=[INF, INF], row_lo=[-INF, -INF], row_hi=[4.0, 12.0], maximize=True))
    return M, np.array([8 / 3, 2 / 3]), np.array([1 / 3, 1 / 6])


def test_residuals_vanish_at_known_optimum():
    M, x, y = tiny_optimum()
    ores = O.original_space_residual(M, -1.0, [1, 1], [1, 1], x, -y, 1e-6)
    assert ores < 1e-15
    rel = O.relative_kkt(M, x, y)
    assert rel[0] < 1e-15 and rel[1] < 1e-15 and rel[2] < 1e-15
    assert abs(rel[3] - 10 / 3) < 1e-14 and abs(rel[4] - 10 / 3) < 1e-14
    # scaled-space residual with internal sign: c_int = -c, y_int = -y
    assert O.kkt_residual(M, [-1.0, -1.0], x, -y, 1e-6) < 1e-15


def test_residuals_see_a_planted_violation():
    M, x, y = tiny_optimum()
    xb = x + np.array([0.1, 0.0])          # row 2 activity 12.4 > 12, row 1 4.1 > 4
    rel = O.relative_kkt(M, xb, y)
    assert abs(rel[0] - math.sqrt(0.1 ** 2 + 0.4 ** 2) / (1 + math.sqrt(16 + 144))) < 1e-14
    ores = O.original_space_residual(M, -1.0, [1, 1], [1, 1], xb, -y, 1e-6)
    assert abs(ores - 0.4) < 1e-12
    yb = y + np.array([0.0, 0.05])          # still dual feasible: only the gap sees it
    rel = O.relative_kkt(M, x, yb)
    assert rel[1] == 0.0
    d = 4 / 3 + (1 / 6 + 0.05) * 12
    assert abs(rel[2] - abs(10 / 3 - d) / (1 + 10 / 3 + d)) < 1e-14
    yw = np.array([-0.1, 1 / 6])            # wrong sign on a <= row: a dual residual
    rel = O.relative_kkt(M, x, yw)
    rc = np.array([-1.0, -1.0]) + M.dense().T @ yw   # internal rc = c_int - A^T y_int, y_int = -y
    lam = np.maximum(rc, 0.0)
    assert abs(rel[1] - math.sqrt(0.1 ** 2 + np.sum((rc - lam) ** 2)) / (1 + math.sqrt(2))) < 1e-14
    # permutation invariance of the relative KKT
    Mp = O.Model(dict(m=2, n=2, row_ptr=[0, 2, 4], col_idx=[0, 1, 0, 1], val=[2.0, 4.0, 2.0, 1.0], c=[1.0, 1.0],
                      col_lo=[0.0, 0.0], col_hi=[INF, INF], row_lo=[-INF, -INF], row_hi=[12.0, 4.0], maximize=True))
    np.testing.assert_allclose(O.relative_kkt(Mp, xb[::-1], y[::-1])[:3], O.relative_kkt(M, xb, y)[:3], rtol=1e-14)


def test_farkas_and_cs_hand_cases():
    # x >= 2 and x <= 1 (benchmarks/smoke/tiny_infeasible.mps): y = (1, -1) proves it
    M = O.Model(dict(m=2, n=1, row_ptr=[0, 1, 2], col_idx=[0, 0], val=[1.0, 1.0], c=[1.0], col_lo=[0.0],
                     col_hi=[INF], row_lo=[2.0, -INF], row_hi=[INF, 1.0]))
    assert O.farkas_certifies(M, [1.0, -1.0])
    assert not O.farkas_certifies(M, [1.0, 1.0])
    assert not O.farkas_certifies(M, [0.0, 0.0])
    Mo, x, y = tiny_optimum()
    rc = np.array([1.0, 1.0]) - Mo.dense().T @ y
    assert O.cs_violation(Mo, x, y, rc, 1e-7) < 1e-15
    assert O.cs_violation(Mo, x * 0.5, y, rc, 1e-7) > 0.1


def test_primal_ray_hand_cases():
    # min -x s.t. x - y <= 5, x, y >= 0 (benchmarks/smoke/tiny_unbounded.mps): d = (1, 1) is a ray
    M = O.Model(dict(m=1, n=2, row_ptr=[0, 2], col_idx=[0, 1], val=[1.0, -1.0], c=[-1.0, 0.0], col_lo=[0.0, 0.0],
                     col_hi=[INF, INF], row_lo=[-INF], row_hi=[5.0]))
    assert O.primal_ray_certifies(M, [1.0, 1.0])
    assert O.primal_ray_certifies(M, [3.0, 7.0])          # any scale, A d <= 0 with room
    assert not O.primal_ray_certifies(M, [1.0, 0.0])      # A d = 1 leaves the row's recession cone
    assert not O.primal_ray_certifies(M, [0.0, 1.0])      # c'd = 0: not improving
    assert not O.primal_ray_certifies(M, [-1.0, -1.0])    # leaves x >= 0
    assert not O.primal_ray_certifies(M, [0.0, 0.0])
    assert not O.primal_ray_certifies(M, [np.nan, 1.0])
    # the sense flips the improving direction: max x has the same ray
    Mx = O.Model(dict(m=1, n=2, row_ptr=[0, 2], col_idx=[0, 1], val=[1.0, -1.0], c=[1.0, 0.0], col_lo=[0.0, 0.0],
                      col_hi=[INF, INF], row_lo=[-INF], row_hi=[5.0], maximize=True))
    assert O.primal_ray_certifies(Mx, [1.0, 1.0])
    # a bounded LP has none: min x + y, x + y >= 1, x, y >= 0
    Mb = O.Model(dict(m=1, n=2, row_ptr=[0, 2], col_idx=[0, 1], val=[1.0, 1.0], c=[1.0, 1.0], col_lo=[0.0, 0.0],
                      col_hi=[INF, INF], row_lo=[1.0], row_hi=[INF]))
    rng = np.random.default_rng(0)
    assert not any(O.primal_ray_certifies(Mb, rng.standard_normal(2)) for _ in range(200))


def check_same(res, g):
    assert res["status"] == g["status"]
    assert res["iterations"] == g["iterations"]
    assert res["restarts"] == g["restarts"]
    assert res["operator_norm"] == g["operator_norm"]
    if isinstance(g["objective"], float):
        assert res["objective"] == g["objective"]
    np.testing.assert_array_equal(res["trace"], g["trace"])
    if "x" in g:
        np.testing.assert_array_equal(res["x"], g["x"])
        np.testing.assert_array_equal(res["y"], g["y"])


def test_pdhg_bitwise_vs_reference_on_hand_lps(golden):
    for name, case in golden["lps"].items():
        M = model(case["model"])
        cap = 20000 if name.startswith("smoke_") else 1000000
        r = O.pdhg(M, period=case.get("period", 40), max_iterations=cap, trace_cap=100000)
        check_same(r, case["result"])


@pytest.mark.parametrize("key", ["c0_200x500_s0_tol0.0001", "c0_200x500_s1_tol0.0001", "c0_200x500_s2_tol0.0001"])
def test_pdhg_bitwise_vs_reference_on_random(golden, key):
    import sovereign_amd as S
    case = golden["random"][key]
    M = O.Model(S.LP.random(case["m"], case["n"], case["k"], case["seed"]).arrays())
    r = O.pdhg(M, tolerance=case["tolerance"], max_iterations=case["max_iterations"], trace_cap=100000)
    check_same(r, case["result"])


def test_primal_weight_off_is_the_reference(golden):
    # the extension at omega == 1 without updates must not move a bit
    case = golden["lps"]["two_variable_optimum_period5"]
    M = model(case["model"])
    a = O.pdhg(M, period=5, primal_weight=False, trace_cap=100000)
    check_same(a, case["result"])
    b = O.pdhg(M, period=5, primal_weight=True, omega=1.0)
    assert b["status"] == "optimal" and abs(b["objective"] - 10 / 3) < 5e-4
