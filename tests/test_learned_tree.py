"""The spanning-tree recovery of sovereign_amd.learned (no GPU): from the
optimal basis of a balanced transportation LP, transport_point must return
the optimal vertex and a dual that prices every basic arc at zero reduced
cost and no arc below zero; from the optimal dual, transport_primal must
return the optimal vertex.  The optimum comes from HiGHS (scipy's linprog),
the exact solver the reference's own PDHG-vs-oracle check uses."""
import numpy as np
import pytest

import sovereign_amd as S
from sovereign_amd import learned as W


def highs(a):
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    A = csr_matrix((a["val"][: a["nnz"]], a["col_idx"][: a["nnz"]], a["row_ptr"]), shape=(a["m"], a["n"]))
    r = linprog(a["c"], A_eq=A, b_eq=a["row_lo"], bounds=(0, None), method="highs")
    assert r.status == 0
    return A, r.x, r.eqlin.marginals, r.fun


@pytest.mark.parametrize("index", [0, 1, 2])
def test_tree_of_the_optimal_basis_is_the_optimal_vertex(index):
    fam = S.Family(20, 40, 4, 5)
    lp, _ = fam.instance(3, index)
    a = lp.arrays()
    A, xs, ys, obj = highs(a)
    # the optimal flows rank every basic arc first
    x, y = W.transport_point(a, xs, 20, 40)
    assert np.allclose(x, xs, rtol=0, atol=1e-9 * xs.max())
    assert np.max(np.abs(A @ x - a["row_lo"])) <= 1e-9 * a["row_lo"].max()
    d = a["c"] - A.T @ y
    assert d.min() >= -1e-9 * np.abs(a["c"]).max()                 # dual feasible
    assert np.max(np.abs(d[xs > 1e-9])) <= 1e-9 * np.abs(a["c"]).max()   # basic arcs priced at zero
    assert abs(a["c"] @ x - obj) <= 1e-12 * abs(obj)
    assert abs(a["row_lo"] @ y - obj) <= 1e-9 * abs(obj)             # strong duality
    # the optimal dual's spanning tree under reduced cost is the same vertex
    xp = W.transport_primal(a, ys, 20, 40)
    assert np.allclose(xp, xs, rtol=0, atol=1e-9 * xs.max())


def test_a_wrong_tree_gives_a_worse_point():
    fam = S.Family(20, 40, 4, 5)
    lp, _ = fam.instance(3, 0)
    a = lp.arrays()
    A, xs, ys, obj = highs(a)
    score = xs.copy()
    score[np.argmax(xs)] = -1.0          # the largest basic arc dropped from the tree
    x, y = W.transport_point(a, score, 20, 40)
    assert x[np.argmax(xs)] == 0.0
    assert x.min() >= 0.0
    # clipped flows: either infeasible or feasible but no better than optimal
    assert np.max(np.abs(A @ x - a["row_lo"])) > 1e-6 or a["c"] @ x >= obj * (1 - 1e-12)
