"""The live batch session (sov::PdhgBatchSession through sov_batch_session_*):
related LPs re-solved on the device with their data overwritten in place.

  * node bounding: eight copies of one LP, each child tightening one column
    bound around the parent's solution, re-solved from the parent's point;
  * right-hand sides and objectives overwritten in place, and an instance
    started from a neighbour's last solution;
  * a re-solve with nothing changed starts at the optimum.

Every objective is held to 1e-6 relative of the exact optimum of the LP as
the session holds it (HiGHS through scipy, the solver the reference's own
PDHG-vs-oracle check uses; tests/golden/make_exact_optima.highs_optimum),
and every relative KKT to the tolerance.  The reference has no counterpart
(its build map plans the engine, docs/build-map.md:266-276)."""
import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from golden.make_exact_optima import highs_optimum

pytestmark = pytest.mark.gpu

TOL = 1e-6


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def with_data(a, **kw):
    """The LP of arrays `a` with some vectors replaced."""
    d = dict(a, **kw)
    return S.LP.from_csr(d["m"], d["n"], d["row_ptr"], d["col_idx"], d["val"], d["c"], d["col_lo"], d["col_hi"],
                         d["row_lo"], d["row_hi"], d["maximize"], d["obj_const"])


def check_against_highs(lp, r, label):
    a = lp.arrays()
    st, exact, _ = highs_optimum(a)
    host = O.relative_kkt(O.Model(a), r.x, r.y)[:3]
    print(f"{label}: {r.status} it={r.iterations} obj {r.objective:.10g} exact {exact:.10g} host rel {host}",
          flush=True)
    assert st == 0 and r.status == "optimal", (label, r.status)
    assert max(r.relative_kkt) <= TOL * 1.0001 and max(host) <= TOL * 1.001
    assert rel(r.objective, exact) < 1e-6, label
    return exact


def test_node_bounding_re_solves_children_from_the_parent(gpu):
    base = S.LP.random(300, 800, 5, 3)
    a = base.arrays()
    B = 8
    sess = S.BatchSession([base] * B, tolerance=TOL, precision="f32")
    parent, _ = sess.solve()
    root = check_against_highs(base, parent[0], "root")
    for r in parent[1:]:
        assert r.status == "optimal" and rel(r.objective, root) < 1e-6
    # children: one column bound each, around the root's point
    x = parent[0].x
    inner = np.argsort(-np.minimum(x - a["col_lo"], a["col_hi"] - x))[:B]
    kids = []
    for k, j in enumerate(inner):
        lo, hi = a["col_lo"].copy(), a["col_hi"].copy()
        if k % 2 == 0:
            hi[j] = 0.5 * x[j]                          # down branch
        else:
            lo[j] = x[j] + 0.5 * (a["col_hi"][j] - x[j])   # up branch
        sess.set_col_bounds(k, lo, hi)
        kids.append(with_data(a, col_lo=lo, col_hi=hi))
    warm, _ = sess.solve()
    for k in range(B):
        exact = check_against_highs(kids[k], warm[k], f"child {k}")
        assert exact >= root - 1e-9 * (1 + abs(root))   # a tighter box cannot lower a minimum
    cold, _ = S.solve_batch(kids, tolerance=TOL, precision="f32")
    wi, ci = np.median([r.iterations for r in warm]), np.median([r.iterations for r in cold])
    print(f"children: median iterations warm {wi:.0f} vs cold {ci:.0f}", flush=True)
    for k in range(B):
        assert cold[k].status == "optimal" and rel(cold[k].objective, warm[k].objective) < 2e-6
    assert wi < ci
    sess.close()


def test_rhs_and_objective_overwrites_and_a_neighbour_start(gpu):
    base = S.LP.transportation(20, 30, 0)
    a = base.arrays()
    sess = S.BatchSession([base] * 4, tolerance=TOL, precision="f64")
    first, _ = sess.solve()
    for r in first:
        check_against_highs(base, r, "first")
    # instance 1 and 3: supplies and demands 5 % larger (still balanced);
    # instance 2: perturbed costs; instance 3 starts from instance 0's solution
    rlo, rhi = a["row_lo"] * 1.05, a["row_hi"] * 1.05
    c2 = a["c"] * (1.0 + 0.1 * np.sin(np.arange(a["n"])))
    sess.set_row_bounds(1, rlo, rhi)
    sess.set_row_bounds(3, rlo, rhi)
    sess.set_objective(2, c2)
    sess.warm_start_from(3, 0)
    second, _ = sess.solve()
    check_against_highs(base, second[0], "unchanged")
    check_against_highs(with_data(a, row_lo=rlo, row_hi=rhi), second[1], "rhs, own start")
    check_against_highs(with_data(a, c=c2), second[2], "objective")
    check_against_highs(with_data(a, row_lo=rlo, row_hi=rhi), second[3], "rhs, neighbour start")
    # the unchanged instance restarts at its optimum: it stops within a few checks
    assert second[0].iterations <= max(120, first[0].iterations // 10), (second[0].iterations, first[0].iterations)
    # an update that breaks the instance is refused and changes nothing
    bad_lo = a["col_lo"].copy()
    bad_lo[0] = 5.0
    bad_hi = a["col_hi"].copy()
    bad_hi[0] = 1.0
    with pytest.raises(RuntimeError):
        sess.set_col_bounds(0, bad_lo, bad_hi)
    third, _ = sess.solve()
    check_against_highs(base, third[0], "after a refused update")
    sess.close()
