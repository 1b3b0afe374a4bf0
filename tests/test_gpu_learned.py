"""The learned warm start end to end on a small transportation family
(20 sources x 40 sinks, 4 cost parameters): cold GPU solves make the data,
the network trains on the GPU with this project's kernels, and on held-out
instances its warm start (dual + primal weight) must cut the median
iterations to 1e-6 relative KKT against the cold start, reach the same
objective (1e-6 relative of the exact optimum from HiGHS), and keep every
instance optimal.  Also reported: the gain of a constant primal weight (the
training set's mean, no network), the baseline the network has to match.
A start at the optimal vertex (the spanning tree of the optimal basis,
learned.transport_point) must stop at the first check."""
import numpy as np
import pytest

import oracle_ffi as O
import sovereign_amd as S
from sovereign_amd.device import Ctx
from sovereign_amd import learned as W
from test_learned_tree import highs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    yield c
    c.close()


def test_trained_network_cuts_iterations(ctx):
    fam = S.Family(20, 40, 4, 5)
    X, Y, LW, its = W.dataset(fam, 1, 400, chunk=400)
    assert len(X) == 400 and X.shape[1] == 20 + 40 + 4
    net, xn, yn, losses = W.train(ctx, X, np.column_stack([Y, LW]), hidden=(128, 128, 128), epochs=80, batch=64,
                                  lr=2e-3)
    assert losses[-1] < 0.5 * losses[0]
    rep = W.evaluate(ctx, fam, net, xn, yn, 120, seed=2, const_logw=float(LW.mean()), chunk=120)
    print(rep, flush=True)
    for k in ("cold", "constant", "learned"):
        assert rep[k]["optimal"] == 120
    assert rep["median_iteration_ratio_cold_over_learned"] > 1.3
    assert rep["learned"]["median_iterations"] < 0.8 * rep["cold"]["median_iterations"]
    assert rep["max_objective_rel_diff"] < 1e-4
    # warm-started answers against the exact optimum (HiGHS) on a few
    # instances, and the same instances started at their optimal vertex (the
    # spanning tree of the optimal basis): a start at an optimum stops at the
    # first check
    for i in range(3):
        lp, f = fam.instance(2, i)
        a = lp.arrays()
        _, xs, _, exact = highs(a)
        pr = W.predict(net, xn, yn, f[None, :])[0]
        kw = dict(backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative", restarts="adaptive")
        r = S.solve(lp, warm_start_dual=pr[:60], primal_weight=True, initial_primal_weight=float(np.exp(pr[60])),
                    primal_weight_smoothing=0.0, **kw)
        host = O.relative_kkt(O.Model(a), r.x, r.y)[:3]
        print(f"instance {i}: warm {r.status} it={r.iterations} obj {r.objective:.12g} exact {exact:.12g} "
              f"host rel {host}", flush=True)
        assert r.status == "optimal" and max(host) <= 1e-6 * 1.001
        assert abs(r.objective - exact) <= 1e-6 * (1.0 + abs(exact))
        xv, yv = W.transport_point(a, xs, 20, 40)
        v = S.solve(lp, warm_start_primal=xv, warm_start_dual=yv, **kw)
        assert v.status == "optimal" and v.iterations == 40, (v.status, v.iterations)
        assert abs(v.objective - exact) <= 1e-9 * (1.0 + abs(exact))
    net.close()
