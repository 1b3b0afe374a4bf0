"""The warm-start network's device kernels (BF16 MFMA products, GELU, Adam)
against the CPU FP64 reference of tests/mlp_ref.py, which rounds to BF16
exactly where the kernels do, so the two differ only by FP32 vs FP64
accumulation.  The reference's own gradients are pinned by finite
differences (tests/test_mlp_ref.py)."""
import numpy as np
import pytest

import mlp_ref as R
from sovereign_amd.device import Ctx
from sovereign_amd.learned import Mlp, init_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    c = Ctx(0)
    yield c
    c.close()


def rel_err(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("widths,batch", [([316, 512, 512, 512, 301], 64), ([37, 70, 5], 13), ([8, 3], 1)])
def test_forward_and_backward_match_the_cpu_reference(ctx, widths, batch):
    rng = np.random.default_rng(len(widths) + batch)
    net = Mlp(ctx, widths, max(batch, 64))
    p = init_params(widths, rng)
    p += (0.01 * rng.standard_normal(p.shape)).astype(np.float32)   # nonzero biases too
    net.set_params(p)
    X = rng.standard_normal((batch, widths[0])).astype(np.float32)
    dY = rng.standard_normal((batch, widths[-1])).astype(np.float32)
    y = net.forward(X)
    g = net.backward(dY)
    y_ref, cache = R.forward(p.astype(np.float64), widths, X.astype(np.float64), emulate=True)
    g_ref = R.backward(p.astype(np.float64), widths, cache, dY.astype(np.float64), emulate=True)
    print(widths, batch, "forward rel err", rel_err(y, y_ref), "gradient rel err", rel_err(g, g_ref), flush=True)
    # same BF16 roundings; FP32 accumulation vs FP64 (and an occasional
    # one-ulp BF16 flip of an activation that FP32 puts on the other side of a tie)
    assert rel_err(y, y_ref) < 2e-3
    assert rel_err(g, g_ref) < 1e-2
    # every layer's block of the gradient, not just the total
    off = 0
    for l in range(1, len(widths)):
        n = widths[l] * widths[l - 1] + widths[l]
        assert rel_err(g[off:off + n], g_ref[off:off + n]) < 2e-2, l
        off += n
    # and the unrounded FP64 network is within BF16 reach
    y64, _ = R.forward(p.astype(np.float64), widths, X.astype(np.float64), emulate=False)
    assert rel_err(y, y64) < 3e-2
    net.close()


def test_adam_training_fits_a_smooth_map(ctx):
    rng = np.random.default_rng(7)
    widths = [16, 128, 128, 4]
    M = rng.standard_normal((16, 4)) / 4.0
    X = rng.standard_normal((1024, 16)).astype(np.float32)
    T = np.tanh(X @ M).astype(np.float32)
    net = Mlp(ctx, widths, 256)
    net.set_params(init_params(widths, rng))
    first = None
    for step in range(300):
        sel = rng.integers(0, len(X), 256)
        loss = net.train_step(X[sel], T[sel], 3e-3)
        first = first if first is not None else loss
    final = np.mean((net.predict(X) - T) ** 2)
    print("training loss", first, "->", final, flush=True)
    assert final < first / 20
    # the device Adam matches one host Adam step from the same state
    p0, _ = net.params()
    net2 = Mlp(ctx, widths, 256)
    net2.set_params(p0)
    net2.train_step(X[:256], T[:256], 1e-3)
    p1, g = net2.params()
    m = 0.1 * g.astype(np.float64)
    v = 0.001 * g.astype(np.float64) ** 2
    want = p0 - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.max(np.abs(p1 - want)) < 1e-6
    net.close()
    net2.close()
