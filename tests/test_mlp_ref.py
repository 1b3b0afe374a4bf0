"""The CPU reference network itself, pinned without the device: its
gradients against central finite differences of its own FP64 forward pass,
BF16 rounding against hand cases, GELU against its closed forms."""
import numpy as np

import mlp_ref as R


def test_bf16_rounding_hand_cases():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert R.bf16(1.0 + 2.0 ** -8) == 1.0
    assert R.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert R.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7      # representable
    assert R.bf16(-3.0) == -3.0 and R.bf16(0.0) == 0.0
    assert np.isnan(R.bf16(np.nan))
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.max(np.abs(R.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_gelu_closed_forms():
    assert R.gelu(0.0) == 0.0
    assert abs(R.gelu(1.0) - 0.8413447460685429) < 1e-15          # Phi(1)
    x = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (R.gelu(x + h) - R.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - R.gelu_grad(x))) < 1e-8


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    widths = [5, 7, 6, 3]
    p = R.init_params(widths, rng) + 0.05 * rng.standard_normal(sum(
        widths[l] * widths[l - 1] + widths[l] for l in range(1, len(widths))))
    X = rng.standard_normal((4, 5))
    T = rng.standard_normal((4, 3))

    def loss(q):
        y, _ = R.forward(q, widths, X, emulate=False)
        return 0.5 * np.sum((y - T) ** 2)

    y, cache = R.forward(p, widths, X, emulate=False)
    g = R.backward(p, widths, cache, y - T, emulate=False)
    fd = np.zeros_like(p)
    h = 1e-6
    for k in range(len(p)):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (loss(p + e) - loss(p - e)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
