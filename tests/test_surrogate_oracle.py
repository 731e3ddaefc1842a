"""The CPU restatement of the surrogate network (oracle/mlp.py), pinned by central finite
differences of its loss and by the BF16 rounding rule (CPU only)."""

import numpy as np

from oracle import mlp


def test_gradient_against_finite_differences():
    p = mlp.init(3)
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(16, 4))
    Y = rng.normal(size=(16, 4))
    loss, g = mlp.loss_and_gradient(p, X, Y)
    assert abs(loss - ((mlp.forward(p, X) - Y) ** 2).mean()) < 1e-14
    idx = rng.choice(p.size, 200, replace=False)
    idx = np.concatenate([idx, [0, mlp.SIZE - 1, 512, 640, 33664]])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = mlp.loss_and_gradient(q, X, Y)
        q[i] -= 2 * eps
        lm, _ = mlp.loss_and_gradient(q, X, Y)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), i


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3])
    r = mlp.to_bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie, even
    assert r[2] == 1.0 + 2 ** -6  # tie, even upwards
    assert r[3] == 1.0            # below half
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) / 3.0e-3 < 2 ** -8
