"""Pin the NumPy proposal-network oracle (oracle/mlp.py) by central finite differences of its loss."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, 4481)
    x = rng.uniform(-1, 1, (37, 3))
    t = rng.normal(0, 1, 37)
    loss, g = mlp.loss_and_gradient(p, x, t)
    assert np.isclose(loss, np.mean((mlp.forward(p, x) - t) ** 2))
    idx = np.concatenate([np.arange(0, 192, 17), np.arange(192, 256, 9), rng.choice(np.arange(256, 4352), 60),
                          np.arange(4352, 4416, 7), np.arange(4416, 4481, 6), [4480]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss_and_gradient(p + e, x, t)[0] - mlp.loss_and_gradient(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
