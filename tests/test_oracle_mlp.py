"""Pin the NumPy proposal-network oracle (oracle/mlp.py) by central finite differences of its loss."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    H, IN = mlp.H, mlp.IN
    b1, w2, b2, w3, b3 = IN * H, IN * H + H, IN * H + H + H * H, IN * H + 2 * H + H * H, IN * H + 3 * H + H * H
    p = rng.normal(0, 0.3, b3 + 1)
    x = rng.uniform(-1, 1, (37, IN))
    t = rng.normal(0, 1, 37)
    loss, g = mlp.loss_and_gradient(p, x, t)
    assert g.shape == p.shape == (4801,)
    assert np.isclose(loss, np.mean((mlp.forward(p, x) - t) ** 2))
    idx = np.concatenate([np.arange(0, b1, 23), np.arange(b1, w2, 9), rng.choice(np.arange(w2, b2), 60),
                          np.arange(b2, w3, 7), np.arange(w3, b3, 6), [b3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss_and_gradient(p + e, x, t)[0] - mlp.loss_and_gradient(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
