"""Host-side bookkeeping of the learned refinement (no GPU): mesh order, sub-cell centres that
reproduce the uniform refined mesh, and the neighbourhood variance."""
import numpy as np

from wanniernlqg_amd import refine


def test_mesh_points_axis1_fastest():
    k = refine.mesh_points(3)
    assert np.allclose(k[1], [1 / 3, 0, 0]) and np.allclose(k[3], [0, 1 / 3, 0]) and np.allclose(k[9], [0, 0, 1 / 3])


def test_full_refinement_is_the_uniform_mesh():
    n, r = 4, 3
    sub = refine.subcell_points(np.arange(n ** 3), n, r)
    dense = refine.mesh_points(n * r)
    key = lambda a: set(map(tuple, np.round(np.mod(a, 1.0) * n * r).astype(int).tolist()))
    assert sub.shape == dense.shape
    assert key(sub) == key(dense)
    # the centre of every cell is its middle sub-cell centre
    mid = sub.reshape(n ** 3, r ** 3, 3)[:, (r ** 3) // 2]
    assert np.allclose(mid, refine.mesh_points(n))


def test_local_log_variance_brute_force():
    n = 5
    rng = np.random.default_rng(1)
    v = rng.normal(size=n ** 3)
    got = refine.local_log_variance(v, n)
    f = v.reshape(n, n, n)
    for (i3, i2, i1) in [(0, 0, 0), (4, 2, 1), (2, 4, 4)]:
        nb = [f[(i3 + a) % n, (i2 + b) % n, (i1 + c) % n] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
        assert np.isclose(got[i1 + n * (i2 + n * i3)], np.log10(np.var(nb) + 1e-30))
