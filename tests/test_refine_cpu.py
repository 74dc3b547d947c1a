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


def test_neighbourhood_min_max_brute_force():
    n = 4
    v = np.random.default_rng(2).normal(size=n ** 3)
    lo = refine.neighbourhood(v, n, np.minimum)
    hi = refine.neighbourhood(v, n, np.maximum)
    f = v.reshape(n, n, n)
    for (i3, i2, i1) in [(0, 0, 0), (3, 1, 2), (1, 3, 3)]:
        nb = [f[(i3 + a) % n, (i2 + b) % n, (i1 + c) % n] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
        idx = i1 + n * (i2 + n * i3)
        assert lo[idx] == min(nb) and hi[idx] == max(nb)


def test_point_spectral_features():
    # levels (eV) around mu = 0.1: adjacent spacings 0.3, 0.002, 0.5; only the pair with midpoint within
    # 0.25 eV of mu counts for the gap (0.002 at midpoint 0.101); nearest level 0.1 - 0.1 = 0.0 / 0.2 = 0
    e = np.array([[-0.2, 0.1, 0.102, 0.602], [-3.0, -2.0, 2.0, 3.0]])
    f = refine.point_spectral_features(e, mu=0.1, scale=0.01)
    assert np.isclose(f[0, 0], np.log10(0.002)) and np.isclose(f[0, 1], 0.0)
    # no pair near mu: the gap feature saturates at the window; distance 1.9 eV / 0.01 eV
    assert np.isclose(f[1, 0], np.log10(refine.GAP_WINDOW_EV)) and np.isclose(f[1, 1], 190.0)
    # order of the input levels does not matter
    g = refine.point_spectral_features(e[:, ::-1], mu=0.1, scale=0.01)
    assert np.array_equal(f, g)


def test_cell_features_are_bounded_and_monotone():
    n = 3
    rng = np.random.default_rng(3)
    e = np.sort(rng.uniform(-2, 2, (n ** 3, 6)), axis=1)
    norms = 10.0 ** rng.uniform(-12, 0, n ** 3)
    x = refine.cell_features(e, norms, n, 0.0, 0.005, norms.max())
    assert x.shape == (n ** 3, 3) and x.dtype == np.float32
    assert np.all(x >= -1.0) and np.all(x <= 1.0)
    # every cell of a 3^3 periodic mesh sees all 27 points: the features are the global extrema
    assert np.allclose(x[:, 2], 1.0)
    assert np.allclose(x[:, 0], x[0, 0]) and np.allclose(x[:, 1], x[0, 1])
