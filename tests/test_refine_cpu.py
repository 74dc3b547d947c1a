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


def test_band_slopes_closed_form():
    # E(k) = cos(2 pi k1) + 0.5 cos(2 pi k3): central differences on the n^3 mesh are
    # -sin(2 pi i/n) sin(2 pi/n) along axis 1 and -0.5 sin(2 pi i3/n) sin(2 pi/n) along axis 3
    n = 8
    k = refine.mesh_points(n)
    e = (np.cos(2 * np.pi * k[:, 0]) + 0.5 * np.cos(2 * np.pi * k[:, 2]))[:, None]
    got = refine.band_slopes(e, n)[:, 0]
    s = np.sin(2 * np.pi / n)
    want = np.hypot(np.sin(2 * np.pi * k[:, 0]) * s, 0.5 * np.sin(2 * np.pi * k[:, 2]) * s)
    assert np.allclose(got, want, atol=1e-14)


def test_fermi_surface_features_of_a_cosine_band():
    # E = cos(2 pi k1) - mu with mu = 0 on n = 8: the Fermi surface is the planes k1 = 1/4, 3/4
    n, scale = 8, 0.01
    k = refine.mesh_points(n)
    e = np.cos(2 * np.pi * k[:, 0])[:, None]
    f = refine.fermi_surface_features(e, n, 0.0, scale)
    i1 = np.round(k[:, 0] * n).astype(int)
    slope = np.sin(2 * np.pi / n) * np.abs(np.sin(2 * np.pi * i1 / n))
    on = (i1 == 2) | (i1 == 6)
    assert np.allclose(f[on, 0], 0.0, atol=1e-12)                       # the surface passes the centre
    assert np.allclose(f[on, 1], scale / slope[on])                      # shell = scale / slope
    assert np.all(f[on, 2] > 0.0)                                        # straddled inside the cell
    far = (i1 == 0) | (i1 == 4)                                          # flat extrema: slope 0, capped
    assert np.all(f[far, 0] == refine.SURFACE_CAP)
    mid = (i1 == 1)
    assert np.allclose(f[mid, 0], np.abs(np.cos(2 * np.pi / n)) / slope[mid])
    # a cell whose interpolated level stays on one side of mu has a negative depth
    assert np.all(f[(i1 == 0) | (i1 == 4), 2] < 0.0)


def test_crossing_depth_brute_force():
    n = 4
    rng = np.random.default_rng(5)
    e = np.sort(rng.uniform(-1, 1, (n ** 3, 3)), axis=1)
    got = refine.fermi_surface_features(e, n, 0.1, 0.01)[:, 2]
    slope = refine.band_slopes(e, n) + 1e-12
    f = e.reshape(n, n, n, 3) - 0.1                                      # [i3, i2, i1, band]
    for idx in (0, 17, 63):
        i1, i2, i3 = idx % n, (idx // n) % n, idx // (n * n)
        vals = [f[i3, i2, i1]]
        for s3 in (-1, 1):
            for s2 in (-1, 1):
                for s1 in (-1, 1):
                    pts = [f[(i3 + a * s3) % n, (i2 + b * s2) % n, (i1 + c * s1) % n]
                           for a in (0, 1) for b in (0, 1) for c in (0, 1)]
                    vals.append(np.mean(pts, axis=0))
        vals = np.array(vals)
        depth = (np.minimum(vals.max(axis=0), -vals.min(axis=0)) / slope[idx]).max()
        assert np.isclose(got[idx], np.clip(depth, -5.0, 2.0))


def test_cell_features_shape_and_order():
    n = 4
    rng = np.random.default_rng(3)
    e = np.sort(rng.uniform(-2, 2, (n ** 3, 6)), axis=1)
    norms = 10.0 ** rng.uniform(-14, 0, n ** 3)
    x = refine.cell_features(e, norms, n, 0.0, 0.005, norms.max())
    assert x.shape == (n ** 3, len(refine.FEATURES)) and np.all(np.isfinite(x))
    ln = np.maximum(np.log10(norms / norms.max()), refine.NORM_FLOOR)
    assert np.allclose(x[:, 5], ln) and np.allclose(x[:, 4], refine.neighbourhood(ln, n, np.maximum))
    pf = refine.point_spectral_features(e, 0.0, 0.005)
    assert np.allclose(x[:, 2], pf[:, 0]) and np.allclose(x[:, 3], refine.neighbourhood(pf[:, 0], n, np.minimum))


def test_error_curve_and_cells_needed_brute_force():
    rng = np.random.default_rng(7)
    base = rng.normal(size=(9, 1, 3, 3, 3))
    refined = base + rng.normal(size=base.shape) * np.logspace(0, -8, 9)[:, None, None, None, None]
    order = rng.permutation(9)
    curve = refine.error_curve(order, base, refined)
    total = refined.sum(axis=0)
    for b in range(10):
        est = base.copy()
        est[order[:b]] = refined[order[:b]]
        assert np.isclose(curve[b], np.linalg.norm(est.sum(axis=0) - total) / np.linalg.norm(total), rtol=1e-9,
                          atol=1e-15)
    assert curve[-1] < 1e-14
    for tol in (1e-1, 1e-4, 1e-9):
        b = refine.cells_needed(curve, tol)
        assert np.all(curve[b:] < tol) and (b == 0 or curve[b - 1] >= tol)


def test_variance_order_ranks_the_spike_first():
    n = 6
    base = np.full((n ** 3, 1, 3, 3, 3), 1e-3)
    base[100] *= 1e4
    order = refine.variance_order(n, base)
    # the 27 cells whose neighbourhood holds the spike come first
    hot = set(order[:27].tolist())
    f = np.zeros(n ** 3)
    f[100] = 1.0
    assert hot == set(np.flatnonzero(refine.neighbourhood(f, n, np.maximum) > 0).tolist())
