"""Pin the BCD / AHC oracle (oracle/geometry.py) by what the mathematics fixes.

* Band Berry curvature == Berry phase of a small k-space loop / its area (gauge-invariant link
  products of the eigenvectors plus the loop integral of A_H,nn), with and without a position
  operator.  A wrong sign of the Wannier curvature, a transposed connection product or a dropped
  commutator fails it.
* AHC of decoupled Qi-Wu-Zhang layers == -(e^2/h) C / c, with the Chern number C from an
  independent Fukui-Hatsugai-Suzuki lattice computation (sign and every unit factor).
* -f'(E) against the reference's own closed-form weight values and a finite-difference
  derivative of the Fermi function.
* BCD vanishes with inversion symmetry; the Matrix AHC integrand equals the occupied sum of
  band curvatures; the spectral isolation check fires on exact degeneracies.
"""
import math

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kloop
from wanniernlqg_amd.model import TightBindingModel, synthetic_model

HBAR = G.RESPONSE_HBAR_JS
E = G.RESPONSE_CHARGE_C


def cart_to_frac(model, kappa):
    return model.lattice @ kappa / (2 * math.pi)


def _eig_at(model, kappa):
    k = cart_to_frac(model, np.asarray(kappa, dtype=np.float64))[None]
    d = G.geometry_data(model, k, 0.0)
    return d


def _loop_phase(model, kappa, a, b, h, band, sub=8):
    """Berry phase of band around the counter-clockwise square of side h in the (a, b) plane:
    -Im log of the link product (gauge part) plus the Simpson loop integral of A_H,nn."""
    ea = np.zeros(3)
    eb = np.zeros(3)
    ea[a] = h
    eb[b] = h
    c0 = np.asarray(kappa) - 0.5 * ea - 0.5 * eb
    corners = [c0, c0 + ea, c0 + ea + eb, c0 + eb, c0]
    pts = []
    for s in range(4):
        for t in range(sub):
            pts.append(corners[s] + (corners[s + 1] - corners[s]) * t / sub)
    pts.append(c0)
    pts = np.array(pts)
    mids = 0.5 * (pts[:-1] + pts[1:])
    allk = np.concatenate([pts, mids])
    d = G.geometry_data(model, np.array([cart_to_frac(model, p) for p in allk]), 0.0)
    npts = pts.shape[0]
    u = d.eigenvectors[:npts, :, band]
    prod = 1.0 + 0j
    for i in range(npts - 1):
        prod *= np.vdot(u[i], u[i + 1])
    phase = -np.angle(prod)
    adiag = np.real(d.internal_connection[:, :, band, band])            # [K, 3]
    for i in range(npts - 1):
        dk = pts[i + 1] - pts[i]
        simpson = (adiag[i] + 4 * adiag[npts + i] + adiag[i + 1]) / 6.0
        phase += simpson @ dk
    return phase


@pytest.mark.parametrize("with_position", [False, True])
def test_band_curvature_equals_loop_berry_phase(with_position):
    m = synthetic_model(4, 1, 31, position_scale=0.15 if with_position else None, lattice_constant=3.0)
    kappa = np.array([0.31, -0.47, 0.22])
    d = _eig_at(m, kappa)
    assert np.min(np.diff(d.energies[0])) > 0.05
    omega = G.band_berry_curvature(d)[0]                                 # [n, a, b]
    h = 0.02
    for band in range(4):
        for a, b in ((0, 1), (1, 2), (0, 2)):
            # loop phase / area = Omega + O(h^2) + O(h^4): two Richardson levels over h, h/2, h/4
            e = [_loop_phase(m, kappa, a, b, h / s, band) / (h / s) ** 2 for s in (1, 2, 4)]
            r1 = [(4 * e[i + 1] - e[i]) / 3 for i in range(2)]
            est = (16 * r1[1] - r1[0]) / 15
            scale = max(1.0, np.max(np.abs(omega[band])))
            assert abs(omega[band, a, b] - est) < 1e-7 * scale, (band, a, b, omega[band, a, b], est)
            assert abs(omega[band, a, b] + omega[band, b, a]) < 1e-9 * scale
        assert np.all(np.abs(np.diagonal(omega[band])) < 1e-9)


def qwz_layers(mass: float, a: float = 2.5, c: float = 3.0) -> TightBindingModel:
    """H(k) = sin(k_x a) sx + sin(k_y a) sy + (mass + cos k_x a + cos k_y a) sz per layer, layers
    stacked along z with spacing c and no interlayer hopping; point orbitals (A = 0)."""
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]], dtype=complex)
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    rv = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]
    hr = np.zeros((5, 2, 2), dtype=complex)
    hr[0] = mass * sz
    # sin(k.R) = (e^{ikR} - e^{-ikR})/(2i), cos = (e^{ikR} + e^{-ikR})/2
    hr[1] = sx / 2j + sz / 2
    hr[2] = -sx / 2j + sz / 2
    hr[3] = sy / 2j + sz / 2
    hr[4] = -sy / 2j + sz / 2
    lat = np.diag([a, a, c])
    return TightBindingModel(lat, np.array(rv), np.ones(5, dtype=np.int32), hr, None)


def fhs_chern(model, n: int, band: int = 0) -> int:
    """Fukui-Hatsugai-Suzuki lattice Chern number of one band on an n x n k_z = 0 mesh."""
    k = np.array([[i / n, j / n, 0.0] for j in range(n) for i in range(n)])
    d = G.geometry_data(model, k, 0.0)
    u = d.eigenvectors[:, :, band].reshape(n, n, -1)       # [j, i, orbital]
    total = 0.0
    for j in range(n):
        for i in range(n):
            u0 = u[j, i]
            u1 = u[j, (i + 1) % n]
            u2 = u[(j + 1) % n, (i + 1) % n]
            u3 = u[(j + 1) % n, i]
            p = np.vdot(u0, u1) * np.vdot(u1, u2) * np.vdot(u2, u3) * np.vdot(u3, u0)
            total += -np.angle(p)
    return total / (2 * math.pi)


@pytest.mark.parametrize("mass,expect_nonzero", [(1.0, True), (-1.2, True), (2.6, False)])
def test_ahc_quantized_chern_layers(mass, expect_nonzero):
    c = 3.0
    m = qwz_layers(mass, c=c)
    chern = fhs_chern(m, 12)
    assert abs(chern - round(chern)) < 1e-9
    chern = round(chern)
    assert (chern != 0) == expect_nonzero
    task = G.GeometryTask(fermi_energies=[0.0, 5.0])
    _, total = G.integrate_geometry_mesh(m, task, (48, 48, 1), "ahc")
    quantum = E * E / (2 * math.pi * HBAR) / (c * 1e-10)
    assert abs(total[0, 0, 1] - (-quantum * chern)) < 1e-7 * quantum
    assert abs(total[0, 1, 0] - quantum * chern) < 1e-7 * quantum
    assert abs(total[0, 0, 0]) < 1e-9 * quantum
    # both bands occupied: the total Chern number of a two-band model is zero
    assert abs(total[1, 0, 1]) < 1e-7 * quantum


def test_fermi_surface_weight_reference_values():
    # reference test/berry_multipole_dc_unit.jl:39-43
    assert G.fermi_surface_weight(0.0, 0.0, 0.0, 0.02) == pytest.approx(1 / (math.sqrt(math.pi) * 0.02), rel=1e-14)
    w = G.fermi_surface_weight(0.16, 0.0, 0.0, 0.02)
    assert 0 < w < 1e-14
    assert G.fermi_surface_weight(0.0, 0.0, 300.0, None) == pytest.approx(1 / (4 * 8.617333262145e-5 * 300), rel=1e-14)
    with pytest.raises(ValueError):
        G.fermi_surface_weight(0.0, 0.0, 0.0, None)


def test_fermi_surface_weight_is_minus_fermi_derivative():
    kt = G.RESPONSE_KB_EV * 75.0
    f = lambda x: 1.0 / (np.exp((x - 0.1) / kt) + 1.0)
    g = lambda x: 1.0 / (np.exp((0.1 - x) / kt) + 1.0)     # 1 - f, free of cancellation below mu
    h = 1e-6
    for e in (-0.05, 0.08, 0.1, 0.13, 0.3):
        fd = -(f(e + h) - f(e - h)) / (2 * h) if e >= 0.1 else (g(e + h) - g(e - h)) / (2 * h)
        assert G.fermi_surface_weight(e, 0.1, 75.0, None) == pytest.approx(fd, rel=1e-6, abs=1e-12)


def _inversion_symmetric(model):
    idx = {tuple(r): i for i, r in enumerate(model.r_vectors.tolist())}
    hr = model.hamiltonian_r.copy()
    for i, r in enumerate(model.r_vectors.tolist()):
        j = idx[tuple(-x for x in r)]
        if i < j:
            s = 0.5 * (hr[i] + hr[j])
            s = 0.5 * (s + s.conj().T)
            hr[i] = s
            hr[j] = s
    z = idx[(0, 0, 0)]
    hr[z] = 0.5 * (hr[z] + hr[z].conj().T)
    return TightBindingModel(model.lattice, model.r_vectors, model.r_degeneracies, hr, None)


def test_bcd_vanishes_with_inversion():
    base = synthetic_model(4, 1, 8)
    task = G.GeometryTask(fermi_energies=[-0.5, 0.0, 0.7], temperature=600.0, degeneracy_threshold=1e-9)
    _, d0 = G.integrate_geometry_mesh(base, task, (6, 6, 6), "bcd")
    _, d1 = G.integrate_geometry_mesh(_inversion_symmetric(base), task, (6, 6, 6), "bcd")
    assert np.linalg.norm(d0) > 0
    assert np.linalg.norm(d1) < 1e-10 * np.linalg.norm(d0)
    # Omega_ab is antisymmetric, hence so is D in (a, b)
    assert np.allclose(d0, -np.swapaxes(d0, 1, 2), atol=1e-12 * np.abs(d0).max())


def test_ahc_integrand_equals_occupied_band_curvature():
    m = synthetic_model(5, 1, 13, position_scale=0.1)
    k = np.random.default_rng(3).uniform(0, 1, (4, 3))
    d = G.geometry_data(m, k, 0.0)
    omega = G.band_berry_curvature(d)
    mus = [-50.0, -0.4, 0.3, 50.0]
    task = G.GeometryTask(fermi_energies=mus)
    got = G.ahc_point_values(m, task, k, d)
    scale = -E * E / HBAR * 1e-20
    for i, mu in enumerate(mus):
        occ = d.energies <= mu
        ref = scale * np.einsum("kn,knab->kab", occ.astype(float), omega)
        assert np.allclose(got[:, i], ref, rtol=1e-9, atol=1e-12 * abs(scale) * np.abs(omega).max())
    assert np.all(got[:, 0] == 0.0)


def test_bcd_gap_error_on_degenerate_bands():
    m = synthetic_model(2, 1, 4)
    hr = np.zeros((m.num_r_vectors, 4, 4), dtype=complex)
    hr[:, :2, :2] = m.hamiltonian_r
    hr[:, 2:, 2:] = m.hamiltonian_r      # two decoupled copies: every level doubly degenerate
    dup = TightBindingModel(m.lattice, m.r_vectors, m.r_degeneracies, hr, None)
    task = G.GeometryTask(fermi_energies=[0.0], temperature=300.0)
    with pytest.raises(G.GapError):
        G.bcd_point_moments(dup, task, np.array([[0.1, 0.2, 0.3]]))


def test_bcd_dc_conductivity_index_map():
    rng = np.random.default_rng(0)
    d = rng.standard_normal((3, 3, 3))
    s1 = G.bcd_dc_conductivity(d, 1.0)
    s2 = G.bcd_dc_conductivity(d, 2.0)
    assert np.array_equal(s2, 2 * s1)
    assert np.allclose(s1, np.transpose(s1, (1, 0, 2)))     # symmetric in the two field indices
    assert np.allclose(G.bcd_dc_conductivity(d, 1.0, 2), 1e-10 * s1)
    pref = E ** 3 * 1e-15 / (2 * HBAR ** 2)
    assert s1[1, 2, 0] == pytest.approx(pref * (d[0, 1, 2] + d[0, 2, 1]), rel=1e-14)
