"""Pin the oracle (oracle/kloop.py) by what the mathematics fixes: brute-force Fourier sums,
periodicity, finite differences, LAPACK residuals, the covariant-derivative identity of the
generalized derivative, inversion zeros and gauge invariance.  Each check is chosen so that a
plausible slip (sign, index, transposed operand, dropped term) fails it."""
import cmath
import math
import os

import numpy as np
import pytest

from oracle import kloop
from wanniernlqg_amd.model import (TightBindingModel, read_wannier_tb, synthetic_model, write_wannier_tb)

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")


def cart_to_frac(model, kappa):
    """Cartesian wavevector kappa (1/Angstrom) -> fractional k with 2 pi R.k = kappa . R_cart."""
    return model.lattice @ kappa / (2 * math.pi)


def test_fixture_reader_values():
    m = read_wannier_tb(FIXTURE)
    assert m.num_orbitals == 4 and m.num_r_vectors == 2
    assert list(m.r_degeneracies) == [2, 4]
    assert m.hamiltonian_r[0, 0, 0] == -2.5 and m.hamiltonian_r[1, 3, 3] == 0.4
    assert abs(m.hamiltonian_r[0, 1, 0] - 0.15) < 1e-15   # row 2 col 1 of R=(0,0,0): 0.05*(2+1)
    assert m.position_r[0, 0, 0, 0] == 0.5 and abs(m.position_r[1, 2, 3, 2] - 0.25 * 0.01 * 7) < 1e-15


def test_tb_roundtrip(tmp_path):
    m = synthetic_model(5, 1, 3, position_scale=0.1)
    p = write_wannier_tb(m, str(tmp_path / "m_tb.dat"))
    r = read_wannier_tb(p)
    assert np.array_equal(r.r_vectors, m.r_vectors)
    assert np.array_equal(r.hamiltonian_r, m.hamiltonian_r)
    assert np.array_equal(r.position_r, m.position_r)
    assert np.array_equal(r.lattice, m.lattice)


def test_synthetic_model_hermitian_pairs():
    m = synthetic_model(6, 2, 4, position_scale=0.1)
    idx = {tuple(r): i for i, r in enumerate(m.r_vectors.tolist())}
    for i, r in enumerate(m.r_vectors.tolist()):
        j = idx[tuple(-x for x in r)]
        assert np.array_equal(m.hamiltonian_r[j], m.hamiltonian_r[i].conj().T)
        assert np.array_equal(m.position_r[j], np.conj(np.swapaxes(m.position_r[i], 1, 2)))


def test_interpolation_brute_force():
    m = read_wannier_tb(FIXTURE)
    k = np.array([[0.17, 0.23, 0.31], [0.5, -0.25, 0.125]])
    w = kloop.interpolate(m, k)
    n = m.num_orbitals
    for ki in range(2):
        for i in range(n):
            for j in range(n):
                h = 0j
                dh = [0j, 0j, 0j]
                for r in range(m.num_r_vectors):
                    R = m.r_vectors[r]
                    ph = cmath.exp(2j * math.pi * sum(int(R[a]) * k[ki, a] for a in range(3))) / m.r_degeneracies[r]
                    h += ph * m.hamiltonian_r[r, i, j]
                    rc = [sum(int(R[x]) * m.lattice[x, a] for x in range(3)) for a in range(3)]
                    for a in range(3):
                        dh[a] += ph * 1j * rc[a] * m.hamiltonian_r[r, i, j]
                assert abs(w["H"][ki, i, j] - h) < 1e-14
                for a in range(3):
                    assert abs(w["dH"][ki, a, i, j] - dh[a]) < 1e-14


def test_interpolation_periodic_and_hermitian():
    m = synthetic_model(5, 2, 1)
    k = np.random.default_rng(0).uniform(0, 1, (4, 3))
    h0 = kloop.interpolate(m, k)["H"]
    h1 = kloop.interpolate(m, k + np.array([1.0, -2.0, 3.0]))["H"]
    assert np.max(np.abs(h0 - h1)) < 1e-12
    assert np.max(np.abs(h0 - np.conj(np.swapaxes(h0, 1, 2)))) < 1e-13


def test_derivative_channels_finite_difference():
    m = synthetic_model(4, 2, 2, position_scale=0.1, lattice_constant=3.0)
    kappa = np.array([0.31, -0.44, 0.57])
    hstep = 1e-5
    w0 = kloop.interpolate(m, cart_to_frac(m, kappa)[None])
    for b in range(3):
        e = np.zeros(3)
        e[b] = hstep
        wp = kloop.interpolate(m, cart_to_frac(m, kappa + e)[None])
        wm = kloop.interpolate(m, cart_to_frac(m, kappa - e)[None])
        fd_h = (wp["H"] - wm["H"]) / (2 * hstep)
        assert np.max(np.abs(fd_h - w0["dH"][:, b])) < 1e-7
        for a in range(3):
            fd_dh = (wp["dH"][:, a] - wm["dH"][:, a]) / (2 * hstep)
            assert np.max(np.abs(fd_dh - w0["ddH"][:, a, b])) < 1e-6
            fd_a = (wp["A"][:, a] - wm["A"][:, a]) / (2 * hstep)
            assert np.max(np.abs(fd_a - w0["dA"][:, a, b])) < 1e-7


def test_eigensystem_residual_and_phase():
    m = read_wannier_tb(FIXTURE)   # non-Hermitian H(R): only the upper triangle counts
    k = np.random.default_rng(3).uniform(0, 1, (6, 3))
    hk = kloop.interpolate(m, k)["H"]
    e, u = kloop.hermitian_spectrum(hk)
    hu = np.triu(hk) + np.conj(np.swapaxes(np.triu(hk, 1), 1, 2))
    d = np.arange(4)
    hu[:, d, d] = hk[:, d, d].real
    assert np.max(np.abs(hu @ u - u * e[:, None, :])) < 1e-13
    assert np.all(np.diff(e, axis=1) >= 0)
    piv = np.argmax(np.abs(u), axis=1)
    pv = np.take_along_axis(u, piv[:, None, :], axis=1)[:, 0, :]
    assert np.max(np.abs(pv.imag)) < 1e-15 and np.min(pv.real) > 0


def test_canonicalize_first_on_ties():
    v = np.array([[0.5j, 0.0], [-0.5, 1.0], [0.5, 0.0]], dtype=complex).reshape(1, 3, 2)
    c = kloop.canonicalize_eigenvectors(v)
    assert c[0, 0, 0] == 0.5 and c[0, 1, 0] == 0.5j   # pivot = first of the tied magnitudes
    assert np.array_equal(c[0, :, 1], v[0, :, 1])


def test_generalized_derivative_scalar_equals_commutator_form():
    """Restates the reference's scalar == GEMM identity (test/generalized_derivative_gemm_unit.jl:47-191)."""
    m = synthetic_model(7, 1, 9, position_scale=0.1)
    k = np.random.default_rng(4).uniform(0, 1, (2, 3))
    d = kloop.compute_kpoint_data(m, k, 1e-3)
    G = kloop.generalized_derivative(d)
    for ki in range(2):
        for a in range(3):
            for b in range(3):
                for n in range(7):
                    for mm in range(7):
                        v = kloop.generalized_derivative_element(d, ki, n, mm, a, b)
                        assert abs(v - G[ki, a, b, n, mm]) <= 1e-10 + 1e-8 * abs(v)


def test_generalized_derivative_is_covariant_derivative():
    """For isolated bands at eta = 0: r^a_{nm;b} = d_b r^a_nm - i (Ab^b_nn - Ab^b_mm) r^a_nm with the full
    diagonal connection Ab_nn = A_H,nn + i (U^+ d_b U)_nn, by central differences in a fixed gauge."""
    m = synthetic_model(5, 1, 21, position_scale=0.1, lattice_constant=3.0)
    kappa = np.array([0.4, 0.9, -0.7])
    h = 1e-5
    d0 = kloop.compute_kpoint_data(m, cart_to_frac(m, kappa)[None], 0.0)
    assert np.min(np.diff(d0.energies[0])) > 1e-2
    G = kloop.generalized_derivative(d0)[0]
    u0 = d0.eigenvectors[0]
    n = 5
    off = ~np.eye(n, dtype=bool)
    for b in range(3):
        e = np.zeros(3)
        e[b] = h
        dp = kloop.compute_kpoint_data(m, cart_to_frac(m, kappa + e)[None], 0.0)
        dm = kloop.compute_kpoint_data(m, cart_to_frac(m, kappa - e)[None], 0.0)
        # keep the gauge of the central point (canonical phases may flip between points)
        for dd in (dp, dm):
            ov = np.conj(u0.T) @ dd.eigenvectors[0]
            ph = np.diagonal(ov) / np.abs(np.diagonal(ov))
            fix = np.conj(ph)
            dd.berry_connection[0] = dd.berry_connection[0] * np.conj(fix)[None, :, None] * fix[None, None, :]
            dd.eigenvectors[0] = dd.eigenvectors[0] * fix[None, :]
        dudb = (dp.eigenvectors[0] - dm.eigenvectors[0]) / (2 * h)
        conn_diag = np.diagonal(d0.internal_connection[0][b]) + 1j * np.diagonal(np.conj(u0.T) @ dudb)
        for a in range(3):
            r = d0.berry_connection[0][a]
            dr = (dp.berry_connection[0][a] - dm.berry_connection[0][a]) / (2 * h)
            cov = dr - 1j * (conn_diag[:, None] - conn_diag[None, :]) * r
            err = np.max(np.abs((cov - G[a, b])[off]))
            assert err < 1e-5 * max(1.0, np.max(np.abs(G[a, b][off]))), (a, b, err)


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
    hr[idx[(0, 0, 0)]] = 0.5 * (hr[idx[(0, 0, 0)]] + hr[idx[(0, 0, 0)]].conj().T)
    return TightBindingModel(model.lattice, model.r_vectors, model.r_degeneracies, hr, None)


def test_shift_and_injection_vanish_with_inversion():
    base = synthetic_model(6, 1, 5)
    task = kloop.Task(photon_energies=np.linspace(0, 5, 60), broadening=0.1, shift_current=True,
                      injection_current=True)
    _, t0 = kloop.integrate_mesh(base, task, (6, 6, 6), batch=64)
    _, t1 = kloop.integrate_mesh(_inversion_symmetric(base), task, (6, 6, 6), batch=64)
    for key in ("sc", "ic"):
        assert np.linalg.norm(t0[key]) > 0
        assert np.linalg.norm(t1[key]) < 1e-9 * np.linalg.norm(t0[key])


def test_point_values_gauge_invariant():
    m = synthetic_model(6, 1, 12, position_scale=0.1)
    k = np.random.default_rng(9).uniform(0, 1, (3, 3))
    task = kloop.Task(photon_energies=np.linspace(0, 6, 80), broadening=0.1, shift_current=True,
                      injection_current=True)
    ref = kloop.point_values(m, task, k)
    w = kloop.interpolate(m, k)
    e, u = kloop.hermitian_spectrum(w["H"])
    phases = np.exp(2j * np.pi * np.random.default_rng(1).uniform(0, 1, u.shape[::2][:1] + (u.shape[2],)))
    u2 = u * phases[:, None, :]
    v = kloop.rotate(u2, w["dH"])
    ah = kloop.rotate(u2, w["A"])
    de = e[:, :, None] - e[:, None, :]
    inv = kloop.regularized_energy_inverse(de, 1e-3)
    dd = -v * inv[:, None]
    data = kloop.KPointData(e, u2, v, kloop.rotate(u2, w["ddH"]), ah, kloop.rotate(u2, w["dA"]), inv, dd, ah + 1j * dd)
    got = kloop.point_values(m, task, k, data)
    for key in ("sc", "ic"):
        assert np.linalg.norm(got[key] - ref[key]) <= 1e-10 * np.linalg.norm(ref[key])


def test_injection_vanishes_for_flat_bands():
    """v_mm = v_nn (all R != 0 hoppings removed): the injection kernel is identically zero."""
    m = synthetic_model(5, 1, 2, position_scale=0.1)
    hr = np.zeros_like(m.hamiltonian_r)
    z = [i for i, r in enumerate(m.r_vectors.tolist()) if r == [0, 0, 0]][0]
    hr[z] = m.hamiltonian_r[z]
    flat = TightBindingModel(m.lattice, m.r_vectors, m.r_degeneracies, hr, m.position_r)
    task = kloop.Task(photon_energies=np.linspace(0, 6, 50), broadening=0.2, shift_current=False,
                      injection_current=True)
    v = kloop.point_values(flat, task, np.random.default_rng(2).uniform(0, 1, (4, 3)))
    assert np.max(np.abs(v["ic"])) == 0.0


@pytest.mark.parametrize("nk", [1, 7, 63, 64, 65, 1000, 2097152])
def test_lane_partition(nk):
    nl = kloop.lane_count(nk)
    assert nl == min(nk, 64)
    ranges = [kloop.lane_range(nk, nl, l) for l in range(nl)]
    assert ranges[0][0] == 0 and ranges[-1][1] == nk
    for (a, b), (c, d) in zip(ranges, ranges[1:]):
        assert b == c and b > a
    sizes = [b - a for a, b in ranges]
    assert max(sizes) - min(sizes) <= 1


def test_mesh_order_axis1_fastest():
    k = kloop.integral_kpoints((4, 3, 2))
    assert np.allclose(k[1], [0.25, 0, 0]) and np.allclose(k[4], [0, 1 / 3, 0]) and np.allclose(k[12], [0, 0, 0.5])


def test_factored_interpolation_equals_channel_sums():
    """interpolate_factored (R factors on the Fourier factors) == interpolate (R-space channels) to rounding."""
    m = synthetic_model(6, 2, 13, position_scale=0.1, lattice_constant=3.3)
    k = np.random.default_rng(8).uniform(0, 1, (3, 3))
    a, b = kloop.interpolate(m, k), kloop.interpolate_factored(m, k)
    for key in ("H", "dH", "ddH", "A", "dA"):
        assert a[key].shape == b[key].shape
        assert np.max(np.abs(a[key] - b[key])) <= 1e-12 * max(1.0, np.max(np.abs(a[key]))), key
