"""The C++/OpenMP restatement of the shift / injection current k-loop (oracle/kloop_omp.cpp, the CPU baseline
of bench.py) against the NumPy oracle (oracle/kloop.py), which the other test_oracle_* files pin: two
implementations that share no code.  Its Householder + QL eigensolver against LAPACK (numpy.linalg.eigh on the
upper triangle, canonical phase).  Bar: 1e-12 relative (eigenvalues and vectors), 1e-11 for the integrals."""
import numpy as np
import pytest

from oracle import kloop
from oracle import kloop_omp as O
from wanniernlqg_amd.model import config_model, read_wannier_tb, synthetic_model
import os

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")


@pytest.fixture(scope="module", autouse=True)
def built():
    O.build()


@pytest.mark.parametrize("n", [1, 2, 5, 16, 33, 48])
def test_eigensystem_matches_lapack_upper_triangle(n):
    rng = np.random.default_rng(n)
    m = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))      # not Hermitian: only the upper triangle counts
    e, v = O.hermitian_eigen(m)
    er, vr = kloop.hermitian_spectrum(m[None])
    scale = max(1.0, np.abs(er).max())
    assert np.abs(e - er[0]).max() <= 1e-12 * scale
    assert np.abs(v - vr[0]).max() <= 1e-11


def test_eigensystem_with_a_degenerate_pair():
    """Two equal eigenvalues: the eigenvectors span the same space (the basis inside it is free)."""
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)))
    h = q @ np.diag([-1.0, 0.5, 0.5, 1.0, 2.0, 3.0]) @ q.conj().T
    e, v = O.hermitian_eigen(h)
    assert np.abs(e - [-1.0, 0.5, 0.5, 1.0, 2.0, 3.0]).max() < 1e-13
    assert np.abs(h @ v - v * e[None, :]).max() < 1e-13
    assert np.abs(v.conj().T @ v - np.eye(6)).max() < 1e-13


def _plane_sum(model, task, mesh, p0, p1):
    kp = kloop.integral_kpoints(mesh)
    k3 = np.rint(kp[:, 2] * mesh[2]).astype(int)
    kp = kp[(k3 >= p0) & (k3 < p1)]
    vals = kloop.point_values(model, task, kp)
    return {key: v.sum(axis=0) / int(np.prod(mesh)) for key, v in vals.items()}


CASES = {
    "fixture_gauss_T0": (lambda: read_wannier_tb(FIXTURE), (5, 4, 3), 0, 3,
                         dict(broadening=0.05)),
    "n8_pos_T300": (lambda: synthetic_model(8, 1, 3, position_scale=0.1), (6, 5, 4), 1, 3,
                    dict(broadening=0.05, temperature=300.0)),
    "n6_lorentz_eta0": (lambda: synthetic_model(6, 1, 5), (4, 4, 4), 2, 4,
                        dict(broadening=0.04, broadening_type="Lorentzian", denominator_regularization=0.0)),
    "benchmark_n32": (lambda: config_model(1), (4, 3, 2), 0, 1, dict(broadening=0.05)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_plane_integrals_match_numpy_oracle(case):
    make, mesh, p0, p1, kw = CASES[case]
    model = make()
    task = kloop.Task(photon_energies=np.linspace(0.0, 4.0, 60), shift_current=True, injection_current=True, **kw)
    got = O.integrate_planes(model, task, mesh, p0, p1, threads=4)
    ref = _plane_sum(model, task, mesh, p0, p1)
    for key in ("sc", "ic"):
        assert np.abs(ref[key]).max() > 0
        assert kloop.spectrum_relative_error(got[key], ref[key]).max() < 1e-11, key


def test_whole_mesh_equals_oracle_integral_and_threads_agree():
    model = synthetic_model(8, 1, 3, position_scale=0.1)
    mesh = (5, 4, 4)
    task = kloop.Task(photon_energies=np.linspace(0.0, 4.0, 60), broadening=0.05, shift_current=True,
                      injection_current=True)
    _, ref = kloop.integrate_mesh(model, task, mesh)
    one = O.integrate_planes(model, task, mesh, 0, mesh[2], threads=1)
    many = O.integrate_planes(model, task, mesh, 0, mesh[2], threads=8)
    for key in ("sc", "ic"):
        assert kloop.spectrum_relative_error(one[key], ref[key]).max() < 1e-11
        assert kloop.spectrum_relative_error(many[key], one[key]).max() < 1e-13
