"""GPU checks on the models of the benchmark configurations themselves (through the C ABI), against the
oracle:

* n_w = 96, |R_i| <= 6 (2197 R vectors) with a position operator (the large SHG configuration): the
  interpolated channels and the eigensystem at generic k points against the oracle's Fourier sums and
  LAPACK, and the per-point SHG densities against the oracle's values stored by
  tests/golden/make_cfg4_shg_points.py (the oracle takes ~30 s per point at this size);
* n_w = 16, |R_i| <= 4 with four near-degenerate Weyl-like pairs (1-5 meV gaps; the learned-proposal
  configuration): per-point BCD moments at mu = 0, T = 50 K next to every hot spot against the oracle.

Bar: FP64, channels and energies 1e-12 relative to the largest entry, responses 1e-10 relative."""
import os
import sys

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kloop
from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import config_model

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CFG4 = os.path.join(HERE, "golden", "cfg4_shg", "points.npz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def cfg4_engine(fix):
    task = E.EngineTask(second_harmonic=True, photon_energies=fix["frequencies"], broadening=0.02,
                        shg_components=[tuple(int(x) for x in c) for c in fix["components"]], fermi_energy=0.0,
                        temperature=0.0, shg_intermediate_regularization=1e-3, shg_low_frequency_broadening=0.025,
                        shg_eta_correction=True, degeneracy_threshold=0.002)
    return E.Engine(config_model(4), task)


def test_large_configuration_channels_and_spectrum():
    fix = np.load(CFG4)
    k = fix["kpoints"]
    model = config_model(4)
    say(f"[cfg4] n_w={model.num_orbitals} N_R={model.num_r_vectors}: channels and spectrum at {k.shape[0]} points")
    with cfg4_engine(fix) as eng:
        ch = eng.fourier_points(k)
        e, v = eng.spectrum_points(k)
    w = kloop.interpolate_factored(model, k)
    sym = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    ref = [w["H"][:, None]] + [w["dH"][:, a][:, None] for a in range(3)]
    ref += [w["ddH"][:, a, b][:, None] for a, b in sym]
    ref += [w["A"][:, a][:, None] for a in range(3)]
    ref += [w["dA"][:, a, b][:, None] for a in range(3) for b in range(3)]
    ref = np.concatenate(ref, axis=1)
    assert ch.shape == ref.shape
    for c in range(ref.shape[1]):
        assert np.max(np.abs(ch[:, c] - ref[:, c])) <= 1e-12 * np.max(np.abs(ref[:, c])), c
    ew, _ = kloop.hermitian_spectrum(w["H"])
    assert np.max(np.abs(e - ew)) <= 1e-12 * np.max(np.abs(ew))
    hu = w["H"]
    for i in range(k.shape[0]):
        assert np.max(np.abs(hu[i] @ v[i] - v[i] * e[i][None, :])) <= 1e-11 * np.max(np.abs(ew))
        assert np.max(np.abs(v[i].conj().T @ v[i] - np.eye(model.num_orbitals))) <= 1e-12


def test_large_configuration_shg_points_match_stored_oracle_values():
    fix = np.load(CFG4)
    say("[cfg4] SHG densities at the stored points")
    with cfg4_engine(fix) as eng:
        got = eng.eval_points(fix["kpoints"])["shg"]
    ref = fix["densities"]
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-10, rel(got, ref)


def test_hotspot_configuration_bcd_points_against_oracle():
    model = config_model(3)
    spots = np.array([s["kstar"] for s in model.meta["spots"]])
    rng = np.random.default_rng(17)
    # three points within 1/64 of every hot spot (the base-mesh cell size) and three random points
    near = (spots[:, None, :] + rng.uniform(-1 / 64, 1 / 64, (spots.shape[0], 3, 3))).reshape(-1, 3)
    k = np.concatenate([near, rng.uniform(0, 1, (3, 3))])
    say(f"[cfg3] n_w={model.num_orbitals}: BCD at {k.shape[0]} points near the hot spots, mu 0, T 50 K")
    with E.Engine(model, E.EngineTask(berry_curvature_dipole=True, fermi_energies=[0.0], temperature=50.0,
                                      denominator_regularization=0.0, degeneracy_threshold=1e-6)) as eng:
        got = eng.eval_points(k)["bcd"]
    task = G.GeometryTask(fermi_energies=[0.0], temperature=50.0, denominator_regularization=0.0,
                          degeneracy_threshold=1e-6)
    ref = G.bcd_point_moments(model, task, k)
    assert np.max(np.abs(ref)) > 0
    for i in range(k.shape[0]):
        scale = max(np.max(np.abs(ref[i])), 1e-300)
        assert np.max(np.abs(got[i] - ref[i])) <= 1e-10 * scale or np.max(np.abs(ref[i])) < 1e-12 * np.max(np.abs(ref)), i
