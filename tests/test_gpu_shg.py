"""GPU parity of the SHG stage (a per-k optical vertex kernel, a pair-coefficient kernel and the
frequency contraction) through the C ABI: against the reference's own SHG golden files (4-band
cases, the GPU computes every band of the model) and against the oracle (oracle/shg.py) for point
values and mesh lanes.  Bar: goldens atol 1e-10 rtol 1e-8 (the reference test's tolerance); oracle
1e-10 relative (Frobenius norm)."""
import os
import sys

import numpy as np
import pytest

from oracle import shg as S
from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import read_wannier_tb, synthetic_model

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "second_harmonic")
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")
FREQS = [0.0, 0.1, 0.73, 1.7]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def shg_engine(model, freqs, **kw):
    t = dict(photon_energies=np.asarray(freqs, dtype=np.float64), second_harmonic=True, shift_current=False)
    t.update(kw)
    return E.Engine(model, E.EngineTask(**t))


@pytest.mark.parametrize("gamma", [False, True])
@pytest.mark.parametrize("gaussian", [True, False])
@pytest.mark.parametrize("corrected", [True, False])
def test_shg_golden_four_band(gamma, gaussian, corrected):
    model = read_wannier_tb(FIXTURE)
    point = [0.0, 0.0, 0.0] if gamma else [0.17, 0.23, 0.31]
    name = ("gamma" if gamma else "generic") + "_" + ("gaussian" if gaussian else "lorentzian") + "_" + \
        ("corrected" if corrected else "uncorrected") + ".dat"
    say(f"[shg golden] {name}")
    with shg_engine(model, FREQS, broadening=0.025, broadening_type="Gaussian" if gaussian else "Lorentzian",
                    shg_low_frequency_broadening=0.025, shg_intermediate_regularization=0.04,
                    shg_eta_correction=corrected, fermi_energy=0.0, temperature=0.0) as eng:
        got = eng.eval_points(np.array([point]))["shg"][0]          # [w, comp, 7, 2]
    ref = np.loadtxt(os.path.join(GOLD, name))
    expected = (ref[:, 0] + 1j * ref[:, 1]).reshape((4, 18, 2), order="F")
    total = got.sum(axis=2)
    assert np.all(np.abs(total - expected) <= 1e-10 + 1e-8 * np.abs(expected))
    assert np.all(got[:, :, 5:7, :] == 0)


CASES = {
    "fixture_T300": (lambda: read_wannier_tb(FIXTURE), dict(temperature=300.0, fermi_energy=0.1)),
    "n6_pos_T0": (lambda: synthetic_model(6, 1, 3, position_scale=0.1), dict()),
    "n6_nopos_lorentz": (lambda: synthetic_model(6, 1, 5), dict(broadening_type="Lorentzian", temperature=150.0)),
    "n10_pos_T500_sub": (lambda: synthetic_model(10, 1, 7, position_scale=0.1),
                         dict(temperature=500.0, shg_components=[(0, 1, 1), (2, 2, 2), (1, 0, 2)])),
}


def oracle_kw(kw):
    return dict(fermi_energy=kw.get("fermi_energy", 0.0), temperature=kw.get("temperature", 0.0),
                broadening=0.03, low_frequency_broadening=0.025,
                gaussian=kw.get("broadening_type", "Gaussian") == "Gaussian", intermediate_regularization=0.02,
                eta_correction=True, degeneracy_threshold=0.002)


def engine_kw(kw):
    t = dict(broadening=0.03, shg_low_frequency_broadening=0.025, shg_intermediate_regularization=0.02)
    t.update(kw)
    return t


@pytest.mark.parametrize("case", sorted(CASES))
def test_shg_points_vs_oracle(case):
    make, kw = CASES[case]
    m = make()
    freqs = np.linspace(0.0, 3.0, 37)
    k = np.random.default_rng(11).uniform(0, 1, (5, 3))
    say(f"[shg points] {case}")
    with shg_engine(m, freqs, **engine_kw(kw)) as eng:
        got = eng.eval_points(k)["shg"]
        comps = eng.shg_components
    ref = S.shg_point_values(m, k, freqs, comps, **oracle_kw(kw))
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-10, rel(got, ref)


@pytest.mark.parametrize("case,backend", [("fixture_T300", "planes"), ("n6_pos_T0", "direct"),
                                          ("n10_pos_T500_sub", "auto")])
def test_shg_mesh_vs_oracle(case, backend):
    make, kw = CASES[case]
    m = make()
    freqs = np.linspace(0.0, 3.0, 70)
    mesh = (3, 4, 3)
    say(f"[shg mesh] {case} {backend}")
    with shg_engine(m, freqs, fourier_backend=backend, **engine_kw(kw)) as eng:
        lanes, total, _ = eng.run_mesh(mesh)
        comps = eng.shg_components
    rl, rt = S.integrate_shg_mesh(m, mesh, freqs, comps, **oracle_kw(kw))
    assert rel(lanes["shg"], rl) < 1e-10
    assert rel(total["shg"], rt) < 1e-10
