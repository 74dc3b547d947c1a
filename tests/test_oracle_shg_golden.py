"""Pin the oracle's interpolation -> eigensystem -> rotation -> vertex chain and its SHG kernel
against the reference's 24 SHG golden files (reference test/second_harmonic_unit.jl:29-105,
fixtures test/fixtures/second_harmonic/*.dat, model examples/fixtures/synthetic_runtime/synthetic_tb.dat).
Tolerance as in the reference test: atol 1e-10, rtol 1e-8."""
import os

import numpy as np
import pytest

from oracle import kloop, shg
from wanniernlqg_amd.model import read_wannier_tb

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "second_harmonic")
MODEL = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")
FREQS = [0.0, 0.1, 0.73, 1.7]


def golden_name(bands, gamma, gaussian, corrected):
    return (("" if bands == 4 else f"{bands}band_") + ("gamma" if gamma else "generic") + "_"
            + ("gaussian" if gaussian else "lorentzian") + "_" + ("corrected" if corrected else "uncorrected") + ".dat")


@pytest.mark.parametrize("bands", [2, 3, 4])
@pytest.mark.parametrize("gamma", [False, True])
@pytest.mark.parametrize("gaussian", [True, False])
@pytest.mark.parametrize("corrected", [True, False])
def test_shg_golden(bands, gamma, gaussian, corrected):
    model = read_wannier_tb(MODEL)
    point = [0.0, 0.0, 0.0] if gamma else [0.17, 0.23, 0.31]
    d = kloop.compute_kpoint_data(model, np.array([point]), 1e-3)
    sel = slice(1, 3) if bands == 2 else slice(0, bands)
    out = shg.second_harmonic_response(
        d.energies[0, sel], d.internal_connection[0][:, sel, sel], d.hamiltonian_derivatives[0][:, sel, sel],
        d.internal_connection_derivatives[0][:, :, sel, sel], d.hamiltonian_second_derivatives[0][:, :, sel, sel],
        FREQS, fermi_energy=0.0, temperature=0.0, broadening=0.025, low_frequency_broadening=0.025,
        gaussian=gaussian, intermediate_regularization=0.04, eta_correction=corrected, degeneracy_threshold=0.002)
    ref = np.loadtxt(os.path.join(GOLD, golden_name(bands, gamma, gaussian, corrected)))
    expected = (ref[:, 0] + 1j * ref[:, 1]).reshape((4, 18, 2), order="F")
    got = out.sum(axis=2)
    assert np.all(np.abs(got - expected) <= 1e-10 + 1e-8 * np.abs(expected))
    # zero temperature: one-band terms vanish (reference second_harmonic_unit.jl:108-109)
    assert np.all(out[:, :, 5:7, :] == 0)


def test_shg_symmetric_incident_indices():
    """Swapping the incident-field indices (b, c) leaves the response unchanged
    (reference second_harmonic_unit.jl:110-118)."""
    model = read_wannier_tb(MODEL)
    d = kloop.compute_kpoint_data(model, np.array([[0.17, 0.23, 0.31]]), 1e-3)
    args = (d.energies[0], d.internal_connection[0], d.hamiltonian_derivatives[0],
            d.internal_connection_derivatives[0], d.hamiltonian_second_derivatives[0], FREQS)
    opts = dict(fermi_energy=0.0, temperature=0.0, broadening=0.025, low_frequency_broadening=0.025,
                gaussian=True, intermediate_regularization=0.04, eta_correction=True, degeneracy_threshold=0.002)
    a = shg.second_harmonic_response(*args, **opts)
    b = shg.second_harmonic_response(*args, [(x, z, y) for (x, y, z) in shg.SECOND_HARMONIC_COMPONENTS], **opts)
    assert np.allclose(a, b, atol=1e-10, rtol=1e-8)
