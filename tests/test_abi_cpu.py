"""CPU-side checks of the C-ABI library: it loads, exports exactly what include/wnlqg.h declares,
and rejects invalid input before touching the GPU (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import synthetic_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wnlqg.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*int32_t\s+(wnlqg_\w+)\s*\(", text, flags=re.M)))


def test_library_loads_and_version():
    lib = E.load_library()
    declared = int(re.search(r"#define WNLQG_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert lib.wnlqg_version() == declared == E.ABI_VERSION


def test_status_codes_match_header():
    """Every WNLQG_E_* code of the header (including the eigensolver's sweep-cap failure) has the same
    value and name in the binding's error map."""
    codes = dict((int(v), name) for name, v in
                 re.findall(r"#define (WNLQG_E_\w+) \((-\d+)\)", open(HEADER).read()))
    assert codes == E.ERRORS
    assert codes[-7] == "WNLQG_E_NOT_CONVERGED"


def test_exports_match_header():
    declared = header_functions()
    assert declared == sorted(E.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r"\bT (wnlqg_\w+)", out)))
    assert exported == declared


def test_fourier_backends_match_header():
    """The binding's fourier_backend names map to the header's WNLQG_FOURIER_* values (the reference's auto,
    direct and mixed, FourierExecutionPlan.jl:63-67, plus planes); an unknown name is rejected by the library."""
    text = open(HEADER).read()
    values = {name.lower(): int(v) for name, v in re.findall(r"#define WNLQG_FOURIER_(\w+) (\d+)", text)}
    assert values == E.FOURIER_BACKENDS
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), E.EngineTask(photon_energies=np.linspace(0, 4, 20), broadening=0.02,
                                                         fourier_backend="fft"))
    assert ei.value.code == -1


def _task(**kw):
    base = dict(photon_energies=np.linspace(0, 4, 20), broadening=0.02)
    base.update(kw)
    return E.EngineTask(**base)


def test_rejects_nonfinite_model():
    m = synthetic_model(4, 1, 0)
    m.hamiltonian_r[3, 1, 2] = np.nan
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(m, _task())
    assert ei.value.code == -2


def test_rejects_unsorted_frequencies():
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), _task(photon_energies=np.array([1.0, 0.5])))
    assert ei.value.code == -6


def test_rejects_unknown_smearing():
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), _task(broadening_type="Voigt"))
    assert ei.value.code == -1


def test_rejects_partial_band_window_and_large_n():
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(8, 1, 0), _task(band_window_size=2))
    assert ei.value.code == -6
    big = synthetic_model(65, 0, 0)
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(big, _task())
    assert ei.value.code == -6
    # more than 64 orbitals: SHG only (up to 128)
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(big, _geom())
    assert ei.value.code == -6
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(129, 0, 0), E.EngineTask(photon_energies=np.linspace(0, 3, 10), second_harmonic=True,
                                                          broadening=0.02))
    assert ei.value.code == -6


def _geom(**kw):
    base = dict(berry_curvature_dipole=True, fermi_energies=[0.0], temperature=300.0,
                denominator_regularization=0.0)
    base.update(kw)
    return E.EngineTask(**base)


@pytest.mark.parametrize("kw,code", [
    (dict(temperature=0.0), -1),                                    # T = 0 Fermi surface needs eta_fs
    (dict(temperature=300.0, fermi_surface_broadening=0.02), -1),   # eta_fs only at T = 0
    (dict(fermi_energies=np.zeros(65)), -1),                        # at most 64 Fermi energies
    (dict(fermi_energies=[np.nan]), -2),
    (dict(degeneracy_threshold=-1.0), -1),
    (dict(precision="mixed"), -6),
    (dict(berry_curvature_dipole=False, intrinsic_hall=True, temperature=10.0), -1),  # AHC needs T = 0
])
def test_rejects_invalid_geometry_tasks(kw, code):
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), _geom(**kw))
    assert ei.value.code == code


def test_rejects_mixed_driver_families():
    # SC with BCD: two reference drivers, refused before any device work
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), E.EngineTask(photon_energies=np.linspace(0, 1, 5), shift_current=True,
                                                        berry_curvature_dipole=True, fermi_energies=[0.0]))
    assert ei.value.code == -6
    # BCD with AHC (one engine sharing the front end): AHC still requires T = 0
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), _geom(intrinsic_hall=True))
    assert ei.value.code == -1


@pytest.mark.parametrize("kw,code", [
    (dict(shg_components=[(0, 1, 3)]), -1),
    (dict(shg_low_frequency_broadening=0.0), -1),
    (dict(shg_intermediate_regularization=-1.0), -1),
    (dict(precision="mixed"), -6),
    (dict(berry_curvature_dipole=True, fermi_energies=[0.0]), -6),   # SHG with BCD: two drivers
])
def test_rejects_invalid_shg_tasks(kw, code):
    base = dict(photon_energies=np.linspace(0, 3, 10), second_harmonic=True, broadening=0.02)
    base.update(kw)
    with pytest.raises(E.WnlqgError) as ei:
        E.Engine(synthetic_model(4, 1, 0), E.EngineTask(**base))
    assert ei.value.code == code
