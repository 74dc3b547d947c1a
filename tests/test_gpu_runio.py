"""A TOML run file through the GPU engine to result files, read back and checked against the oracle.

The run file has the layout of the reference's response examples (a Wannier model, a generated photon
axis, a bz_mesh, shift + injection current and SHG tasks).  The written numbers are compared with the
oracle's mesh integrals (oracle/kloop.py, oracle/shg.py) at the written precision: 17 digits for the
SC / IC tensors, 12 (the SHG default) for the SHG spectra; bar 1e-10 relative."""
import os
import shutil
import sys

import numpy as np
import pytest

from oracle import kloop
from oracle import shg as S
from wanniernlqg_amd import runio as R
from wanniernlqg_amd.model import read_wannier_tb

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")

RUN = """
[model]
source = "wannier"
model_file = "four_band_tb.dat"

[response]
photon_energy_min_ev = 0.0
photon_energy_max_ev = 3.0
photon_energy_step_ev = 0.25
fermi_energy_ev = 0.1
temperature_k = 300.0

[response.sampling]
kind = "bz_mesh"
k_mesh = [4, 3, 3]
spatial_dimension = 3

[response.tasks.sc]
kind = "SC"
method = "Conventional"

[response.tasks.ic]
kind = "IC"
method = "Conventional"

[response.tasks.shg]
kind = "SHG"
method = "Conventional"
output = "both"
[response.tasks.shg.numerics]
broadening = 0.03
intermediate_regularization = 0.02
degeneracy_threshold = 0.002

[output]
output_root = "results"
response_output_digits = 17
"""


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def read_tensor(path):
    rows = [l.split() for l in open(path).read().splitlines() if not l.startswith("#")]
    f = np.array([[float(x) for x in r] for r in rows])
    z = f[:, 2::2] + 1j * f[:, 3::2]
    return f[:, 0], f[:, 1], z.reshape(len(rows), 3, 3, 3)


def test_run_file_to_result_files(tmp_path):
    shutil.copy(FIXTURE, tmp_path / "four_band_tb.dat")
    cfg_path = tmp_path / "run.toml"
    cfg_path.write_text(RUN)
    say("[runio] run file: fixture model, mesh 4x3x3, SC + IC + SHG")
    written = R.run(str(cfg_path))
    names = sorted(os.path.relpath(p, tmp_path) for p in written)
    comps = S.SECOND_HARMONIC_COMPONENTS
    labels = ["".join("xyz"[a] for a in c) for c in comps]
    assert names == sorted([os.path.join("results", "sc", "four_band_sc_conv.dat"),
                            os.path.join("results", "ic", "four_band_ic_conv.dat")] +
                           [os.path.join("results", "shg", f"shg_{q}_{l}.dat") for q in ("chi", "sigma")
                            for l in labels])
    omega = R.energy_range(0.0, 3.0, 0.25)
    model = read_wannier_tb(FIXTURE)
    task = kloop.Task(photon_energies=omega, broadening=0.040, broadening_type="Gaussian", fermi_energy=0.1,
                      temperature=300.0, denominator_regularization=1e-3, transition_window_factor=5.0,
                      shift_current=True, injection_current=True)
    _, ot = kloop.integrate_mesh(model, task, (4, 3, 3), batch=64)
    for key in ("sc", "ic"):
        mu, om, z = read_tensor(str(tmp_path / "results" / key / f"four_band_{key}_conv.dat"))
        assert np.all(mu == 0.1) and np.array_equal(om, omega)
        r = rel(z, ot[key])
        say(f"  {key} file vs oracle rel {r:.3e}")
        assert r < 1e-10
    _, st = S.integrate_shg_mesh(model, (4, 3, 3), omega, comps, fermi_energy=0.1, temperature=300.0,
                                 broadening=0.03, low_frequency_broadening=0.025, gaussian=True,
                                 intermediate_regularization=0.02, eta_correction=True, degeneracy_threshold=0.002)
    for q, name in ((0, "chi"), (1, "sigma")):
        for c, label in enumerate(labels):
            rows = [l.split() for l in open(tmp_path / "results" / "shg" / f"shg_{name}_{label}.dat")
                    if not l.startswith("#")]
            f = np.array([[float(x) for x in r] for r in rows])
            assert np.array_equal(f[:, 0], omega)
            want = st[:, c, :, q].sum(axis=1)
            got = f[:, 1] + 1j * f[:, 2]
            assert rel(got, want) < 1e-10, (name, label, rel(got, want))
            # output = "both": the seven terms follow the total, each against the oracle
            terms = f[:, 3::2] + 1j * f[:, 4::2]
            assert terms.shape[1] == 7
            assert rel(terms, st[:, c, :, q]) < 1e-10
