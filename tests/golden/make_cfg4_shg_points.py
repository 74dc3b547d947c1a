"""Write tests/golden/cfg4_shg/points.npz: the oracle's per-point SHG densities of the n_w = 96,
|R_i| <= 6 (2197 R vectors) model with position operator (config_model(4)) at two generic k points.

The NumPy SHG oracle loops over band pairs in Python (about 30 s per point at n_w = 96), too slow for a
GPU test run, so its values are computed once here and stored as a fixture (inputs and outputs only).
Task: chi / sigma(-2w; w, w), components (a, b, c) = (x, y, y), (y, x, z), (z, z, z), 24 photon energies in
[0, 3] eV, Gaussian 0.02 eV, mu = 0, T = 0, eta = 1e-3 eV, eta-corrected vertices, low-frequency width
0.025 eV, degeneracy threshold 0.002 eV.

    python tests/golden/make_cfg4_shg_points.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import shg  # noqa: E402
from wanniernlqg_amd.model import config_model  # noqa: E402

COMPONENTS = [(0, 1, 1), (1, 0, 2), (2, 2, 2)]


def main():
    model = config_model(4)
    kpts = np.random.default_rng(2024).uniform(0.0, 1.0, (2, 3))
    freqs = np.linspace(0.0, 3.0, 24)
    vals = shg.shg_point_values(model, kpts, freqs, COMPONENTS, fermi_energy=0.0, temperature=0.0,
                                broadening=0.02, low_frequency_broadening=0.025, gaussian=True,
                                intermediate_regularization=1e-3, eta_correction=True, degeneracy_threshold=0.002)
    out = os.path.join(HERE, "cfg4_shg", "points.npz")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    np.savez_compressed(out, kpoints=kpts, frequencies=freqs, components=np.array(COMPONENTS, dtype=np.int32),
                        densities=vals)
    print(out, vals.shape, np.abs(vals).max())


if __name__ == "__main__":
    main()
