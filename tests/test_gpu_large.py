"""Engines with more than 64 orbitals (the SHG family, up to 128): the Jacobi with its eigenvector
matrix in global memory and the rotation with U^+ X staged in global memory.

The NumPy SHG oracle loops in Python (about 30 s per k point at n = 96), so the large-n SHG path is
checked through an identity instead: for a direct sum of two decoupled models (block-diagonal H(R)
and A(R)) every dipole and vertex between the blocks vanishes, so at T = 0 the per-point SHG
densities are exactly the sums of the two blocks' densities.  The blocks themselves (n <= 64) run on
the oracle-pinned path.  Interpolation and the eigensystem are compared with the oracle directly."""
import sys

import numpy as np
import pytest

from oracle import kloop
from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import TightBindingModel, synthetic_model

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def shg_task(**kw):
    base = dict(second_harmonic=True, photon_energies=np.linspace(0.0, 3.0, 40), broadening=0.05,
                shg_components=[(0, 1, 1), (1, 0, 2), (2, 2, 2)], fermi_energy=0.0, temperature=0.0)
    base.update(kw)
    return E.EngineTask(**base)


def direct_sum(m1, m2):
    assert np.array_equal(m1.r_vectors, m2.r_vectors)
    n1, n2 = m1.num_orbitals, m2.num_orbitals
    nr = m1.r_vectors.shape[0]
    h = np.zeros((nr, n1 + n2, n1 + n2), dtype=np.complex128)
    h[:, :n1, :n1] = m1.hamiltonian_r
    h[:, n1:, n1:] = m2.hamiltonian_r
    a = np.zeros((nr, 3, n1 + n2, n1 + n2), dtype=np.complex128)
    a[:, :, :n1, :n1] = m1.position_r
    a[:, :, n1:, n1:] = m2.position_r
    return TightBindingModel(m1.lattice, m1.r_vectors, m1.r_degeneracies, h, a, name=f"sum_{n1}_{n2}")


@pytest.mark.parametrize("n", [72, 96, 128])
def test_large_spectrum_and_channels(n):
    model = synthetic_model(n, 1, 50 + n, position_scale=0.1)
    k = np.random.default_rng(n).uniform(0, 1, (8, 3))
    say(f"[large] spectrum + channels n={n}")
    with E.Engine(model, shg_task()) as eng:
        e, v = eng.spectrum_points(k)
        ch = eng.fourier_points(k[:2])
    w = kloop.interpolate(model, k)
    ew, vw = kloop.hermitian_spectrum(w["H"])
    scale = np.max(np.abs(ew))
    say(f"  max |dE| {np.max(np.abs(e - ew)):.3e}, max |dU| {np.max(np.abs(v - vw)):.3e}")
    assert np.max(np.abs(e - ew)) < 1e-12 * scale
    # residual and orthonormality of the GPU eigenvectors themselves
    for i in range(k.shape[0]):
        res = w["H"][i] @ v[i] - v[i] * e[i][None, :]
        assert np.linalg.norm(res) < 1e-11 * scale * n
        assert np.linalg.norm(v[i].conj().T @ v[i] - np.eye(n)) < 1e-11 * n
    # canonical phases: the same vectors as the oracle wherever the levels are separated
    gaps = np.min(np.diff(ew, axis=1), axis=1)
    assert np.max(np.abs(v - vw)) < 1e-13 / max(gaps.min(), 1e-6) * 1e3
    # channels 0 (H), 1-3 (dH) and the position channels against the oracle's Fourier sums
    for c, ref in [(0, w["H"][:2]), (1, w["dH"][:2, 0]), (10, w["A"][:2, 0]), (13, w["dA"][:2, 0, 0])]:
        assert np.max(np.abs(ch[:, c] - ref)) < 1e-12 * max(np.max(np.abs(ref)), 1.0)


@pytest.mark.parametrize("n1,n2,mesh,batch", [(48, 48, (3, 2, 2), 0), (40, 40, (3, 2, 2), 0), (64, 64, (3, 2, 2), 0),
                                               (48, 48, (16, 8, 3), 128)])
def test_large_shg_direct_sum(n1, n2, mesh, batch):
    """SHG of a direct sum of two decoupled models equals the sum of the parts (n > 64 against the n <= 64
    path).  The 3 x 2 x 2 mesh runs in one batch (the eigensolver's warm-start groups of 4); the 16 x 8 x 3
    mesh with 128-point batches runs one k3 plane per batch, so planes 2 and 3 start every point from the same
    point of the previous plane."""
    m1 = synthetic_model(n1, 1, 3, position_scale=0.1)
    m2 = synthetic_model(n2, 1, 4, position_scale=0.1)
    m = direct_sum(m1, m2)
    k = np.random.default_rng(9).uniform(0, 1, (5, 3))
    say(f"[large] SHG direct sum {n1}+{n2}, mesh {mesh}, batch {batch}")
    with E.Engine(m, shg_task(batch_size=batch)) as eng:
        lanes, total, st = eng.run_mesh(mesh)
        pts = eng.eval_points(k)["shg"]
    parts = []
    for part in (m1, m2):
        with E.Engine(part, shg_task(batch_size=batch)) as eng:
            parts.append((eng.run_mesh(mesh), eng.eval_points(k)["shg"]))
    lane_sum = parts[0][0][0]["shg"] + parts[1][0][0]["shg"]
    total_sum = parts[0][0][1]["shg"] + parts[1][0][1]["shg"]
    pt_sum = parts[0][1] + parts[1][1]
    rl = np.linalg.norm(lanes["shg"] - lane_sum) / np.linalg.norm(lane_sum)
    rt = np.linalg.norm(total["shg"] - total_sum) / np.linalg.norm(total_sum)
    rp = np.linalg.norm(pts - pt_sum) / np.linalg.norm(pt_sum)
    say(f"  lanes {rl:.3e} total {rt:.3e} points {rp:.3e}; kpoints {st['kpoints']}")
    assert rl < 1e-10 and rt < 1e-10 and rp < 1e-10
    assert st["kpoints"] == int(np.prod(mesh))
