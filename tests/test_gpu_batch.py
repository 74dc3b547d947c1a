"""GPU checks of the batching form: several models of uneven n_w (BCD + AHC, SC + IC) captured in one
HIP graph give the results of the individual engines, replays are bitwise identical, and the
autotuner picks a supported eigensolver shape."""
import sys

import numpy as np
import pytest

from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import config_model, synthetic_model

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def geometry_task():
    return E.EngineTask(berry_curvature_dipole=True, intrinsic_hall=True, fermi_energies=[-0.5, 0.0, 0.5],
                        temperature=0.0, fermi_surface_broadening=0.05, denominator_regularization=0.0,
                        degeneracy_threshold=0.0, batch_size=2048)


def test_batch_matches_individual_engines():
    entries = [(config_model(2, seed), geometry_task(), mesh)
               for seed, mesh in ((0, (6, 6, 6)), (1, (8, 7, 5)), (4, (5, 5, 5)), (7, (6, 4, 6)))]
    entries.append((synthetic_model(12, 1, 3, position_scale=0.1),
                    E.EngineTask(photon_energies=np.linspace(0, 4, 60), broadening=0.05, injection_current=True,
                                 shift_current=True, batch_size=2048), (5, 6, 4)))
    say(f"[batch] {len(entries)} entries, n_w = {[e[0].num_orbitals for e in entries]}")
    with E.Batch(entries, autotune=True) as b:
        tuning = b.tuning()
        r1, ms1 = b.run()
        r2, ms2 = b.run()
    assert all(t in (64, 128, 256) for t in tuning)
    for (l1, t1), (l2, t2) in zip(r1, r2):
        for key in t1:
            assert np.array_equal(l1[key], l2[key]) and np.array_equal(t1[key], t2[key])
    for (model, task, mesh), (lanes, total) in zip(entries, r1):
        with E.Engine(model, task) as eng:
            l0, t0, _ = eng.run_mesh(mesh)
        for key in t0:
            assert rel(lanes[key], l0[key]) < 1e-11, key
            assert rel(total[key], t0[key]) < 1e-11, key
    say(f"  graph {ms1:.2f} ms / {ms2:.2f} ms, eigensolver threads {tuning}")
