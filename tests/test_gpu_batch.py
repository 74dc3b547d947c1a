"""GPU checks of the batching form: several models of uneven n_w (BCD + AHC, SC + IC) captured in one
HIP graph give the results of the individual engines, replays are bitwise identical, and the
autotuner picks a supported eigensolver shape."""
import sys

import numpy as np
import pytest

from oracle import geometry as G
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
    assert all(t in (1, 2, 64, 128, 256) for t in tuning)
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


def test_batch_with_large_shg_entry_and_autotune():
    """An n_w = 72 SHG entry (V of the eigensolver in global memory, not tuned) beside an n_w = 24 SC/IC entry,
    autotune on: each entry equals wnlqg_run_mesh of the same engine settings."""
    shg = E.EngineTask(second_harmonic=True, photon_energies=np.linspace(0.2, 2.0, 12), broadening=0.05,
                       shg_components=[(0, 1, 1), (2, 2, 2)], shg_intermediate_regularization=1e-3)
    sc = E.EngineTask(photon_energies=np.linspace(0, 4, 40), broadening=0.05, shift_current=True,
                      injection_current=True)
    entries = [(synthetic_model(72, 1, 17, position_scale=0.1), shg, (3, 2, 2)),
               (synthetic_model(24, 1, 5, position_scale=0.1), sc, (4, 3, 3))]
    say("[batch] n_w = 72 SHG + n_w = 24 SC/IC, autotune on")
    with E.Batch(entries, autotune=True) as b:
        tuning = b.tuning()
        res, _ = b.run()
    assert tuning[0] == 0 and tuning[1] in (1, 2, 64, 128, 256)
    for (model, task, mesh), (lanes, total) in zip(entries, res):
        with E.Engine(model, task) as eng:
            l0, t0, _ = eng.run_mesh(mesh)
        for key in t0:
            assert rel(lanes[key], l0[key]) < 1e-11, key
            assert rel(total[key], t0[key]) < 1e-11, key


@pytest.mark.parametrize("seeds", [(0, 5, 11)])
def test_batch_configuration_models_against_oracle(seeds):
    """Batching-form models (n_w = 4, 24, 48: seeds 0, 5, 11 of the 64, |R_i| <= 4, 729 R vectors) with BCD +
    intrinsic AHC at mu = -0.5, 0, 0.5 eV run through the batch graph and compared with the oracle's
    dense-mesh integrals (oracle/geometry.py) on small meshes.  Bar: FP64, 1e-10 relative."""
    meshes = {0: (6, 5, 4), 5: (4, 4, 3), 11: (3, 3, 2)}
    entries = [(config_model(2, sd), geometry_task(), meshes[sd]) for sd in seeds]
    say(f"[batch] configuration models n_w = {[e[0].num_orbitals for e in entries]} against the oracle")
    with E.Batch(entries, autotune=True) as b:
        res, _ = b.run()
    for (model, _, mesh), (_, total) in zip(entries, res):
        task = G.GeometryTask(fermi_energies=[-0.5, 0.0, 0.5], temperature=0.0, fs_broadening=0.05,
                              denominator_regularization=0.0, degeneracy_threshold=0.0)
        for q in ("bcd", "ahc"):
            _, ref = G.integrate_geometry_mesh(model, task, mesh, q)
            assert rel(total[q], ref) < 1e-10, (model.num_orbitals, q, rel(total[q], ref))
