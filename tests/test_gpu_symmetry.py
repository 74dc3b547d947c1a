"""Symmetry-reduced meshes on the GPU: the weighted point-list k-loop (wnlqg_run_points) against the
oracle, and the reduced integral (representatives x multiplicities, projected) against the engine's own
full-mesh integral, which the mesh tests hold to the oracle.  Bar: 1e-10 relative (FP64)."""
import sys

import numpy as np
import pytest

from oracle import kloop
from wanniernlqg_amd import engine as E
from wanniernlqg_amd import symmetry as S
from wanniernlqg_amd.model import synthetic_model

pytestmark = pytest.mark.gpu

C4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
MX = np.diag([-1, 1, 1])
C2 = np.diag([-1, -1, 1])
E3 = np.eye(3, dtype=int)
GENERATORS = {"C4v": [(C4, False), (MX, False)], "C2v+T": [(C2, False), (MX, False), (E3, True)]}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def task_kw():
    return dict(photon_energies=np.linspace(0.2, 5.0, 60), broadening=0.05, fermi_energy=0.05, shift_current=True,
                injection_current=True)


def test_run_points_against_oracle():
    model = synthetic_model(12, 1, 4, position_scale=0.1)
    rng = np.random.default_rng(1)
    k = rng.uniform(-1, 1, (301, 3))
    w = rng.uniform(0.1, 2.0, 301)
    nl = 7
    say(f"[run_points] n=12, m=301, lanes={nl}")
    with E.Engine(model, E.EngineTask(**task_kw())) as eng:
        lanes, total = eng.run_points(k, w, nl)
    kw = task_kw()
    task = kloop.Task(photon_energies=kw["photon_energies"], broadening=0.05, fermi_energy=0.05, shift_current=True,
                      injection_current=True)
    vals = kloop.point_values(model, task, k)
    for key in ("sc", "ic"):
        for lane in range(nl):
            a, b = kloop.lane_range(301, nl, lane)
            want = np.einsum("k,k...->...", w[a:b], vals[key][a:b])
            assert rel(lanes[key][lane], want) < 1e-10, (key, lane)
        assert rel(total[key], np.einsum("k,k...->...", w, vals[key])) < 1e-10


@pytest.mark.parametrize("group,n,rmax,mesh,batch", [("C4v", 8, 2, (8, 8, 6), 0), ("C4v", 16, 1, (6, 6, 4), 0),
                                                     ("C2v+T", 10, 1, (6, 4, 4), 0), ("C4v", 8, 1, (16, 16, 3), 128),
                                                     ("C4v", 16, 1, (16, 16, 4), 256)])
def test_reduced_mesh_equals_full_mesh(group, n, rmax, mesh, batch):
    """batch 0: whole planes per batch (several planes in one); batch 128: a 16 x 16 plane split over
    two batches (the second reuses the first's first contraction stage); batch 256 with n = 16: one plane
    per batch, eigensolver started from the previous plane's representatives."""
    base = synthetic_model(n, rmax, 9, position_scale=0.1)
    ops = S.group_from_generators(GENERATORS[group], np.asarray(base.lattice))
    model = S.symmetrize_model(base, ops)
    say(f"[symmetry] {group} n={n} mesh={mesh} batch={batch}")
    with E.Engine(model, E.EngineTask(batch_size=batch, **task_kw())) as eng:
        cov = S.check_covariance(lambda kk: eng.eval_points(kk), ops, samples=3)
        _, full, _ = eng.run_mesh(mesh)
        lp, red, orb = S.integrate_symmetric(eng, mesh, ops, n_lanes=5, path="planes")
        lq, red_pts, _ = S.integrate_symmetric(eng, mesh, ops, n_lanes=5, path="points")
    say(f"  covariance residual {cov:.2e}; {len(orb.representatives)} of {orb.full_count} points evaluated")
    assert cov < 1e-9
    assert len(orb.representatives) * 2 < orb.full_count
    for key in ("sc", "ic"):
        r = rel(red[key], full[key])
        r2 = rel(red_pts[key], full[key])
        say(f"  {key}: reduced vs full rel {r:.3e} (planes), {r2:.3e} (points)")
        assert r < 1e-10 and r2 < 1e-10
        # the two interpolations of the same representatives, lane by lane
        assert rel(lp[key], lq[key]) < 1e-10


def test_run_points_rejects_shg_engines():
    model = synthetic_model(6, 1, 8, position_scale=0.1)
    with E.Engine(model, E.EngineTask(second_harmonic=True, shift_current=False, photon_energies=[0.5, 1.0],
                                      broadening=0.05, shg_components=[(0, 1, 1)], fermi_energy=0.0)) as eng:
        with pytest.raises(E.WnlqgError):
            eng.run_points(np.zeros((4, 3)), np.ones(4))


def test_identity_group_reproduces_mesh_lanes():
    """With the identity alone every point represents itself with weight 1/N_k: the reduced path's lanes
    are the mesh lanes (same partition), up to where the weight enters the product."""
    model = synthetic_model(16, 1, 2, position_scale=0.1)
    ops = S.group_from_generators([], np.asarray(model.lattice))
    mesh = (6, 5, 4)
    say("[symmetry] identity group, n=16, mesh (6, 5, 4)")
    with E.Engine(model, E.EngineTask(**task_kw())) as eng:
        lanes, total, _ = eng.run_mesh(mesh)
        lr, tr, orb = S.integrate_symmetric(eng, mesh, ops, n_lanes=lanes["sc"].shape[0], path="planes")
        lp, tp, _ = S.integrate_symmetric(eng, mesh, ops, n_lanes=lanes["sc"].shape[0], path="points")
        # one point, one lane
        l1, t1 = eng.run_points(np.array([[0.1, 0.2, 0.3]]), np.array([0.5]), 1)
        v1 = eng.eval_points(np.array([[0.1, 0.2, 0.3]]))
    assert len(orb.representatives) == int(np.prod(mesh))
    for key in ("sc", "ic"):
        assert rel(lr[key], lanes[key]) < 1e-12
        assert rel(lp[key], lanes[key]) < 1e-12
        assert rel(tr[key], total[key]) < 1e-12
        assert rel(t1[key], 0.5 * v1[key][0]) < 1e-12
