"""Symmetry-reduced meshes (wanniernlqg_amd/symmetry.py) on the CPU.

Pinned by: orbits against brute-force enumeration of every point's images; the projector's algebra
(idempotent, symmetric, the component count a point group allows: 4 for SC under C4v); covariance of
symmetrized models' point values under every operation, computed by the oracle (oracle/kloop.py), which
fixes the action conventions (rotation of k, field-index exchange and sign of antiunitary operations);
and the reduction identity itself on the oracle: the full-mesh sum equals the projected, multiplicity-
weighted sum over the representatives."""
import numpy as np
import pytest

from oracle import kloop
from wanniernlqg_amd import symmetry as S
from wanniernlqg_amd.model import synthetic_model

C4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
MX = np.diag([-1, 1, 1])
C2 = np.diag([-1, -1, 1])
E3 = np.eye(3, dtype=int)


def groups(lattice):
    return {
        "C4v": S.group_from_generators([(C4, False), (MX, False)], lattice),
        "C2v+T": S.group_from_generators([(C2, False), (MX, False), (E3, True)], lattice),
        "T": S.group_from_generators([(E3, True)], lattice),
    }


def test_group_closure_and_validation():
    lat = 4.0 * np.eye(3)
    g = groups(lat)
    assert [len(g[k]) for k in ("C4v", "C2v+T", "T")] == [8, 8, 2]
    for ops in g.values():
        S.validate_group(ops)
        assert np.array_equal(ops[0].rotation_fractional, np.eye(3)) and not ops[0].antiunitary
    with pytest.raises(ValueError):
        S.validate_group(g["C4v"][:5])
    # a rotation that does not close a 4 x 3 mesh
    with pytest.raises(ValueError):
        S.integer_action(S.operation(C4, lat), (4, 3, 2))


@pytest.mark.parametrize("name,mesh", [("C4v", (4, 4, 3)), ("C2v+T", (5, 4, 3)), ("T", (3, 3, 2)),
                                       ("C4v", (6, 6, 1))])
def test_orbits_match_brute_force(name, mesh):
    ops = groups(4.0 * np.eye(3))[name]
    orb = S.kmesh_orbits(mesh, ops)
    perms = [S.mesh_permutation(op, mesh) for op in ops]
    N = int(np.prod(mesh))
    seen = {}
    for i in range(N):
        orbit = {i}
        frontier = [i]
        while frontier:
            j = frontier.pop()
            for p in perms:
                if p[j] not in orbit:
                    orbit.add(int(p[j]))
                    frontier.append(int(p[j]))
        seen[min(orbit)] = len(orbit)
    assert list(orb.representatives) == sorted(seen)
    assert list(orb.multiplicities) == [seen[r] for r in sorted(seen)]
    assert orb.full_count == N and orb.multiplicities.sum() == N


def test_projector_algebra():
    ops = groups(4.0 * np.eye(3))["C4v"]
    for q in ("sc", "ic"):
        P = S.response_projector(q, ops)
        assert np.allclose(P @ P, P, atol=1e-14) and np.allclose(P, P.T, atol=1e-14)
        # C4v allows zzz, zxx = zyy, xxz = yyz, xzx = yzy
        assert abs(np.trace(P) - 4) < 1e-12
    for M1, M2 in [(S.action_matrix("sc", a), S.action_matrix("sc", b)) for a in ops[:3] for b in ops[:3]]:
        assert np.allclose(M1 @ M1.T, np.eye(27), atol=1e-14)


def _model(name, seed):
    base = synthetic_model(4, 1, seed, position_scale=0.1)
    ops = groups(np.asarray(base.lattice))[name]
    return S.symmetrize_model(base, ops), ops


def _evaluate(model, task):
    return lambda k: kloop.point_values(model, task, np.asarray(k))


@pytest.mark.parametrize("name", ["C4v", "C2v+T", "T"])
def test_covariance_of_symmetrized_models(name):
    model, ops = _model(name, 3)
    task = kloop.Task(photon_energies=np.linspace(0.5, 4.0, 12), broadening=0.1, fermi_energy=0.05,
                      temperature=300.0, shift_current=True, injection_current=True)
    assert S.check_covariance(_evaluate(model, task), ops, samples=3) < 1e-9
    # the unsymmetrized model is not covariant
    raw = synthetic_model(4, 1, 3, position_scale=0.1)
    assert S.check_covariance(_evaluate(raw, task), ops, samples=3) > 1e-3


@pytest.mark.parametrize("name,mesh", [("C4v", (4, 4, 3)), ("C2v+T", (4, 3, 2)), ("T", (3, 3, 2))])
def test_reduction_identity_on_the_oracle(name, mesh):
    model, ops = _model(name, 5)
    task = kloop.Task(photon_energies=np.linspace(0.5, 4.0, 10), broadening=0.1, fermi_energy=0.05,
                      shift_current=True, injection_current=True)
    _, full = kloop.integrate_mesh(model, task, mesh)
    orb = S.kmesh_orbits(mesh, ops)
    k = kloop.integral_kpoints(mesh)[orb.representatives]
    vals = kloop.point_values(model, task, k)
    w = orb.multiplicities / orb.full_count
    assert len(orb.representatives) < np.prod(mesh)
    for q in ("sc", "ic"):
        red = S.project(np.einsum("k,k...->...", w, vals[q]), S.response_projector(q, ops))
        err = np.linalg.norm(red - full[q]) / np.linalg.norm(full[q])
        assert err < 1e-12, (q, err)
