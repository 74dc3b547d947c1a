"""Symmetry-reduced dense meshes for the shift and injection current: orbit representatives of the
Gamma-centred mesh, integrated with their multiplicities on the GPU, then projected onto the tensors the
point group allows.

Follows the reference's response-symmetry path for Conventional SC / IC integrals:

* ``kmesh_orbits``: the exact modular action of each operation on the mesh indices
  (``_response_symmetry_integer_action`` / ``_response_symmetry_permutation``,
  ``src/Runtime/ResponseSymmetryRuntime.jl:383-442``: reciprocal rotation inv(W)^T, negated for an
  antiunitary operation, required to map the mesh onto itself) and the orbits it generates, each
  represented by its smallest linear index, with its size as multiplicity
  (``response_symmetry_kmesh_orbits``, ``:463-496``).  The orbits are found by label propagation over
  the permutations (vectorized) instead of a serial union-find; both give the connected components
  with their minimum index;
* the k-loop over the representatives with weight multiplicity / N_k (``consume_point!``,
  ``src/Runtime/Execution/IntegralDriver.jl:822-831``) runs on the GPU, either with whole k3 planes
  interpolated by the separable contraction and only the representatives' rows diagonalized and
  contracted (``Engine.run_mesh_reduced``), or with the representatives' k points interpolated alone
  (``Engine.run_points``);
* ``response_projector``: the group average of the final-tensor actions R (x) R (x) R, with the two field
  indices exchanged and the sign +1 (SC) / -1 (IC) for antiunitary operations
  (``response_symmetry_action_matrix`` / ``response_symmetry_projector``,
  ``src/Responses/Shared/ResponseSymmetry.jl:44-147``), applied to the real and imaginary parts of the
  total (``apply_response_symmetry!``, ``ResponseSymmetryRuntime.jl:1121-1146``).  The projection is
  applied directly (P x) rather than through an orthonormal basis of its range (B B^T x); for a group
  average of orthogonal actions both are the same orthogonal projector.

The reduction is exact when the integrand is covariant, value(g k) = A_g value(k) for every operation
(``check_covariance`` measures it on sample points, the role of
``validate_response_integrand_covariance``).  ``symmetrize_model`` builds covariant test models (orbitals
at the origin) by averaging H(R) and A(R) over the group.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

# antiunitary field-exchange sign of the final tensor (response_symmetry_sign, ResponseSymmetry.jl:4-10)
ANTIUNITARY_SIGN = {"sc": 1.0, "ic": -1.0}


@dataclass(frozen=True)
class SymmetryOperation:
    """A point-group operation: W acts on fractional real-space coordinates (x' = W x), R on Cartesian
    vectors; antiunitary marks the operation combined with time reversal (``ResponseSymmetryOperation``,
    ``src/IO/ResponseSymmetryArtifacts.jl:10-15``; the translation does not enter the integrals)."""
    rotation_fractional: np.ndarray
    rotation_cartesian: np.ndarray
    antiunitary: bool = False

    def key(self):
        return (tuple(int(v) for v in np.asarray(self.rotation_fractional).ravel()), bool(self.antiunitary))


def operation(rotation_fractional, lattice, antiunitary: bool = False) -> SymmetryOperation:
    """Operation from its fractional rotation; R = L^T W L^-T for lattice vectors as the rows of L."""
    W = np.asarray(rotation_fractional, dtype=np.int64).reshape(3, 3)
    L = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
    R = L.T @ W @ np.linalg.inv(L.T)
    return SymmetryOperation(W, R, bool(antiunitary))


def group_from_generators(generators, lattice) -> List[SymmetryOperation]:
    """Closure of (fractional rotation, antiunitary) generators under products, identity first."""
    elems = {(tuple(np.eye(3, dtype=np.int64).ravel()), False)}
    gens = [(tuple(int(v) for v in np.asarray(W, dtype=np.int64).ravel()), bool(a)) for W, a in generators]
    frontier = list(elems)
    while frontier:
        nxt = []
        for (w1, a1) in frontier:
            for (w2, a2) in gens:
                prod = (tuple(int(v) for v in (np.array(w1).reshape(3, 3) @ np.array(w2).reshape(3, 3)).ravel()),
                        a1 != a2)
                if prod not in elems:
                    elems.add(prod)
                    nxt.append(prod)
        frontier = nxt
        if len(elems) > 96:
            raise ValueError("generators do not close into a finite point group")
    ordered = sorted(elems, key=lambda e: (e != (tuple(np.eye(3, dtype=np.int64).ravel()), False), e[1], e[0]))
    return [operation(np.array(w).reshape(3, 3), lattice, a) for w, a in ordered]


def validate_group(ops: Sequence[SymmetryOperation]) -> None:
    """Identity present, every product present, no duplicates, orthogonal Cartesian rotations."""
    keys = [op.key() for op in ops]
    if len(set(keys)) != len(keys):
        raise ValueError("duplicate symmetry operations")
    if (tuple(np.eye(3, dtype=np.int64).ravel()), False) not in keys:
        raise ValueError("the operations must contain the identity")
    for op in ops:
        R = np.asarray(op.rotation_cartesian, dtype=np.float64)
        if np.max(np.abs(R.T @ R - np.eye(3))) > 1e-8:
            raise ValueError("rotation_cartesian is not orthogonal")
    table = set(keys)
    for a, b in itertools.product(ops, ops):
        prod = (tuple(int(v) for v in (np.asarray(a.rotation_fractional) @ np.asarray(b.rotation_fractional)).ravel()),
                a.antiunitary != b.antiunitary)
        if prod not in table:
            raise ValueError("the operations do not form a group (a product is missing)")


def integer_action(op: SymmetryOperation, mesh: Sequence[int]) -> np.ndarray:
    """Integer matrix acting on zero-based mesh indices: i' = A i mod N."""
    W = np.asarray(op.rotation_fractional, dtype=np.int64)
    recip = np.rint(np.linalg.inv(W).T).astype(np.int64)
    if not np.array_equal(W.T @ recip, np.eye(3, dtype=np.int64)):
        raise ValueError("fractional rotation has no exact integer reciprocal inverse")
    if op.antiunitary:
        recip = -recip
    mesh = [int(n) for n in mesh]
    action = np.zeros((3, 3), dtype=np.int64)
    for o in range(3):
        for i in range(3):
            num = mesh[o] * recip[o, i]
            if num % mesh[i]:
                raise ValueError(f"operation does not close k_mesh={tuple(mesh)}: coefficient ({o},{i}) = "
                                 f"{num}/{mesh[i]} is not an integer")
            action[o, i] = num // mesh[i]
    return action


def mesh_indices(mesh: Sequence[int]) -> np.ndarray:
    """Zero-based mesh coordinates [3, N] of the linear indices, axis 1 fastest."""
    n1, n2, n3 = (int(n) for n in mesh)
    idx = np.arange(n1 * n2 * n3, dtype=np.int64)
    return np.stack([idx % n1, (idx // n1) % n2, idx // (n1 * n2)])


def mesh_permutation(op: SymmetryOperation, mesh: Sequence[int]) -> np.ndarray:
    action = integer_action(op, mesh)
    n = np.array([int(v) for v in mesh], dtype=np.int64)
    mapped = np.mod(action @ mesh_indices(mesh), n[:, None])
    perm = mapped[0] + n[0] * mapped[1] + n[0] * n[1] * mapped[2]
    if np.any(np.bincount(perm, minlength=perm.shape[0]) != 1):
        raise ValueError("operation is not a permutation of the mesh")
    return perm


@dataclass
class MeshOrbits:
    representatives: np.ndarray   # sorted linear indices (0-based)
    multiplicities: np.ndarray
    full_count: int


def kmesh_orbits(mesh: Sequence[int], ops: Sequence[SymmetryOperation]) -> MeshOrbits:
    mesh = tuple(int(v) for v in mesh)
    if any(v < 1 for v in mesh):
        raise ValueError("k mesh entries must be positive")
    perms = [mesh_permutation(op, mesh) for op in ops]
    N = int(np.prod(mesh))
    label = np.arange(N, dtype=np.int64)
    inverses = []
    for p in perms:
        inv = np.empty_like(p)
        inv[p] = np.arange(N, dtype=np.int64)
        inverses.append(inv)
    while True:
        old = label.copy()
        for p, inv in zip(perms, inverses):
            label = np.minimum(label, label[p])
            label = np.minimum(label, label[inv])
        while True:
            jumped = label[label]
            if np.array_equal(jumped, label):
                break
            label = jumped
        if np.array_equal(label, old):
            break
    reps = np.unique(label)
    mult = np.bincount(label, minlength=N)[reps]
    if int(mult.sum()) != N:
        raise ValueError("orbit multiplicities do not sum to the mesh")
    return MeshOrbits(reps, mult, N)


def action_matrix(quantity: str, op: SymmetryOperation) -> np.ndarray:
    """27 x 27 action on the [a, b, c] components (rightmost fastest) of an SC / IC tensor."""
    R = np.asarray(op.rotation_cartesian, dtype=np.float64)
    M = np.einsum("aA,bB,cC->abcABC", R, R, R)
    if op.antiunitary:
        M = ANTIUNITARY_SIGN[quantity] * np.transpose(M, (0, 2, 1, 3, 4, 5))
    M = M.reshape(27, 27)
    M[np.abs(M) < 1e-15] = 0.0
    return M


def response_projector(quantity: str, ops: Sequence[SymmetryOperation]) -> np.ndarray:
    if not ops:
        raise ValueError("at least one point-group operation is required")
    P = sum(action_matrix(quantity, op) for op in ops) / len(ops)
    P[np.abs(P) < 1e-15] = 0.0
    return P


def project(tensor: np.ndarray, P: np.ndarray) -> np.ndarray:
    """Apply the projector to [..., 3, 3, 3] (real and imaginary parts alike)."""
    t = np.asarray(tensor)
    flat = t.reshape(t.shape[:-3] + (27,))
    return (flat @ P.T).reshape(t.shape)


def integrate_symmetric(engine, mesh: Sequence[int], ops: Sequence[SymmetryOperation],
                        n_lanes: Optional[int] = None, path: str = "auto"):
    """Integrate the mesh over its orbit representatives only.  Returns (lanes, total, orbits): lanes over
    the representatives (unprojected, weights multiplicity / N_k), and totals projected onto the allowed
    tensors, equal to the full-mesh totals for a covariant integrand.  path "planes": whole k3 planes
    contracted by the separable interpolation, the last stage only for the representatives
    (``Engine.run_mesh_reduced``);
    "points": the representatives' k points interpolated alone (``Engine.run_points``); "auto": planes
    unless the engine's Fourier backend is "direct"."""
    validate_group(ops)
    if path not in ("auto", "planes", "points"):
        raise ValueError("path must be auto, planes or points")
    orb = kmesh_orbits(mesh, ops)
    w = orb.multiplicities / float(orb.full_count)
    if path == "auto":
        path = "points" if getattr(engine.task, "fourier_backend", "auto") == "direct" else "planes"
    if path == "planes":
        lanes, total = engine.run_mesh_reduced(mesh, orb.representatives, w, n_lanes)
    else:
        n = np.array([int(v) for v in mesh], dtype=np.float64)
        k = (mesh_indices(mesh)[:, orb.representatives] / n[:, None]).T
        lanes, total = engine.run_points(k, w, n_lanes)
    total = {s: project(v, response_projector(s, ops)) for s, v in total.items()}
    return lanes, total, orb


def image_points(k: np.ndarray, op: SymmetryOperation) -> np.ndarray:
    """Fractional k -> g k (inv(W)^T k, negated for an antiunitary operation)."""
    W = np.asarray(op.rotation_fractional, dtype=np.float64)
    recip = np.linalg.inv(W).T
    out = np.asarray(k, dtype=np.float64) @ recip.T
    return -out if op.antiunitary else out


def check_covariance(evaluate, ops: Sequence[SymmetryOperation], samples: int = 4, seed: int = 0) -> float:
    """Largest relative residual of value(g k) = A_g value(k) over random points and all operations;
    evaluate(k [m,3]) -> {state: [m, n_w, 3, 3, 3]}."""
    k = np.random.default_rng(seed).uniform(0.0, 1.0, (samples, 3))
    base = evaluate(k)
    worst = 0.0
    for op in ops:
        img = evaluate(image_points(k, op))
        for s, v in base.items():
            want = project(v, action_matrix(s, op))
            scale = max(np.max(np.abs(v)), 1e-300)
            worst = max(worst, float(np.max(np.abs(img[s] - want)) / scale))
    return worst


def symmetrize_model(model, ops: Sequence[SymmetryOperation]):
    """Average H(R) and A(R) over the group for orbitals at the origin: H'(R) = mean_g c_g H(W_g R),
    A'_a(R) = mean_g sum_b R_g[b, a] c_g A_b(W_g R), c_g complex conjugation for antiunitary g.  The
    R-vector set must be closed under the operations, with equal degeneracies on each orbit."""
    from wanniernlqg_amd.model import TightBindingModel

    validate_group(ops)
    rv = np.asarray(model.r_vectors, dtype=np.int64)
    index = {tuple(r): i for i, r in enumerate(rv)}
    deg = np.asarray(model.r_degeneracies)
    H = np.asarray(model.hamiltonian_r)
    A = None if model.position_r is None else np.asarray(model.position_r)
    Hs = np.zeros_like(H)
    As = None if A is None else np.zeros_like(A)
    for op in ops:
        W = np.asarray(op.rotation_fractional, dtype=np.int64)
        R = np.asarray(op.rotation_cartesian, dtype=np.float64)
        for i, r in enumerate(rv):
            j = index.get(tuple(W @ r))
            if j is None or deg[j] != deg[i]:
                raise ValueError("R-vector set is not closed under the operations")
            h = H[j].conj() if op.antiunitary else H[j]
            Hs[i] += h
            if A is not None:
                a = A[j].conj() if op.antiunitary else A[j]
                As[i] += np.einsum("ba,bmn->amn", R, a)
    Hs /= len(ops)
    if As is not None:
        As /= len(ops)
    return TightBindingModel(np.asarray(model.lattice), rv.astype(np.int32), deg, Hs, As,
                             name=getattr(model, "name", None))
