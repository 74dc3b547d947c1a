"""Learned adaptive refinement of the dense-mesh Berry-curvature dipole (the learned-proposal form).

The reference has no learned component and no local refinement; this module is new and is judged
only against the engine's own uniform dense mesh.  Scheme:

* Cells.  A Gamma-centred N^3 mesh (``integral_kpoint!``, reference
  ``src/Core/CalculationUtils.jl:280-292``) is read as N^3 cells centred on the mesh points; the
  plain estimate is the mean of the integrand over the cell centres.
* Refinement.  A refined cell is replaced by the mean over its r^3 sub-cell centres, r odd, which
  are exactly the points of the uniform (N r)^3 mesh inside the cell (r = 3, 5, ..., 15).  Refining
  every cell therefore reproduces the uniform (N r)^3 result, which serves as both the fallback and
  the judge of the error.
* Features.  Every evaluation of the integrand also returns the band energies of the same point
  (``Engine.eval_points(..., energies=True)``).  A cell is described by three spectral features of
  its periodic 3 x 3 x 3 neighbourhood of mesh points: the smallest spacing of adjacent levels near
  the chemical potential, the smallest distance of a level to the chemical potential in units of the
  Fermi-surface energy scale (k_B T, or the Gaussian width at T = 0), and the largest integrand norm.
  They are what makes a cell "hot" for a Fermi-surface integrand (near-degenerate pairs at the Fermi
  level), stated in units that do not depend on the mesh spacing.
* Proposal.  The 3 -> 64 -> 64 -> 1 SiLU network (``ProposalNetwork``, trained and evaluated by the
  library's HIP kernels) maps the scaled features to log10 of the cell's integrand variance.  Training
  targets come from plain runs on coarser meshes (variance of ||D(k)|| over each cell's periodic
  3 x 3 x 3 neighbourhood).  Cells of the base mesh are refined in order of the predicted score.
* The integrand is always the exact per-point BCD moment of the engine.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from .mlp import ProposalNetwork

KB_EV = 8.617333262145e-5          # Boltzmann constant, eV/K (Core/Constants.jl)
GAP_WINDOW_EV = 0.25               # levels whose midpoint lies within this of mu count for the gap feature
DISTANCE_CAP = 60.0                # |E - mu| / energy scale beyond which a level is "far"
NORM_DECADES = 8.0                 # decades of ||D|| below the training maximum mapped to [-1, 1]


def mesh_points(n: int) -> np.ndarray:
    """Gamma-centred n^3 mesh, axis 1 fastest: k = (i1, i2, i3)/n."""
    idx = np.arange(n ** 3)
    return np.stack([idx % n, (idx // n) % n, idx // (n * n)], axis=1) / float(n)


def subcell_points(cells: np.ndarray, n: int, r: int) -> np.ndarray:
    """The r^3 sub-cell centres (r odd) of each listed cell: [len(cells) * r^3, 3], cell-major."""
    if r % 2 != 1:
        raise ValueError("refinement factors are odd so that sub-cell centres lie on the (n r)^3 mesh")
    base = mesh_points(n)[np.asarray(cells)]
    off = (np.arange(r) - (r - 1) // 2) / float(n * r)
    o3 = np.stack(np.meshgrid(off, off, off, indexing="ij"), axis=-1).reshape(-1, 3)
    return (base[:, None, :] + o3[None, :, :]).reshape(-1, 3)


def point_scalar(moments: np.ndarray, mu_index: int = 0) -> np.ndarray:
    """Per-point ranking scalar ||D(k)||_F over the 27 components at one chemical potential."""
    m = moments[:, mu_index].reshape(moments.shape[0], -1)
    return np.sqrt(np.sum(m * m, axis=1))


def neighbourhood(values: np.ndarray, n: int, op: Callable) -> np.ndarray:
    """op (np.minimum or np.maximum) over each point's periodic 3 x 3 x 3 neighbourhood of the n^3 mesh."""
    f = values.reshape(n, n, n)              # [i3, i2, i1]
    out = f.copy()
    for d3 in (-1, 0, 1):
        for d2 in (-1, 0, 1):
            for d1 in (-1, 0, 1):
                out = op(out, np.roll(f, (d3, d2, d1), axis=(0, 1, 2)))
    return out.reshape(-1)


def local_log_variance(values: np.ndarray, n: int, floor: float = 1.0e-30, relative_floor: float = 0.0) -> np.ndarray:
    """log10 of the variance over each cell's periodic 3 x 3 x 3 neighbourhood (axis 1 fastest), with
    the floor max(floor, relative_floor * largest variance) added before the logarithm."""
    f = values.reshape(n, n, n)             # [i3, i2, i1]
    s1 = np.zeros_like(f)
    s2 = np.zeros_like(f)
    for d3 in (-1, 0, 1):
        for d2 in (-1, 0, 1):
            for d1 in (-1, 0, 1):
                g = np.roll(f, (d3, d2, d1), axis=(0, 1, 2))
                s1 += g
                s2 += g * g
    mean = s1 / 27.0
    var = np.maximum(s2 / 27.0 - mean * mean, 0.0)
    fl = max(floor, relative_floor * float(var.max()))
    return np.log10(var + fl).reshape(-1)


def point_spectral_features(energies: np.ndarray, mu: float, scale: float) -> np.ndarray:
    """Per point [m, 2]: log10 of the smallest spacing of adjacent levels whose midpoint lies within
    GAP_WINDOW_EV of mu (GAP_WINDOW_EV when there is none), and min_n |E_n - mu| / scale."""
    e = np.sort(np.asarray(energies, dtype=np.float64), axis=1)
    mid = 0.5 * (e[:, 1:] + e[:, :-1])
    gaps = np.where(np.abs(mid - mu) < GAP_WINDOW_EV, e[:, 1:] - e[:, :-1], np.inf)
    gap = np.minimum(gaps.min(axis=1) if gaps.shape[1] else np.full(e.shape[0], np.inf), GAP_WINDOW_EV)
    dist = np.min(np.abs(e - mu), axis=1) / scale
    return np.stack([np.log10(np.maximum(gap, 1e-300)), dist], axis=1)


def cell_features(energies: np.ndarray, norms: np.ndarray, n: int, mu: float, scale: float,
                  norm_ref: float) -> np.ndarray:
    """Scaled network inputs [n^3, 3] of the n^3 cells: neighbourhood minimum of the log gap and of
    the distance to mu, neighbourhood maximum of log10 ||D|| (relative to log10 norm_ref)."""
    pf = point_spectral_features(energies, mu, scale)
    gap = neighbourhood(pf[:, 0], n, np.minimum)
    dist = neighbourhood(pf[:, 1], n, np.minimum)
    lnorm = neighbourhood(np.log10(np.maximum(norms, 1e-300)), n, np.maximum)
    x = np.empty((n ** 3, 3), dtype=np.float32)
    x[:, 0] = np.clip((gap + 2.0) / 2.0, -1.0, 1.0)                        # 1e-4 eV -> -1, 1 eV -> 1
    x[:, 1] = np.clip(dist, 0.0, DISTANCE_CAP) / (0.5 * DISTANCE_CAP) - 1.0
    x[:, 2] = np.clip((lnorm - np.log10(norm_ref)) / NORM_DECADES + 1.0, -1.0, 1.0)
    return x


@dataclass
class AdaptiveResult:
    estimate: np.ndarray      # [n_mu, 3, 3, 3] in the units of Engine.run_mesh's total
    evaluations: int
    refined_cells: int


class LearnedRefinement:
    """BCD engine + proposal network."""

    def __init__(self, engine, network: Optional[ProposalNetwork] = None, seed: int = 0, mu_index: int = 0):
        if getattr(engine, "states", None) != ["bcd"]:
            raise ValueError("the learned refinement drives a Berry-curvature dipole engine")
        self.engine = engine
        self.network = network if network is not None else ProposalNetwork(seed)
        self.volume = abs(float(np.linalg.det(engine.model.lattice)))
        self.evaluations = 0
        self.mu_index = mu_index
        task = engine.task
        self.mu = float(np.asarray(task.fermi_energies, dtype=np.float64).reshape(-1)[mu_index])
        if task.temperature > 0.0:
            self.scale = KB_EV * float(task.temperature)
        elif task.fermi_surface_broadening > 0.0:
            self.scale = float(task.fermi_surface_broadening)
        else:
            raise ValueError("the Fermi-surface energy scale needs temperature > 0 or fermi_surface_broadening > 0")
        self.norm_ref = None
        self.target_shift = 0.0

    def point_data(self, k: np.ndarray):
        """Exact BCD moments and band energies at k (one evaluation per point)."""
        self.evaluations += k.shape[0]
        vals, energies = self.engine.eval_points(k, energies=True)
        return vals["bcd"], energies

    def point_moments(self, k: np.ndarray) -> np.ndarray:
        self.evaluations += k.shape[0]
        return self.engine.eval_points(k)["bcd"]

    def features(self, n: int, moments: np.ndarray, energies: np.ndarray) -> np.ndarray:
        if self.norm_ref is None:
            raise RuntimeError("train the proposal network first (it fixes the integrand scale)")
        return cell_features(energies, point_scalar(moments, self.mu_index), n, self.mu, self.scale, self.norm_ref)

    def coarse_dataset(self, n: int, relative_floor: float = 1.0e-12):
        """Plain n^3 run: raw data for the features and log10 local-variance targets."""
        vals, energies = self.point_data(mesh_points(n))
        target = local_log_variance(point_scalar(vals, self.mu_index), n, relative_floor=relative_floor)
        return vals, energies, target.astype(np.float32)

    def train(self, meshes: Sequence[int] = (32, 48), steps: int = 6000, batch: int = 4096, lr: float = 2e-3,
              seed: int = 0):
        runs = [(n,) + self.coarse_dataset(n) for n in meshes]
        self.norm_ref = max(float(point_scalar(v, self.mu_index).max()) for _, v, _, _ in runs)
        if not self.norm_ref > 0.0:
            raise RuntimeError("the integrand vanishes on the training meshes")
        x = np.concatenate([self.features(n, v, e) for n, v, e, _ in runs])
        t = np.concatenate([t for _, _, _, t in runs])
        self.target_shift = float(np.median(t))
        return self.network.train(x, t - self.target_shift, steps, batch, lr, seed)

    def scores(self, n: int, moments: np.ndarray, energies: np.ndarray) -> np.ndarray:
        """Predicted log10 variance (minus the training median) of the n^3 cells from the base data."""
        return self.network.predict(self.features(n, moments, energies))

    def estimate(self, n: int, r: int, budget_cells: int, order: Optional[np.ndarray] = None,
                 base: Optional[np.ndarray] = None, base_energies: Optional[np.ndarray] = None) -> AdaptiveResult:
        """Base n^3 cell centres; the first ``budget_cells`` cells of ``order`` (default: the network's
        ranking of the base cells) refined by r; the mean over the cells, divided by |det lattice| like
        run_mesh's total.  ``base``/``base_energies`` reuse an earlier evaluation of the base mesh (still
        counted as n^3 evaluations)."""
        before = self.evaluations
        if base is None:
            base, base_energies = self.point_data(mesh_points(n))
        else:
            self.evaluations += base.shape[0]
        if order is None:
            if base_energies is None:
                raise ValueError("the network's ranking needs the base energies")
            order = np.argsort(-self.scores(n, base, base_energies), kind="stable")
        cells = np.asarray(order[:budget_cells])
        vals = base.copy()
        if cells.size:
            sub = self.point_moments(subcell_points(cells, n, r))
            r3 = r ** 3
            vals[cells] = sub.reshape((cells.size, r3) + sub.shape[1:]).mean(axis=1)
            self.evaluations -= cells.size      # each refined cell's centre is one of its sub-cell centres
        est = vals.mean(axis=0) / self.volume
        return AdaptiveResult(est, self.evaluations - before, int(cells.size))
