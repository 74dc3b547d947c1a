"""Learned adaptive refinement of the dense-mesh Berry-curvature dipole (the learned-proposal form).

The reference has no learned component and no local refinement; this module is new and is judged
against the engine's own uniform dense mesh.  Scheme:

* Cells.  A Gamma-centred N^3 mesh (``integral_kpoint!``, reference
  ``src/Core/CalculationUtils.jl:280-292``) is read as N^3 cells centred on the mesh points; the
  plain estimate is the mean of the integrand over the cell centres.
* Refinement.  A refined cell is replaced by the mean over its r^3 sub-cell centres, r odd, which
  are exactly the points of the uniform (N r)^3 mesh inside the cell (r = 3, 5, ..., 15).  Refining
  every cell therefore reproduces the uniform (N r)^3 result, which serves as both the fallback and
  the judge of the error.  The sub-cell sums run on the GPU (``Engine.run_points`` with one lane per
  cell), so refinements of 15 per axis keep the sub-cell moments on the device.
* Features.  Every evaluation of the integrand also returns the band energies of the same point
  (``Engine.eval_points(..., energies=True)``).  A cell is described by eight numbers from the base
  mesh alone, in units that do not depend on the mesh spacing: where the nearest Fermi surface lies
  (|E_n - mu| over the band's central-difference slope, in cell units), the thickness of its thermal
  shell (the Fermi-surface energy scale over that slope), how deep the trilinear interpolation of
  E_n - mu over the cell's centre and corners straddles mu, the smallest spacing of adjacent levels
  near mu at the centre and over the 3 x 3 x 3 neighbourhood, the integrand norm at the centre and
  its neighbourhood maximum, and the neighbourhood variance of the norm (the non-learned ranking's
  own signal).
* Proposal.  The 8 -> 64 -> 64 -> 1 SiLU network (``ProposalNetwork``, trained and evaluated by the
  library's HIP kernels) maps the standardized features to log10 of the cell's refinement gain
  |mean over its sub-cells - centre value| relative to the integral.  Training targets come from
  coarse meshes refined everywhere (cheap: a few percent of the judged dense mesh).  Cells of the
  base mesh are refined in order of the predicted gain.
* The integrand is always the exact per-point BCD moment of the engine.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from .mlp import ProposalNetwork

KB_EV = 8.617333262145e-5          # Boltzmann constant, eV/K (Core/Constants.jl)
GAP_WINDOW_EV = 0.25               # levels whose midpoint lies within this of mu count for the gap feature
SURFACE_CAP = 20.0                 # Fermi-surface distances (cell units) beyond this are "far"
NORM_FLOOR = -12.0                 # log10 ||D|| / training maximum below this counts as zero
VARIANCE_FLOOR = -30.0             # the same for the neighbourhood variance (relative to the maximum^2)
TARGET_DECADES = 15.0              # training gains more than this many decades below the largest are floored
FEATURES = ("log10 Fermi-surface distance", "log10 thermal shell", "log10 gap at the centre",
            "log10 gap in the neighbourhood", "log10 norm, neighbourhood maximum", "log10 norm at the centre",
            "log10 neighbourhood variance", "Fermi-surface crossing depth")


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


def shifted(values: np.ndarray, n: int, d) -> np.ndarray:
    """values at the mesh point i + d (periodic), d = (d1, d2, d3) in mesh steps; values [n^3, ...], axis 1
    fastest."""
    f = values.reshape((n, n, n) + values.shape[1:])      # [i3, i2, i1, ...]
    return np.roll(f, (-d[2], -d[1], -d[0]), axis=(0, 1, 2)).reshape(values.shape)


def band_slopes(energies: np.ndarray, n: int) -> np.ndarray:
    """|grad E_n| times the mesh step (eV per cell) of every band at every point: central differences over
    the periodic mesh, [n^3, n_bands] (levels in ascending order)."""
    e = np.asarray(energies, dtype=np.float64)
    sq = np.zeros_like(e)
    for a in range(3):
        d = [0, 0, 0]
        d[a] = 1
        g = 0.5 * (shifted(e, n, d) - shifted(e, n, [-x for x in d]))
        sq += g * g
    return np.sqrt(sq)


def fermi_surface_features(energies: np.ndarray, n: int, mu: float, scale: float) -> np.ndarray:
    """Per cell [n^3, 3]: the distance of the nearest Fermi surface from the centre in cell units
    (min_n |E_n - mu| / slope_n, capped at SURFACE_CAP), the thermal shell thickness of that band in cell
    units (scale / slope), and the crossing depth max_n min(max_c (E_n - mu), max_c (mu - E_n)) / slope_n
    over the centre and the 8 corners c of the cell (a corner value is the mean of the 8 mesh points
    around it: trilinear interpolation), positive when the interpolated level straddles mu inside the
    cell, clipped to [-5, 2]."""
    e = np.asarray(energies, dtype=np.float64) - mu
    slope = band_slopes(energies, n) + 1e-12
    ratio = np.abs(e) / slope
    j = ratio.argmin(axis=1)
    rows = np.arange(e.shape[0])
    dist = np.minimum(ratio[rows, j], SURFACE_CAP)
    shell = scale / slope[rows, j]
    lo = e.copy()
    hi = e.copy()
    for s3 in (-1, 1):
        for s2 in (-1, 1):
            for s1 in (-1, 1):
                corner = np.zeros_like(e)
                for a in (0, 1):
                    for b in (0, 1):
                        for c in (0, 1):
                            corner += shifted(e, n, (c * s1, b * s2, a * s3))
                corner *= 0.125
                lo = np.minimum(lo, corner)
                hi = np.maximum(hi, corner)
    depth = np.clip((np.minimum(hi, -lo) / slope).max(axis=1), -5.0, 2.0)
    return np.stack([dist, shell, depth], axis=1)


def cell_features(energies: np.ndarray, norms: np.ndarray, n: int, mu: float, scale: float,
                  norm_ref: float) -> np.ndarray:
    """Raw network inputs [n^3, 8] of the n^3 cells, in the order of FEATURES (see the module notes);
    the network sees them standardized with the training set's mean and spread."""
    fs = fermi_surface_features(energies, n, mu, scale)
    pf = point_spectral_features(energies, mu, scale)
    norms = np.asarray(norms, dtype=np.float64)
    ln = np.maximum(np.log10(np.maximum(norms, 1e-300)) - np.log10(norm_ref), NORM_FLOOR)
    lv = np.maximum(local_log_variance(norms, n, relative_floor=1e-12) - 2.0 * np.log10(norm_ref), VARIANCE_FLOOR)
    x = np.empty((n ** 3, len(FEATURES)))
    x[:, 0] = np.log10(fs[:, 0] + 1e-3)
    x[:, 1] = np.log10(fs[:, 1])
    x[:, 2] = pf[:, 0]
    x[:, 3] = neighbourhood(pf[:, 0], n, np.minimum)
    x[:, 4] = neighbourhood(ln, n, np.maximum)
    x[:, 5] = ln
    x[:, 6] = lv
    x[:, 7] = fs[:, 2]
    return x


def refinement_gains(base: np.ndarray, refined: np.ndarray, mu_index: int = 0) -> np.ndarray:
    """|refined mean - centre value| per cell (Frobenius norm over the 27 components at one chemical
    potential)."""
    d = (np.asarray(refined) - np.asarray(base))[:, mu_index].reshape(base.shape[0], -1)
    return np.sqrt(np.sum(d * d, axis=1))


def error_curve(order: np.ndarray, base: np.ndarray, refined: np.ndarray) -> np.ndarray:
    """Relative error (Frobenius norm over all components) of the estimate that refines the first b cells
    of ``order``, against the estimate that refines every cell (the uniform (N r)^3 mesh), for
    b = 0 .. N: [N + 1]."""
    n = base.shape[0]
    delta = (np.asarray(refined) - np.asarray(base)).reshape(n, -1)[np.asarray(order)]
    total = np.asarray(refined).reshape(n, -1).sum(axis=0)
    rem = delta.sum(axis=0)[None, :] - np.vstack([np.zeros((1, delta.shape[1])), np.cumsum(delta, axis=0)])
    return np.linalg.norm(rem, axis=1) / np.linalg.norm(total)


def cells_needed(curve: np.ndarray, tol: float) -> int:
    """Smallest refined-cell count b from which the error stays below tol for every larger count."""
    bad = np.flatnonzero(~(curve < tol))
    return 0 if bad.size == 0 else int(bad[-1] + 1)


def variance_order(n: int, base: np.ndarray, mu_index: int = 0) -> np.ndarray:
    """The non-learned ranking: cells in decreasing neighbourhood variance of the base-mesh integrand norm."""
    return np.argsort(-local_log_variance(point_scalar(base, mu_index), n), kind="stable")


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
        self.x_mean = np.zeros(len(FEATURES))
        self.x_std = np.ones(len(FEATURES))
        self.target_shift = 0.0

    def point_data(self, k: np.ndarray):
        """Exact BCD moments and band energies at k (one evaluation per point)."""
        self.evaluations += k.shape[0]
        vals, energies = self.engine.eval_points(k, energies=True)
        return vals["bcd"], energies

    def point_moments(self, k: np.ndarray) -> np.ndarray:
        self.evaluations += k.shape[0]
        return self.engine.eval_points(k)["bcd"]

    def refined_cell_means(self, cells: np.ndarray, n: int, r: int, max_points: int = 1 << 24) -> np.ndarray:
        """Mean BCD moments over the r^3 sub-cell centres of each listed cell of the n^3 mesh:
        [len(cells), n_mu, 3, 3, 3].  The sums run on the device (``Engine.run_points``, one lane per cell,
        weights 1/r^3), so the sub-cell moments never leave the GPU; at most ``max_points`` nodes per call.
        Counted as len(cells) r^3 evaluations."""
        cells = np.asarray(cells, dtype=np.int64).reshape(-1)
        r3 = r ** 3
        per = max(1, max_points // r3)
        out = []
        for c0 in range(0, cells.size, per):
            cc = cells[c0:c0 + per]
            k = subcell_points(cc, n, r)
            lanes, _ = self.engine.run_points(k, np.full(k.shape[0], 1.0 / r3), cc.size)
            out.append(lanes["bcd"])
        self.evaluations += cells.size * r3
        if not out:
            nmu = np.asarray(self.engine.task.fermi_energies).reshape(-1).shape[0]
            return np.zeros((0, nmu, 3, 3, 3))
        return np.concatenate(out)

    def features(self, n: int, moments: np.ndarray, energies: np.ndarray) -> np.ndarray:
        """Standardized network inputs [n^3, 8] (float32) of the n^3 cells from the base-mesh data."""
        if self.norm_ref is None:
            raise RuntimeError("train the proposal network first (it fixes the integrand scale)")
        x = cell_features(energies, point_scalar(moments, self.mu_index), n, self.mu, self.scale, self.norm_ref)
        return ((x - self.x_mean) / self.x_std).astype(np.float32)

    def coarse_dataset(self, n: int, r: int):
        """Coarse n^3 run refined everywhere by r: base moments, energies and the log10 refinement gain of
        every cell relative to the norm of the refined integral (n^3 r^3 evaluations, on the device)."""
        vals, energies = self.point_data(mesh_points(n))
        means = self.refined_cell_means(np.arange(n ** 3), n, r)
        self.evaluations -= n ** 3          # the refined means contain the centres
        gains = refinement_gains(vals, means, self.mu_index)
        total = np.linalg.norm(means[:, self.mu_index].mean(axis=0))
        if not total > 0.0:
            raise RuntimeError("the integrand vanishes on the training meshes")
        return vals, energies, np.log10(gains / total + 1e-300)

    def train(self, meshes: Sequence[int] = (16, 20, 24), r: int = 5, steps: int = 6000, batch: int = 4096,
              lr: float = 2e-3, seed: int = 0):
        """Fit the network to the refinement gains of coarse meshes refined everywhere by r; targets more
        than TARGET_DECADES below the largest are floored there (they never decide a ranking).  Returns the
        per-step loss history."""
        runs = [(n,) + self.coarse_dataset(n, r) for n in meshes]
        self.norm_ref = max(float(point_scalar(v, self.mu_index).max()) for _, v, _, _ in runs)
        if not self.norm_ref > 0.0:
            raise RuntimeError("the integrand vanishes on the training meshes")
        x = np.concatenate([cell_features(e, point_scalar(v, self.mu_index), n, self.mu, self.scale, self.norm_ref)
                            for n, v, e, _ in runs])
        self.x_mean = x.mean(axis=0)
        self.x_std = x.std(axis=0) + 1e-9
        t = np.concatenate([t for _, _, _, t in runs])
        t = np.maximum(t, t.max() - TARGET_DECADES)
        self.target_shift = float(np.median(t))
        xs = ((x - self.x_mean) / self.x_std).astype(np.float32)
        return self.network.train(xs, (t - self.target_shift).astype(np.float32), steps, batch, lr, seed)

    def scores(self, n: int, moments: np.ndarray, energies: np.ndarray) -> np.ndarray:
        """Predicted log10 refinement gain (minus the training median) of the n^3 cells from the base data."""
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
            vals[cells] = self.refined_cell_means(cells, n, r)
            self.evaluations -= cells.size      # each refined cell's centre is one of its sub-cell centres
        est = vals.mean(axis=0) / self.volume
        return AdaptiveResult(est, self.evaluations - before, int(cells.size))
