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
* Proposal.  The 3 -> 64 -> 64 -> 1 network (``ProposalNetwork``, trained and evaluated by the
  library's HIP kernels) maps a cell centre (fractional coordinates scaled to [-1, 1)) to
  log10 of the local integrand variance.  Training targets come from plain runs on coarser meshes
  (variance of ||D(k)|| over each cell's periodic 3 x 3 x 3 neighbourhood).  Cells are refined in
  order of the predicted score.
* The integrand is always the exact per-point BCD moment of the engine (``Engine.eval_points``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from .mlp import ProposalNetwork


def mesh_points(n: int) -> np.ndarray:
    """Gamma-centred n^3 mesh, axis 1 fastest: k = (i1, i2, i3)/n."""
    idx = np.arange(n ** 3)
    return np.stack([idx % n, (idx // n) % n, idx // (n * n)], axis=1) / float(n)


def cell_features(n: int) -> np.ndarray:
    """Cell centres of the n^3 mesh as network inputs in [-1, 1)."""
    return (2.0 * mesh_points(n) - 1.0).astype(np.float32)


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


@dataclass
class AdaptiveResult:
    estimate: np.ndarray      # [n_mu, 3, 3, 3] in the units of Engine.run_mesh's total
    evaluations: int
    refined_cells: int


class LearnedRefinement:
    """BCD engine + proposal network."""

    def __init__(self, engine, network: Optional[ProposalNetwork] = None, seed: int = 0):
        if getattr(engine, "states", None) != ["bcd"]:
            raise ValueError("the learned refinement drives a Berry-curvature dipole engine")
        self.engine = engine
        self.network = network if network is not None else ProposalNetwork(seed)
        self.volume = abs(float(np.linalg.det(engine.model.lattice)))
        self.evaluations = 0

    def point_moments(self, k: np.ndarray) -> np.ndarray:
        self.evaluations += k.shape[0]
        return self.engine.eval_points(k)["bcd"]

    def coarse_dataset(self, n: int, mu_index: int = 0, relative_floor: float = 1.0e-12):
        vals = self.point_moments(mesh_points(n))
        target = local_log_variance(point_scalar(vals, mu_index), n, relative_floor=relative_floor)
        return cell_features(n), target.astype(np.float32)

    def train(self, meshes: Sequence[int] = (32, 48), steps: int = 3000, batch: int = 4096, lr: float = 3e-3,
              seed: int = 0, mu_index: int = 0):
        xs, ts = [], []
        for n in meshes:
            x, t = self.coarse_dataset(n, mu_index)
            xs.append(x)
            ts.append(t)
        x = np.concatenate(xs)
        t = np.concatenate(ts)
        self.target_shift = float(np.median(t))
        hist = self.network.train(x, t - self.target_shift, steps, batch, lr, seed)
        return hist

    def scores(self, n: int) -> np.ndarray:
        return self.network.predict(cell_features(n))

    def estimate(self, n: int, r: int, budget_cells: int, order: Optional[np.ndarray] = None,
                 base: Optional[np.ndarray] = None) -> AdaptiveResult:
        """Base n^3 cell centres; the first ``budget_cells`` cells of ``order`` (default: the network's
        ranking) refined by r; the mean over the cells, divided by |det lattice| like run_mesh's total."""
        before = self.evaluations
        if base is None:
            base = self.point_moments(mesh_points(n))
        else:
            self.evaluations += base.shape[0]
        if order is None:
            order = np.argsort(-self.scores(n), kind="stable")
        cells = np.asarray(order[:budget_cells])
        vals = base.copy()
        if cells.size:
            sub = self.point_moments(subcell_points(cells, n, r))
            r3 = r ** 3
            vals[cells] = sub.reshape((cells.size, r3) + sub.shape[1:]).mean(axis=1)
            self.evaluations -= cells.size      # each refined cell's centre is one of its sub-cell centres
        est = vals.mean(axis=0) / self.volume
        return AdaptiveResult(est, self.evaluations - before, int(cells.size))
