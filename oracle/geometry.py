"""NumPy FP64 restatement of the Berry-curvature dipole (BCD, Fermi-surface form) and the
intrinsic anomalous Hall conductivity (AHC, Matrix method) k-loops (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module.

Restated reference functions (paths relative to the reference root):

* Wannier curvature F_ab = dA_b/dk_a - dA_a/dk_b - i[A_a, A_b], rotated U^+ F U
  -- ``src/MatrixElements/Interpolated/PositionMatrixElements.jl:491-540`` (channel order of
  ``_pair_channel``: position direction first, derivative direction second,
  ``MatrixElementPlan.jl:201-215``; gradient channel = i R_d A_p, ``PositionMatrixElements.jl:20-80``).
* Band Berry curvature Omega^n_ab = 2 sum_{m != n} -Im(A^a_nm A^b_mn) + Re F^H_nn,ab
  -- ``src/Responses/QuantumGeometry/ConventionalQuantumGeometry.jl:167-181``.
* Spectral isolation check -- ``MatrixElements/Interpolated/OrthogonalBandDerivatives.jl:66-82``
  (called for every BCD point, ``BerryMultipoleGeometry.jl:81``).
* -f'(E): ``src/Responses/Shared/ResponseOccupations.jl:18-29`` (T > 0: z/(kT(1+z)^2),
  T = 0: Gaussian of width eta_fs); FS weight ``BerryMultipoleDC.jl:66-78``.
* BCD moments D[mu,a,b,c] += -f'(E_n) Omega^n_ab dE_n/dk_c
  -- ``BerryMultipoleMoments.jl:7-45``, ``BerryMultipoleGeometry.jl:164-216``; lane sums of raw
  moments, / N_k, * 1/|det lattice| -- ``Runtime/Execution/BerryMultipoleDriver.jl:2-29,279-303``;
  DC map sigma[b,c,a] = e^3 tau/(2 hbar^2) (D_abc + D_acb) -- ``BerryMultipoleDC.jl:95-113``.
* AHC -- ``src/MatrixElements/IntrinsicHallWorkspace.jl:6-23`` (v = dH_H + i(E_n - E_m) A_H),
  ``IntrinsicHallGeometry.jl:117-136`` (r = -i v/(E_l - E_m) across the occupation boundary),
  ``src/Responses/IntrinsicHall/IntrinsicHallCurvature.jl:2-29``,
  ``IntrinsicHallOccupations.jl:8-44`` (occupied = #{E_n <= mu}); per-point scale
  -e^2/hbar * 1e-20 (``Runtime/Execution/SpectralResponseDriver.jl:388-400``); lane sums, / N_k,
  / (|det lattice| * 1e-30) (``SpectralResponseDriver.jl:541-552``).

Pinning (tests/test_oracle_geometry.py): Omega^n against the gauge-invariant plaquette
(Fukui-Hatsugai-Suzuki link) Berry phase plus the loop integral of A_H,nn, for models with and
without a position operator; AHC of stacked Qi-Wu-Zhang Chern layers against -(e^2/h) C / c
with C from an independent link-variable Chern number; -f' against the reference's own
closed-form weight values (``test/berry_multipole_dc_unit.jl:39-43``) and a finite-difference
derivative of the Fermi function; BCD vanishing under inversion symmetry; the DC map against
its index symmetries (``berry_multipole_dc_unit.jl:64-79``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import kloop

RESPONSE_CHARGE_C = 1.602176634e-19
RESPONSE_HBAR_JS = 1.054571817e-34
RESPONSE_KB_EV = 8.617333262145e-5


class GapError(ValueError):
    """GEOMETRY_SINGLE_BAND_SPECTRUM_NOT_ISOLATED (``OrthogonalBandDerivatives.jl:74-78``)."""


@dataclass
class GeometryTask:
    """BCD / AHC task parameters (``BerryMultipoleDCParameters``, ``BerryMultipoleDC.jl:6-55``;
    ``TaskConfiguration.jl:568-578`` defaults)."""
    fermi_energies: np.ndarray
    temperature: float = 0.0
    fs_broadening: Optional[float] = None      # eta_fs_ev, required for the T = 0 Fermi surface
    denominator_regularization: float = 0.0
    degeneracy_threshold: float = 0.002
    relaxation_time_fs: float = 1.0

    def __post_init__(self):
        self.fermi_energies = np.atleast_1d(np.asarray(self.fermi_energies, dtype=np.float64))


def geometry_data(model, kpts: np.ndarray, eta: float, wannier: Optional[dict] = None):
    """compute_kpoint! with the Wannier-gauge channels kept (for F^W)."""
    return kloop.compute_kpoint_data(model, kpts, eta, wannier=wannier, keep_wannier=True)


def hamiltonian_curvature(data) -> np.ndarray:
    """F^H[K, a, b, n, m] = U^+ F^W_ab U, F^W_ab = d_a A_b - d_b A_a - i[A_a, A_b]
    (``PositionMatrixElements.jl:491-540``); dA[a, b] holds d_b A_a."""
    w = data.wannier
    a = w["A"]                                   # [K, 3, n, n]
    da = w["dA"]                                 # [K, a, b, n, n] = d_b A_a
    K, _, n, _ = a.shape
    f = np.empty((K, 3, 3, n, n), dtype=np.complex128)
    for x in range(3):
        for y in range(3):
            comm = a[:, x] @ a[:, y] - a[:, y] @ a[:, x]
            f[:, x, y] = da[:, y, x] - da[:, x, y] - 1j * comm
    return kloop.rotate(data.eigenvectors, f)


def band_berry_curvature(data, fh: Optional[np.ndarray] = None) -> np.ndarray:
    """Omega[K, n, a, b] = 2 sum_{m != n} -Im(A^a_nm A^b_mn) + Re F^H_nn,ab
    (``ConventionalQuantumGeometry.jl:167-181``)."""
    if fh is None:
        fh = hamiltonian_curvature(data)
    ber = data.berry_connection                  # [K, 3, n, n]
    K, _, n, _ = ber.shape
    off = ~np.eye(n, dtype=bool)
    out = np.empty((K, n, 3, 3))
    for x in range(3):
        for y in range(3):
            prod = ber[:, x] * np.swapaxes(ber[:, y], -1, -2)      # [K, n, m] = A^x_nm A^y_mn
            s = np.where(off[None], -prod.imag, 0.0).sum(axis=2)
            out[:, :, x, y] = 2.0 * s + np.diagonal(fh[:, x, y], axis1=-2, axis2=-1).real
    return out


def isolated_minimum_gap(energies: np.ndarray, tolerance: float) -> float:
    """``geometry_isolated_spectrum_minimum_gap`` (``OrthogonalBandDerivatives.jl:66-82``) over one k."""
    e = np.asarray(energies, dtype=np.float64)
    if not np.all(np.isfinite(e)):
        raise ValueError("empty/nonfinite geometry spectrum")
    gap = math.inf
    for m in range(e.shape[0]):
        for l in range(m):
            g = abs(e[m] - e[l])
            if not g > tolerance:
                raise GapError(f"GEOMETRY_SINGLE_BAND_SPECTRUM_NOT_ISOLATED: bands=({l + 1},{m + 1}), "
                               f"gap_ev={g}, threshold_ev={tolerance}")
            gap = min(gap, g)
    return gap


def fermi_surface_weight(energy, mu: float, temperature: float, eta_fs: Optional[float]):
    """Minus the Fermi derivative in 1/eV (``ResponseOccupations.jl:18-29``), Gaussian at T = 0."""
    energy = np.asarray(energy, dtype=np.float64)
    if temperature > 0:
        scale = RESPONSE_KB_EV * temperature
        z = np.exp(-np.abs((energy - mu) / scale))
        return z / (scale * (1.0 + z) ** 2)
    if not (eta_fs is not None and eta_fs > 0):
        raise ValueError("T=0 requires positive eta_fs_ev")
    x = (energy - mu) / eta_fs
    return np.exp(-x * x) / (math.sqrt(math.pi) * eta_fs)


def bcd_point_moments(model, task: GeometryTask, kpts: np.ndarray, data=None) -> np.ndarray:
    """Per-k raw moments D[K, n_mu, a, b, c] (``BerryMultipoleMoments.jl:7-45``, FS form, order 0)."""
    if data is None:
        data = geometry_data(model, kpts, task.denominator_regularization)
    for e in data.energies:
        isolated_minimum_gap(e, task.degeneracy_threshold)
    omega = band_berry_curvature(data)                                           # [K, n, a, b]
    grad = np.diagonal(data.hamiltonian_derivatives, axis1=-2, axis2=-1).real    # [K, 3, n]
    K = omega.shape[0]
    out = np.zeros((K, task.fermi_energies.shape[0], 3, 3, 3))
    for i, mu in enumerate(task.fermi_energies):
        w = fermi_surface_weight(data.energies, mu, task.temperature, task.fs_broadening)   # [K, n]
        out[:, i] = np.einsum("kn,knab,kcn->kabc", w, omega, grad)
    return out


def ahc_point_values(model, task: GeometryTask, kpts: np.ndarray, data=None) -> np.ndarray:
    """Per-k AHC values [K, n_mu, a, b] = (-e^2/hbar 1e-20) Omega_ab(mu) (Matrix method, T = 0)."""
    if task.temperature != 0:
        raise ValueError("INTRINSIC_HALL_REQUIRES_ZERO_TEMPERATURE")
    if data is None:
        data = geometry_data(model, kpts, task.denominator_regularization)
    fh = hamiltonian_curvature(data)
    e = data.energies
    de = e[:, :, None] - e[:, None, :]
    vel = data.hamiltonian_derivatives + 1j * de[:, None] * data.internal_connection       # [K, 3, n, n]
    K, n = e.shape
    scale = -RESPONSE_CHARGE_C ** 2 / RESPONSE_HBAR_JS * 1e-20
    out = np.zeros((K, task.fermi_energies.shape[0], 3, 3))
    for k in range(K):
        for i, mu in enumerate(task.fermi_energies):
            occ = int(np.searchsorted(e[k], mu, side="right"))
            curv = np.zeros((3, 3))
            if occ > 0:
                lo, hi = slice(0, occ), slice(occ, n)
                r_lm = -1j * (vel[k][:, lo, hi] / de[k][lo, hi][None])               # [3, occ, n-occ]
                r_ml = -1j * (vel[k][:, hi, lo] / de[k][hi, lo][None])               # [3, n-occ, occ]
                for x in range(3):
                    for y in range(3):
                        left = r_lm[x] * r_ml[y].T
                        right = r_lm[y] * r_ml[x].T
                        curv[x, y] = (np.diagonal(fh[k, x, y]).real[:occ].sum()
                                      - (left - right).imag.sum())
            out[k, i] = scale * curv
    return out


def integrate_geometry_mesh(model, task: GeometryTask, mesh: Sequence[int], quantity: str,
                            batch: int = 2048, k_limit: Optional[int] = None):
    """Dense-mesh BCD or AHC k-loop.  Returns (lanes, total): lanes[l] = sum of raw per-point
    values over lane l's k range (what consume_point! leaves in the lane buffer), total = the
    lanes summed in lane order, / N_k, times the reference's measure (BCD: 1/|det L|, giving
    the moment D_abc; AHC: 1/(|det L| 1e-30), giving sigma_ab in S/m)."""
    nk = int(np.prod(mesh))
    nl = kloop.lane_count(nk)
    stop = nk if k_limit is None else min(nk, int(k_limit))
    nmu = task.fermi_energies.shape[0]
    shape = (nmu, 3, 3, 3) if quantity == "bcd" else (nmu, 3, 3)
    lanes = np.zeros((nl,) + shape)
    starts = np.array([kloop.lane_range(nk, nl, l)[0] for l in range(nl)] + [nk])
    chans = kloop.real_space_channels(model)
    for k0 in range(0, stop, batch):
        k1 = min(stop, k0 + batch)
        kp = kloop.integral_kpoints(mesh, k0, k1)
        data = geometry_data(model, kp, task.denominator_regularization, kloop.interpolate(model, kp, chans))
        vals = (bcd_point_moments if quantity == "bcd" else ahc_point_values)(model, task, kp, data)
        lane_of = np.searchsorted(starts, np.arange(k0, k1), side="right") - 1
        for lane in np.unique(lane_of):
            lanes[lane] += vals[lane_of == lane].sum(axis=0)
    vol = abs(np.linalg.det(model.lattice))
    total = lanes.sum(axis=0) / nk
    total = total / vol if quantity == "bcd" else total / (vol * 1e-30)
    return lanes, total


def bcd_dc_conductivity(moment: np.ndarray, relaxation_time_fs: float, spatial_dimension: int = 3) -> np.ndarray:
    """sigma[b, c, a] = e^3 tau/(2 hbar^2) (D_abc + D_acb) (1e-10)^(3-d) (``BerryMultipoleDC.jl:95-113``)."""
    tau = relaxation_time_fs * 1e-15
    pref = RESPONSE_CHARGE_C ** 3 * tau / (2 * RESPONSE_HBAR_JS ** 2) * (1e-10) ** (3 - spatial_dimension)
    out = np.zeros((3, 3, 3))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                out[b, c, a] = pref * (moment[a, b, c] + moment[a, c, b])
    return out
