"""NumPy FP64 restatement of the reference's dense-mesh k-loop (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this module.  It is the checker, never the thing shipped.

Restated reference functions (all paths relative to the reference root):

* Fourier factors / interpolation  -- ``MatrixElements/Interpolated/HamiltonianMatrixElements.jl:28-58,92-177,427-462``,
  ``PositionMatrixElements.jl:8-84,239-409`` (Convention II: displacement = Cartesian R).
* Eigensystem                      -- ``Core/ESKernel/DenseSpectra.jl:14-16`` (LAPACK, upper triangle),
  ``HamiltonianMatrixElements.jl:6-25`` (canonical column phase), ``:223-236``.
* Hamiltonian-gauge rotation        -- ``HamiltonianMatrixElements.jl:61-71``.
* Energy differences, D, Berry connection -- ``HamiltonianMatrixElements.jl:385-390``,
  ``PositionMatrixElements.jl:444-489``, ``PositionVertexRegularization.jl:17-29``.
* Screening                         -- ``Core/CalculationUtils.jl:21-45,71-89``,
  ``Runtime/Execution/TransitionScreening.jl:32-94``.
* Generalized derivative / SC kernel -- ``Responses/ShiftCurrent/ConventionalShiftCurrent.jl:40-84,191-229,527-767``.
* Injection kernel                  -- ``Responses/InjectionCurrent/ConventionalInjectionCurrent.jl:23-97``.
* Frequency contraction             -- ``ConventionalShiftCurrent.jl:774-884``, ``ConventionalInjectionCurrent.jl:104-…``.
* Prefactors                        -- ``Runtime/Execution/ConventionalCurrentIntegralExecution.jl:243-245,296-298``,
  constants ``Core/Constants.jl:1-4``; per-point N_k = 1 (``IntegralDriver.jl:644``), then 1/N_k in
  ``consume_point!`` (``IntegralDriver.jl:820-874``).
* Mesh and lanes                    -- ``Core/CalculationUtils.jl:280-292``, ``Runtime/Setup/BundleAssembly.jl:35-47``,
  ``Runtime/Reduction/DeterministicReduction.jl:2-37``.

Pinning (tests/test_oracle_*.py):

* interpolation, eigensystem, rotation, A, dA, ddH, D chain + SHG: the 24 SHG golden files
  (tests/golden/second_harmonic, from the reference's test fixtures), atol 1e-10 rtol 1e-8;
* interpolation: brute-force per-element sums, H(k+G) = H(k), Hermiticity, central finite
  differences of H(k) for dH and ddH and of A(k) for dA;
* eigensystem: residual/orthonormality, canonical phase rule;
* generalized derivative: scalar element == commutator (GEMM) form (reference's own
  ``generalized_derivative_gemm_unit.jl`` identity) and, for isolated bands at eta -> 0,
  a finite-difference derivative of the phase-fixed r_nm;
* shift current: vanishes for inversion-symmetric models, invariant under random per-band
  phases; injection: vanishes when v_mm = v_nn;
* lanes: integer partition identities.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

ELEMENTARY_CHARGE_C = 1.602176634e-19
REDUCED_PLANCK_CONSTANT_J_S = 1.054571817e-34
REDUCED_PLANCK_CONSTANT_EV_S = 6.582119569e-16
BOLTZMANN_CONSTANT_EV_K = 8.617333262e-5

LANE_LIMIT = 64


@dataclass
class Task:
    """Subset of ``EffectiveTaskConfig`` (reference ``Runtime/Config.jl:4-92``) used on this path."""
    photon_energies: np.ndarray
    broadening: float = 0.040
    broadening_type: str = "Gaussian"
    fermi_energy: float = 0.0
    temperature: float = 0.0
    denominator_regularization: float = 0.001
    transition_window_factor: float = 5.0
    band_window_size: int = -1
    shift_current: bool = True
    injection_current: bool = False

    def __post_init__(self):
        self.photon_energies = np.ascontiguousarray(self.photon_energies, dtype=np.float64)


# ----------------------------------------------------------------------------------------
# Mesh and lanes
# ----------------------------------------------------------------------------------------

def integral_kpoints(mesh: Sequence[int], start: int = 0, stop: Optional[int] = None) -> np.ndarray:
    """Gamma-centred mesh k_i = i/N_i with axis 1 fastest (``CalculationUtils.jl:280-292``)."""
    n1, n2, n3 = (int(x) for x in mesh)
    total = n1 * n2 * n3
    stop = total if stop is None else stop
    idx = np.arange(start, stop, dtype=np.int64)
    i1 = idx % n1
    i2 = (idx // n1) % n2
    i3 = idx // (n1 * n2)
    return np.stack([i1 / n1, i2 / n2, i3 / n3], axis=1)


def lane_count(num_kpoints: int) -> int:
    """``BundleAssembly.jl:37-38``: min(N_k, 64)."""
    return min(num_kpoints, LANE_LIMIT)


def lane_range(num_kpoints: int, num_lanes: int, lane: int) -> tuple[int, int]:
    """0-based half-open k range of 0-based ``lane`` (``BundleAssembly.jl:41-47``)."""
    first = ((lane) * num_kpoints) // num_lanes
    last = ((lane + 1) * num_kpoints) // num_lanes
    return first, last


# ----------------------------------------------------------------------------------------
# Interpolation (Convention II)
# ----------------------------------------------------------------------------------------

def r_cartesian(model) -> np.ndarray:
    """``_r_vector_cartesian`` (``HamiltonianMatrixElements.jl:74-85``): sum_i R_i * lattice[i, :]."""
    return model.r_vectors.astype(np.float64) @ model.lattice


def fourier_factors(model, kpts: np.ndarray) -> np.ndarray:
    """cis(2 pi R.k)/deg(R) (``HamiltonianMatrixElements.jl:28-39``); returns [K, N_R]."""
    phases = np.asarray(kpts, dtype=np.float64) @ model.r_vectors.astype(np.float64).T
    return np.exp(2j * np.pi * phases) / model.r_degeneracies[None, :].astype(np.float64)


def real_space_channels(model) -> dict:
    """R-space channels (``HamiltonianMatrixElements.jl:92-177``, ``PositionMatrixElements.jl:8-84``).

    dH_a(R) = i R_a H(R); ddH_ab(R) = -R_a R_b H(R); dA_ab(R) = i R_b A_a(R)
    (a = position direction, b = derivative direction).
    """
    rc = r_cartesian(model)
    hr = model.hamiltonian_r
    ar = model.position_or_zero()
    dh = 1j * rc.T[:, :, None, None] * hr[None]                          # [3, N_R, n, n]
    ddh = -(rc.T[:, None, :, None, None] * rc.T[None, :, :, None, None]) * hr[None, None]  # [3,3,N_R,n,n]
    a = np.transpose(ar, (1, 0, 2, 3))                                      # [3, N_R, n, n]
    da = 1j * rc.T[None, :, :, None, None] * a[:, None]                    # [a, b, N_R, n, n]
    return dict(H=hr, dH=dh, ddH=ddh, A=a, dA=da)


def interpolate(model, kpts: np.ndarray, channels: Optional[dict] = None) -> dict:
    """Wannier-gauge X(k) = sum_R f_R(k) X(R) for all channels (``HamiltonianMatrixElements.jl:42-58``)."""
    ch = real_space_channels(model) if channels is None else channels
    f = fourier_factors(model, kpts)
    n = model.num_orbitals
    nr = model.num_r_vectors

    def contract(x):
        lead = x.shape[:-3]
        flat = np.moveaxis(x.reshape((-1, nr, n, n)), 1, 0).reshape(nr, -1)
        out = f @ flat
        return out.reshape((f.shape[0],) + lead + (n, n))

    return {key: contract(val) for key, val in ch.items()}


def interpolate_factored(model, kpts: np.ndarray) -> dict:
    """The channels of :func:`interpolate` with the R-dependent factors moved onto the Fourier factors:
    dH_a(k) = sum_R (i R_a f_R) H(R), ddH_ab(k) = sum_R (-R_a R_b f_R) H(R), dA_ab(k) = sum_R (i R_b f_R) A_a(R).
    The same sums, without materialising the R-space channels (large models: n = 96, N_R = 2197)."""
    f = fourier_factors(model, kpts)                   # [K, N_R]
    rc = r_cartesian(model)                            # [N_R, 3]
    n, nr = model.num_orbitals, model.num_r_vectors
    hr = model.hamiltonian_r.reshape(nr, n * n)
    K = f.shape[0]
    out = {"H": (f @ hr).reshape(K, n, n)}
    out["dH"] = np.stack([((1j * rc[None, :, a] * f) @ hr).reshape(K, n, n) for a in range(3)], axis=1)
    out["ddH"] = np.stack([np.stack([((-rc[None, :, a] * rc[None, :, b] * f) @ hr).reshape(K, n, n)
                                     for b in range(3)], axis=1) for a in range(3)], axis=1)
    ar = model.position_or_zero()                      # [N_R, 3, n, n]
    out["A"] = np.stack([(f @ ar[:, a].reshape(nr, n * n)).reshape(K, n, n) for a in range(3)], axis=1)
    out["dA"] = np.stack([np.stack([((1j * rc[None, :, b] * f) @ ar[:, a].reshape(nr, n * n)).reshape(K, n, n)
                                    for b in range(3)], axis=1) for a in range(3)], axis=1)
    return out


# ----------------------------------------------------------------------------------------
# Eigensystem
# ----------------------------------------------------------------------------------------

def canonicalize_eigenvectors(vectors: np.ndarray) -> np.ndarray:
    """Largest-|.| entry of each column real >= 0, first on ties (``HamiltonianMatrixElements.jl:6-25``)."""
    v = np.array(vectors, dtype=np.complex128, copy=True)
    mag = np.abs(v)
    pivot = np.argmax(mag, axis=-2)                      # first occurrence of the maximum
    pv = np.take_along_axis(v, pivot[..., None, :], axis=-2)[..., 0, :]
    pm = np.abs(pv)
    phase = np.where(pm > 0, np.conj(pv) / np.where(pm > 0, pm, 1.0), 1.0)
    return v * phase[..., None, :]


def hermitian_spectrum(hk: np.ndarray):
    """eigen(Hermitian(H, :U)) -- LAPACK on the upper triangle (``DenseSpectra.jl:14-16``)."""
    e, v = np.linalg.eigh(hk, UPLO="U")
    return e, canonicalize_eigenvectors(v)


# ----------------------------------------------------------------------------------------
# Hamiltonian-gauge k-point data
# ----------------------------------------------------------------------------------------

def regularized_energy_inverse(d: np.ndarray, eta: float) -> np.ndarray:
    """d/(d^2+eta^2) with the scaled branch for over/underflow (``PositionVertexRegularization.jl:17-29``)."""
    d = np.asarray(d, dtype=np.float64)
    den = d * d + eta * eta
    ok = np.isfinite(den) & (den != 0)
    out = np.empty_like(d)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        out[ok] = d[ok] / den[ok]
        bad = ~ok
        if np.any(bad):
            s = np.maximum(np.abs(d[bad]), eta)
            sd, sw = d[bad] / s, eta / s
            out[bad] = (sd / (sd * sd + sw * sw)) / s
    return out


def bare_energy_inverse(d: np.ndarray) -> np.ndarray:
    """1/d, zero strictly below 1e-7 eV (``PositionVertexRegularization.jl:1-8``)."""
    d = np.asarray(d, dtype=np.float64)
    with np.errstate(divide="ignore"):
        return np.where(np.abs(d) < 1.0e-7, 0.0, 1.0 / np.where(d == 0, 1.0, d))


def regularized_energy_weight(d: np.ndarray, eta: float) -> np.ndarray:
    """eta^2/(d^2+eta^2) (``PositionVertexRegularization.jl:31-48``)."""
    d = np.asarray(d, dtype=np.float64)
    return (eta * eta) / (d * d + eta * eta)


@dataclass
class KPointData:
    """Batched ``KPointMatrixData`` (reference ``MatrixElementWorkspace.jl:161-178``); leading axis = k."""
    energies: np.ndarray                 # [K, n]
    eigenvectors: np.ndarray             # [K, n, n] columns
    hamiltonian_derivatives: np.ndarray  # V  [K, 3, n, n]
    hamiltonian_second_derivatives: np.ndarray  # W [K, 3, 3, n, n]
    internal_connection: np.ndarray      # A_H [K, 3, n, n]
    internal_connection_derivatives: np.ndarray  # [K, a, b, n, n] = U^+ (i R_b A_a) U
    inverse_energy_differences: np.ndarray  # [K, n, n]
    gauge_correction: np.ndarray         # D [K, 3, n, n]
    berry_connection: np.ndarray         # A_H + i D  [K, 3, n, n]
    wannier: dict = field(default_factory=dict)


def rotate(u: np.ndarray, x: np.ndarray) -> np.ndarray:
    """U^+ X U over trailing (n, n) axes (``HamiltonianMatrixElements.jl:61-71``)."""
    ud = np.conj(np.swapaxes(u, -1, -2))
    extra = x.ndim - u.ndim
    ue = u.reshape(u.shape[:1] + (1,) * extra + u.shape[1:])
    ude = ud.reshape(ud.shape[:1] + (1,) * extra + ud.shape[1:])
    return ude @ x @ ue


def compute_kpoint_data(model, kpts: np.ndarray, eta: float = 1.0e-3, wannier: Optional[dict] = None,
                        keep_wannier: bool = False) -> KPointData:
    """``compute_kpoint!`` (``HamiltonianMatrixElements.jl:568``) for a batch of k points."""
    w = interpolate(model, kpts) if wannier is None else wannier
    e, u = hermitian_spectrum(w["H"])
    v = rotate(u, w["dH"])
    ww = rotate(u, w["ddH"])
    ah = rotate(u, w["A"])
    dah = rotate(u, w["dA"])
    de = e[:, :, None] - e[:, None, :]                                   # E_n - E_m
    if eta == 0.0:
        inv = np.where(de == 0, 0.0, 1.0 / np.where(de == 0, 1.0, de))
    else:
        inv = regularized_energy_inverse(de, eta)
    d = -v * inv[:, None]
    berry = ah + 1j * d
    return KPointData(e, u, v, ww, ah, dah, inv, d, berry, w if keep_wannier else {})


# ----------------------------------------------------------------------------------------
# Screening
# ----------------------------------------------------------------------------------------

def fermi_dirac(energies: np.ndarray, mu: float, temperature: float) -> np.ndarray:
    """``CalculationUtils.jl:21-45``: T=0 -> E <= mu gives 1; else clamped logistic."""
    kt = BOLTZMANN_CONSTANT_EV_K * temperature
    if kt == 0:
        return (energies <= mu).astype(np.float64)
    arg = np.clip((energies - mu) / kt, -700.0, 700.0)
    return 1.0 / (np.exp(arg) + 1.0)


def transition_screen(energies: np.ndarray, task: Task):
    """``_fill_transition_screen!`` (``TransitionScreening.jl:32-94``) for a batch.

    Returns (occupation_differences [K,n,n], delta [K,n_w,n,n], active_pair [K,n,n], active_k [K]).
    Band window -1 (all bands) only.
    """
    if task.band_window_size != -1 and 2 * task.band_window_size <= energies.shape[1]:
        raise NotImplementedError("band windows other than -1 are not restated")
    f = fermi_dirac(energies, task.fermi_energy, task.temperature)
    df = f[:, :, None] - f[:, None, :]                                  # f_n - f_m
    trans = energies[:, None, :] - energies[:, :, None]                 # E_m - E_n at [n, m]
    w = task.broadening
    x = trans[:, None, :, :] - task.photon_energies[None, :, None, None]
    occ = np.abs(df) > 1.0e-10
    if task.broadening_type == "Gaussian":
        cutoff = w * math.sqrt(200.0)
        inside = np.abs(x) < cutoff
        delta = np.where(inside, (1.0 / (math.sqrt(math.pi) * w)) * np.exp(-(x / w) ** 2), 0.0)
        pair = occ & np.any(inside, axis=1)
        active_k = np.any(pair, axis=(1, 2))
    elif task.broadening_type == "Lorentzian":
        delta = (w / math.pi) / (x * x + w * w)
        pair = occ.copy()
        thr = task.transition_window_factor * w
        if thr <= 0:
            active_k = np.any(occ, axis=(1, 2))
        else:
            active_k = np.any(occ[:, None] & (np.abs(x) < thr), axis=(1, 2, 3))
    else:
        raise ValueError(f"Unknown smearing type: {task.broadening_type}")
    return df, delta, pair, active_k


# ----------------------------------------------------------------------------------------
# Shift current and injection current
# ----------------------------------------------------------------------------------------

def generalized_derivative_element(data: KPointData, k: int, n: int, m: int, a: int, b: int) -> complex:
    """Scalar ``conventional_generalized_position_derivative_element`` (``ConventionalShiftCurrent.jl:191-229``)."""
    A = data.internal_connection[k]
    D = data.gauge_correction[k]
    V = data.hamiltonian_derivatives[k]
    nb = A.shape[-1]
    sum_ad = 0j
    sum_hd = 0j
    for l in range(nb):
        if l != n and l != m:
            sum_ad += A[a, n, l] * D[b, l, m] - D[b, n, l] * A[a, l, m]
            sum_hd += V[a, n, l] * D[b, l, m] - D[b, n, l] * V[a, l, m]
    ad_bit = (A[b, n, n] - A[b, m, m]) * D[a, n, m] + (A[a, n, n] - A[a, m, m]) * D[b, n, m]
    aa_bit = (A[b, n, n] - A[b, m, m]) * A[a, n, m]
    dv_bit = (V[b, n, n] - V[b, m, m]) * D[a, n, m] + (V[a, n, n] - V[a, m, m]) * D[b, n, m]
    return (data.internal_connection_derivatives[k, a, b, n, m] + ad_bit - 1j * aa_bit + sum_ad
            + 1j * (data.hamiltonian_second_derivatives[k, a, b, n, m] + sum_hd + dv_bit)
            * (-data.inverse_energy_differences[k, n, m]))


def generalized_derivative(data: KPointData) -> np.ndarray:
    """All r^b_{nm;a} via commutators (``ConventionalShiftCurrent.jl:40-84,86-164``).

    Returns G[K, a, b, n, m] with the reference's (a, b) argument order.
    """
    A = data.internal_connection          # [K,3,n,n]
    D = data.gauge_correction
    V = data.hamiltonian_derivatives
    nb = A.shape[-1]
    eye = np.eye(nb, dtype=bool)
    Ad = np.diagonal(A, axis1=-2, axis2=-1)   # [K,3,n]
    Dd = np.diagonal(D, axis1=-2, axis2=-1)
    Vd = np.diagonal(V, axis1=-2, axis2=-1)
    K = A.shape[0]
    out = np.empty((K, 3, 3, nb, nb), dtype=np.complex128)
    notdiag = (~eye).astype(np.float64)
    for a in range(3):
        for b in range(3):
            comm_a = A[:, a] @ D[:, b] - D[:, b] @ A[:, a]
            comm_h = V[:, a] @ D[:, b] - D[:, b] @ V[:, a]
            end_ad = Ad[:, a, :, None] * D[:, b] - Dd[:, b, :, None] * A[:, a]
            end_hd = Vd[:, a, :, None] * D[:, b] - Dd[:, b, :, None] * V[:, a]
            end_ad = end_ad + notdiag * (A[:, a] * Dd[:, b, None, :] - D[:, b] * Ad[:, a, None, :])
            end_hd = end_hd + notdiag * (V[:, a] * Dd[:, b, None, :] - D[:, b] * Vd[:, a, None, :])
            sum_ad = comm_a - end_ad
            sum_hd = comm_h - end_hd
            dA_b = Ad[:, b, :, None] - Ad[:, b, None, :]
            dA_a = Ad[:, a, :, None] - Ad[:, a, None, :]
            dV_b = Vd[:, b, :, None] - Vd[:, b, None, :]
            dV_a = Vd[:, a, :, None] - Vd[:, a, None, :]
            ad_bit = dA_b * D[:, a] + dA_a * D[:, b]
            aa_bit = dA_b * A[:, a]
            dv_bit = dV_b * D[:, a] + dV_a * D[:, b]
            out[:, a, b] = (data.internal_connection_derivatives[:, a, b] + ad_bit - 1j * aa_bit + sum_ad
                            + 1j * (data.hamiltonian_second_derivatives[:, a, b] + sum_hd + dv_bit)
                            * (-data.inverse_energy_differences))
    return out


def shift_current_kernel(data: KPointData, gderiv: Optional[np.ndarray] = None) -> np.ndarray:
    """Q[K,n,m,a,b,c] = -i (G[n,m,c,a] Ab[m,n,b] - G[m,n,b,a] Ab[n,m,c]) (``ConventionalShiftCurrent.jl:710-729``)."""
    G = generalized_derivative(data) if gderiv is None else gderiv   # [K, x, y, n, m]
    Ab = data.berry_connection                                       # [K, 3, n, m]
    # term1[n,m,a,b,c] = G[c,a][n,m] * Ab[b][m,n]
    Gnm = np.transpose(G, (0, 3, 4, 2, 1))        # [K, n, m, a(y), c(x)]
    AbT = np.transpose(Ab, (0, 3, 2, 1))          # [K, n, m, b] = Ab[b][m, n]
    AbN = np.transpose(Ab, (0, 2, 3, 1))          # [K, n, m, c] = Ab[c][n, m]
    GmnT = np.transpose(G, (0, 4, 3, 2, 1))       # [K, n, m, a(y), b(x)] = G[b,a][m,n]
    t1 = Gnm[:, :, :, :, None, :] * AbT[:, :, :, None, :, None]   # [K,n,m,a,b,c]
    t2 = GmnT[:, :, :, :, :, None] * AbN[:, :, :, None, None, :]
    return -1j * (t1 - t2)


def injection_current_kernel(data: KPointData) -> np.ndarray:
    """eta[K,n,m,a,b,c] = (V_mm^a - V_nn^a) Ab[n,m,c] Ab[m,n,b] (``ConventionalInjectionCurrent.jl:23-97``)."""
    Vd = np.diagonal(data.hamiltonian_derivatives, axis1=-2, axis2=-1)   # [K, 3, n]
    vdiff = Vd[:, :, None, :] - Vd[:, :, :, None]                        # [K, a, n, m] = V_mm - V_nn
    Ab = data.berry_connection
    AbN = np.transpose(Ab, (0, 2, 3, 1))          # [K,n,m,c]
    AbT = np.transpose(Ab, (0, 3, 2, 1))          # [K,n,m,b] = Ab[b][m,n]
    vd = np.transpose(vdiff, (0, 2, 3, 1))        # [K,n,m,a]
    return vd[:, :, :, :, None, None] * AbT[:, :, :, None, :, None] * AbN[:, :, :, None, None, :]


def shift_current_prefactor(cell_volume: float) -> float:
    """``ConventionalCurrentIntegralExecution.jl:243-245`` with kpoint_weight = num_kpoints = 1."""
    return (REDUCED_PLANCK_CONSTANT_EV_S * math.pi * (-ELEMENTARY_CHARGE_C) ** 3
            / (2.0 * REDUCED_PLANCK_CONSTANT_J_S ** 2 * 1 * cell_volume))


def injection_current_prefactor(cell_volume: float) -> float:
    """``ConventionalCurrentIntegralExecution.jl:296-298`` with kpoint_weight = num_kpoints = 1."""
    return -math.pi * ELEMENTARY_CHARGE_C ** 3 / (REDUCED_PLANCK_CONSTANT_J_S ** 2 * 1 * cell_volume)


def point_values(model, task: Task, kpts: np.ndarray, data: Optional[KPointData] = None) -> dict:
    """Per-k ``point_value!`` for SC/IC (``IntegralDriver.jl:596``) with N_k = 1.

    Returns {'sc': [K, n_w, 3, 3, 3], 'ic': ...} complex, [a, b, c] last.
    """
    if data is None:
        data = compute_kpoint_data(model, kpts, task.denominator_regularization)
    df, delta, pair, active_k = transition_screen(data.energies, task)
    weight = np.where(pair[:, None], df[:, None] * delta, 0.0) * active_k[:, None, None, None]
    K, nw = weight.shape[:2]
    nb = data.energies.shape[1]
    out = {}
    vol = model.cell_volume
    wflat = weight.reshape(K, nw, nb * nb)
    if task.shift_current:
        q = shift_current_kernel(data).reshape(K, nb * nb, 27)
        q = np.where(pair.reshape(K, nb * nb, 1), q, 0.0)
        out["sc"] = (shift_current_prefactor(vol) * (wflat @ q)).reshape(K, nw, 3, 3, 3)
    if task.injection_current:
        q = injection_current_kernel(data).reshape(K, nb * nb, 27)
        q = np.where(pair.reshape(K, nb * nb, 1), q, 0.0)
        out["ic"] = (injection_current_prefactor(vol) * (wflat @ q)).reshape(K, nw, 3, 3, 3)
    return out


def integrate_mesh(model, task: Task, mesh: Sequence[int], batch: int = 4096, k_limit: Optional[int] = None):
    """Integral k-loop with deterministic lanes (``IntegralDriver.jl:92-1066``, ``BundleAssembly.jl:35-47``).

    Returns (lanes, total) where lanes[key] is [n_lanes, n_w, 3, 3, 3] of sum_k (1/N_k) value_k
    over each lane's k range, and total[key] sums the lanes in lane order
    (``DeterministicReduction.jl:27``).  ``k_limit`` evaluates only the first k_limit points
    (for bounded CPU timing); lanes then still use the full-mesh partition and 1/N_k.
    """
    nk = int(np.prod(mesh))
    nl = lane_count(nk)
    stop = nk if k_limit is None else min(nk, int(k_limit))
    keys = [k for k, on in (("sc", task.shift_current), ("ic", task.injection_current)) if on]
    nw = task.photon_energies.shape[0]
    lanes = {key: np.zeros((nl, nw, 3, 3, 3), dtype=np.complex128) for key in keys}
    starts = np.array([lane_range(nk, nl, l)[0] for l in range(nl)] + [nk])
    scale = 1.0 / nk
    chans = real_space_channels(model)
    for k0 in range(0, stop, batch):
        k1 = min(stop, k0 + batch)
        kp = integral_kpoints(mesh, k0, k1)
        data = compute_kpoint_data(model, kp, task.denominator_regularization,
                                   wannier=interpolate(model, kp, chans))
        vals = point_values(model, task, kp, data)
        lane_of = np.searchsorted(starts, np.arange(k0, k1), side="right") - 1
        for key in keys:
            for lane in np.unique(lane_of):
                sel = lane_of == lane
                lanes[key][lane] += scale * vals[key][sel].sum(axis=0)
    total = {key: lanes[key].sum(axis=0) for key in keys}
    return lanes, total


def spectrum_relative_error(x: np.ndarray, ref: np.ndarray, floor: float = 1.0e-12) -> np.ndarray:
    """Per-component ||x - ref||_2 / ||ref||_2 over the omega axis (axis 0); absolute when
    ||ref|| < floor * max component norm."""
    x = x.reshape(x.shape[0], -1)
    ref = ref.reshape(ref.shape[0], -1)
    num = np.linalg.norm(x - ref, axis=0)
    den = np.linalg.norm(ref, axis=0)
    big = den >= floor * max(den.max(), 1e-300)
    return np.where(big, num / np.where(big, den, 1.0), num)
