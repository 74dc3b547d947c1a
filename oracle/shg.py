"""NumPy restatement of the reference's SHG kernel (TEST INFRASTRUCTURE ONLY).

Restates, for one k point:

* ``cutoff_optical_position_vertices`` / ``optical_position_vertices!``
  (reference ``MatrixElements/Interpolated/OpticalPositionDerivatives.jl:13-130``);
* ``second_harmonic_response!`` and ``second_harmonic_one_band!``
  (reference ``Responses/SecondHarmonicGeneration/SecondHarmonicKernels.jl:71-294``).

Pinned by the 24 golden files of the reference (tests/golden/second_harmonic), which in turn
pin the whole interpolation -> eigensystem -> rotation -> vertex chain of ``oracle.kloop``.
"""

from __future__ import annotations

import math

import numpy as np

from .kloop import bare_energy_inverse, regularized_energy_inverse, regularized_energy_weight

SECOND_HARMONIC_COMPONENTS = [(a, b, c) for a in range(3) for (b, c) in ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))]


def optical_position_vertices(E, A, V, dA, W, eta, eta_correction, active_pairs=None):
    """Return (dipole r0, regular_dipole r_eta, derivative Gcut, bare D0).

    Inputs for one k: E[n], A[3,n,n] (internal connection), V[3,n,n] (dH in H gauge),
    dA[a,b,n,n] (internal connection derivatives), W[a,b,n,n] (second derivatives).
    Arrays returned as [3,n,n] and derivative as [a,b,n,n].
    """
    n = E.shape[0]
    diff = E[None, :] - E[:, None]                      # [n, m] = E_m - E_n
    bare = V * bare_energy_inverse(diff)[None]          # D0_nm = h_nm/(E_m - E_n) with cutoff
    regular = V * regularized_energy_inverse(diff, eta)[None]
    dipole = A + 1j * bare
    regular_dipole = A + 1j * regular
    deriv = np.zeros((3, 3, n, n), dtype=np.complex128)
    Ad = np.diagonal(A, axis1=1, axis2=2)               # [3, n]
    Vd = np.diagonal(V, axis1=1, axis2=2)
    inv = bare_energy_inverse(diff)
    for a in range(3):
        for b in range(3):
            comm_a = A[a] @ regular[b] - regular[b] @ A[a]
            comm_h = V[a] @ regular[b] - regular[b] @ V[a]
            for m in range(n):
                for nn in range(n):
                    if nn == m:
                        continue
                    if active_pairs is not None and not active_pairs[nn, m]:
                        continue
                    sum_a = comm_a[nn, m] - (Ad[a, nn] - Ad[a, m]) * regular[b, nn, m]
                    sum_h = comm_h[nn, m] - (Vd[a, nn] - Vd[a, m]) * regular[b, nn, m]
                    finite = (dA[a, b, nn, m] + sum_a
                              + (Ad[a, nn] - Ad[a, m]) * bare[b, nn, m]
                              + (Ad[b, nn] - Ad[b, m]) * bare[a, nn, m]
                              - 1j * A[a, nn, m] * (Ad[b, nn] - Ad[b, m]))
                    numerator = 1j * (W[a, b, nn, m] + sum_h
                                      + bare[a, nn, m] * (Vd[b, nn] - Vd[b, m]).real
                                      + bare[b, nn, m] * (Vd[a, nn] - Vd[a, m]).real)
                    if eta_correction:
                        for p in range(n):
                            if p == nn or p == m:
                                continue
                            numerator += regularized_energy_weight(E[p] - E[m], eta) * (
                                A[a, nn, p] * V[b, p, m]
                                - (V[a, nn, p] + 1j * (E[nn] - E[p]) * A[a, nn, p]) * A[b, p, m])
                            numerator -= regularized_energy_weight(E[nn] - E[p], eta) * (
                                V[b, nn, p] * A[a, p, m]
                                - A[b, nn, p] * (V[a, p, m] + 1j * (E[p] - E[m]) * A[a, p, m]))
                    iv = inv[nn, m]
                    deriv[a, b, nn, m] = finite if iv == 0 else finite + iv * numerator
    return dipole, regular_dipole, deriv, bare


def _delta(x, width, gaussian):
    y = x / width
    return np.exp(-y * y) / (math.sqrt(math.pi) * width) if gaussian else 1.0 / (math.pi * width * (1 + y * y))


def _principal(x, width):
    return x / (x * x + width * width)


def _resonance(x, width, gaussian):
    return _principal(x, width) + 1j * math.pi * _delta(x, width, gaussian)


def _occupation(e, mu, temperature):
    if temperature <= 1.0e-7:
        return 1.0 if e <= mu else 0.0
    x = (e - mu) / (8.617333262145e-5 * temperature)
    return math.exp(-x) / (1 + math.exp(-x)) if x >= 0 else 1.0 / (1 + math.exp(x))


def second_harmonic_response(E, A, V, dA, W, frequencies, components=None, *, fermi_energy, temperature,
                             broadening, low_frequency_broadening, gaussian, intermediate_regularization,
                             eta_correction, degeneracy_threshold):
    """``second_harmonic_response!`` for one k point; returns output[freq, comp, term(7), quantity(2)]."""
    comps = SECOND_HARMONIC_COMPONENTS if components is None else components
    freqs = np.asarray(frequencies, dtype=np.float64)
    n = E.shape[0]
    eta = intermediate_regularization
    out = np.zeros((freqs.shape[0], len(comps), 7, 2), dtype=np.complex128)
    occ = np.array([_occupation(e, fermi_energy, temperature) for e in E])
    active = np.array([[nn != m and abs(occ[nn] - occ[m]) >= 1.0e-10 for m in range(n)] for nn in range(n)])
    dip, rdip, der, bare = optical_position_vertices(E, A, V, dA, W, eta, eta_correction, active)
    Vd = np.diagonal(V, axis1=1, axis2=2)
    low = _principal(freqs, low_frequency_broadening)
    for nn in range(n):
        for m in range(n):
            if nn == m:
                continue
            o = occ[nn] - occ[m]
            if abs(o) < 1.0e-10:
                continue
            diff = E[m] - E[nn]
            inverse = float(bare_energy_inverse(np.array(diff)))
            d1 = _resonance(diff - freqs, broadening, gaussian)
            d2 = _resonance(diff - 2 * freqs, broadening, gaussian)
            dn = _resonance(-diff - freqs, broadening, gaussian)
            factor = 0.5j * o
            for ci, (a, b, c) in enumerate(comps):
                first = dip[a, nn, m] * (der[b, c, m, nn] + der[c, b, m, nn])
                second = der[a, b, nn, m] * dip[c, m, nn] + der[a, c, nn, m] * dip[b, m, nn]
                third = dip[a, nn, m] * (dip[b, m, nn] * (Vd[c, m] - Vd[c, nn]).real
                                         + dip[c, m, nn] * (Vd[b, m] - Vd[b, nn]).real) * inverse
                fourth = der[b, a, nn, m] * dip[c, m, nn] + der[c, a, nn, m] * dip[b, m, nn]
                fifth = (Vd[a, nn] - Vd[a, m]).real * (dip[b, nn, m] * dip[c, m, nn] + dip[c, nn, m] * dip[b, m, nn])
                tf = 0j
                to = 0j
                for p in range(n):
                    if p == nn or p == m:
                        continue
                    tf += dip[a, nn, m] * (rdip[b, m, p] * rdip[c, p, nn] + rdip[c, m, p] * rdip[b, p, nn]) \
                        * _principal(2 * E[p] - E[nn] - E[m], eta)
                    to += rdip[a, m, p] * (rdip[b, p, nn] * dip[c, nn, m] + rdip[c, p, nn] * dip[b, nn, m]) \
                        * _principal(2 * E[nn] - E[m] - E[p], eta)
                    to += rdip[a, p, nn] * (dip[b, nn, m] * rdip[c, m, p] + dip[c, nn, m] * rdip[b, m, p]) \
                        * _principal(2 * E[m] - E[p] - E[nn], eta)
                out[:, ci, 0, 0] += factor * inverse * (2 * first * d2 + second * d1 + (first + second) * low)
                out[:, ci, 1, 0] += factor * inverse * third * (d1 - 4 * d2 - low)
                out[:, ci, 2, 0] += 0.5 * o * (2 * tf * d2 + to * dn)
                out[:, ci, 3, 0] -= factor * inverse * 0.5 * fourth * (d1 + low)
                out[:, ci, 4, 0] -= factor * 0.25 * fifth * d1 * low ** 2
                out[:, ci, 0, 1] += o * (first * d2 + second * d1)
                out[:, ci, 1, 1] += o * third * (d1 - 2 * d2)
                out[:, ci, 2, 1] -= 1j * o * diff * (tf * d2 - to * dn)
                out[:, ci, 3, 1] -= o * 0.5 * fourth * d1
                out[:, ci, 4, 1] -= o * 0.25 * fifth * d1 * low
    if temperature <= 1.0e-7:
        return out
    _one_band(out, E, A, V, W, freqs, comps, bare, fermi_energy, temperature, low_frequency_broadening,
              gaussian, eta, degeneracy_threshold, dA)
    return out


def _one_band(out, E, A, V, W, freqs, comps, bare, mu, temperature, lowb, gaussian, eta, tol, dA):
    """``second_harmonic_one_band!`` (``SecondHarmonicKernels.jl:245-294``)."""
    n = E.shape[0]
    kt = 8.617333262145e-5 * temperature
    low = _principal(freqs, lowb)
    kernel = low - 1j * _delta(freqs, lowb, gaussian)
    for nn in range(n):
        off = E[nn] - mu
        if abs(off) > 5 * kt:
            continue
        d1 = -math.exp(-0.5 * (off / kt) ** 2) / (kt * math.sqrt(2 * math.pi))
        d2 = -off / kt ** 2 * d1
        isolated = (nn == 0 or E[nn] - E[nn - 1] > tol) and (nn == n - 1 or E[nn + 1] - E[nn] > tol)
        curv = None
        if isolated:
            curv = np.zeros((3, 3))
            for a in range(3):
                for b in range(3):
                    if a == b:
                        continue
                    val = dA[b, a, nn, nn] - dA[a, b, nn, nn]
                    for p in range(n):
                        val += (A[b, nn, p] * bare[a, p, nn] - bare[a, nn, p] * A[b, p, nn]
                                - A[a, nn, p] * bare[b, p, nn] + bare[b, nn, p] * A[a, p, nn]
                                - 1j * (bare[a, nn, p] * bare[b, p, nn] - bare[b, nn, p] * bare[a, p, nn]))
                    curv[a, b] = val.real
        for ci, (a, b, c) in enumerate(comps):
            hess = W[b, c, nn, nn].real
            for p in range(n):
                if p == nn:
                    continue
                hess += 2 * (V[b, nn, p] * V[c, p, nn]).real * _principal(E[nn] - E[p], eta)
            va, vb, vc = V[a, nn, nn].real, V[b, nn, nn].real, V[c, nn, nn].real
            semi = d2 * va * vb * vc + d1 * va * hess
            bcd = d1 * (vb * curv[a, c] + vc * curv[a, b]) if isolated else 0.0
            out[:, ci, 5, 0] -= 0.25 * bcd * kernel * low
            out[:, ci, 5, 1] += 0.5j * bcd * kernel
            out[:, ci, 6, 0] -= 0.25j * semi * kernel * low ** 2
            out[:, ci, 6, 1] -= 0.5 * semi * kernel * low


SHG_ELECTRIC_CONSTANT = 8.8541878128e-12
SHG_CHARGE_C = 1.602176634e-19
SHG_HBAR_JS = 1.054571817e-34


def shg_point_values(model, kpts, frequencies, components=None, *, fermi_energy=0.0, temperature=0.0,
                     broadening=0.02, low_frequency_broadening=0.025, gaussian=True,
                     intermediate_regularization=1e-3, eta_correction=True, degeneracy_threshold=0.002):
    """Raw per-point SHG densities [K, n_w, n_comp, 7, 2] (``point_value!`` of
    ``Runtime/Execution/SecondHarmonicDriver.jl:146-168``)."""
    from . import kloop
    data = kloop.compute_kpoint_data(model, np.asarray(kpts, dtype=np.float64), 1e-3)
    K = data.energies.shape[0]
    comps = SECOND_HARMONIC_COMPONENTS if components is None else components
    out = np.zeros((K, len(frequencies), len(comps), 7, 2), dtype=np.complex128)
    for k in range(K):
        out[k] = second_harmonic_response(
            data.energies[k], data.internal_connection[k], data.hamiltonian_derivatives[k],
            data.internal_connection_derivatives[k], data.hamiltonian_second_derivatives[k], frequencies, comps,
            fermi_energy=fermi_energy, temperature=temperature, broadening=broadening,
            low_frequency_broadening=low_frequency_broadening, gaussian=gaussian,
            intermediate_regularization=intermediate_regularization, eta_correction=eta_correction,
            degeneracy_threshold=degeneracy_threshold)
    return out


def shg_factors(model):
    """(chi, sigma) unit factors e/(eps0 V) 1e12 and e^2/(hbar V), V = |det lattice|
    (``SecondHarmonicDriver.jl:135-139``)."""
    vol = abs(np.linalg.det(model.lattice))
    return (SHG_CHARGE_C / (SHG_ELECTRIC_CONSTANT * vol) * 1e12, SHG_CHARGE_C ** 2 / (SHG_HBAR_JS * vol))


def integrate_shg_mesh(model, mesh, frequencies, components=None, **kw):
    """Dense-mesh SHG: lanes[l] = sum of raw point values over lane l (``consume_point!``,
    ``SecondHarmonicDriver.jl:178-185``); total = lanes summed in order, quantity q scaled by
    factors[q] / N_k (``finish!``, ``SecondHarmonicDriver.jl:221-242``)."""
    from . import kloop
    nk = int(np.prod(mesh))
    nl = kloop.lane_count(nk)
    comps = SECOND_HARMONIC_COMPONENTS if components is None else components
    lanes = np.zeros((nl, len(frequencies), len(comps), 7, 2), dtype=np.complex128)
    kp = kloop.integral_kpoints(mesh)
    vals = shg_point_values(model, kp, frequencies, comps, **kw)
    for lane in range(nl):
        a, b = kloop.lane_range(nk, nl, lane)
        lanes[lane] = vals[a:b].sum(axis=0)
    total = lanes.sum(axis=0)
    f = shg_factors(model)
    for q in range(2):
        total[..., q] *= f[q] / nk
    return lanes, total
