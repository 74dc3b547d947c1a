"""ctypes binding of ``libwnlqg_hip.so`` (include/wnlqg.h).

The binding is the host-side mirror a Julia ``ccall`` shim would be: it marshals a
``TightBindingModel`` into the reference's column-major layouts and a resolved task into
``wnlqg_task``.  Loading fails loudly if the library is missing; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from .model import TightBindingModel

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwnlqg_hip.so")

WNLQG_SHIFT_CURRENT = 1
WNLQG_INJECTION_CURRENT = 2
WNLQG_BERRY_CURVATURE_DIPOLE = 4
WNLQG_INTRINSIC_HALL = 8
WNLQG_SECOND_HARMONIC = 16
SECOND_HARMONIC_COMPONENTS = [(a, b, c) for a in range(3) for (b, c) in ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))]
WNLQG_GAUSSIAN = 0
WNLQG_LORENTZIAN = 1
WNLQG_PRECISION_FP64 = 0
WNLQG_PRECISION_MIXED = 1
WNLQG_PRECISION_MIXED_HIGH = 2
WNLQG_PRECISION_SPLIT_BF16 = 3

ERRORS = {-1: "WNLQG_E_INVALID", -2: "WNLQG_E_NONFINITE", -3: "WNLQG_E_GAP", -4: "WNLQG_E_HIP",
          -5: "WNLQG_E_OOM", -6: "WNLQG_E_UNSUPPORTED", -7: "WNLQG_E_NOT_CONVERGED"}

EXPORTED = ["wnlqg_version", "wnlqg_last_error", "wnlqg_create", "wnlqg_destroy", "wnlqg_num_states",
            "wnlqg_num_channels", "wnlqg_run_mesh", "wnlqg_run_points", "wnlqg_run_mesh_reduced",
            "wnlqg_eval_points",
            "wnlqg_eval_points_spectrum",
            "wnlqg_fourier_points",
            "wnlqg_spectrum_points", "wnlqg_synchronize", "wnlqg_mlp_create", "wnlqg_mlp_destroy",
            "wnlqg_mlp_last_error", "wnlqg_mlp_get_params", "wnlqg_mlp_set_params", "wnlqg_mlp_predict",
            "wnlqg_mlp_gradient", "wnlqg_mlp_gradient_noalloc", "wnlqg_mlp_train", "wnlqg_batch_create",
            "wnlqg_batch_run", "wnlqg_batch_tuning",
            "wnlqg_batch_destroy"]


class WnlqgError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class _Model(C.Structure):
    _fields_ = [("num_orbitals", C.c_int32), ("num_r_vectors", C.c_int32),
                ("lattice", C.POINTER(C.c_double)), ("r_vectors", C.POINTER(C.c_int32)),
                ("r_degeneracies", C.POINTER(C.c_int32)), ("hamiltonian_r", C.POINTER(C.c_double)),
                ("position_r", C.POINTER(C.c_double))]


class _Task(C.Structure):
    _fields_ = [("quantities", C.c_int32), ("num_photon_energies", C.c_int32),
                ("photon_energies", C.POINTER(C.c_double)), ("broadening_type", C.c_int32),
                ("broadening", C.c_double), ("fermi_energy", C.c_double), ("temperature", C.c_double),
                ("denominator_regularization", C.c_double), ("transition_window_factor", C.c_double),
                ("band_window_size", C.c_int32), ("precision", C.c_int32), ("batch_size", C.c_int32),
                ("fourier_backend", C.c_int32), ("num_fermi_energies", C.c_int32),
                ("fermi_energies", C.POINTER(C.c_double)), ("fermi_surface_broadening", C.c_double),
                ("degeneracy_threshold", C.c_double), ("num_shg_components", C.c_int32),
                ("shg_components", C.POINTER(C.c_int32)), ("shg_low_frequency_broadening", C.c_double),
                ("shg_intermediate_regularization", C.c_double), ("shg_eta_correction", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("interp_ms", C.c_double), ("eig_ms", C.c_double),
                ("rotate_ms", C.c_double), ("pairs_ms", C.c_double), ("accum_ms", C.c_double),
                ("kpoints", C.c_int64), ("active_pairs", C.c_int64), ("batches", C.c_int32),
                ("interp_launches", C.c_int32), ("interp_flops", C.c_double), ("max_sweeps", C.c_int32),
                ("separable", C.c_int32), ("interp_bytes", C.c_double), ("mean_sweeps", C.c_double),
                ("warm_ms", C.c_double), ("eig_launches", C.c_int32), ("rotate_row_points", C.c_int32),
                ("rotate_element_points", C.c_int64), ("rotate_full_points", C.c_int64),
                ("rotate_element_bands", C.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


ABI_VERSION = 9   # WNLQG_ABI_VERSION of include/wnlqg.h
# fourier_backend names -> WNLQG_FOURIER_*: the reference's "auto", "direct", "mixed" and this build's "planes"
FOURIER_BACKENDS = {"auto": 0, "direct": 1, "planes": 2, "mixed": 3}


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m wanniernlqg_amd.build` "
                           "(the engine has no CPU fallback)")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    lib.wnlqg_version.restype = C.c_int32
    lib.wnlqg_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.wnlqg_last_error.restype = C.c_int32
    lib.wnlqg_create.argtypes = [C.POINTER(_Model), C.POINTER(_Task), C.POINTER(C.c_void_p)]
    lib.wnlqg_destroy.argtypes = [C.c_void_p]
    lib.wnlqg_num_states.argtypes = [C.c_void_p]
    lib.wnlqg_num_channels.argtypes = [C.c_void_p]
    lib.wnlqg_run_mesh.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, dp, dp, C.POINTER(Stats)]
    lib.wnlqg_run_points.argtypes = [C.c_void_p, dp, dp, C.c_int64, C.c_int32, dp, dp]
    lib.wnlqg_run_mesh_reduced.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), dp, C.c_int64,
                                           C.c_int32, dp, dp]
    lib.wnlqg_eval_points.argtypes = [C.c_void_p, dp, C.c_int32, dp]
    lib.wnlqg_eval_points_spectrum.argtypes = [C.c_void_p, dp, C.c_int32, dp, dp]
    lib.wnlqg_fourier_points.argtypes = [C.c_void_p, dp, C.c_int32, dp]
    lib.wnlqg_spectrum_points.argtypes = [C.c_void_p, dp, C.c_int32, dp, dp]
    lib.wnlqg_synchronize.argtypes = [C.c_void_p]
    fp = C.POINTER(C.c_float)
    lib.wnlqg_mlp_create.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
    lib.wnlqg_mlp_destroy.argtypes = [C.c_void_p]
    lib.wnlqg_mlp_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.wnlqg_mlp_get_params.argtypes = [C.c_void_p, fp]
    lib.wnlqg_mlp_set_params.argtypes = [C.c_void_p, fp]
    lib.wnlqg_mlp_predict.argtypes = [C.c_void_p, fp, C.c_int64, fp]
    lib.wnlqg_mlp_gradient.argtypes = [C.c_void_p, fp, fp, C.c_int64, fp, fp]
    lib.wnlqg_mlp_gradient_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                               C.c_void_p]
    lib.wnlqg_mlp_train.argtypes = [C.c_void_p, fp, fp, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, fp]
    lib.wnlqg_batch_create.argtypes = [C.c_int32, C.POINTER(_Model), C.POINTER(_Task), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_void_p)]
    lib.wnlqg_batch_run.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double)),
                                    C.POINTER(C.c_double)]
    lib.wnlqg_batch_tuning.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    lib.wnlqg_batch_destroy.argtypes = [C.c_void_p]
    for name in EXPORTED:
        getattr(lib, name).restype = C.c_int32
    if lib.wnlqg_version() != ABI_VERSION:
        raise RuntimeError(f"{path} reports ABI version {lib.wnlqg_version()}, these bindings expect {ABI_VERSION}: "
                           "rebuild it with `python -m wanniernlqg_amd.build`")
    _lib = lib
    return lib


def _check(lib, rc: int):
    if rc != 0:
        buf = C.create_string_buffer(2048)
        lib.wnlqg_last_error(buf, 2048)
        raise WnlqgError(rc, buf.value.decode(errors="replace"))


def _ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class EngineTask:
    """Resolved task parameters (subset of ``EffectiveTaskConfig``, reference ``src/Runtime/Config.jl:4-170``;
    defaults from ``Config.jl:126-155``).  One engine serves either SC and/or IC (the Integral driver), or
    BCD alone (``BerryMultipoleDriver.jl``; ``fermi_energies``, ``temperature``, ``fermi_surface_broadening``,
    ``degeneracy_threshold``, ``denominator_regularization`` = 0 by default for BCD,
    ``TaskConfiguration.jl:568-578``), or intrinsic AHC alone (``SpectralResponseDriver.jl``, T = 0)."""
    photon_energies: Optional[np.ndarray] = None
    shift_current: Optional[bool] = None      # None: on unless another quantity is requested
    injection_current: bool = False
    berry_curvature_dipole: bool = False
    intrinsic_hall: bool = False
    fermi_energies: Optional[np.ndarray] = None
    fermi_surface_broadening: float = 0.0
    degeneracy_threshold: float = 0.002
    second_harmonic: bool = False             # SHG alone (SecondHarmonicDriver.jl)
    shg_components: Optional[Sequence] = None  # (a, b, c) triples, 0-based; None = all 18
    shg_low_frequency_broadening: float = 0.025
    shg_intermediate_regularization: float = 1.0e-3
    shg_eta_correction: bool = True
    broadening: float = 0.040
    broadening_type: str = "Gaussian"
    fermi_energy: float = 0.0
    temperature: float = 0.0
    denominator_regularization: float = 0.001
    transition_window_factor: float = 5.0
    band_window_size: int = -1
    precision: str = "fp64"
    batch_size: int = 0
    fourier_backend: str = "auto"      # meshes: "auto"/"planes"/"mixed" (separable) or "direct"


class Engine:
    """One model + one resolved task on the current HIP device."""

    def __init__(self, model: TightBindingModel, task: EngineTask):
        m, t = self._marshal(model, task)
        h = C.c_void_p()
        _check(self.lib, self.lib.wnlqg_create(C.byref(m), C.byref(t), C.byref(h)))
        self.handle = h
        self.num_states = self.lib.wnlqg_num_states(h)
        self.num_channels = self.lib.wnlqg_num_channels(h)

    def _marshal(self, model: TightBindingModel, task: EngineTask):
        """Julia-layout copies of the model and the wnlqg_model / wnlqg_task structs (kept alive by self)."""
        self.lib = load_library()
        self.model = model
        self.task = task
        self.handle = None
        n, nr = model.num_orbitals, model.num_r_vectors
        # Julia column-major layouts
        self._lat = np.ascontiguousarray(model.lattice.T).ravel()                     # lattice[i,a] at i+3a
        self._rv = np.ascontiguousarray(model.r_vectors, dtype=np.int32).ravel()     # [3,N_R]
        self._deg = np.ascontiguousarray(model.r_degeneracies, dtype=np.int32)
        self._hr = np.ascontiguousarray(np.transpose(model.hamiltonian_r, (0, 2, 1))).view(np.float64).ravel()
        self._ar = None
        if model.position_r is not None:
            self._ar = np.ascontiguousarray(np.transpose(model.position_r, (0, 1, 3, 2))).view(np.float64).ravel()
        geometry = task.berry_curvature_dipole or task.intrinsic_hall
        shg = task.second_harmonic
        comps = SECOND_HARMONIC_COMPONENTS if task.shg_components is None else [tuple(c) for c in task.shg_components]
        self._comps = np.ascontiguousarray(np.array(comps, dtype=np.int32).ravel())
        self._om = np.ascontiguousarray(np.zeros(1) if task.photon_energies is None else task.photon_energies,
                                        dtype=np.float64)
        self._mu = np.ascontiguousarray(np.atleast_1d(np.zeros(1) if task.fermi_energies is None
                                                      else task.fermi_energies), dtype=np.float64)
        m = _Model(n, nr, _ptr(self._lat), _ptr(self._rv, C.c_int32), _ptr(self._deg, C.c_int32), _ptr(self._hr),
                   _ptr(self._ar) if self._ar is not None else C.POINTER(C.c_double)())
        sc = task.shift_current if task.shift_current is not None else not (geometry or shg or task.injection_current)
        q = ((WNLQG_SHIFT_CURRENT if sc else 0) | (WNLQG_INJECTION_CURRENT if task.injection_current else 0) |
             (WNLQG_BERRY_CURVATURE_DIPOLE if task.berry_curvature_dipole else 0) |
             (WNLQG_INTRINSIC_HALL if task.intrinsic_hall else 0) | (WNLQG_SECOND_HARMONIC if shg else 0))
        bt = {"Gaussian": WNLQG_GAUSSIAN, "Lorentzian": WNLQG_LORENTZIAN}.get(task.broadening_type, 99)
        prec = {"fp64": WNLQG_PRECISION_FP64, "mixed": WNLQG_PRECISION_MIXED,
                "mixed6": WNLQG_PRECISION_MIXED_HIGH, "bf16x3": WNLQG_PRECISION_SPLIT_BF16}[task.precision]
        t = _Task(q, self._om.shape[0], _ptr(self._om), bt, task.broadening, task.fermi_energy, task.temperature,
                  task.denominator_regularization, task.transition_window_factor, task.band_window_size, prec,
                  task.batch_size, FOURIER_BACKENDS.get(task.fourier_backend, -1),
                  self._mu.shape[0] if geometry else 0, _ptr(self._mu), task.fermi_surface_broadening,
                  task.degeneracy_threshold, 0 if task.shg_components is None else len(comps),
                  _ptr(self._comps, C.c_int32), task.shg_low_frequency_broadening,
                  task.shg_intermediate_regularization, 1 if task.shg_eta_correction else 0)
        self.shg = shg
        self.shg_components = comps
        if shg:
            self.states = ["shg"]
        elif geometry:
            self.states = (["bcd"] if task.berry_curvature_dipole else []) + (["ahc"] if task.intrinsic_hall else [])
        else:
            self.states = [s for s, on in (("sc", sc), ("ic", task.injection_current)) if on]
        self.geometry = geometry
        self._axis = self._mu.shape[0] if geometry else self._om.shape[0]
        # complex elements per k point (or lane) of each state
        if geometry:
            self._sizes = [(s, (27 if s == "bcd" else 9) * self._axis) for s in self.states]
        else:
            self._sizes = [(s, 27 * self._axis) for s in self.states]
        self.num_states = len(self.states)
        self._structs = (m, t)
        return m, t

    def _unpack_mesh(self, lane_out: np.ndarray, total_out: np.ndarray, nl: int):
        if self.shg:
            return {"shg": self._unpack_shg(lane_out, (nl,))}, {"shg": self._unpack_shg(total_out, ())}
        if self.geometry:
            return self._unpack_geometry(lane_out, (nl,)), self._unpack_geometry(total_out, ())
        nw = self._axis
        lanes = _unpack(lane_out, (nl,), self.num_states, nw)
        total = _unpack(total_out, (), self.num_states, nw)
        return ({s: lanes[..., i, :, :, :, :] for i, s in enumerate(self.states)},
                {s: total[i] for i, s in enumerate(self.states)})

    def close(self):
        if getattr(self, "handle", None):
            self.lib.wnlqg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- entry points ---------------------------------------------------------------------
    def run_mesh(self, mesh: Sequence[int], n_lanes: Optional[int] = None):
        """Integrate the Gamma-centred mesh.  Returns (lanes, total, stats) with
        lanes[state] of shape [n_lanes, n_w, 3, 3, 3] and total[state] [n_w, 3, 3, 3] ([a, b, c] last)."""
        mesh_a = np.ascontiguousarray(mesh, dtype=np.int32)
        nk = int(np.prod(mesh_a.astype(np.int64)))
        nl = min(nk, 64) if n_lanes is None else int(n_lanes)
        nw = self._axis
        per = self._result_elems()
        lane_out = np.zeros(nl * per * 2)
        total_out = np.zeros(per * 2)
        st = Stats()
        _check(self.lib, self.lib.wnlqg_run_mesh(self.handle, _ptr(mesh_a, C.c_int32), nl, _ptr(lane_out),
                                                 _ptr(total_out), C.byref(st)))
        lanes, total = self._unpack_mesh(lane_out, total_out, nl)
        return lanes, total, st.as_dict()

    def run_points(self, kpoints: np.ndarray, weights: np.ndarray, n_lanes: Optional[int] = None):
        """Weighted point list (SC / IC, BCD and AHC engines): returns (lanes, total) like run_mesh, each
        lane the sum of weights[i] * value(k_i) over its items (the k-loop of a symmetry-reduced mesh, or
        sub-cell nodes of a refinement: weights 1/N_k over a whole mesh give run_mesh's total)."""
        k = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        if w.shape[0] != k.shape[0]:
            raise ValueError("one weight per k point required")
        m = k.shape[0]
        nl = min(m, 64) if n_lanes is None else int(n_lanes)
        per = self._result_elems()
        lane_out = np.zeros(nl * per * 2)
        total_out = np.zeros(per * 2)
        _check(self.lib, self.lib.wnlqg_run_points(self.handle, _ptr(k), _ptr(w), m, nl, _ptr(lane_out),
                                                   _ptr(total_out)))
        return self._unpack_mesh(lane_out, total_out, nl)

    def run_mesh_reduced(self, mesh: Sequence[int], reps: np.ndarray, weights: np.ndarray,
                         n_lanes: Optional[int] = None):
        """The mesh points reps (sorted 0-based linear indices) with weights, interpolated by whole planes
        (SC / IC engines, separable backend): returns (lanes, total) like run_points."""
        mesh_a = np.ascontiguousarray(mesh, dtype=np.int32)
        r = np.ascontiguousarray(reps, dtype=np.int64).reshape(-1)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        if w.shape[0] != r.shape[0]:
            raise ValueError("one weight per representative required")
        m = r.shape[0]
        nl = min(m, 64) if n_lanes is None else int(n_lanes)
        per = self._result_elems()
        lane_out = np.zeros(nl * per * 2)
        total_out = np.zeros(per * 2)
        _check(self.lib, self.lib.wnlqg_run_mesh_reduced(self.handle, _ptr(mesh_a, C.c_int32), _ptr(r, C.c_int64), _ptr(w),
                                                         m, nl, _ptr(lane_out), _ptr(total_out)))
        return self._unpack_mesh(lane_out, total_out, nl)

    def _result_elems(self) -> int:
        if self.shg:
            return self._om.shape[0] * len(self.shg_components) * 14
        return sum(size for _, size in self._sizes)

    def _unpack_shg(self, flat: np.ndarray, lead: tuple) -> np.ndarray:
        """Complex column-major [n_w, n_comp, 7, 2] -> [..., n_w, n_comp, 7, 2]."""
        nc = len(self.shg_components)
        z = flat.view(np.complex128).reshape(lead + (2, 7, nc, self._om.shape[0]))
        L = len(lead)
        return np.transpose(z, tuple(range(L)) + (L + 3, L + 2, L + 1, L)).copy()

    def _unpack_geometry(self, flat: np.ndarray, lead: tuple) -> dict:
        """Per state: complex column-major [n_mu,3,3,3] (BCD) then [n_mu,3,3] (AHC) -> real [..., n_mu, a, b(, c)]."""
        total = self._result_elems()
        z = flat.view(np.complex128).reshape(lead + (total,))
        out, off = {}, 0
        L = len(lead)
        for s, size in self._sizes:
            blk = z[..., off:off + size]
            off += size
            if s == "bcd":
                blk = blk.reshape(lead + (3, 3, 3, self._axis))
                perm = (L + 3, L + 2, L + 1, L)
            else:
                blk = blk.reshape(lead + (3, 3, self._axis))
                perm = (L + 2, L + 1, L)
            out[s] = np.transpose(blk, tuple(range(L)) + perm).real.copy()
        return out

    def eval_points(self, kpoints: np.ndarray, energies: bool = False):
        """Per-point values at arbitrary k; with energies=True also the ascending band energies
        [m, n] of the same evaluation (returned as a second value)."""
        k = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
        nw = self._axis
        out = np.zeros(k.shape[0] * self._result_elems() * 2)
        if energies:
            en = np.zeros((k.shape[0], self.model.num_orbitals))
            _check(self.lib, self.lib.wnlqg_eval_points_spectrum(self.handle, _ptr(k), k.shape[0], _ptr(out),
                                                                 _ptr(en)))
        else:
            _check(self.lib, self.lib.wnlqg_eval_points(self.handle, _ptr(k), k.shape[0], _ptr(out)))
        if self.shg:
            res = {"shg": self._unpack_shg(out, (k.shape[0],))}
        elif self.geometry:
            res = self._unpack_geometry(out, (k.shape[0],))
        else:
            vals = _unpack(out, (k.shape[0],), self.num_states, nw)
            res = {s: vals[:, i] for i, s in enumerate(self.states)}
        return (res, en) if energies else res

    def fourier_points(self, kpoints: np.ndarray) -> np.ndarray:
        k = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
        n = self.model.num_orbitals
        out = np.zeros(k.shape[0] * self.num_channels * n * n * 2)
        _check(self.lib, self.lib.wnlqg_fourier_points(self.handle, _ptr(k), k.shape[0], _ptr(out)))
        return out.view(np.complex128).reshape(k.shape[0], self.num_channels, n, n)

    def spectrum_points(self, kpoints: np.ndarray):
        k = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
        n = self.model.num_orbitals
        e = np.zeros(k.shape[0] * n)
        v = np.zeros(k.shape[0] * n * n * 2)
        _check(self.lib, self.lib.wnlqg_spectrum_points(self.handle, _ptr(k), k.shape[0], _ptr(e), _ptr(v)))
        return e.reshape(k.shape[0], n), v.view(np.complex128).reshape(k.shape[0], n, n)

    def synchronize(self):
        _check(self.lib, self.lib.wnlqg_synchronize(self.handle))


def _unpack(flat: np.ndarray, lead: tuple, nq: int, nw: int) -> np.ndarray:
    """Complex [..., state, (n_w,3,3,3 column-major)] -> [..., state, n_w, a, b, c]."""
    z = flat.view(np.complex128).reshape(lead + (nq, 3, 3, 3, nw))   # c, b, a, w (C order of Julia [w,a,b,c])
    return np.transpose(z, tuple(range(len(lead))) + (len(lead), len(lead) + 4, len(lead) + 3, len(lead) + 2,
                                                       len(lead) + 1))


class Batch:
    """Batching form (``wnlqg_batch_*``): many (model, task, mesh) integrals of uneven n_w captured in
    one HIP graph and replayed by one launch."""

    def __init__(self, entries, autotune: bool = True):
        """entries: sequence of (model, EngineTask, mesh[, n_lanes])."""
        self.lib = load_library()
        self.shells = []
        models, tasks, meshes, lanes = [], [], [], []
        for ent in entries:
            model, task, mesh = ent[0], ent[1], ent[2]
            shell = Engine.__new__(Engine)
            m, t = shell._marshal(model, task)
            nk = int(np.prod(np.asarray(mesh, dtype=np.int64)))
            nl = min(nk, 64) if len(ent) < 4 or ent[3] is None else int(ent[3])
            self.shells.append((shell, nl))
            models.append(m)
            tasks.append(t)
            meshes.extend(int(x) for x in mesh)
            lanes.append(nl)
        count = len(self.shells)
        self._models = (_Model * count)(*models)
        self._tasks = (_Task * count)(*tasks)
        self._meshes = np.ascontiguousarray(meshes, dtype=np.int32)
        self._lanes = np.ascontiguousarray(lanes, dtype=np.int32)
        h = C.c_void_p()
        _check(self.lib, self.lib.wnlqg_batch_create(count, self._models, self._tasks, _ptr(self._meshes, C.c_int32),
                                                     _ptr(self._lanes, C.c_int32), 1 if autotune else 0,
                                                     C.byref(h)))
        self.handle = h

    def tuning(self):
        out = []
        for i in range(len(self.shells)):
            v = C.c_int32()
            _check(self.lib, self.lib.wnlqg_batch_tuning(self.handle, i, C.byref(v)))
            out.append(v.value)
        return out

    def run(self):
        """Replay the graph; returns ([(lanes, total) per entry], device milliseconds)."""
        bufs = []
        for shell, nl in self.shells:
            per = shell._result_elems()
            bufs.append((np.zeros(nl * per * 2), np.zeros(per * 2)))
        lp = (C.POINTER(C.c_double) * len(bufs))(*[_ptr(b[0]) for b in bufs])
        tp = (C.POINTER(C.c_double) * len(bufs))(*[_ptr(b[1]) for b in bufs])
        ms = C.c_double()
        _check(self.lib, self.lib.wnlqg_batch_run(self.handle, lp, tp, C.byref(ms)))
        results = [shell._unpack_mesh(b[0], b[1], nl) for (shell, nl), b in zip(self.shells, bufs)]
        return results, ms.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.wnlqg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
