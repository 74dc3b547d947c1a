"""ctypes wrapper of ``oracle/kloop_omp.cpp``, the C++/OpenMP restatement of the shift / injection current
k-loop (TEST INFRASTRUCTURE ONLY: the CPU baseline of ``bench.py`` and an independent second implementation
that ``tests/test_oracle_cpp.py`` checks against ``oracle/kloop.py``).  Nothing in the product path imports it.

``build()`` compiles ``liboracle_kloop.so`` next to this file with g++ (-O3, OpenMP, x86-64-v3);
``integrate_planes()`` returns the same total as ``kloop.integrate_mesh`` restricted to whole k3 planes
(the lanes' sum with the prefactors and 1/N_k applied)."""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import kloop

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCE = os.path.join(HERE, "kloop_omp.cpp")
LIB = os.path.join(HERE, "liboracle_kloop.so")
_lib = None


def build(verbose: bool = False) -> str:
    """Compile the library (always, like the HIP build); returns its path."""
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-o", LIB + ".tmp", SOURCE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with oracle.kloop_omp.build()")
        lib = C.CDLL(LIB)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        lib.oracle_kloop_planes.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, ip, C.c_int, C.c_int, dp, C.c_int,
                                            C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_int, C.c_int, C.c_int, dp, dp, dp]
        lib.oracle_kloop_planes.restype = C.c_int
        lib.oracle_hermitian_eigen.argtypes = [C.c_int, dp, dp, dp]
        lib.oracle_hermitian_eigen.restype = C.c_int
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def hermitian_eigen(h: np.ndarray):
    """Eigenvalues (ascending) and canonical-phase eigenvectors of the upper triangle of one matrix."""
    lib = load()
    h = np.ascontiguousarray(h, dtype=np.complex128)
    n = h.shape[0]
    e = np.empty(n)
    v = np.empty((n, n), dtype=np.complex128)
    lib.oracle_hermitian_eigen(n, _dp(h.view(np.float64)), _dp(e), _dp(v.view(np.float64)))
    return e, v


def integrate_planes(model, task: kloop.Task, mesh, plane_begin: int, plane_end: int, threads: int = 0,
                     timing: dict = None):
    """Sum over the k3 planes [plane_begin, plane_end) of the Gamma-centred mesh of (1/N_k) x prefactor x
    point value, per quantity: {'sc', 'ic'} -> [n_w, 3, 3, 3] complex (Gaussian or Lorentzian, band window -1).
    ``timing`` (optional dict) receives 'loop_s', the wall time of the k-loop after the R-space set-up."""
    lib = load()
    if task.band_window_size != -1 and 2 * task.band_window_size <= model.num_orbitals:
        raise NotImplementedError("band windows other than -1 are not restated")
    n, nr = model.num_orbitals, model.num_r_vectors
    rv = np.ascontiguousarray(model.r_vectors, dtype=np.int32)
    deg = np.ascontiguousarray(model.r_degeneracies, dtype=np.int32)
    lat = np.ascontiguousarray(model.lattice, dtype=np.float64)
    hr = np.ascontiguousarray(model.hamiltonian_r, dtype=np.complex128)
    ar = None if model.position_r is None else np.ascontiguousarray(model.position_r, dtype=np.complex128)
    om = np.ascontiguousarray(task.photon_energies, dtype=np.float64)
    me = np.ascontiguousarray(mesh, dtype=np.int32)
    nw = om.shape[0]
    sc = np.zeros((nw, 27), dtype=np.complex128)
    ic = np.zeros((nw, 27), dtype=np.complex128)
    loop_s = np.zeros(1)
    rc = lib.oracle_kloop_planes(n, nr, _ip(rv), _ip(deg), _dp(lat), _dp(hr.view(np.float64)),
                                 _dp(ar.view(np.float64)) if ar is not None else C.POINTER(C.c_double)(), _ip(me),
                                 int(plane_begin), int(plane_end), _dp(om), nw, float(task.broadening),
                                 1 if task.broadening_type == "Gaussian" else 0, float(task.fermi_energy),
                                 float(task.temperature), float(task.denominator_regularization),
                                 float(task.transition_window_factor), 1 if task.shift_current else 0,
                                 1 if task.injection_current else 0, int(threads), _dp(sc.view(np.float64)),
                                 _dp(ic.view(np.float64)), _dp(loop_s))
    if timing is not None:
        timing["loop_s"] = float(loop_s[0])
    if rc != 0:
        raise ValueError("the R vectors do not fit a box of span <= 64 per axis")
    nk = int(np.prod(mesh))
    vol = model.cell_volume
    out = {}
    if task.shift_current:
        out["sc"] = (kloop.shift_current_prefactor(vol) / nk * sc).reshape(nw, 3, 3, 3)
    if task.injection_current:
        out["ic"] = (kloop.injection_current_prefactor(vol) / nk * ic).reshape(nw, 3, 3, 3)
    return out
