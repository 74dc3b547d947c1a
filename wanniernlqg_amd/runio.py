"""Run files and result files of the reference's CLI for the integral tasks the engine serves.

A user of the reference runs ``scripts/wanniernlqg.jl run <cfg>.toml`` and reads
``<system>_sc_conv.dat`` / ``<system>_ic_conv.dat`` (``shg_<chi|sigma>_<abc>.dat`` for SHG).  This
module reads the same TOML run file (the subset of the schema of ``src/Runtime/TaskInput.jl`` that a
dense-mesh shift / injection current or SHG integral of a Wannier ``_tb.dat`` model uses), runs the HIP
engine on the mesh, and writes the result files in the reference's formats:

* ``write_response_tensor`` follows ``src/IO/ResultWriters.jl:8-80`` (header, rightmost-index-fastest
  columns, ``%.<digits>e`` numbers, digits in 7..17);
* ``result_filename`` follows ``src/Runtime/ResultFilenames.jl:30-60`` with the observable suffixes of
  ``src/Runtime/Planning/TaskRegistry.jl`` (``sc``, ``ic``, ``shg``) and the method suffix ``conv``;
* ``write_second_harmonic`` follows ``src/IO/SecondHarmonicWriters.jl:8-70`` (integral rows);
* ``energy_range`` follows ``_task_input_energy_range`` (``src/Runtime/TaskInputEnergyAxes.jl``).

Supported run-file keys (others are rejected, so that a run never silently ignores a setting):
``[model] source = "wannier", model_file``; ``[response] photon_energies_ev`` or
``photon_energy_{min,max}_ev`` with ``photon_energy_step_ev`` or ``photon_energy_num_points``,
``fermi_energy_ev``, ``temperature_k``; ``[response.sampling] kind = "bz_mesh", k_mesh,
spatial_dimension = 3``; ``[response.execution] fourier_backend`` = "auto", "direct" or "mixed" (the reference's
names, FourierExecutionPlan.jl:63-67; "mixed" runs the engine's separable contraction, which evaluates the same
uniform-mesh Fourier sums as the reference's Mixed-FFT); ``[response.tasks.<id>] kind`` in SC / shift_current, IC /
injection_current, SHG / second_harmonic_generation (with the selectors ``output`` = total / terms /
both and ``response`` = conductivity / susceptibility / both), ``method = "Conventional"``, optional
``[response.tasks.<id>.numerics]`` with ``broadening``, ``broadening_type``,
``denominator_regularization``, ``transition_window_factor``, ``band_window_size``,
``degeneracy_threshold``, ``low_frequency_broadening``, ``intermediate_regularization``, ``eta_correction``; ``[output]
output_root``, ``system_name``, ``response_output_digits``, ``progress_enabled``.  Each task's result
files go to ``<output_root>/<task id>/`` (``output_root`` relative to the run file, default the run file's
directory; ``src/Runtime/TaskConfiguration.jl:1711-1750``), named with the system name derived from the
model file when none is given (``src/Runtime/TaskInput.jl:934-958``: the file stem without ``_tb``).

    python -m wanniernlqg_amd.runio run <cfg>.toml
"""

from __future__ import annotations

import math
import os
import re
import sys
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Sequence

import numpy as np

# quantity -> observable suffix of the integral result file (TaskRegistry.jl: "sc", "ic", "shg")
OBSERVABLE_SUFFIX = {"shift_current": "sc", "injection_current": "ic", "second_harmonic_generation": "shg"}
TASK_KINDS = {"sc": "shift_current", "shift_current": "shift_current", "ic": "injection_current",
              "injection_current": "injection_current", "shg": "second_harmonic_generation",
              "second_harmonic_generation": "second_harmonic_generation"}
# ordered algebraic SHG contributions (SecondHarmonicKernels.jl:71-79)
SECOND_HARMONIC_TERMS = ("interband_derivative", "interband_velocity", "three_band", "transport_derivative",
                         "transport_velocity", "berry_curvature_dipole", "semiclassical")
MAX_GENERATED_AXIS_POINTS = 1_000_000
OPTICAL_DEFAULTS = dict(broadening=0.040, broadening_type="Gaussian", transition_window_factor=5.0,
                        denominator_regularization=0.001, degeneracy_threshold=0.002, band_window_size=-1)
SHG_DEFAULTS = dict(broadening=0.025, broadening_type="Gaussian", low_frequency_broadening=0.025,
                    intermediate_regularization=0.001, eta_correction=True, degeneracy_threshold=1.0e-4)
NUMERICS_KEYS = {"broadening", "broadening_type", "denominator_regularization", "transition_window_factor",
                 "band_window_size", "degeneracy_threshold", "low_frequency_broadening",
                 "intermediate_regularization", "eta_correction"}


def energy_range(lower, upper, step=None, count=None) -> np.ndarray:
    """Photon-energy axis from endpoints and a step or a point count (``_task_input_energy_range``,
    ``src/Runtime/TaskInputEnergyAxes.jl:34-90``): with a count, ``range(lo, hi; length = count)`` (each
    point the correctly rounded lo + i (hi - lo) / (count - 1)); with a step, lo + i * step for
    i = 0 .. n - 1, n = floor((hi - lo + tol) / step) + 1 in exact rational arithmetic,
    tol = min(8 eps(max(|lo|, |hi|, |step|)), 1e-9 step), the last point clamped to hi when it overshoots
    by at most tol.  Booleans, non-finite values, collapsing endpoints and more than 1,000,000 points are
    rejected with ValueError."""
    def real(x):
        return isinstance(x, (int, float, np.integer, np.floating)) and not isinstance(x, (bool, np.bool_))

    for v in (lower, upper):
        if not real(v) or not math.isfinite(v) or not math.isfinite(float(v)):
            raise ValueError("photon energy endpoints must be finite representable real numbers")
    lo, hi = float(lower), float(upper)
    if upper < lower:
        raise ValueError("photon energy maximum must be at least its minimum")
    if upper != lower and not hi > lo:
        raise ValueError("distinct photon energy endpoints collapse at Float64 resolution")
    if count is not None:
        if not (isinstance(count, (int, np.integer)) and not isinstance(count, (bool, np.bool_))
                and 1 <= count <= MAX_GENERATED_AXIS_POINTS):
            raise ValueError(f"photon energy point count must be an integer in 1:{MAX_GENERATED_AXIS_POINTS}")
        if not (lo == hi if count == 1 else hi > lo):
            raise ValueError("one point requires equal endpoints; multiple points require distinct endpoints")
        if count == 1:
            return np.array([lo])
        span, m = Fraction(hi) - Fraction(lo), int(count) - 1
        out = np.array([float(Fraction(lo) + span * i / m) for i in range(m + 1)])
    else:
        if not (real(step) and math.isfinite(step) and step > 0 and float(step) > 0):
            raise ValueError("photon energy step must be finite, representable and positive")
        delta = float(step)
        tol = min(8 * float(np.spacing(max(abs(lo), abs(hi), abs(delta)))), delta * 1e-9)
        n = math.floor((Fraction(hi) - Fraction(lo) + Fraction(tol)) / Fraction(delta)) + 1
        if not 1 <= n <= MAX_GENERATED_AXIS_POINTS:
            raise ValueError("generated photon energy point count exceeds the allocation limit")
        out = np.array([lo + i * delta for i in range(n)])
        if out[-1] > hi:
            if out[-1] - hi > tol:
                raise ValueError("photon energy endpoint exceeds maximum")
            out[-1] = hi
    if not (np.all(np.isfinite(out)) and np.all(np.diff(out) > 0)):
        raise ValueError("photon energy range is not strictly increasing at Float64 resolution")
    return out


def _fmt(value: float, digits: int) -> str:
    """C ``%.<digits>e``; non-finite values spelled as Julia's Printf spells them."""
    if math.isnan(value):
        return "NaN"
    if math.isinf(value):
        return "Inf" if value > 0 else "-Inf"
    return f"%.{digits}e" % value


def write_response_tensor(fermi_energy: float, photon_energies: Sequence[float], tensor: np.ndarray,
                          spatial_dimension: int, directory: str, filename: str, digits: int = 7,
                          region_metadata: Optional[Dict[str, object]] = None) -> str:
    """Photon-energy rows, then real and imaginary parts of every Cartesian component with the rightmost
    index fastest (``write_response_tensor``, ``src/IO/ResultWriters.jl:8-80``).  ``tensor`` is
    [n_energy, axes...]; returns the written path."""
    if not 7 <= digits <= 17:
        raise ValueError(f"response output digits must lie in 7:17; got {digits}")
    tensor = np.asarray(tensor, dtype=np.complex128)
    omega = np.asarray(photon_energies, dtype=np.float64)
    axes = list(tensor.shape[1:])
    if tensor.shape[0] != omega.shape[0]:
        raise ValueError(f"Dimension mismatch: tensor has {tensor.shape[0]} energy points, expected {omega.shape[0]}.")
    if not (all(a == spatial_dimension for a in axes)
            or (len(axes) == 4 and axes == [spatial_dimension, 3, spatial_dimension, spatial_dimension])):
        raise ValueError(f"Unsupported response tensor axis sizes={tuple(axes)} for "
                         f"spatial_dimension={spatial_dimension}.")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, filename)
    coords = "xyz"
    comps = list(np.ndindex(*axes))          # C order: rightmost index fastest
    with open(path, "w") as fh:
        fh.write("#### WannierNLQG response tensor\n")
        for key, value in (region_metadata or {}).items():
            fh.write(f"# {key}={value}\n")
        fh.write("#    Efermi     Omega" + "".join("     " + "".join(coords[i] for i in idx) for idx in comps) + "\n")
        for w in range(omega.shape[0]):
            line = _fmt(fermi_energy, digits) + "    " + _fmt(omega[w], digits)
            for idx in comps:
                v = tensor[(w,) + idx]
                line += "    " + _fmt(v.real, digits) + "    " + _fmt(v.imag, digits) + "    "
            fh.write(line + "\n")
    return path


def result_filename(system_name: str, quantity: str, method: str = "conventional",
                    calculation: str = "integral") -> str:
    """``<system>_<observable>_<method>.dat`` for integral results (``ResultFilenames.jl:30-60``)."""
    if calculation != "integral":
        raise ValueError("only integral result filenames are produced here")
    if method.lower() != "conventional":
        raise ValueError(f"No result filename suffix for method={method}.")
    if quantity not in OBSERVABLE_SUFFIX:
        raise ValueError(f"No result filename observable for quantity={quantity}, calculation={calculation}.")
    return f"{system_name}_{OBSERVABLE_SUFFIX[quantity]}_conv.dat"


def write_second_harmonic(directory: str, coordinates: np.ndarray, values: np.ndarray,
                          components: Sequence[Sequence[int]], terms: Sequence[str] = SECOND_HARMONIC_TERMS,
                          output: str = "total", response: str = "both", digits: int = 12,
                          domain_metadata: Optional[Dict[str, object]] = None) -> List[str]:
    """One file per quantity and component, ``shg_<chi|sigma>_<abc>.dat``, rows of energy, total and/or
    the seven terms (``write_second_harmonic``, ``src/IO/SecondHarmonicWriters.jl:8-70``, integral rows).
    ``values`` is [row, component, term, quantity]; components are 1-based (a, b, c)."""
    values = np.asarray(values, dtype=np.complex128)
    coordinates = np.asarray(coordinates, dtype=np.float64).reshape(values.shape[0], -1)
    paths = []
    os.makedirs(directory, exist_ok=True)
    for quantity, name, unit in ((0, "chi", "pm/V"), (1, "sigma", "A/V^2")):
        if response == "susceptibility" and quantity == 1:
            continue
        if response == "conductivity" and quantity == 0:
            continue
        for index, comp in enumerate(components):
            label = "".join("xyz"[a - 1] for a in comp)
            path = os.path.join(directory, f"shg_{name}_{label}.dat")
            with open(path, "w") as fh:
                fh.write(f"# SHG {name} units={unit} component={label}\n")
                fh.write("# full-BZ normalized susceptibility/conductivity\n")
                for key, value in (domain_metadata or {}).items():
                    fh.write(f"# {key}={value}\n")
                labels = ["energy_eV"]
                if output != "terms":
                    labels += ["total_re", "total_im"]
                if output != "total":
                    labels += [t + s for t in terms for s in ("_re", "_im")]
                fh.write("# " + " ".join(labels) + "\n")
                for row in range(values.shape[0]):
                    nums = list(coordinates[row])
                    if output != "terms":
                        total = values[row, index, :, quantity].sum()
                        nums += [total.real, total.imag]
                    if output != "total":
                        for term in range(len(terms)):
                            v = values[row, index, term, quantity]
                            nums += [v.real, v.imag]
                    fh.write(" ".join(_fmt(x, digits) for x in nums) + "\n")
            paths.append(path)
    return paths


def full_zone_metadata(k_mesh: Sequence[int]) -> Dict[str, object]:
    """Comment lines the SHG files carry for a full-zone mesh of a periodic model
    (``integration_domain_metadata``, ``src/Runtime/NativeModelRuntime.jl:535-557``)."""
    return {"integration_bounds": "nothing", "integration_dimension": len(k_mesh), "model_periodic": "true",
            "integration_measure": "fractional full-BZ average",
            "boundary_corrections": "not included for finite-domain formulas derived under periodic-BZ assumptions"}


@dataclass
class RunTask:
    task_id: str
    quantity: str
    numerics: Dict[str, object] = field(default_factory=dict)
    selectors: Dict[str, str] = field(default_factory=dict)   # SHG output / response


@dataclass
class RunConfig:
    model_file: str
    system_name: str
    output_root: str
    photon_energies: np.ndarray
    fermi_energy: float
    temperature: float
    k_mesh: tuple
    spatial_dimension: int
    tasks: List[RunTask]
    digits: Optional[int] = None
    fourier_backend: str = "auto"


def _check_keys(table: dict, allowed: set, path: str):
    extra = set(table) - allowed
    if extra:
        raise ValueError(f"{path}: unsupported keys {sorted(extra)} (this runner covers dense-mesh SC / IC / SHG "
                         "integrals of Wannier models)")


def load_run_config(path: str) -> RunConfig:
    """Parse the supported subset of a reference TOML run file (``src/Runtime/TaskInput.jl``)."""
    try:
        import tomllib as toml   # Python >= 3.11
    except ImportError:  # pragma: no cover - Python 3.10
        import tomli as toml
    with open(path, "rb") as fh:
        doc = toml.load(fh)
    base = os.path.dirname(os.path.abspath(path))
    _check_keys(doc, {"model", "response", "output"}, "run file")
    model = doc.get("model", {})
    _check_keys(model, {"source", "model_file"}, "model")
    if model.get("source") != "wannier" or "model_file" not in model:
        raise ValueError('model: source = "wannier" with a model_file is required')
    model_file = os.path.normpath(os.path.join(base, model["model_file"]))
    resp = doc.get("response", {})
    _check_keys(resp, {"photon_energies_ev", "photon_energy_min_ev", "photon_energy_max_ev", "photon_energy_step_ev",
                       "photon_energy_num_points", "fermi_energy_ev", "temperature_k", "sampling", "execution",
                       "tasks"}, "response")
    if "photon_energies_ev" in resp:
        if any(k.startswith("photon_energy_") for k in resp):
            raise ValueError("response: give photon_energies_ev or a generated range, not both")
        omega = np.asarray(resp["photon_energies_ev"], dtype=np.float64)
    else:
        # a generated axis names both endpoints and exactly one of step / count (_task_input_expand_axes!)
        if ("photon_energy_min_ev" not in resp or "photon_energy_max_ev" not in resp
                or ("photon_energy_step_ev" in resp) == ("photon_energy_num_points" in resp)):
            raise ValueError("response: a generated photon energy axis needs photon_energy_min_ev, "
                             "photon_energy_max_ev and exactly one of photon_energy_step_ev / "
                             "photon_energy_num_points")
        omega = energy_range(resp["photon_energy_min_ev"], resp["photon_energy_max_ev"],
                             resp.get("photon_energy_step_ev"), resp.get("photon_energy_num_points"))
    samp = resp.get("sampling", {})
    _check_keys(samp, {"kind", "k_mesh", "spatial_dimension"}, "response.sampling")
    if samp.get("kind") != "bz_mesh" or len(samp.get("k_mesh", [])) != 3:
        raise ValueError('response.sampling: kind = "bz_mesh" with a three-entry k_mesh is required')
    dim = int(samp.get("spatial_dimension", 2))
    if dim != 3:
        raise ValueError("response.sampling.spatial_dimension must be 3 for this runner (the reference's default is 2)")
    _check_keys(resp.get("execution", {}), {"fourier_backend"}, "response.execution")
    # normalized_fourier_backend (FourierExecutionPlan.jl:63-67) of the key canonicalized as _canonical_key
    # (Config.jl:437-442): lower case, runs of blanks and hyphens to one underscore, outer underscores removed
    backend = re.sub(r"_+", "_", re.sub(r"[\s\-]+", "_", str(resp.get("execution", {}).get("fourier_backend", "auto"))
                                       .strip().lower())).strip("_")
    if backend not in ("auto", "direct", "mixed"):
        raise ValueError(f"fourier_backend={backend!r} is invalid; use direct, mixed, or auto.")
    tasks = []
    for tid, tab in sorted(resp.get("tasks", {}).items()):
        _check_keys(tab, {"kind", "method", "numerics", "output", "response"}, f"response.tasks.{tid}")
        kind = str(tab.get("kind", "")).lower()
        if kind not in TASK_KINDS:
            raise ValueError(f"response.tasks.{tid}: unsupported kind {tab.get('kind')!r}")
        # SHG selectors (_task_input_explicit_selectors, src/Runtime/TaskInputParameters.jl:50-68)
        selectors = {}
        for key, choices in (("output", ("total", "terms", "both")),
                             ("response", ("conductivity", "susceptibility", "both"))):
            if key in tab:
                if TASK_KINDS[kind] != "second_harmonic_generation":
                    raise ValueError(f"response.tasks.{tid}.{key} is not used by this observation.")
                if tab[key] not in choices:
                    raise ValueError(f"response.tasks.{tid}.{key} has an invalid explicit selector.")
                selectors[key] = tab[key]
        if str(tab.get("method", "Conventional")).lower() != "conventional":
            raise ValueError(f"response.tasks.{tid}: only method = \"Conventional\" is supported")
        num = dict(tab.get("numerics", {}))
        _check_keys(num, NUMERICS_KEYS, f"response.tasks.{tid}.numerics")
        tasks.append(RunTask(tid, TASK_KINDS[kind], num, selectors))
    if not tasks:
        raise ValueError("response.tasks: at least one task is required")
    out = doc.get("output", {})
    _check_keys(out, {"output_root", "system_name", "response_output_digits", "progress_enabled"}, "output")
    stem = os.path.splitext(os.path.basename(model_file))[0]
    stem = re.sub(r"[^\w.-]", "_", re.sub(r"_tb$", "", stem))
    digits = out.get("response_output_digits")
    if digits is not None and not 7 <= int(digits) <= 17:
        raise ValueError("output.response_output_digits must lie in 7:17")
    temperature = float(resp.get("temperature_k", 0.0))
    if not (math.isfinite(temperature) and temperature >= 0):
        raise ValueError("response.temperature_k must be finite and nonnegative")
    if any(int(v) < 1 for v in samp["k_mesh"]):
        raise ValueError("response.sampling.k_mesh entries must be positive")
    return RunConfig(model_file=model_file, system_name=str(out.get("system_name", stem)),
                     output_root=os.path.normpath(os.path.join(base, out.get("output_root", "") or ".")),
                     photon_energies=omega, fermi_energy=float(resp.get("fermi_energy_ev", 0.0)),
                     temperature=temperature, k_mesh=tuple(int(v) for v in samp["k_mesh"]), spatial_dimension=dim, tasks=tasks,
                     digits=None if digits is None else int(digits), fourier_backend=backend)


def run(path: str) -> List[str]:
    """Run every task of the run file on the GPU engine and write its result files; returns the paths."""
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import read_wannier_tb

    cfg = load_run_config(path)
    model = read_wannier_tb(cfg.model_file)
    written: List[str] = []
    # SC and IC with equal numerics share one engine (one interpolation and eigensolve)
    groups: Dict[tuple, List[RunTask]] = {}
    for t in cfg.tasks:
        key = ("shg" if t.quantity == "second_harmonic_generation" else "conv", tuple(sorted(t.numerics.items())))
        groups.setdefault(key, []).append(t)
    for (family, _), tasks in groups.items():
        # quantity defaults of resolved_numerical_values (src/Runtime/TaskConfiguration.jl:511-567)
        num = dict(SHG_DEFAULTS if family == "shg" else OPTICAL_DEFAULTS)
        num.update(tasks[0].numerics)
        common = dict(photon_energies=cfg.photon_energies, fermi_energy=cfg.fermi_energy, temperature=cfg.temperature,
                      broadening=float(num["broadening"]), broadening_type=str(num["broadening_type"]),
                      degeneracy_threshold=float(num["degeneracy_threshold"]))
        if family == "shg":
            task = E.EngineTask(second_harmonic=True, fourier_backend=cfg.fourier_backend,
                                shg_low_frequency_broadening=float(num["low_frequency_broadening"]),
                                shg_intermediate_regularization=float(num["intermediate_regularization"]),
                                shg_eta_correction=bool(num["eta_correction"]), **common)
        else:
            quantities = {t.quantity for t in tasks}
            task = E.EngineTask(shift_current="shift_current" in quantities, fourier_backend=cfg.fourier_backend,
                                injection_current="injection_current" in quantities,
                                denominator_regularization=float(num["denominator_regularization"]),
                                transition_window_factor=float(num["transition_window_factor"]),
                                band_window_size=int(num["band_window_size"]), **common)
        with E.Engine(model, task) as eng:
            _, total, _ = eng.run_mesh(cfg.k_mesh)
            if family == "shg":
                comps = [tuple(a + 1 for a in c) for c in eng.shg_components]
                for t in tasks:
                    written += write_second_harmonic(os.path.join(cfg.output_root, t.task_id),
                                                     cfg.photon_energies[:, None], total["shg"], comps,
                                                     output=t.selectors.get("output", "total"),
                                                     response=t.selectors.get("response", "both"),
                                                     digits=cfg.digits or 12,
                                                     domain_metadata=full_zone_metadata(cfg.k_mesh))
                continue
            for t in tasks:
                key = "sc" if t.quantity == "shift_current" else "ic"
                written.append(write_response_tensor(cfg.fermi_energy, cfg.photon_energies, total[key],
                                                     cfg.spatial_dimension,
                                                     os.path.join(cfg.output_root, t.task_id),
                                                     result_filename(cfg.system_name, t.quantity),
                                                     digits=cfg.digits or 7))
    return written


def main(argv: Optional[Sequence[str]] = None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if len(argv) != 2 or argv[0] != "run":
        print("usage: python -m wanniernlqg_amd.runio run <cfg>.toml", file=sys.stderr)
        return 2
    for p in run(argv[1]):
        print(p)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
