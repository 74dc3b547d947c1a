"""Benchmark: k-points/s of the dense-mesh shift + injection current k-loop on one MI355X.

Workload (the configuration the headline metric is quoted on): synthetic orthogonal TB model with
n_w = 32 orbitals, R cube |R_i| <= 5 (1331 R vectors), position operator amplitude 0.1 Angstrom,
a = 4 Angstrom cubic cell; conventional shift current + injection current, all 27 abc components,
400 photon energies in [0, 6] eV, Gaussian width 0.01 eV, mu = 0, T = 0, on the 128^3 Gamma-centred
mesh (2,097,152 k points).  One step = one full-mesh integral (all stages on the device, model and
R-space channels already resident).  The headline runs the mixed form on the separable mesh path (FP64
plane contraction with the channels after H stored in FP32; FP64 eigensolver, rotation, vertices and
compensated k-sum) and, beside it in the same run, the FP64 path on the same mesh: its speed and the
headline's relative spectrum error per component (the metric's bar: 1e-6).  The plain FP64 k x R
contraction and the split-FP16 / split-BF16 GEMM forms run beside them; the CPU baseline times the NumPy
FP64 oracle on a bounded k sample.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64] [--backend auto|direct|planes]
  python bench.py --form geometry|shg|batch|learned|large|symmetry  (the other forms, one JSON line each)
  python bench.py --form copy                            (HBM copy used to calibrate the PMC byte counters)
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """Cores of the host: this process's CPU share (its affinity set), nproc, and physical cores from lscpu."""
    info = {"affinity": len(os.sched_getaffinity(0)), "nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        fields = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        cores = int(fields.get("Core(s) per socket", "0").strip()) * int(fields.get("Socket(s)", "0").strip())
        info["physical_cores"] = cores or None
        info["model"] = fields.get("Model name", "").strip()
    except Exception:
        info["physical_cores"] = None
    return info


def cpu_baseline(model, omega, width, planes: int):
    """Time the C++/OpenMP restatement of the reference's CPU k-loop (oracle/kloop_omp.cpp: plane-by-plane
    Fourier sums as the reference's Mixed-FFT mesh backend, Householder + QL eigensystem, full rotation of every
    channel, the SC/IC pair kernels; a port, not the reference: Julia is absent) on whole k3 planes of the same
    128^3 mesh with one OpenMP thread per core of this process's CPU share; the k-loop's own wall time (the
    R-space set-up excluded, like the reference's k-loop timer).  Returns k-points/s, threads, seconds, points."""
    from oracle import kloop, kloop_omp
    threads = len(os.sched_getaffinity(0))
    task = kloop.Task(photon_energies=omega, broadening=width, shift_current=True, injection_current=True)
    tm = {}
    p0 = 5
    kloop_omp.integrate_planes(model, task, (128, 128, 128), p0, p0 + planes, threads=threads, timing=tm)
    points = planes * 128 * 128
    return points / tm["loop_s"], threads, tm["loop_s"], points


def spectrum_relative_error(x, ref, floor=1.0e-12):
    """Per component ||x - ref||_2 / ||ref||_2 over the omega axis (absolute for components whose norm is below
    floor times the largest component norm)."""
    x = np.asarray(x).reshape(x.shape[0], -1)
    ref = np.asarray(ref).reshape(ref.shape[0], -1)
    num = np.linalg.norm(x - ref, axis=0)
    den = np.linalg.norm(ref, axis=0)
    big = den >= floor * max(den.max(), 1e-300)
    return np.where(big, num / np.where(big, den, 1.0), num)


def jacobi_flops_per_k(n: int, sweeps: float) -> float:
    """Algorithmic flops of cyclic Jacobi on an n x n Hermitian matrix, 16 n^3 per sweep: per rotation
    (n (n - 1)/2 of them) one Hermitian row/column pair of H (the other triangle is its conjugate) and one
    column pair of V, n complex 2x2 updates each at about 8 flops per updated element."""
    return sweeps * 16.0 * n ** 3


def rotate_sc_bytes(st, n: int, precision: str) -> float:
    """Algorithmic HBM bytes of the SC/IC rotation over a mesh.  Every point with pairs reads its C - 1 = 21
    channels after H (n^2 complex each: 16 bytes per entry in FP64, 8 in the mixed form's FP32 storage) and
    its eigenvectors U (16 n^2).  Writes: points rotated in full write 21 channels (16 n^2 each); element
    points write the pairs' elements of dd H and dA (2 directions x 15 complex per pair) and dH, A (6
    channels) either in full (16 n^2) or, for Hermitian models, the rows at the pairs' bands B (|B| n entries
    each; the pair kernel reads the columns as their conjugates)."""
    pts = st["rotate_full_points"] + st["rotate_element_points"]
    ch = 8.0 if precision == "mixed" else 16.0
    rd = pts * (21 * n * n * ch + 16.0 * n * n)
    el, rows = st["rotate_element_points"], st["rotate_row_points"]
    mb = st["rotate_element_bands"] / max(1, el)
    wr = (st["rotate_full_points"] * 21 * 16.0 * n * n + st["active_pairs"] * 2 * 15 * 16.0
          + (el - rows) * 6 * 16.0 * n * n + rows * 6 * 16.0 * mb * n)
    return rd + wr


def rotate_sc_flops(st, n: int) -> float:
    """Algorithmic flops of the SC/IC rotation over a mesh (8 flops per complex multiply-add): points rotated
    in full transform 21 channels (16 n^3 each); element points form U_B^+ X (8 |B| n^2) for the 15 dd H and
    dA channels, |B| the bands of the point's pairs, and transform dH and A (6 channels) in full (16 n^3) or,
    for Hermitian models, as the rows U_B^+ X U (16 |B| n^2); dot products for the pairs' elements are not
    counted; points without pairs are skipped."""
    el, rows = st["rotate_element_points"], st["rotate_row_points"]
    mb = st["rotate_element_bands"] / max(1, el)
    return (st["rotate_full_points"] * 21 * 16.0 * n ** 3 + el * 15 * 8.0 * mb * n ** 2
            + (el - rows) * 6 * 16.0 * n ** 3 + rows * 6 * 16.0 * mb * n ** 2)


def pairs_flops_per_pair(n: int) -> float:
    """Algorithmic flops of one active band pair in the SC/IC pair kernel: six generalized-derivative rows
    (two directions x three components), each an intermediate-state sum over n - 2 bands of, for the three
    derivative directions, two scaled dipoles (4 flops) and two complex multiply-add/subtract pairs for the
    A and dH commutator terms (32 flops), plus ~200 flops of closing terms per row; then the 27 shift
    kernels (two complex products and a difference, 14 flops) and 27 injection kernels (12 flops)."""
    return 6.0 * (108.0 * (n - 2) + 200.0) + 27 * 14.0 + 27 * 12.0


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "kloop_n32_64cube_hbm_traffic.md")


def measured_traffic_per_k(precision="fp64"):
    """HBM bytes per k point of each k-loop kernel of the given precision form, from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the reading of a plain copy of known size); {} when absent."""
    tag = "traffic_bytes_per_k_%s:" % ("mixed" if precision == "mixed" else "fp64")
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith(tag):
                    return {kv.split("=")[0]: float(kv.split("=")[1]) for kv in line.split(":", 1)[1].split()}
    except OSError:
        pass
    return {}


def roofline(st, n, nk, precision):
    """Roofline of the single kernel with the largest device time among those the run times with HIP events on
    their stream (k_rotate_sc, k_contract, or the interpolation GEMM): algorithmic work per launch over the
    average launch duration (in a two-stream step the launches share the GPU with the other stream's).  The
    eigensolver stage (two kernels, H sweeps and the V kernel, plus the warm-start transform) is reported
    beside it in 'jacobi_stage'."""
    nb = max(1, st["batches"])
    kern = {"rotate": st["rotate_ms"], "interp": st["interp_ms"]}
    dominant = max(kern, key=kern.get)
    traffic = measured_traffic_per_k(precision)
    if dominant == "rotate":
        launches = nb
        avg_s = 1e-3 * st["rotate_ms"] / launches
        byts = rotate_sc_bytes(st, n, precision) / launches
        roof = {"kernel": "k_rotate_sc (reads the 21 channels after H and U per point with pairs; dH and A as rows "
                          "at the pairs' bands for Hermitian models, dd H and dA reduced to the pairs' elements; FP64 "
                          "MFMA products)",
                "bound": "hbm", "achieved": byts / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "mfma_tflops": rotate_sc_flops(st, n) / launches / avg_s / 1e12}
        tkey = "k_rotate_sc"
    elif st["separable"]:
        launches = 3 * nb
        avg_s = 1e-3 * st["interp_ms"] / launches
        roof = {"kernel": "k_contract (separable interpolation in three stages: the coefficient box read once per "
                          "plane, every interpolated matrix written, the channels after H in FP32 in the mixed form; "
                          "the partial sums between stages not counted)", "bound": "hbm",
                "achieved": st["interp_bytes"] / launches / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        tkey = "k_contract"
    else:
        launches = nb
        avg_s = 1e-3 * st["interp_ms"] / launches
        mult = 1.0 if precision == "fp64" else 3.0
        roof = {"kernel": "k_interp (k x R contraction)", "bound": "mfma",
                "achieved": mult * st["interp_flops"] / launches / avg_s / 1e12,
                "peak": FP64_PEAK_TFLOPS if precision == "fp64" else BF16_PEAK_TFLOPS, "unit": "TFLOP/s"}
        tkey = "k_interp"
    roof["frac"] = roof["achieved"] / roof["peak"] if roof["achieved"] is not None else None
    roof["traffic"] = traffic[tkey] * nk / launches if tkey in traffic else None
    if roof["traffic"] is not None:
        roof["traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, ROOT)
    roof["avg_launch_ms"] = 1e3 * avg_s
    roof["launches"] = launches
    # the eigensolver stage: sweeps * 16 n^3 flops per point over its event-timed span (both kernels)
    jl = max(1, st["eig_launches"])
    roof["jacobi_stage"] = {
        "kernels": "k_eigd_sweeps + k_eigd_vapply (deferred-eigenvector Jacobi, FP64 VALU)" if n <= 32 else
                   "k_eig_f64 (Jacobi, FP64 VALU)",
        "ms": st["eig_ms"], "launches": jl, "mean_sweeps": st["mean_sweeps"],
        "tflops": jacobi_flops_per_k(n, st["mean_sweeps"]) * nk / (1e-3 * st["eig_ms"]) / 1e12,
        "frac_of_fp64_peak": jacobi_flops_per_k(n, st["mean_sweeps"]) * nk / (1e-3 * st["eig_ms"]) / 1e12
                             / FP64_PEAK_TFLOPS}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64", "bf16x3"])
    ap.add_argument("--backend", default="auto", choices=["auto", "direct", "planes"])
    ap.add_argument("--mesh", type=int, default=128)
    ap.add_argument("--no-compare", action="store_true", help="skip the plain-FP64 and mixed direct comparison runs")
    ap.add_argument("--cpu-planes", type=int, default=1, help="k3 planes of the 128^3 mesh timed on the CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-meshes", default="16,20,24",
                    help="learned form: coarse meshes refined everywhere by 5 for the network's training targets")
    ap.add_argument("--form", default="headline", choices=["headline", "geometry", "shg", "batch", "learned", "large", "symmetry", "copy",
                                                         "fft"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the k-mesh would shard by "
              "deterministic lanes across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)
    if args.form != "headline":
        FORMS[args.form](args)
        return

    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model

    model = config_model(1)
    omega = np.linspace(0.0, 6.0, 400)
    width = 0.01
    mesh = (args.mesh,) * 3
    nk = int(np.prod(mesh))
    n = model.num_orbitals

    def make(prec, backend):
        return E.Engine(model, E.EngineTask(photon_energies=omega, broadening=width, shift_current=True,
                                            injection_current=True, precision=prec, fourier_backend=backend))

    log(f"[bench] model n={n} N_R={model.num_r_vectors} mesh={mesh} precision={args.precision} backend={args.backend}")
    t0 = time.perf_counter()
    eng = make(args.precision, args.backend)
    log(f"[bench] engine created in {time.perf_counter() - t0:.2f} s (model upload + R-space channels)")
    for w in range(args.warmup):
        _, tot, st = eng.run_mesh(mesh)
        log(f"[bench] warmup {w}: {st['total_ms']:.1f} ms")
    eng.synchronize()
    t0 = time.perf_counter()
    stats = []
    for s in range(args.steps):
        _, tot, st = eng.run_mesh(mesh)
        stats.append(st)
        log(f"[bench] step {s}: device {st['total_ms']:.1f} ms (interp {st['interp_ms']:.1f}, eig {st['eig_ms']:.1f}, "
            f"warm {st['warm_ms']:.1f}, rot {st['rotate_ms']:.1f}, pairs {st['pairs_ms']:.1f}, "
            f"accum {st['accum_ms']:.1f}; separable={st['separable']}, sweeps mean {st['mean_sweeps']:.2f} "
            f"max {st['max_sweeps']})")
    eng.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = nk * args.steps / wall
    head_total = tot
    # after the timed region: one step in one stream (WNLQG_MESH_OVERLAP=0), each kernel timed alone
    os.environ["WNLQG_MESH_OVERLAP"] = "0"
    try:
        _, _, st_alone = eng.run_mesh(mesh)
    finally:
        os.environ.pop("WNLQG_MESH_OVERLAP", None)
    eng.close()

    # roofline of the dominant kernel, from its device event timings inside run_mesh: in the timed steps
    # (two streams, the kernels of batch b - 1's response sharing the GPU with batch b's contraction and
    # eigensystem) and in the one-stream step after them
    st = stats[-1]
    roof = roofline(st, n, nk, args.precision)
    roof_alone = roofline(st_alone, n, nk, args.precision)
    roof["one_stream"] = {k: roof_alone[k] for k in ("kernel", "achieved", "frac", "avg_launch_ms", "launches",
                                                     "traffic")}
    roof["one_stream"]["jacobi_stage"] = roof_alone["jacobi_stage"]

    # the FP64 GPU path on the same mesh: the reference of the metric's "mixed vs FP64 at rel err <= 1e-6"
    if args.precision != "fp64":
        ref = make("fp64", args.backend)
        ref.run_mesh(mesh)
        _, fp64_total, fst = ref.run_mesh(mesh)
        ref.close()
        head_err = {k: spectrum_relative_error(head_total[k], fp64_total[k]) for k in head_total}
        fp64 = {"kpoints_per_s": nk / (1e-3 * fst["total_ms"]), "device_ms": fst["total_ms"],
                "stages_ms": {k: fst[k] for k in ("interp_ms", "eig_ms", "warm_ms", "rotate_ms", "pairs_ms",
                                                  "accum_ms")},
                "headline_speedup": fst["total_ms"] / st["total_ms"],
                "headline_rel_spectrum_err": {k: float(v.max()) for k, v in head_err.items()},
                "headline_rel_spectrum_err_per_component": {k: [float(x) for x in v] for k, v in head_err.items()}}
        log(f"[bench] fp64 separable: {fst['total_ms']:.1f} ms; headline error {fp64['headline_rel_spectrum_err']}")
        if max(fp64["headline_rel_spectrum_err"].values()) > 1e-6:
            log("[bench] WARNING: the headline exceeds 1e-6 relative spectrum error against FP64")
    else:
        fp64_total, fp64 = head_total, None

    compare = {}
    if not args.no_compare:
        # the k x R contraction forms beside the headline: the plain FP64 GPU path and the split-FP16 /
        # split-BF16 mixed GEMMs, errors against the FP64 separable mesh
        for label, prec, backend in (("plain_fp64_direct", "fp64", "direct"), ("mixed_f16_direct", "mixed", "direct"),
                                     ("split_bf16x3_direct", "bf16x3", "direct")):
            ref = make(prec, backend)
            _, rtot, rst = ref.run_mesh(mesh)
            ref.close()
            errs = {k: float(spectrum_relative_error(rtot[k], fp64_total[k]).max()) for k in rtot}
            compare[label] = {"kpoints_per_s": nk / (1e-3 * rst["total_ms"]), "device_ms": rst["total_ms"],
                              "interp_ms": rst["interp_ms"], "max_rel_spectrum_err_vs_fp64": errs}
            log(f"[bench] {label}: {compare[label]}")
        compare["headline_speedup_vs_plain_fp64"] = compare["plain_fp64_direct"]["device_ms"] / st["total_ms"]

    cpu = None
    if not args.no_cpu:
        kps, threads, dt, points = cpu_baseline(model, omega, width, args.cpu_planes)
        cpu = {"value": kps, "unit": "k-points/s", "cores": threads, "kind": "port",
               "sample": f"{points} k points (k3 planes 5..{4 + args.cpu_planes} of the 128^3 mesh), same model and task; "
                         f"restated oracle, not the reference: C++/OpenMP FP64 (oracle/kloop_omp.cpp), k-loop "
                         f"{dt:.1f} s on {threads} threads",
               "host": host_cpu()}
        log(f"[bench] cpu baseline {kps:.1f} k/s on {threads} threads")

    out = {
        "metric": "k-points/s (n_w=32,128\u00b3 shift current) mixed vs FP64 GPU at rel err \u22641e-6",
        "value": value,
        "unit": "k-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f64 (interpolated channels after H stored f32)",
                  "bf16x3": "bf16x2-split-mfma+f64"}[args.precision],
        "data": "synthetic (seeded TB model, no external data)",
        "config": {"workload": f"shift+injection current, n_w=32, N_R=1331, {args.mesh}^3 Gamma-centred mesh, "
                               f"400 omega in [0,6] eV, Gaussian 0.01 eV, 27 components",
                   "kpoints": nk, "n_w": 32, "num_r_vectors": model.num_r_vectors, "n_omega": 400,
                   "precision": args.precision, "fourier_backend": args.backend,
                   "separable": bool(st["separable"])},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": {k: st[k] for k in ("interp_ms", "eig_ms", "warm_ms", "rotate_ms", "pairs_ms", "accum_ms",
                                         "total_ms")},
        "stage_note": "timed steps: two streams (batch b - 1's screening, rotation, pairs, frequency contraction "
                      "and lane sums beside batch b's contraction, warm transform and Jacobi); each figure is the "
                      "event-timed span of its kernels on their stream, so spans share GPU time and their sum "
                      "exceeds total_ms; stages_one_stream_ms: the step after the timed region on one stream, "
                      "each kernel alone",
        "stages_one_stream_ms": {k: st_alone[k] for k in ("interp_ms", "eig_ms", "warm_ms", "rotate_ms", "pairs_ms",
                                                          "accum_ms", "total_ms")},
        "rotation_points": {"elements": st["rotate_element_points"], "full": st["rotate_full_points"],
                            "rows": st["rotate_row_points"],
                            "mean_bands": st["rotate_element_bands"] / max(1, st["rotate_element_points"])},
        "active_pairs_per_k": st["active_pairs"] / nk,
    }
    if fp64 is not None:
        out["fp64_gpu"] = fp64
    if compare:
        out["compare"] = compare
    print(json.dumps(out), flush=True)


def _timed_runs(eng, mesh, steps, warmup):
    for _ in range(warmup):
        eng.run_mesh(mesh)
    eng.synchronize()
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        _, tot, st = eng.run_mesh(mesh)
    eng.synchronize()
    return (time.perf_counter() - t0) / steps, tot, st


def form_geometry(args):
    """BCD (Fermi surface) + intrinsic AHC in one engine on the n_w = 32 model, 3 chemical potentials."""
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model
    model = config_model(1)
    mesh = (args.mesh,) * 3 if args.mesh != 128 else (96, 96, 96)
    task = E.EngineTask(berry_curvature_dipole=True, intrinsic_hall=True, fermi_energies=[-0.5, 0.0, 0.5],
                        temperature=0.0, fermi_surface_broadening=0.02, denominator_regularization=0.0,
                        degeneracy_threshold=0.0)
    with E.Engine(model, task) as eng:
        wall, _, st = _timed_runs(eng, mesh, args.steps, args.warmup)
    nk = int(np.prod(mesh))
    out = {"metric": "k-points/s (BCD + intrinsic AHC, n_w=32)", "value": nk / wall, "unit": "k-points/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded TB model)",
           "config": {"workload": f"BCD (Fermi surface, eta_fs 0.02 eV) + AHC at mu = -0.5, 0, 0.5 eV, T = 0, "
                                  f"n_w=32, N_R=1331, {mesh[0]}^3 mesh", "kpoints": nk},
           "stages_ms": {k: st[k] for k in ("interp_ms", "eig_ms", "rotate_ms", "pairs_ms", "accum_ms", "total_ms")}}
    print(json.dumps(out), flush=True)


def form_shg(args):
    """SHG chi/sigma(-2w; w, w), 3 components, 400 w in [0, 3] eV, on the n_w = 32 model."""
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model
    model = config_model(1)
    mesh = (args.mesh,) * 3 if args.mesh != 128 else (48, 48, 48)
    task = E.EngineTask(second_harmonic=True, photon_energies=np.linspace(0.0, 3.0, 400), broadening=0.02,
                        shg_components=[(0, 1, 1), (1, 0, 2), (2, 2, 2)], fermi_energy=0.0, temperature=0.0,
                        shg_intermediate_regularization=1e-3)
    with E.Engine(model, task) as eng:
        wall, _, st = _timed_runs(eng, mesh, args.steps, args.warmup)
    nk = int(np.prod(mesh))
    out = {"metric": "k-points/s (SHG, 3 components, 400 omega, n_w=32)", "value": nk / wall, "unit": "k-points/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded TB model)",
           "config": {"workload": f"SHG xyy, yxz, zzz; 400 omega in [0,3] eV; Gaussian 0.02 eV; n_w=32, N_R=1331; "
                                  f"{mesh[0]}^3 mesh", "kpoints": nk},
           "stages_ms": {k: st[k] for k in ("interp_ms", "eig_ms", "rotate_ms", "pairs_ms", "accum_ms", "total_ms")}}
    print(json.dumps(out), flush=True)


def form_batch(args):
    """Batching form: 64 models (n_w = 4 (s mod 12 + 1), |R_i| <= 4, seeds 0..63) with meshes drawn from
    {16,32,48,64,96}^3, BCD + AHC at 3 chemical potentials, one captured graph, against the same
    integrals run one engine after another."""
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model
    rng = np.random.default_rng(2)
    sizes = [16, 32, 48, 64, 96]
    entries = []
    for seed in range(64):
        m = config_model(2, seed)
        n = sizes[int(rng.integers(0, len(sizes)))]
        task = E.EngineTask(berry_curvature_dipole=True, intrinsic_hall=True, fermi_energies=[-0.5, 0.0, 0.5],
                            temperature=0.0, fermi_surface_broadening=0.02, denominator_regularization=0.0,
                            degeneracy_threshold=0.0)
        entries.append((m, task, (n, n, n)))
    total_k = sum(int(np.prod(e[2])) for e in entries)
    t0 = time.perf_counter()
    b = E.Batch(entries, autotune=True)
    setup = time.perf_counter() - t0
    b.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, dev_ms = b.run()
    wall = (time.perf_counter() - t0) / args.steps
    tuning = b.tuning()
    b.close()
    # the same integrals one engine at a time (plain path)
    t0 = time.perf_counter()
    seq_dev = 0.0
    for m, task, mesh in entries:
        with E.Engine(m, task) as eng:
            _, _, st = eng.run_mesh(mesh)
            seq_dev += st["total_ms"]
    seq_wall = time.perf_counter() - t0
    out = {"metric": "k-points/s (batching form: 64 models, BCD + AHC)", "value": total_k / wall,
           "unit": "k-points/s", "n_gpus": 1, "steps": args.steps, "ms_per_step": 1e3 * wall, "higher_is_better": True,
           "dtype": "f64", "data": "synthetic (seeded TB models)",
           "config": {"workload": "64 models n_w in {4..48}, |R_i|<=4, meshes {16..96}^3, BCD + AHC at 3 mu",
                      "kpoints": total_k},
           "graph_device_ms": dev_ms, "setup_s": setup, "eig_threads_autotuned": sorted(set(tuning)),
           "sequential_engines": {"wall_s_including_setup": seq_wall, "device_ms": seq_dev,
                                  "kpoints_per_s_device": total_k / (1e-3 * seq_dev)}}
    print(json.dumps(out), flush=True)


def form_learned(args):
    """Learned refinement of the BCD on the hot-spot model (n_w = 16, |R_i| <= 4, four Weyl-like pairs with
    1-5 meV gaps, mu = 0, T = 50 K): the network, trained on coarse runs refined everywhere by 5
    (--train-meshes, default 16^3, 20^3, 24^3), ranks the 64^3 base cells; top cells refined by 15 per axis (sub-cell sums on the device).  Judge:
    the uniform 960^3 mesh (refining every cell), whose own convergence is stated against 1088^3.
    Comparators: the non-learned base-variance ranking, the reference's h-adaptive cubature and a
    uniform mesh, each at the learned estimate's evaluation count.  value: dense-mesh evaluations over
    the learned estimate's evaluations (training included) for a 1e-3 reproduction of the judge."""
    from wanniernlqg_amd import cubature
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd import refine
    from wanniernlqg_amd.model import config_model
    m = config_model(3)
    n, r, fine = 64, 15, 1088
    top = n ** 3 // 5                   # cells refined per ranking for the error curves
    tols = (1e-2, 1e-3, 1e-4)
    task = E.EngineTask(berry_curvature_dipole=True, fermi_energies=[0.0], temperature=50.0,
                        denominator_regularization=0.0, degeneracy_threshold=0.0)
    with E.Engine(m, task) as eng:
        lr = refine.LearnedRefinement(eng, seed=1)
        t0 = time.perf_counter()
        train_meshes = tuple(int(v) for v in args.train_meshes.split(","))
        hist = lr.train(meshes=train_meshes, r=5, steps=6000, batch=4096, lr=2e-3, seed=3)
        train_s, train_evals = time.perf_counter() - t0, lr.evaluations
        log(f"[learned] trained on {train_evals} evaluations in {train_s:.1f} s")
        base, ben = lr.point_data(refine.mesh_points(n))
        t0 = time.perf_counter()
        orders = {"learned": np.argsort(-lr.scores(n, base, ben), kind="stable"),
                  "base_variance": refine.variance_order(n, base)}
        rank_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, judge, _ = eng.run_mesh((n * r,) * 3)
        dense_s = time.perf_counter() - t0
        log(f"[learned] judge {n * r}^3 in {dense_s:.1f} s")
        _, finer, _ = eng.run_mesh((fine,) * 3)
        log(f"[learned] convergence mesh {fine}^3 done")
        ref = judge["bcd"]
        err = lambda v: float(np.linalg.norm(v - ref) / np.linalg.norm(ref))
        # refined means of the union of both rankings' top cells, on the device, in slices
        cells = np.unique(np.concatenate([o[:top] for o in orders.values()]))
        means = np.empty((cells.size,) + base.shape[1:])
        t0 = time.perf_counter()
        step = 16384
        for c0 in range(0, cells.size, step):
            means[c0:c0 + step] = lr.refined_cell_means(cells[c0:c0 + step], n, r)
            log(f"[learned] refined {min(c0 + step, cells.size)} of {cells.size} cells")
        refine_rate = cells.size * r ** 3 / (time.perf_counter() - t0)
        slot = np.full(n ** 3, -1)
        slot[cells] = np.arange(cells.size)
        total = n ** 3 * lr.volume * ref.reshape(-1)          # sum of every cell's refined mean
        all_delta = total - base.reshape(n ** 3, -1).sum(axis=0)
        rows = {}
        for name, o in orders.items():
            d = means[slot[o[:top]]].reshape(top, -1) - base[o[:top]].reshape(top, -1)
            rem = all_delta[None, :] - np.vstack([np.zeros((1, d.shape[1])), np.cumsum(d, axis=0)])
            curve = np.linalg.norm(rem, axis=1) / np.linalg.norm(total)
            need = [refine.cells_needed(curve, t) if curve[-1] < t else None for t in tols]
            rows[name] = {"cells_needed": need, "err_at_1pct": float(curve[n ** 3 // 100]),
                          "err_at_5pct": float(curve[n ** 3 // 20]), "err_at_20pct": float(curve[top])}
        evals = lambda c: None if c is None else n ** 3 + c * (r ** 3 - 1)
        dense_evals = (n * r) ** 3
        comparators = []
        for i, t in enumerate(tols):
            le = evals(rows["learned"]["cells_needed"][i])
            if le is None:
                continue
            t0 = time.perf_counter()
            cub = cubature.integrate_engine(eng, cubature.HAdaptiveCubature(atol=0.0, rtol=1e-4, max_evaluations=le),
                                            boxes_per_step=256)
            cub_s = time.perf_counter() - t0
            nu = int(np.ceil(le ** (1 / 3)))
            _, uni, _ = eng.run_mesh((nu,) * 3)
            comparators.append({"tol": t, "learned_evaluations": le, "hadaptive_evaluations": cub.evaluations,
                                "hadaptive_err": err(cub.value[0]), "hadaptive_s": cub_s,
                                "uniform_mesh": nu, "uniform_err": err(uni["bcd"])})
            log(f"[learned] tol {t}: cubature err {comparators[-1]['hadaptive_err']:.3g}, uniform {nu}^3 err "
                f"{comparators[-1]['uniform_err']:.3g}")
    for name in rows:
        rows[name]["evaluations"] = [evals(c) for c in rows[name]["cells_needed"]]
        rows[name]["evaluations_vs_dense"] = [None if e is None else dense_evals / e for e in rows[name]["evaluations"]]
    le3 = rows["learned"]["evaluations"][1]
    out = {"metric": "k evaluations for dense-mesh accuracy (learned adaptive BCD)",
           "value": None if le3 is None else dense_evals / (le3 + train_evals),
           "unit": "x fewer evaluations than the dense mesh (training included)", "higher_is_better": True,
           "n_gpus": 1, "dtype": "f64 integrand, f32 network", "data": "synthetic hot-spot model",
           "config": {"workload": "BCD mu=0 T=50K, n_w=16, |R_i|<=4, 4 near-degenerate pairs (1-5 meV); base 64^3, "
                                  "refinement 15 per axis, dense 960^3", "dense_kpoints": dense_evals},
           "accuracy_criterion": "relative L2 error vs the 960^3 result (every cell refined) <= 1e-3",
           "judge_convergence_vs_1088": err(finer["bcd"]), "training_meshes": list(train_meshes),
           "training_evaluations": train_evals,
           "train_loss_first_last": [float(hist[:100].mean()), float(hist[-100:].mean())], "train_s": train_s,
           "rank_s": rank_s, "dense_s": dense_s, "refine_points_per_s": refine_rate, "tolerances": list(tols),
           "rankings": rows, "comparators_at_learned_budget": comparators}
    print(json.dumps(out), flush=True)


def form_symmetry(args):
    """Symmetry-reduced mesh: the headline model (n_w = 32, |R_i| <= 5, A amplitude 0.1 Angstrom) averaged over
    C4v (8 operations, orbitals at the origin), shift + injection current, 400 omega in [0, 6] eV, Gaussian
    0.01 eV, on the --mesh^3 mesh: the full separable mesh against the orbit representatives (multiplicity
    weights; planes contracted, the last stage only for the representatives: wnlqg_run_mesh_reduced) followed by the
    tensor projection; the representatives interpolated alone (wnlqg_run_points) reported beside.  value: full-mesh k-points per second of
    the reduced integral (orbit planning on the host timed apart)."""
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd import symmetry as S
    from wanniernlqg_amd.model import config_model
    base = config_model(1)
    c4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
    mx = np.diag([-1, 1, 1])
    ops = S.group_from_generators([(c4, False), (mx, False)], np.asarray(base.lattice))
    model = S.symmetrize_model(base, ops)
    mesh = (args.mesh,) * 3
    nk = int(np.prod(mesh))
    task = E.EngineTask(photon_energies=np.linspace(0.0, 6.0, 400), broadening=0.01, shift_current=True,
                        injection_current=True)
    with E.Engine(model, task) as eng:
        wall_full, full, _ = _timed_runs(eng, mesh, args.steps, args.warmup)
        t0 = time.perf_counter()
        orb = S.kmesh_orbits(mesh, ops)
        kp = (S.mesh_indices(mesh)[:, orb.representatives] / np.array(mesh, dtype=float)[:, None]).T
        w = orb.multiplicities / float(orb.full_count)
        proj = {q: S.response_projector(q, ops) for q in ("sc", "ic")}
        plan_s = time.perf_counter() - t0
        for _ in range(args.warmup):
            eng.run_mesh_reduced(mesh, orb.representatives, w)
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, tot = eng.run_mesh_reduced(mesh, orb.representatives, w)
            red = {q: S.project(v, proj[q]) for q, v in tot.items()}
        eng.synchronize()
        wall_red = (time.perf_counter() - t0) / args.steps
        # the representatives' points interpolated alone (direct k x R products)
        t0 = time.perf_counter()
        _, tot_pts = eng.run_points(kp, w)
        eng.synchronize()
        wall_pts = time.perf_counter() - t0
        err_pts = max(float(np.linalg.norm(S.project(tot_pts[q], proj[q]) - full[q]) / np.linalg.norm(full[q]))
                      for q in ("sc", "ic"))
        cov = S.check_covariance(lambda kk: eng.eval_points(kk), ops, samples=2)
    err = max(float(np.linalg.norm(red[q] - full[q]) / np.linalg.norm(full[q])) for q in ("sc", "ic"))
    out = {"metric": "k-points/s (symmetry-reduced mesh: C4v, n_w=32, SC + IC, 400 omega)", "value": nk / wall_red,
           "unit": "full-mesh k-points/s", "higher_is_better": True, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall_red, "dtype": "f64", "data": "synthetic C4v model",
           "config": {"workload": "SC+IC, n_w=32, |R_i|<=5 averaged over C4v, 400 omega in [0,6] eV, Gaussian 0.01 eV",
                      "mesh": list(mesh)},
           "representatives": int(orb.representatives.shape[0]), "full_kpoints": nk,
           "full_mesh_kps": nk / wall_full, "full_mesh_ms": 1e3 * wall_full, "speedup_vs_full_mesh": wall_full / wall_red,
           "orbit_planning_s": plan_s, "rel_err_vs_full_mesh": err, "covariance_residual": cov,
           "points_path_ms": 1e3 * wall_pts, "points_path_rel_err": err_pts}
    print(json.dumps(out), flush=True)


def form_large(args):
    """n_w = 96 SHG (|R_i| <= 6, 2197 R vectors, A amplitude 0.1 Angstrom; xyy, yxz, zzz; 400 omega in [0, 3] eV,
    Gaussian 0.02 eV) on a --mesh^3 sub-mesh (default 16^3) of the 256^3 configuration.  The engine keeps
    the Jacobi's eigenvectors and the rotation's U^+ X in global memory at this size."""
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model
    model = config_model(4)
    mesh = (args.mesh,) * 3 if args.mesh != 128 else (16, 16, 16)
    task = E.EngineTask(second_harmonic=True, photon_energies=np.linspace(0.0, 3.0, 400), broadening=0.02,
                        shg_components=[(0, 1, 1), (1, 0, 2), (2, 2, 2)], fermi_energy=0.0, temperature=0.0)
    t0 = time.perf_counter()
    eng = E.Engine(model, task)
    setup = time.perf_counter() - t0
    wall, _, st = _timed_runs(eng, mesh, args.steps, args.warmup)
    eng.close()
    nk = int(np.prod(mesh))
    out = {"metric": "k-points/s (SHG, n_w=96, 3 components, 400 omega)", "value": nk / wall, "unit": "k-points/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded TB model)",
           "config": {"workload": f"SHG xyy, yxz, zzz; 400 omega in [0,3] eV; Gaussian 0.02 eV; n_w=96, N_R=2197; "
                                  f"{mesh[0]}^3 sub-mesh of the 256^3 configuration", "kpoints": nk},
           "setup_s": setup, "projected_256cube_s": 256 ** 3 * wall / nk, "mean_sweeps": st["mean_sweeps"],
           "stages_ms": {k: st[k] for k in ("interp_ms", "eig_ms", "rotate_ms", "pairs_ms", "accum_ms", "total_ms")}}
    print(json.dumps(out), flush=True)


def form_copy(args):
    """HBM calibration: a plain device-to-device copy of 256 MiB (32 Mi doubles read, 32 Mi written), five
    launches.  Its FETCH_SIZE / WRITE_SIZE counter readings against the known 256 MiB per launch give the
    scale factors applied to the same counters of the k-loop kernels."""
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("the copy calibration runs on the GPU")
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    nbytes = x.numel() * 8
    out = {"metric": "HBM copy bandwidth (calibration)", "value": 2 * nbytes / dt / 1e9, "unit": "GB/s",
           "n_gpus": 1, "steps": 5, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "device copy, 256 MiB read + 256 MiB written per launch", "bytes_per_launch":
                      nbytes}}
    print(json.dumps(out), flush=True)


def form_fft(args):
    """Interpolation by Mixed-FFT against the separable contraction on the headline mesh (n_w = 32, |R_i| <= 5,
    22 channels, 128^3): the reference's Mixed-FFT (MixedFourierMatrixElements.jl:954-1123) splits the mesh
    into NKdiv^3 = 8^3 offsets of an NKFFT^3 = 16^3 grid; per offset it multiplies the R-space channels by
    the offset's phases, places them on the 16^3 grid (R mod 16) and transforms.  Timed here on the GPU with
    torch.fft (rocFFT) per offset, including the transposition of the result to per-k matrices that the
    eigensolver reads (the separable contraction writes them in that layout directly), and multiplied by the
    512 offsets; the separable contraction is the engine's own interpolation time on the same mesh (FP64)."""
    import torch
    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model
    if not torch.cuda.is_available():
        raise SystemExit("the FFT comparison runs on the GPU")
    model = config_model(1)
    n, nr = model.num_orbitals, model.num_r_vectors
    nkfft, nkdiv = 16, 8
    chans = 22
    dev = torch.device("cuda")
    rv = torch.as_tensor(model.r_vectors, dtype=torch.int64, device=dev)
    cell = ((rv[:, 0] % nkfft) * nkfft + (rv[:, 1] % nkfft)) * nkfft + (rv[:, 2] % nkfft)
    g = torch.Generator(device="cpu").manual_seed(5)
    fr = torch.randn(chans * n * n, nr, dtype=torch.complex128, generator=g).to(dev)
    grid = torch.empty(chans * n * n, nkfft ** 3, dtype=torch.complex128, device=dev)

    def offset(d):
        ph = torch.exp(2j * np.pi * (rv.to(torch.float64) @ torch.tensor(d, dtype=torch.float64, device=dev))
                       / (nkdiv * nkfft))
        grid.zero_()
        grid[:, cell] = fr * ph[None, :]
        out = torch.fft.fftn(grid.view(-1, nkfft, nkfft, nkfft), dim=(1, 2, 3))
        return out.view(chans * n * n, -1).transpose(0, 1).contiguous()     # per-k matrices

    offsets = [(i, j, k) for i in range(2) for j in range(2) for k in range(2)]
    offset(offsets[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for d in offsets:
        offset(d)
    torch.cuda.synchronize()
    per_offset = (time.perf_counter() - t0) / len(offsets)
    fft_mesh_ms = 1e3 * per_offset * nkdiv ** 3
    task = E.EngineTask(photon_energies=np.linspace(0.0, 6.0, 400), broadening=0.01, shift_current=True,
                        injection_current=True, precision="fp64")
    with E.Engine(model, task) as eng:
        _, _, st = _timed_runs(eng, (128, 128, 128), 1, 1)
    out = {"metric": "interpolation ms per 128^3 mesh (n_w=32, N_R=1331, 22 channels)", "value": fft_mesh_ms,
           "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "Mixed-FFT NKdiv=8 x NKFFT=16 per axis (torch.fft / rocFFT, 8 offsets timed, x512) "
                                  "against the engine's separable contraction", "kpoints": 128 ** 3},
           "fft_ms_per_offset": 1e3 * per_offset, "fft_mesh_ms": fft_mesh_ms,
           "separable_mesh_ms": st["interp_ms"], "ratio_fft_over_separable": fft_mesh_ms / st["interp_ms"]}
    print(json.dumps(out), flush=True)


FORMS = {"geometry": form_geometry, "shg": form_shg, "batch": form_batch, "learned": form_learned,
         "large": form_large, "symmetry": form_symmetry, "copy": form_copy,
         "fft": form_fft}


if __name__ == "__main__":
    main()
