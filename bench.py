"""Benchmark: k-points/s of the dense-mesh shift + injection current k-loop on one MI355X.

Workload (the configuration the headline metric is quoted on): synthetic orthogonal TB model with
n_w = 32 orbitals, R cube |R_i| <= 5 (1331 R vectors), position operator amplitude 0.1 Angstrom,
a = 4 Angstrom cubic cell; conventional shift current + injection current, all 27 abc components,
400 photon energies in [0, 6] eV, Gaussian width 0.01 eV, mu = 0, T = 0, on the 128^3 Gamma-centred
mesh (2,097,152 k points).  One step = one full-mesh integral (all stages on the device, model and
R-space channels already resident).  The FP64 GPU path runs beside the headline path for the
relative spectrum error; the CPU baseline times the NumPy FP64 oracle on a bounded k sample.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64] [--mesh N]
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


def cpu_baseline(model, omega, width, k_sample: int):
    """Time the NumPy FP64 oracle (restatement of the reference's CPU k-loop) on the first
    k_sample points of the same mesh; returns k-points/s and the thread count used."""
    from oracle import kloop
    try:
        from threadpoolctl import threadpool_info
        threads = max([int(p.get("num_threads", 1)) for p in threadpool_info()] + [1])
    except Exception:
        threads = os.cpu_count() or 1
    task = kloop.Task(photon_energies=omega, broadening=width, shift_current=True, injection_current=True)
    t0 = time.perf_counter()
    kloop.integrate_mesh(model, task, (128, 128, 128), batch=64, k_limit=k_sample)
    dt = time.perf_counter() - t0
    return k_sample / dt, threads, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["mixed", "fp64"])
    ap.add_argument("--mesh", type=int, default=128)
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--cpu-sample", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the k-mesh would shard by "
              "deterministic lanes across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    from wanniernlqg_amd import engine as E
    from wanniernlqg_amd.model import config_model

    model = config_model(1)
    omega = np.linspace(0.0, 6.0, 400)
    width = 0.01
    mesh = (args.mesh,) * 3
    nk = int(np.prod(mesh))

    def make(prec):
        return E.Engine(model, E.EngineTask(photon_energies=omega, broadening=width, shift_current=True,
                                            injection_current=True, precision=prec))

    log(f"[bench] model n=32 N_R={model.num_r_vectors} mesh={mesh} precision={args.precision}")
    t0 = time.perf_counter()
    eng = make(args.precision)
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
            f"rot {st['rotate_ms']:.1f}, pairs {st['pairs_ms']:.1f}, accum {st['accum_ms']:.1f})")
    eng.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = nk * args.steps / wall
    head_total = tot

    # roofline of the dominant kernel (interpolation GEMM), from its device event timings
    st = stats[-1]
    launches = max(1, st["interp_launches"])
    interp_avg_s = 1e-3 * st["interp_ms"] / launches
    flops_per_launch = st["interp_flops"] / launches
    if args.precision == "fp64":
        roof = {"bound": "mfma", "achieved": flops_per_launch / interp_avg_s / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s"}
    else:
        roof = {"bound": "mfma", "achieved": 3.0 * flops_per_launch / interp_avg_s / 1e12, "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    eng.close()

    compare = None
    if args.precision != "fp64" and not args.no_compare:
        ref = make("fp64")
        _, rtot, rst = ref.run_mesh(mesh)
        ref.close()
        from oracle.kloop import spectrum_relative_error  # noqa: E402  (checker only)
        errs = {k: float(spectrum_relative_error(head_total[k], rtot[k]).max()) for k in rtot}
        compare = {"fp64_kpoints_per_s": nk / (1e-3 * rst["total_ms"]), "speedup_vs_fp64":
                   (nk / value) and (1e-3 * rst["total_ms"]) / (wall / args.steps), "max_rel_err": errs}
        log(f"[bench] fp64 compare: {compare}")

    cpu = None
    if not args.no_cpu:
        kps, threads, dt = cpu_baseline(model, omega, width, args.cpu_sample)
        cpu = {"value": kps, "unit": "k-points/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_sample} k points of the 128^3 mesh, same model/task, NumPy FP64 oracle "
                         f"({dt:.1f} s)"}
        log(f"[bench] cpu baseline {kps:.1f} k/s on {threads} threads")

    out = {
        "metric": "k-points/s (n_w=32,128^3 shift current) mixed vs FP64 GPU at rel err <=1e-6",
        "value": value,
        "unit": "k-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "bf16x3-mfma+f64",
        "data": "synthetic (seeded TB model, no external data)",
        "config": {"workload": f"cfg1 shift+injection current n_w=32 N_R=1331 {args.mesh}^3 mesh, 400 omega, "
                               f"27 components, precision={args.precision}",
                   "kpoints": nk, "n_w": 32, "num_r_vectors": model.num_r_vectors, "n_omega": 400},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": {k: st[k] for k in ("interp_ms", "eig_ms", "rotate_ms", "pairs_ms", "accum_ms", "total_ms")},
    }
    if compare:
        out["fp64_compare"] = compare
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
