"""Throughput of the wavepacket propagator on one MI355X.

Metric: grid-point-steps/s = sum over blocks of n1*n2*ntheta x steps / seconds,
with psi, V and all tables resident in HBM when the timed region starts.

`value` is the batched mixed-precision path (FP32 psi/FFT/phases, BF16 split
theta GEMM with FP32 accumulation, FP64 flux accumulation) on the F+H2 J=0
256x192x120 grid.  The plain FP64 path (one kernel per reference operator)
and the fused FP64 path are timed on the same grid and reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per grid point and step of each kernel class (M = ntheta):
# theta-fwd reads psi and V and writes the FBR array; the R and r passes and
# theta-bwd each read and write one complex array.
B_ALG = {
    "mixed": {"theta": 8 + 4 + 8 + 8 + 8, "theta_fwd": 8 + 4 + 8, "fft_R": 16, "fft_r": 16, "total": 68},
    "fp64": {"theta": 16 + 8 + 16 + 16 + 16, "theta_fwd": 16 + 8 + 16, "fft_R": 32, "fft_r": 32, "total": 136},
}
TK_THETA, TK_FFTR1, TK_FFTR2, TK_FLUX, TK_OTHER = 1, 2, 3, 4, 5


def timed_run(qd, prob, precision, pipeline, steps, warmup, profile=0):
    p = prob.copy()
    with qd.Propagator(p, precision=precision, pipeline=pipeline, profile=profile) as prop:
        if warmup:
            prop.propagate(warmup)
        t0 = time.perf_counter()
        prop.propagate(steps)
        wall = time.perf_counter() - t0
        step_ms = prop.timing(0)
        kern = {k: prop.timing(i) for k, i in
                (("theta", TK_THETA), ("fft_R", TK_FFTR1), ("fft_r", TK_FFTR2), ("flux", TK_FLUX),
                 ("other", TK_OTHER))}
    return wall, step_ms, kern


def cpu_baseline(prob, seconds_target=15.0):
    """The oracle (C++/OpenMP restatement of timeEvol.C, reference-faithful:
    P(E) every step and the per-step energy diagnostics on) on a bounded
    number of steps of the same grid."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding as oracle

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, 16)
    # one step to size the sample
    t0 = time.perf_counter()
    oracle.time_evolution(prob, total_steps=1, crp_every=1, diagnostics=1, nthreads=cores)
    t1 = time.perf_counter() - t0
    n = int(max(2, min(200, seconds_target / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.time_evolution(prob, total_steps=n, crp_every=1, diagnostics=1, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": prob.grid_points * n / dt, "unit": "grid-point-steps/s", "cores": cores, "kind": "port",
            "sample": "%d steps of the same grid (P(E) and energies every step), C++/OpenMP oracle, %.1f s"
                      % (n, dt)}


def calibrate_copy(mib=1024, reps=5):
    """Copies `mib` MiB device to device `reps` times: the reference reading
    for the FETCH_SIZE / WRITE_SIZE counters (each launch moves exactly
    mib MiB in and mib MiB out)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_MiB": mib, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a device-to-device copy of known size (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the propagator on a single MI355X; "
              "J blocks would shard across GPUs by J (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    qd = importlib.import_module("quantum-dynamics_amd")
    prob = qd.problem.config(args.workload)
    npt = prob.grid_points

    wall_m, step_m, _ = timed_run(qd, prob, "mixed", "fused", args.steps, args.warmup)
    _, _, kern_m = timed_run(qd, prob, "mixed", "fused", max(20, args.steps // 4), 5, profile=1)
    wall_p, step_p, _ = timed_run(qd, prob, "fp64", "plain", max(20, args.steps // 2), 5)
    wall_f, step_f, _ = timed_run(qd, prob, "fp64", "fused", max(20, args.steps // 2), 5)
    _, _, kern_f = timed_run(qd, prob, "fp64", "fused", max(20, args.steps // 4), 5, profile=1)

    value = npt / (step_m * 1e-3)
    plain = npt / (step_p * 1e-3)
    fused64 = npt / (step_f * 1e-3)
    # dominant kernel class of the mixed step
    dom = max(("theta", "fft_R", "fft_r"), key=lambda k: kern_m[k])
    per_launch_bytes = {
        "theta": B_ALG["mixed"]["theta"] / 2.0,  # two theta launches per step
        "fft_R": B_ALG["mixed"]["fft_R"],
        "fft_r": B_ALG["mixed"]["fft_r"],
    }[dom] * npt
    achieved = per_launch_bytes / (kern_m[dom] * 1e-3) / 1e9
    out = {
        "metric": "grid-point-steps/s (plain FP64 vs batched mixed) at |dP(E)|<=1e-4",
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_m,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bf16x3 split MFMA theta GEMM, f64 flux)",
        "data": "synthetic (LEPS surface, analytic packet)",
        "config": {"workload": prob.name + " mixed precision", "n1": prob.n1, "n2": prob.n2,
                   "ntheta": prob.ntheta, "grid_points": npt, "dt": prob.dt},
        "plain_fp64": {"value": plain, "ms_per_step": step_p},
        "fused_fp64": {"value": fused64, "ms_per_step": step_f},
        "speedup_vs_plain_fp64": value / plain,
        "kernel_ms": {"mixed": kern_m, "fp64_fused": kern_f},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "step_achieved_GBs": B_ALG["mixed"]["total"] * value / 1e9},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
