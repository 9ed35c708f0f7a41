"""Throughput of the wavepacket propagator on one MI355X.

Metric: grid-point-steps/s = sum over blocks of n1*n2*n_theta x steps / seconds,
with psi, V and all tables resident in HBM when the timed region starts.

`value` is the batched mixed-precision step (FP32 psi/FFT/phases, BF16 split
theta GEMM with FP32 accumulation, FP64 flux accumulation) on the batched
F+H2 workload: J = 0..20, every K = 0..min(J, 60), j_max = 60 -- 231
Coriolis-coupled (J, K) blocks of 256 x 192 x (61 - K) points, 616.9 M
grid points, advanced together every step.

Beside it, on the same grids:
  plain_fp64   the reference's operator sequence in FP64, one kernel per
               operator (V phase, theta-fwd, rotation, FFTs, kinetic, inverse
               FFTs, rotation, theta-bwd, dump), run block by block: the time
               per step is the sum over the 231 blocks of the measured step
               time of a 256 x 192 x (61 - K) single-block propagation
               (no Coriolis coupling, which only favours it);
  fused_fp64   the batched step in FP64;
  single_grid  the F+H2 J = 0 256 x 192 x 120 grid: mixed, fused FP64, plain FP64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--workload batched|single] [--calibrate]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per grid point of one launch of each kernel class
# (B = s_c (2 + 6 M / n_theta) + s_v per grid-point-step: 68 B mixed, 136 B
# FP64): theta-fwd reads psi (s_c) and V (s_v) and writes the FBR array; the R
# and r passes and theta-bwd each read and write one complex array.
# With the fused theta step (mixed precision, basis of at most 128) one kernel
# per step reads the FBR array and exp(-i V dt) (8 B each), writes the FBR
# array and the 11 r rows of the flux stencil (half-width 5) around the
# dividing surface, 8 B x 11 / 192 r points: 24.46 B per point of a single
# block.  Blocks that share V (the J blocks of one K) read it once per step:
# 16 B + 8 B x (V columns / DVR columns) + the flux rows, 17.2 B per point on
# the J = 0..20 batch (1071 V columns, 12551 DVR columns).
FLUX_ROWS_B = 8.0 * 11 / 192
B_ALG = {
    "mixed": {"theta_fwd": 8 + 4 + 8, "theta_bwd": 16, "theta_fused": 24 + FLUX_ROWS_B, "fft_R": 16, "fft_r": 16,
              "total": 68},
    "fp64": {"theta_fwd": 16 + 8 + 16, "theta_bwd": 32, "fft_R": 32, "fft_r": 32, "total": 136},
}
# HBM bytes per launch and grid point, measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE (fetch scaled x2 by a 1 GiB copy kernel) over the launches of
# timed steps (the theta step writes the flux rows there, and the r pass also
# extracts the previous step's flux surface: + 0.63 B); profiles/*.md.
MEASURED_B_PER_POINT = {
    "batched": {"theta": 17.2, "fft_R": 16.4, "fft_r": 17.0},  # batched_step_kernels_FH2_J0-20.md (fused theta)
    "single": {"theta": 24.65, "fft_R": 16.7, "fft_r": 17.2},  # hbm_traffic_FH2_256x192x120.md (wave-pair theta)
}
TK_THETA, TK_FFTR1, TK_FFTR2, TK_FLUX, TK_OTHER, TK_ACC = 1, 2, 3, 4, 5, 6
KCLASSES = (("theta", TK_THETA), ("fft_R", TK_FFTR1), ("fft_r", TK_FFTR2), ("flux", TK_FLUX), ("other", TK_OTHER),
            ("flux_accumulate", TK_ACC))
FLUX_BLOCK = 64  # steps per time block of the flux accumulation (the library's default)


def _run(prop, steps, warmup, profile):
    """Times `steps` steps and everything they owe: the flux slices of the
    timed steps are folded into phi(E), dphi(E) and P(E) is evaluated
    (qd_get_crp) before the clock stops.  The warm-up's slices are folded
    before the clock starts, so the timed region begins on an empty time
    block: whole blocks inside it run as replayed hipGraphs, the rest eagerly."""
    if warmup:
        prop.propagate(warmup)
    prop.crp()
    t0 = time.perf_counter()
    prop.propagate(steps)          # returns after the device has finished
    prop.crp()                     # the last (partial) time block's accumulation and P(E)
    wall = time.perf_counter() - t0
    out = {"wall_ms_per_step": 1e3 * wall / steps, "event_ms_per_step": prop.timing(0),
           "tuning": prop.tuning(), "flux_block": FLUX_BLOCK,
           "flux_accumulations_timed": (steps + FLUX_BLOCK - 1) // FLUX_BLOCK}
    if profile:
        out["kernel_ms"] = {k: prop.timing(i) for k, i in KCLASSES}
    return out


def timed_single(qd, prob, precision, pipeline, steps, warmup, profile=0, autotune=0):
    # whole flux blocks replayed as hipGraphs (the profiled runs launch eagerly,
    # with events between the kernels)
    with qd.Propagator(prob.copy(), precision=precision, pipeline=pipeline, profile=profile,
                       autotune=autotune, use_graph=0 if profile else 1, flux_block=FLUX_BLOCK) as prop:
        return _run(prop, steps, warmup, profile)


def timed_batched(qd, jk, precision, steps, warmup, profile=0, autotune=0):
    with qd.propagator.JKPropagator(jk, precision=precision, profile=profile, autotune=autotune,
                                    use_graph=0 if profile else 1, flux_block=FLUX_BLOCK) as prop:
        return _run(prop, steps, warmup, profile)


def batched_problem(qd, Jmax=20, jmax=60):
    pb = qd.problem
    base = pb.config(2, ntheta=jmax + 1)
    return pb.make_jk_problem(base, pb.F_H2, list(range(Jmax + 1)), jmax=jmax)


def plain_fp64_same_grids(qd, jk, steps, warmup):
    """Plain FP64 step time summed over the blocks of `jk` (one single-block
    run per distinct theta size)."""
    pb = qd.problem
    count = {}
    for b in jk.blocks:
        count[b.n] = count.get(b.n, 0) + 1
    per_n = {}
    for n in sorted(count):
        prob = pb.config(2, ntheta=n)
        per_n[n] = timed_single(qd, prob, "fp64", "plain", steps, warmup)["wall_ms_per_step"]
    ms = sum(count[n] * per_n[n] for n in count)
    return ms, per_n, count


def _host_cpus():
    """Threads the oracle may use: the CPUs this process may run on, limited
    by the cgroup CPU quota where one is set; with the CPU model."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return (min(n, quota) if quota else n), n, quota, model


def _fftw_present():
    import subprocess
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return False
    return "libfftw3" in out


def cpu_baseline(qd, block_sizes, seconds_target=20.0):
    """The oracle (C++/OpenMP restatement of timeEvol.C, reference-faithful:
    P(E) and the energy diagnostics every step) timed on a sample of the
    batched workload's blocks: one 256 x 192 x n grid for each of three theta
    sizes n (the smallest, middle and largest block size), a few steps each;
    the per-step time of every block is interpolated linearly in n from the
    sample and summed over the blocks.  The reference propagates J = 0
    only, so each block is timed as an uncoupled grid of its size (no K <-> l
    mixing, which only favours the CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding as oracle
    import numpy as np

    threads, nproc, quota, model = _host_cpus()
    sizes = sorted(set(block_sizes))
    sample = sorted({sizes[0], sizes[len(sizes) // 2], sizes[-1]})
    per_n, desc, total_s = {}, [], 0.0
    for n in sample:
        prob = qd.problem.config(2, ntheta=n)
        t0 = time.perf_counter()
        oracle.time_evolution(prob, total_steps=1, crp_every=1, diagnostics=1, nthreads=threads)
        t1 = time.perf_counter() - t0
        k = int(max(2, min(200, (seconds_target / len(sample)) / max(t1, 1e-3))))
        t0 = time.perf_counter()
        oracle.time_evolution(prob, total_steps=k, crp_every=1, diagnostics=1, nthreads=threads)
        dt = time.perf_counter() - t0
        total_s += dt
        per_n[n] = dt / k
        desc.append("%d steps of 256x192x%d" % (k, n))
    xs = np.array(sample, dtype=float)
    ys = np.array([per_n[n] for n in sample])
    step_s = float(sum(np.interp(n, xs, ys) for n in block_sizes))
    points = 256 * 192 * sum(block_sizes)
    return {"value": points / step_s, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "s_per_step_extrapolated": step_s, "cpu_model": model, "nproc": nproc, "cgroup_cpu_quota": quota,
            "fft": "FFTW present on the host but not used: the oracle's own mixed-radix FFT" if _fftw_present()
                   else "the oracle's own mixed-radix FFT (no libfftw3 on the host)",
            "sample": "C++/OpenMP oracle on %d threads (%s), P(E) and energies every step: %s (%.1f s); "
                      "per-step time interpolated in n_theta and summed over the workload's %d block(s)"
                      % (threads, model, ", ".join(desc), total_s, len(block_sizes))}


def accuracy_report(qd, out_path=None):
    """Per-energy |P_mixed(E) - P_fp64(E)| over whole propagations (10000
    steps of 5 au): the F+H2 J = 0 256 x 192 x 120 grid, and J = 0, 10, 20 of
    the batched workload (all K blocks, Coriolis-coupled, j_max = 60)."""
    import numpy as np

    pb = qd.problem
    rep = {}
    prob = pb.config(2)
    P = {prec: qd.run(prob, precision=prec)[1] for prec in ("fp64", "mixed")}
    d = np.abs(P["mixed"] - P["fp64"])
    rep["single_grid"] = {"workload": "F+H2 J=0 256x192x120, 10000 x 5 au", "max_abs_dP": float(d.max()),
                          "at_E_eV": float(prob.collision_energies_ev[d.argmax()]), "max_P": float(P["fp64"].max()),
                          "E_eV": prob.collision_energies_ev.tolist(), "P_fp64": P["fp64"].tolist(),
                          "abs_dP": d.tolist()}
    jk = pb.make_jk_problem(pb.config(2, ntheta=61), pb.F_H2, [0, 10, 20], jmax=60)
    PJ = {}
    for prec in ("fp64", "mixed"):
        with qd.propagator.JKPropagator(jk, precision=prec) as prop:
            for k in range(10):
                prop.propagate(jk.base.total_steps // 10)
                print("accuracy: batched %s %d0%%" % (prec, k + 1), file=sys.stderr, flush=True)
            PJ[prec] = prop.crp_by_J()
    rep["batched"] = {"workload": "F+H2 J=0,10,20, K=0..J, j_max=60, 256x192x(61-K), 10000 x 5 au",
                      "per_J": {}}
    for J in PJ["fp64"]:
        dj = np.abs(PJ["mixed"][J] - PJ["fp64"][J])
        rep["batched"]["per_J"][int(J)] = {"max_abs_dP": float(dj.max()), "max_P": float(PJ["fp64"][J].max()),
                                           "abs_dP": dj.tolist(), "P_fp64": PJ["fp64"][J].tolist()}
    if out_path:
        with open(out_path, "w") as f:
            json.dump(rep, f)
    return {"metric": "per-energy |P_mixed - P_fp64|", "single_grid_max_abs_dP": rep["single_grid"]["max_abs_dP"],
            "batched_max_abs_dP": {J: v["max_abs_dP"] for J, v in rep["batched"]["per_J"].items()}}


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


def roofline(kern, npt, workload, fused_theta, prec="mixed"):
    """Dominant kernel class of the step, its algorithmic bytes per launch
    over its event-timed launch duration (the theta class is one fused launch
    per step, or the mean of the forward and backward launches)."""
    dom = max(("theta", "fft_R", "fft_r"), key=lambda k: kern[k])
    b = B_ALG[prec]
    theta = b["theta_fused"] if fused_theta else (b["theta_fwd"] + b["theta_bwd"]) / 2.0
    per_point = {"theta": theta, "fft_R": b["fft_R"], "fft_r": b["fft_r"]}[dom]
    achieved = per_point * npt / (kern[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_B_PER_POINT[workload][dom] * npt,
            "traffic_unit": "bytes per launch (rocprof)", "algorithmic_bytes": per_point * npt,
            "note": ("the R pass with K<->l mixing is not HBM-bound: about 127 VALU instructions per point (62 the "
                     "register FFT with its phases, 65 the mixing in both directions), VALU issuing on about 58 % of "
                     "its cycles, and waits at the bundle barriers around the mixing; it moves its 16 algorithmic "
                     "bytes per point (16.4 measured)" if (dom == "fft_R" and workload == "batched") else
                     "the fused theta step on one 256x192x120 grid (wave pairs, one per component) is latency-bound: "
                     "256 work groups of eight waves, one per CU (110.6 KB of theta matrices in LDS), waiting on "
                     "about 46 % of their wave cycles; it moves its 24.5 algorithmic bytes per point (24.65 "
                     "measured)" if (dom == "theta" and workload == "single") else "")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("batched", "single", "jinterp", "accuracy"), default="batched",
                    help="jinterp: the learned J interpolation end to end (F+H2, every 4th J of 0..40 propagated, "
                         "the network trained on the GPU, checked against propagating all 41 J)")
    ap.add_argument("--jinterp-steps", type=int, default=10000, help="propagation steps of the jinterp workload")
    ap.add_argument("--out", help="jinterp: also write the full report (per-energy sums) as JSON here")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a device-to-device copy of known size (HBM counter calibration)")
    ap.add_argument("--only", choices=("batched-mixed", "single-mixed"),
                    help="run just this step (autotuned) for K steps after W warm-up steps and print its "
                         "timing; for profiler runs")
    args = ap.parse_args()
    if args.calibrate:
        calibrate_copy()
        return
    if args.workload == "accuracy":
        print(json.dumps(accuracy_report(importlib.import_module("quantum-dynamics_amd"), args.out)))
        return
    if args.workload == "jinterp":
        qd = importlib.import_module("quantum-dynamics_amd")
        ji = importlib.import_module("quantum-dynamics_amd.jinterp")
        pb = qd.problem
        base = pb.config(2, ntheta=61)
        rep = ji.learned_partial_waves(qd, base, pb.F_H2, J_max=40, stride=4, jmax=60, steps=args.jinterp_steps,
                                       reference="all", log=lambda m: print(m, file=sys.stderr, flush=True))
        if args.out:
            with open(args.out, "w") as f:
                json.dump(rep, f)
        rep.pop("per_energy", None)
        print(json.dumps({"metric": "learned J interpolation: partial-wave-sum error and wall time vs all 41 J",
                          "workload": "F+H2 J=0..40 (every 4th propagated), K=0..min(J,60), j_max=60, "
                                      "256x192x(61-K), mixed precision", **rep}))
        return
    if args.only:
        qd = importlib.import_module("quantum-dynamics_amd")
        if args.only == "batched-mixed":
            jk = batched_problem(qd)
            r = timed_batched(qd, jk, "mixed", args.steps, args.warmup, profile=1, autotune=1)
            npt = jk.grid_points
        else:
            prob = qd.problem.config(2)
            r = timed_single(qd, prob, "mixed", "fused", args.steps, args.warmup, profile=1, autotune=1)
            npt = prob.grid_points
        print(json.dumps({"only": args.only, "grid_points": npt, "steps": args.steps, **r}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the propagator runs on a single MI355X here; the batched "
              "workload would shard by J across GPUs (the K blocks of one J stay together; P^J(E) is "
              "gathered once at the end)", file=sys.stderr)
        sys.exit(2)
    qd = importlib.import_module("quantum-dynamics_amd")
    K, W = args.steps, args.warmup

    # single grid (F+H2 J=0, 256 x 192 x 120)
    prob = qd.problem.config(2)
    npt1 = prob.grid_points
    s_mixed = timed_single(qd, prob, "mixed", "fused", K, W, autotune=1)
    s_mixed["kernel_ms"] = timed_single(qd, prob, "mixed", "fused", max(20, K // 4), 5, profile=1,
                                        autotune=1)["kernel_ms"]
    # every leg that enters a ratio is timed over the same window of K steps,
    # with its folds and P(E) inside (a shorter window would charge a leg
    # more fold per step)
    s_f64 = timed_single(qd, prob, "fp64", "fused", K, W, profile=1)
    s_plain = timed_single(qd, prob, "fp64", "plain", K, W)
    single = {
        "workload": "F+H2 J=0 256x192x120", "grid_points": npt1,
        "mixed": {"value": npt1 / (s_mixed["wall_ms_per_step"] * 1e-3), **s_mixed},
        "fused_fp64": {"value": npt1 / (s_f64["wall_ms_per_step"] * 1e-3), **s_f64},
        "plain_fp64": {"value": npt1 / (s_plain["wall_ms_per_step"] * 1e-3), **s_plain},
    }
    single["speedup_mixed_vs_plain_fp64"] = single["mixed"]["value"] / single["plain_fp64"]["value"]
    single["roofline"] = roofline(s_mixed["kernel_ms"], npt1, "single", s_mixed["tuning"]["fused_theta"])

    if args.workload == "batched":
        jk = batched_problem(qd)
        npt = jk.grid_points
        label = ("F+H2 J=0..20, K=0..min(J,60), j_max=60: 231 Coriolis-coupled (J,K) blocks of 256x192x(61-K), "
                 "mixed precision")
        m = timed_batched(qd, jk, "mixed", K, W, autotune=1)
        prof = timed_batched(qd, jk, "mixed", max(10, K // 10), 2, profile=1, autotune=1)
        f64 = timed_batched(qd, jk, "fp64", max(10, K // 10), 2, autotune=1)
        plain_ms, plain_per_n, counts = plain_fp64_same_grids(qd, jk, K, W)
        step_ms = m["wall_ms_per_step"]
        value = npt / (step_ms * 1e-3)
        plain = {"value": npt / (plain_ms * 1e-3), "ms_per_step": plain_ms,
                 "how": "sum over blocks of single-block plain FP64 step times (per theta size, ms)",
                 "per_theta_size_ms": plain_per_n, "blocks_per_theta_size": counts}
        fused = {"value": npt / (f64["wall_ms_per_step"] * 1e-3), "ms_per_step": f64["wall_ms_per_step"]}
        kern = prof["kernel_ms"]
        rf = roofline(kern, npt, "batched", prof["tuning"]["fused_theta"])
        cfg = {"workload": label, "blocks": len(jk.blocks), "n1": jk.base.n1, "n2": jk.base.n2,
               "theta_sizes": "41..61", "grid_points": npt, "dt": jk.base.dt}
    else:
        npt = npt1
        step_ms = s_mixed["wall_ms_per_step"]
        value = single["mixed"]["value"]
        plain = {"value": single["plain_fp64"]["value"], "ms_per_step": s_plain["wall_ms_per_step"]}
        fused = {"value": single["fused_fp64"]["value"], "ms_per_step": s_f64["wall_ms_per_step"]}
        kern = s_mixed["kernel_ms"]
        rf = single["roofline"]
        m = s_mixed
        cfg = {"workload": "F+H2 J=0 256x192x120 mixed precision", "n1": prob.n1, "n2": prob.n2,
               "ntheta": prob.ntheta, "grid_points": npt, "dt": prob.dt}

    out = {
        "metric": "grid-point-steps/s (plain FP64 vs batched mixed) at |dP(E)|<=1e-4",
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bf16x3 split MFMA theta GEMM, f64 flux)",
        "data": "synthetic (analytic LEPS surface on the Jacobi grid, analytic initial packet)",
        "config": cfg,
        "event_ms_per_step": m["event_ms_per_step"],
        "plain_fp64": plain,
        "fused_fp64": fused,
        "speedup_vs_plain_fp64": value / plain["value"],
        "kernel_ms": kern,
        "tuning": m["tuning"],
        "roofline": rf,
        "single_grid": single,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(qd, [b.n for b in jk.blocks] if args.workload == "batched"
                                           else [prob.ntheta])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
