#!/usr/bin/env python3
"""Benchmark of the CCSD amplitude-residual hot path on one MI355X.

Default workload (one "step" = one launch of the hot kernel over resident data):
  pp-ladder  R[a,b,i,j] += sum_cd tau[c,d,i,j] V[a,b,c,d], o=32, v=256, FP64
             (ccsd_r1_r2.h:275).  V ~ N(0, 0.05) symmetrised to the 8-fold
             (ac|bd) symmetry, tau ~ N(0, 1e-2) / (e_c + e_d - e_i - e_j),
             e_occ ~ U[-1.5, -0.3], e_vir ~ U[0.1, 3.0]; V and tau resident in
             HBM before timing.

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(MFMA-bound, algorithmic FLOPs 2 o^2 v^4 per launch) and the CPU baseline (the
oracle's contraction on the host cores over a bounded row sample).

Other workloads: --workload ccsd-h2o8 runs the full CCSD iteration on the
(H2O)8 cc-pVDZ system (o=40, v=152) in FP64 and mixed precision.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0}  # TFLOP/s dense MFMA, MI355X


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="pp-ladder", choices=["pp-ladder", "copy"])
    p.add_argument("--copy-gib", type=float, default=4.0,
                   help="size of the plain copy used to calibrate rocprofv3 byte counters")
    p.add_argument("--precision", default="f64", choices=["f64", "f32", "bf16x3"])
    p.add_argument("--o", type=int, default=32)
    p.add_argument("--v", type=int, default=256)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample-rows", type=int, default=8192)
    return p.parse_args()


def make_pp_ladder_inputs(o, v, seed=2):
    """Synthetic pp-ladder operands, generated on the GPU (V) / host (tau)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    e_occ = rng.uniform(-1.5, -0.3, o)
    e_vir = rng.uniform(0.1, 3.0, v)
    tau = rng.normal(0.0, 1e-2, (v, v, o, o))
    tau /= (e_vir[:, None, None, None] + e_vir[None, :, None, None]
            - e_occ[None, None, :, None] - e_occ[None, None, None, :])
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    # X[p,q,r,s] = (pq|rs) ~ N(0, 0.05), symmetrised over the 8 real-orbital
    # permutations; V[a,b,c,d] = (ac|bd)
    X = torch.empty((v, v, v, v), dtype=torch.float64, device="cuda")
    X.normal_(0.0, 0.05, generator=g)
    S = X.clone()
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2), (2, 3, 0, 1), (3, 2, 0, 1),
                 (2, 3, 1, 0), (3, 2, 1, 0)]:
        S.add_(X.permute(*perm))
    del X
    S.mul_(0.125)
    V = S.permute(0, 2, 1, 3).contiguous()
    del S
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return V, torch.from_numpy(tau).to("cuda"), e_occ, e_vir, tau


def cpu_baseline_pp(V, tau_np, rows):
    """The oracle's pp-ladder contraction (oracle.ccsd.pp_ladder, a BLAS dgemm
    through numpy) on a bounded sample of rows (a,b) of V."""
    import numpy as np
    from oracle import ccsd as O  # noqa: F401  (the oracle defines the contraction)
    v = tau_np.shape[0]
    o = tau_np.shape[2]
    rows = min(rows, v * v)
    Vs = V.reshape(v * v, v * v)[:rows].cpu().numpy()
    T = tau_np.reshape(v * v, o * o)
    Vs @ T[:, :8]  # warm the BLAS pool
    t0 = time.perf_counter()
    R = Vs @ T
    dt = time.perf_counter() - t0
    flops = 2.0 * rows * v * v * o * o
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "seconds": dt,
            "sample": f"oracle pp-ladder GEMM on rows 0..{rows - 1} of V ({rows}x{v * v} @ "
                      f"{v * v}x{o * o}, {flops:.3e} FLOPs, numpy/OpenBLAS fp64)",
            "full_kernel_seconds_extrapolated": dt * (v * v) / rows}, R


def bench_pp_ladder(args):
    import numpy as np
    import torch
    from mpqc_amd.ccsd import pp_ladder
    o, v = args.o, args.v
    V, tau, e_occ, e_vir, tau_np = make_pp_ladder_inputs(o, v)
    R = torch.zeros((v, v, o, o), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def step():
        pp_ladder(V.data_ptr(), tau.data_ptr(), R.data_ptr(), o, v, 1.0, 0.0, args.precision,
                  stream=sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        step()
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms = sum(kms) / len(kms)
    flops = 2.0 * o * o * float(v) ** 4
    tflops = flops / (ms * 1e-3) / 1e12
    # parity spot check against float64 host products on a few rows
    rows = np.random.default_rng(5).choice(v * v, 16, replace=False)
    Vr = V.reshape(v * v, v * v)[torch.from_numpy(rows).to("cuda")].cpu().numpy()
    ref = Vr @ tau_np.reshape(v * v, o * o)
    got = R.reshape(v * v, o * o)[torch.from_numpy(rows).to("cuda")].cpu().numpy()
    max_rel = float(np.abs(got - ref).max() / np.abs(ref).max())
    cpu = None
    if not args.no_cpu_baseline:
        cpu, _ = cpu_baseline_pp(V, tau_np, args.cpu_sample_rows)
    line = {
        "metric": "CCSD iter wall-time + TFLOP/s (frac MFMA peak), mixed vs FP64 at <=1 uEh",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "wall_s_timed_region": round(wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32-mfma", "bf16x3": "bf16x3-split"}[args.precision],
        "data": "synthetic (seeded: V torch CUDA generator seed 2, tau numpy PCG64 seed 2)",
        "config": {"workload": f"pp-ladder R[ab,ij]+=sum_cd tau[cd,ij] V[ab,cd] o={o} v={v}",
                   "o": o, "v": v, "flops_per_step": flops, "precision": args.precision,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3),
                     "peak": PEAK[args.precision], "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK[args.precision], 4), "traffic": None,
                     "kernel": "dgemm_f64_kernel (pp-ladder GEMM M=v^2 N=o^2 K=v^2)"},
        "parity_max_rel_err_vs_fp64_host": max_rel,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["cpu_baseline"]["gpu_speedup"] = round(tflops / cpu["value"], 1)
    return line


def bench_copy(args):
    """A plain device copy of known size: the reference point for scaling the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = int(args.copy_gib * (1 << 30) / 8)
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(args.steps):
        y.copy_(x)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / args.steps
    nbytes = 2 * n * 8
    return {"metric": "device copy bandwidth (counter calibration)", "value": nbytes / ms / 1e6,
            "unit": "GB/s", "ms_per_step": ms, "bytes_read_per_step": n * 8,
            "bytes_written_per_step": n * 8, "steps": args.steps, "warmup": args.warmup}


def main():
    args = _args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "designed but not built (see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if args.workload == "copy":
        line = bench_copy(args)
    else:
        line = bench_pp_ladder(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
