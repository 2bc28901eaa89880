#!/usr/bin/env python3
"""Benchmark of the CCSD amplitude-residual hot path on one MI355X.

Default workload (one "step" = one CCSD iteration over resident data):
  ccsd-h2o8  (H2O)8 on a 2x2x2 lattice (O-O 2.8 A, seeded 0.1 A jitter, random
             orientations), cc-pVDZ, o=40, v=152: RHF and AO->MO on the host,
             then the converged solve to 1e-8 and K timed iterations
             (update + tau + residual + energy/error) in FP64 and in each
             mixed-precision form; the headline is the fastest form within
             1 uEh of FP64.

Prints ONE JSON line with the metric, the roofline of the dominant kernel (the
pp-ladder contraction, algorithmic FLOPs 2 o^2 v^4 per launch) and the CPU
baseline (one oracle residual on the host cores).

Other workloads: pp-ladder (R[a,b,i,j] += sum_cd tau[c,d,i,j] V[a,b,c,d],
o=32, v=256, FP64, ccsd_r1_r2.h:275; V ~ N(0, 0.05) symmetrised to the 8-fold
(ac|bd) symmetry, tau ~ N(0, 1e-2) / (e_c + e_d - e_i - e_j)); ccsd-h2o6
(o=30, v=114); ccsd-h20 (H20 chain, o=10, v=90); copy (counter calibration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# TFLOP/s dense MFMA, MI355X, of the unit each form's contractions run on
PEAK = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0, "f16x3": 2500.0, "bf16x6": 2500.0}
DTYPE = {"f64": "f64",
         "f32": "f32-mfma contractions / f64 update, DIIS, energy",
         "bf16x3": "bf16 hi+lo split contractions (3 products) / f64 update, DIIS, energy",
         "f16x3": "fp16 hi+lo split contractions (3 products) / f64 update, DIIS, energy",
         "bf16x6": "bf16 hi+mid+lo split contractions (6 products) / f64 update, DIIS, energy"}


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="ccsd-h2o8",
                   choices=["ccsd-h2o8", "ccsd-h2o6", "ccsd-h20", "pp-ladder", "copy"])
    p.add_argument("--forms", default="f64,f16x3",
                   help="contraction precisions to run for the CCSD workloads (first = reference)")
    p.add_argument("--max-iter", type=int, default=40)
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


SYSTEMS = {
    # name: (description, builder)
    "ccsd-h2o8": "(H2O)8 on a 2x2x2 cubic lattice, O-O 2.8 A +- U[0.1] A jitter, random "
                 "orientations (seed 4), cc-pVDZ, all electrons (o=40, v=152)",
    "ccsd-h2o6": "(H2O)6 on a 3x2x1 lattice, same generator (seed 5), cc-pVDZ, all electrons "
                 "(o=30, v=114)",
    "ccsd-h20": "linear H20 chain, 1.8 bohr spacing, cc-pVDZ, all electrons (o=10, v=90)",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_system(name):
    from mpqc_amd.chem.molecule import h_chain, water_cluster
    from mpqc_amd.chem.scf import rhf
    from mpqc_amd.chem.mo import ccsd_integrals
    if name == "ccsd-h2o8":
        mol = water_cluster((2, 2, 2), seed=4)
    elif name == "ccsd-h2o6":
        mol = water_cluster((3, 2, 1), seed=5)
    else:
        mol = h_chain(20, 1.8)
    log(f"{name}: {len(mol.symbols)} atoms, RHF/cc-pVDZ")
    t0 = time.perf_counter()
    scf = rhf(mol, "cc-pVDZ")
    t1 = time.perf_counter()
    log(f"RHF E={scf.energy:.10f} in {t1 - t0:.1f}s; AO->MO")
    inp = ccsd_integrals(scf, frozen_core=False)
    t2 = time.perf_counter()
    log(f"MO integrals o={inp.o} v={inp.v} in {t2 - t1:.1f}s")
    del scf
    return inp, {"rhf_s": round(t1 - t0, 2), "ao2mo_s": round(t2 - t1, 2)}


def cpu_baseline_ccsd(inp):
    """One residual evaluation (cs_ccsd_r1 + cs_ccsd_r2) of the oracle on the
    host cores at the first-iteration (MP2) amplitudes."""
    import numpy as np
    from oracle import ccsd as O
    ints = O.Integrals(**inp.blocks)
    r2 = ints.Gijab.transpose(2, 3, 0, 1)
    t1 = O.jacobi_update_t1(ints.Fia.T, inp.eps_occ, inp.eps_vir)
    t2 = O.jacobi_update_t2(r2, inp.eps_occ, inp.eps_vir)
    tau = O.make_tau(t1, t2)
    t0 = time.perf_counter()
    O.cs_ccsd_r1(t1, t2, tau, ints)
    O.cs_ccsd_r2(t1, t2, tau, ints)
    dt = time.perf_counter() - t0
    flops = O.ccsd_flops(inp.o, inp.v)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "seconds_per_iteration": dt,
            "sample": f"one oracle residual (numpy/OpenBLAS fp64 einsum) at o={inp.o} v={inp.v}, "
                      f"{flops:.3e} algorithmic FLOPs"}


def bench_ccsd(args):
    import numpy as np
    from mpqc_amd.ccsd import CCSD, Integrals
    from oracle.ccsd import ccsd_flops
    inp, setup = build_system(args.workload)
    o, v = inp.o, inp.v
    flops = ccsd_flops(o, v)
    ints = Integrals(**inp.blocks)
    forms = {}
    for prec in args.forms.split(","):
        c = CCSD(o, v, precision=prec, max_iter=args.max_iter)
        c.set_integrals(ints)
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        # converged solve (target precision 1e-8, as the workload specifies)
        E = c.compute_ccsd(target_precision=1e-8)
        rows = c.trace()
        log(f"{prec}: E_corr={E:.12f} after {len(rows)} iterations")
        # timed iterations: W untimed, then K timed; one iteration = Jacobi/DIIS
        # update + tau + residual + energy/error (the last call synchronises)
        c.init_guess()
        c.energy_error()
        for _ in range(args.warmup):
            c.update()
            c.residual()
            c.energy_error()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            c.update()
            c.residual()
            c.energy_error()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        # the same iterations again with per-contraction HIP-event timing
        c.set_option("kernel_timing", 1)
        c.init_guess()
        c.energy_error()
        for _ in range(args.warmup + args.steps):
            c.update()
            c.residual()
            c.energy_error()
        kt = c.kernel_times()
        c.close()
        n_it = args.warmup + args.steps
        tot = sum(x["ms"] for x in kt.values()) / n_it
        log(f"{prec}: {ms:.2f} ms/iteration; event-timed phases sum to {tot:.2f} ms/iteration")
        for k, x in sorted(kt.items(), key=lambda kv: -kv[1]["ms"])[:40]:
            log(f"   {k:24s} {x['ms'] / n_it:8.3f} ms/it  {x['calls'] / n_it:5.1f} calls/it"
                + (f"  {x['flops'] / max(1e-12, x['ms'] * 1e-3) / 1e12:7.2f} TF" if x["flops"] else ""))
        forms[prec] = {"E_corr": E, "iterations": len(rows), "ms_per_iteration": ms,
                       "tflops": flops / (ms * 1e-3) / 1e12,
                       "solve_iteration_ms": [round(r.seconds * 1e3, 3) for r in rows],
                       "kernels": {k: {"calls": x["calls"], "avg_ms": x["ms"] / max(1, x["calls"]),
                                       "tflops": x["flops"] / max(1e-12, x["ms"] * 1e-3) / 1e12}
                                   for k, x in kt.items()}}
    ref = forms[args.forms.split(",")[0]]
    for prec, f in forms.items():
        f["dE_vs_f64"] = f["E_corr"] - ref["E_corr"]
        f["speedup_vs_f64"] = ref["ms_per_iteration"] / f["ms_per_iteration"]
    # headline form: the fastest one within 1 uEh of FP64
    ok = [p for p, f in forms.items() if abs(f["dE_vs_f64"]) <= 1e-6]
    head = min(ok, key=lambda p: forms[p]["ms_per_iteration"])
    hf = forms[head]
    pp = hf["kernels"].get("r2:pp_ladder", {"avg_ms": float("nan"), "tflops": float("nan")})
    log("CPU baseline (oracle residual)" if not args.no_cpu_baseline else "no CPU baseline")
    cpu = None if args.no_cpu_baseline else cpu_baseline_ccsd(inp)
    line = {
        "metric": "CCSD iter wall-time + TFLOP/s (frac MFMA peak), mixed vs FP64 at <=1 uEh",
        "value": round(hf["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hf["ms_per_iteration"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic molecule (seeded geometry), RHF + AO->MO on the host",
        "config": {"workload": SYSTEMS[args.workload], "o": o, "v": v,
                   "algorithmic_flops_per_iteration": flops, "form": head,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "achieved": round(pp["tflops"], 3),
                     "peak": PEAK[head], "unit": "TFLOP/s",
                     "frac": round(pp["tflops"] / PEAK[head], 4), "traffic": None,
                     "kernel": f"pp-ladder GEMM (M=v^2, N=o^2, K=v^2) avg {pp['avg_ms']:.3f} ms"},
        "forms": {p: {k: (round(x, 6) if isinstance(x, float) else x) for k, x in f.items()
                      if k != "kernels"} for p, f in forms.items()},
        "kernels": {p: f["kernels"] for p, f in forms.items()},
        "setup": setup,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["cpu_baseline"]["gpu_speedup"] = round(hf["tflops"] / cpu["value"], 1)
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
    elif args.workload == "pp-ladder":
        line = bench_pp_ladder(args)
    else:
        line = bench_ccsd(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
