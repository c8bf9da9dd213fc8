#!/usr/bin/env python3
"""Throughput benchmark: eigen-solved walks/s on one MI355X.

Workload (the headline configuration): Johansen model 4, dim 12, T = 10,000
steps, 1,000,000 walks per step (seeds 1..1,000,000, Philox stream 0).  One
"step" is one pass of the walk path over that batch: device normals ->
one-pass moments -> model algebra -> FP64 eigen solve, with the seeds already
resident in HBM and the eigenvalues left in HBM.

Prints one JSON line with walks/s, the roofline figure of the dominant kernel
(timed with HIP events on its own stream) and the CPU baseline (the C
restatement of the reference walk in oracle/, on a bounded sample of the same
workload, on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "eigen-solved walks/s @dim12 T=10k model4 vs FP64 GPU; p50/p99 query latency"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)


def algorithmic_flops_per_walk(model: int, dim: int, steps: int) -> int:
    """2*T*nF*(nF+dim): the S11 and S01 contractions as the reference forms
    them (src/johansen_statistics.rs:32-33)."""
    nf = dim + 1 if model in (1, 3) else dim
    return 2 * steps * nf * (nf + dim)


def cpu_baseline(model, dim, steps, threads, target_s=12.0):
    import numpy as np

    import oracle

    # calibrate on a small sample, then time ~target_s of work
    cal_n = 2 * threads
    t0 = time.perf_counter()
    oracle.walks_by_seed_c(model, dim, steps, np.arange(1, cal_n + 1, dtype=np.uint32), threads=threads)
    rate = cal_n / (time.perf_counter() - t0)
    n = int(min(max(rate * target_s, 4 * threads), 200_000))
    seeds = np.arange(1, n + 1, dtype=np.uint32)
    t0 = time.perf_counter()
    eig = oracle.walks_by_seed_c(model, dim, steps, seeds, threads=threads)
    dt = time.perf_counter() - t0
    return n / dt, n, dt, eig


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--walks", type=int, default=1_000_000, help="walks per step")
    ap.add_argument("--model", type=int, default=4)
    ap.add_argument("--dim", type=int, default=12)
    ap.add_argument("--T", type=int, default=10_000, help="time steps per walk")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this engine runs on one GPU per process and this build measures one GPU only "
              f"(--gpus {args.gpus}, WORLD_SIZE {world}); walks shard by seed range across GPUs with no "
              f"exchange until the final .dat/quantile merge. Run with --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import __graft_entry__

    pkg = __graft_entry__.import_package()
    prec = pkg.PRECISIONS[args.precision]
    model, dim, T, n = args.model, args.dim, args.T, args.walks
    nf = pkg.num_eigenvalues(model, dim)

    torch.cuda.init()
    dev = torch.device("cuda:0")
    eng = pkg.Engine(0)
    seeds = torch.arange(1, n + 1, dtype=torch.int32, device=dev)  # uint32 bit pattern
    out = torch.empty(n * nf, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step():
        eng.eigenvalues_by_seed_device(model, dim, T, seeds.data_ptr(), n, out.data_ptr(), precision=prec,
                                       hip_stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    eng.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mom_ms, mom_launches, mom_walks = eng.profile_read(pkg.KERNEL_MOMENTS)
    eig_ms, eig_launches, eig_walks = eng.profile_read(pkg.KERNEL_EIGEN)
    eng.profile_enable(False)

    elapsed = t1 - t0
    walks_per_s = args.steps * n / elapsed
    fpw = algorithmic_flops_per_walk(model, dim, T)
    avg_ms = mom_ms / max(mom_launches, 1)
    walks_per_launch = mom_walks / max(mom_launches, 1)
    achieved = fpw * walks_per_launch / (avg_ms * 1e-3) / 1e12

    gpu_eigs = out.view(n, nf)[:64].cpu().numpy()
    assert np.all(np.isfinite(gpu_eigs)), "non-finite eigenvalues"

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    key = f"walk_moments_{args.precision}_m{model}_d{dim}_T{T}"
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(key)

    line = {
        "metric": METRIC,
        "value": walks_per_s,
        "unit": "walks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else args.precision,
        "data": "synthetic (device Philox4x32-10 normals, seeds 1..N)",
        "config": {
            "workload": f"Johansen model {model}, dim {dim}, T={T:,} steps, {n:,} walks per step, "
                        f"{args.precision} moment accumulation, FP64 eigen solve",
            "johansen_model": model,
            "dim": dim,
            "T": T,
            "walks_per_step": n,
            "precision": args.precision,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": f"walk_moments_fp64<{dim}>",
            "algorithmic_flop_per_walk": fpw,
            "avg_launch_ms": avg_ms,
            "walks_per_launch": walks_per_launch,
        },
        "kernels_ms": {"moments": mom_ms, "eigen": eig_ms, "launches": mom_launches + eig_launches},
    }

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        rate, cn, cdt, ceig = cpu_baseline(model, dim, T, threads)
        # the GPU walks of seeds 1..64 against the CPU restatement
        k = min(64, cn)
        rel = np.max(np.abs(gpu_eigs[:k] - ceig[:k]) / np.maximum(np.abs(ceig[:k]), 1e-9 * ceig[:k].max()))
        line["cpu_baseline"] = {
            "value": rate,
            "unit": "walks/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{cn:,} walks (seeds 1..{cn:,}) of the same workload in {cdt:.1f} s; C restatement of "
                      f"the reference CPU walk (oracle/jne_oracle.c), one walk per thread",
            "gpu_vs_cpu_max_rel_diff_first_64": float(rel),
        }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
