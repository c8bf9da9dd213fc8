#!/usr/bin/env python3
"""Throughput benchmark: eigen-solved walks/s on one MI355X.

Workload (the headline configuration): Johansen model 4, dim 12, T = 10,000
steps, 1,000,000 walks per step (seeds 1..1,000,000, Philox stream 0).  One
"step" is one pass of the walk path over that batch: device normals ->
one-pass moments -> model algebra -> FP64 eigen solve, with the seeds already
resident in HBM and the eigenvalues left in HBM.

The headline path is the mixed-precision one (moment Gram on the matrix cores,
fp16 hi/lo operands, FP32 accumulate flushed to FP64 every 64 steps; FP64
state, model algebra and eigen solve).  The plain FP64 path is timed on the
same batch in the same run, and the two outputs are compared walk by walk.

Prints one JSON line with walks/s, the roofline figure of the dominant kernel
(timed with HIP events on the stream it runs on), two rates of the normal
generator on the same GPU and the CPU baseline (the C restatement of the
reference walk in oracle/, on a bounded sample of the same workload, on this
host's cores).  The generator rates: roofline.rng_normals_per_s is the
generator alone (bin/jne_rng_rate: every normal drawn, nothing stored), the
ceiling of a walk path that draws the same normals, and frac_of_rng is the
moment kernel's normals/s against it; roofline.stream_normals_per_s is the
normals stream of jne_walk_normals_device, which writes 4 B per normal to HBM
and is bound by those writes (frac_of_stream against it).
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
# MI355X published dense peaks, TFLOP/s
PEAK_TFLOPS = {"fp64": 78.6, "mfma_fp16x2": 2500.0, "mfma_bf16x2": 2500.0, "mfma_bf16x3": 2500.0,
               "mfma_fp32": 157.3, "fp32_kahan": 157.3}
DTYPE = {"fp64": "f64", "mfma_fp16x2": "fp16 hi/lo MFMA (fp32 accumulate) + f64",
         "mfma_bf16x2": "bf16 hi/lo MFMA (fp32 accumulate) + f64",
         "mfma_bf16x3": "bf16 hi/mid/lo MFMA (fp32 accumulate) + f64", "mfma_fp32": "fp32 MFMA + f64",
         "fp32_kahan": "fp32 products, Kahan-compensated fp32 sums + f64 state"}
# MFMA operations issued per 16-step chunk of one walk (32x32x16 products)
MFMA_PER_CHUNK = {"mfma_fp16x2": 3, "mfma_bf16x2": 3, "mfma_bf16x3": 6}


def algorithmic_flops_per_walk(model: int, dim: int, steps: int) -> int:
    """2*T*nF*(nF+dim): the S11 and S01 contractions as the reference forms
    them (src/johansen_statistics.rs:32-33)."""
    nf = dim + 1 if model in (1, 3) else dim
    return 2 * steps * nf * (nf + dim)


def relative_floor(got, ref, floor=1e-5):
    """|d_i| / |rho_i|, with rho_max for entries below floor * rho_max."""
    import numpy as np

    rmax = np.max(np.abs(ref), axis=-1, keepdims=True)
    den = np.where(np.abs(ref) >= floor * rmax, np.abs(ref), rmax)
    return np.abs(got - ref) / den


def host_cpu_info():
    """The host's CPU model (lscpu "Model name"), logical CPUs (nproc), the CPUs
    this process may run on (affinity), its cgroup CPU quota and OMP_NUM_THREADS
    (the job's CPU share where the host is shared), and the threads the CPU
    baseline uses: every CPU of that share."""
    import subprocess

    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        info["model_name"] = next((l.split(":", 1)[1].strip() for l in out.splitlines()
                                   if l.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        info["model_name"] = None
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    omp = os.environ.get("OMP_NUM_THREADS")
    info["omp_num_threads"] = int(omp) if omp and omp.isdigit() else None
    share = info["affinity_cpus"]
    if quota:
        share = min(share, max(1, int(quota)))
    if info["omp_num_threads"]:
        share = min(share, info["omp_num_threads"])
    info["cpu_share"] = share
    return info


def cpu_baseline(model, dim, steps, threads, target_s=15.0):
    import numpy as np

    import oracle

    # calibrate on a small sample (16 walks per thread, so that thread start-up
    # does not dominate it), then time ~target_s of work
    cal_n = 16 * threads
    t0 = time.perf_counter()
    oracle.walks_by_seed_c(model, dim, steps, np.arange(1, cal_n + 1, dtype=np.uint32), threads=threads)
    rate = cal_n / (time.perf_counter() - t0)
    n = int(min(max(rate * target_s, 4 * threads), 200_000))
    seeds = np.arange(1, n + 1, dtype=np.uint32)
    t0 = time.perf_counter()
    eig = oracle.walks_by_seed_c(model, dim, steps, seeds, threads=threads)
    dt = time.perf_counter() - t0
    return n / dt, n, dt, eig


def serving_requests(n, rng):
    """Critical-value queries: model U{0..4}, dim U{1..12}, T log-uniform in
    [100, 10,000], runs log-uniform in [10^4, 10^6], each with its own seed."""
    out = []
    for _ in range(n):
        out.append((int(rng.integers(0, 5)), int(rng.integers(1, 13)),
                    int(round(10 ** rng.uniform(2.0, 4.0))), int(round(10 ** rng.uniform(4.0, 6.0))),
                    int(rng.integers(1, 2**31))))
    return out


def serve_at_rate(server, reqs, rate, precision, rng):
    """Submits reqs with Poisson arrivals at `rate` per second; returns
    per-query latencies (ms, submit to completion) and the wall time from the
    first submission to the last completion."""
    import numpy as np

    gaps = rng.exponential(1.0 / rate, len(reqs))
    t0 = time.perf_counter()
    due = t0
    ids = []
    for gap, r in zip(gaps, reqs):
        due += gap
        wait = due - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        ids.append(server.submit(*r, precision=precision))
    lat = np.array([server.wait(i)[1] for i in ids])
    return lat, time.perf_counter() - t0


# query size classes by the normals a query draws (runs x T x dim): the
# largest class holds the 10^6-walk, dim-12, long queries (up to 1.2e11)
SIZE_CLASSES = [("normals<1e8", 0.0, 1e8), ("1e8<=normals<1e9", 1e8, 1e9), ("1e9<=normals<1e10", 1e9, 1e10),
                ("normals>=1e10", 1e10, float("inf"))]


def latency_by_class(reqs, lat):
    import numpy as np

    size = np.array([r[1] * r[2] * r[3] for r in reqs], dtype=np.float64)
    out = {}
    for name, lo, hi in SIZE_CLASSES:
        sel = lat[(size >= lo) & (size < hi)]
        out[name] = {"queries": int(len(sel)),
                     "p50_ms": float(np.percentile(sel, 50)) if len(sel) else None,
                     "p99_ms": float(np.percentile(sel, 99)) if len(sel) else None}
    return out


def serving_sweep(pkg, rates, n_req, precision, seed=2024, lanes=4):
    import numpy as np

    rng = np.random.default_rng(seed)
    reqs = serving_requests(n_req, rng)
    walk_steps = sum(r[2] * r[3] for r in reqs)
    walks = sum(r[3] for r in reqs)
    out = {"precision": precision, "requests_per_rate": n_req,
           "mix": "model U{0..4}, dim U{1..12}, T log-U[100,1e4], runs log-U[1e4,1e6], own seed",
           "walks_per_rate": walks, "rates": []}
    with pkg.Server(0, hbm_budget_bytes=8 << 30, lanes=lanes) as server:
        server.wait(server.submit(4, 12, 1000, 70000, 1, precision=precision))  # warm the lanes
        for rate in rates:
            lat, wall = serve_at_rate(server, reqs, rate, precision, np.random.default_rng(seed + int(rate)))
            out["rates"].append({"offered_req_per_s": rate, "p50_ms": float(np.percentile(lat, 50)),
                                 "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
                                 "walks_per_s": walks / wall, "walk_steps_per_s": walk_steps / wall,
                                 "achieved_req_per_s": len(reqs) / wall,
                                 "by_size_class": latency_by_class(reqs, lat)})
            print(f"bench: served {len(reqs)} queries at {rate:g} req/s ({precision}) in {wall:.1f} s",
                  file=sys.stderr, flush=True)
    return out


def precision_ladder(pkg, eng, model, dim, T, n, oracle_walks=64):
    """Each reduced precision against the FP64 path on walks 1..n: per-eigenvalue
    max/mean relative error (floor rule) and the shift of the trace and
    max-eigenvalue quantiles; the FP64 path itself against the CPU oracle on
    the first oracle_walks walks."""
    import numpy as np

    import oracle

    seeds = np.arange(1, n + 1, dtype=np.uint32)
    ref = eng.eigenvalues_by_seed(model, dim, T, seeds)
    cpu = oracle.walks_by_seed_c(model, dim, T, seeds[:oracle_walks], threads=min(16, os.cpu_count() or 1))
    anchor = np.abs(ref[:oracle_walks] - cpu) / (np.abs(cpu) + 1e-13 * cpu.max(axis=1, keepdims=True))
    ps = [0.90, 0.95, 0.99]
    rng = np.random.default_rng(0)

    def boot_se(x):
        q = np.array([np.quantile(x[rng.integers(0, len(x), len(x))], ps) for _ in range(200)])
        return q.std(axis=0)

    stats_ref = {"trace": ref.sum(1), "maxeig": ref.max(1)}
    out = {"workload": f"model {model}, dim {dim}, T={T:,}, walks 1..{n:,}",
           "fp64_vs_cpu_oracle_max_rel": float(anchor.max()),
           "mc_standard_error": {k: (boot_se(v) / np.quantile(v, ps)).tolist() for k, v in stats_ref.items()},
           "precisions": {}}
    for name in ["mfma_bf16x2", "mfma_fp16x2", "mfma_bf16x3", "mfma_fp32", "fp32_kahan"]:
        got = eng.eigenvalues_by_seed(model, dim, T, seeds, precision=pkg.PRECISIONS[name])
        rel = relative_floor(got, ref)
        row = {"max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
               "per_eigenvalue_max_rel": rel.max(axis=0).tolist(),
               "per_eigenvalue_mean_rel": rel.mean(axis=0).tolist()}
        for k, f in [("trace", lambda e: e.sum(1)), ("maxeig", lambda e: e.max(1))]:
            qa, qb = np.quantile(f(got), ps), np.quantile(stats_ref[k], ps)
            row[f"{k}_quantile_rel_shift"] = ((qa - qb) / qb).tolist()
        out["precisions"][name] = row
    return out


def time_path(pkg, eng, prec_name, model, dim, T, seeds, out, n, steps, warmup, stream, torch):
    """Runs warmup + steps passes; returns (walks/s, s/step, moments kernel (ms, launches, walks), eigen ms)."""
    prec = pkg.PRECISIONS[prec_name]

    def step():
        eng.eigenvalues_by_seed_device(model, dim, T, seeds.data_ptr(), n, out.data_ptr(), precision=prec,
                                       hip_stream=stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    eng.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mom = eng.profile_read(pkg.KERNEL_MOMENTS)
    eig = eng.profile_read(pkg.KERNEL_EIGEN)
    eng.profile_enable(False)
    return steps * n / (t1 - t0), (t1 - t0) / steps, mom, eig[0]


def unit_rates_for(key, walks_per_launch, avg_ms):
    """Achieved TFLOP/s and fraction of peak of the matrix cores, the FP32 VALU
    and the FP64 VALU: per-walk FLOP by unit from the committed counter runs
    (profiles/unit_rates.json) over this run's kernel time."""
    path = os.path.join(ROOT, "profiles", "unit_rates.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(key)
    if not rec:
        return None
    per_s = walks_per_launch / (avg_ms * 1e-3) / 1e12
    out = {}
    for unit, field, peak in [("mfma_f16", "mfma_f16_flop_per_walk", 2500.0),
                              ("fp32_valu", "fp32_valu_flop_per_walk", 157.3),
                              ("fp64_valu", "fp64_valu_flop_per_walk", 78.6)]:
        out[unit] = {"tflops": rec[field] * per_s, "peak": peak, "frac": rec[field] * per_s / peak}
    if rec.get("mfma_busy_cycles"):
        out["mfma_coexec_frac"] = rec["mfma_coexec_cycles"] / rec["mfma_busy_cycles"]
    out["source"] = "FLOP per walk from rocprofv3 counters (profiles/unit_rates.json) / HIP-event kernel time"
    return out


def stream_rate(pkg, eng, dim, T, torch, stream, walks=16384, reps=5):
    """Normals/s of the random stream alone (jne_walk_normals_device: Philox +
    Box-Muller written to HBM, 4 B per normal) at the workload's shape: the
    ceiling of any walk path that draws the same normals."""
    dev = torch.device("cuda:0")
    seeds = torch.arange(1, walks + 1, dtype=torch.int32, device=dev)
    buf = torch.empty(walks * T * dim, dtype=torch.float32, device=dev)
    eng.walk_normals_device(dim, T, seeds.data_ptr(), walks, buf.data_ptr(), hip_stream=stream)
    torch.cuda.synchronize()
    eng.profile_enable(True)
    for _ in range(reps):
        eng.walk_normals_device(dim, T, seeds.data_ptr(), walks, buf.data_ptr(), hip_stream=stream)
    ms, launches, _ = eng.profile_read(pkg.KERNEL_NORMALS)
    eng.profile_enable(False)
    del buf
    avg_ms = ms / max(launches, 1)
    return walks * T * dim / (avg_ms * 1e-3), avg_ms


def rng_rate(timeout_s=120):
    """Normals/s of the generator alone, from bin/jne_rng_rate (a child process
    on the same GPU): the JSON line it prints, or None if it is not built."""
    import subprocess

    exe = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "jne_rng_rate")
    if not os.path.exists(exe):
        return None
    res = subprocess.run([exe], capture_output=True, text=True, timeout=timeout_s, check=True)
    return json.loads(res.stdout.strip().splitlines()[-1])


def valu_issue_for(key):
    """VALU issue of this moment kernel and of the generator alone
    (bin/jne_rng_rate) from the committed counter runs (profiles/valu_busy.json:
    rocprofv3 SQ_ACTIVE_INST_VALU, SQ_BUSY_CU_CYCLES): VALU wave instructions
    issued per CU-cycle and per normal drawn.  frac_of_rng factors into
    issue_rate_vs_rng (how fast the kernel issues against the generator alone)
    times insts_vs_rng (the generator's instructions per normal over the
    kernel's: the share of the kernel's issue that is the normals)."""
    path = os.path.join(ROOT, "profiles", "valu_busy.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    k, g = rec.get(key), rec.get("rng_rate_kernel")
    if not k or not g:
        return None
    return {"kernel_valu_insts_per_cu_cycle": k["valu_insts_per_cu_cycle"],
            "rng_valu_insts_per_cu_cycle": g["valu_insts_per_cu_cycle"],
            "kernel_valu_insts_per_normal": k["valu_insts_per_normal"],
            "rng_valu_insts_per_normal": g["valu_insts_per_normal"],
            "issue_rate_vs_rng": k["valu_insts_per_cu_cycle"] / g["valu_insts_per_cu_cycle"],
            "insts_vs_rng": g["valu_insts_per_normal"] / k["valu_insts_per_normal"],
            "counter_frac_of_rng": (k["valu_insts_per_cu_cycle"] / g["valu_insts_per_cu_cycle"])
            * (g["valu_insts_per_normal"] / k["valu_insts_per_normal"]),
            "source": "rocprofv3 --pmc SQ_ACTIVE_INST_VALU SQ_BUSY_CU_CYCLES (profiles/valu_busy.json, "
                      "commands in profiles/valu_busy.md)"}


def kernel_name(precision: str, dim: int) -> str:
    if precision == "fp64":
        return f"walk_moments_fp64<{dim}>"
    if precision == "fp32_kahan":
        return f"walk_moments_mfma<{dim}, fp32_kahan>"
    return f"walk_moments_mfma<{dim}, {precision}>"


def moment_roofline(precision, model, dim, T, mom, wps):
    """The roofline object of the moment kernel: algorithmic FLOP per launch
    over its HIP-event launch time against the unit's peak; HBM bytes per
    launch from the committed counter runs (per-walk, copy-calibrated) times
    this run's walks per launch; the per-unit rates from the committed
    instruction counts."""
    mom_ms, mom_launches, mom_walks = mom
    fpw = algorithmic_flops_per_walk(model, dim, T)
    avg_ms = mom_ms / max(mom_launches, 1)
    walks_per_launch = mom_walks / max(mom_launches, 1)
    achieved = fpw * walks_per_launch / (avg_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[precision]
    key = f"walk_moments_{precision}_m{model}_d{dim}_T{T}"
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            rec = json.load(f).get(key)
        if rec:
            traffic = (rec["fetch_bytes_per_walk"] + rec["write_bytes_per_walk"]) * walks_per_launch
    roofline = {
        # both moment passes are bound by VALU instruction issue (the normal
        # generator is most of it); achieved/peak/frac are the algorithmic
        # contraction against the peak of the unit the Gram runs on, and
        # valu_issue gives the counter-measured VALU busy fractions
        "bound": "valu",
        "achieved": achieved,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "gram_unit": "mfma_f16" if precision.startswith("mfma") and precision != "mfma_fp32" else
                     ("mfma_f32" if precision == "mfma_fp32" else ("fp32_valu" if precision == "fp32_kahan"
                                                                   else "fp64_valu")),
        "traffic": traffic,
        "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)",
        "algorithmic_bytes_per_launch": (4 + 8 * (2 * dim * dim + 6 * dim)) * walks_per_launch,
        "kernel": kernel_name(precision, dim),
        "algorithmic_flop_per_walk": fpw,
        "avg_launch_ms": avg_ms,
        "walks_per_launch": walks_per_launch,
        "normals_per_s": wps * T * dim,
        "unit_rates": unit_rates_for(key, walks_per_launch, avg_ms),
        "valu_issue": valu_issue_for(key),
    }
    if precision in MFMA_PER_CHUNK:
        issued = MFMA_PER_CHUNK[precision] * 2 * 32 * 32 * 16 * ((T + 15) // 16)
        roofline["issued_mfma_tflops"] = issued * walks_per_launch / (avg_ms * 1e-3) / 1e12
    return roofline


def cpu_baseline_block(model, dim, T, threads, head, fp64_head, k, host):
    """cpu_baseline object: the C port on `threads` host threads over a bounded
    sample, and its agreement with the GPU on the first walks."""
    import numpy as np

    rate, cn, cdt, ceig = cpu_baseline(model, dim, T, threads)
    block = {
        "value": rate,
        "unit": "walks/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{cn:,} walks (seeds 1..{cn:,}) of the same workload in {cdt:.1f} s; C restatement of "
                  f"the reference CPU walk (oracle/jne_oracle.c), one walk per thread, on every CPU of "
                  f"this job's share of the host",
        "host": host,
    }
    m = min(64, cn, k)
    if fp64_head is not None:
        d = np.abs(fp64_head[:m] - ceig[:m])
        block["fp64_gpu_vs_cpu_max_rel_diff_first_64"] = float(
            (d / (np.abs(ceig[:m]) + 1e-13 * ceig[:m].max(axis=1, keepdims=True))).max())
    block["headline_vs_cpu_max_rel_diff_first_64"] = float(relative_floor(head[:m], ceig[:m]).max())
    return block


def config3_sweep(pkg, eng, runs=100_000, T=10_000, dims=range(1, 13), models=(0, 1, 2, 3, 4)):
    """The reference's production sweep (models 0-4 x dims 1-12, T = 10^4) at `runs` runs per
    pair, streamed to .dat files: the quiet CLI (the models of a dim share one moment pass per walk,
    jne_run_simulation_models) against the same sweep model by model (one jne_run_simulation-
    equivalent call per pair), with the 60 files compared byte for byte."""
    import filecmp
    import subprocess
    import tempfile

    cli = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "johansen-null-eigenspectra")
    with tempfile.TemporaryDirectory() as tmp:
        fused_dir, single_dir = os.path.join(tmp, "fused"), os.path.join(tmp, "single")
        os.makedirs(fused_dir)
        os.makedirs(os.path.join(single_dir, "data"))
        t0 = time.perf_counter()
        subprocess.run([cli, "--runs", f"{runs:,}", "--steps", f"{T:,}", "--dim-start", str(min(dims)),
                        "--dim-end", str(max(dims)), "--precision", "auto", "--quiet"], cwd=fused_dir, check=True,
                       capture_output=True, timeout=600)
        fused_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for dim in dims:
            for m in models:
                eng.run_simulation_models([m], dim, T, runs,
                                          [os.path.join(single_dir, "data", f"eigenvalues_model{m}_dim{dim}_steps{T}.dat")],
                                          precision=pkg.PRECISIONS["auto"])
        single_s = time.perf_counter() - t0
        names = sorted(os.listdir(os.path.join(fused_dir, "data")))
        same = len(names) == len(dims) * len(models) and all(
            filecmp.cmp(os.path.join(fused_dir, "data", nm), os.path.join(single_dir, "data", nm), shallow=False)
            for nm in names)
        nbytes = sum(os.path.getsize(os.path.join(fused_dir, "data", nm)) for nm in names)
    walks = runs * len(dims) * len(models)
    return {"workload": f"production sweep: models 0-4 x dims {min(dims)}-{max(dims)}, T={T:,}, {runs:,} runs per "
                        f"pair, auto precision, streamed to .dat", "walks": walks, "files": len(names),
            "bytes": nbytes, "fused_cli_s": fused_s, "fused_walks_per_s": walks / fused_s,
            "model_by_model_s": single_s, "model_by_model_walks_per_s": walks / single_s,
            "files_identical": same}


def config1_full(pkg, eng, threads, host, model=0, dim=2, T=1000, runs=100_000):
    """The reference's CPU-native configuration in full (model 0, dim 2,
    T = 1,000, 100,000 runs; src/data_storage/parallel_compute.rs:14-41): the
    C port on the CPU share, and on the GPU both the batch path (seeds in,
    eigenvalues out, host buffers) and the streamed run that writes
    eigenvalues_model0_dim2_steps1000.dat (jne_run_simulation, from an empty
    directory, end to end); the GPU and CPU results are compared walk by walk."""
    import tempfile

    import numpy as np

    import oracle

    seeds = np.arange(1, runs + 1, dtype=np.uint32)
    t0 = time.perf_counter()
    cpu = oracle.walks_by_seed_c(model, dim, T, seeds, threads=threads)
    cpu_s = time.perf_counter() - t0
    eng.eigenvalues_by_seed(model, dim, T, seeds[:1000])  # warm the kernels and the buffers
    t0 = time.perf_counter()
    gpu = eng.eigenvalues_by_seed(model, dim, T, seeds)
    gpu_s = time.perf_counter() - t0
    with tempfile.TemporaryDirectory() as tmp:
        sim = pkg.EigenvalueSimulation(model, dim, T, runs, engine=eng, data_dir=tmp)
        t0 = time.perf_counter()
        sim.run_simulation_quiet()
        sim_s = time.perf_counter() - t0
        path = sim.get_filename()
        dat_bytes = os.path.getsize(path)
        fs, fe, *_ = pkg.read_dat(path)
        same = bool(np.array_equal(fe, gpu[fs.astype(np.int64) - 1]))
    # the FP64 parity bar, |d| <= 1e-9 |rho| + 1e-13 rho_max: a ratio <= 1 meets it.  The smallest
    # eigenvalue of some walks is ~1e-10 rho_max, where the problem's own conditioning leaves
    # ~1e-7 relative between any two FP64 solvers (LAPACK dggev and the C oracle included)
    d = np.abs(gpu - cpu)
    cmax = cpu.max(axis=1, keepdims=True)
    bar = d / (1e-9 * np.abs(cpu) + 1e-13 * cmax)
    return {"workload": f"Johansen model {model}, dim {dim}, T={T:,} steps, {runs:,} runs (seeds 1..{runs:,}), "
                        f"FP64 path", "cpu_port_walks_per_s": runs / cpu_s, "cpu_s": cpu_s, "cpu_threads": threads,
            "cpu_host": host.get("model_name"), "gpu_batch_walks_per_s": runs / gpu_s, "gpu_batch_s": gpu_s,
            "gpu_streamed_dat_s": sim_s, "gpu_streamed_walks_per_s": runs / sim_s, "dat_file_bytes": dat_bytes,
            "dat_equals_batch": same, "gpu_vs_cpu_max_normwise_diff": float((d / cmax).max()),
            "gpu_vs_cpu_max_diff_over_fp64_bar": float(bar.max())}


def wide_dims(pkg, eng, torch, stream, dims=(20, 32), model=4, T=10_000, n=65_536, steps=3):
    """The one-wave-per-walk FP64 path above dim 12 (the reference's operator
    takes any dim, src/johansen_statistics.rs:59-85): walks/s of the whole path
    and the time per launch of its moment and eigen kernels, at model 4,
    T = 10^4, n walks per launch."""
    dev = torch.device("cuda:0")
    seeds = torch.arange(1, n + 1, dtype=torch.int32, device=dev)
    rows = []
    for dim in dims:
        out = torch.empty(n * pkg.num_eigenvalues(model, dim), dtype=torch.float64, device=dev)
        wps, _sps, mom, eig_ms = time_path(pkg, eng, "fp64", model, dim, T, seeds, out, n, steps, 1, stream, torch)
        assert bool(torch.isfinite(out).all()), "non-finite eigenvalues"
        launches = max(mom[1], 1)
        rows.append({"dim": dim, "walks_per_s": wps, "moments_ms_per_launch": mom[0] / launches,
                     "eigen_ms_per_launch": eig_ms / launches})
    return {"workload": f"Johansen model {model}, dims {', '.join(map(str, dims))}, T={T:,} steps, {n:,} walks per "
                        f"launch, FP64 path with one wave per walk", "dims": rows}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--walks", type=int, default=1_000_000, help="walks per step")
    ap.add_argument("--model", type=int, default=4)
    ap.add_argument("--dim", type=int, default=12)
    ap.add_argument("--T", type=int, default=10_000, help="time steps per walk")
    ap.add_argument("--precision", default="mfma_fp16x2", help="headline path")
    ap.add_argument("--compare", default="fp64", help="path timed beside it ('' to skip)")
    ap.add_argument("--check-walks", type=int, default=10_000, help="walks compared between the two paths")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rng-rate", action="store_true", help="skip the generator-alone rate (bin/jne_rng_rate)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = every CPU of this job's share of the host (affinity, cgroup quota, OMP_NUM_THREADS)")
    ap.add_argument("--no-config3", action="store_true",
                    help="skip the production sweep sample (models 0-4 x dims 1-12, T=10^4, 10^5 runs per pair)")
    ap.add_argument("--no-wide", action="store_true", help="skip the dims-20 and 32 timing of the wide-dim path")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the plain-CPU configuration (model 0, dim 2, T=1,000, 100,000 runs) on CPU and GPU")
    ap.add_argument("--serve-rates", default="10,30,100,300,1000",
                    help="Poisson query rates (req/s) for the serving sweep ('' to skip)")
    ap.add_argument("--serve-requests", type=int, default=1000, help="queries per rate")
    ap.add_argument("--serve-precision", default="auto", help="precision of the served queries")
    ap.add_argument("--serve-lanes", type=int, default=4, help="launch lanes of the query server")
    ap.add_argument("--no-serve-fp64", action="store_true",
                    help="skip the FP64 serving run at the highest rate (the saturation comparison)")
    ap.add_argument("--ladder", type=int, default=0,
                    help="print the precision ladder on this many walks instead of timing")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration under rocprofv3)")
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
    model, dim, T, n = args.model, args.dim, args.T, args.walks
    nf = pkg.num_eigenvalues(model, dim)

    torch.cuda.init()
    dev = torch.device("cuda:0")
    eng = pkg.Engine(0)
    if args.ladder:
        print(json.dumps(precision_ladder(pkg, eng, model, dim, T, args.ladder)), flush=True)
        eng.close()
        return 0
    seeds = torch.arange(1, n + 1, dtype=torch.int32, device=dev)  # uint32 bit pattern
    out = torch.empty(n * nf, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    if args.copy_probe_mb:
        src = torch.empty(args.copy_probe_mb << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    wps, sps, mom, eig_ms = time_path(pkg, eng, args.precision, model, dim, T, seeds, out, n, args.steps,
                                      args.warmup, stream, torch)
    k = min(args.check_walks, n)
    head = out.view(n, nf)[:k].cpu().numpy()
    assert np.all(np.isfinite(head)), "non-finite eigenvalues"

    mom_ms, mom_launches, _ = mom
    roofline = moment_roofline(args.precision, model, dim, T, mom, wps)

    line = {
        "metric": METRIC,
        "value": wps,
        "unit": "walks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
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
        "roofline": roofline,
        "kernels_ms": {"moments": mom_ms, "eigen": eig_ms, "launches": mom_launches},
    }

    srate, sms = stream_rate(pkg, eng, dim, T, torch, stream)
    roofline["stream_normals_per_s"] = srate
    roofline["stream_write_GBps"] = srate * 4 / 1e9
    roofline["frac_of_stream"] = roofline["normals_per_s"] / srate
    rng = None if args.no_rng_rate else rng_rate()
    roofline["rng_normals_per_s"] = rng["normals_per_s"] if rng else None
    roofline["frac_of_rng"] = roofline["normals_per_s"] / rng["normals_per_s"] if rng else None
    roofline["rng_source"] = ("bin/jne_rng_rate: " + json.dumps(rng)) if rng else "not measured in this run"

    if args.compare:
        cwps, csps, cmom, ceig_ms = time_path(pkg, eng, args.compare, model, dim, T, seeds, out, n, args.steps,
                                              args.warmup, stream, torch)
        ref = out.view(n, nf)[:k].cpu().numpy()
        rel = relative_floor(head, ref)
        c_avg = cmom[0] / max(cmom[1], 1)
        line["compare"] = {
            "precision": args.compare,
            "value": cwps,
            "ms_per_step": csps * 1e3,
            "moments_avg_launch_ms": c_avg,
            "unit_rates": unit_rates_for(f"walk_moments_{args.compare}_m{model}_d{dim}_T{T}",
                                         cmom[2] / max(cmom[1], 1), c_avg),
            "valu_issue": valu_issue_for(f"walk_moments_{args.compare}_m{model}_d{dim}_T{T}"),
            "frac_of_rng": cwps * T * dim / rng["normals_per_s"] if rng else None,
            "speedup": wps / cwps,
            "max_rel_diff": float(rel.max()),
            "mean_rel_diff": float(rel.mean()),
            "walks_compared": k,
        }
        fp64_head = ref if args.compare == "fp64" else None
    else:
        fp64_head = head if args.precision == "fp64" else None

    if args.serve_rates:
        rates = [float(r) for r in args.serve_rates.split(",")]
        sv = serving_sweep(pkg, rates, args.serve_requests, args.serve_precision, lanes=args.serve_lanes)
        if not args.no_serve_fp64:
            # the same query sequence on the FP64 path at the highest rate: both saturated
            top = serving_sweep(pkg, [rates[-1]], args.serve_requests, "fp64", lanes=args.serve_lanes)["rates"][0]
            sv["fp64_at_top_rate"] = top
            sv["saturated_mixed_vs_fp64"] = sv["rates"][-1]["walk_steps_per_s"] / top["walk_steps_per_s"]
        line["serving"] = sv

    if not args.no_wide:
        line["wide_dims"] = wide_dims(pkg, eng, torch, stream)
    host = host_cpu_info()
    threads = args.cpu_threads or host["cpu_share"]
    if not args.no_config1:
        line["config1"] = config1_full(pkg, eng, threads, host)
    if not args.no_config3:
        line["config3"] = config3_sweep(pkg, eng)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_block(model, dim, T, threads, head, fp64_head, k, host)
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
