"""bench.py on the CPU: the one-GPU contract and the helpers that shape its
JSON line (no GPU work); and bin/jne_rng_rate, the generator-alone rate that
bench.py reports beside the moment kernel (argument checks on the CPU, its
arithmetic against the oracle's stream on the GPU)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and r.stdout == ""
    assert "one GPU" in r.stderr or "--gpus 1" in r.stderr


def test_relative_floor():
    ref = np.array([[10.0, 1.0, 1e-9]])
    got = np.array([[10.0 + 1e-4, 1.0 - 1e-6, 2e-9]])
    rel = bench.relative_floor(got, ref)
    # entries below 1e-5 * rho_max are measured against rho_max
    assert np.allclose(rel, [[1e-5, 1e-6, 1e-9 / 10.0]])


def test_unit_rates_from_the_committed_counters():
    with open(os.path.join(ROOT, "profiles", "unit_rates.json")) as f:
        rec = json.load(f)["walk_moments_mfma_fp16x2_m4_d12_T10000"]
    u = bench.unit_rates_for("walk_moments_mfma_fp16x2_m4_d12_T10000", 1000.0, 1.0)  # 1000 walks per ms
    assert abs(u["mfma_f16"]["tflops"] - rec["mfma_f16_flop_per_walk"] * 1e6 / 1e12) < 1e-9
    assert u["fp64_valu"]["peak"] == 78.6 and u["fp32_valu"]["peak"] == 157.3
    assert 0.0 < u["mfma_coexec_frac"] <= 1.0
    assert bench.unit_rates_for("walk_moments_none", 1.0, 1.0) is None


def test_host_cpu_info_names_the_share_it_uses(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    h = bench.host_cpu_info()
    assert h["nproc"] == os.cpu_count() and h["affinity_cpus"] >= 1
    assert h["omp_num_threads"] == 3 and 1 <= h["cpu_share"] <= 3
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_cpu_info()["cpu_share"] <= len(os.sched_getaffinity(0))


def test_latency_by_size_class():
    # (model, dim, T, runs, seed): normals = dim * T * runs
    reqs = [(0, 1, 100, 10_000, 1), (4, 12, 10_000, 1_000_000, 2), (4, 12, 10_000, 900_000, 3),
            (2, 5, 1000, 100_000, 4)]
    lat = np.array([1.0, 300.0, 200.0, 10.0])
    c = bench.latency_by_class(reqs, lat)
    assert c["normals<1e8"] == {"queries": 1, "p50_ms": 1.0, "p99_ms": 1.0}
    assert c["normals>=1e10"]["queries"] == 2 and c["normals>=1e10"]["p50_ms"] == 250.0
    assert c["1e8<=normals<1e9"]["queries"] == 1 and c["1e9<=normals<1e10"]["queries"] == 0
    assert c["1e9<=normals<1e10"]["p99_ms"] is None


def test_algorithmic_flops():
    # 2 T nF (nF + d): the reference's S11 and S01 contractions
    assert bench.algorithmic_flops_per_walk(4, 12, 10000) == 2 * 10000 * 12 * 24
    assert bench.algorithmic_flops_per_walk(1, 3, 100) == 2 * 100 * 4 * 7


RNG_RATE = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "jne_rng_rate")
needs_tool = pytest.mark.skipif(not os.path.exists(RNG_RATE), reason="bin/jne_rng_rate is not built")


@needs_tool
@pytest.mark.parametrize("args", [["--calls", "4"], ["--calls", "0"], ["--blocks", "0"], ["--reps", "101"],
                                  ["--bogus", "1"]])
def test_rng_rate_tool_rejects_bad_arguments(args):
    # checked before any HIP call: exit 2 with a message, no GPU needed
    res = subprocess.run([RNG_RATE] + args, capture_output=True, text=True, timeout=60)
    assert res.returncode == 2
    assert "jne_rng_rate" in res.stderr


@needs_tool
@pytest.mark.gpu
def test_rng_rate_tool_draws_the_walk_normals(O):
    """Thread t of the tool draws the normals of walk t + 1 (stream 0) in
    order and sums them in FP32, (v0 + v1) + (v2 + v3) per Philox call; the
    printed total (a double sum over threads) is replayed from the oracle's
    stream with the same FP32 operations."""
    blocks, calls = 2, 30
    res = subprocess.run([RNG_RATE, "--blocks", str(blocks), "--calls", str(calls), "--reps", "2"],
                         capture_output=True, text=True, timeout=120, check=True)
    out = json.loads(res.stdout.strip().splitlines()[-1])
    threads = blocks * 256
    assert out["threads"] == threads and out["calls_per_thread"] == calls
    assert out["normals_per_launch"] == 4 * threads * calls
    assert out["normals_per_s"] > 0
    total = 0.0
    for t in range(threads):
        v = O.walk_normals(t + 1, 0, 1, 4 * calls).reshape(calls, 4).astype(np.float32)
        acc = np.float32(0.0)
        for q in range(calls):
            acc = np.float32(acc + np.float32(np.float32(v[q, 0] + v[q, 1]) + np.float32(v[q, 2] + v[q, 3])))
        total += float(acc)
    assert abs(out["sum_of_outputs"] - total) <= 5e-7, (out["sum_of_outputs"], total)
