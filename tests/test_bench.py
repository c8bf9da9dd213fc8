"""bench.py on the CPU: the one-GPU contract and the helpers that shape its
JSON line (no GPU work)."""
import json
import os
import subprocess
import sys

import numpy as np

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


def test_algorithmic_flops():
    # 2 T nF (nF + d): the reference's S11 and S01 contractions
    assert bench.algorithmic_flops_per_walk(4, 12, 10000) == 2 * 10000 * 12 * 24
    assert bench.algorithmic_flops_per_walk(1, 3, 100) == 2 * 100 * 4 * 7
