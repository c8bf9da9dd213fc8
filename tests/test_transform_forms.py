"""The device normal transform's operation forms equal the oracle's on every
input of their ranges (CPU).

jne_rng.h forms 2 ln(u1) at twice the scale with a bit-arithmetic exponent
split and turns the Box-Muller pair by sign flips instead of a select; the
GPU stream tests compare sampled normals bit for bit, and this test covers
the boundary inputs a sample misses (u1 = 1, exponent-split edges, angles
where the turned cosine or sine is zero): transform_forms.c restates the
device forms and compares them with jo_normal_pair / ln_poly of
oracle/jne_oracle.c over every u1 and a dense cover of the angle words.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_device_transform_forms_equal_oracle(tmp_path):
    exe = str(tmp_path / "transform_forms")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-I", ORACLE,
                    os.path.join(HERE, "transform_forms.c"), "-o", exe, "-lm"], check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert " 0/" in res.stdout and res.stdout.count(" 0/") == 2, res.stdout
