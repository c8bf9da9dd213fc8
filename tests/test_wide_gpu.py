"""Dims above 12 on the GPU (the runtime-dim FP64 path), through the C ABI.

The reference's operator takes any dim (src/johansen_statistics.rs:59-85)
and its CLI checks only dim > 0 (src/cli.rs:201-225).  Dims 13..32 run the
one-wave-per-walk kernels; they are held to the FP64 tolerance against the C
restatement of the reference walk and against the numpy restatement with
LAPACK dggev: |d| <= 1e-9 |rho| + 1e-13 rho_max.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import assert_eigs_close

pytestmark = pytest.mark.gpu

RTOL = 1e-9
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "johansen-null-eigenspectra")


@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dim", [13, 16, 17, 20, 24, 25, 32])  # both ends of each padded class
def test_wide_dims_match_oracle(engine, O, model, dim):
    steps = 301 if dim % 2 else 256  # a partial last round
    seeds = np.array([1, 2, 0xFFFFFFFF], dtype=np.uint32)
    got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    assert got.shape == (3, O.num_eigenvalues(model, dim))
    ref = O.walks_by_seed_c(model, dim, steps, seeds, threads=3)
    assert_eigs_close(got, ref, RTOL, f"model {model} dim {dim} steps {steps} (C oracle)")
    lap = O.walk_eigenvalues_by_seed(model, dim, steps, 2)
    assert_eigs_close(got[1], lap, RTOL, f"model {model} dim {dim} steps {steps} (dggev oracle)")
    assert (np.diff(got, axis=1) <= 0).all()


@pytest.mark.parametrize("model,dim,steps", [(4, 16, 5000), (2, 24, 3001), (3, 13, 10000), (1, 32, 2000)])
def test_wide_dims_long_walks(engine, O, model, dim, steps):
    seeds = np.array([7, 8], dtype=np.uint32)
    got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    ref = O.walks_by_seed_c(model, dim, steps, seeds, threads=2)
    assert_eigs_close(got, ref, RTOL, f"model {model} dim {dim} steps {steps}")


def test_wide_tiny_smallest_eigenvalue(engine, O):
    """Walks whose smallest eigenvalue is 4e-11 .. 2e-9 of the largest (model 4,
    T = 10^4): the Householder + bisection solve keeps to the bar there, where
    the 1e-13 rho_max term dominates (dggev itself is off by ~1e-6 relative on
    such an eigenvalue against a 40-digit solve of the same moments)."""
    for dim, seeds in ((32, [2710, 1180]), (20, [2230, 2274])):
        seeds = np.array(seeds, dtype=np.uint32)
        got = engine.eigenvalues_by_seed(4, dim, 10000, seeds)
        ref = O.walks_by_seed_c(4, dim, 10000, seeds, threads=2)
        assert (np.abs(ref).min(axis=1) < 3e-9 * np.abs(ref).max(axis=1)).all()
        assert_eigs_close(got, ref, RTOL, f"model 4 dim {dim} steps 10000 (C oracle)")
        lap = O.walk_eigenvalues_by_seed(4, dim, 10000, int(seeds[0]))
        assert_eigs_close(got[0], lap, RTOL, f"model 4 dim {dim} steps 10000 (dggev oracle)")


def test_wide_arbitrary_increments_and_from_normals(engine, O):
    rng = np.random.default_rng(13)
    for model in range(5):
        z = rng.standard_normal((2, 97, 14))
        got = engine.eigenvalues_from_normals(model, z)
        ref = np.stack([O.walk_eigenvalues_lapack(model, z[k]) for k in range(2)])
        assert_eigs_close(got, ref, RTOL, f"model {model}")
    # float-representable increments: the from-normals and by-seed entry points agree bit for bit
    seeds = np.arange(1, 9, dtype=np.uint32)
    z = engine.walk_normals(21, 150, seeds).astype(np.float64)
    assert np.array_equal(engine.eigenvalues_from_normals(4, z), engine.eigenvalues_by_seed(4, 21, 150, seeds))


def test_wide_batch_invariance_and_precisions(engine, jne):
    seeds = np.arange(1, 131, dtype=np.uint32)
    full = engine.eigenvalues_by_seed(2, 17, 333, seeds)
    perm = np.random.default_rng(2).permutation(len(seeds))
    assert np.array_equal(full[perm], engine.eigenvalues_by_seed(2, 17, 333, seeds[perm]))
    assert np.array_equal(full[99], engine.eigenvalues_by_seed(2, 17, 333, seeds[99:100])[0])
    # auto runs the FP64 path above dim 12; the reduced precisions cover dims 1..12 only
    assert np.array_equal(full, engine.eigenvalues_by_seed(2, 17, 333, seeds, precision=jne.PRECISIONS["auto"]))
    for prec in ["mfma_fp16x2", "fp32_kahan", "mfma_bf16x3"]:
        with pytest.raises(jne.JneError) as e:
            engine.eigenvalues_by_seed(2, 17, 333, seeds, precision=jne.PRECISIONS[prec])
        assert e.value.code == jne.JNE_ERR_UNSUPPORTED
    with pytest.raises(jne.JneError) as e:
        engine.eigenvalues_by_seed(0, 33, 10, [1])
    assert e.value.code == jne.JNE_ERR_INVALID_ARG


def test_wide_cli_run(jne, engine, tmp_path):
    p = subprocess.run([CLI, "--dim-start", "12", "--dim-end", "14", "--steps", "400", "--runs", "3,000",
                        "--model", "1,4", "--quiet"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0 and p.stdout == "", p.stderr
    for model in (1, 4):
        for dim in (12, 13, 14):
            seeds, eigs, m, d, st = jne.read_dat(tmp_path / "data" / f"eigenvalues_model{model}_dim{dim}_steps400.dat")
            assert (m, d, st) == (model, dim, 400) and sorted(seeds.tolist()) == list(range(1, 3001))
            ref = engine.eigenvalues_by_seed(model, dim, 400, seeds)
            assert np.array_equal(eigs, ref)
