"""The reference production sweep at its own length: models 0-4 x dims 1..12,
T = 10,000 steps (src/cli.rs:23-35 defaults, src/main.rs:85-127), on the GPU.

For every (model, dim) pair:

* the FP64 path against the C restatement of the reference walk on four
  seeds spread over the u32 range, and against the numpy restatement with
  LAPACK dggev (the reference's own F construction and solve,
  src/johansen_statistics.rs:24-197) on one seed: |d| <= 1e-9 |rho| +
  1e-13 rho_max;
* the mixed (MFMA fp16 hi/lo) pass, the other 1e-5 rungs and the compensated
  FP32 VALU pass against the FP64 path on 256 walks: batch-normwise
  max|d| <= 1e-5 max(rho_max, median rho_max), and per eigenvalue
  |d| <= 1e-5 |rho| with the structural-zero floor (1e-5 rho_max for entries
  below it) -- the stated mixed-path tolerance;
* `auto` equal bit for bit to the path it selects.

Then the streamed .dat path at T = 10,000 (model 3, dim 12, 20,000 runs,
interrupted mid-record and resumed) walk by walk against the batch path, and
the command-line tool over all 60 pairs at T = 10,000, each file against the
batch path.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import assert_eigs_close

pytestmark = pytest.mark.gpu

T = 10_000
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "johansen-null-eigenspectra")
ORACLE_SEEDS = np.array([1, 2, 2_654_435_761, 0xFFFFFFFF], dtype=np.uint32)
N_MIXED = 256
TOL_MIXED = {"mfma_fp16x2": 1e-5, "mfma_bf16x3": 2e-5, "mfma_fp32": 1e-5, "fp32_kahan": 1e-5}
PAIRS = [(m, d) for m in range(5) for d in range(1, 13)]


def normwise(got, ref):
    rmax = np.max(np.abs(ref), axis=-1)
    return np.max(np.abs(got - ref), axis=-1) / np.maximum(rmax, np.median(rmax))


def relative_floor(got, ref, floor=1e-5):
    rmax = np.max(np.abs(ref), axis=-1, keepdims=True)
    den = np.where(np.abs(ref) >= floor * rmax, np.abs(ref), rmax)
    return np.abs(got - ref) / den


@pytest.mark.parametrize("model,dim", PAIRS)
def test_sweep_fp64_matches_oracle(engine, O, model, dim):
    got = engine.eigenvalues_by_seed(model, dim, T, ORACLE_SEEDS)
    ref = O.walks_by_seed_c(model, dim, T, ORACLE_SEEDS, threads=4)
    assert_eigs_close(got, ref, 1e-9, f"model {model} dim {dim} T {T} (C oracle)")
    lap = O.walk_eigenvalues_by_seed(model, dim, T, int(ORACLE_SEEDS[2]))
    assert_eigs_close(got[2], lap, 1e-9, f"model {model} dim {dim} T {T} (dggev oracle)")


@pytest.fixture(scope="module")
def fp64_sweep(engine):
    seeds = np.arange(1, N_MIXED + 1, dtype=np.uint32)
    return {(m, d): engine.eigenvalues_by_seed(m, d, T, seeds) for m, d in PAIRS}


@pytest.mark.parametrize("model,dim", PAIRS)
def test_sweep_reduced_precisions_match_fp64(engine, jne, fp64_sweep, model, dim):
    seeds = np.arange(1, N_MIXED + 1, dtype=np.uint32)
    ref = fp64_sweep[(model, dim)]
    for prec, tol in TOL_MIXED.items():
        got = engine.eigenvalues_by_seed(model, dim, T, seeds, precision=jne.PRECISIONS[prec])
        assert np.isfinite(got).all()
        nw = normwise(got, ref).max()
        rel = relative_floor(got, ref).max()
        assert nw <= tol, f"{prec} model {model} dim {dim}: normwise {nw:.2e}"
        assert rel <= tol, f"{prec} model {model} dim {dim}: relative {rel:.2e}"
    auto = engine.eigenvalues_by_seed(model, dim, T, seeds, precision=jne.PRECISIONS["auto"])
    mixed = engine.eigenvalues_by_seed(model, dim, T, seeds, precision=jne.PRECISIONS["mfma_fp16x2"])
    assert np.array_equal(auto, ref) or np.array_equal(auto, mixed)


@pytest.mark.parametrize("model,dim", [(0, 1), (1, 4), (2, 7), (3, 9), (4, 12)])
def test_sweep_mixed_matches_oracle(engine, jne, O, model, dim):
    got = engine.eigenvalues_by_seed(model, dim, T, ORACLE_SEEDS, precision=jne.PRECISIONS["mfma_fp16x2"])
    ref = O.walks_by_seed_c(model, dim, T, ORACLE_SEEDS, threads=4)
    assert normwise(got, ref).max() <= 1e-5
    assert relative_floor(got, ref).max() <= 1e-5


def test_streamed_resume_at_full_length(jne, engine, O, tmp_path):
    """jne_run_simulation at T = 10,000 (model 3, dim 12): 8,000 runs, then the
    file is cut as a crash mid-write leaves it (no trailer, half a record), then
    resumed to 20,000 runs; every record equals the batch path's walk, the bytes
    parse, and the trailer counts 20,000."""
    model, dim, runs = 3, 12, 20_000
    s1 = jne.EigenvalueSimulation(model, dim, T, 8_000, precision="auto", engine=engine,
                                  data_dir=str(tmp_path / "data"))
    s1.run_simulation_quiet()
    path = s1.get_filename()
    size = os.path.getsize(path)
    nf = dim + 1
    with open(path, "r+b") as f:  # drop the 17-byte trailer and half of the last record
        f.truncate(size - 17 - (1 + 8 * nf) // 2)
    info = jne.dat_info(path)
    assert not info["has_trailer"] and info["records"] == 7_999
    s2 = jne.EigenvalueSimulation(model, dim, T, runs, precision="auto", engine=engine,
                                  data_dir=str(tmp_path / "data"))
    s2.run_simulation_quiet()
    seeds, eigs, m, d, st = jne.read_dat(path)
    assert (m, d, st) == (model, dim, T)
    info = jne.dat_info(path)
    assert info["has_trailer"] and info["records"] == runs and info["nf"] == nf
    assert sorted(seeds.tolist()) == list(range(1, runs + 1))
    batch = engine.eigenvalues_by_seed(model, dim, T, np.arange(1, runs + 1, dtype=np.uint32),
                                       precision=jne.PRECISIONS["auto"])
    assert np.array_equal(eigs, batch[seeds.astype(np.int64) - 1])
    ref = O.walks_by_seed_c(model, dim, T, np.array([1, 8_000, 20_000], dtype=np.uint32), threads=3)
    got = batch[[0, 7_999, 19_999]]
    assert normwise(got, ref).max() <= 1e-5


def test_cli_sweep_all_pairs_at_full_length(jne, engine, tmp_path):
    """The command-line tool over models 0-4 x dims 1..12 at T = 10,000 (the
    reference's defaults but --runs), every file against the batch path."""
    runs = 2_000
    p = subprocess.run([CLI, "--steps", "10,000", "--runs", f"{runs:,}", "--precision", "auto", "--quiet"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0 and p.stdout == "", p.stderr
    all_seeds = np.arange(1, runs + 1, dtype=np.uint32)
    for model, dim in PAIRS:
        path = tmp_path / "data" / f"eigenvalues_model{model}_dim{dim}_steps{T}.dat"
        info = jne.dat_info(path)
        assert info["has_trailer"] and info["records"] == runs, (model, dim, info)
        seeds, eigs, *_ = jne.read_dat(path)
        assert sorted(seeds.tolist()) == list(range(1, runs + 1))
        batch = engine.eigenvalues_by_seed(model, dim, T, all_seeds, precision=jne.PRECISIONS["auto"])
        assert np.array_equal(eigs, batch[seeds.astype(np.int64) - 1]), (model, dim)
