"""Several models of the same walks sharing one moment pass, through the C ABI.

A walk's moments do not depend on the model, so
jne_eigenvalues_by_seed_models_device runs the moment pass once and the model
algebra and eigen solve per model; each model's output must be bit for bit
that of its own jne_eigenvalues_by_seed_device call.  The quiet sweep driver
(jne_run_simulation_models, used by the CLI with --quiet) must write each
.dat file with the same records as a per-model jne_run_simulation, also when
the models' files are at different stages of a resumed run.
"""
import filecmp
import subprocess
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "johansen-null-eigenspectra")


@pytest.mark.parametrize("prec", ["fp64", "mfma_fp16x2", "auto"])
@pytest.mark.parametrize("dim,steps", [(1, 333), (4, 1000), (9, 700), (12, 2000), (13, 300), (20, 257)])
def test_models_device_equals_single(engine, jne, prec, dim, steps):
    from conftest import DeviceArray
    if dim > 12 and prec == "mfma_fp16x2":
        pytest.skip("reduced precisions cover dims 1..12")
    p = jne.PRECISIONS[prec]
    seeds = np.arange(5, 5 + 2000, dtype=np.uint32)
    n = len(seeds)
    models = [4, 0, 3, 1, 2]
    d_seeds = DeviceArray(seeds.nbytes)
    d_seeds.from_host(seeds)
    outs = [DeviceArray(n * jne.num_eigenvalues(m, dim) * 8) for m in models]
    engine.eigenvalues_by_seed_models_device(models, dim, steps, d_seeds.ptr, n, [o.ptr for o in outs], precision=p)
    for m, o in zip(models, outs):
        nf = jne.num_eigenvalues(m, dim)
        got = o.to_host((n, nf), np.float64)
        ref = engine.eigenvalues_by_seed(m, dim, steps, seeds, precision=p)
        assert np.array_equal(got, ref), (m, dim, steps, prec)


def test_models_driver_fresh_and_resumed(jne, engine, tmp_path):
    dim, steps, runs = 5, 400, 3000
    models = [0, 1, 2, 3, 4]
    fused = [tmp_path / f"fused_m{m}.dat" for m in models]
    single = [tmp_path / f"single_m{m}.dat" for m in models]
    # resumed: model 1's file holds a prefix, model 3's is complete, the rest are empty
    engine_simulate = lambda m, path, r: engine.run_simulation_models([m], dim, steps, r, [path])
    engine_simulate(1, fused[1], 1000)
    engine_simulate(3, fused[3], runs)
    engine.run_simulation_models(models, dim, steps, runs, fused, precision=jne.PRECISIONS["auto"])
    for m, path in zip(models, single):
        engine.run_simulation_models([m], dim, steps, runs, [path], precision=jne.PRECISIONS["auto"])
    for m in models:
        s1, e1, *_ = jne.read_dat(fused[m])
        s2, e2, *_ = jne.read_dat(single[m])
        o1, o2 = np.argsort(s1), np.argsort(s2)
        assert np.array_equal(s1[o1], s2[o2]) and sorted(s1.tolist()) == list(range(1, runs + 1)), m
        assert np.array_equal(e1[o1], e2[o2]), m
    # a fresh fused sweep writes each file byte for byte as the per-model run
    fresh = [tmp_path / f"fresh_m{m}.dat" for m in models]
    engine.run_simulation_models(models, dim, steps, runs, fresh, precision=jne.PRECISIONS["auto"])
    for m in models:
        assert filecmp.cmp(fresh[m], single[m], shallow=False), m


def test_models_driver_rejects_bad_arguments(jne, engine, tmp_path):
    with pytest.raises(jne.JneError) as e:
        engine.run_simulation_models([0, 7], 3, 100, 10, [tmp_path / "a.dat", tmp_path / "b.dat"])
    assert e.value.code == jne.JNE_ERR_INVALID_ARG
    with pytest.raises(jne.JneError) as e:
        engine.run_simulation_models([0, 1, 2, 3, 4, 0], 3, 100, 10, [tmp_path / f"{k}.dat" for k in range(6)])
    assert e.value.code == jne.JNE_ERR_INVALID_ARG


def test_cli_quiet_sweep_matches_loud(tmp_path):
    """--quiet runs the models of a dim together; the files equal those of the
    message-printing (model by model) run."""
    args = ["--dim-start", "2", "--dim-end", "3", "--steps", "300", "--runs", "2,500", "--precision", "auto"]
    (tmp_path / "q").mkdir()
    (tmp_path / "l").mkdir()
    p = subprocess.run([CLI, *args, "--quiet"], cwd=tmp_path / "q", capture_output=True, text=True, timeout=300)
    assert p.returncode == 0 and p.stdout == "", p.stderr
    p = subprocess.run([CLI, *args], cwd=tmp_path / "l", capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    for dim in (2, 3):
        for m in range(5):
            name = f"eigenvalues_model{m}_dim{dim}_steps300.dat"
            assert filecmp.cmp(tmp_path / "q" / "data" / name, tmp_path / "l" / "data" / name, shallow=False), name
