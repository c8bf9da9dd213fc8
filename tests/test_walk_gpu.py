"""GPU parity of the FP64 walk path against the oracle, through the C ABI.

Tolerances: |rho_gpu - rho_oracle| <= 1e-9 |rho| + 1e-13 rho_max per entry
against the oracle (reference algorithm + LAPACK dggev on the same
increments).  The absolute term covers the structural zero eigenvalue of
models 1 and 3 and eigenvalues many orders below rho_max, which two
backward-stable solvers agree on only to about eps * rho_max.  Normals are
compared bit for bit.
"""
import numpy as np
import pytest

from conftest import assert_eigs_close

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def test_normals_bit_exact(engine, O):
    for dim, steps in [(1, 7), (3, 33), (5, 101), (12, 257)]:
        seeds = np.array([1, 2, 1000, 0xFFFFFFFF], dtype=np.uint32)
        got = engine.walk_normals(dim, steps, seeds)
        for k, s in enumerate(seeds):
            ref = O.walk_normals(int(s), 0, dim, steps)
            assert np.array_equal(got[k].view(np.uint32), ref.view(np.uint32)), (dim, steps, s)
    # a non-zero rng stream gives a different, equally exact stream
    got = engine.walk_normals(4, 50, [7], rng_stream=12345)
    assert np.array_equal(got[0].view(np.uint32), O.walk_normals(7, 12345, 4, 50).view(np.uint32))
    assert not np.array_equal(got[0], O.walk_normals(7, 0, 4, 50))


def test_normals_device_entry_point(engine, O):
    """jne_walk_normals_device equals the host entry point and the CPU stream:
    more walks than one launch covers (65,535), a row of T*d floats that is
    not a multiple of 4 (scalar stores), a walk longer than one pass of the
    grid (strided calls), and an output pointer that is not 16-byte aligned."""
    from conftest import DeviceArray

    for dim, steps, n, offset in [(3, 5, 70000, 0), (12, 50000, 2, 0), (4, 25, 300, 4), (12, 100, 64, 0)]:
        seeds = np.arange(1, n + 1, dtype=np.uint32)
        seeds[-1] = 0xFFFFFFFF
        d_seeds = DeviceArray(seeds.nbytes)
        d_out = DeviceArray(n * steps * dim * 4 + offset)
        try:
            d_seeds.from_host(seeds)
            engine.walk_normals_device(dim, steps, d_seeds.ptr, n, d_out.ptr + offset)
            got = d_out.to_host((n, steps, dim), np.float32, offset)
        finally:
            d_seeds.free()
            d_out.free()
        host = engine.walk_normals(dim, steps, seeds)
        assert np.array_equal(got.view(np.uint32), host.view(np.uint32)), (dim, steps, n, offset)
        for k in [0, n // 2, n - 1]:
            ref = O.walk_normals(int(seeds[k]), 0, dim, steps)
            assert np.array_equal(got[k].view(np.uint32), ref.view(np.uint32)), (dim, steps, k)


@pytest.mark.parametrize("prec", ["fp64", "mfma_fp16x2"])
def test_eigenvalues_device_entry_point(engine, jne, prec):
    """jne_eigenvalues_by_seed_device equals the host entry point bit for bit,
    across the 2^18-walk chunks of one call."""
    from conftest import DeviceArray

    model, dim, steps, n = 4, 3, 70, (1 << 18) + 1000
    seeds = np.arange(7, 7 + n, dtype=np.uint32)
    d_seeds = DeviceArray(seeds.nbytes)
    d_out = DeviceArray(n * dim * 8)
    p = jne.PRECISIONS[prec]
    try:
        d_seeds.from_host(seeds)
        engine.eigenvalues_by_seed_device(model, dim, steps, d_seeds.ptr, n, d_out.ptr, precision=p)
        got = d_out.to_host((n, dim), np.float64)
    finally:
        d_seeds.free()
        d_out.free()
    assert np.array_equal(got, engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=p))


@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dim", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12])
def test_by_seed_matches_oracle_small(engine, O, model, dim):
    steps = 97 if dim % 2 else 128  # includes steps not a multiple of the 4L block
    seeds = np.arange(1, 6, dtype=np.uint32)
    got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    assert got.shape == (5, O.num_eigenvalues(model, dim))
    ref = np.stack([O.walk_eigenvalues_by_seed(model, dim, steps, int(s)) for s in seeds])
    assert_eigs_close(got, ref, RTOL, f"model {model} dim {dim} steps {steps}")


@pytest.mark.parametrize("model,dim,steps", [(2, 5, 5000), (4, 12, 10000), (3, 12, 2000), (1, 7, 3001)])
def test_by_seed_matches_oracle_large(engine, O, model, dim, steps):
    seeds = np.array([1, 17, 999], dtype=np.uint32)
    got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    ref = np.stack([O.walk_eigenvalues_by_seed(model, dim, steps, int(s)) for s in seeds])
    assert_eigs_close(got, ref, RTOL, f"model {model} dim {dim} steps {steps}")


def test_device_calls_on_two_streams(engine, jne):
    """Device calls on one context alternating between two caller streams
    (several chunks each, enqueued without synchronising): the context's
    shared scratch is ordered between them, so both results are exact."""
    from conftest import DeviceArray

    model, dim, steps, n = 4, 12, 300, (1 << 18) + 5000
    seeds = [np.arange(1 + k * n, 1 + (k + 1) * n, dtype=np.uint32) for k in range(2)]
    streams = [DeviceArray.stream(), DeviceArray.stream()]
    d_seeds = [DeviceArray(n * 4), DeviceArray(n * 4)]
    d_out = [DeviceArray(n * dim * 8), DeviceArray(n * dim * 8)]
    try:
        for k in range(2):
            d_seeds[k].from_host(seeds[k])
        for _ in range(3):
            for k in range(2):
                engine.eigenvalues_by_seed_device(model, dim, steps, d_seeds[k].ptr, n, d_out[k].ptr,
                                                  precision=jne.PRECISIONS["mfma_fp16x2"], hip_stream=streams[k])
        got = [d_out[k].to_host((n, dim), np.float64) for k in range(2)]
    finally:
        for a in d_seeds + d_out:
            a.free()
        for st in streams:
            DeviceArray.destroy_stream(st)
    for k in range(2):
        ref = engine.eigenvalues_by_seed(model, dim, steps, seeds[k], precision=jne.PRECISIONS["mfma_fp16x2"])
        assert np.array_equal(got[k], ref), k


def test_device_call_on_torchs_default_stream(engine, jne):
    """A device call on torch's default stream (handle 0: HIP's null stream)
    is ordered after the torch work that produced its inputs and before the
    torch copy that reads its outputs, with no device synchronisation in
    between."""
    import torch

    model, dim, steps, n = 4, 12, 2000, 50_000
    dev = torch.device("cuda:0")
    for _ in range(2):
        seeds = torch.arange(1, n + 1, dtype=torch.int32, device=dev) * 3  # enqueued on the null stream
        out = torch.full((n, dim), float("nan"), dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        assert stream == 0
        engine.eigenvalues_by_seed_device(model, dim, steps, seeds.data_ptr(), n, out.data_ptr(),
                                          precision=jne.PRECISIONS["mfma_fp16x2"], hip_stream=stream)
        got = out.cpu().numpy()  # a null-stream copy: waits for the walk kernels
        ref = engine.eigenvalues_by_seed(model, dim, steps, np.arange(1, n + 1, dtype=np.uint32) * 3,
                                         precision=jne.PRECISIONS["mfma_fp16x2"])
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("model,dim,steps,n", [(0, 2, 1000, 20000), (2, 5, 5000, 2000), (4, 12, 10000, 400)])
def test_workload_shapes_many_walks(engine, O, model, dim, steps, n):
    """The plain-CPU, plain-GPU and headline workload shapes on many walks
    (seeds spread over the u32 range) against the C restatement of the
    reference walk, at the FP64 tolerance."""
    seeds = np.unique(np.linspace(1, 0xFFFFFFFF, n).astype(np.uint64)).astype(np.uint32)
    got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    ref = O.walks_by_seed_c(model, dim, steps, seeds, threads=8)
    assert_eigs_close(got, ref, RTOL, f"model {model} dim {dim} steps {steps}")


def test_from_normals_equals_by_seed(engine, O):
    model, dim, steps = 4, 6, 300
    seeds = np.arange(1, 40, dtype=np.uint32)
    z = engine.walk_normals(dim, steps, seeds).astype(np.float64)
    a = engine.eigenvalues_from_normals(model, z)
    b = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    assert np.array_equal(a, b)


def test_from_normals_arbitrary_increments(engine, O):
    rng = np.random.default_rng(5)
    for model in range(5):
        z = rng.standard_normal((3, 211, 4))
        got = engine.eigenvalues_from_normals(model, z)
        ref = np.stack([O.walk_eigenvalues_lapack(model, z[k]) for k in range(3)])
        assert_eigs_close(got, ref, RTOL, f"model {model}")


def test_batch_invariance(engine):
    """A walk's result depends only on its seed, not on its slot in the batch."""
    model, dim, steps = 4, 12, 500
    seeds = np.arange(1, 301, dtype=np.uint32)
    full = engine.eigenvalues_by_seed(model, dim, steps, seeds)
    perm = np.random.default_rng(0).permutation(len(seeds))
    shuffled = engine.eigenvalues_by_seed(model, dim, steps, seeds[perm])
    assert np.array_equal(full[perm], shuffled)
    alone = engine.eigenvalues_by_seed(model, dim, steps, seeds[123:124])
    assert np.array_equal(alone[0], full[123])


def test_ragged_batch_sizes(engine, O):
    model, dim, steps = 2, 11, 64
    for n in [1, 15, 16, 17, 63, 65]:
        seeds = np.arange(100, 100 + n, dtype=np.uint32)
        got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
        ref = O.walks_by_seed_c(model, dim, steps, seeds)
        assert_eigs_close(got, ref, RTOL, f"n={n}")


def test_empty_and_invalid(engine, jne):
    out = engine.eigenvalues_by_seed(0, 2, 10, np.zeros(0, dtype=np.uint32))
    assert out.shape == (0, 2)
    with pytest.raises(jne.JneError) as e:
        engine.eigenvalues_by_seed(5, 2, 10, [1])
    assert e.value.code == jne.JNE_ERR_INVALID_ARG and "invalid model" in str(e.value)
    with pytest.raises(jne.JneError):
        engine.eigenvalues_by_seed(0, 33, 10, [1])
    with pytest.raises(jne.JneError):
        engine.eigenvalues_by_seed(0, 2, 1, [1])
    # the context stays usable after an error
    assert engine.eigenvalues_by_seed(0, 1, 10, [1]).shape == (1, 1)


def test_golden_fixtures(engine):
    """Committed fixtures (oracle: reference algorithm + LAPACK dggev)."""
    import json
    import os

    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walk_eigenvalues.json")))
    for c in fx["cases"]:
        got = engine.eigenvalues_by_seed(c["model"], c["dim"], c["steps"], [c["seed"]])[0]
        assert_eigs_close(got, np.array(c["eigenvalues"]), RTOL, str(c))
    sn = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stream_normals.json")))
    for c in sn["cases"]:
        z = engine.walk_normals(c["dim"], c["steps"], [c["seed"]], rng_stream=c["rng_stream"])[0]
        assert z.ravel().view(np.uint32).tolist() == c["bits"]


def test_calculate_eigenvalues_mirror(jne, O):
    got = jne.calculate_eigenvalues(3, 200, 42, jne.JohansenModel.InterceptTrendUnrestrictedBoth)
    ref = O.walk_eigenvalues_by_seed(4, 3, 200, 42)
    assert_eigs_close(np.array(got), ref, RTOL)
    assert got == sorted(got, reverse=True)
