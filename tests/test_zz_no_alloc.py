"""The allocation-free entry point (jne_eigenvalues_by_seed_device_noalloc)
through the C ABI: its outputs equal jne_eigenvalues_by_seed_device's bit for
bit on two calls in a row, the second with fewer walks (the fixed scratch
keeps the first call's contents, which must not leak into the second), at
the FP64 and the mixed precision; a batch above the scratch capacity is
refused with a message and leaves the context usable."""
import numpy as np
import pytest

from conftest import DeviceArray

pytestmark = pytest.mark.gpu


def run_both(engine, model, dim, steps, seeds, prec):
    n = len(seeds)
    nf = dim + 1 if model in (1, 3) else dim
    d_seeds = DeviceArray(seeds.nbytes)
    d_a, d_b = DeviceArray(n * nf * 8), DeviceArray(n * nf * 8)
    try:
        d_seeds.from_host(seeds)
        engine.eigenvalues_by_seed_device(model, dim, steps, d_seeds.ptr, n, d_a.ptr, precision=prec)
        engine.synchronize()
        engine.eigenvalues_by_seed_device_noalloc(model, dim, steps, d_seeds.ptr, n, d_b.ptr, precision=prec)
        engine.synchronize()
        return d_a.to_host((n, nf), np.float64), d_b.to_host((n, nf), np.float64)
    finally:
        for a in (d_seeds, d_a, d_b):
            a.free()


@pytest.mark.parametrize("prec", ["fp64", "mfma_fp16x2", "auto"])
def test_noalloc_equals_the_allocating_entry_point(engine, jne, prec):
    p = jne.PRECISIONS[prec]
    model, dim, steps = 4, 12, 2000
    n1 = jne.noalloc_max_walks(dim)  # the whole scratch
    first = np.arange(1, n1 + 1, dtype=np.uint32)
    a, b = run_both(engine, model, dim, steps, first, p)
    assert np.isfinite(a).all() and np.array_equal(a, b)
    second = np.arange(500_001, 500_001 + 777, dtype=np.uint32)  # fewer walks, other seeds
    a, b = run_both(engine, 3, 7, 333, second, p)
    assert np.isfinite(a).all() and np.array_equal(a, b)


def test_noalloc_wide_dim(engine, jne):
    """Above dim 12 (the runtime-dim FP64 path, no M scratch): the capacity
    formula without the M term, and the same results as the allocating call."""
    assert jne.noalloc_max_walks(32) == (64 << 20) // 8 // (2 * 32 * 32 + 6 * 32)
    seeds = np.arange(1, 1001, dtype=np.uint32)
    a, b = run_both(engine, 4, 20, 200, seeds, jne.PRECISIONS["auto"])
    assert np.isfinite(a).all() and np.array_equal(a, b)


def test_noalloc_small_dim_and_capacity(engine, jne):
    cap = jne.noalloc_max_walks(1)
    assert cap == (64 << 20) // 8 // 10
    seeds = np.arange(1, 300_001, dtype=np.uint32)  # above one 2^18-walk chunk of the other entry point
    a, b = run_both(engine, 0, 1, 100, seeds, jne.PRECISIONS["fp64"])
    assert np.array_equal(a, b)
    n = jne.noalloc_max_walks(12) + 1
    d_seeds, d_out = DeviceArray(n * 4), DeviceArray(n * 12 * 8)
    try:
        with pytest.raises(jne.JneError) as e:
            engine.eigenvalues_by_seed_device_noalloc(4, 12, 100, d_seeds.ptr, n, d_out.ptr)
        assert e.value.code == jne.JNE_ERR_INVALID_ARG and "exceed the scratch" in str(e.value)
    finally:
        d_seeds.free()
        d_out.free()
    # the context stays usable
    assert engine.eigenvalues_by_seed(0, 1, 10, [1]).shape == (1, 1)
