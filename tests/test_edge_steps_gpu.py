"""Walk lengths at the edges of the kernels' step blocking, through the C ABI.

The FP64 pass at dims 1-6 keeps a walk in one lane and consumes its steps in
batches of 12, 6, 4, 3, 4 and 2 steps (dims 1..6); the mixed pass works in
64-step segments.  Lengths here leave every possible remainder of those
blocks (29 = 2 x 12 + 5, 37 = 3 x 12 + 1, 63 = 0 x 64 + 63, 65 = 64 + 1,
131 = 2 x 64 + 3, 193 = 3 x 64 + 1, odd lengths for the 2-step batches).
FP64 is held to the oracle (C restatement of the reference walk) at
|d| <= 1e-9 |rho| + 1e-13 rho_max; the reduced precisions to the FP64 path
at the short-walk normwise bound, and the from-normals entry point to the
by-seed one bit for bit.  A walk of zero increments (S11 = 0, not positive
definite) gives a NaN row on each form of the eigen solve.
"""
import numpy as np
import pytest

from conftest import assert_eigs_close

pytestmark = pytest.mark.gpu

# the short-walk normwise bounds of the mixed-pass sweep
TOL_SHORT = {"mfma_fp16x2": 2e-3, "fp32_kahan": 2e-3}


def normwise(got, ref):
    """max_i |d_i| / max(rho_max(walk), median over the batch of rho_max)."""
    assert got.shape == ref.shape and np.isfinite(got).all()
    rmax = np.max(np.abs(ref), axis=-1)
    return np.max(np.abs(got - ref), axis=-1) / np.maximum(rmax, np.median(rmax))

LENGTHS = [29, 37, 63, 65, 131, 193]


@pytest.mark.parametrize("dim", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("steps", LENGTHS)
def test_fp64_one_lane_per_walk_lengths(engine, O, dim, steps):
    seeds = np.array([3, 77, 4096], dtype=np.uint32)
    for model in range(5):
        got = engine.eigenvalues_by_seed(model, dim, steps, seeds)
        ref = O.walks_by_seed_c(model, dim, steps, seeds, threads=3)
        assert_eigs_close(got, ref, 1e-9, f"model {model} dim {dim} steps {steps}")
    z = engine.walk_normals(dim, steps, seeds).astype(np.float64)
    assert np.array_equal(engine.eigenvalues_from_normals(4, z), engine.eigenvalues_by_seed(4, dim, steps, seeds))


@pytest.mark.parametrize("prec", ["mfma_fp16x2", "fp32_kahan"])
@pytest.mark.parametrize("dim", [1, 2, 3, 4, 5, 6, 12])
def test_mixed_segment_remainders(engine, jne, prec, dim):
    p = jne.PRECISIONS[prec]
    seeds = np.arange(1, 14, dtype=np.uint32)  # a partly filled last wave at every packing
    for steps in [63, 65, 131, 193]:
        for model in (0, 4):
            ref = engine.eigenvalues_by_seed(model, dim, steps, seeds)
            got = engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=p)
            err = normwise(got, ref).max()
            assert err <= TOL_SHORT[prec], f"{prec} model {model} dim {dim} steps {steps}: {err:.2e}"
        z = engine.walk_normals(dim, steps, seeds).astype(np.float64)
        assert np.array_equal(engine.eigenvalues_from_normals(4, z, precision=p), got), (prec, dim, steps)


@pytest.mark.parametrize("dim", [3, 12, 20])
def test_degenerate_walk_gives_nan_row(jne, engine, O, dim):
    """A walk whose increments are all zero has S11 = 0 (not positive
    definite): the call returns JNE_ERR_NUMERIC naming that walk and its row
    is NaN, on the fused (dim 3), split (dim 12) and one-wave-per-walk (dim 20)
    solves alike, and the walks beside it in the batch keep their values."""
    import ctypes
    rng = np.random.default_rng(dim)
    for model in range(5):
        z = np.ascontiguousarray(rng.standard_normal((3, 64, dim)))
        z[1] = 0.0
        out = np.full((3, jne.num_eigenvalues(model, dim)), 7.0)
        rc = engine._lib.jne_eigenvalues_from_normals(engine._ctx, model, dim, 64, z.ctypes.data, 3,
                                                      jne.PRECISIONS["fp64"], out.ctypes.data)
        assert rc == jne.JNE_ERR_NUMERIC
        assert engine._lib.jne_last_error(engine._ctx).decode().startswith("walk 1:")
        assert np.isnan(out[1]).all()
        assert np.isfinite(out[[0, 2]]).all()
        ref = np.stack([O.walk_eigenvalues_lapack(model, z[k]) for k in (0, 2)])
        assert_eigs_close(out[[0, 2]], ref, 1e-9, f"model {model} dim {dim}")
