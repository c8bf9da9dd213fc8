"""GPU parity of the reduced-precision moment passes, through the C ABI.

The mixed path draws the same Philox increments as the FP64 path and keeps the
state B and every per-segment term in FP64; only the within-segment Gram is
accumulated on the matrix cores (FP32 accumulate, flushed to FP64 every 64
steps).  The compensated FP32 VALU check (fp32_kahan) runs the same segment
split with the within-segment Gram formed on the VALU: FP32 products of the
same rows added into Kahan-compensated FP32 sums.  Their eigenvalues are
compared with the oracle / the FP64 path on the same seeds.

Two measures:

* relative with the structural-zero floor (the stated mixed-path tolerance):
  |d_i| <= 1e-5 |rho_i|, or <= 1e-5 rho_max for an entry below 1e-5 rho_max
  (the zero eigenvalue of models 1 and 3).  Checked at model 4, dim 12,
  T = 10,000 for fp16x2, bf16x3, fp32 and the compensated FP32 check
  (measured worst case 5e-6 over 10,000 walks for fp16x2).
* batch-normwise, for the sweep over models, dims and lengths:
  max_i |d_i| <= TOL * max(rho_max(walk), median over the batch of rho_max).
  The batch floor keeps walks whose whole spectrum is tiny (dim 1, where a
  near-cancelling sum e b makes rho small) from turning an absolutely small
  error into a large relative one.  TOL_LONG applies for steps >= 1000;
  short walks (steps < 1000) are dominated by the within-segment part, and
  TOL_SHORT still catches any structural mistake (a wrong row or a dropped
  tail step gives O(1) differences).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_LONG = {"mfma_fp16x2": 1e-5, "mfma_bf16x3": 1e-5, "mfma_fp32": 1e-5, "mfma_bf16x2": 1e-4,
            "fp32_kahan": 1e-5}
TOL_SHORT = {"mfma_fp16x2": 2e-3, "mfma_bf16x3": 2e-3, "mfma_fp32": 2e-3, "mfma_bf16x2": 5e-2,
             "fp32_kahan": 2e-3}
PRECS = list(TOL_LONG)
# relative-with-floor bound at the headline configuration (None: normwise only)
REL_HEADLINE = {"mfma_fp16x2": 1e-5, "mfma_bf16x3": 1e-5, "mfma_fp32": 1e-5, "mfma_bf16x2": None,
                "fp32_kahan": 1e-5}


def normwise(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape
    assert np.isfinite(got).all()
    rmax = np.max(np.abs(ref), axis=-1)
    return np.max(np.abs(got - ref), axis=-1) / np.maximum(rmax, np.median(rmax))


def relative_floor(got, ref, floor=1e-5):
    """|d_i| / |rho_i|, with rho_max for entries below floor * rho_max."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    rmax = np.max(np.abs(ref), axis=-1, keepdims=True)
    den = np.where(np.abs(ref) >= floor * rmax, np.abs(ref), rmax)
    return np.abs(got - ref) / den


def tol_for(prec, steps):
    return (TOL_LONG if steps >= 1000 else TOL_SHORT)[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dim", list(range(1, 13)))
def test_mixed_matches_fp64_path(engine, jne, prec, model, dim):
    # steps: < one segment, exactly one, several plus a tail
    for steps in [33, 64, 200, 1000]:
        seeds = np.arange(1, 33, dtype=np.uint32)
        ref = engine.eigenvalues_by_seed(model, dim, steps, seeds)
        got = engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=jne.PRECISIONS[prec])
        err = normwise(got, ref).max()
        assert err <= tol_for(prec, steps), f"{prec} model {model} dim {dim} steps {steps}: {err:.2e}"


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_matches_oracle_headline(engine, jne, O, prec):
    """Model 4, dim 12, T = 10,000 (the precision-ladder configuration)."""
    seeds = np.array([1, 2, 3, 1000, 77777], dtype=np.uint32)
    ref = O.walks_by_seed_c(4, 12, 10000, seeds, threads=5)
    got = engine.eigenvalues_by_seed(4, 12, 10000, seeds, precision=jne.PRECISIONS[prec])
    err = normwise(got, ref).max()
    assert err <= TOL_LONG[prec], f"{prec}: {err:.2e}"
    if REL_HEADLINE[prec] is not None:
        rel = relative_floor(got, ref).max()
        assert rel <= REL_HEADLINE[prec], f"{prec}: relative {rel:.2e}"


@pytest.mark.parametrize("prec", ["mfma_fp16x2", "mfma_bf16x3", "fp32_kahan"])
def test_mixed_relative_tolerance_headline_batch(engine, jne, prec):
    """The stated tolerance on 2,000 walks of the headline configuration."""
    seeds = np.arange(1, 2001, dtype=np.uint32)
    ref = engine.eigenvalues_by_seed(4, 12, 10000, seeds)
    got = engine.eigenvalues_by_seed(4, 12, 10000, seeds, precision=jne.PRECISIONS[prec])
    rel = relative_floor(got, ref)
    assert rel.max() <= REL_HEADLINE[prec], f"{prec}: max relative {rel.max():.2e}"


@pytest.mark.parametrize("model,dim,steps", [(2, 5, 5000), (1, 7, 3001), (3, 12, 2047), (0, 9, 1100)]
                         # each dim auto sends to the MFMA pass, with a partial last segment
                         + [(m, 8, 3037) for m in range(5)] + [(m, 10, 333) for m in range(5)]
                         + [(m, 11, 301) for m in range(5)] + [(m, 9, 517) for m in range(5)])
def test_mixed_matches_oracle_long(engine, jne, O, model, dim, steps):
    seeds = np.array([5, 6, 7], dtype=np.uint32)
    ref = O.walks_by_seed_c(model, dim, steps, seeds, threads=3)
    got = engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=jne.PRECISIONS["mfma_fp16x2"])
    err = normwise(got, ref).max()
    assert err <= TOL_LONG["mfma_fp16x2"], f"model {model} dim {dim} steps {steps}: {err:.2e}"


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_from_normals_equals_by_seed(engine, jne, prec):
    """Float-representable increments give bit-identical results on both entry points."""
    model, dim, steps = 4, 6, 300
    seeds = np.arange(1, 20, dtype=np.uint32)
    z = engine.walk_normals(dim, steps, seeds).astype(np.float64)
    p = jne.PRECISIONS[prec]
    assert np.array_equal(engine.eigenvalues_from_normals(model, z, precision=p),
                          engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=p))


def test_mixed_batch_invariance(engine, jne):
    p = jne.PRECISIONS["mfma_fp16x2"]
    model, dim, steps = 3, 12, 777
    seeds = np.arange(1, 201, dtype=np.uint32)
    full = engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=p)
    perm = np.random.default_rng(1).permutation(len(seeds))
    assert np.array_equal(full[perm], engine.eigenvalues_by_seed(model, dim, steps, seeds[perm], precision=p))
    assert np.array_equal(full[57], engine.eigenvalues_by_seed(model, dim, steps, seeds[57:58], precision=p)[0])


@pytest.mark.parametrize("prec", ["mfma_fp16x2", "fp32_kahan", "mfma_bf16x3"])
@pytest.mark.parametrize("dim", [1, 2, 3, 4, 6])
def test_mixed_packed_walks_are_independent(engine, jne, prec, dim):
    """Dims 1, 2, 3, 4 and 6 carry 12 / dim walks per wave: each walk's result is
    that of the walk alone, whatever its neighbours and wherever the batch ends
    (a batch of 25 leaves a partly filled last wave at every packing)."""
    p = jne.PRECISIONS[prec]
    model, steps = 4, 333
    seeds = np.arange(1, 26, dtype=np.uint32)
    full = engine.eigenvalues_by_seed(model, dim, steps, seeds, precision=p)
    for k in (0, 5, 11, 12, 24):
        assert np.array_equal(full[k], engine.eigenvalues_by_seed(model, dim, steps, seeds[k:k + 1], precision=p)[0]), k
    perm = np.random.default_rng(dim).permutation(len(seeds))
    assert np.array_equal(full[perm], engine.eigenvalues_by_seed(model, dim, steps, seeds[perm], precision=p))
    z = engine.walk_normals(dim, steps, seeds).astype(np.float64)
    assert np.array_equal(full, engine.eigenvalues_from_normals(model, z, precision=p))


def test_mixed_deterministic_and_sorted(engine, jne):
    p = jne.PRECISIONS["mfma_fp16x2"]
    seeds = np.arange(1, 65, dtype=np.uint32)
    a = engine.eigenvalues_by_seed(4, 12, 3000, seeds, precision=p)
    b = engine.eigenvalues_by_seed(4, 12, 3000, seeds, precision=p)
    assert np.array_equal(a, b)
    assert (np.diff(a, axis=1) <= 0).all() and (a >= 0).all()


def test_fp32_kahan_all_dims(engine, jne):
    """Every dim's compensated kernel (each has its own entry layout)."""
    p = jne.PRECISIONS["fp32_kahan"]
    seeds = np.arange(1, 65, dtype=np.uint32)
    for dim in range(1, 13):
        ref = engine.eigenvalues_by_seed(2, dim, 2500, seeds)
        got = engine.eigenvalues_by_seed(2, dim, 2500, seeds, precision=p)
        err = normwise(got, ref).max()
        assert err <= TOL_LONG["fp32_kahan"], f"dim {dim}: {err:.2e}"


def test_auto_precision_picks_the_faster_path(engine, jne):
    """auto runs the MFMA pass for dim 12 from 200 steps, dims 10-11 from 300,
    dim 9 from 500, dim 8 from 1,000 and dim 7 from 3,000 (the measured
    crossover), the FP64 pass elsewhere."""
    seeds = np.arange(1, 33, dtype=np.uint32)
    for dim, steps, same_as in [(12, 200, "mfma_fp16x2"), (12, 199, "fp64"), (11, 300, "mfma_fp16x2"),
                                (11, 299, "fp64"), (10, 300, "mfma_fp16x2"), (10, 299, "fp64"),
                                (9, 500, "mfma_fp16x2"), (9, 499, "fp64"), (8, 1000, "mfma_fp16x2"), (8, 999, "fp64"),
                                (7, 3000, "mfma_fp16x2"), (7, 2999, "fp64"), (6, 10000, "fp64"), (3, 700, "fp64"),
                                (13, 10000, "fp64")]:
        a = engine.eigenvalues_by_seed(4, dim, steps, seeds, precision=jne.PRECISIONS["auto"])
        b = engine.eigenvalues_by_seed(4, dim, steps, seeds, precision=jne.PRECISIONS[same_as])
        assert np.array_equal(a, b), (dim, steps)
