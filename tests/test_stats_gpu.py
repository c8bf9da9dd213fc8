"""Statistical parity on the GPU.

The device stream (Philox4x32-10 + Box-Muller) replaces the reference's
Xoshiro256++/Ziggurat draw, so parity of the statistics is distributional:

* the device normals against N(0,1) (Kolmogorov-Smirnov), with no
  correlation between consecutive draws or between components;
* the trace and max-eigenvalue statistics of walks driven by the device
  stream against walks driven by an independent generator (numpy's PCG64
  normals through the from-normals entry point, i.e. the reference algorithm
  on other increments): two-sample KS, and the 0.90/0.95/0.99 quantiles
  within 4 Monte Carlo standard errors (bootstrap).
Fixed seeds make every run identical; the thresholds leave room for chance
but not for a wrong distribution.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_normals_are_standard_normal(engine):
    from scipy import stats

    z = engine.walk_normals(4, 50000, np.arange(1, 6, dtype=np.uint32)).astype(np.float64)  # 1e6 draws
    flat = z.reshape(-1)
    assert stats.kstest(flat, "norm").pvalue > 1e-3
    assert abs(flat.mean()) < 5 / np.sqrt(flat.size) and abs(flat.var() - 1) < 5 * np.sqrt(2 / flat.size)
    w = z.reshape(-1, 4)
    lag = np.corrcoef(w[:-1, 0], w[1:, 0])[0, 1]
    comp = np.corrcoef(w[:, 0], w[:, 3])[0, 1]
    assert abs(lag) < 5 / np.sqrt(len(w)) and abs(comp) < 5 / np.sqrt(len(w))


def bootstrap_se(x, ps, rng, reps=200):
    n = len(x)
    qs = np.array([np.quantile(x[rng.integers(0, n, n)], ps) for _ in range(reps)])
    return qs.std(axis=0)


@pytest.mark.parametrize("model,dim,steps", [(2, 3, 300), (4, 2, 200)])
def test_statistics_match_an_independent_generator(engine, model, dim, steps):
    from scipy import stats

    n = 12000
    a = engine.eigenvalues_by_seed(model, dim, steps, np.arange(1, n + 1, dtype=np.uint32))
    z = np.random.default_rng(20240601 + model).standard_normal((n, steps, dim))
    b = engine.eigenvalues_from_normals(model, z)
    rng = np.random.default_rng(7)
    ps = [0.90, 0.95, 0.99]
    for fa, fb in [(a.sum(1), b.sum(1)), (a.max(1), b.max(1))]:
        assert stats.ks_2samp(fa, fb).pvalue > 1e-3
        qa, qb = np.quantile(fa, ps), np.quantile(fb, ps)
        se = np.hypot(bootstrap_se(fa, ps, rng), bootstrap_se(fb, ps, rng))
        assert np.all(np.abs(qa - qb) <= 4 * se), (qa, qb, se)
