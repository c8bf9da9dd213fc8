"""The query server on the GPU: results against the batch path and the
percentile definition, the same result alone and in a mix of concurrent
queries, HBM-budget admission, timeouts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PS = [0.90, 0.95, 0.99]


def expected(engine, O, model, dim, steps, runs, seed, precision):
    seeds = np.arange(1, runs + 1, dtype=np.uint32)
    eigs = engine.eigenvalues_by_seed(model, dim, steps, seeds, rng_stream=seed, precision=precision)
    trace = np.cumsum(eigs, axis=1)[:, -1]  # left-to-right sum, as the statistics kernel adds
    maxeig = eigs.max(axis=1)
    return np.array(O.percentiles(trace, PS) + O.percentiles(maxeig, PS))


@pytest.fixture(scope="module")
def server(jne):
    s = jne.Server(0, hbm_budget_bytes=1 << 30, lanes=4)
    yield s
    s.close()


def test_query_matches_batch_path(jne, engine, O, server):
    # the third and fourth span several chunks: 65,536 walks at most, and at
    # dim 12, T = 10,000 at most 2^31 normals (17,664 walks), the last one
    # partial; the last two run the wide-dim kernels (one wave per walk)
    for model, dim, steps, runs, seed, prec in [(2, 3, 500, 20000, 7, "fp64"), (4, 12, 1000, 3000, 11, "mfma_fp16x2"),
                                               (1, 1, 100, 70001, 3, "mfma_fp16x2"),
                                               (4, 12, 10000, 40000, 5, "mfma_fp16x2"),
                                               (0, 13, 300, 20000, 21, "auto"), (4, 32, 500, 9000, 23, "fp64")]:
        q, lat = server.wait(server.submit(model, dim, steps, runs, seed, prec))
        assert np.array_equal(q, expected(engine, O, model, dim, steps, runs, seed, jne.PRECISIONS[prec]))
        assert lat > 0


def test_same_alone_and_in_the_mix(jne, server):
    a = (3, 7, 800, 50000, 1234)
    alone, _ = server.wait(server.submit(*a))
    rng = np.random.default_rng(0)
    ids, specs = [], []
    for k in range(12):
        spec = (int(rng.integers(0, 5)), int(rng.integers(1, 13)), int(rng.integers(100, 2000)),
                int(rng.integers(1000, 80000)), int(rng.integers(1, 2**31)))
        if k == 6:
            ids.append(server.submit(*a))
            specs.append(a)
        ids.append(server.submit(*spec))
        specs.append(spec)
    results = [server.wait(i)[0] for i in ids]
    assert np.array_equal(results[6], alone)
    # two of the others, re-run alone
    for k in (0, 9):
        assert np.array_equal(server.wait(server.submit(*specs[k]))[0], results[k])


def test_mixed_precision_quantiles_close_to_fp64(server):
    q64, _ = server.wait(server.submit(4, 12, 3000, 20000, 99, "fp64"))
    q16, _ = server.wait(server.submit(4, 12, 3000, 20000, 99, "mfma_fp16x2"))
    assert np.all(np.abs(q16 - q64) <= 1e-5 * np.abs(q64))


def test_admission_by_hbm_budget(jne):
    with jne.Server(0, hbm_budget_bytes=2 << 20, lanes=2) as s:  # room for 131,072 walks of statistics
        ids = [s.submit(0, 2, 100, 100000, k) for k in range(4)]  # one at a time fits
        for i in ids:
            q, _ = s.wait(i)
            assert np.isfinite(q).all()
        big = s.submit(0, 2, 100, 200000, 5)
        with pytest.raises(jne.JneError) as e:
            s.wait(big)
        assert e.value.code == jne.JNE_ERR_NOMEM


def test_wait_timeout_and_unknown_id(jne, server):
    i = server.submit(4, 12, 10000, 100000, 5, "mfma_fp16x2")
    with pytest.raises(jne.JneError) as e:
        server.wait(i, timeout_ms=0)
    assert e.value.code == jne.JNE_ERR_TIMEOUT
    q, lat = server.wait(i)
    assert np.isfinite(q).all()
    with pytest.raises(jne.JneError):
        server.wait(i)
    with pytest.raises(jne.JneError):
        server.submit(4, 33, 100, 10, 1)
