"""GPU parity of jne_percentiles_device (bucket selection on the device)
against the oracle's percentile rule (rank p*(n-1), linear interpolation;
src/simulation_analyzers.rs:4-18), through the C ABI.

Exact order statistics and the same interpolation arithmetic, so the values
must be equal exactly (bit for bit but for the sign of a zero: -0.0 and 0.0
are equal values and their order in a sort is not defined).  Cases: sizes from 1 to 2^21, smooth and heavy-tailed
data, negative values and signed zeros, ties (every value equal; three
distinct values, whose buckets overflow the selection and go to the radix
select), a few extreme outliers (a stretched bucket width), infinities (the
radix select),
NaN (an error) and invalid arguments.
"""
import numpy as np
import pytest

from conftest import DeviceArray

pytestmark = pytest.mark.gpu

PS = [0.90, 0.95, 0.99]


def on_device(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    d = DeviceArray(x.nbytes)
    d.from_host(x)
    return d


def check(engine, O, x, ps=PS):
    d = on_device(x)
    got = engine.percentiles_device(d.ptr, len(x), ps)
    ref = O.percentiles(np.asarray(x, dtype=np.float64), ps)
    assert np.array_equal(np.array(got), np.array(ref, dtype=np.float64)), (got, ref)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 101, 4097, 123457, 1 << 21])
def test_sizes(engine, O, n):
    rng = np.random.default_rng(n)
    check(engine, O, rng.standard_normal(n) * 3.0 + 20.0)


@pytest.mark.parametrize("dist", ["lognormal", "chisq", "uniform", "cauchy", "negative"])
def test_distributions(engine, O, dist):
    rng = np.random.default_rng(7)
    n = 200_000
    x = {"lognormal": rng.lognormal(2.0, 1.0, n), "chisq": rng.chisquare(3, n) * 10.0,
         "uniform": rng.uniform(-1e-3, 1e-3, n), "cauchy": rng.standard_cauchy(n),
         "negative": -rng.exponential(5.0, n)}[dist]
    check(engine, O, x, [0.0, 0.5, 0.9, 0.95, 0.99, 0.999, 1.0])


def test_ties_and_signed_zeros(engine, O):
    check(engine, O, np.full(50_000, 3.25))                       # all equal
    check(engine, O, np.repeat([1.0, 2.0, 5.0], [40_000, 5_000, 5_000]))  # crowded buckets: the radix select
    check(engine, O, np.array([0.0, -0.0, 0.0, -0.0, 1.0]), [0.0, 0.25, 0.5, 1.0])


def test_outliers_and_infinities(engine, O):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(300_000)
    x[:3] = [1e300, -1e300, 5e299]          # a stretched bucket width: the radix select
    check(engine, O, x)
    y = rng.standard_normal(100_000)
    y[10] = np.inf                         # not finite: the radix select
    check(engine, O, y)
    y[11] = -np.inf
    check(engine, O, y, [0.0, 0.5, 1.0])


def test_errors(engine, jne):
    d = on_device(np.array([1.0, np.nan, 2.0]))
    with pytest.raises(jne.JneError):
        engine.percentiles_device(d.ptr, 3, PS)
    ok = on_device(np.arange(10.0))
    for ps in ([], [0.5] * 9, [-0.1], [1.5], [float("nan")]):
        with pytest.raises(jne.JneError):
            engine.percentiles_device(ok.ptr, 10, ps)
    with pytest.raises(jne.JneError):
        engine.percentiles_device(ok.ptr, 1 << 31, PS)
    assert all(np.isnan(engine.percentiles_device(ok.ptr, 0, PS)))   # as jne_percentiles
    assert engine.percentiles_device(ok.ptr, 10, [0.5]) == [4.5]


def test_radix_select_fallback(engine, O):
    """Arrays the bucket selection hands to the radix select (an infinity, or
    a stretched bucket width), built so that the order statistics differ only
    in their last key digits, sit in long runs of ties, or lie among negative
    values and signed zeros."""
    rng = np.random.default_rng(11)
    ulps = 1.0 + np.arange(60_000) * np.finfo(np.float64).eps  # keys equal but for the low digits
    rng.shuffle(ulps)
    check(engine, O, np.concatenate([ulps, [np.inf]]), [0.0, 0.1, 0.5, 0.9, 0.95, 0.99])
    ties = np.repeat(rng.standard_normal(40), 2_500)  # 40 values, 2,500 copies each
    check(engine, O, np.concatenate([ties, [1e308, -1e308]]), [0.0, 0.05, 0.5, 0.9, 0.95, 0.99, 1.0])
    neg = -np.abs(rng.standard_normal(200_000))
    neg[:1000] = -0.0
    neg[1000:2000] = 0.0
    check(engine, O, np.concatenate([neg, [-np.inf]]), [0.0, 0.5, 0.9, 0.995, 1.0])
    big = rng.standard_normal(1 << 21)
    big[123] = np.inf
    check(engine, O, big)
