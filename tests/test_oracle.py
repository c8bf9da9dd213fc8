"""Pins of the CPU oracle (runs without a GPU).

Each check is anchored outside the oracle's own formula: Random123 known
answers (Philox), libm Box-Muller and the reference's own N(0,1) quantile
fixture (normals), the reference's known-answer tests (cumsum, outer
products, ULEB128, .dat bytes), closed forms and invariants (Brownian path,
F construction, d=1 eigenvalue, Ito identity) and LAPACK dggev -- the routine
the reference calls -- against the C Jacobi solve.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import assert_eigs_close

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _interpolated_cdf_sorted(data_sorted, query_sorted):
    """Empirical CDF with linear interpolation, as the reference's RNG tests
    compute it (src/tests/rng_matrix_test/mod.rs:137-167)."""
    n = len(data_sorted)
    out, i = [], 0
    for x in query_sorted:
        if x <= data_sorted[0]:
            out.append(0.0)
            continue
        if x >= data_sorted[-1]:
            out.append(1.0)
            continue
        while i + 1 < n and data_sorted[i + 1] <= x:
            i += 1
        x0, x1 = data_sorted[i], data_sorted[i + 1]
        w = (x - x0) / (x1 - x0)
        out.append((i / n) * (1 - w) + ((i + 1) / n) * w)
    return np.array(out)


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(O, ctr, key, expect):
    assert O.philox4x32_10(ctr, key).tolist() == expect


# ---------------------------------------------------------------- normals
def test_normal_pair_matches_libm_box_muller(O):
    rng = np.random.default_rng(1)
    words = rng.integers(0, 2**32, size=(20000, 2), dtype=np.uint64)
    worst = 0.0
    for wa, wb in words:
        z = O.normal_pair(int(wa), int(wb)).astype(np.float64)
        u1 = (float(np.float32(float(wa))) * 2.0**-32) + 2.0**-33
        r = math.sqrt(-2.0 * math.log(u1))
        q = int(wb) >> 30
        g = (int(wb) & 0x3FFFFFFF) - 0x20000000
        theta = (q + 0.5 + g * 2.0**-30) * math.pi / 2
        ref = np.array([r * math.cos(theta), r * math.sin(theta)])
        worst = max(worst, float(np.max(np.abs(z - ref))) / max(r, 1e-3))
    assert worst < 2e-6, worst


def test_normals_reference_cdf_fixture(O):
    """The reference's gen_normal_matrix CDF test (gen_normal_matrix_test.rs:7-16)
    on the build's stream: a 200 x 300 draw, seed 42, within 1e-2."""
    fx = json.load(open(os.path.join(GOLDEN, "reference_normal_quantiles.json")))
    z = np.sort(O.walk_normals(42, 0, 200, 300).ravel().astype(np.float64))
    cdf = _interpolated_cdf_sorted(z, fx["quantiles"])
    assert np.max(np.abs(cdf - np.array(fx["expected_cdf"]))) < 1e-2


def test_brownian_increments_reference_cdf_fixture(O):
    """brownian_motion_test.rs:47-96: per-dimension increments / sqrt(dt) pass
    the CDF check within 0.05 (dim 3, 1000 steps, seed 42)."""
    fx = json.load(open(os.path.join(GOLDEN, "reference_normal_quantiles.json")))
    z = O.walk_normals(42, 0, 3, 1000)
    B = O.brownian(z.astype(np.float64))
    sqrt_dt = math.sqrt(1.0 / 1000)
    for d in range(3):
        inc = np.sort(np.diff(B[d]) / sqrt_dt)
        cdf = _interpolated_cdf_sorted(inc, fx["quantiles"])
        assert np.max(np.abs(cdf - np.array(fx["expected_cdf"]))) < 0.05


def test_normals_ks_and_tail(O):
    from scipy import stats

    z = np.concatenate([O.walk_normals(s, 0, 12, 10000).ravel() for s in (1, 2, 3)]).astype(np.float64)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 5 / math.sqrt(len(z))
    assert abs(z.std() - 1) < 5 / math.sqrt(2 * len(z))
    # u1 >= 2^-33 bounds |z| by sqrt(66 ln 2)
    assert np.max(np.abs(z)) <= math.sqrt(66 * math.log(2)) + 1e-6


def test_stream_golden(O):
    fx = json.load(open(os.path.join(GOLDEN, "stream_normals.json")))
    for c in fx["cases"]:
        z = O.walk_normals(c["seed"], c["rng_stream"], c["dim"], c["steps"])
        assert z.ravel().view(np.uint32).tolist() == c["bits"]


def test_stream_reproducible_and_seed_dependent(O):
    a = O.walk_normals(123, 0, 2, 50)
    assert np.array_equal(a, O.walk_normals(123, 0, 2, 50))
    assert not np.array_equal(a, O.walk_normals(456, 0, 2, 50))
    assert not np.array_equal(a, O.walk_normals(123, 1, 2, 50))


# ---------------------------------------------------------------- Brownian path, cumsum, outer products
def test_brownian_shape_and_cumsum(O):
    """Shape d x (T+1), zero start column, and the reference's RowWise cumsum
    known answer (dmatrix_cumsum_test.rs:5-11) after undoing the scaling."""
    z = np.array([[1.0, 4.0], [2.0, 5.0], [3.0, 6.0]])  # T=3, d=2; rows of the reference 2x3 matrix
    B = O.brownian(z)
    assert B.shape == (2, 4)
    assert np.all(B[:, 0] == 0.0)
    np.testing.assert_allclose(B[:, 1:] / math.sqrt(1.0 / 3), [[1, 3, 6], [4, 9, 15]], rtol=1e-14)


def test_brownian_c_matches_numpy(O):
    z = O.walk_normals(9, 0, 4, 777).astype(np.float64)
    Bn = O.brownian(z)
    Bc = np.zeros((778, 4))
    import ctypes

    O.lib().jo_brownian(4, 777, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        Bc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert np.array_equal(Bn, Bc.T)


def _sum_outer(O, a, b):
    import ctypes

    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    out = np.zeros((a.shape[0], b.shape[0]), order="F")
    P = ctypes.POINTER(ctypes.c_double)
    O.lib().jo_sum_outer(a.shape[0], b.shape[0], a.shape[1], a.ctypes.data_as(P), b.ctypes.data_as(P),
                         out.ctypes.data_as(P))
    return out


def test_sum_of_outer_products_known_answers(O):
    """sum_of_outer_products_test.rs:5-159."""
    a = np.array([[1.0, 3.0], [2.0, 4.0]])
    b = np.array([[5.0, 7.0], [6.0, 8.0]])
    exp = np.outer(a[:, 0], b[:, 0]) + np.outer(a[:, 1], b[:, 1])
    assert np.max(np.abs(_sum_outer(O, a, b) - exp)) < 1e-10
    a = np.array([[1.0, 4.0], [2.0, 5.0], [3.0, 6.0]])
    b = np.array([[7.0, 9.0], [8.0, 10.0]])
    r = _sum_outer(O, a, b)
    assert r.shape == (3, 2)
    assert np.max(np.abs(r - (np.outer(a[:, 0], b[:, 0]) + np.outer(a[:, 1], b[:, 1])))) < 1e-10
    r = _sum_outer(O, np.array([[2.0], [3.0]]), np.array([[4.0], [5.0]]))
    assert np.max(np.abs(r - [[8, 10], [12, 15]])) < 1e-10
    assert np.all(_sum_outer(O, np.zeros((3, 2)), np.zeros((2, 2))) == 0)
    assert np.max(np.abs(_sum_outer(O, np.eye(2), np.eye(2)) - np.eye(2))) < 1e-10
    a = np.array([[1e-10, 3e-10], [2e-10, 4e-10]])
    b = np.array([[5e-10, 7e-10], [6e-10, 8e-10]])
    exp = np.outer(a[:, 0], b[:, 0]) + np.outer(a[:, 1], b[:, 1])
    assert np.max(np.abs(_sum_outer(O, a, b) - exp)) < 1e-25


# ---------------------------------------------------------------- F construction
@pytest.mark.parametrize("d", [1, 2, 4])
def test_construct_f_invariants(O, d):
    T = 257
    z = O.walk_normals(5, 0, d, T).astype(np.float64)
    B = O.brownian(z)
    bprev = B[:, :T]
    u = np.arange(1, T + 1) / T
    F0 = O.construct_f(0, bprev)
    assert np.array_equal(F0, bprev)
    F1 = O.construct_f(1, bprev)
    assert F1.shape == (d + 1, T) and np.all(F1[d] == 1.0)
    F2 = O.construct_f(2, bprev)
    assert F2.shape == (d, T)
    assert np.all(np.abs(F2[: d - 1].mean(axis=1)) < 1e-13)  # demeaned rows
    np.testing.assert_allclose(F2[d - 1], u - 0.5, rtol=0, atol=1e-15)  # trend row, not demeaned
    F3 = O.construct_f(3, bprev)
    assert F3.shape == (d + 1, T)
    assert np.all(np.abs(F3[:d].mean(axis=1)) < 1e-13)
    F4 = O.construct_f(4, bprev)
    assert F4.shape == (d, T)
    Z = np.vstack([np.ones(T), u])
    assert np.max(np.abs(F4 @ Z.T)) < 1e-9  # residuals orthogonal to 1 and u
    # numpy restatement == C restatement
    import ctypes

    for m in range(5):
        nf = O.num_eigenvalues(m, d)
        Fc = np.zeros((T, nf))
        bp = np.ascontiguousarray(bprev.T)
        P = ctypes.POINTER(ctypes.c_double)
        assert O.lib().jo_construct_f(m, d, T, bp.ctypes.data_as(P), Fc.ctypes.data_as(P)) == 0
        np.testing.assert_allclose(Fc.T, O.construct_f(m, bprev), rtol=0, atol=1e-13)


def test_ito_identity(O):
    """sum b e^T + e b^T = B_T B_T^T - sum e e^T (discrete Ito)."""
    z = O.walk_normals(11, 0, 3, 5000).astype(np.float64)
    B = O.brownian(z)
    b, e = B[:, :-1], np.diff(B, axis=1)
    lhs = b @ e.T + e @ b.T
    rhs = np.outer(B[:, -1], B[:, -1]) - e @ e.T
    assert np.max(np.abs(lhs - rhs)) < 1e-12


# ---------------------------------------------------------------- eigen solve
def test_closed_form_dim1_model0(O):
    """d = 1, model 0: rho = (sum b e)^2 / (dt * sum b^2)."""
    for seed in (1, 2, 3):
        T = 1000
        z = O.walk_normals(seed, 0, 1, T).astype(np.float64)
        B = O.brownian(z)[0]
        b, e = B[:-1], np.diff(B)
        rho = (b @ e) ** 2 / ((1.0 / T) * (b @ b))
        for ev in (O.walk_eigenvalues_lapack(0, z), O.walk_eigenvalues_c(0, z)):
            assert ev.shape == (1,)
            assert abs(ev[0] - rho) <= 1e-12 * rho


@pytest.mark.parametrize("model", range(5))
@pytest.mark.parametrize("d", [1, 2, 3, 6, 12])
def test_dggev_vs_jacobi(O, model, d):
    z = O.walk_normals(100 + d, 0, d, 400).astype(np.float64)
    a = O.walk_eigenvalues_lapack(model, z)
    c = O.walk_eigenvalues_c(model, z)
    assert a.shape == c.shape == (O.num_eigenvalues(model, d),)
    assert_eigs_close(c, a, 1e-10, f"model {model} d {d}")
    assert np.all(a >= 0) and np.all(np.diff(a) <= 0)
    if model in (1, 3):  # one structural zero eigenvalue
        assert a[-1] < 1e-10 * a[0]


@pytest.mark.parametrize("model", [0, 1, 3])
def test_rotation_invariance_full(O, model):
    rng = np.random.default_rng(3)
    d, T = 4, 300
    z = O.walk_normals(21, 0, d, T).astype(np.float64)
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    a = O.walk_eigenvalues_lapack(model, z)
    b = O.walk_eigenvalues_lapack(model, z @ Q.T)
    assert_eigs_close(b, a, 1e-9)


@pytest.mark.parametrize("model", [2, 4])
def test_rotation_invariance_leading_block(O, model):
    """Models 2 and 4 replace the last component; rotations of the first d-1
    components leave the eigenvalues unchanged."""
    rng = np.random.default_rng(4)
    d, T = 4, 300
    z = O.walk_normals(22, 0, d, T).astype(np.float64)
    Q = np.eye(d)
    Q[: d - 1, : d - 1], _ = np.linalg.qr(rng.standard_normal((d - 1, d - 1)))
    assert_eigs_close(O.walk_eigenvalues_lapack(model, z @ Q.T), O.walk_eigenvalues_lapack(model, z), 1e-9)


def test_walk_golden(O):
    fx = json.load(open(os.path.join(GOLDEN, "walk_eigenvalues.json")))
    for c in fx["cases"][::7]:
        ev = O.walk_eigenvalues_by_seed(c["model"], c["dim"], c["steps"], c["seed"])
        assert_eigs_close(ev, np.array(c["eigenvalues"]), 1e-12, str(c))
        cv = O.walks_by_seed_c(c["model"], c["dim"], c["steps"], [c["seed"]])[0]
        assert_eigs_close(cv, np.array(c["eigenvalues"]), 1e-10, str(c))


# ---------------------------------------------------------------- statistics
def test_percentile_hand_cases(O):
    """get_percentile_value (simulation_analyzers.rs:4-18): linear interpolation
    at rank p*(n-1)."""
    v = [1.0, 2.0, 3.0, 4.0]
    assert O.percentile_sorted(v, 0.5) == 2.5
    assert O.percentile_sorted(v, 0.0) == 1.0
    assert O.percentile_sorted(v, 1.0) == 4.0
    assert abs(O.percentile_sorted(v, 0.9) - 3.7) < 1e-15
    assert O.percentile_sorted([7.0], 0.3) == 7.0
    assert math.isnan(O.percentile_sorted([], 0.5))
    x = np.random.default_rng(0).standard_normal(1001)
    for p in (0.5, 0.9, 0.95, 0.99):
        assert abs(O.percentiles(x, [p])[0] - np.quantile(x, p, method="linear")) < 1e-12


# ---------------------------------------------------------------- ULEB128 and .dat
def test_uleb128_known_answers(O):
    """uleb128_unit_test.rs / uleb128.rs:62-102."""
    assert O.uleb128_encode(0) == b"\x00"
    assert O.uleb128_encode(127) == b"\x7f"
    assert O.uleb128_encode(128) == b"\x80\x01"
    assert O.uleb128_encode(300) == b"\xac\x02"
    assert O.uleb128_encode(16384) == b"\x80\x80\x01"
    assert O.uleb128_encode(2**32 - 1) == b"\xff\xff\xff\xff\x0f"
    for v in (0, 1, 127, 128, 300, 16383, 16384, 2**21, 2**28, 2**32 - 1):
        assert O.uleb128_decode(O.uleb128_encode(v)) == (v, len(O.uleb128_encode(v)))
    with pytest.raises(ValueError, match="IncompleteEncoding"):
        O.uleb128_decode(b"\x80")
    with pytest.raises(ValueError, match="EncodingTooLong"):
        O.uleb128_decode(b"\x80\x80\x80\x80\x80\x01")
    with pytest.raises(ValueError, match="ValueTooLarge"):
        O.uleb128_decode(b"\xff\xff\xff\xff\x10")


def test_dat_layout_known_answers(O):
    """DATA_FORMAT.md:100-149 examples (eigenvalues little-endian, as
    writer.rs:247-249 writes them with to_le_bytes)."""
    data = O.dat_bytes(0, 1, 10, [(1, [1.0]), (300, [2.0])])
    assert data[:18] == bytes.fromhex("45494745 4E56414C 535F5636 00 01 0A000000".replace(" ", ""))
    assert data[18:28] == b"\x01\x01" + np.float64(1.0).tobytes()
    assert data[28:39] == b"\xac\x02\x01" + np.float64(2.0).tobytes()
    assert data[-17:] == b"EOF_MARK" + (2).to_bytes(8, "little") + b"\x01"
    eof_example = bytes.fromhex("454F465F4D41524B 8096980000000000 01".replace(" ", ""))
    assert int.from_bytes(eof_example[8:16], "little") == 10_000_000
    recs, m, d, s = O.dat_parse(data)
    assert (m, d, s) == (0, 1, 10) and recs == [(1, [1.0]), (300, [2.0])]
    # no trailer -> scan path (append_writer_test.rs:66-90)
    recs, *_ = O.dat_parse(O.dat_bytes(0, 1, 10, [(1, [1.0]), (2, [3.0])], finish=False))
    assert recs == [(1, [1.0]), (2, [3.0])]


def test_expected_file_size_formula(O):
    """file_format.rs:15-26 on a real byte image, and the 10M-run, 1-eigenvalue
    size the code implies (DATA_FORMAT.md's 47,886,371 omits the 8-byte values)."""
    def expected(n, nf):
        seeds = sum(len(O.uleb128_encode(s)) for s in range(1, n + 1)) if n < 100000 else None
        return 18 + seeds + n + 8 * nf * n + 17

    recs = [(s, [float(s), 0.5]) for s in range(1, 400)]
    assert len(O.dat_bytes(2, 2, 77, recs)) == expected(399, 2)
    n = 10_000_000
    uleb_total = 127 * 1 + (16383 - 127) * 2 + (2097151 - 16383) * 3 + (n - 2097151) * 4
    assert 18 + uleb_total + n + 8 * n + 17 == 127_886_374
