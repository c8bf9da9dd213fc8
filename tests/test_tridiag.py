"""The eigenvalue solves of csrc/jne_tridiag.h built for the host (CPU).

The tridiagonal QL solve (the device's solve of M = W^T W for dims 7-12) is
checked against LAPACK's symmetric eigensolver (numpy.linalg.eigvalsh) on
matrices of every size 1-12: Wishart-like, graded scales, clustered and
repeated eigenvalues, rank-deficient, already diagonal or tridiagonal, and
eleven decades of dynamic range.  The Sturm-count bisection (the device's
solve of the tridiagonal form for dims 13-32) is checked against LAPACK's
tridiagonal solver (scipy.linalg.eigvalsh_tridiagonal) on tridiagonal
matrices of every size 1-32, among them the reductions of the same kinds of
matrices.  The bar is the engine's eigenvalue parity bar,
|d lambda| <= 1e-9 |lambda| + 1e-13 lambda_max."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "johansen-null-eigenspectra_amd", "csrc")


@pytest.fixture(scope="module")
def tri(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    so = str(tmp_path_factory.mktemp("tri") / "libtri.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, os.path.join(HERE, "tridiag_host.cpp"),
                    "-o", so], check=True)
    lib = ctypes.CDLL(so)
    lib.eig.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.sturm_eigs.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]

    def solve(m):
        n = m.shape[0]
        a = np.array([m[i, j] for i in range(n) for j in range(i, n)], dtype=np.float64)
        out = np.zeros(n)
        assert lib.eig(n, a.ctypes.data, out.ctypes.data) == 0
        return out

    def sturm(d, e):
        d = np.ascontiguousarray(d, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        out = np.zeros(len(d))
        assert lib.sturm_eigs(len(d), d.ctypes.data, e.ctypes.data, out.ctypes.data) == 0
        return out
    solve.sturm = sturm
    return solve


def _matrices(rng, n, kind):
    if kind == "wishart":
        w = rng.standard_normal((n + int(rng.integers(0, 3)), n))
        return w.T @ w
    if kind == "graded":
        w = rng.standard_normal((n + 1, n)) * np.logspace(0, -6, n)
        return w.T @ w
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if kind == "clustered":
        ev = 1.0 + 1e-9 * rng.standard_normal(n)
        ev[: n // 2] = 3.0
    elif kind == "range":
        ev = np.logspace(3, -8, n)
    elif kind == "rank":
        ev = np.where(np.arange(n) < max(1, n // 2), rng.random(n) + 0.5, 0.0)
    else:  # "tridiagonal": already reduced
        m = np.diag(rng.random(n) * 10)
        if n > 1:
            off = rng.random(n - 1) * (rng.random(n - 1) < 0.7)
            m += np.diag(off, 1) + np.diag(off, -1)
        return m
    m = (q * ev) @ q.T
    return (m + m.T) / 2


KINDS = ["wishart", "graded", "clustered", "range", "rank", "tridiagonal"]


@pytest.mark.parametrize("kind", KINDS)
def test_tridiag_ql_matches_lapack(tri, kind):
    rng = np.random.default_rng(1000 + KINDS.index(kind))
    worst = 0.0
    for trial in range(400):
        n = 1 + trial % 12
        m = _matrices(rng, n, kind)
        ref = np.sort(np.linalg.eigvalsh(m))
        got = np.sort(tri(m))
        assert np.isfinite(got).all()
        bar = 1e-9 * np.abs(ref) + 1e-13 * np.abs(ref).max() + 1e-300
        worst = max(worst, float((np.abs(got - ref) / bar).max()))
    assert worst <= 1.0, worst


def _tridiagonals(rng, n, kind):
    """(d, e) of an n x n symmetric tridiagonal matrix."""
    if kind in ("wishart", "graded", "clustered", "range", "rank"):
        from scipy.linalg import hessenberg
        h = hessenberg(_matrices(rng, n, kind))  # symmetric: tridiagonal up to rounding
        return np.diag(h).copy(), np.diag(h, -1).copy()
    if kind == "split":  # zero couplings: independent blocks
        d = rng.standard_normal(n)
        e = rng.standard_normal(n - 1) * (rng.random(n - 1) < 0.6)
        return d, e
    if kind == "wilkinson":  # pairs of eigenvalues agreeing to many digits
        m = (n - 1) / 2
        return np.abs(np.arange(n) - m), np.ones(n - 1)
    # "tiny": couplings from 1 down to 1e-12, diagonal over eight decades
    return np.logspace(0, -8, n) * rng.random(n), np.logspace(0, -12, max(n - 1, 0)) * rng.standard_normal(n - 1)


TRI_KINDS = ["wishart", "graded", "clustered", "range", "rank", "split", "wilkinson", "tiny"]


@pytest.mark.parametrize("kind", TRI_KINDS)
def test_sturm_bisection_matches_lapack(tri, kind):
    from scipy.linalg import eigvalsh_tridiagonal
    rng = np.random.default_rng(2000 + TRI_KINDS.index(kind))
    worst = 0.0
    for trial in range(160):
        n = 1 + trial % 32
        d, e = _tridiagonals(rng, n, kind)
        ref = eigvalsh_tridiagonal(d, e) if n > 1 else d.copy()
        got = tri.sturm(d, e)
        assert np.isfinite(got).all() and (np.diff(got) >= 0).all()
        bar = 1e-9 * np.abs(ref) + 1e-13 * np.abs(ref).max() + 1e-300
        worst = max(worst, float((np.abs(got - ref) / bar).max()))
    assert worst <= 1.0, worst
