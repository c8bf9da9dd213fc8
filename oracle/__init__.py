"""CPU oracle for the johansen-null-eigenspectra walk.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline).  The shipped library never calls it.

Two restatements of the reference's per-walk algorithm live here:

* ``walk_eigenvalues_lapack`` -- numpy restatement of
  ``src/rng_matrix.rs:57-141`` (Brownian path), ``src/johansen_statistics.rs``
  ``:59-85`` (walk driver), ``:102-197`` (``construct_f_matrix``),
  ``src/matrix_utils.rs:67-85`` (``sum_of_outer_products``) and the eigen
  solve ``src/johansen_statistics.rs:35-45``, which calls LAPACK ``dggev``
  exactly as the reference does (through scipy's LAPACK binding) and returns
  ``|alpha|/beta`` sorted descending.
* ``walk_eigenvalues_c`` -- the plain-C restatement in ``jne_oracle.c`` (same
  construction, Cholesky + cyclic Jacobi eigen solve), used for speed and as
  the CPU baseline.

It also holds the build's random stream (Philox4x32-10 + Box-Muller with
polynomial transcendentals) in C, and restatements of the reference's
statistics (``src/simulation_analyzers.rs:4-65``), ULEB128 codec
(``src/data_storage/uleb128.rs:68-138``) and ``.dat`` V6 byte format
(``src/data_storage/writer.rs:61-64,241-249,262-289``,
``src/data_storage/reader.rs:23-216``).

Parity pins (see tests/test_oracle.py): Random123 known-answer vectors for
Philox; the reference's own N(0,1) quantile fixture for the normals; closed
forms (d=1 model 0, Ito identity, demeaning/detrending invariants); LAPACK
dggev vs the C Jacobi solve; ULEB128 and .dat known answers from the
reference's tests and DATA_FORMAT.md.
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libjne_oracle.so")
_lib = None


def build() -> str:
    """Compile the C oracle (gcc) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        f32p = ctypes.POINTER(ctypes.c_float)
        f64p = ctypes.POINTER(ctypes.c_double)
        L.jo_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.jo_normal_pair.argtypes = [ctypes.c_uint32, ctypes.c_uint32, f32p]
        L.jo_walk_normals.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, f32p]
        L.jo_num_eigenvalues.argtypes = [ctypes.c_int, ctypes.c_int]
        L.jo_brownian.argtypes = [ctypes.c_int, ctypes.c_uint32, f64p, f64p]
        L.jo_construct_f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, f64p, f64p]
        L.jo_sum_outer.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, f64p, f64p, f64p]
        L.jo_gen_eig_sym.argtypes = [ctypes.c_int, f64p, f64p, f64p]
        L.jo_walk_from_normals.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, f64p, f64p]
        L.jo_walk_moments.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, f64p, f64p, f64p]
        L.jo_walks_by_seed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                       u32p, ctypes.c_size_t, ctypes.c_int, f64p]
        L.jo_percentile_sorted.argtypes = [f64p, ctypes.c_size_t, ctypes.c_double]
        L.jo_percentile_sorted.restype = ctypes.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


# --------------------------------------------------------------------------
# random stream
# --------------------------------------------------------------------------
JNE_KEY1 = 0x4A4E4531


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().jo_philox4x32_10(_ptr(c, ctypes.c_uint32), _ptr(k, ctypes.c_uint32), _ptr(out, ctypes.c_uint32))
    return out


def normal_pair(wa: int, wb: int) -> np.ndarray:
    out = np.zeros(2, dtype=np.float32)
    lib().jo_normal_pair(int(wa), int(wb), _ptr(out, ctypes.c_float))
    return out


def walk_normals(seed: int, stream: int, d: int, T: int) -> np.ndarray:
    """Normals of one walk as a (T, d) float32 array (z[t, i])."""
    z = np.zeros(T * d, dtype=np.float32)
    lib().jo_walk_normals(int(seed), int(stream), int(d), int(T), _ptr(z, ctypes.c_float))
    return z.reshape(T, d)


def num_eigenvalues(model: int, d: int) -> int:
    """dim+1 for models 1 and 3, else dim (thread_manager.rs:40-44)."""
    return d + 1 if model in (1, 3) else d


# --------------------------------------------------------------------------
# numpy restatement of the reference walk + LAPACK dggev
# --------------------------------------------------------------------------
def brownian(z: np.ndarray) -> np.ndarray:
    """B (d x (T+1)) from normals z (T x d): rng_matrix.rs:135-140.

    scaled = [0 | z^T] * sqrt(1/T); B = cumsum along time (sequential)."""
    T, d = z.shape
    sqrt_dt = np.sqrt(1.0 / T)
    scaled = np.zeros((d, T + 1), dtype=np.float64)
    scaled[:, 1:] = z.T.astype(np.float64) * sqrt_dt
    return np.cumsum(scaled, axis=1)


def construct_f(model: int, bprev: np.ndarray) -> np.ndarray:
    """construct_f_matrix (johansen_statistics.rs:102-197) on bm_previous (d x T)."""
    rows, cols = bprev.shape
    t = float(cols)
    if model == 0:
        return bprev.copy()
    if model == 1:
        return np.vstack([bprev, np.ones((1, cols))])
    y_centered = (np.arange(1, cols + 1, dtype=np.float64) / t) - 0.5
    if model == 2:
        x = bprev[: rows - 1]
        x = x - x.sum(axis=1, keepdims=True) / cols
        return np.vstack([x, y_centered[None, :]])
    if model == 3:
        x = bprev - bprev.sum(axis=1, keepdims=True) / cols
        return np.vstack([x, y_centered[None, :]])
    if model == 4:
        y = np.arange(1, cols + 1, dtype=np.float64) / t
        xw = np.vstack([bprev[: rows - 1], (y * y)[None, :]])
        z = np.vstack([np.ones(cols), y])
        zzt = z @ z.T
        det = zzt[0, 0] * zzt[1, 1] - zzt[1, 0] * zzt[0, 1]
        inv = np.array([[zzt[1, 1], -zzt[0, 1]], [-zzt[1, 0], zzt[0, 0]]]) / det
        return xw - ((xw @ z.T) @ inv) @ z
    raise ValueError(f"invalid model {model}")


def walk_moments(model: int, z: np.ndarray):
    """(S01raw, S11) of one walk, johansen_statistics.rs:30-33 and :80-82."""
    T, d = z.shape
    bm = brownian(z)
    bprev = bm[:, :T]
    dbm = bm[:, 1:] - bm[:, :T]
    fm = construct_f(model, bprev)
    s01raw = dbm @ fm.T                     # sum_t dB_t F_t^T  (d x nF)
    s11 = (fm @ fm.T) * (1.0 / T)           # (sum_t F_t F_t^T) * dt
    return s01raw, s11


def eig_dggev(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """|alpha|/beta from LAPACK dggev, sorted descending (johansen_statistics.rs:35-45)."""
    from scipy.linalg import lapack

    alphar, alphai, beta, _vl, _vr, _work, info = lapack.dggev(
        np.asfortranarray(A, dtype=np.float64), np.asfortranarray(B, dtype=np.float64),
        compute_vl=0, compute_vr=0)
    if info != 0:
        raise RuntimeError(f"dggev failed: info={info}")
    rho = np.hypot(alphar, alphai) / beta
    return np.sort(rho)[::-1].copy()


def walk_eigenvalues_lapack(model: int, z: np.ndarray) -> np.ndarray:
    """One walk from normals z (T x d) through the reference algorithm + dggev."""
    s01raw, s11 = walk_moments(model, np.asarray(z, dtype=np.float64))
    return eig_dggev(s01raw.T @ s01raw, s11)


# --------------------------------------------------------------------------
# C restatement
# --------------------------------------------------------------------------
def walk_eigenvalues_c(model: int, z: np.ndarray) -> np.ndarray:
    z = np.ascontiguousarray(z, dtype=np.float64)
    T, d = z.shape
    out = np.zeros(num_eigenvalues(model, d), dtype=np.float64)
    rc = lib().jo_walk_from_normals(model, d, T, _ptr(z, ctypes.c_double), _ptr(out, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"jo_walk_from_normals rc={rc}")
    return out


def walk_moments_c(model: int, z: np.ndarray):
    z = np.ascontiguousarray(z, dtype=np.float64)
    T, d = z.shape
    nf = num_eigenvalues(model, d)
    s01 = np.zeros((nf, d), dtype=np.float64)  # column-major d x nF
    s11 = np.zeros((nf, nf), dtype=np.float64)
    rc = lib().jo_walk_moments(model, d, T, _ptr(z, ctypes.c_double), _ptr(s01, ctypes.c_double),
                               _ptr(s11, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"jo_walk_moments rc={rc}")
    return s01.T.copy(), s11.T.copy()


def gen_eig_sym_c(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.float64)
    b = np.asfortranarray(B, dtype=np.float64)
    out = np.zeros(n, dtype=np.float64)
    rc = lib().jo_gen_eig_sym(n, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(out, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"jo_gen_eig_sym rc={rc}")
    return out


def walks_by_seed_c(model: int, d: int, T: int, seeds, stream: int = 0, threads: int = 1) -> np.ndarray:
    s = np.ascontiguousarray(seeds, dtype=np.uint32)
    nf = num_eigenvalues(model, d)
    out = np.zeros((len(s), nf), dtype=np.float64)
    rc = lib().jo_walks_by_seed(model, d, T, stream, _ptr(s, ctypes.c_uint32), len(s), threads,
                                _ptr(out, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"jo_walks_by_seed rc={rc}")
    return out


def walk_eigenvalues_by_seed(model: int, d: int, T: int, seed: int, stream: int = 0) -> np.ndarray:
    """Reference algorithm + dggev on the build's stream for one seed."""
    return walk_eigenvalues_lapack(model, walk_normals(seed, stream, d, T))


# --------------------------------------------------------------------------
# statistics (simulation_analyzers.rs:4-65)
# --------------------------------------------------------------------------
def percentile_sorted(sorted_values, p: float) -> float:
    n = len(sorted_values)
    if n == 0:
        return float("nan")
    rank = p * (n - 1)
    lo, hi = int(np.floor(rank)), int(np.ceil(rank))
    if lo == hi:
        return float(sorted_values[lo])
    w = rank - lo
    return float(sorted_values[lo] * (1.0 - w) + sorted_values[hi] * w)


def trace_stat(eigs: np.ndarray) -> np.ndarray:
    return np.asarray(eigs).sum(axis=-1)


def maxeig_stat(eigs: np.ndarray) -> np.ndarray:
    return np.asarray(eigs).max(axis=-1)


def percentiles(values, ps):
    v = np.sort(np.asarray(values, dtype=np.float64))
    return [percentile_sorted(v, p) for p in ps]


# --------------------------------------------------------------------------
# ULEB128 (uleb128.rs:68-138) and .dat V6 (writer.rs / reader.rs)
# --------------------------------------------------------------------------
MAGIC = b"EIGENVALS_V6"
EOF_MARK = b"EOF_MARK"


def uleb128_encode(value: int) -> bytes:
    out = bytearray()
    while True:
        byte = value & 0x7F
        value >>= 7
        if value:
            byte |= 0x80
        out.append(byte)
        if not value:
            return bytes(out)


def uleb128_decode(data: bytes):
    """(value, bytes_read) or raises ValueError with the reference's error names."""
    result, shift = 0, 0
    for i, byte in enumerate(data):
        if i + 1 > 5:
            raise ValueError("EncodingTooLong")
        if shift >= 32:
            raise ValueError("ValueTooLarge")
        bits = byte & 0x7F
        if shift == 28 and bits > 0x0F:
            raise ValueError("ValueTooLarge")
        result |= bits << shift
        shift += 7
        if not byte & 0x80:
            return result, i + 1
    raise ValueError("IncompleteEncoding")


def dat_bytes(model: int, dim: int, steps: int, records, finish: bool = True) -> bytes:
    """Byte image of a V6 file written record by record (writer.rs)."""
    out = bytearray(MAGIC + struct.pack("<BBI", model, dim & 0xFF, steps))
    for seed, eigs in records:
        out += uleb128_encode(seed) + struct.pack("<B", len(eigs))
        out += struct.pack(f"<{len(eigs)}d", *eigs)
    if finish:
        nf = len(records[0][1]) if records else 0
        out += EOF_MARK + struct.pack("<QB", len(records), nf)
    return bytes(out)


def dat_parse(data: bytes):
    """(records, model, dim, steps) following reader.rs:23-216 (fast or scan path)."""
    if data[:12] != MAGIC:
        raise ValueError("File format error: magic header mismatch")
    model, dim, steps = struct.unpack("<BBI", data[12:18])
    if len(data) < 35:
        return [], model, dim, steps
    records = []
    if data[-17:-9] == EOF_MARK:
        count, nf = struct.unpack("<QB", data[-9:])
        pos = 18
        for _ in range(count):
            seed, n = uleb128_decode(data[pos:pos + 5])
            pos += n
            k = data[pos]
            pos += 1
            if k == 0:
                raise ValueError("Invalid eigenvalue count: cannot be zero")
            if k != nf:
                raise ValueError("Eigenvalue count mismatch")
            records.append((seed, list(struct.unpack(f"<{k}d", data[pos:pos + 8 * k]))))
            pos += 8 * k
        return records, model, dim, steps
    pos = 18
    while True:
        try:
            seed, n = uleb128_decode(data[pos:pos + 5])
        except ValueError:
            break
        pos += n
        if pos >= len(data):
            break
        k = data[pos]
        pos += 1
        if k == 0 and data[pos:pos + 8] == EOF_MARK:
            break
        if pos + 8 * k > len(data):
            break
        records.append((seed, list(struct.unpack(f"<{k}d", data[pos:pos + 8 * k]))))
        pos += 8 * k
    return records, model, dim, steps
