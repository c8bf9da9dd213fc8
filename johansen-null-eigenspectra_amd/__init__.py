"""Python binding of the MI355X johansen-null-eigenspectra engine.

A thin ctypes layer over ``lib/libjne.so`` (the C ABI of ``include/jne.h``).
The compute runs only in the HIP kernels of that library: if the library is
missing or no GPU is visible, the calls raise -- there is no CPU fallback.

Mirrors the reference's interfaces around the walk path:

* ``calculate_eigenvalues(dim, steps, seed, model)`` -- the per-walk operator
  of ``src/johansen_statistics.rs:59-85`` (returns |rho| sorted descending);
* ``Engine`` -- the batched forms used in place of the per-seed rayon loop of
  ``src/data_storage/parallel_compute.rs:14-41`` (host or device pointers, by
  seed or from caller-supplied normals, the device stream itself, and
  percentiles of a device array by the rule of ``src/simulation_analyzers.rs:4-18``);
* ``JohansenModel`` -- the model enum of ``src/johansen_models.rs:6-141``;
* ``EigenvalueSimulation`` -- the library facade of
  ``src/data_storage/simulation.rs`` (runs streamed to the ``.dat`` V6 file
  with resume, ``read_data``/``read_all_data``, trace and max-eigenvalue
  percentiles of ``src/simulation_analyzers.rs``);
* ``DatWriter``, ``read_dat``, ``dat_info``, ``remaining_seeds``,
  ``percentiles`` -- the ``.dat`` store of ``src/data_storage/``;
* ``Server`` -- critical-value queries (no counterpart in the reference):
  continuous batching on the device, quantiles selected on the device.

Precisions (``PRECISIONS``): ``fp64`` (the reference-equivalent path),
``mfma_fp16x2`` / ``mfma_bf16x3`` / ``mfma_fp32`` / ``mfma_bf16x2`` (moment
Gram on the matrix cores), ``fp32_kahan`` (compensated FP32 VALU check) and
``auto`` (the faster of fp64 and mfma_fp16x2 for each dim and length).
"""
from __future__ import annotations

import ctypes
import enum
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libjne.so")

JNE_OK = 0
JNE_ERR_INVALID_ARG = -1
JNE_ERR_HIP = -2
JNE_ERR_UNSUPPORTED = -3
JNE_ERR_NOMEM = -4
JNE_ERR_NUMERIC = -5
JNE_ERR_IO = -6
JNE_ERR_FORMAT = -7
JNE_ERR_TIMEOUT = -8

STAT_TRACE = 0
STAT_MAXEIG = 1

PREC_FP64 = 0
PREC_FP32_KAHAN = 1
PREC_MFMA_FP16X2 = 2
PREC_MFMA_BF16X2 = 3
PREC_MFMA_BF16X3 = 4
PREC_MFMA_FP32 = 5
PREC_AUTO = 6
PRECISIONS = {
    "fp64": PREC_FP64,
    "fp32_kahan": PREC_FP32_KAHAN,
    "mfma_fp16x2": PREC_MFMA_FP16X2,
    "mfma_bf16x2": PREC_MFMA_BF16X2,
    "mfma_bf16x3": PREC_MFMA_BF16X3,
    "mfma_fp32": PREC_MFMA_FP32,
    "auto": PREC_AUTO,
}

MAX_DIM = 32        # JNE_MAX_DIM: dims the device path runs
MAX_MIXED_DIM = 12  # JNE_MAX_MIXED_DIM: dims of the reduced precisions

# JNE_NOALLOC_SCRATCH_BYTES: the fixed per-GPU scratch of the no-alloc entry point
NOALLOC_SCRATCH_BYTES = 64 << 20


def noalloc_max_walks(dim: int) -> int:
    """Walks one no-alloc call takes at `dim`: the scratch over 8 (2.5 dim^2 + 6.5 dim + 1) bytes per
    walk up to dim 12, 8 (2 dim^2 + 6 dim) above (whose eigen solve needs no scratch)."""
    d = int(dim)
    m = d * (d + 1) // 2 + 1 if d <= MAX_MIXED_DIM else 0
    return (NOALLOC_SCRATCH_BYTES // 8) // (2 * d * d + 6 * d + m)

KERNEL_MOMENTS = 0
KERNEL_EIGEN = 1
KERNEL_NORMALS = 2

# JNE_STREAM_CONTEXT: the context's own stream (a stream handle of 0 is HIP's
# null stream, as in any HIP call)
STREAM_CONTEXT = ctypes.c_void_p(-1).value


def _stream_arg(hip_stream):
    """None: the context's own stream; an int: that hipStream_t (0 = HIP's null stream)."""
    return ctypes.c_void_p(STREAM_CONTEXT if hip_stream is None else int(hip_stream))

_lib = None
_lib_lock = threading.Lock()


class JneError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[jne error {code}] {message}")
        self.code = code


def load_library() -> ctypes.CDLL:
    """Load lib/libjne.so; raises if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (run __graft_entry__.build() or make -C {_HERE})")
        L = ctypes.CDLL(LIB_PATH)
        c_void_pp = ctypes.POINTER(ctypes.c_void_p)
        L.jne_create.argtypes = [ctypes.c_int, c_void_pp]
        L.jne_create.restype = ctypes.c_int
        L.jne_destroy.argtypes = [ctypes.c_void_p]
        L.jne_destroy.restype = None
        L.jne_last_error.argtypes = [ctypes.c_void_p]
        L.jne_last_error.restype = ctypes.c_char_p
        L.jne_num_eigenvalues.argtypes = [ctypes.c_int, ctypes.c_int]
        L.jne_num_eigenvalues.restype = ctypes.c_int
        L.jne_eigenvalues_by_seed.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                              ctypes.c_void_p]
        L.jne_eigenvalues_by_seed.restype = ctypes.c_int
        L.jne_eigenvalues_by_seed_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                                     ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t,
                                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.jne_eigenvalues_by_seed_device.restype = ctypes.c_int
        L.jne_eigenvalues_by_seed_device_noalloc.argtypes = L.jne_eigenvalues_by_seed_device.argtypes
        L.jne_eigenvalues_by_seed_device_noalloc.restype = ctypes.c_int
        L.jne_eigenvalues_by_seed_models_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                            ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                            ctypes.c_void_p, ctypes.c_void_p]
        L.jne_eigenvalues_by_seed_models_device.restype = ctypes.c_int
        L.jne_eigenvalues_from_normals.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                                   ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        L.jne_eigenvalues_from_normals.restype = ctypes.c_int
        L.jne_walk_normals.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.jne_walk_normals.restype = ctypes.c_int
        L.jne_walk_normals_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        L.jne_walk_normals_device.restype = ctypes.c_int
        L.jne_percentiles_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                             ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        L.jne_percentiles_device.restype = ctypes.c_int
        L.jne_synchronize.argtypes = [ctypes.c_void_p]
        L.jne_synchronize.restype = ctypes.c_int
        L.jne_profile_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jne_profile_enable.restype = ctypes.c_int
        L.jne_profile_read.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        L.jne_profile_read.restype = ctypes.c_int
        # storage, percentiles and the simulation driver
        L.jne_host_last_error.argtypes = []
        L.jne_host_last_error.restype = ctypes.c_char_p
        L.jne_dat_writer_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                          ctypes.c_uint64, ctypes.c_int, c_void_pp]
        L.jne_dat_writer_open.restype = ctypes.c_int
        L.jne_dat_writer_count.argtypes = [ctypes.c_void_p]
        L.jne_dat_writer_count.restype = ctypes.c_uint64
        L.jne_dat_writer_set_progress.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.jne_dat_writer_set_progress.restype = ctypes.c_int
        L.jne_dat_writer_append.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                            ctypes.c_int]
        L.jne_dat_writer_append.restype = ctypes.c_int
        L.jne_dat_writer_finish.argtypes = [ctypes.c_void_p]
        L.jne_dat_writer_finish.restype = ctypes.c_int
        L.jne_dat_writer_abandon.argtypes = [ctypes.c_void_p]
        L.jne_dat_writer_abandon.restype = None
        c_int_p = ctypes.POINTER(ctypes.c_int)
        c_u32_p = ctypes.POINTER(ctypes.c_uint32)
        c_u64_p = ctypes.POINTER(ctypes.c_uint64)
        L.jne_dat_read_info.argtypes = [ctypes.c_char_p, c_int_p, c_int_p, c_u32_p, c_u64_p, c_int_p, c_int_p]
        L.jne_dat_read_info.restype = ctypes.c_int
        L.jne_dat_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, c_u64_p,
                                   c_int_p]
        L.jne_dat_read.restype = ctypes.c_int
        L.jne_dat_remaining_seeds.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                              ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, c_u64_p, c_u64_p]
        L.jne_dat_remaining_seeds.restype = ctypes.c_int
        L.jne_percentiles.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_void_p]
        L.jne_percentiles.restype = ctypes.c_int
        L.jne_dat_statistics.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_uint64, c_u64_p]
        L.jne_dat_statistics.restype = ctypes.c_int
        L.jne_run_simulation.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                         ctypes.c_uint64, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.jne_run_simulation.restype = ctypes.c_int
        L.jne_run_simulation_models.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_int]
        L.jne_run_simulation_models.restype = ctypes.c_int
        L.jne_server_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, c_void_pp]
        L.jne_server_create.restype = ctypes.c_int
        L.jne_server_destroy.argtypes = [ctypes.c_void_p]
        L.jne_server_destroy.restype = None
        L.jne_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_int, c_u64_p]
        L.jne_server_submit.restype = ctypes.c_int
        L.jne_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double)]
        L.jne_server_wait.restype = ctypes.c_int
        L.jne_server_counters.argtypes = [ctypes.c_void_p, c_u64_p, ctypes.POINTER(ctypes.c_double)]
        L.jne_server_counters.restype = ctypes.c_int
        _lib = L
        return L


def _host_check(rc: int):
    if rc != JNE_OK:
        raise JneError(rc, load_library().jne_host_last_error().decode())


def _path(p) -> bytes:
    return os.fsencode(p)


def num_eigenvalues(model: int, dim: int) -> int:
    """dim+1 for models 1 and 3, else dim (src/data_storage/thread_manager.rs:40-44)."""
    n = load_library().jne_num_eigenvalues(int(model), int(dim))
    if n < 0:
        raise JneError(n, f"invalid model {model} or dim {dim}")
    return n


class JohansenModel(enum.IntEnum):
    """Model enum of src/johansen_models.rs:6-141 (ids 0..4, default 2)."""

    NoInterceptNoTrend = 0
    InterceptNoTrendWithInterceptInCoint = 1
    InterceptNoTrendUnrestrictedIntercept = 2
    InterceptTrendUnrestrictedInterceptRestrictedTrend = 3
    InterceptTrendUnrestrictedBoth = 4

    def to_number(self) -> int:
        return int(self)

    @classmethod
    def from_number(cls, n: int):
        try:
            return cls(int(n))
        except ValueError:
            return None

    def description(self) -> str:
        return _DESCRIPTIONS[int(self)]

    def has_intercept(self) -> bool:
        return self != JohansenModel.NoInterceptNoTrend

    def has_trend(self) -> bool:
        return int(self) in (3, 4)

    def intercept_fully_explained_by_cointegration(self) -> bool:
        return int(self) == 1

    def trend_fully_explained_by_cointegration(self) -> bool:
        return int(self) == 3

    @classmethod
    def all_models(cls):
        return [cls(i) for i in range(5)]

    @classmethod
    def default(cls):
        return cls.InterceptNoTrendUnrestrictedIntercept

    def __str__(self) -> str:  # Display impl, src/johansen_models.rs:137-141
        return f"Model {int(self)}: {self.description()}"


_DESCRIPTIONS = (
    "No intercept, no trend",
    "Intercept, no trend, intercept in cointegration",
    "Intercept, no trend, intercept not fully explained by cointegration",
    "Intercept, trend, trend in cointegration",
    "Intercept, trend, intercept and trend not fully explained by cointegration",
)


class Engine:
    """One jne_ctx: a HIP device + stream; calls on one Engine are serialised."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        ctx = ctypes.c_void_p()
        rc = self._lib.jne_create(int(device), ctypes.byref(ctx))
        if rc != JNE_OK:
            raise JneError(rc, self._lib.jne_last_error(None).decode())
        self._ctx = ctx

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.jne_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != JNE_OK:
            raise JneError(rc, self._lib.jne_last_error(self._ctx).decode())

    def eigenvalues_by_seed(self, model: int, dim: int, steps: int, seeds, rng_stream: int = 0,
                            precision: int = PREC_FP64) -> np.ndarray:
        s = np.ascontiguousarray(seeds, dtype=np.uint32)
        nf = num_eigenvalues(model, dim)
        out = np.empty((len(s), nf), dtype=np.float64)
        self._check(self._lib.jne_eigenvalues_by_seed(self._ctx, int(model), int(dim), int(steps), int(rng_stream),
                                                      s.ctypes.data, len(s), int(precision), out.ctypes.data))
        return out

    def eigenvalues_by_seed_device(self, model: int, dim: int, steps: int, d_seeds: int, n: int, d_out: int,
                                   rng_stream: int = 0, precision: int = PREC_FP64, hip_stream=None):
        """Device pointers (ints); enqueues on hip_stream (None: the context's stream; 0: HIP's null
        stream, e.g. torch's default stream) without synchronising."""
        self._check(self._lib.jne_eigenvalues_by_seed_device(
            self._ctx, int(model), int(dim), int(steps), int(rng_stream), ctypes.c_void_p(d_seeds), int(n),
            int(precision), ctypes.c_void_p(d_out), _stream_arg(hip_stream)))

    def eigenvalues_by_seed_models_device(self, models, dim: int, steps: int, d_seeds: int, n: int, d_outs,
                                          rng_stream: int = 0, precision: int = PREC_FP64, hip_stream=None):
        """Several models of the same walks (device pointers): d_outs[k] gets model models[k]'s
        eigenvalues; one moment pass per walk, one eigen solve per model."""
        m = np.ascontiguousarray(models, dtype=np.int32)
        outs = (ctypes.c_void_p * len(m))(*[ctypes.c_void_p(int(p)) for p in d_outs])
        self._check(self._lib.jne_eigenvalues_by_seed_models_device(
            self._ctx, m.ctypes.data, len(m), int(dim), int(steps), int(rng_stream), ctypes.c_void_p(d_seeds),
            int(n), int(precision), outs, _stream_arg(hip_stream)))

    def run_simulation_models(self, models, dim: int, steps: int, runs: int, paths, precision: int = PREC_FP64,
                              quiet: bool = True):
        """jne_run_simulation_models: the models of one dim to their .dat files (quiet: the models
        needing the same seeds share one moment pass per walk)."""
        m = np.ascontiguousarray(models, dtype=np.int32)
        enc = [os.fsencode(str(p)) for p in paths]
        cp = (ctypes.c_char_p * len(enc))(*enc)
        _host_check(self._lib.jne_run_simulation_models(self._ctx, m.ctypes.data, len(m), int(dim), int(steps),
                                                        int(runs), int(precision), cp, 1 if quiet else 0))

    def eigenvalues_by_seed_device_noalloc(self, model: int, dim: int, steps: int, d_seeds: int, n: int, d_out: int,
                                           rng_stream: int = 0, precision: int = PREC_FP64, hip_stream=None):
        """eigenvalues_by_seed_device without any allocation: the scratch is the library's fixed
        per-GPU array (NOALLOC_SCRATCH_BYTES), shared by every such call on the GPU (overlapping calls
        race); n is at most noalloc_max_walks(dim)."""
        self._check(self._lib.jne_eigenvalues_by_seed_device_noalloc(
            self._ctx, int(model), int(dim), int(steps), int(rng_stream), ctypes.c_void_p(d_seeds), int(n),
            int(precision), ctypes.c_void_p(d_out), _stream_arg(hip_stream)))

    def eigenvalues_from_normals(self, model: int, z, precision: int = PREC_FP64) -> np.ndarray:
        """z: (n, steps, dim) float64 N(0,1) draws."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        if z.ndim == 2:
            z = z[None]
        n, steps, dim = z.shape
        nf = num_eigenvalues(model, dim)
        out = np.empty((n, nf), dtype=np.float64)
        self._check(self._lib.jne_eigenvalues_from_normals(self._ctx, int(model), int(dim), int(steps),
                                                           z.ctypes.data, n, int(precision), out.ctypes.data))
        return out

    def walk_normals(self, dim: int, steps: int, seeds, rng_stream: int = 0) -> np.ndarray:
        s = np.ascontiguousarray(seeds, dtype=np.uint32)
        out = np.empty((len(s), steps, dim), dtype=np.float32)
        self._check(self._lib.jne_walk_normals(self._ctx, int(dim), int(steps), int(rng_stream), s.ctypes.data,
                                               len(s), out.ctypes.data))
        return out

    def walk_normals_device(self, dim: int, steps: int, d_seeds: int, n: int, d_out: int, rng_stream: int = 0,
                            hip_stream=None):
        """Device pointers (ints): d_out receives n*steps*dim floats; enqueues on hip_stream (None: the
        context's stream; 0: HIP's null stream)."""
        self._check(self._lib.jne_walk_normals_device(
            self._ctx, int(dim), int(steps), int(rng_stream), ctypes.c_void_p(d_seeds), int(n),
            ctypes.c_void_p(d_out), _stream_arg(hip_stream)))

    def percentiles_device(self, d_values: int, n: int, ps, hip_stream=None) -> list:
        """Percentiles ps (1..8 values in [0, 1]) of the device array d_values[0..n) (a device pointer,
        int), selected on the device; the rule and values of percentiles()."""
        p = np.ascontiguousarray(ps, dtype=np.float64)
        out = np.empty(max(len(p), 1), dtype=np.float64)
        self._check(self._lib.jne_percentiles_device(self._ctx, ctypes.c_void_p(d_values), int(n), p.ctypes.data,
                                                     len(p), out.ctypes.data, _stream_arg(hip_stream)))
        return [float(x) for x in out[:len(p)]]

    def synchronize(self):
        self._check(self._lib.jne_synchronize(self._ctx))

    def profile_enable(self, enable: bool = True):
        self._check(self._lib.jne_profile_enable(self._ctx, 1 if enable else 0))

    def profile_read(self, kernel: int):
        """(total_ms, launches, walks) of kernel KERNEL_MOMENTS, KERNEL_EIGEN or KERNEL_NORMALS."""
        ms = ctypes.c_double()
        nl = ctypes.c_uint64()
        nw = ctypes.c_uint64()
        self._check(self._lib.jne_profile_read(self._ctx, int(kernel), ctypes.byref(ms), ctypes.byref(nl),
                                               ctypes.byref(nw)))
        return ms.value, nl.value, nw.value


class DatWriter:
    """Append-only .dat V6 writer with resume (src/data_storage/writer.rs:15-302)."""

    def __init__(self, path, model: int, dim: int, steps: int, expected_records: int = 0, quiet: bool = True):
        self._lib = load_library()
        w = ctypes.c_void_p()
        _host_check(self._lib.jne_dat_writer_open(_path(path), int(model), int(dim), int(steps),
                                                  int(expected_records), 1 if quiet else 0, ctypes.byref(w)))
        self._w = w

    @property
    def count(self) -> int:
        return int(self._lib.jne_dat_writer_count(self._w))

    def set_progress(self, total_runs: int):
        _host_check(self._lib.jne_dat_writer_set_progress(self._w, int(total_runs)))

    def append(self, seeds, eigenvalues):
        s = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(-1)
        e = np.ascontiguousarray(eigenvalues, dtype=np.float64)
        if e.ndim == 1:
            e = e.reshape(len(s), -1) if len(s) else e.reshape(0, 0)
        _host_check(self._lib.jne_dat_writer_append(self._w, s.ctypes.data, e.ctypes.data, len(s),
                                                    e.shape[1] if e.ndim == 2 else 0))

    def finish(self):
        w, self._w = self._w, None
        _host_check(self._lib.jne_dat_writer_finish(w))

    def abandon(self):
        w, self._w = self._w, None
        if w:
            self._lib.jne_dat_writer_abandon(w)


def dat_info(path) -> dict:
    L = load_library()
    m, d, nf, tr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    st, n = ctypes.c_uint32(), ctypes.c_uint64()
    _host_check(L.jne_dat_read_info(_path(path), ctypes.byref(m), ctypes.byref(d), ctypes.byref(st), ctypes.byref(n),
                                    ctypes.byref(nf), ctypes.byref(tr)))
    return {"model": m.value, "dim": d.value, "steps": st.value, "records": n.value, "nf": nf.value,
            "has_trailer": bool(tr.value)}


def read_dat(path):
    """(seeds uint32[n], eigenvalues float64[n, nF], model, dim, steps): read_append_file (reader.rs:23-216)."""
    info = dat_info(path)
    n, nf = info["records"], max(info["nf"], 0)
    seeds = np.empty(n, dtype=np.uint32)
    eigs = np.empty((n, nf), dtype=np.float64)
    got, k = ctypes.c_uint64(), ctypes.c_int()
    _host_check(load_library().jne_dat_read(_path(path), seeds.ctypes.data, eigs.ctypes.data, n, ctypes.byref(got),
                                            ctypes.byref(k)))
    return seeds[:got.value], eigs[:got.value], info["model"], info["dim"], info["steps"]


def remaining_seeds(path, model: int, dim: int, steps: int, runs: int):
    """(remaining seeds 1..=runs not in the file, records in the file): progress.rs:10-61."""
    L = load_library()
    n, done = ctypes.c_uint64(), ctypes.c_uint64()
    _host_check(L.jne_dat_remaining_seeds(_path(path), int(model), int(dim), int(steps), int(runs), None, 0,
                                          ctypes.byref(n), ctypes.byref(done)))
    out = np.empty(n.value, dtype=np.uint32)
    _host_check(L.jne_dat_remaining_seeds(_path(path), int(model), int(dim), int(steps), int(runs), out.ctypes.data,
                                          n.value, ctypes.byref(n), ctypes.byref(done)))
    return out, done.value


def percentiles(values, ps) -> list:
    """Rank p*(n-1) with linear interpolation (src/simulation_analyzers.rs:4-18)."""
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    p = np.ascontiguousarray(ps, dtype=np.float64).reshape(-1)
    out = np.empty(len(p), dtype=np.float64)
    _host_check(load_library().jne_percentiles(v.ctypes.data, len(v), p.ctypes.data, len(p), out.ctypes.data))
    return out.tolist()


class Server:
    """Critical-value query server (continuous batching over `lanes` launch
    streams, HBM-footprint admission).  wait() returns
    (quantiles[6], latency_ms): trace p90, p95, p99, max-eigenvalue p90, p95, p99."""

    QUANTILES = (0.90, 0.95, 0.99)

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 8 << 30, lanes: int = 4):
        self._lib = load_library()
        s = ctypes.c_void_p()
        _host_check(self._lib.jne_server_create(int(device), int(hbm_budget_bytes), int(lanes), ctypes.byref(s)))
        self._s = s

    def submit(self, model: int, dim: int, steps: int, runs: int, seed: int, precision: str = "mfma_fp16x2") -> int:
        i = ctypes.c_uint64()
        _host_check(self._lib.jne_server_submit(self._s, int(model), int(dim), int(steps), int(runs), int(seed),
                                                PRECISIONS[precision], ctypes.byref(i)))
        return i.value

    def wait(self, qid: int, timeout_ms: int = -1):
        q = np.empty(6, dtype=np.float64)
        lat = ctypes.c_double()
        _host_check(self._lib.jne_server_wait(self._s, int(qid), int(timeout_ms), q.ctypes.data, ctypes.byref(lat)))
        return q, lat.value

    def counters(self):
        w, b = ctypes.c_uint64(), ctypes.c_double()
        _host_check(self._lib.jne_server_counters(self._s, ctypes.byref(w), ctypes.byref(b)))
        return w.value, b.value

    def close(self):
        if getattr(self, "_s", None):
            self._lib.jne_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class EigenvalueSimulation:
    """The library facade of src/data_storage/simulation.rs:9-123 and
    src/simulation_analyzers.rs:67-81 (LIBRARY_USAGE.md), on the device.

    Results go to data/eigenvalues_model{m}_dim{d}_steps{T}.dat under the
    current directory (data_dir overrides it); runs resume from that file.
    precision selects the moment accumulation (PRECISIONS); the default is
    the FP64 path."""

    def __init__(self, model, dim: int, steps: int, num_runs: int, precision: str = "fp64", engine=None,
                 data_dir: str = "data"):
        self.model = JohansenModel(int(model))
        self.dim = int(dim)
        self.steps = int(steps)
        self.num_runs = int(num_runs)
        self.precision = precision
        self.data_dir = data_dir
        self._engine = engine

    def __repr__(self):
        return (f"EigenvalueSimulation(model={self.model!r}, dim={self.dim}, steps={self.steps}, "
                f"num_runs={self.num_runs})")

    def _run(self, quiet: bool):
        eng = self._engine or default_engine()
        rc = eng._lib.jne_run_simulation(eng._ctx, int(self.model), self.dim, self.steps, self.num_runs,
                                         PRECISIONS[self.precision], _path(self.get_filename(self.model)),
                                         1 if quiet else 0)
        _host_check(rc)

    def run_simulation(self):
        self._run(False)

    def run_simulation_quiet(self):
        self._run(True)

    def read_all_data(self):
        """Every record of the file as [(seed, [eigenvalues])], file order (simulation.rs:48-53)."""
        seeds, eigs, *_ = read_dat(self.get_filename(self.model))
        return [(int(s), e.tolist()) for s, e in zip(seeds, eigs)]

    def _count_mismatch(self, n: int):
        """The error of read_data when the file holds n != num_runs records with seed <= num_runs."""
        return JneError(JNE_ERR_FORMAT,
                        f"Data count mismatch: expected {self.num_runs} records, found {n} records. "
                        f"This indicates incomplete simulation data.\n"
                        f"Solutions:\n"
                        f"1. Use `read_all_data()` to read available data ({n} records)\n"
                        f"2. Run `run_simulation()` to complete the simulation\n"
                        f"3. Create a new EigenvalueSimulation with num_runs={n} to match existing data")

    def read_data(self):
        """Records with seed <= num_runs; an error unless there are exactly num_runs (simulation.rs:55-89)."""
        data = [r for r in self.read_all_data() if r[0] <= self.num_runs]
        if len(data) != self.num_runs:
            raise self._count_mismatch(len(data))
        return data

    def get_filename(self, model=None) -> str:
        """data/eigenvalues_model{m}_dim{d}_steps{T}.dat; creates the folder (simulation.rs:91-122)."""
        m = int(self.model if model is None else model)
        os.makedirs(self.data_dir, exist_ok=True)
        return os.path.join(self.data_dir, f"eigenvalues_model{m}_dim{self.dim}_steps{self.steps}.dat")

    def _stat_percentiles(self, stat: int, ps):
        """calculate_percentiles (simulation_analyzers.rs:42-64): the records read_data keeps, each
        reduced by the statistic in C (jne_dat_statistics: a sequential sum, or a maximum folded from
        f64::MIN, as the reference's aggregators), then the percentile rule -- without a Python list
        of every record."""
        L = load_library()
        path = _path(self.get_filename(self.model))
        if self.num_runs == 0:  # read_data keeps no record (seeds start at 1): an empty result
            dat_info(self.get_filename(self.model))  # ... but a missing file is still an error
            return []
        n = ctypes.c_uint64()
        _host_check(L.jne_dat_statistics(path, self.num_runs, int(stat), None, 0, ctypes.byref(n)))
        if n.value != self.num_runs:
            raise self._count_mismatch(n.value)
        vals = np.empty(n.value, dtype=np.float64)
        _host_check(L.jne_dat_statistics(path, self.num_runs, int(stat), vals.ctypes.data, n.value,
                                         ctypes.byref(n)))
        return percentiles(vals, ps)

    def calculate_trace_percentiles(self, ps):
        """Percentiles of sum(rho) over the records (simulation_analyzers.rs:26-31,67-73)."""
        return self._stat_percentiles(STAT_TRACE, ps)

    def calculate_maxeig_percentiles(self, ps):
        """Percentiles of max(rho) over the records (simulation_analyzers.rs:34-39,75-80)."""
        return self._stat_percentiles(STAT_MAXEIG, ps)


_default_engine = None


def default_engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


def calculate_eigenvalues(dim: int, steps: int, seed: int, model) -> list:
    """Per-walk operator of src/johansen_statistics.rs:59-85 on the device."""
    return default_engine().eigenvalues_by_seed(int(model), dim, steps, [seed])[0].tolist()
