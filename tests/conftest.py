import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def jne():
    """The package johansen-null-eigenspectra_amd/ (imported as jne_amd)."""
    return __graft_entry__.import_package()


@pytest.fixture(scope="session")
def engine(jne):
    eng = jne.Engine(0)
    yield eng
    eng.close()


class DeviceArray:
    """A device allocation through the HIP runtime (ctypes), for the
    device-pointer entry points: from_host / to_host copy synchronously."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes

            cls._hip = ctypes.CDLL("libamdhip64.so")
        return cls._hip

    def __init__(self, nbytes: int):
        import ctypes

        self.nbytes = int(nbytes)
        self._p = ctypes.c_void_p()
        rc = self.hip().hipMalloc(ctypes.byref(self._p), ctypes.c_size_t(max(self.nbytes, 1)))
        assert rc == 0, f"hipMalloc({self.nbytes}) failed: {rc}"

    @property
    def ptr(self) -> int:
        return self._p.value

    def from_host(self, a, offset: int = 0):
        import ctypes

        import numpy as np

        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        rc = self.hip().hipMemcpy(ctypes.c_void_p(self.ptr + offset), ctypes.c_void_p(a.ctypes.data),
                                  ctypes.c_size_t(a.nbytes), 1)
        assert rc == 0, rc

    def to_host(self, shape, dtype, offset: int = 0):
        import ctypes

        import numpy as np

        out = np.empty(shape, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        assert self.hip().hipDeviceSynchronize() == 0
        rc = self.hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr + offset),
                                  ctypes.c_size_t(out.nbytes), 2)
        assert rc == 0, rc
        return out

    def free(self):
        if self._p.value:
            self.hip().hipFree(self._p)
            self._p.value = None

    @classmethod
    def stream(cls) -> int:
        """A new HIP stream (default priority) as an integer handle."""
        import ctypes

        st = ctypes.c_void_p()
        assert cls.hip().hipStreamCreate(ctypes.byref(st)) == 0
        return st.value

    @classmethod
    def destroy_stream(cls, st: int):
        import ctypes

        cls.hip().hipStreamDestroy(ctypes.c_void_p(st))


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle


def assert_eigs_close(got, ref, rtol, what="", atol=1e-13):
    """Elementwise |got - ref| <= rtol * |ref| + atol * rho_max.

    The absolute term (default 1e-13 rho_max, a few hundred ulps of the
    largest eigenvalue) covers what double precision cannot resolve
    relatively: the structural zero eigenvalue of models 1 and 3 (rounding
    noise) and eigenvalues many orders below rho_max, whose backward-stable
    solves (LAPACK dggev in the oracle, Cholesky + Jacobi on the device) agree
    only to about eps * rho_max."""
    import numpy as np

    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, f"{what}: shape {got.shape} vs {ref.shape}"
    rho_max = np.max(np.abs(ref), axis=-1, keepdims=True)
    tol = rtol * np.abs(ref) + atol * rho_max
    bad = np.abs(got - ref) > tol
    if bad.any():
        idx = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: first mismatch at {tuple(idx)}: got {got[tuple(idx)]!r} "
                             f"ref {ref[tuple(idx)]!r} (rtol {rtol})")
