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


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle


def assert_eigs_close(got, ref, rtol, what=""):
    """|got - ref| <= rtol * max(|ref|, rho_max * 1) elementwise with the
    structural-zero floor: entries are compared against rtol*|ref|, or
    rtol*rho_max for entries below rtol*rho_max (models 1 and 3 carry one
    eigenvalue that is zero up to rounding)."""
    import numpy as np

    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, f"{what}: shape {got.shape} vs {ref.shape}"
    rho_max = np.max(np.abs(ref), axis=-1, keepdims=True)
    tol = np.maximum(rtol * np.abs(ref), rtol * rho_max)
    bad = np.abs(got - ref) > tol
    if bad.any():
        idx = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: first mismatch at {tuple(idx)}: got {got[tuple(idx)]!r} "
                             f"ref {ref[tuple(idx)]!r} (rtol {rtol})")
