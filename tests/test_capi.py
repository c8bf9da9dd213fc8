"""The C-ABI library without a GPU: it loads, exports every entry point that
include/*.h declares, and fails cleanly (no abort) when no device exists."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b(jne_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(jne):
    lib = jne.load_library()
    names = _declared_symbols()
    assert "jne_eigenvalues_by_seed" in names and len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_num_eigenvalues(jne):
    for d in range(1, 13):
        for m in range(5):
            assert jne.num_eigenvalues(m, d) == (d + 1 if m in (1, 3) else d)
    lib = jne.load_library()
    assert lib.jne_num_eigenvalues(5, 2) == jne.JNE_ERR_INVALID_ARG
    assert lib.jne_num_eigenvalues(0, 0) == jne.JNE_ERR_INVALID_ARG


def _gpu_present(jne):
    lib = jne.load_library()
    ctx = ctypes.c_void_p()
    rc = lib.jne_create(0, ctypes.byref(ctx))
    if rc == 0:
        lib.jne_destroy(ctx)
        return True
    return False


def test_create_without_device_reports_error(jne):
    if _gpu_present(jne):
        pytest.skip("a HIP device is present")
    with pytest.raises(jne.JneError) as e:
        jne.Engine(0)
    assert e.value.code == jne.JNE_ERR_HIP
    assert "jne_create" in str(e.value)


def test_null_context_is_rejected(jne):
    lib = jne.load_library()
    out = (ctypes.c_double * 4)()
    seeds = (ctypes.c_uint32 * 1)(1)
    assert lib.jne_eigenvalues_by_seed(None, 0, 2, 10, 0, seeds, 1, 0, out) == jne.JNE_ERR_INVALID_ARG
    assert lib.jne_synchronize(None) == jne.JNE_ERR_INVALID_ARG


def test_model_enum_mirror(jne):
    """JohansenModel mirror of src/johansen_models.rs:6-141 (johansen_models_test.rs)."""
    M = jne.JohansenModel
    assert [m.to_number() for m in M.all_models()] == [0, 1, 2, 3, 4]
    assert M.from_number(5) is None and M.from_number(3) == M.InterceptTrendUnrestrictedInterceptRestrictedTrend
    assert M.default() == M.InterceptNoTrendUnrestrictedIntercept
    assert str(M.NoInterceptNoTrend) == "Model 0: No intercept, no trend"
    assert not M.NoInterceptNoTrend.has_intercept() and M.InterceptNoTrendWithInterceptInCoint.has_intercept()
    assert M.InterceptTrendUnrestrictedBoth.has_trend() and not M.InterceptNoTrendUnrestrictedIntercept.has_trend()
    assert M.InterceptNoTrendWithInterceptInCoint.intercept_fully_explained_by_cointegration()
    assert M.InterceptTrendUnrestrictedInterceptRestrictedTrend.trend_fully_explained_by_cointegration()
