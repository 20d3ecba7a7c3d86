"""The C-ABI library loads on a host without a GPU and exports every symbol
include/mann.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mann.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mann_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("mann_bf_knn", "mann_knn_graph", "mann_kmeans_fit", "mann_ivf_search",
                 "mann_route_distances"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from manifoldann_amd import _lib
    L = _lib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(_declared()) <= bound, sorted(set(_declared()) - bound)


def test_status_calls_without_gpu():
    from manifoldann_amd import _lib
    L = _lib.lib()
    assert L.mann_version() >= 10000
    assert L.mann_device_count() >= 0
    v = ctypes.c_int64()
    assert L.mann_tune_get(b"scan_cfg", ctypes.byref(v)) == 0
    assert L.mann_tune_set(b"no_such_knob", 1) == -2
    assert b"unknown tuning key" in L.mann_last_error()
    assert L.mann_tune_set(b"scan_cfg", 10) == -2


def test_argument_errors_map_to_reference_exceptions():
    from manifoldann_amd import _lib
    L = _lib.lib()
    # null handles are rejected before any device work
    assert L.mann_bf_knn(None, None, 4, 1, 4, 1, 0, None, None) == -6
    with pytest.raises(_lib.ArgumentError):
        _lib.check(-2)
    with pytest.raises(_lib.DimensionMismatch):
        _lib.check(-1)
    with pytest.raises(AssertionError):
        _lib.check(-3)
