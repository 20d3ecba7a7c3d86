"""Loader of the gfx950 HIP library (libmann_hip.so) through its C ABI.

There is no CPU fallback: every call goes to the HIP library, and a missing or
unloadable library raises immediately.
"""
from __future__ import annotations

import atexit
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmann_hip.so")

_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.c_void_p  # host or device pointers are both passed as addresses
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every symbol include/mann.h declares
SIGNATURES = [
    ("mann_version", ctypes.c_int, []),
    ("mann_last_error", ctypes.c_char_p, []),
    ("mann_device_count", ctypes.c_int, []),
    ("mann_dataset_create", ctypes.c_int, [_vp, _i64, _i64, _i64, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("mann_dataset_create_f16", ctypes.c_int, [_vp, _i64, _i64, _i64, ctypes.c_int,
                                                ctypes.POINTER(_vp)]),
    ("mann_dataset_destroy", ctypes.c_int, [_vp]),
    ("mann_dataset_info", ctypes.c_int, [_vp, _i64p, _i64p]),
    ("mann_dataset_set_first_pass", ctypes.c_int, [_vp, ctypes.c_int]),
    ("mann_bf_knn_graph_create", ctypes.c_int,
     [_vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_int, _vp, _vp, _vp]),
    ("mann_graph_launch", ctypes.c_int, [_vp]),
    ("mann_graph_destroy", ctypes.c_int, [_vp]),
    ("mann_ivf_search_graph_create", ctypes.c_int,
     [_vp, _vp, _i64, _i64, _i64, _i64, _i64, ctypes.c_int, _vp, _vp, _vp, _vp]),
    ("mann_autotune_bf_knn", ctypes.c_int,
     [_vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_int, _i64p, ctypes.POINTER(ctypes.c_double)]),
    ("mann_autotune_ivf", ctypes.c_int,
     [_vp, _vp, _i64, _i64, _i64, _i64, _i64, ctypes.c_int, _vp, _i64p,
      ctypes.POINTER(ctypes.c_double)]),
    ("mann_autotune_clear", ctypes.c_int, []),
    ("mann_bf_knn", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_int, _vp, _vp]),
    ("mann_bf_knn_noalloc", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_int, _vp,
                                           _vp]),
    ("mann_knn_graph", ctypes.c_int, [_vp, _i64, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("mann_knn_candidates", ctypes.c_int, [_vp, _i64, ctypes.c_int, _vp, _vp, _vp]),
    ("mann_knn_graph_symmetrize", ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _i64p, _i64p]),
    ("mann_kmeans_assign", ctypes.c_int, [_vp, _vp, _i64, _vp, _vp]),
    ("mann_kmeanspp", ctypes.c_int, [_vp, _i64, ctypes.c_uint64, _vp, _vp]),
    ("mann_kmeans_fit", ctypes.c_int, [_vp, _i64, _i64, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_uint64, _vp, _i64, _vp, _vp, _i64p]),
    ("mann_ivf_build", ctypes.c_int, [_vp, _vp, _i64, _vp, ctypes.POINTER(_vp)]),
    ("mann_ivf_flat_build", ctypes.c_int, [_vp, _vp, _i64, _vp, ctypes.c_int,
                                           ctypes.POINTER(_vp)]),
    ("mann_ivf_destroy", ctypes.c_int, [_vp]),
    ("mann_ivf_info", ctypes.c_int, [_vp, _i64p, _vp]),
    ("mann_ivf_list_ids", ctypes.c_int, [_vp, _vp]),
    ("mann_ivf_route", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _vp]),
    ("mann_ivf_probe", ctypes.c_int, [_vp, _vp, _i64, _i64, _vp]),
    ("mann_ivf_search", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, _i64, ctypes.c_int, _vp,
                                       _vp, _vp]),
    ("mann_ivf_search_ex", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, _i64, ctypes.c_int,
                                          _vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("mann_ivf_hnsw_build", ctypes.c_int, [_vp, _i64, _i64, ctypes.c_int, ctypes.c_uint64]),
    ("mann_ivf_hnsw_search", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp,
                                            ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("mann_ivf_hnsw_set_graph", ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    ("mann_ivf_hnsw_graph", ctypes.c_int, [_vp, _i64p, _i64p, _vp, _vp, _vp, _vp, _vp]),
    ("mann_route_distances", ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    ("mann_router_create", ctypes.c_int, [_i64, _i64, _i64, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("mann_router_destroy", ctypes.c_int, [_vp]),
    ("mann_router_params", ctypes.c_int, [_vp, _vp, _i64p]),
    ("mann_router_set_params", ctypes.c_int, [_vp, _vp]),
    ("mann_router_set_prior", ctypes.c_int, [_vp, _vp, ctypes.c_double]),
    ("mann_router_grad", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _vp,
                                        ctypes.POINTER(ctypes.c_double)]),
    ("mann_router_train", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                                         ctypes.c_double, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_double)]),
    ("mann_router_scores", ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _vp]),
    ("mann_ivf_search_routed", ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                              ctypes.c_int, _vp, _vp]),
    ("mann_ivf_probe_mass", ctypes.c_int, [_vp, _vp, _i64, _i64, ctypes.c_double, _vp]),
    ("mann_ivf_search_routed_mass", ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                                   ctypes.c_double, ctypes.c_int, _vp, _vp]),
    ("mann_last_probe_count", ctypes.c_int, [_i64p]),
    ("mann_last_emitted", ctypes.c_int, [_i64p]),
    ("mann_tune_set", ctypes.c_int, [ctypes.c_char_p, _i64]),
    ("mann_tune_get", ctypes.c_int, [ctypes.c_char_p, _i64p]),
    ("mann_profile_enable", ctypes.c_int, [ctypes.c_int]),
    ("mann_profile_reset", ctypes.c_int, []),
    ("mann_profile_query", ctypes.c_int, [ctypes.c_char_p, _i64p, ctypes.POINTER(ctypes.c_double)]),
    ("mann_last_stats", ctypes.c_int, [_i64p]),
]

_lib = None
_shutting_down = False


def _mark_shutdown():
    global _shutting_down
    _shutting_down = True


atexit.register(_mark_shutdown)


def shutting_down() -> bool:
    """True once the interpreter exits: handles are then left to the process
    teardown instead of freeing device memory under a stopping runtime."""
    return _shutting_down


class MannError(RuntimeError):
    pass


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch."""


class ArgumentError(ValueError):
    """Julia ArgumentError."""


def lib():
    """Load libmann_hip.so (built by manifoldann_amd.build); raise if absent.

    torch ships its own HIP runtime; when torch is already imported its runtime
    is initialised first, so both runtimes see the device.
    """
    global _lib
    if _lib is None:
        import sys
        if "torch" in sys.modules:
            t = sys.modules["torch"]
            try:
                if t.cuda.is_available():
                    t.cuda.init()
            except Exception:
                pass
        if not os.path.exists(LIB_PATH):
            raise MannError(f"{LIB_PATH} is missing: run `python -m manifoldann_amd.build` "
                            "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    """Map a status code to the reference's exception class."""
    if rc == 0:
        return
    msg = lib().mann_last_error().decode(errors="replace")
    if rc == -1:
        raise DimensionMismatch(msg)
    if rc == -2:
        raise ArgumentError(msg)
    if rc == -3:
        raise AssertionError(msg)
    raise MannError(f"status {rc}: {msg}")
