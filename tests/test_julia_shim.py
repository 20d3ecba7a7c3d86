"""The Julia binding (julia/ManifoldANNHIP.jl) against the C ABI.

Julia is not installed here, so the shim cannot run; this test keeps it from
drifting away from the library: every `ccall` in it names a function that
include/mann.h declares and libmann_hip.so exports, with the same number of
arguments and the same argument kinds (pointer, int64, int, uint64, double),
and the ctypes table of the Python mirror agrees with the header for every
declared function.  Each ccall signature is then exercised through ctypes with
null handles: the library answers with a status code (no call reaches a GPU
before its null checks)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "julia", "ManifoldANNHIP.jl")
HEADER = os.path.join(ROOT, "include", "mann.h")

JULIA_KIND = {"Int64": "i64", "Cint": "int", "UInt64": "u64", "Float64": "f64", "Cdouble": "f64",
              "Cstring": "ptr"}


def _julia_kind(t):
    t = t.strip()
    if t.startswith("Ptr{") or t.startswith("Ref{"):
        return "ptr"
    return JULIA_KIND[t]


def _c_kind(t):
    t = t.strip()
    if "*" in t or t.endswith("_t") and t.startswith("mann_"):
        return "ptr"       # pointers and the opaque handle typedefs
    t = t.replace("const ", "").split()[0]
    return {"int64_t": "i64", "int": "int", "uint64_t": "u64", "double": "f64"}[t]


def _split_top(s):
    """Split s at commas outside braces/parentheses."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "{(":
            depth += 1
        elif ch in "})":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


def header_decls():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"(int|const char\*)\s+(mann_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, params = m.groups()
        params = params.strip()
        kinds = [] if params in ("", "void") else [
            _c_kind(re.sub(r"\s*\w+$", "", p.strip()) if not p.strip().endswith("*")
                    else p.strip()) for p in params.split(",")]
        decls[name] = ("ptr" if "char" in ret else "int", kinds)
    return decls


def shim_ccalls():
    text = open(SHIM).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(mann_\w+),\s*LIB\),\s*(\w+),\s*\(", text):
        name, ret = m.group(1), m.group(2)
        # the argument-type tuple: from the '(' after the return type to its match
        i = m.end() - 1
        depth, j = 0, i
        while True:
            if text[j] == "(":
                depth += 1
            elif text[j] == ")":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        types = _split_top(text[i + 1:j])
        calls.append((name, "ptr" if ret == "Cstring" else "int", [_julia_kind(t) for t in types]))
    return calls


def test_shim_ccalls_match_the_header():
    decls = header_decls()
    calls = shim_ccalls()
    names = {c[0] for c in calls}
    # the methods the shim overrides reach these entry points
    for must in ("mann_bf_knn", "mann_knn_graph", "mann_knn_graph_symmetrize", "mann_kmeans_fit",
                 "mann_knn_candidates", "mann_ivf_build", "mann_ivf_flat_build",
                 "mann_ivf_search_ex", "mann_dataset_create", "mann_dataset_destroy"):
        assert must in names, must
    for name, ret, kinds in calls:
        assert name in decls, f"{name} is not declared in include/mann.h"
        hret, hkinds = decls[name]
        assert ret == hret, name
        assert kinds == hkinds, (name, kinds, hkinds)


def test_python_ctypes_table_matches_the_header():
    from manifoldann_amd._lib import SIGNATURES
    decls = header_decls()
    table = {name: args for name, _, args in SIGNATURES}
    assert set(table) == set(decls), set(table) ^ set(decls)
    for name, (_, kinds) in decls.items():
        assert len(table[name]) == len(kinds), name


def test_library_exports_every_declared_function():
    from manifoldann_amd._lib import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(LIB_PATH)
    for name in header_decls():
        assert hasattr(L, name), name


CT = {"ptr": ctypes.c_void_p, "i64": ctypes.c_int64, "int": ctypes.c_int, "u64": ctypes.c_uint64,
      "f64": ctypes.c_double}


def test_every_shim_signature_is_callable_with_null_handles():
    from manifoldann_amd._lib import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(LIB_PATH)
    seen = set()
    for name, ret, kinds in shim_ccalls():
        if name in seen:
            continue
        seen.add(name)
        f = getattr(L, name)
        f.restype = ctypes.c_char_p if ret == "ptr" else ctypes.c_int
        f.argtypes = [CT[k] for k in kinds]
        rc = f(*[CT[k](0) for k in kinds])
        if ret == "ptr":
            assert rc is not None, name
        else:
            # 0 (a null handle is a no-op to destroy) or a status: -6 null,
            # -2 argument, -1 dimension, -3 assertion, -4 no device here
            assert rc in (0, -1, -2, -3, -4, -6), (name, rc)
