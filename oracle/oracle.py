"""ctypes front end of the CPU oracle (mann_oracle.c) plus NumPy FP64 helpers.

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it as the checker / CPU baseline.  The product
package ``manifoldann_amd`` never imports it.

Conventions match the reference's Julia API (ids 1-based, ``Neighbor.id``) and
the C ABI: point-major rows (Julia's d x n column-major), results k x nq
column-major i.e. a (nq, k) C-order array here.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmann_oracle.so")
_lib = None

EUCLIDEAN, SQEUCLIDEAN, COSINE = 0, 1, 2

_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_sumsq_diff.restype = ctypes.c_float
        L.orc_sumsq_diff.argtypes = [_f32p, _f32p, _i64]
        L.orc_dot.restype = ctypes.c_float
        L.orc_dot.argtypes = [_f32p, _f32p, _i64]
        L.orc_metric.restype = ctypes.c_float
        L.orc_metric.argtypes = [ctypes.c_int, _f32p, _f32p, _i64]
        L.orc_bf_knn.argtypes = [_f32p, _i64, _i64, _i64, _f32p, _i64, _i64, _i64,
                                 ctypes.c_int, _i64p, _f32p, ctypes.c_int]
        L.orc_knn_graph.argtypes = [_f32p, _i64, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int,
                                    _i64p, ctypes.c_int]
        L.orc_symmetrize.restype = _i64
        L.orc_symmetrize.argtypes = [_i64p, _i64, _i64, _i64p, _i64p]
        L.orc_norms.restype = None
        L.orc_norms.argtypes = [_f32p, _i64, _i64, _i64, _f32p]
        L.orc_kmeans_assign.argtypes = [_f32p, _i64, _i64, _i64, _f32p, _i64, _i64p, _f32p,
                                        ctypes.c_int]
        L.orc_lloyd.argtypes = [_f32p, _i64, _i64, _i64, _f32p, _i64, _i64, ctypes.c_double,
                                _i64p, _i64p, ctypes.c_int]
        L.orc_kmeanspp.argtypes = [_f32p, _i64, _i64, _i64, _i64, ctypes.c_uint64, _f32p, _i64p,
                                   ctypes.c_int]
        L.orc_kmeanspp_draws.argtypes = [_f32p, _i64, _i64, _i64, _i64,
                                         ctypes.POINTER(ctypes.c_double), _f32p, _i64p,
                                         ctypes.c_int]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_uniform.restype = ctypes.c_double
        L.orc_uniform.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_route_distances.argtypes = [_f32p, _i64, _i64, _i64, _f32p, _i64, _f32p,
                                          ctypes.c_int]
        L.orc_select_topk.restype = _i64
        L.orc_select_topk.argtypes = [_f32p, _i64, _i64, _i64, _i64p]
        L.orc_partition.restype = _i64
        L.orc_partition.argtypes = [_i64p, _i64, _i64p, _i64p]
        L.orc_simple_merge.restype = _i64
        L.orc_simple_merge.argtypes = [_i64p, _f32p, _i64, _i64, _i64p, _f32p]
        L.orc_disjoint_merge.restype = _i64
        L.orc_disjoint_merge.argtypes = [_i64p, _f32p, _i64, _i64, _i64p, _f32p]
        L.orc_ivf_search.argtypes = [_f32p, _i64, _i64, _i64p, _i64p, _i64, _f32p, _i64, _i64,
                                     _f32p, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _i64p, _f32p, ctypes.c_int]
        L.orc_ivf_flat_query.restype = ctypes.c_int
        L.orc_ivf_flat_query.argtypes = [_f32p, _i64, _i64, _i64p, _i64p, _i64, _f32p, _i64,
                                         _i64, _f32p, _i64, _i64, ctypes.c_int, _i64p, _f32p,
                                         _i64p]
        _i32p = ctypes.POINTER(ctypes.c_int32)
        L.orc_hnsw_levels.argtypes = [_i64, _i64, ctypes.c_uint64, _i32p]
        L.orc_ivf_hnsw_build.argtypes = [_f32p, _i64, _i64, _i64p, _i64p, _i64, _i64, _i64,
                                         ctypes.c_int, _i32p, _i64, _i32p, _i32p, _i32p, _i32p,
                                         ctypes.c_int]
        L.orc_ivf_hnsw_search.argtypes = [_f32p, _i64, _i64, _i64p, _i64p, _i64, _i64, _i32p,
                                          _i32p, _i32p, _i32p, _f32p, _i64, _i64, _f32p, _i64,
                                          _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, _i64p,
                                          _f32p, ctypes.c_int]
        L.orc_recall_at_k.restype = ctypes.c_double
        L.orc_recall_at_k.argtypes = [_i64p, _i64, _i64p, _i64, _i64]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _p(a):
    if a.dtype == np.float32:
        return a.ctypes.data_as(_f32p)
    return a.ctypes.data_as(_i64p)


def sumsq_diff(x, q):
    x, q = _f32(x), _f32(q)
    return np.float32(lib().orc_sumsq_diff(_p(x), _p(q), x.shape[-1]))


def dot(x, y):
    x, y = _f32(x), _f32(y)
    return np.float32(lib().orc_dot(_p(x), _p(y), x.shape[-1]))


def metric(m, x, q):
    x, q = _f32(x), _f32(q)
    return np.float32(lib().orc_metric(m, _p(x), _p(q), x.shape[-1]))


def bf_knn(base, queries, k, metric=EUCLIDEAN, nthreads=0):
    """base (n, d), queries (nq, d) -> ids (nq, k) 1-based int64, dist (nq, k) f32."""
    base, queries = _f32(base), _f32(queries)
    n, d = base.shape
    nq = queries.shape[0]
    ids = np.zeros((nq, max(k, 0)), np.int64)
    dist = np.full((nq, max(k, 0)), np.inf, np.float32)
    if nq and k > 0:
        lib().orc_bf_knn(_p(base), d, n, d, _p(queries), nq, d, k, metric, _p(ids), _p(dist),
                         nthreads)
    return ids, dist


def knn_graph(data, k, include_self=False, metric=EUCLIDEAN, nthreads=0):
    data = _f32(data)
    n, d = data.shape
    out = np.zeros((n, k), np.int64)
    rc = lib().orc_knn_graph(_p(data), d, n, d, k, int(include_self), metric, _p(out), nthreads)
    if rc != 0:
        raise ValueError("oracle knn_graph: ArgumentError")
    return out


def symmetrize(adj):
    adj = _i64a(adj)
    n, k = adj.shape
    off = np.zeros(n + 1, np.int64)
    ids = np.zeros(max(2 * n * k, 1), np.int64)
    w = lib().orc_symmetrize(_p(adj), n, k, _p(off), _p(ids))
    return [ids[off[i]:off[i + 1]].copy() for i in range(n)], int(w)


def norms(X):
    X = _f32(X)
    out = np.zeros(X.shape[0], np.float32)
    lib().orc_norms(_p(X), X.shape[1], X.shape[0], X.shape[1], _p(out))
    return out


def kmeans_assign(X, C, nthreads=0):
    """X (n, d), C (k, d) -> assign (n,) 1-based, mindist (n,) f32."""
    X, C = _f32(X), _f32(C)
    n, d = X.shape
    a = np.zeros(n, np.int64)
    md = np.zeros(n, np.float32)
    lib().orc_kmeans_assign(_p(X), d, n, d, _p(C), C.shape[0], _p(a), _p(md), nthreads)
    return a, md


def lloyd(X, C_init, max_iters, tol=1e-6, nthreads=0):
    X = _f32(X)
    C = _f32(C_init).copy()
    n, d = X.shape
    a = np.zeros(n, np.int64)
    it = np.zeros(1, np.int64)
    rc = lib().orc_lloyd(_p(X), d, n, d, _p(C), C.shape[0], max_iters, tol, _p(a), _p(it),
                         nthreads)
    if rc != 0:
        raise AssertionError("Cannot fit more clusters than points")
    return C, a, int(it[0])


def kmeanspp(X, k, seed, nthreads=0):
    X = _f32(X)
    n, d = X.shape
    C = np.zeros((k, d), np.float32)
    idx = np.zeros(k, np.int64)
    rc = lib().orc_kmeanspp(_p(X), d, n, d, k, seed, _p(C), _p(idx), nthreads)
    if rc != 0:
        raise AssertionError("Cannot initialize more clusters than points")
    return C, idx


def kmeanspp_draws(X, k, draws, nthreads=0):
    """kmeans++ with explicit uniform draws u_0..u_{k-1} (u_0 picks the first
    centre, u_i the threshold fraction of round i) -> (C (k, d), idx (k,) 1-based)."""
    X = _f32(X)
    n, d = X.shape
    u = np.ascontiguousarray(draws, np.float64)
    assert u.shape[0] >= k
    C = np.zeros((k, d), np.float32)
    idx = np.zeros(k, np.int64)
    rc = lib().orc_kmeanspp_draws(_p(X), d, n, d, k,
                                  u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _p(C),
                                  _p(idx), nthreads)
    if rc != 0:
        raise AssertionError("Cannot initialize more clusters than points")
    return C, idx


def splitmix64_stream(seed, count):
    st = ctypes.c_uint64(seed)
    return [int(lib().orc_splitmix64(ctypes.byref(st))) for _ in range(count)]


def uniform_stream(seed, count):
    st = ctypes.c_uint64(seed)
    return np.array([lib().orc_uniform(ctypes.byref(st)) for _ in range(count)])


def route_distances(Q, C, nthreads=0):
    Q, C = _f32(Q), _f32(C)
    nq, d = Q.shape
    out = np.zeros((nq, C.shape[0]), np.float32)
    lib().orc_route_distances(_p(Q), d, nq, d, _p(C), C.shape[0], _p(out), nthreads)
    return out


def select_topk(dist, n_domain, nprobe):
    dist = _f32(dist)
    out = np.zeros(max(nprobe, 1), np.int64)
    kk = lib().orc_select_topk(_p(dist), dist.shape[0], n_domain, nprobe, _p(out))
    if kk < 0:
        raise ValueError("Transform produced no routing distances for TopKRouting")
    return out[:kk].copy()


def partition(assign):
    assign = _i64a(assign)
    n = assign.shape[0]
    nb_max = int(assign.max()) if n else 0
    off = np.zeros(max(nb_max, 0) + 1, np.int64)
    ids = np.zeros(max(n, 1), np.int64)
    nb = lib().orc_partition(_p(assign), n, _p(off), _p(ids))
    if nb < 0:
        raise ValueError("Transform produced invalid bucket assignments")
    return off, ids[:n].copy(), int(nb)


def simple_merge(ids, dist, k):
    ids, dist = _i64a(ids), _f32(dist)
    oi = np.zeros(max(k, 1), np.int64)
    od = np.zeros(max(k, 1), np.float32)
    m = lib().orc_simple_merge(_p(ids), _p(dist), ids.shape[0], k, _p(oi), _p(od))
    return oi[:m].copy(), od[:m].copy()


def disjoint_merge(ids, dist, k):
    ids, dist = _i64a(ids), _f32(dist)
    oi = np.zeros(max(k, 1), np.int64)
    od = np.zeros(max(k, 1), np.float32)
    m = lib().orc_disjoint_merge(_p(ids), _p(dist), ids.shape[0], k, _p(oi), _p(od))
    return oi[:m].copy(), od[:m].copy()


def ivf_search(X, offsets, ids, nb, Q, route, k, nprobe, metric=EUCLIDEAN, merge="simple",
               probe_order=False, nthreads=0):
    X, Q, route = _f32(X), _f32(Q), _f32(route)
    offsets, ids = _i64a(offsets), _i64a(ids)
    n, d = X.shape
    nq = Q.shape[0]
    oi = np.zeros((nq, k), np.int64)
    od = np.full((nq, k), np.inf, np.float32)
    rc = lib().orc_ivf_search(_p(X), d, d, _p(offsets), _p(ids), nb, _p(Q), nq, d, _p(route),
                              route.shape[1], k, nprobe, metric,
                              0 if merge == "simple" else 1, int(probe_order), _p(oi), _p(od),
                              nthreads)
    if rc != 0:
        raise ValueError("Transform produced no routing distances for TopKRouting")
    return oi, od


def ivf_flat_query(X, offsets, ids, Q, route, k, nprobe, metric=EUCLIDEAN):
    """query(::IVFFlatIndex) per query: bounded-heap merge in probe order.
    offsets has nlist + 1 entries (empty lists included).
    Returns (ids (nq, k) padded with 0, dist (nq, k) padded with inf, k_eff (nq,))."""
    X, Q, route = _f32(X), _f32(Q), _f32(route)
    offsets, ids = _i64a(offsets), _i64a(ids)
    n, d = X.shape
    nq = Q.shape[0]
    nlist = offsets.shape[0] - 1
    oi = np.zeros((nq, max(k, 0)), np.int64)
    od = np.full((nq, max(k, 0)), np.inf, np.float32)
    cnt = np.zeros(max(nq, 1), np.int64)
    rc = lib().orc_ivf_flat_query(_p(X), d, d, _p(offsets), _p(ids), nlist, _p(Q), nq, d,
                                  _p(route), max(k, 0), nprobe, metric, _p(oi), _p(od), _p(cnt))
    if rc != 0:
        raise ValueError("nlist must be positive")
    return oi, od, cnt[:nq].copy()


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


HNSW_HEURISTIC, HNSW_DIVERSIFIED = 0, 1


def hnsw_levels(n, M, seed):
    """HNSW levels of n points from one splitmix64 stream (planner.jl:19-30)."""
    out = np.zeros(max(n, 1), np.int32)
    rc = lib().orc_hnsw_levels(n, M, seed, _p32(out))
    if rc != 0:
        raise ValueError("M must be positive")
    return out[:n].copy()


class HnswGraphs:
    """The per-list HNSW graphs of an IVF-HNSW index, list-contiguous rows:
    nbr (layers, n, M + 1) local 0-based ids, deg (layers, n), entry and
    max_layer per bucket (-1 when empty), level per row."""

    def __init__(self, nbr, deg, entry, max_layer, level, M):
        self.nbr, self.deg, self.entry, self.max_layer = nbr, deg, entry, max_layer
        self.level, self.M = level, M


def ivf_hnsw_build(X, offsets, ids, nb, M, ef_construction, policy=HNSW_DIVERSIFIED, seed=0,
                   level=None, nthreads=0):
    """Serial HNSW build of every non-empty list (the terminals of
    build_ivf_hnsw_index).  level: per list-contiguous row (default: drawn from
    the seed's stream)."""
    X = _f32(X)
    offsets, ids = _i64a(offsets), _i64a(ids)
    n, d = X.shape
    n_all = int(offsets[nb])
    lev = hnsw_levels(n_all, M, seed) if level is None else _i32(level)
    layers = int(lev.max()) + 1 if n_all else 1
    nbr = np.zeros((layers, max(n_all, 1), M + 1), np.int32)
    deg = np.zeros((layers, max(n_all, 1)), np.int32)
    entry = np.zeros(max(nb, 1), np.int32)
    maxl = np.zeros(max(nb, 1), np.int32)
    rc = lib().orc_ivf_hnsw_build(_p(X), d, d, _p(offsets), _p(ids), nb, M, ef_construction,
                                  policy, _p32(lev), layers, _p32(nbr), _p32(deg), _p32(entry),
                                  _p32(maxl), nthreads)
    if rc != 0:
        raise ValueError(f"HNSW build failed ({rc})")
    return HnswGraphs(nbr[:, :n_all], deg[:, :n_all], entry[:nb], maxl[:nb], lev, M)


def ivf_hnsw_search(X, offsets, ids, nb, graphs, Q, route, k, nprobe, ef_search,
                    merge="disjoint", probe_order=False, nthreads=0):
    """query(::MultiLevelIndex{KMeans, TopK, HNSW}): ids (nq, k) 1-based padded
    with 0, dist (nq, k) padded with inf."""
    X, Q, route = _f32(X), _f32(Q), _f32(route)
    offsets, ids = _i64a(offsets), _i64a(ids)
    n, d = X.shape
    nq = Q.shape[0]
    g = graphs
    nbr = np.ascontiguousarray(g.nbr, np.int32)
    deg = np.ascontiguousarray(g.deg, np.int32)
    oi = np.zeros((nq, max(k, 0)), np.int64)
    od = np.full((nq, max(k, 0)), np.inf, np.float32)
    if nq == 0 or k <= 0:
        return oi, od
    rc = lib().orc_ivf_hnsw_search(_p(X), d, d, _p(offsets), _p(ids), nb, g.M, _p32(nbr),
                                   _p32(deg), _p32(_i32(g.entry)), _p32(_i32(g.max_layer)),
                                   _p(Q), nq, d, _p(route), route.shape[1], k, nprobe,
                                   ef_search, 0 if merge == "simple" else 1, int(probe_order),
                                   _p(oi), _p(od), nthreads)
    if rc != 0:
        raise ValueError(f"IVF-HNSW search failed ({rc})")
    return oi, od


def recall_at_k(found, truth):
    found, truth = _i64a(found), _i64a(truth)
    return float(lib().orc_recall_at_k(_p(found), found.shape[1], _p(truth), truth.shape[1],
                                       found.shape[0]))


# ---------------------------------------------------------------------------
# NumPy FP64 helpers: an independent exact-arithmetic check of the C oracle.
# ---------------------------------------------------------------------------

def fp64_knn(base, queries, k, metric=EUCLIDEAN):
    """Exact (FP64) brute force, ties to the lower index; ids 1-based."""
    B = np.asarray(base, np.float64)
    Qm = np.asarray(queries, np.float64)
    if metric == COSINE:
        nb = np.linalg.norm(B, axis=1)
        nq = np.linalg.norm(Qm, axis=1)
        D = 1.0 - (Qm @ B.T) / (nq[:, None] * nb[None, :])
        D = np.maximum(D, 0.0)
    else:
        D = ((Qm[:, None, :] - B[None, :, :]) ** 2).sum(-1) if B.shape[0] * Qm.shape[0] < 4e6 \
            else np.maximum((Qm ** 2).sum(1)[:, None] + (B ** 2).sum(1)[None, :] - 2 * Qm @ B.T, 0)
        if metric == EUCLIDEAN:
            D = np.sqrt(D)
    order = np.lexsort((np.broadcast_to(np.arange(B.shape[0]), D.shape), D), axis=1)[:, :k]
    return order + 1, np.take_along_axis(D, order, 1)


def fp32_rounding_gap_ambiguous(D64_sorted, k, d):
    """Count queries whose FP64 gap between the k-th and (k+1)-th neighbour is
    below the FP32 rounding bound of a d-term sum (ties FP32 may order either way)."""
    u = 2.0 ** -24
    gap = D64_sorted[:, k] - D64_sorted[:, k - 1]
    bound = (d + 4) * u * D64_sorted[:, k]
    return int((gap <= bound).sum())
