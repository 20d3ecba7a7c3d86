"""Host-side mirror of ManifoldANN.jl's exact-search API over the HIP library.

Shapes follow Julia: a data matrix is d x n with one point per column (numpy
arrays are taken column-major; a torch tensor must have unit stride along d).
Ids are 1-based like ``Neighbor.id``.  Results of a batch query are
``list[list[Neighbor]]``; the array form (``query_arrays``) returns (nq, k)
id/distance matrices (0 / +inf in empty slots).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any, NamedTuple, Optional

import numpy as np

from ._lib import ArgumentError, DimensionMismatch, check, lib, shutting_down

EUCLIDEAN, SQEUCLIDEAN, COSINE = 0, 1, 2


class Euclidean:
    code = EUCLIDEAN

    def __eq__(self, o):
        return type(o) is type(self)

    def __hash__(self):
        return hash(type(self))


class SqEuclidean(Euclidean):
    code = SQEUCLIDEAN


class CosineDist(Euclidean):
    code = COSINE


# utils/distances.jl:15-18
default_distance = Euclidean()
default_squared_distance = SqEuclidean()
euclidean_distance = Euclidean()
cosine_distance = CosineDist()


class Neighbor(NamedTuple):
    """utils/neighbor_heaps.jl:11-14 (id is 1-based)."""
    id: int
    dist: float


def neighbor_ids(neighbors):
    """utils/neighbor_heaps.jl:22-44."""
    if neighbors and isinstance(neighbors[0], list):
        return [[n.id for n in row] for row in neighbors]
    return [n.id for n in neighbors]


# ---------------------------------------------------------------------------
# array plumbing
# ---------------------------------------------------------------------------

def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ready(t):
    """The library reads device memory on its own stream: the torch work that
    produced `t` (on torch's current stream) must be complete first."""
    if t.is_cuda:
        import torch
        torch.cuda.current_stream(t.device).synchronize()
    return t


@dataclass
class _Mat:
    ptr: int
    d: int
    n: int
    ld: int
    keep: Any
    device: bool


def as_matrix(x, d_expected: Optional[int] = None) -> _Mat:
    """View a Julia-shaped d x n Float32 matrix (or a length-d vector)."""
    if _is_torch(x):
        import torch
        t = x
        if t.dim() == 1:
            t = t.reshape(-1, 1)
        if t.dtype != torch.float32:
            t = t.to(torch.float32)
        d, n = t.shape
        if n > 1 and t.stride(0) != 1:
            t = t.t().contiguous().t()
        ld = t.stride(1) if n > 1 else d
        return _Mat(_ready(t).data_ptr(), d, n, ld, t, t.is_cuda)
    a = np.asarray(x, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.ndim != 2:
        raise DimensionMismatch("expected a vector or a d x n matrix")
    a = np.asfortranarray(a)
    d, n = a.shape
    return _Mat(a.ctypes.data, d, n, d, a, False)


def _is_f16(x) -> bool:
    if _is_torch(x):
        import torch
        return x.dtype == torch.float16
    return isinstance(x, np.ndarray) and x.dtype == np.float16


def _as_matrix_f16(x) -> _Mat:
    """View a d x n Float16 matrix (numpy or torch) as rows of d halves."""
    if _is_torch(x):
        t = x if x.dim() == 2 else x.reshape(-1, 1)
        d, n = t.shape
        if n > 1 and t.stride(0) != 1:
            t = t.t().contiguous().t()
        ld = t.stride(1) if n > 1 else d
        return _Mat(_ready(t).data_ptr(), d, n, ld, t, t.is_cuda)
    a = np.asfortranarray(x if x.ndim == 2 else x.reshape(-1, 1))
    d, n = a.shape
    return _Mat(a.ctypes.data, d, n, d, a, False)


class Dataset:
    """Device-resident copy of a data matrix (explicit, caller-owned).

    The reference passes ``data`` to every call and indices never store it
    (ann_index.jl:1-10); on the GPU the prepared copy (FP32 + bf16 + norms)
    is this handle, created once and reused across calls.
    """

    FIRST_PASS = {"bf16": 0, "fp8": 1}

    def __init__(self, data, metric=default_distance, first_pass="bf16"):
        h = ctypes.c_void_p()
        self.metric = metric
        if _is_f16(data):  # Matrix{Float16}: upcast exactly on the device
            m = _as_matrix_f16(data)
            check(lib().mann_dataset_create_f16(m.ptr, m.d, m.n, m.ld, metric.code,
                                                ctypes.byref(h)))
        else:
            m = as_matrix(data)
            check(lib().mann_dataset_create(m.ptr, m.d, m.n, m.ld, metric.code, ctypes.byref(h)))
        self.d, self.n = m.d, m.n
        self.handle = h
        self.first_pass = "bf16"
        if first_pass != "bf16":
            self.set_first_pass(first_pass)

    def set_first_pass(self, kind):
        """Arithmetic of the exact searches' first pass: "bf16" or "fp8" (OCP
        E4M3 copy of the rows, one byte per element).  Results are identical
        either way; only the speed differs."""
        if kind not in self.FIRST_PASS:
            raise ArgumentError(f"first_pass must be one of {sorted(self.FIRST_PASS)}")
        check(lib().mann_dataset_set_first_pass(self.handle, self.FIRST_PASS[kind]))
        self.first_pass = kind

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        try:
            h = getattr(self, "handle", None)
            if h is not None and h.value and not shutting_down():
                lib().mann_dataset_destroy(h)
                self.handle = None
        except Exception:
            pass


def _dataset_for(data, metric) -> Dataset:
    if isinstance(data, Dataset):
        if type(data.metric) is not type(metric):
            raise ArgumentError("dataset was prepared for another metric")
        return data
    return Dataset(data, metric)


def _data_shape(data):
    if isinstance(data, Dataset):
        return data.d, data.n
    if _is_torch(data) or isinstance(data, np.ndarray):
        s = data.shape
        return (s[0], s[1]) if len(s) == 2 else (s[0], 1)
    a = np.asarray(data)
    return a.shape[0], (a.shape[1] if a.ndim == 2 else 1)


# ---------------------------------------------------------------------------
# BruteForceIndex (indices/bruteforce.jl)
# ---------------------------------------------------------------------------

@dataclass
class BruteForceIndex:
    dimension: int
    n_points: int
    distance: Any = field(default_factory=Euclidean)


def build_index(index_type, data, distance=default_distance, **kw):
    """build_index(BruteForceIndex, data; distance) (bruteforce.jl:31-35)."""
    if index_type is BruteForceIndex:
        d, n = _data_shape(data)
        if d <= 0:
            raise ArgumentError("Dataset must have at least one dimension")
        if n <= 0:
            raise ArgumentError("Dataset must contain at least one point")
        return BruteForceIndex(d, n, distance)
    from .ivf import build_index as ivf_build_index
    return ivf_build_index(index_type, data, distance=distance, **kw)


def insert(index: BruteForceIndex, points):
    """insert!(index, point|points) (bruteforce.jl:82-95)."""
    a = np.asarray(points)
    if a.ndim == 1:
        if a.shape[0] != index.dimension:
            raise DimensionMismatch(f"Expected point dimension {index.dimension}")
        index.n_points += 1
    else:
        if a.shape[0] != index.dimension:
            raise DimensionMismatch(f"Expected point dimension {index.dimension}")
        index.n_points += a.shape[1]
    return index


def _validate(index, data, d_q):
    """validate_index_dimensions (utils/validation.jl:8-19)."""
    d, n = _data_shape(data)
    if d != index.dimension:
        raise DimensionMismatch(f"Expected data with {index.dimension} rows")
    if d_q != index.dimension:
        raise DimensionMismatch(f"Expected query length {index.dimension}")
    if n < index.n_points:
        raise ArgumentError(f"Data contains {n} points but index tracks {index.n_points}")


def query_arrays(index, data, Q, k, out=None, nprobe=None, ef_search=None):
    """Batch exact kNN; returns (ids (nq, k) int64 1-based, dist (nq, k) f32).

    Device (torch) queries give device results.  Equivalent to the generic
    batch query (ann_index.jl:71-101) over bruteforce.jl:46-71.
    """
    if hasattr(index, "root"):
        from .ivf import query_arrays as ivf_query_arrays
        return ivf_query_arrays(index, data, Q, k, nprobe=nprobe, ef_search=ef_search)
    if ef_search is not None:
        raise ArgumentError("ef_search is a keyword of HNSW queries")
    if nprobe is not None:
        raise ArgumentError("nprobe is a keyword of IVF index queries")
    qm = as_matrix(Q)
    d, n = _data_shape(data)
    if qm.d != d:
        raise DimensionMismatch(f"Expected queries with {d} rows, got {qm.d}")
    nq = qm.n
    if qm.n > 0:
        _validate(index, data, qm.d)
    k_out = max(int(k), 0)
    if qm.device:
        import torch
        ids = torch.zeros((nq, k_out), dtype=torch.int64, device=qm.keep.device)
        dist = torch.full((nq, k_out), float("inf"), dtype=torch.float32, device=qm.keep.device)
        pi, pd = _ready(ids).data_ptr(), _ready(dist).data_ptr()
    else:
        ids = np.zeros((nq, k_out), np.int64)
        dist = np.full((nq, k_out), np.inf, np.float32)
        pi, pd = ids.ctypes.data, dist.ctypes.data
    if nq == 0 or k_out == 0:
        return ids, dist
    ds = _dataset_for(data, index.distance)
    check(lib().mann_bf_knn(ds.handle, qm.ptr, qm.d, nq, qm.ld, k_out, index.distance.code, pi, pd))
    return ids, dist


def _to_neighbors(ids, dist):
    ids = ids.cpu().numpy() if _is_torch(ids) else ids
    dist = dist.cpu().numpy() if _is_torch(dist) else dist
    out = []
    for r in range(ids.shape[0]):
        row = [Neighbor(int(i), float(v)) for i, v in zip(ids[r], dist[r]) if i != 0]
        out.append(row)
    return out


def query(index, data, q, k, nprobe=None, ef_search=None):
    """query(index, data, q|Q, k[; nprobe]) -> Vector{Neighbor} or
    Vector{Vector{Neighbor}}; nprobe (IVF indexes only) as in ivf_flat.jl:138-144."""
    if isinstance(q, (list, tuple)) and q and not np.isscalar(q[0]):
        # vector-of-vectors batch (ann_index.jl:110-120)
        q = np.stack([np.asarray(v, np.float32) for v in q], axis=1)
    single = (np.ndim(q) == 1) if not _is_torch(q) else q.dim() == 1
    if hasattr(index, "root"):
        # IVF: one vector takes the single-query path, whose per-list results
        # merge in probe order (query.jl:25-106); a matrix the batched one
        from .ivf import query_arrays as ivf_query_arrays
        ids, dist = ivf_query_arrays(index, data, q, k, nprobe=nprobe, probe_order=single,
                                     ef_search=ef_search)
        res = _to_neighbors(ids, dist)
        return (res[0] if k > 0 else []) if single else res
    if single:
        d_q = q.shape[0] if _is_torch(q) else np.asarray(q).shape[0]
        _validate(index, data, d_q) if not hasattr(index, "root") else None
        ids, dist = query_arrays(index, data, q, k, nprobe=nprobe)
        return _to_neighbors(ids, dist)[0] if k > 0 else []
    ids, dist = query_arrays(index, data, q, k, nprobe=nprobe)
    return _to_neighbors(ids, dist)


def last_stats():
    """[reranked, overflowed, second-pass, exact-fallback] of the last call."""
    out = (ctypes.c_int64 * 4)()
    check(lib().mann_last_stats(out))
    return list(out)


def autotune(index, data, Q=None, k=None):
    """Time the exact search's knobs on this batch and keep the best for its
    shape (device, dataset size, metric, first pass, k, query count rounded
    up to a power of two); Q=None tunes the kNN graph self-join.  Later
    searches of that shape use the choice.  No reference counterpart."""
    ds = _dataset_for(data, index.distance)
    if k is None or int(k) <= 0:
        raise ArgumentError("k must be positive")
    choice = (ctypes.c_int64 * 6)()
    ms = ctypes.c_double()
    if Q is None:
        check(lib().mann_autotune_bf_knn(ds.handle, None, ds.d, ds.n, ds.d, int(k),
                                         index.distance.code, choice, ctypes.byref(ms)))
    else:
        qm = as_matrix(Q)
        check(lib().mann_autotune_bf_knn(ds.handle, qm.ptr, qm.d, qm.n, qm.ld, int(k),
                                         index.distance.code, choice, ctypes.byref(ms)))
    keys = ("scan_cfg", "emit_per_k", "shortlist_mult", "scan8_cfg", "fp8_emit_per_k", "trials")
    out = dict(zip(keys, list(choice)))
    out["ms"] = ms.value
    return out


class SearchGraph:
    """An exact batch query captured in a HIP graph: one shape, fixed device
    buffers.  Q is a d x nq torch CUDA matrix whose columns are contiguous
    (e.g. rows.t() of an nq x d tensor); refill it in place between launches.
    launch() replays the search and returns (ids, dist) tensors, which the
    next launch overwrites.  No reference counterpart."""

    def __init__(self, index, data, Q, k):
        import torch
        if not _is_torch(Q) or not Q.is_cuda or Q.dtype != torch.float32 or Q.dim() != 2:
            raise ArgumentError("SearchGraph takes a d x nq float32 CUDA tensor")
        if Q.shape[1] > 1 and Q.stride(0) != 1:
            raise ArgumentError("SearchGraph needs columns contiguous in memory (rows.t())")
        self.ds = _dataset_for(data, index.distance)
        d, nq = Q.shape
        self.Q = Q
        self.ids = torch.zeros((nq, int(k)), dtype=torch.int64, device=Q.device)
        self.dist = torch.zeros((nq, int(k)), dtype=torch.float32, device=Q.device)
        _ready(self.ids)
        _ready(self.dist)
        _ready(Q)
        h = ctypes.c_void_p()
        ld = Q.stride(1) if nq > 1 else d
        check(lib().mann_bf_knn_graph_create(self.ds.handle, Q.data_ptr(), d, nq, ld, int(k),
                                             index.distance.code, self.ids.data_ptr(),
                                             self.dist.data_ptr(), ctypes.byref(h)))
        self.handle = h

    def launch(self):
        _ready(self.Q)
        check(lib().mann_graph_launch(self.handle))
        return self.ids, self.dist

    def __del__(self):
        try:
            h = getattr(self, "handle", None)
            if h is not None and h.value and not shutting_down():
                lib().mann_graph_destroy(h)
                self.handle = None
        except Exception:
            pass


def last_emitted():
    """First-pass candidates the selection read in the last search (all passes)."""
    out = ctypes.c_int64()
    check(lib().mann_last_emitted(ctypes.byref(out)))
    return out.value


# ---------------------------------------------------------------------------
# kNN graph (graphs/knn_graph.jl)
# ---------------------------------------------------------------------------

@dataclass
class KNNGraph:
    neighbors: list
    k: int
    include_self: bool
    metadata: Any

    def __len__(self):
        return len(self.neighbors)

    def __getitem__(self, i):
        return self.neighbors[i]

    def __iter__(self):
        return iter(self.neighbors)


def configured_k(index):
    return getattr(index, "k", None)


def _resolve_graph_k(index, k):
    """knn_graph.jl:83-103."""
    configured = configured_k(index)
    if k is None:
        if configured is None:
            raise ArgumentError(f"Index {type(index).__name__} does not expose a default k; pass `k=` explicitly")
        return configured
    if k <= 0:
        raise ArgumentError("k must be positive")
    if configured is not None and k > configured:
        raise ArgumentError(f"Requested k={k} exceeds index configuration k={configured}")
    return int(k)


def build_knn_graph(index, data, k=None, include_self=False, directed=True, metadata=None,
                    return_arrays=False):
    """build_knn_graph (knn_graph.jl:42-81) on the GPU self-join."""
    d, n = _data_shape(data)
    if n <= 0:
        raise ArgumentError("Dataset must contain at least one point")
    rk = _resolve_graph_k(index, k)
    maxnb = n - (0 if include_self else 1)
    if not (0 < rk <= maxnb):
        raise ArgumentError(f"k must satisfy 0 < k <= {maxnb} for dataset with {n} points")
    if metadata is not None and len(metadata) != n:
        raise ArgumentError(f"Metadata length {len(metadata)} must equal number of points {n}")
    ds = _dataset_for(data, index.distance)
    adj = np.zeros((n, rk), np.int64)
    dist = np.zeros((n, rk), np.float32)
    check(lib().mann_knn_graph(ds.handle, rk, int(include_self), index.distance.code,
                               adj.ctypes.data, dist.ctypes.data))
    original_k = rk
    if return_arrays and directed:
        return adj, dist
    if directed:
        neighbors = [row.tolist() for row in adj]
    else:
        neighbors, rk = symmetrize_adjacency(adj)
    meta = {"original_k": original_k, "directed": bool(directed),
            "node_metadata": list(metadata) if metadata is not None else None}
    return KNNGraph(neighbors, rk, bool(include_self), meta)


def knn_candidates(index, data, candidate_k, return_arrays=False):
    """The candidate lists of build_weighted_graph with candidate_k > k
    (weighted_knn_graph.jl:497-503): for every point, query candidate_k + 1,
    drop the point itself by id, keep the first candidate_k; short lists are
    returned as they are.  -> list of id lists (1-based), or with
    return_arrays=True (ids (n, candidate_k) 0-padded, dist, counts)."""
    d, n = _data_shape(data)
    if n <= 0:
        raise ArgumentError("Dataset must contain at least one point")
    ds = _dataset_for(data, index.distance)
    ids = np.zeros((n, candidate_k), np.int64)
    dist = np.zeros((n, candidate_k), np.float32)
    cnt = np.zeros(n, np.int64)
    check(lib().mann_knn_candidates(ds.handle, int(candidate_k), index.distance.code,
                                    ids.ctypes.data, dist.ctypes.data, cnt.ctypes.data))
    if return_arrays:
        return ids, dist, cnt
    return [ids[i, :cnt[i]].tolist() for i in range(n)]


def symmetrize_adjacency(adj):
    """_symmetrize_adjacency (knn_graph.jl:196-210) -> (lists, max degree)."""
    adj = np.ascontiguousarray(adj, np.int64)
    n, k = adj.shape
    off = np.zeros(n + 1, np.int64)
    ids = np.zeros(max(2 * n * k, 1), np.int64)
    tot = ctypes.c_int64()
    mx = ctypes.c_int64()
    check(lib().mann_knn_graph_symmetrize(adj.ctypes.data, n, k, off.ctypes.data, ids.ctypes.data,
                                          ctypes.byref(tot), ctypes.byref(mx)))
    return [ids[off[i]:off[i + 1]].tolist() for i in range(n)], int(mx.value)


def has_metadata(graph):
    return graph.metadata is not None and graph.metadata["node_metadata"] is not None


def node_metadata(graph, i):
    if graph.metadata is None:
        raise ArgumentError("Graph does not store metadata")
    if graph.metadata["node_metadata"] is None:
        raise ArgumentError("Graph does not store node-level metadata")
    return graph.metadata["node_metadata"][i - 1]


def graph_metadata(graph):
    return graph.metadata


# ---------------------------------------------------------------------------
# recall (utils/eval_utils.jl:7-35)
# ---------------------------------------------------------------------------

def recall_at_k(found, truth):
    if len(found) and np.ndim(found[0]) == 0:
        k = len(truth)
        if k == 0:
            return 0.0
        gt = set(int(t) for t in truth)
        return sum(1 for f in found if int(f) in gt) / k
    if len(found) != len(truth):
        raise ArgumentError("found and ground truth batches must match in length")
    if len(found) == 0:
        return 0.0
    return float(np.mean([recall_at_k(f, t) for f, t in zip(found, truth)]))
