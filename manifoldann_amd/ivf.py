"""Host-side mirror of ManifoldANN.jl's k-means transform and IVF
(MultiLevelIndex{KMeansTransform, TopKRouting, BruteForceIndex}) over the HIP
library.

Reference files: transforms/KMeansTransform.jl, transforms/kmeans/*.jl,
transforms/utils.jl, indices/multilevel/{config,builder,routing,query,merge}.jl.
Randomness (kmeans++ draws, :random init, subsampling) comes from the
library's splitmix64 stream of an integer seed: Julia RNG streams cannot be
reproduced outside Julia.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from ._lib import ArgumentError, DimensionMismatch, check, lib, shutting_down
from .index import (BruteForceIndex, Dataset, Euclidean, Neighbor, _data_shape, _dataset_for,
                    _is_torch, _ready, _to_neighbors, as_matrix, default_distance)

KMEANS_DEFAULT_MAX_ITERATIONS = 100   # utils/constants.jl:127
KMEANS_DEFAULT_TOLERANCE = 1e-6       # utils/constants.jl:135
IVF_DEFAULT_NPROBE = 5                # utils/constants.jl:154


def _seed_of(rng) -> int:
    if rng is None:
        return 0x5EED
    if isinstance(rng, (int, np.integer)):
        return int(rng) & (2**64 - 1)
    if isinstance(rng, np.random.Generator):
        return int(rng.integers(0, 2**63))
    raise ArgumentError("rng must be an integer seed or a numpy Generator")


# ---------------------------------------------------------------------------
# transforms (transforms/Transform.jl, KMeansTransform.jl)
# ---------------------------------------------------------------------------

@dataclass
class KMeansAssignment:
    """KMeansTransform.jl:85-87: distances to every centroid."""
    distances: np.ndarray


@dataclass
class TransformResult:
    """Transform.jl:36-39."""
    data: Any
    assignment: Any


class KMeansTransform:
    """KMeansTransform.jl:41-79 (assertions as in the reference constructor)."""

    def __init__(self, k: int, distance=default_distance, init: str = "kmeans_plus_plus",
                 max_iters: int = KMEANS_DEFAULT_MAX_ITERATIONS,
                 tol: float = KMEANS_DEFAULT_TOLERANCE, subsample_size: Optional[int] = None):
        assert k > 0, "Number of clusters must be positive"
        assert init in ("random", "kmeans_plus_plus"), "init must be :random or :kmeans_plus_plus"
        assert max_iters > 0, "max_iters must be positive"
        assert tol > 0, "tol must be positive"
        if subsample_size is not None and subsample_size <= 0:
            raise ArgumentError("subsample_size must be positive or nothing")
        if not isinstance(distance, Euclidean) or type(distance) is not Euclidean:
            raise ArgumentError("the GPU k-means supports the Euclidean distance")
        self.k = k
        self.distance = distance
        self.init = init
        self.max_iters = max_iters
        self.tol = tol
        self.subsample_size = subsample_size
        self.centroids = None           # d x k float32 (Julia layout)
        self.centroid_norms = None
        self.pending_assignments = None
        self.n_iters = None
        self.init_centroids = None      # optional d x k injected initial centroids


def _point_dataset(X, distance):
    """The device dataset the list scan reads: unit bf16 rows for CosineDist,
    plain rows for Euclidean / SqEuclidean."""
    from .index import CosineDist
    return _dataset_for(X, CosineDist() if isinstance(distance, CosineDist) else Euclidean())


def _centroids_rows(C_dk):
    """d x k (Julia) -> k x d row-major float32 for the C ABI."""
    return np.ascontiguousarray(np.asarray(C_dk, np.float32).T)


def fit(t: KMeansTransform, X, rng=None):
    """fit!(t, X; rng) (KMeansTransform.jl:106-162)."""
    d, n = _data_shape(X)
    assert t.k <= n, f"Cannot fit {t.k} clusters with only {n} points"
    seed = _seed_of(rng)
    # the subsample draw, the fit on it and the full assignment pass all run in
    # the library (subsample = 0: fit on every point)
    use_sub = t.subsample_size is not None and t.subsample_size < n
    ds = _dataset_for(X, Euclidean())
    init_code = {"kmeans_plus_plus": 1, "random": 2}[t.init]
    init_ptr = None
    if t.init_centroids is not None:
        init_rows = _centroids_rows(t.init_centroids)
        init_code, init_ptr = 0, init_rows.ctypes.data
    out_c = np.zeros((t.k, d), np.float32)
    assign = np.zeros(n, np.int64)
    iters = ctypes.c_int64()
    check(lib().mann_kmeans_fit(ds.handle, t.k, t.max_iters, t.tol, init_code, seed, init_ptr,
                                int(t.subsample_size) if use_sub else 0, out_c.ctypes.data,
                                assign.ctypes.data, ctypes.byref(iters)))
    t.centroids = np.asfortranarray(out_c.T)
    t.centroid_norms = (out_c.astype(np.float64) ** 2).sum(1).astype(np.float32)
    t.pending_assignments = assign
    t.n_iters = int(iters.value)
    return t


def kmeans_assign(X, centroids_dk):
    """assign_clusters!(pairwise_euclidean!) -> (assign 1-based, mindist)."""
    ds = _dataset_for(X, Euclidean())
    C = _centroids_rows(centroids_dk)
    k = C.shape[0]
    a = np.zeros(ds.n, np.int64)
    md = np.zeros(ds.n, np.float32)
    check(lib().mann_kmeans_assign(ds.handle, C.ctypes.data, k, a.ctypes.data, md.ctypes.data))
    return a, md


def kmeans_plus_plus(X, k, seed):
    """init_kmeans_plus_plus with the splitmix64 stream -> (d x k, 1-based idx)."""
    ds = _dataset_for(X, Euclidean())
    out = np.zeros((k, ds.d), np.float32)
    idx = np.zeros(k, np.int64)
    check(lib().mann_kmeanspp(ds.handle, k, seed, out.ctypes.data, idx.ctypes.data))
    return np.asfortranarray(out.T), idx


def take_pending_assignments(t: KMeansTransform):
    a = t.pending_assignments
    t.pending_assignments = None
    return a


def preserves_data(t) -> bool:
    return True


# ---------------------------------------------------------------------------
# routing / merging (indices/multilevel/routing.jl, merge.jl)
# ---------------------------------------------------------------------------

class TopKRouting:
    """routing.jl:36-44."""

    def __init__(self, k: int):
        if k <= 0:
            raise ArgumentError("k must be positive")
        self.k = k


class ExhaustiveRouting:
    pass


def select_indices(strategy, assignment, indices):
    """routing.jl:94-127 (1-based positions)."""
    if isinstance(strategy, ExhaustiveRouting):
        return list(range(1, len(indices) + 1))
    n_children = len(indices)
    if n_children <= 0:
        return []
    dist = np.asarray(assignment.distances, np.float32)
    avail = min(n_children, dist.shape[0])
    if avail <= 0:
        raise ArgumentError("Transform produced no routing distances for TopKRouting")
    kk = min(strategy.k, avail)
    order = np.lexsort((np.arange(avail), dist[:avail]))[:kk]
    return [int(i) + 1 for i in order]


class SimpleMerge:
    pass


class DisjointMerge:
    pass


def merge_results(strategy, result_lists, k):
    """merge.jl:79-169 on lists of Neighbor."""
    allv = [nb for lst in result_lists for nb in lst]
    if not allv:
        return []
    if isinstance(strategy, SimpleMerge):
        srt = sorted(allv, key=lambda nb: nb.dist)   # Python's sort is stable
        out, seen = [], set()
        for nb in srt:
            if nb.id in seen:
                continue
            seen.add(nb.id)
            out.append(nb)
            if len(out) >= min(k, len(allv)):
                break
        return out
    # DisjointMerge: bounded max-heap with strict '<' eviction, then (dist, id)
    cap = min(k, len(allv))
    heap = []
    for nb in allv:
        if len(heap) < cap:
            heap.append(nb)
            x = len(heap) - 1
            while x > 0:
                p = (x - 1) // 2
                if heap[x].dist > heap[p].dist:
                    heap[x], heap[p] = heap[p], heap[x]
                    x = p
                else:
                    break
        elif nb.dist < heap[0].dist:
            heap[0] = nb
            x = 0
            while True:
                l, r, g = 2 * x + 1, 2 * x + 2, x
                if l < len(heap) and heap[l].dist > heap[g].dist:
                    g = l
                if r < len(heap) and heap[r].dist > heap[g].dist:
                    g = r
                if g == x:
                    break
                heap[x], heap[g] = heap[g], heap[x]
                x = g
    return sorted(heap, key=lambda nb: (nb.dist, nb.id))


# ---------------------------------------------------------------------------
# MultiLevelIndex (indices/multilevel/*)
# ---------------------------------------------------------------------------

@dataclass
class TerminalConfig:
    index_type: Any
    params: tuple = ()


@dataclass
class TransformedConfig:
    transform: Any
    routing: Any
    child_config: Any


class HNSWIndex:
    """The HNSWIndex terminal type (hnsw/types.jl:170-216) for
    TerminalConfig(HNSWIndex, params): the GPU builds and searches the HNSW
    graph of every list of a MultiLevelIndex (the IVF-HNSW of
    multilevel/ivf_hnsw.jl:46-88).  params: M, ef_construction, ef_search,
    neighbor_policy (:heuristic / :diversified) as in hnsw/index.jl:10-67."""


HNSW_DEFAULT_M = 16                  # constants.jl:21
HNSW_DEFAULT_EF_CONSTRUCTION = 200   # constants.jl:30
HNSW_DEFAULT_EF_SEARCH = 64          # constants.jl:38
IVF_DEFAULT_NPROBE = 5               # constants.jl:154


@dataclass
class HnswTerminal:
    M: int = HNSW_DEFAULT_M
    ef_construction: int = HNSW_DEFAULT_EF_CONSTRUCTION
    ef_search: int = HNSW_DEFAULT_EF_SEARCH
    neighbor_policy: str = "diversified"

    @classmethod
    def from_params(cls, params):
        kw = dict(params) if params else {}
        unknown = set(kw) - {"M", "ef_construction", "ef_search", "neighbor_policy"}
        if unknown:
            raise ArgumentError(f"unsupported HNSWIndex parameters {sorted(unknown)}")
        t = cls(**kw)
        # build_index(HNSWIndex) argument checks (hnsw/index.jl:25-31, :69-88)
        if t.M <= 0:
            raise ArgumentError("M must be positive")
        if t.ef_construction < 1:
            raise ArgumentError("ef_construction must be at least 1")
        if t.ef_search < 1:
            raise ArgumentError("ef_search must be at least 1")
        if t.neighbor_policy not in ("heuristic", "diversified"):
            raise ArgumentError(f"Unknown neighbor_policy symbol: {t.neighbor_policy}")
        return t

    @property
    def policy_code(self):
        return 0 if self.neighbor_policy == "heuristic" else 1


def _attach_hnsw(handle, terminal: HnswTerminal, seed: int):
    check(lib().mann_ivf_hnsw_build(handle, terminal.M, terminal.ef_construction,
                                    terminal.policy_code, seed & 0xFFFFFFFFFFFFFFFF))


class MultiLevelIndex:
    """One KMeans level over BruteForce (or HNSW) leaves, held on the GPU."""

    def __init__(self, transform, routing, handle, d, n, nlist, merge_strategy, distance):
        self.transform = transform
        self.routing = routing
        self.handle = handle
        self.d, self.n, self.nlist = d, n, nlist
        self.merge_strategy = merge_strategy
        self.distance = distance
        nb = ctypes.c_int64()
        check(lib().mann_ivf_info(handle, ctypes.byref(nb), None))
        sizes = np.zeros(nb.value, np.int64)
        check(lib().mann_ivf_info(handle, ctypes.byref(nb), sizes.ctypes.data))
        self.n_buckets = int(nb.value)
        self.list_sizes = sizes
        # bucket_lookup (builder.jl:133-185): bucket -> child position, 0 if empty
        self.bucket_lookup = np.zeros(self.n_buckets, np.int64)
        pos = 0
        for b in range(self.n_buckets):
            if sizes[b] > 0:
                pos += 1
                self.bucket_lookup[b] = pos
        self.root = self
        self.terminal = None    # HnswTerminal for HNSW leaves

    @property
    def id_mappings(self):
        """id_mappings[child] = ascending global ids (transforms/utils.jl:160-171)."""
        ids = np.zeros(self.n, np.int64)
        check(lib().mann_ivf_list_ids(self.handle, ids.ctypes.data))
        off = np.concatenate([[0], np.cumsum(self.list_sizes)])
        return [ids[off[b]:off[b + 1]] for b in range(self.n_buckets) if self.list_sizes[b] > 0]

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        try:
            h = getattr(self, "handle", None)
            if h is not None and h.value and not shutting_down():
                lib().mann_ivf_destroy(h)
                self.handle = None
        except Exception:
            pass


def build_index(index_type, X, config=None, merge_strategy=None, distance=default_distance,
                rng=None, **kw):
    """build_index(MultiLevelIndex, X, config; merge_strategy, distance, rng)
    (builder.jl:61-76) for a KMeans transform over BruteForce terminals, or
    build_index(IVFFlatIndex, X; nlist, ...) (ivf_flat.jl:30-93)."""
    if index_type is IVFFlatIndex:
        if config is not None or merge_strategy is not None:
            raise ArgumentError("IVFFlatIndex takes keyword arguments, not a config")
        return build_ivf_flat(X, distance=distance, rng=rng, **kw)
    if kw:
        raise ArgumentError(f"unexpected keyword arguments {sorted(kw)}")
    if index_type is not MultiLevelIndex:
        raise ArgumentError(f"unsupported index type {index_type}")
    if not isinstance(config, TransformedConfig) or not isinstance(config.transform, KMeansTransform):
        raise ArgumentError("the GPU MultiLevelIndex needs a KMeansTransform root")
    child = config.child_config
    if not (isinstance(child, TerminalConfig)
            and child.index_type in (BruteForceIndex, HNSWIndex)):
        raise ArgumentError("the GPU MultiLevelIndex needs BruteForce or HNSW terminals")
    hnsw = HnswTerminal.from_params(child.params) if child.index_type is HNSWIndex else None
    import copy
    t = copy.deepcopy(config.transform)     # builder.jl:103
    fit(t, X, rng=rng)
    assign = take_pending_assignments(t)
    # HNSW leaves measure the Euclidean distance (the HNSWIndex default)
    ds = _point_dataset(X, Euclidean() if hnsw else distance)
    C = _centroids_rows(t.centroids)
    h = ctypes.c_void_p()
    check(lib().mann_ivf_build(ds.handle, C.ctypes.data, t.k, assign.ctypes.data, ctypes.byref(h)))
    idx = MultiLevelIndex(t, config.routing, h, ds.d, ds.n, t.k,
                          merge_strategy or SimpleMerge(), distance)
    if hnsw is not None:
        # levels: an independent splitmix64 stream next to the k-means draws
        _attach_hnsw(h, hnsw, _seed_of(rng) * 0x9E3779B97F4A7C15 + 2)
        idx.terminal = hnsw
    return idx


def build_ivf_hnsw_index(X, nlist, routing_k=IVF_DEFAULT_NPROBE, kmeans_distance=None,
                         kmeans_init="kmeans_plus_plus", kmeans_max_iters=25, kmeans_tol=1e-4,
                         kmeans_subsample_size="default", hnsw_M=HNSW_DEFAULT_M,
                         hnsw_ef_construction=HNSW_DEFAULT_EF_CONSTRUCTION,
                         hnsw_ef_search=HNSW_DEFAULT_EF_SEARCH,
                         hnsw_neighbor_policy="diversified", merge_strategy=None,
                         distance=default_distance, rng=None):
    """build_ivf_hnsw_index (multilevel/ivf_hnsw.jl:46-88): KMeans(nlist) routing
    to per-list HNSW graphs, DisjointMerge by default."""
    if nlist <= 0:
        raise ArgumentError(f"nlist must be positive, got {nlist}")
    if kmeans_subsample_size == "default":
        kmeans_subsample_size = max(256 * nlist, 32_768)
    routing = TopKRouting(min(max(routing_k, 1), nlist))
    transform = KMeansTransform(k=nlist, distance=kmeans_distance or Euclidean(),
                                init=kmeans_init, max_iters=kmeans_max_iters, tol=kmeans_tol,
                                subsample_size=kmeans_subsample_size)
    params = dict(M=hnsw_M, ef_construction=hnsw_ef_construction, ef_search=hnsw_ef_search,
                  neighbor_policy=hnsw_neighbor_policy)
    config = TransformedConfig(transform, routing, TerminalConfig(HNSWIndex, params))
    return build_index(MultiLevelIndex, X, config,
                       merge_strategy=merge_strategy or DisjointMerge(), distance=distance,
                       rng=rng)


def ivf_hnsw_from_assignment(X, centroids_dk, assign, routing, hnsw: HnswTerminal, seed=0,
                             merge_strategy=None):
    """An IVF-HNSW index from an existing k-means fit (centroids d x k,
    1-based assignment) with level seed `seed`."""
    idx = build_index_from_assignment(X, centroids_dk, assign, routing,
                                      merge_strategy=merge_strategy or DisjointMerge())
    _attach_hnsw(idx.handle, hnsw, seed)
    idx.terminal = hnsw
    return idx


def set_hnsw_graph(index: MultiLevelIndex, nbr, deg, entry, max_layer, level=None,
                   ef_search=HNSW_DEFAULT_EF_SEARCH):
    """Install graphs built elsewhere (the layout of hnsw_graph) as the index's
    HNSW terminals."""
    nbr = np.ascontiguousarray(nbr, np.int32)
    deg = np.ascontiguousarray(deg, np.int32)
    layers, n, cap = nbr.shape
    if n != index.n or deg.shape != (layers, n):
        raise DimensionMismatch("graph arrays do not match the index")
    lv = None if level is None else np.ascontiguousarray(level, np.int32)
    check(lib().mann_ivf_hnsw_set_graph(index.handle, layers, cap - 1, nbr.ctypes.data,
                                        deg.ctypes.data,
                                        np.ascontiguousarray(entry, np.int32).ctypes.data,
                                        np.ascontiguousarray(max_layer, np.int32).ctypes.data,
                                        None if lv is None else lv.ctypes.data))
    index.terminal = HnswTerminal(M=cap - 1, ef_search=ef_search)


def hnsw_graph(index: MultiLevelIndex):
    """The HNSW graphs of an IVF-HNSW index: (nbr (layers, n, M + 1) local 0-based
    ids, deg (layers, n), entry and max_layer per bucket, level per row), rows in
    list-contiguous order."""
    L, M = ctypes.c_int64(), ctypes.c_int64()
    check(lib().mann_ivf_hnsw_graph(index.handle, ctypes.byref(L), ctypes.byref(M), None, None,
                                    None, None, None))
    L, M, n, nb = L.value, M.value, index.n, index.n_buckets
    nbr = np.zeros((L, n, M + 1), np.int32)
    deg = np.zeros((L, n), np.int32)
    entry = np.zeros(nb, np.int32)
    maxl = np.zeros(nb, np.int32)
    level = np.zeros(n, np.int32)
    check(lib().mann_ivf_hnsw_graph(index.handle, None, None, nbr.ctypes.data, deg.ctypes.data,
                                    entry.ctypes.data, maxl.ctypes.data, level.ctypes.data))
    return nbr, deg, entry, maxl, level


def build_index_from_assignment(X, centroids_dk, assign, routing, merge_strategy=None,
                                distance=default_distance):
    """The IVF layout from an existing k-means fit (centroids d x k, 1-based assignment)."""
    t = KMeansTransform(k=np.asarray(centroids_dk).shape[1])
    t.centroids = np.asfortranarray(np.asarray(centroids_dk, np.float32))
    ds = _point_dataset(X, distance)
    C = _centroids_rows(t.centroids)
    a = np.ascontiguousarray(assign, np.int64)
    h = ctypes.c_void_p()
    check(lib().mann_ivf_build(ds.handle, C.ctypes.data, t.k, a.ctypes.data, ctypes.byref(h)))
    return MultiLevelIndex(t, routing, h, ds.d, ds.n, t.k, merge_strategy or SimpleMerge(),
                           distance)


class IVFFlatIndex:
    """IVFFlatIndex (ivf_flat.jl:11-19): k-means centroids and one inverted list
    of point ids per centroid (empty lists included), held on the GPU with a
    list-contiguous copy of the base; the data is still passed to every query,
    as the shared ANN contract asks."""

    def __init__(self, kmeans, handle, d, n, nlist, default_nprobe, distance, centroid_metric):
        self.kmeans = kmeans
        self.centroids = kmeans.centroids
        self.centroid_norms = kmeans.centroid_norms
        self.handle = handle
        self.dimension = d
        self.d, self.n, self.nlist = d, n, nlist
        self.default_nprobe = default_nprobe
        self.distance = distance
        self.centroid_metric = centroid_metric
        self.routing = TopKRouting(default_nprobe)
        self.root = self
        sizes = np.zeros(nlist, np.int64)
        nb = ctypes.c_int64()
        check(lib().mann_ivf_info(handle, ctypes.byref(nb), sizes.ctypes.data))
        self.list_sizes = sizes

    @property
    def lists(self):
        """lists[c] = ascending ids of the points assigned to centroid c (ivf_flat.jl:66-77)."""
        ids = np.zeros(self.n, np.int64)
        check(lib().mann_ivf_list_ids(self.handle, ids.ctypes.data))
        off = np.concatenate([[0], np.cumsum(self.list_sizes)])
        return [ids[off[c]:off[c + 1]] for c in range(self.nlist)]

    def __del__(self):
        try:
            h = getattr(self, "handle", None)
            if h is not None and h.value and not shutting_down():
                lib().mann_ivf_destroy(h)
                self.handle = None
        except Exception:
            pass


def build_ivf_flat(X, nlist, nprobe=10, distance=default_distance, centroid_metric=None,
                   kmeans_max_iters=25, kmeans_tol=1e-4, kmeans_subsample_size="default",
                   rng=None):
    """build_index(IVFFlatIndex, data; nlist, nprobe=10, distance, centroid_metric,
    kmeans_max_iters=25, kmeans_tol=1e-4, kmeans_subsample_size=max(256 nlist, 32768))
    (ivf_flat.jl:30-93): k-means++ Lloyd on the subsample, one full assignment
    pass, one list per centroid.  The routing metric is Euclidean or
    SqEuclidean (same order; other centroid metrics are not offered here)."""
    d, n = _data_shape(X)
    if d <= 0:
        raise ArgumentError("Dataset must have at least one dimension")
    if n <= 0:
        raise ArgumentError("Dataset must contain at least one point")
    if nlist <= 0:
        raise ArgumentError("nlist must be positive")
    centroid_metric = Euclidean() if centroid_metric is None else centroid_metric
    if type(centroid_metric).__name__ not in ("Euclidean", "SqEuclidean", "CosineDist"):
        raise ArgumentError("centroid_metric must be Euclidean, SqEuclidean or CosineDist")
    if type(distance).__name__ not in ("Euclidean", "SqEuclidean", "CosineDist"):
        raise ArgumentError("distance must be Euclidean, SqEuclidean or CosineDist")
    if kmeans_subsample_size == "default":
        kmeans_subsample_size = max(256 * nlist, 32768)
    # SqEuclidean orders the centroids as Euclidean does: one k-means serves both
    t = KMeansTransform(k=nlist, distance=Euclidean(), init="kmeans_plus_plus",
                        max_iters=kmeans_max_iters, tol=kmeans_tol,
                        subsample_size=kmeans_subsample_size)
    fit(t, X, rng=rng)
    assign = take_pending_assignments(t)
    if assign is None or len(assign) != n:
        raise ArgumentError("KMeansTransform did not return assignments")
    return ivf_flat_from_assignment(X, t, assign, nprobe, distance, centroid_metric)


def ivf_flat_from_assignment(X, t, assign, nprobe=10, distance=default_distance,
                             centroid_metric=None):
    """IVFFlatIndex from a fitted KMeansTransform and its 1-based assignment.
    The lists hold the rows for the point metric `distance`; routing uses
    `centroid_metric` (ivf_flat.jl:103-131)."""
    centroid_metric = centroid_metric or Euclidean()
    ds = _point_dataset(X, distance)
    C = _centroids_rows(t.centroids)
    a = np.ascontiguousarray(assign, np.int64)
    h = ctypes.c_void_p()
    check(lib().mann_ivf_flat_build(ds.handle, C.ctypes.data, t.k, a.ctypes.data,
                                    centroid_metric.code, ctypes.byref(h)))
    return IVFFlatIndex(t, h, ds.d, ds.n, t.k, max(1, min(int(nprobe), t.k)), distance,
                        centroid_metric)


def transform(t, x):
    """transform(::KMeansTransform, x) (KMeansTransform.jl:181-190): the FP32
    norms-form distances to every centroid, computed on the GPU."""
    if t.centroids is None:
        raise ArgumentError("KMeansTransform must be fitted before transforming")
    qm = as_matrix(x)
    C = _centroids_rows(t.centroids)
    if qm.d != C.shape[1]:
        raise DimensionMismatch(f"Expected query length {C.shape[1]}")
    out = np.zeros((qm.n, C.shape[0]), np.float32)
    check(lib().mann_route_distances(C.ctypes.data, C.shape[0], qm.d, qm.ptr, qm.n, qm.ld,
                                     out.ctypes.data))
    if np.ndim(x) == 1:
        return TransformResult(x, KMeansAssignment(out[0]))
    return [TransformResult(None, KMeansAssignment(r)) for r in out]


def route_distances(index: MultiLevelIndex, Q):
    """Routing distances (nq, nlist) of the index's KMeans transform."""
    qm = as_matrix(Q)
    if qm.d != index.d:
        raise DimensionMismatch(f"Expected queries with {index.d} rows, got {qm.d}")
    out = np.zeros((qm.n, index.nlist), np.float32)
    check(lib().mann_ivf_route(index.handle, qm.ptr, qm.d, qm.n, qm.ld, out.ctypes.data))
    return out


def probe_lists(index: MultiLevelIndex, route):
    """TopKRouting + empty-bucket removal per query -> (nq, nprobe) 1-based buckets, 0 pad."""
    route = np.ascontiguousarray(route, np.float32)
    nq = route.shape[0]
    nprobe = index.routing.k
    out = np.zeros((nq, nprobe), np.int64)
    check(lib().mann_ivf_probe(index.handle, route.ctypes.data, nq, nprobe, out.ctypes.data))
    return out


def query_arrays(index: MultiLevelIndex, data, Q, k, route=None, nprobe=None,
                 probe_order=False, ef_search=None):
    """Batched query (query.jl:279-300) -> (ids (nq, k), dist (nq, k)).
    probe_order=True merges each query's per-list results in probe order, as
    the single-query path does (query.jl:25-106); the index's merge strategy
    (SimpleMerge / DisjointMerge) applies either way."""
    qm = as_matrix(Q)
    d, _ = _data_shape(data)
    if isinstance(index, IVFFlatIndex):
        # _validate_dimensions and the nprobe clamp (ivf_flat.jl:95-101, 146)
        if d != index.dimension:
            raise DimensionMismatch(f"Expected data with {index.dimension} rows")
        if qm.d != index.dimension:
            raise DimensionMismatch(f"Expected query of length {index.dimension}")
        npb = index.default_nprobe if nprobe is None else int(nprobe)
        nprobe = min(max(npb, 1), index.nlist)
    if qm.d != d:
        raise DimensionMismatch(f"Expected queries with {d} rows, got {qm.d}")
    nq = qm.n
    k = max(int(k), 0)
    npb = nprobe if nprobe is not None else index.routing.k
    if qm.device:
        import torch
        ids = torch.zeros((nq, k), dtype=torch.int64, device=qm.keep.device)
        dist = torch.full((nq, k), float("inf"), dtype=torch.float32, device=qm.keep.device)
        pi, pd = _ready(ids).data_ptr(), _ready(dist).data_ptr()
    else:
        ids = np.zeros((nq, k), np.int64)
        dist = np.full((nq, k), np.inf, np.float32)
        pi, pd = ids.ctypes.data, dist.ctypes.data
    if nq == 0 or k == 0:
        return ids, dist
    rp = None
    if route is not None:
        rp = _ready(route).data_ptr() if _is_torch(route) else np.ascontiguousarray(route,
                                                                                  np.float32)
        rp = rp if isinstance(rp, int) else rp.ctypes.data
    merge = 1 if isinstance(getattr(index, "merge_strategy", None), DisjointMerge) else 0
    term = getattr(index, "terminal", None)
    if term is not None:
        # HNSW leaves take ef_search (hnsw/query.jl:63-75, passed down by the
        # multi-level query's keyword arguments)
        ef = term.ef_search if ef_search is None else int(ef_search)
        check(lib().mann_ivf_hnsw_search(index.handle, qm.ptr, qm.d, nq, qm.ld, k, npb, ef, rp,
                                         merge, int(bool(probe_order)), pi, pd))
        return ids, dist
    if ef_search is not None:
        raise ArgumentError("ef_search is a keyword of HNSW leaves")
    check(lib().mann_ivf_search_ex(index.handle, qm.ptr, qm.d, nq, qm.ld, k, npb,
                                   index.distance.code, rp, merge, int(bool(probe_order)), pi,
                                   pd))
    return ids, dist


def query(index, data, q, k, nprobe=None, ef_search=None):
    """query(index, data, q | Q, k): one vector goes through the single-query
    path (per-list results in probe order, query.jl:25-106), a matrix through
    the batched one (child order, query.jl:279-399)."""
    single = (np.ndim(q) == 1) if not _is_torch(q) else q.dim() == 1
    ids, dist = query_arrays(index, data, q, k, nprobe=nprobe, probe_order=single,
                             ef_search=ef_search)
    res = _to_neighbors(ids, dist)
    return res[0] if single else res


def autotune(index: MultiLevelIndex, Q, k, nprobe=None, route=None):
    """Time the batched search's knobs on this batch (list-scan tile, work-item
    rows, first-list threshold tile, shortlist) and keep the fastest for the
    shape; later searches of the shape use it.  Returns the choice.  No
    reference counterpart."""
    if isinstance(index, IVFFlatIndex):
        raise ArgumentError("the IVF autotuner covers MultiLevelIndex searches")
    qm = as_matrix(Q)
    npb = index.routing.k if nprobe is None else int(nprobe)
    rp = None
    if route is not None:
        rp = _ready(route).data_ptr() if _is_torch(route) else \
            np.ascontiguousarray(route, np.float32)
        rp = rp if isinstance(rp, int) else rp.ctypes.data
    choice = (ctypes.c_int64 * 5)()
    ms = ctypes.c_double()
    check(lib().mann_autotune_ivf(index.handle, qm.ptr, qm.d, qm.n, qm.ld, int(k), npb,
                                  index.distance.code, rp, choice, ctypes.byref(ms)))
    out = dict(zip(("ivf_cfg", "ivf_rows_item", "ivf_tau_cfg", "shortlist_mult", "trials"),
                   list(choice)))
    out["ms"] = ms.value
    return out


class IVFSearchGraph:
    """The batched MultiLevelIndex query (query_arrays with the batched
    SimpleMerge order) captured in a HIP graph for one batch shape.  Q is a
    d x nq float32 CUDA matrix with contiguous columns (rows.t() of an nq x d
    tensor), refilled in place between launches; `route`, if given, is an
    nq x nlist CUDA tensor of routing distances read at every launch.
    launch() replays the search and returns (ids, dist), overwritten by the
    next launch.  No reference counterpart."""

    def __init__(self, index: MultiLevelIndex, Q, k, nprobe=None, route=None):
        import torch
        if isinstance(index, IVFFlatIndex):
            raise ArgumentError("IVF-Flat searches (heap merge) are not captured")
        if not _is_torch(Q) or not Q.is_cuda or Q.dtype != torch.float32 or Q.dim() != 2:
            raise ArgumentError("IVFSearchGraph takes a d x nq float32 CUDA tensor")
        if Q.shape[1] > 1 and Q.stride(0) != 1:
            raise ArgumentError("IVFSearchGraph needs columns contiguous in memory (rows.t())")
        if isinstance(getattr(index, "merge_strategy", None), DisjointMerge):
            raise ArgumentError("a captured IVF search merges in the SimpleMerge order only")
        d, nq = Q.shape
        self.index, self.Q, self.route = index, Q, route
        self.ids = torch.zeros((nq, int(k)), dtype=torch.int64, device=Q.device)
        self.dist = torch.zeros((nq, int(k)), dtype=torch.float32, device=Q.device)
        npb = index.routing.k if nprobe is None else int(nprobe)
        rp = None
        if route is not None:
            # the graph reads this buffer at every launch: it must be the
            # caller's own contiguous tensor (no temporary copy)
            if (not _is_torch(route) or not route.is_cuda or route.dtype != torch.float32
                    or tuple(route.shape) != (nq, index.nlist) or not route.is_contiguous()):
                raise ArgumentError("route must be a contiguous nq x nlist float32 CUDA tensor")
            rp = _ready(route).data_ptr()
        _ready(self.ids)
        _ready(self.dist)
        _ready(Q)
        h = ctypes.c_void_p()
        ld = Q.stride(1) if nq > 1 else d
        check(lib().mann_ivf_search_graph_create(index.handle, Q.data_ptr(), d, nq, ld, int(k),
                                                 npb, index.distance.code, rp,
                                                 self.ids.data_ptr(), self.dist.data_ptr(),
                                                 ctypes.byref(h)))
        self.handle = h

    def launch(self):
        _ready(self.Q)
        if self.route is not None:
            _ready(self.route)
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

