"""The benchmark harness's wrapper contract over the GPU library.

The reference times its Julia indexes through Python wrappers
(benchmarking/benchmarking/wrappers/base.py:8-200, manifoldann.py:15-259,
261-287, 740-824) driven by benchmarking/benchmark.py:741-828:

  prepare_data(X)        untimed: numpy (n, d) -> the library's data matrix
  fit(prepared)          timed build
  prepare_queries(Q)     untimed: numpy (nq, d) -> the library's query matrix
  query_batch_raw(Q, k)  timed: one batched search, library-native result
  finalize_batch_ids(r)  untimed: 0-based id lists per query
  query(v, k)            one vector (the single-query path)
  set_query_params(...)  query-time knobs (nprobe) without a rebuild

The classes here keep those method names, argument meanings and errors, with
the GPU taking Julia's place:
- prepare_data uploads the rows to HBM once (the reference converts them to a
  Julia Matrix{Float32}; both stay outside the timed regions);
- fit builds the index and the device copy of the data the searches read
  (bf16 operands and norms; for IVF-Flat the k-means fit and the lists);
- query_batch_raw returns the (ids, distances) device tensors of one batched
  search, complete on return, with 1-based ids as the Julia path returns;
- finalize_batch_ids converts them to 0-based Python lists, as
  _batch_neighbors_to_ids (manifoldann.py:152-164) does.
"""
from __future__ import annotations

from typing import Any, List

import numpy as np

from . import index as _ix
from . import ivf as _ivf
from ._lib import ArgumentError

METRICS = ("angular", "euclidean")


def _device_matrix(X: np.ndarray):
    """numpy (n, d) -> the d x n column-major device view of its rows (one
    contiguous upload; column j is point j, as Julia's Matrix{Float32})."""
    import torch
    a = np.ascontiguousarray(np.asarray(X, dtype=np.float32))
    if a.ndim != 2:
        raise ArgumentError("expected a 2-D array of shape (n_samples, n_features)")
    t = torch.from_numpy(a).to("cuda")
    torch.cuda.synchronize()
    return t.t()


def ids_to_lists(ids) -> List[List[int]]:
    """(nq, k) 1-based ids (0 = no neighbour) -> 0-based id lists per query.

    Same result as the harness's _batch_neighbors_to_ids (manifoldann.py:
    152-164) on a batch of Vector{Neighbor}: the library pads a short result
    (k > available points) with id 0, which a Julia result simply lacks."""
    if hasattr(ids, "cpu"):
        ids = ids.cpu().numpy()
    ids = np.asarray(ids, dtype=np.int64)
    out = []
    for row in ids:
        row = row[row > 0]
        out.append((row - 1).tolist())
    return out


class ManifoldANNHIPWrapper:
    """Base wrapper (manifoldann.py:15-259 over base.py:8-200)."""

    def __init__(self, metric: str):
        if metric not in METRICS:
            raise ValueError(f"Invalid metric: {metric}. Must be 'angular' or 'euclidean'")
        self._metric = metric
        self._index = None
        self._data = None

    # the reference picks cosine_distance for 'angular' and default_distance
    # (Euclidean) otherwise, matching the ground truth (manifoldann.py:178-191)
    def _get_distance_function(self):
        return _ix.CosineDist() if self._metric == "angular" else _ix.Euclidean()

    def prepare_data(self, X: np.ndarray) -> Any:
        return _device_matrix(X)

    def prepare_queries(self, Q: np.ndarray) -> Any:
        return _device_matrix(Q)

    def set_query_params(self, **params) -> None:
        return None

    def set_num_threads(self, n: int) -> None:
        """No host thread pool to pin: the search runs on the GPU."""
        return None

    def warmup_build(self, dim: int, n: int = 200) -> None:
        """Build and query once on a small random matrix of the actual
        dimension, so that the first timed call loads no code objects."""
        rng = np.random.default_rng(0)
        X = rng.standard_normal((max(n, 64), dim)).astype(np.float32)
        saved = (self._index, self._data)
        try:
            self.fit(self.prepare_data(X))
            self.finalize_batch_ids(self.query_batch_raw(self.prepare_queries(X[:4]), 5))
        finally:
            self._index, self._data = saved

    def memory_usage(self):
        return None

    def fit(self, X) -> None:
        raise NotImplementedError

    def _query_arrays(self, Q, n):
        raise NotImplementedError

    def query_batch_raw(self, queries, n):
        if isinstance(queries, np.ndarray):
            queries = _device_matrix(queries)
        return self._query_arrays(queries, n)

    def finalize_batch_ids(self, raw) -> List[List[int]]:
        ids, _dist = raw
        return ids_to_lists(ids)

    def query_batch(self, queries, n) -> List[List[int]]:
        return self.finalize_batch_ids(self.query_batch_raw(queries, n))

    def query(self, v: np.ndarray, n: int) -> List[int]:
        """One vector through the library's single-query path."""
        q = np.asarray(v, dtype=np.float32).reshape(-1)
        res = _ix.query(self._index, self._data, q, n) if not hasattr(self._index, "root") \
            else _ivf.query(self._index, self._data, q, n, nprobe=self._query_nprobe())
        return [nb.id - 1 for nb in res]

    def _query_nprobe(self):
        return None

    @staticmethod
    def is_available() -> bool:
        try:
            from ._lib import lib
            lib()
            return True
        except Exception:
            return False


class ManifoldANN_BruteForce(ManifoldANNHIPWrapper):
    """BruteForceIndex (manifoldann.py:261-287)."""

    def fit(self, X) -> None:
        dist = self._get_distance_function()
        self._index = _ix.build_index(_ix.BruteForceIndex, X, distance=dist)
        # the device copy the exact scan reads (bf16 rows, norms)
        self._data = _ix.Dataset(X, dist)

    def _query_arrays(self, Q, n):
        return _ix.query_arrays(self._index, self._data, Q, n)

    def __str__(self):
        return "ManifoldANN-BruteForce()"

    @staticmethod
    def get_name():
        return "ManifoldANN-BruteForce"


class ManifoldANN_IVFFlat(ManifoldANNHIPWrapper):
    """IVFFlatIndex (manifoldann.py:740-824): nlist lists, nprobe probed,
    centroid metric CosineDist for 'angular', Euclidean otherwise."""

    def __init__(self, metric, nlist=100, nprobe=10, kmeans_init="kmeans_plus_plus",
                 kmeans_max_iters=25, kmeans_tol=1e-4):
        super().__init__(metric)
        self._nlist = int(nlist)
        self._nprobe = int(nprobe)
        self._kmeans_init = kmeans_init
        self._kmeans_max_iters = int(kmeans_max_iters)
        self._kmeans_tol = float(kmeans_tol)

    def _kmeans_metric(self):
        return _ix.CosineDist() if self._metric == "angular" else _ix.Euclidean()

    def fit(self, X) -> None:
        # the reference's fit passes nlist, nprobe and the two metrics only;
        # the k-means settings keep build_index's defaults (25 iterations,
        # tol 1e-4), which are also this wrapper's defaults
        self._data = X
        self._index = _ix.build_index(_ivf.IVFFlatIndex, X, nlist=self._nlist,
                                      nprobe=self._nprobe,
                                      distance=self._get_distance_function(),
                                      centroid_metric=self._kmeans_metric())

    def set_query_params(self, **params) -> None:
        if "nprobe" in params:
            self._nprobe = int(params["nprobe"])

    def _query_nprobe(self):
        return self._nprobe

    def _query_arrays(self, Q, n):
        return _ivf.query_arrays(self._index, self._data, Q, n, nprobe=self._nprobe)

    def __str__(self):
        return ("ManifoldANN-IVF-Flat("
                f"nlist={self._nlist}, nprobe={self._nprobe}, "
                f"kmeans_init={self._kmeans_init}, kmeans_max_iters={self._kmeans_max_iters})")

    @staticmethod
    def get_name():
        return "ManifoldANN-IVF-Flat"


class ManifoldANN_IVFHNSW(ManifoldANNHIPWrapper):
    """build_ivf_hnsw_index (manifoldann.py:649-737): KMeans(nlist) routing to
    per-list HNSW graphs built and searched on the GPU (M, ef_construction,
    ef_search, neighbour policy), routing_k clamped to [1, nlist],
    DisjointMerge.  The HNSW leaves measure the Euclidean distance for both
    metrics, as the reference's HNSWIndex children do (they receive no
    distance); the GPU k-means is Euclidean, where the reference clusters
    'angular' data with CosineDist."""

    def __init__(self, metric, nlist=64, routing_k=8, kmeans_init="kmeans_plus_plus",
                 kmeans_max_iters=50, kmeans_tol=1e-4, M=16, ef_construction=200,
                 ef_search=64, neighbor_policy="diversified"):
        super().__init__(metric)
        self._nlist = int(nlist)
        self._routing_k = int(routing_k)
        self._kmeans_init = kmeans_init
        self._kmeans_max_iters = int(kmeans_max_iters)
        self._kmeans_tol = float(kmeans_tol)
        self._M = int(M)
        self._ef_construction = int(ef_construction)
        self._ef_search = int(ef_search)
        self._neighbor_policy = neighbor_policy

    def fit(self, X) -> None:
        self._data = X
        routing_k = max(1, min(self._routing_k, self._nlist))
        self._index = _ivf.build_ivf_hnsw_index(
            X, self._nlist, routing_k=routing_k, kmeans_init=self._kmeans_init,
            kmeans_max_iters=self._kmeans_max_iters, kmeans_tol=self._kmeans_tol,
            hnsw_M=self._M, hnsw_ef_construction=self._ef_construction,
            hnsw_ef_search=self._ef_search, hnsw_neighbor_policy=self._neighbor_policy,
            distance=self._get_distance_function())

    def set_query_params(self, **params) -> None:
        if "ef_search" in params:
            self._ef_search = int(params["ef_search"])

    def _query_arrays(self, Q, n):
        return _ivf.query_arrays(self._index, self._data, Q, n, ef_search=self._ef_search)

    def query(self, v: np.ndarray, n: int) -> List[int]:
        q = np.asarray(v, dtype=np.float32).reshape(-1)
        res = _ivf.query(self._index, self._data, q, n, ef_search=self._ef_search)
        return [nb.id - 1 for nb in res]

    def __str__(self):
        return ("ManifoldANN-IVF-HNSW("
                f"nlist={self._nlist}, routing_k={self._routing_k}, "
                f"M={self._M}, ef_construction={self._ef_construction}, "
                f"ef_search={self._ef_search}, neighbor_policy={self._neighbor_policy}, "
                f"kmeans_init={self._kmeans_init})")

    @staticmethod
    def get_name():
        return "ManifoldANN-IVF-HNSW"


def compute_recall(predicted: List[int], ground_truth, k: int = None) -> float:
    """|set(predicted[:k]) ∩ set(truth[:k])| / |set(truth[:k])|, 0 for an
    empty truth; k = len(predicted) when None (benchmarking/utils/
    evaluation.py:7-30)."""
    if k is None:
        k = len(predicted)
    true_k = set(np.asarray(ground_truth)[:k].tolist())
    if not true_k:
        return 0.0
    return len(set(predicted[:k]) & true_k) / len(true_k)


def compute_recall_batch(predictions: List[List[int]], ground_truth, k: int = None) -> float:
    """Mean of compute_recall over the queries, 0 for an empty batch
    (evaluation.py:33-52)."""
    if len(predictions) == 0:
        return 0.0
    return float(np.mean([compute_recall(p, ground_truth[i], k)
                          for i, p in enumerate(predictions)]))
