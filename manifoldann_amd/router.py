"""Learned probe router over the HIP library.

No counterpart exists in the reference: the router is a transform that plugs
into its AbstractTransform seam (transforms/Transform.jl:36-83).  Its
`transform(q)` returns KMeansAssignment(-scores), so the unchanged TopKRouting
(indices/multilevel/routing.jl:94-118) probes the highest-scoring lists first.
Training labels are the lists holding each training query's exact top-k,
computed by the exact GPU kNN over the same base.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import ArgumentError, check, lib, shutting_down
from .index import Dataset, _is_torch, _ready, as_matrix
from .ivf import KMeansAssignment, TransformResult, preserves_data  # noqa: F401


class LearnedProbeRouter:
    """MLP d -> hidden (ReLU) -> nlist with FP32 master weights (bf16 GEMMs)."""

    def __init__(self, d: int, nlist: int, hidden: int = 512, seed: int = 0):
        self.d, self.nlist, self.hidden = d, nlist, hidden
        h = ctypes.c_void_p()
        check(lib().mann_router_create(d, hidden, nlist, seed, ctypes.byref(h)))
        self.handle = h
        self.gamma = 0.0

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        try:
            h = getattr(self, "handle", None)
            if h is not None and h.value and not shutting_down():
                lib().mann_router_destroy(h)
                self.handle = None
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        c = ctypes.c_int64()
        check(lib().mann_router_params(self.handle, None, ctypes.byref(c)))
        return int(c.value)

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        check(lib().mann_router_params(self.handle, out.ctypes.data, None))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ArgumentError("parameter vector has the wrong length")
        check(lib().mann_router_set_params(self.handle, p.ctypes.data))

    def set_prior(self, index, gamma: float):
        """Add gamma * (2 q.c - |c|^2) of the index's centroids to the logits
        (gamma = 0 removes it); the MLP then learns a correction of the
        centroid routing."""
        check(lib().mann_router_set_prior(self.handle, index.handle if index is not None else None,
                                          float(gamma)))
        self.gamma = float(gamma)

    def zero_output_layer(self):
        """W2 = 0, b2 = 0: with a prior the untrained router equals it."""
        p = self.get_params()
        o = self.hidden * self.d + self.hidden
        p[o:] = 0.0
        self.set_params(p)

    def fit_prior(self, index, X, labels, grid=None):
        """Scale of the centroid prior minimising the prior-only cross-entropy
        on (X, labels) (golden-section refinement of a log-spaced grid).
        Leaves the prior attached with that scale and returns it."""
        saved = self.get_params()
        self.set_params(np.zeros_like(saved))
        grid = grid if grid is not None else [2.0 ** e for e in range(-12, 3)]

        def loss_at(g):
            self.set_prior(index, g)
            return self.grad(X, labels)[0]

        losses = [loss_at(g) for g in grid]
        i = int(np.argmin(losses))
        lo = np.log(grid[max(i - 1, 0)])
        hi = np.log(grid[min(i + 1, len(grid) - 1)])
        phi = (np.sqrt(5.0) - 1.0) / 2.0
        a, b = hi - phi * (hi - lo), lo + phi * (hi - lo)
        fa, fb = loss_at(np.exp(a)), loss_at(np.exp(b))
        for _ in range(10):
            if fa < fb:
                hi, b, fb = b, a, fa
                a = hi - phi * (hi - lo)
                fa = loss_at(np.exp(a))
            else:
                lo, a, fa = a, b, fb
                b = lo + phi * (hi - lo)
                fb = loss_at(np.exp(b))
        best = float(np.exp(a if fa < fb else b))
        if min(fa, fb) > losses[i]:
            best = float(grid[i])
        self.set_params(saved)
        self.set_prior(index, best)
        return best

    def grad(self, X, labels):
        """(loss, gradient) of the mean cross-entropy on one batch; X is d x n."""
        m = as_matrix(X)
        lab = np.ascontiguousarray(labels, np.int64)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_double()
        check(lib().mann_router_grad(self.handle, m.ptr, m.d, m.n, m.ld, lab.ctypes.data,
                                     lab.shape[1], g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    def train(self, X, labels, epochs=5, batch=1024, lr=2e-3, seed=1):
        m = as_matrix(X)
        if _is_torch(labels):
            lp, L = _ready(labels).data_ptr(), labels.shape[1]
        else:
            lab = np.ascontiguousarray(labels, np.int64)
            lp, L = lab.ctypes.data, lab.shape[1]
        loss = ctypes.c_double()
        check(lib().mann_router_train(self.handle, m.ptr, m.d, m.n, m.ld, lp, L, epochs, batch,
                                      lr, seed, ctypes.byref(loss)))
        return loss.value

    def scores(self, Q):
        m = as_matrix(Q)
        out = np.zeros((m.n, self.nlist), np.float32)
        check(lib().mann_router_scores(self.handle, m.ptr, m.d, m.n, m.ld, out.ctypes.data))
        return out


class LearnedProbeTransform:
    """Transform seam: wraps the fitted KMeansTransform (which still defines the
    partition) and replaces its routing distances by the router's -scores."""

    def __init__(self, kmeans, router: LearnedProbeRouter):
        self.kmeans = kmeans
        self.router = router

    def transform(self, x):
        s = self.router.scores(np.asarray(x, np.float32).reshape(-1, 1))
        return TransformResult(x, KMeansAssignment(-s[0]))

    def take_pending_assignments(self):
        a = self.kmeans.pending_assignments
        self.kmeans.pending_assignments = None
        return a


def neighbour_list_labels(index, data, train_queries, k=10):
    """Labels of each training query: the (1-based) lists of its exact top-k."""
    from .index import BruteForceIndex, query_arrays
    ds = data if isinstance(data, Dataset) else Dataset(data)
    bf = BruteForceIndex(ds.d, ds.n)
    ids, _ = query_arrays(bf, ds, train_queries, k)
    ids = ids.cpu().numpy() if _is_torch(ids) else ids
    bucket_of = np.zeros(index.n + 1, np.int64)
    all_ids = np.zeros(index.n, np.int64)
    check(lib().mann_ivf_list_ids(index.handle, all_ids.ctypes.data))
    off = np.concatenate([[0], np.cumsum(index.list_sizes)])
    for b in range(index.n_buckets):
        bucket_of[all_ids[off[b]:off[b + 1]]] = b + 1
    return bucket_of[ids]


def last_probe_count() -> int:
    """(list, query) pairs scanned by the calling thread's last IVF search."""
    out = ctypes.c_int64()
    check(lib().mann_last_probe_count(ctypes.byref(out)))
    return int(out.value)


def probe_mass(index, route, nprobe_max, mass):
    """Buckets chosen by probability-mass probing of negated logits `route`
    (nq x nlist): nq x nprobe_max, 1-based, 0 = unused."""
    r = np.ascontiguousarray(route, np.float32)
    out = np.zeros((r.shape[0], nprobe_max), np.int64)
    check(lib().mann_ivf_probe_mass(index.handle, r.ctypes.data, r.shape[0], nprobe_max,
                                    float(mass), out.ctypes.data))
    return out


def query_arrays_routed(index, router: LearnedProbeRouter, Q, k, nprobe, mass=None):
    """IVF batched query with the router's routing: TopKRouting over -scores,
    or (mass given) probability-mass probing with at most nprobe lists."""
    m = as_matrix(Q)
    if m.device:
        import torch
        ids = torch.zeros((m.n, k), dtype=torch.int64, device=m.keep.device)
        dist = torch.zeros((m.n, k), dtype=torch.float32, device=m.keep.device)
        pi, pd = _ready(ids).data_ptr(), _ready(dist).data_ptr()
    else:
        ids = np.zeros((m.n, k), np.int64)
        dist = np.zeros((m.n, k), np.float32)
        pi, pd = ids.ctypes.data, dist.ctypes.data
    if mass is None:
        check(lib().mann_ivf_search_routed(index.handle, router.handle, m.ptr, m.d, m.n, m.ld, k,
                                           nprobe, index.distance.code, pi, pd))
    else:
        check(lib().mann_ivf_search_routed_mass(index.handle, router.handle, m.ptr, m.d, m.n,
                                                m.ld, k, nprobe, float(mass),
                                                index.distance.code, pi, pd))
    return ids, dist
