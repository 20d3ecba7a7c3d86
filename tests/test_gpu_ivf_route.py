"""Certified TopKRouting of the IVF searches (routing.jl:94-118 over the norms
form of kmeans/distance.jl:213-242): a bf16 first pass over the centroids, then
the FP32 norms form of the centroids its bound cannot exclude.  The searches
must equal the oracle's (exact routing distances, ties to the lower bucket) and
the searches with every routing distance in FP32 (`ivf_route_cert` 0)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import datasets
from manifoldann_amd import ivf as I
from manifoldann_amd.router import last_probe_count
from manifoldann_amd._lib import check, lib
from oracle import oracle as O


def _banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _tune(key, value):
    L = lib()
    o = ctypes.c_int64()
    check(L.mann_tune_get(key.encode(), ctypes.byref(o)))
    check(L.mann_tune_set(key.encode(), value))
    return o.value


def _vs_oracle_and_fp32(X, C, a, Q, k, nprobe, tag):
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(nprobe))
    ids, dist = I.query_arrays(index, X.T, Q.T, k)
    probes = last_probe_count()
    off, lids, nb = O.partition(a)
    R = O.route_distances(Q, C)
    oi, od = O.ivf_search(X, off, lids, nb, Q, R, k, nprobe)
    assert (ids == oi).all(), (tag, np.argwhere(ids != oi)[:5])
    assert (dist.view(np.uint32) == od.view(np.uint32)).all(), tag
    old = _tune("ivf_route_cert", 0)
    try:
        ids0, dist0 = I.query_arrays(index, X.T, Q.T, k)
        probes0 = last_probe_count()
    finally:
        _tune("ivf_route_cert", old)
    assert (ids0 == ids).all() and (dist0.view(np.uint32) == dist.view(np.uint32)).all(), tag
    assert probes == probes0, (tag, probes, probes0)


@pytest.mark.parametrize("nprobe", [1, 8, 64, 128])
def test_certified_routing_equals_exact(nprobe):
    _banner(f"certified routing 40000x64 nlist 128 nprobe {nprobe}")
    X = datasets.zipf_anisotropic_mixture(40_000, 64, components=40, seed=4)
    C0 = X[np.random.default_rng(4).choice(40_000, 128, replace=False)]
    C, a, _ = O.lloyd(X, C0, 6)
    Q = datasets.mixture_queries(500, 64, components=40, seed=4, query_seed=17)
    _vs_oracle_and_fp32(X, C, a, Q, 10, nprobe, nprobe)


def test_duplicate_centroids_tie_to_the_lower_bucket():
    # centroids 3, 6 and 10 (1-based) coincide and some queries sit on them:
    # their routing distances tie exactly (0), and nprobe cuts through the tie
    _banner("certified routing with tied centroids")
    rng = np.random.default_rng(31)
    d = 32
    C = rng.standard_normal((12, d)).astype(np.float32)
    C[5] = C[2]
    C[9] = C[2]
    X = rng.standard_normal((6000, d)).astype(np.float32)
    a = rng.integers(1, 13, 6000).astype(np.int64)   # every bucket holds rows
    Q = np.concatenate([np.repeat(C[2:3], 4, 0), rng.standard_normal((60, d))]).astype(np.float32)
    for nprobe in (1, 2, 3, 5):
        _vs_oracle_and_fp32(X, C, a, Q, 10, nprobe, ("ties", nprobe))


def test_far_offset_data_widens_the_band():
    # every vector shifted by 40 in each coordinate: the FP32 norms form loses
    # accuracy to cancellation in proportion to the uncentred norms, so the
    # bound must admit far more centroids; the probes stay the exact ones
    _banner("certified routing, data far from the origin")
    rng = np.random.default_rng(32)
    d = 48
    centres = 3.0 * rng.standard_normal((64, d))
    lab = rng.integers(0, 64, 20_000)
    X = (centres[lab] + rng.standard_normal((20_000, d)) + 40.0).astype(np.float32)
    C = (centres + 40.0).astype(np.float32)
    a, _ = O.kmeans_assign(X, C)
    Q = (centres[rng.integers(0, 64, 300)] + rng.standard_normal((300, d)) + 40.0)
    Q = Q.astype(np.float32)
    for nprobe in (2, 9):
        _vs_oracle_and_fp32(X, C, a, Q, 10, nprobe, ("offset", nprobe))


def test_tune_keys_checked():
    L = lib()
    assert L.mann_tune_set(b"ivf_route_cert", 2) != 0
    assert L.mann_tune_set(b"ivf_route_cfg", -1) != 0
