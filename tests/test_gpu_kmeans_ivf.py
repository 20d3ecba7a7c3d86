"""GPU parity of the k-means transform and the IVF (MultiLevelIndex) path
against the CPU oracle: assignments, centroids, routing and neighbour ids
bit-identical."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import datasets
from manifoldann_amd import ivf as I
from manifoldann_amd._lib import check, lib
from oracle import oracle as O


def _banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _mixture(n, d, seed):
    X = datasets.zipf_anisotropic_mixture(n, d, components=40, seed=seed)
    return X


def test_kmeans_assign_bit_exact():
    _banner("kmeans assign 20000x32 k=64")
    X = _mixture(20_000, 32, 1)
    C = X[np.random.default_rng(1).choice(20_000, 64, replace=False)]
    a, md = I.kmeans_assign(X.T, C.T)
    oa, omd = O.kmeans_assign(X, C)
    assert (a == oa).all()
    assert (md.view(np.uint32) == omd.view(np.uint32)).all()


def test_lloyd_injected_init_bit_exact():
    _banner("lloyd 30000x48 k=96 10 iters")
    X = _mixture(30_000, 48, 2)
    C0 = X[np.random.default_rng(2).choice(30_000, 96, replace=False)]
    t = I.KMeansTransform(k=96, max_iters=10)
    t.init_centroids = C0.T
    I.fit(t, X.T)
    oc, oa, oit = O.lloyd(X, C0, 10)
    assert t.n_iters == oit
    assert (t.centroids.T.view(np.uint32) == oc.view(np.uint32)).all()
    assert (t.pending_assignments == oa).all()


def test_lloyd_empty_cluster_known_answer():
    # test_kmeans_rng_and_empty_clusters.jl:14-51
    _banner("lloyd empty clusters")
    X = np.array([[0, 1, 2, 3, 4], [0, 0, 0, 0, 0]], np.float32)
    t = I.KMeansTransform(k=4, max_iters=1)
    t.init_centroids = np.array([[0, 1e9, 1e9, 1e9], [0, 1e9, 1e9, 1e9]], np.float32)
    I.fit(t, X)
    assert t.centroids.T.tolist() == [[2, 0], [4, 0], [3, 0], [2, 0]]


def test_kmeanspp_matches_oracle():
    _banner("kmeans++ 8000x16 k=32")
    X = _mixture(8000, 16, 3)
    C, idx = I.kmeans_plus_plus(X.T, 32, seed=77)
    oc, oidx = O.kmeanspp(X, 32, seed=77)
    assert (idx == oidx).all()
    assert (C.T == oc).all()


def test_kmeans_assertions():
    X = np.random.default_rng(0).standard_normal((4, 10)).astype(np.float32)
    with pytest.raises(AssertionError):
        I.fit(I.KMeansTransform(k=11), X)
    with pytest.raises(AssertionError):
        I.KMeansTransform(k=0)


def _build_ivf(n=40_000, d=64, nlist=128, seed=4, iters=8):
    X = _mixture(n, d, seed)
    C0 = X[np.random.default_rng(seed).choice(n, nlist, replace=False)]
    oc, oa, _ = O.lloyd(X, C0, iters)
    index = I.build_index_from_assignment(X.T, oc.T, oa, I.TopKRouting(8))
    return X, oc, oa, index


def test_ivf_routing_and_probes_match_oracle():
    _banner("ivf routing / probes")
    X, C, a, index = _build_ivf()
    Q = datasets.mixture_queries(300, 64, components=40, seed=4, query_seed=9)
    R = I.route_distances(index, Q.T)
    oR = O.route_distances(Q, C)
    assert (R.view(np.uint32) == oR.view(np.uint32)).all()
    P = I.probe_lists(index, R)
    off, ids, nb = O.partition(a)
    for q in range(300):
        sel = O.select_topk(oR[q], nb, 8)
        sel = [int(b) for b in sel if off[b] > off[b - 1]]
        got = [int(b) for b in P[q] if b]
        assert got == sel


@pytest.mark.parametrize("nprobe", [1, 4, 16, 128])
def test_ivf_search_bit_exact(nprobe):
    _banner(f"ivf search nprobe={nprobe}")
    X, C, a, index = _build_ivf()
    Q = datasets.mixture_queries(400, 64, components=40, seed=4, query_seed=11)
    ids, dist = I.query_arrays(index, X.T, Q.T, 10, nprobe=nprobe)
    st = M.last_stats()
    off, lids, nb = O.partition(a)
    R = O.route_distances(Q, C)
    oi, od = O.ivf_search(X, off, lids, nb, Q, R, 10, nprobe)
    print(f"stats={st}", flush=True)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()
    if nprobe == 128:   # exhaustive routing == brute force (query_behavior_tests.jl:63-96)
        bi, bd = O.bf_knn(X, Q, 10)
        assert (ids == bi).all()


def test_ivf_empty_buckets_and_short_lists():
    # buckets with no points are routed to but dropped (routing_tests.jl:8-24);
    # probed lists holding fewer than k points give short result lists
    _banner("ivf empty buckets")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((300, 16)).astype(np.float32)
    C = rng.standard_normal((12, 16)).astype(np.float32)
    a, _ = O.kmeans_assign(X, C[:8])        # buckets 9..12 never used; max assignment <= 8
    a[a == 3] = 4                           # bucket 3 empty
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(3))
    Q = rng.standard_normal((50, 16)).astype(np.float32)
    ids, dist = I.query_arrays(index, X.T, Q.T, 7)
    off, lids, nb = O.partition(a)
    R = O.route_distances(Q, C)
    oi, od = O.ivf_search(X, off, lids, nb, Q, R, 7, 3)
    assert (ids == oi).all()
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()


def _ivf_vs_oracle(X, C, a, Q, k, nprobe):
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(nprobe))
    ids, dist = I.query_arrays(index, X.T, Q.T, k)
    st = M.last_stats()
    off, lids, nb = O.partition(a)
    R = O.route_distances(Q, C)
    oi, od = O.ivf_search(X, off, lids, nb, Q, R, k, nprobe)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()
    return st


def test_ivf_refinement_pass():
    # the nearest list is a shell of 2000 points at distance ~1 around the query,
    # the true neighbours sit in the second list at ~0.5: the row-sample
    # threshold admits the whole shell (overflow), the refinement pass re-scans
    # with the threshold of the candidates already held
    _banner("ivf refinement pass")
    rng = np.random.default_rng(21)
    d = 16
    shell = rng.standard_normal((2000, d))
    shell /= np.linalg.norm(shell, axis=1, keepdims=True)
    near = 0.5 * np.eye(d)[0] + 0.01 * rng.standard_normal((10, d))
    far = 5.0 * np.eye(d)[0] + 0.3 * rng.standard_normal((490, d))
    X = np.concatenate([shell, near, far]).astype(np.float32)
    a = np.concatenate([np.ones(2000, np.int64), np.full(500, 2, np.int64)])
    C = np.stack([X[:2000].mean(0), X[2000:].mean(0)]).astype(np.float32)
    Q = (0.01 * rng.standard_normal((8, d))).astype(np.float32)
    st = _ivf_vs_oracle(X, C, a, Q, 10, 2)
    print(f"stats={st}", flush=True)
    assert st[2] >= 1, st  # queries that needed the refinement pass
    assert st[1] >= 1, st  # first-pass overflow


def test_ivf_exact_fallback_on_ties():
    # 3000 identical rows in the probed list: every candidate ties, both passes
    # overflow and the segmented exact fallback must return the (dist, bucket,
    # id) order, i.e. the lowest ids
    _banner("ivf exact fallback on ties")
    rng = np.random.default_rng(22)
    d = 24
    X = np.concatenate([np.ones((3000, d)), rng.standard_normal((1500, d)) + 4.0]).astype(np.float32)
    a = np.concatenate([np.ones(3000, np.int64), np.full(1500, 2, np.int64)])
    C = np.stack([X[:3000].mean(0), X[3000:].mean(0)]).astype(np.float32)
    Q = np.ones((5, d), np.float32)
    st = _ivf_vs_oracle(X, C, a, Q, 10, 2)
    print(f"stats={st}", flush=True)
    assert st[3] == 5, st  # all five resolved by the exact fallback


def test_ivf_external_route_scores():
    _banner("ivf with supplied routing distances")
    X, C, a, index = _build_ivf(n=20_000, d=32, nlist=64, seed=6, iters=4)
    Q = datasets.mixture_queries(200, 32, components=40, seed=6, query_seed=3)
    R = np.random.default_rng(3).random((200, 64)).astype(np.float32)
    ids, dist = I.query_arrays(index, X.T, Q.T, 10, route=R, nprobe=6)
    off, lids, nb = O.partition(a)
    oi, od = O.ivf_search(X, off, lids, nb, Q, R, 10, 6)
    assert (ids == oi).all()


@pytest.mark.parametrize("settings", [
    {"ivf_cfg": 0},
    {"ivf_cfg": 8, "ivf_tau_cfg": 8},
    {"ivf_first_rows": 64, "ivf_rows_item": 256},
    {"ivf_tau_mode": 0},
    {"ivf_tau_grp": 0},
    {"ivf_first_rows": 300, "ivf_rows_item": 256},
])
def test_ivf_tuning_keeps_results_exact(settings):
    # tile configs, first-list row caps, item sizes and threshold modes change
    # only the speed: the ids and distance bits stay the oracle's
    _banner(f"ivf tuning {settings}")
    L = lib()
    old = {}
    for key, v in settings.items():
        o = ctypes.c_int64()
        check(L.mann_tune_get(key.encode(), ctypes.byref(o)))
        old[key] = o.value
        check(L.mann_tune_set(key.encode(), v))
    try:
        X, C, a, _ = _build_ivf()
        Q = datasets.mixture_queries(300, 64, components=40, seed=4, query_seed=13)
        st = _ivf_vs_oracle(X, C, a, Q, 10, 8)
        print(f"stats={st}", flush=True)
    finally:
        for key, v in old.items():
            check(L.mann_tune_set(key.encode(), v))
