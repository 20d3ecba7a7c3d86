"""Merge orders of the GPU IVF search against the oracle (merge.jl:79-169,
query.jl:25-106 and :279-399, ivf_flat.jl:138-167).

The batched SimpleMerge order is (dist, bucket, id); the single-query path
concatenates the per-list results in probe order, so ties resolve by (dist,
probe rank, id); DisjointMerge pushes each list's top-k through a bounded max
heap and returns (dist, id).  The orders differ only where distances tie, so
every case here is built on duplicated points spread over several buckets, and
each test checks that the orders it compares really disagree somewhere."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import ivf as I
from manifoldann_amd import router as RT
from oracle import oracle as O


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _dup_data(seed, n_base=150, copies=4, d=12, nlist=8):
    """n_base points, each stored `copies` times under unrelated ids and
    assigned to random buckets (so equal points sit in different lists)."""
    rng = np.random.default_rng(seed)
    base = rng.standard_normal((n_base, d)).astype(np.float32)
    X = base[rng.permutation(np.repeat(np.arange(n_base), copies))]
    C = rng.standard_normal((nlist, d)).astype(np.float32)
    assign = rng.integers(1, nlist + 1, size=X.shape[0]).astype(np.int64)
    Q = (base[rng.integers(0, n_base, size=48)]
         + 0.05 * rng.standard_normal((48, d))).astype(np.float32)
    return X, C, assign, Q


def _index(X, C, assign, merge, nprobe):
    return I.build_index_from_assignment(np.asfortranarray(X.T), np.asfortranarray(C.T), assign,
                                         I.TopKRouting(nprobe), merge_strategy=merge)


def _oracle(X, C, assign, Q, k, nprobe, merge, probe_order):
    off, ids, nb = O.partition(assign)
    route = O.route_distances(Q, C)
    return O.ivf_search(X, off, ids, nb, Q, route, k, nprobe, merge=merge,
                        probe_order=probe_order)


def _same(ids, dist, oi, od, tag):
    assert ids.shape == oi.shape
    for q in range(oi.shape[0]):
        assert ids[q].tolist() == oi[q].tolist(), (tag, q)
        assert (dist[q].view(np.uint32) == od[q].view(np.uint32)).all(), (tag, q)


def test_cross_bucket_ties_follow_bucket_order():
    # Two copies of each point: the copy with the LOWER id goes to the HIGHER
    # bucket.  (dist, bucket, id) then prefers the higher id, (dist, id) the
    # lower one, so a merge by (dist, id) fails this test.
    _banner("cross-bucket duplicates, batched SimpleMerge")
    rng = np.random.default_rng(5)
    n_base, d = 200, 10
    base = rng.standard_normal((n_base, d)).astype(np.float32)
    X = np.concatenate([base, base])                 # ids i and i + n_base
    assign = np.concatenate([np.full(n_base, 2), np.full(n_base, 1)]).astype(np.int64)
    C = np.stack([base.mean(0) + 0.01, base.mean(0) - 0.01]).astype(np.float32)
    Q = (base[rng.integers(0, n_base, size=64)]
         + 0.02 * rng.standard_normal((64, d))).astype(np.float32)
    idx = _index(X, C, assign, I.SimpleMerge(), 2)
    for k in (1, 3, 8):
        ids, dist = I.query_arrays(idx, np.asfortranarray(X.T), np.asfortranarray(Q.T), k)
        oi, od = _oracle(X, C, assign, Q, k, 2, "simple", False)
        _same(ids, dist, oi, od, ("simple-batched", k))
        # the k-th place of an odd k splits a pair: the copy in bucket 1 (the
        # higher id) is the one kept
        if k % 2 == 1:
            assert (ids[:, k - 1] > n_base).all()


@pytest.mark.parametrize("k,nprobe", [(1, 2), (5, 3), (16, 4), (40, 8), (255, 8)])
def test_merge_orders_match_oracle(k, nprobe):
    _banner(f"merge orders, 600 duplicated points, k {k} nprobe {nprobe}")
    X, C, assign, Q = _dup_data(40 + k)
    data = np.asfortranarray(X.T)
    results = {}
    for merge_name, merge in (("simple", I.SimpleMerge()), ("disjoint", I.DisjointMerge())):
        idx = _index(X, C, assign, merge, nprobe)
        for probe_order in (False, True):
            ids, dist = I.query_arrays(idx, data, np.asfortranarray(Q.T), k,
                                       probe_order=probe_order)
            oi, od = _oracle(X, C, assign, Q, k, nprobe, merge_name, probe_order)
            _same(ids, dist, oi, od, (merge_name, probe_order, k, nprobe))
            results[(merge_name, probe_order)] = ids
        # the single-query entry (one vector) takes the probe-order path
        for q in range(0, 48, 7):
            got = M.query(idx, data, Q[q], k)
            oi, od = _oracle(X, C, assign, Q[q:q + 1], k, nprobe, merge_name, True)
            m = int((oi[0] > 0).sum())
            assert [n.id for n in got] == oi[0, :m].tolist(), (merge_name, q)
    if k > 1:
        # the four orders are not all the same answer on these inputs
        keys = list(results)
        assert any((results[a] != results[b]).any() for a in keys for b in keys if a < b)


def test_disjoint_merge_large_lists():
    # lists of several thousand rows: the replay's per-list top-k selection
    # runs its radix pass, and each tied query's distances span many blocks
    _banner("disjoint merge, 12,000 duplicated points in 4 lists")
    X, C, assign, Q = _dup_data(77, n_base=3000, copies=4, d=16, nlist=4)
    idx = _index(X, C, assign, I.DisjointMerge(), 3)
    data = np.asfortranarray(X.T)
    for k in (7, 64):
        for probe_order in (False, True):
            ids, dist = I.query_arrays(idx, data, np.asfortranarray(Q.T), k,
                                       probe_order=probe_order)
            oi, od = _oracle(X, C, assign, Q, k, 3, "disjoint", probe_order)
            _same(ids, dist, oi, od, ("disjoint-large", probe_order, k))


# ---- IVF-Flat --------------------------------------------------------------------------

def _flat_oracle_lists(index):
    lists = index.lists
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    ids = np.concatenate(lists).astype(np.int64)
    return off, ids


def _flat_index(X, C, assign, distance=None, centroid_metric=None, nprobe=4):
    t = I.KMeansTransform(k=C.shape[0])
    t.centroids = np.asfortranarray(C.T)
    t.centroid_norms = (C.astype(np.float64) ** 2).sum(1).astype(np.float32)
    return I.ivf_flat_from_assignment(np.asfortranarray(X.T), t, assign, nprobe,
                                      distance or M.Euclidean(), centroid_metric)


def _route_metric(Q, C, metric):
    return np.array([[O.metric(metric, c, q) for c in C] for q in Q], np.float32)


@pytest.mark.parametrize("pm,cm", [(O.SQEUCLIDEAN, O.EUCLIDEAN), (O.EUCLIDEAN, O.SQEUCLIDEAN),
                                   (O.COSINE, O.COSINE), (O.COSINE, O.EUCLIDEAN)])
def test_ivf_flat_metrics_with_ties(pm, cm):
    names = {O.EUCLIDEAN: M.Euclidean(), O.SQEUCLIDEAN: M.SqEuclidean(), O.COSINE: M.CosineDist()}
    _banner(f"ivf-flat point metric {pm} centroid metric {cm}")
    X, C, assign, Q = _dup_data(300 + 3 * pm + cm)
    idx = _flat_index(X, C, assign, names[pm], names[cm])
    # routing distances of the centroid metric (norms form for the Euclidean
    # pair, the per-pair CosineDist otherwise)
    if cm == O.COSINE:
        route = _route_metric(Q, C, O.COSINE)
    else:
        route = O.route_distances(Q, C)
        if cm == O.SQEUCLIDEAN:
            # the norms form without its square root
            qn = O.norms(Q)
            cn = O.norms(C)
            dots = np.array([[O.dot(c, q) for c in C] for q in Q], np.float32)
            sq = (cn[None, :] + qn[:, None]).astype(np.float32) - np.float32(2) * dots
            route = np.maximum(sq, np.float32(0)).astype(np.float32)
    got_route = I.route_distances(idx, np.asfortranarray(Q.T))
    assert (got_route.view(np.uint32) == route.view(np.uint32)).all()
    off, ids = _flat_oracle_lists(idx)
    data = np.asfortranarray(X.T)
    for k, nprobe in ((5, 3), (16, 5)):
        oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, k, nprobe, metric=pm)
        gi, gd = I.query_arrays(idx, data, np.asfortranarray(Q.T), k, nprobe=nprobe)
        for q in range(Q.shape[0]):
            assert gi[q, :cnt[q]].tolist() == oi[q, :cnt[q]].tolist(), (pm, cm, k, q)
            assert (gd[q, :cnt[q]].view(np.uint32) == od[q, :cnt[q]].view(np.uint32)).all()


def test_ivf_flat_large_lists_with_ties():
    _banner("ivf-flat 12,000 duplicated points in 4 lists")
    X, C, assign, Q = _dup_data(91, n_base=3000, copies=4, d=16, nlist=4)
    idx = _flat_index(X, C, assign)
    off, ids = _flat_oracle_lists(idx)
    route = O.route_distances(Q, C)
    for k in (9, 100):
        oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, k, 3)
        gi, gd = I.query_arrays(idx, np.asfortranarray(X.T), np.asfortranarray(Q.T), k, nprobe=3)
        for q in range(Q.shape[0]):
            assert gi[q, :cnt[q]].tolist() == oi[q, :cnt[q]].tolist(), (k, q)
            assert (gd[q, :cnt[q]].view(np.uint32) == od[q, :cnt[q]].view(np.uint32)).all()


def test_ivf_flat_routed_search_replays_the_probed_lists():
    # A learned router on a flat index: the heap replay walks exactly the lists
    # the scan probed (router order for fixed nprobe; the mass-selected lists,
    # empty ones skipped, for probability-mass probing).
    _banner("ivf-flat routed search with ties")
    X, C, assign, Q = _dup_data(123, n_base=400, copies=4, d=16, nlist=12)
    idx = _flat_index(X, C, assign)
    r = RT.LearnedProbeRouter(16, 12, hidden=32, seed=4)
    r.set_prior(idx, 0.5)
    off, ids = _flat_oracle_lists(idx)
    route = -r.scores(np.asfortranarray(Q.T))
    k = 6
    gi, gd = RT.query_arrays_routed(idx, r, np.asfortranarray(Q.T), k, 4)
    oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, k, 4)
    for q in range(Q.shape[0]):
        assert gi[q, :cnt[q]].tolist() == oi[q, :cnt[q]].tolist(), ("routed", q)
    gi, gd = RT.query_arrays_routed(idx, r, np.asfortranarray(Q.T), k, 8, mass=0.6)
    chosen = RT.probe_mass(idx, route, 8, 0.6)
    for q in range(Q.shape[0]):
        lists = [int(b) - 1 for b in chosen[q] if b > 0]
        rq = np.full((1, C.shape[0]), np.inf, np.float32)
        rq[0, lists] = np.arange(len(lists), dtype=np.float32)
        oi, od, cnt = O.ivf_flat_query(X, off, ids, Q[q:q + 1], rq, k, len(lists))
        assert gi[q, :cnt[0]].tolist() == oi[0, :cnt[0]].tolist(), ("mass", q)
        assert (gd[q, :cnt[0]].view(np.uint32) == od[0, :cnt[0]].view(np.uint32)).all()
