"""IVFFlatIndex on the GPU (ivf_flat.jl:30-168) against the heap-merge oracle.

Without distance ties at the k-th place the reference's bounded heap keeps the
k nearest candidates of the probed lists, the GPU list scan's (dist, id) top-k;
a query with a tie there is replayed through the heap on the GPU.  Continuous
random data exercises the first case, duplicated points the second."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import ivf as I
from manifoldann_amd._lib import check, lib
from oracle import oracle as O


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _oracle_lists(index):
    lists = index.lists
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    ids = np.concatenate(lists).astype(np.int64)
    return off, ids


def _check_vs_oracle(index, Xd, Q, k, nprobe):
    """GPU query == oracle heap merge over the GPU's lists and routing."""
    X = np.ascontiguousarray(Xd.T)
    route = I.route_distances(index, Q.T)
    off, ids = _oracle_lists(index)
    oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, k, nprobe)
    got = M.query(index, Xd, np.asfortranarray(Q.T), k, nprobe=nprobe)
    assert len(got) == Q.shape[0]
    for q in range(Q.shape[0]):
        assert len(got[q]) == cnt[q]
        assert [n.id for n in got[q]] == oi[q, :cnt[q]].tolist(), (q, k, nprobe)
        assert np.array_equal(np.array([n.dist for n in got[q]], np.float32), od[q, :cnt[q]])
    return got


def test_ivf_flat_matches_heap_oracle_and_batch_equals_single():
    # abstract_batch_query_tests.jl:9-61 (IVFFlat, nlist = 8, nprobe = 4, k = 5)
    _banner("ivf-flat 1000x16 nlist 8 nprobe 4")
    rng = np.random.default_rng(0xBA7CE5)
    data = np.asfortranarray(rng.standard_normal((16, 1000)).astype(np.float32))
    queries = rng.standard_normal((80, 16)).astype(np.float32)
    idx = M.build_index(M.IVFFlatIndex, data, nlist=8, nprobe=4)
    assert idx.nlist == 8 and idx.default_nprobe == 4 and len(idx.lists) == 8
    all_ids = np.sort(np.concatenate(idx.lists))
    assert np.array_equal(all_ids, np.arange(1, 1001))
    assert all((np.diff(l) > 0).all() for l in idx.lists)
    batch = _check_vs_oracle(idx, data, queries, 5, 4)
    single = [M.query(idx, data, queries[i], 5) for i in range(80)]
    vv = M.query(idx, data, [queries[i] for i in range(80)], 5)
    for i in range(80):
        assert [n.id for n in single[i]] == [n.id for n in batch[i]]
        assert [n.id for n in vv[i]] == [n.id for n in batch[i]]


@pytest.mark.parametrize("k,nprobe", [(1, 1), (10, 3), (100, 2), (255, 1), (10, 0), (10, 99)])
def test_ivf_flat_probe_and_k_sweep(k, nprobe):
    _banner(f"ivf-flat 3000x24 nlist 16 k {k} nprobe {nprobe}")
    rng = np.random.default_rng(7)
    data = np.asfortranarray(rng.standard_normal((24, 3000)).astype(np.float32))
    Q = rng.standard_normal((64, 24)).astype(np.float32)
    idx = M.build_index(M.IVFFlatIndex, data, nlist=16, nprobe=10, rng=3)
    _check_vs_oracle(idx, data, Q, k, nprobe)


def test_exhaustive_ivf_flat_equals_brute_force():
    _banner("ivf-flat nprobe = nlist vs brute force")
    rng = np.random.default_rng(9)
    data = np.asfortranarray(rng.standard_normal((32, 2000)).astype(np.float32))
    Q = rng.standard_normal((50, 32)).astype(np.float32)
    idx = M.build_index(M.IVFFlatIndex, data, nlist=12)
    got = M.query(idx, data, np.asfortranarray(Q.T), 10, nprobe=12)
    bf = M.build_index(M.BruteForceIndex, data)
    want = M.query(bf, data, np.asfortranarray(Q.T), 10)
    for q in range(50):
        assert [n.id for n in got[q]] == [n.id for n in want[q]]
        assert [n.dist for n in got[q]] == [n.dist for n in want[q]]


def test_empty_trailing_lists_take_probe_slots():
    # lists 6..8 are empty: a query whose nearest centroids include them probes
    # fewer points, as ivf_flat.jl:150-166 does
    _banner("ivf-flat empty trailing lists")
    rng = np.random.default_rng(21)
    data = np.asfortranarray(rng.standard_normal((8, 600)).astype(np.float32))
    C = rng.standard_normal((8, 8)).astype(np.float32)           # d x nlist
    assign = rng.integers(1, 6, size=600)
    t = I.KMeansTransform(k=8)
    t.centroids = np.asfortranarray(C)
    t.centroid_norms = (C.astype(np.float64) ** 2).sum(0).astype(np.float32)
    idx = I.ivf_flat_from_assignment(data, t, assign, nprobe=3)
    assert [len(l) for l in idx.lists][5:] == [0, 0, 0]
    Q = (C.T[rng.integers(0, 8, size=40)] + 0.1 * rng.standard_normal((40, 8))).astype(np.float32)
    _check_vs_oracle(idx, data, Q, 7, 3)


def test_ivf_flat_dimension_checks():
    rng = np.random.default_rng(1)
    data = np.asfortranarray(rng.standard_normal((8, 300)).astype(np.float32))
    idx = M.build_index(M.IVFFlatIndex, data, nlist=4)
    with pytest.raises(M.DimensionMismatch):
        M.query(idx, data, np.zeros(7, np.float32), 3)
    with pytest.raises(M.DimensionMismatch):
        M.query(idx, np.zeros((7, 300), np.float32), np.zeros(8, np.float32), 3)
    bf = M.build_index(M.BruteForceIndex, data)
    with pytest.raises(M.ArgumentError):
        M.query(bf, data, np.zeros(8, np.float32), 3, nprobe=2)


def _heap_replays():
    v = ctypes.c_int64()
    check(lib().mann_tune_get(b"last_heap_replays", ctypes.byref(v)))
    return v.value


@pytest.mark.parametrize("k,nprobe", [(1, 2), (5, 3), (16, 4), (255, 8)])
def test_distance_ties_follow_the_heap(k, nprobe):
    # 150 distinct points, each stored 4 times under different ids: every
    # k-th place falls in a run of equal distances, where the bounded heap's
    # survivors depend on push order and heap layout (neighbor_heaps.jl:84-190)
    _banner(f"ivf-flat duplicated points k {k} nprobe {nprobe}")
    rng = np.random.default_rng(31 + k)
    base = rng.standard_normal((150, 12)).astype(np.float32)
    X = base[rng.permutation(np.repeat(np.arange(150), 4))]
    data = np.asfortranarray(X.T)
    C = rng.standard_normal((12, 8)).astype(np.float32)
    assign = rng.integers(1, 9, size=600)
    t = I.KMeansTransform(k=8)
    t.centroids = np.asfortranarray(C)
    t.centroid_norms = (C.astype(np.float64) ** 2).sum(0).astype(np.float32)
    idx = I.ivf_flat_from_assignment(data, t, assign)
    Q = (base[rng.integers(0, 150, size=48)]
         + 0.05 * rng.standard_normal((48, 12))).astype(np.float32)
    _check_vs_oracle(idx, data, Q, k, nprobe)
    reps = _heap_replays()
    print(f"heap replays {reps} of 48", flush=True)
    assert reps > 0
    if k > 1:
        # the heap's answer is not the (dist, id) one on some query
        route = I.route_distances(idx, Q.T)
        off, ids = _oracle_lists(idx)
        oi, _, cnt = O.ivf_flat_query(X, off, ids, Q, route, k, nprobe)
        differs = 0
        for q in range(48):
            probed = np.argsort(route[q], kind="stable")[:nprobe]
            cand = np.concatenate([ids[off[c]:off[c + 1]] for c in probed])
            d32 = np.array([O.metric(O.EUCLIDEAN, X[c - 1], Q[q]) for c in cand], np.float32)
            order = np.lexsort((cand, d32))
            differs += cand[order][:cnt[q]].tolist() != oi[q, :cnt[q]].tolist()
        assert differs > 0


def test_ivf_flat_k_limit():
    rng = np.random.default_rng(2)
    data = np.asfortranarray(rng.standard_normal((8, 600)).astype(np.float32))
    idx = M.build_index(M.IVFFlatIndex, data, nlist=2)
    # the merge holds the top-(k+1) in a 256-entry buffer: k > 255 is an ArgumentError
    with pytest.raises(M.ArgumentError):
        M.query(idx, data, np.zeros(8, np.float32), 256, nprobe=2)
