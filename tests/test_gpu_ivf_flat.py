"""IVFFlatIndex on the GPU (ivf_flat.jl:30-168) against the heap-merge oracle.

Without distance ties at the k-th place the reference's bounded heap keeps the
k nearest candidates of the probed lists, which is what the GPU list scan
returns (the lower id wins an exact tie there).  The inputs here are continuous
random data, on which exact FP32 ties at the k-th place do not occur; one would
show as an id mismatch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import ivf as I
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


@pytest.mark.parametrize("k,nprobe", [(1, 1), (10, 3), (100, 2), (256, 1), (10, 0), (10, 99)])
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
