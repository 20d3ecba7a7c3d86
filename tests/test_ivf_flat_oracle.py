"""CPU checks of the IVF-Flat query oracle (ivf_flat.jl:138-167 with the
BoundedMaxHeap of neighbor_heaps.jl:84-190) and of the host-side argument
checks of build_index(IVFFlatIndex)."""
import numpy as np
import pytest

import manifoldann_amd as M
from manifoldann_amd import ivf as I
from oracle import oracle as O


def _tie_case():
    # ids 1, 2, 3 lie at distance 1 from the origin, id 4 far away; lists
    # [3], [1, 2], [4] are probed in that order (routing 0 < 1 < 2)
    X = np.array([[1, 0], [0, 1], [-1, 0], [5, 5]], np.float32)
    off = np.array([0, 1, 3, 4])
    ids = np.array([3, 1, 2, 4])
    Q = np.zeros((1, 2), np.float32)
    route = np.array([[0.0, 1.0, 2.0]], np.float32)
    return X, off, ids, Q, route


def test_heap_keeps_first_pushed_on_equal_distances():
    # k=1: 3 enters the empty heap; 1 and 2 are not strictly closer, 4 is farther
    X, off, ids, Q, route = _tie_case()
    oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, 1, 3)
    assert oi.tolist() == [[3]] and od.tolist() == [[1.0]] and cnt.tolist() == [1]
    # k=2: 3 and 1 fill the heap, 2 ties the root and is rejected; the
    # survivors come out by (dist, id).  The (dist, id) top-2 would be [1, 2].
    oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, 2, 3)
    assert oi.tolist() == [[1, 3]] and cnt.tolist() == [2]


def test_heap_evicts_the_root_among_equal_distances():
    # k=2, pushes A(1), B(1), C(0.5): no sift-up between equal distances keeps
    # A at the root, so C evicts A and B survives
    X = np.array([[0, 1], [1, 0], [0.5, 0]], np.float32)     # ids 1 (A), 2 (B), 3 (C)
    off = np.array([0, 3])
    ids = np.array([1, 2, 3])
    Q = np.zeros((1, 2), np.float32)
    oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, np.zeros((1, 1), np.float32), 2, 1)
    assert oi.tolist() == [[3, 2]]
    assert od.tolist() == [[0.5, 1.0]]


def test_k_eff_truncates_to_the_probed_candidates():
    X, off, ids, Q, route = _tie_case()
    oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, 9, 2)
    assert cnt.tolist() == [3]
    assert oi[0, :3].tolist() == [1, 2, 3] and (oi[0, 3:] == 0).all()
    assert np.isinf(od[0, 3:]).all()
    # nprobe clamps to [1, nlist]: 0 probes the nearest list only, 99 all of them
    assert O.ivf_flat_query(X, off, ids, Q, route, 9, 0)[2].tolist() == [1]
    assert O.ivf_flat_query(X, off, ids, Q, route, 9, 99)[2].tolist() == [4]


def _random_lists(rng, n, nlist, empty=()):
    live = [c for c in range(1, nlist + 1) if c not in empty]
    assign = rng.choice(live, size=n)
    off = np.zeros(nlist + 1, np.int64)
    for c in range(1, nlist + 1):
        off[c] = off[c - 1] + int((assign == c).sum())
    ids = np.concatenate([np.flatnonzero(assign == c) + 1 for c in range(1, nlist + 1)])
    return off, ids


def test_exhaustive_probe_equals_brute_force():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((400, 12)).astype(np.float32)
    Q = rng.standard_normal((30, 12)).astype(np.float32)
    off, ids = _random_lists(rng, 400, 7)
    route = rng.random((30, 7)).astype(np.float32)
    oi, od, _ = O.ivf_flat_query(X, off, ids, Q, route, 10, 7)
    bi, bd = O.bf_knn(X, Q, 10)
    np.testing.assert_array_equal(oi, bi)
    np.testing.assert_array_equal(od, bd)


def test_union_of_probed_lists_in_fp64():
    # without distance ties the heap keeps the k nearest of the probed lists;
    # the probed lists are the nprobe smallest routing entries, lower index
    # first on ties, empty lists taking their slots
    rng = np.random.default_rng(11)
    X = rng.standard_normal((300, 8)).astype(np.float32)
    Q = rng.standard_normal((25, 8)).astype(np.float32)
    off, ids = _random_lists(rng, 300, 9, empty=(2, 9))
    route = rng.integers(0, 4, size=(25, 9)).astype(np.float32)    # many routing ties
    for k, nprobe in [(5, 3), (40, 2), (200, 4)]:
        oi, od, cnt = O.ivf_flat_query(X, off, ids, Q, route, k, nprobe)
        for q in range(25):
            probed = np.argsort(route[q], kind="stable")[:nprobe]
            cand = np.concatenate([ids[off[c]:off[c + 1]] for c in probed])
            d64 = np.sqrt(((X[cand - 1].astype(np.float64) - Q[q]) ** 2).sum(1))
            want = cand[np.lexsort((cand, d64))][:min(k, len(cand))]
            assert cnt[q] == len(want)
            assert oi[q, :cnt[q]].tolist() == want.tolist(), (k, nprobe, q)


def test_build_arguments_are_checked():
    X = np.zeros((4, 10), np.float32)
    with pytest.raises(M.ArgumentError):
        M.build_index(M.IVFFlatIndex, X, nlist=0)
    with pytest.raises(M.ArgumentError):
        M.build_index(M.IVFFlatIndex, np.zeros((0, 10), np.float32), nlist=2)
    with pytest.raises(M.ArgumentError):
        M.build_index(M.IVFFlatIndex, np.zeros((4, 0), np.float32), nlist=2)
    class Chebyshev:  # a metric the GPU path does not offer
        pass
    with pytest.raises(M.ArgumentError):
        M.build_index(M.IVFFlatIndex, X, nlist=2, centroid_metric=Chebyshev())
    with pytest.raises(M.ArgumentError):
        M.build_index(M.IVFFlatIndex, X, nlist=2, distance=Chebyshev())
    with pytest.raises(M.ArgumentError):
        M.build_index(M.IVFFlatIndex, X, nlist=2, merge_strategy=I.SimpleMerge())
