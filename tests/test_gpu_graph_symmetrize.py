"""Undirected kNN graph on the GPU (_symmetrize_adjacency, knn_graph.jl:196-210)
against the oracle's set union, on host and device arrays."""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd._lib import check, lib
from oracle import oracle as O


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _compare(adj):
    lists, mx = M.symmetrize_adjacency(adj)
    want, _ = O.symmetrize(adj)
    assert len(lists) == len(want)
    for i in range(len(want)):
        assert lists[i] == want[i].tolist(), i
    assert mx == max((len(w) for w in want), default=0)
    return lists


@pytest.mark.parametrize("n,k,seed", [(1, 1, 0), (7, 3, 1), (5000, 20, 2), (60000, 21, 3)])
def test_symmetrize_random_adjacency(n, k, seed):
    _banner(f"symmetrize random n {n} k {k}")
    rng = np.random.default_rng(seed)
    adj = rng.integers(1, n + 1, size=(n, k)).astype(np.int64)
    # a hub and repeated entries: long rows and duplicates to drop
    if n > 10:
        adj[: n // 2, 0] = 1
        adj[:, -1] = adj[:, 0]
    t0 = time.perf_counter()
    _compare(adj)
    print(f"n={n} k={k} {1e3 * (time.perf_counter() - t0):.1f} ms (incl. copies)", flush=True)


def test_symmetrize_collinear_known_answer():
    # graph_tests.jl:4-48: directed lists of reshape(1:12, 3, 4) with k = 2
    adj = np.array([[2, 3], [1, 3], [2, 4], [3, 2]], np.int64)
    lists = _compare(adj)
    assert lists == [[2, 3], [1, 3, 4], [1, 2, 4], [2, 3]]


def test_symmetrize_device_arrays():
    import torch
    _banner("symmetrize device arrays")
    n, k = 3000, 12
    rng = np.random.default_rng(4)
    adj = rng.integers(1, n + 1, size=(n, k)).astype(np.int64)
    dadj = torch.from_numpy(adj).cuda()
    doff = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    dids = torch.zeros(2 * n * k, dtype=torch.int64, device="cuda")
    tot, mx = ctypes.c_int64(), ctypes.c_int64()
    torch.cuda.synchronize()
    check(lib().mann_knn_graph_symmetrize(dadj.data_ptr(), n, k, doff.data_ptr(), dids.data_ptr(),
                                          ctypes.byref(tot), ctypes.byref(mx)))
    off, ids = doff.cpu().numpy(), dids.cpu().numpy()
    want, _ = O.symmetrize(adj)
    assert tot.value == off[n] == sum(len(w) for w in want)
    for i in range(n):
        assert ids[off[i]:off[i + 1]].tolist() == want[i].tolist()
    assert mx.value == max(len(w) for w in want)


def test_symmetrize_knn_graph_and_errors():
    _banner("undirected kNN graph 4000x32 k 10")
    rng = np.random.default_rng(5)
    data = np.asfortranarray(rng.standard_normal((32, 4000)).astype(np.float32))
    idx = M.build_index(M.BruteForceIndex, data)
    adj, _ = M.build_knn_graph(idx, data, k=10, return_arrays=True)
    g = M.build_knn_graph(idx, data, k=10, directed=False)
    want, _ = O.symmetrize(adj)
    assert [list(r) for r in g.neighbors] == [w.tolist() for w in want]
    assert g.k == max(len(w) for w in want)
    with pytest.raises(M.ArgumentError):
        M.symmetrize_adjacency(np.array([[2], [3], [4]], np.int64))
    with pytest.raises(M.ArgumentError):
        M.symmetrize_adjacency(np.array([[0], [1]], np.int64))
    lists, mx = M.symmetrize_adjacency(np.zeros((5, 0), np.int64))
    assert lists == [[]] * 5 and mx == 0
