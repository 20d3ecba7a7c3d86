"""CPU pins of the oracle's IVF-HNSW restatement (oracle/mann_oracle_hnsw.c).

Reference: src/indices/hnsw/{query.jl, neighbor_policy.jl, traversal.jl,
candidate_heap.jl, planner.jl}, src/indices/multilevel/ivf_hnsw.jl:46-88, and
the reference's own tests test/unit/indices/hnsw/{hnsw_tests.jl,
heap_tests.jl, batch_query_tests.jl}.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def one_list(X, M, ef_c, level=None, policy=O.HNSW_DIVERSIFIED):
    n = X.shape[0]
    off = np.array([0, n], np.int64)
    ids = np.arange(1, n + 1, dtype=np.int64)
    lev = np.zeros(n, np.int32) if level is None else np.asarray(level, np.int32)
    return off, ids, O.ivf_hnsw_build(X, off, ids, 1, M, ef_c, policy, level=lev)


def lists0(g, layer=0):
    return [list(g.nbr[layer, i, :g.deg[layer, i]]) for i in range(g.deg.shape[1])]


def test_known_graph_on_a_line_diversified():
    # points 0, 1, 3, 6, 10 on a line, M = 2, one layer; traced by hand from
    # insert! / _search_layer / select_neighbors / _prune_slot! (query.jl:264-603):
    #  - node 3 (x=6) gets [2, 1]: 1 and 0 are dominated by 2, the fallback
    #    (keepPrunedConnections) refills with 1; node 2's list [1, 0, 3]
    #    prunes to [1, 3] (0 is dominated by 1, 3 is not: d(3,1)=5 >= 3);
    #    node 1's [0, 2, 3] prunes to [0, 2]
    #  - node 4 (x=10) gets [3, 2]; node 3's [2, 1, 4] prunes to [2, 4], node
    #    2's [1, 3, 4] to [1, 3]
    X = np.array([[0], [1], [3], [6], [10]], np.float32)
    _, _, g = one_list(X, M=2, ef_c=10)
    assert lists0(g) == [[1, 2], [0, 2], [1, 3], [2, 4], [3, 2]]
    assert g.entry[0] == 0 and g.max_layer[0] == 0


def test_known_graph_on_a_line_heuristic():
    # the same points with the heuristic policy (the M nearest,
    # neighbor_policy.jl:22-36, _select_into! query.jl:531-548), traced by hand:
    #  - node 2 (x=3) gets [1, 0]; node 3 (x=6) gets [2, 1] (results 2, 1, 0 at
    #    d = 3, 5, 6); node 2's list [1, 0, 3] prunes to the 2 nearest of x=3:
    #    1 (d=2), then 0 and 3 tie at d=3 and the stable restatement keeps 0
    #    (input order; the reference's partialsort! leaves this tie unpinned);
    #    node 1's [0, 2, 3] prunes to [0, 2]
    #  - node 3 keeps no incoming edge, so node 4 (x=10) finds only 2, 1, 0 and
    #    gets [2, 1]; nodes 2 and 1 prune it away again
    X = np.array([[0], [1], [3], [6], [10]], np.float32)
    _, _, g = one_list(X, M=2, ef_c=10, policy=O.HNSW_HEURISTIC)
    assert lists0(g) == [[1, 2], [0, 2], [1, 0], [2, 1], [2, 1]]


def test_dominance_is_strict():
    # q = (0,0) inserted last; candidates by distance: b=(1,0) (1), c=(0.5,1)
    # (sqrt 1.25), e=(-3,0) (3).  c is as far from b as from q: with the strict
    # test (d_bc < cand.dist, neighbor_policy.jl:82) c is kept -> [b, c]; a <=
    # test would drop it and take e -> [b, e].
    X = np.array([[1, 0], [0.5, 1], [-3, 0], [0, 0]], np.float32)
    _, _, g = one_list(X, M=2, ef_c=10)
    assert lists0(g)[3] == [0, 1]


def test_levels_follow_the_exponential_law():
    # level >= l with probability M^-l for lambda = 1/log(M) (planner.jl:19-30)
    lev = O.hnsw_levels(200_000, 16, 12345)
    for l, tol in ((1, 0.004), (2, 0.0006)):
        frac = float((lev >= l).mean())
        assert abs(frac - 16.0 ** -l) < tol, (l, frac)
    # the stream: level = floor(-log(u) / log(M)) on splitmix64 uniforms
    u = O.uniform_stream(7, 64)
    want = np.floor(-np.log(np.maximum(u, np.finfo(np.float64).eps)) / math.log(4.0))
    assert np.array_equal(O.hnsw_levels(64, 4, 7), want.astype(np.int32))
    # M = 1 uses max(M, 2)
    assert np.array_equal(O.hnsw_levels(64, 1, 7), O.hnsw_levels(64, 2, 7))


def graph_invariants(g, off, nb):
    layers, n, cap = g.nbr.shape
    M = cap - 1
    for b in range(nb):
        lo, hi = off[b], off[b + 1]
        if hi == lo:
            assert g.entry[b] == -1
            continue
        lev = g.level[lo:hi]
        assert g.max_layer[b] == lev.max()
        assert lev[g.entry[b]] == lev.max()
        for layer in range(layers):
            for i in range(hi - lo):
                dg = g.deg[layer, lo + i]
                assert 0 <= dg <= M
                if lev[i] < layer:
                    assert dg == 0
                ids = g.nbr[layer, lo + i, :dg]
                assert len(set(ids.tolist())) == dg
                assert i not in ids
                assert np.all((ids >= 0) & (ids < hi - lo))
                assert np.all(lev[ids] >= layer)


def test_invariants_multilayer():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3000, 12)).astype(np.float32)
    assign = rng.integers(1, 7, 3000)
    assign[assign == 4] = 5            # one empty bucket
    off, ids, nb = O.partition(assign)
    g = O.ivf_hnsw_build(X, off, ids, nb, 6, 30, seed=11)
    assert g.nbr.shape[0] >= 3         # levels reach layer 2 with M = 6
    graph_invariants(g, off, nb)


def connected(g, lo, n):
    seen = {0}
    todo = [0]
    while todo:
        v = todo.pop()
        for w in g.nbr[0, lo + v, :g.deg[0, lo + v]]:
            if int(w) not in seen:
                seen.add(int(w))
                todo.append(int(w))
    return len(seen) == n


def test_large_ef_search_is_exact():
    # with ef >= |list| on a connected single-layer graph the best-first search
    # reaches every node from the entry point, so the result is the list's exact
    # top-k: the IVF-HNSW search then equals the IVF brute force with
    # DisjointMerge, ids and distance bits
    rng = np.random.default_rng(5)
    X = rng.standard_normal((600, 16)).astype(np.float32)
    assign = rng.integers(1, 5, 600)
    off, ids, nb = O.partition(assign)
    g = O.ivf_hnsw_build(X, off, ids, nb, 12, 40, level=np.zeros(600, np.int32))
    for b in range(nb):
        assert connected(g, off[b], off[b + 1] - off[b])
    Q = rng.standard_normal((40, 16)).astype(np.float32)
    C = np.stack([X[ids[off[b]:off[b + 1]] - 1].mean(0) for b in range(nb)])
    R = O.route_distances(Q, C)
    hi, hd = O.ivf_hnsw_search(X, off, ids, nb, g, Q, R, 7, 2, 1000, merge="disjoint")
    bi, bd = O.ivf_search(X, off, ids, nb, Q, R, 7, 2, merge="disjoint")
    assert np.array_equal(hi, bi)
    assert np.array_equal(hd.view(np.int32), bd.view(np.int32))


def test_entry_point_is_marked_before_the_descent():
    # _query_search (query.jl:95-123) marks the entry point visited, descends the
    # upper layers, then searches the base layer from the node reached: when the
    # descent leaves the entry point, the entry point is never scored.  With ef
    # >= |list| on a connected base layer the result is therefore the exact top-k
    # of the list without its entry point, or of the whole list when the entry
    # point is where the descent ended -- and both cases occur
    rng = np.random.default_rng(5)
    n, d, k = 300, 16, 10
    X = rng.standard_normal((n, d)).astype(np.float32)
    off = np.array([0, n], np.int64)
    ids = np.arange(1, n + 1, dtype=np.int64)
    g = O.ivf_hnsw_build(X, off, ids, 1, 12, 40, seed=2)
    assert g.max_layer[0] >= 1 and connected(g, 0, n)
    ep = int(g.entry[0]) + 1
    Q = rng.standard_normal((200, d)).astype(np.float32)
    hi, hd = O.ivf_hnsw_search(X, off, ids, 1, g, Q, np.zeros((200, 1), np.float32), k, 1,
                               1000)
    bi, _ = O.bf_knn(X, Q, k + 1)
    cases = set()
    for q in range(200):
        full = list(bi[q, :k])
        without = [i for i in bi[q] if i != ep][:k]
        if list(hi[q]) == full and ep in full:
            cases.add("entry found")
        else:
            assert list(hi[q]) == without
            if ep in full:
                cases.add("entry skipped")
    assert cases == {"entry found", "entry skipped"}


def test_reference_recall_threshold():
    # hnsw_tests.jl:5-27: M = 8, ef_construction = 80, ef_search = 80 on a
    # 4 x 60 randn set; average recall over k = 1..5 at least 0.6
    rng = np.random.default_rng(2025)
    X = rng.standard_normal((60, 4)).astype(np.float32)
    off = np.array([0, 60], np.int64)
    ids = np.arange(1, 61, dtype=np.int64)
    g = O.ivf_hnsw_build(X, off, ids, 1, 8, 80, seed=2025)
    R = np.zeros((10, 1), np.float32)
    for k in range(1, 6):
        Q = rng.standard_normal((10, 4)).astype(np.float32)
        hi, _ = O.ivf_hnsw_search(X, off, ids, 1, g, Q, R, k, 1, 80)
        ti, _ = O.bf_knn(X, Q, k)
        rec = np.mean([len(set(a) & set(b)) / k for a, b in zip(hi, ti)])
        assert rec >= 0.6


def test_batch_equals_single_and_merges():
    # batch_query_tests.jl:5-33 (batch == single); the single-query path differs
    # only in the order the per-list lists reach the merge (probe order), which
    # DisjointMerge's (dist, id) output does not see
    rng = np.random.default_rng(8)
    X = rng.standard_normal((800, 8)).astype(np.float32)
    assign = rng.integers(1, 9, 800)
    off, ids, nb = O.partition(assign)
    g = O.ivf_hnsw_build(X, off, ids, nb, 6, 24, seed=4)
    Q = rng.standard_normal((30, 8)).astype(np.float32)
    C = np.stack([X[ids[off[b]:off[b + 1]] - 1].mean(0) for b in range(nb)])
    R = O.route_distances(Q, C)
    a = O.ivf_hnsw_search(X, off, ids, nb, g, Q, R, 5, 3, 12, merge="disjoint")
    b = O.ivf_hnsw_search(X, off, ids, nb, g, Q, R, 5, 3, 12, merge="disjoint",
                          probe_order=True)
    assert np.array_equal(a[0], b[0])
    # k <= 0 and empty batches
    z = O.ivf_hnsw_search(X, off, ids, nb, g, Q[:0], R[:0], 5, 3, 12)
    assert z[0].shape == (0, 5)
    e = O.ivf_hnsw_search(X, off, ids, nb, g, Q, R, 0, 3, 12)
    assert e[0].shape == (30, 0)
