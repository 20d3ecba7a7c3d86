"""IVF-HNSW on the GPU against the oracle (multilevel/ivf_hnsw.jl:46-88 with the
HNSW terminals of hnsw/query.jl): the graphs of every list (levels, entry
points, every live adjacency entry of every layer) and the searches (ids and
distance bits) for both merges, both query forms, both neighbour policies, on
tie-heavy data and on a list long enough for the global visited bitmap."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as MA
from manifoldann_amd import ivf as I
from oracle import oracle as O


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _data(seed, n, d, nlist, copies=1, nq=64):
    rng = np.random.default_rng(seed)
    base = rng.standard_normal((n // copies, d)).astype(np.float32)
    X = base[rng.permutation(np.repeat(np.arange(n // copies), copies))] if copies > 1 else base
    C = rng.standard_normal((nlist, d)).astype(np.float32)
    assign = rng.integers(1, nlist + 1, size=X.shape[0]).astype(np.int64)
    Q = (X[rng.integers(0, X.shape[0], size=nq)]
         + 0.1 * rng.standard_normal((nq, d))).astype(np.float32)
    return X, C, assign, Q


def _index(X, C, assign, nprobe, term, seed, merge):
    return I.ivf_hnsw_from_assignment(np.asfortranarray(X.T), np.asfortranarray(C.T), assign,
                                      I.TopKRouting(nprobe), term, seed=seed,
                                      merge_strategy=merge)


def _check_graph(idx, X, assign, term, seed):
    nbr, deg, entry, maxl, level = I.hnsw_graph(idx)
    off, ids, nb = O.partition(assign)
    g = O.ivf_hnsw_build(X, off, ids, nb, term.M, term.ef_construction, term.policy_code,
                         seed=seed)
    assert np.array_equal(level, g.level)
    assert np.array_equal(entry, g.entry) and np.array_equal(maxl, g.max_layer)
    assert deg.shape == g.deg.shape
    assert np.array_equal(deg, g.deg)
    L, n, _ = nbr.shape
    for layer in range(L):
        for r in range(n):
            dg = deg[layer, r]
            assert np.array_equal(nbr[layer, r, :dg], g.nbr[layer, r, :dg]), (layer, r)
    return off, ids, nb, g


def _same(ids, dist, oi, od, tag):
    assert ids.shape == oi.shape, tag
    for q in range(oi.shape[0]):
        assert ids[q].tolist() == oi[q].tolist(), (tag, q, ids[q], oi[q])
        assert (dist[q].view(np.uint32) == od[q].view(np.uint32)).all(), (tag, q)


@pytest.mark.parametrize("policy,M,efc", [("diversified", 6, 24), ("heuristic", 6, 24),
                                          ("diversified", 16, 64)])
def test_graph_and_search_match_oracle(policy, M, efc):
    _banner(f"IVF-HNSW 3000 x 24, 8 lists, policy={policy}, M={M}, ef_c={efc}")
    X, C, assign, Q = _data(11, 3000, 24, 8)
    term = I.HnswTerminal(M=M, ef_construction=efc, ef_search=16, neighbor_policy=policy)
    seed = 99
    idx = _index(X, C, assign, 3, term, seed, I.DisjointMerge())
    off, ids, nb, g = _check_graph(idx, X, assign, term, seed)
    route = O.route_distances(Q, C)
    data = np.asfortranarray(X.T)
    for merge, mname in ((I.DisjointMerge(), "disjoint"), (I.SimpleMerge(), "simple")):
        idx.merge_strategy = merge
        for po in (False, True):
            for k, ef in ((1, 16), (10, 16), (10, 40), (40, 8)):
                ids_g, dist_g = I.query_arrays(idx, data, np.asfortranarray(Q.T), k,
                                               probe_order=po, ef_search=ef)
                oi, od = O.ivf_hnsw_search(X, off, ids, nb, g, Q, route, k, 3, ef, merge=mname,
                                           probe_order=po)
                _same(ids_g, dist_g, oi, od, (mname, po, k, ef))


def test_ties_on_duplicated_points():
    # every point stored 4 times under unrelated ids: equal distances inside the
    # heaps, the stable sorts, the dominance test (d_bc = 0 < d) and the merges
    _banner("IVF-HNSW on duplicated points, 2400 x 8, 6 lists")
    X, C, assign, Q = _data(3, 2400, 8, 6, copies=4)
    term = I.HnswTerminal(M=8, ef_construction=32, ef_search=24)
    idx = _index(X, C, assign, 4, term, 5, I.DisjointMerge())
    off, ids, nb, g = _check_graph(idx, X, assign, term, 5)
    route = O.route_distances(Q, C)
    data = np.asfortranarray(X.T)
    for merge, mname in ((I.DisjointMerge(), "disjoint"), (I.SimpleMerge(), "simple")):
        idx.merge_strategy = merge
        for po in (False, True):
            for k in (3, 12):
                ids_g, dist_g = I.query_arrays(idx, data, np.asfortranarray(Q.T), k,
                                               probe_order=po)
                oi, od = O.ivf_hnsw_search(X, off, ids, nb, g, Q, route, k, 4, 24, merge=mname,
                                           probe_order=po)
                _same(ids_g, dist_g, oi, od, ("dup", mname, po, k))


def test_long_list_global_bitmap_and_short_lists():
    # one list of 36,000 rows (its visited bitmap outgrows LDS in the build and
    # the search), lists shorter than k, an empty bucket, nprobe above the lists
    _banner("IVF-HNSW with a 36,000-row list, 3 buckets (one empty), k = 30")
    rng = np.random.default_rng(21)
    n, d = 36_005, 16
    X = rng.standard_normal((n, d)).astype(np.float32)
    assign = np.full(n, 1, np.int64)
    assign[-5:] = 3                       # bucket 2 empty, bucket 3 has 5 points
    C = np.stack([X[:-5].mean(0), X[:2].mean(0) + 50, X[-5:].mean(0)]).astype(np.float32)
    Q = (X[rng.integers(0, n, size=24)] + 0.05 * rng.standard_normal((24, d))).astype(np.float32)
    term = I.HnswTerminal(M=8, ef_construction=16, ef_search=48)
    idx = _index(X, C, assign, 5, term, 8, I.DisjointMerge())
    off, ids, nb, g = _check_graph(idx, X, assign, term, 8)
    route = O.route_distances(Q, C)
    data = np.asfortranarray(X.T)
    for k, ef in ((30, 48), (3, 400)):
        ids_g, dist_g = I.query_arrays(idx, data, np.asfortranarray(Q.T), k, ef_search=ef)
        oi, od = O.ivf_hnsw_search(X, off, ids, nb, g, Q, route, k, 5, ef, merge="disjoint")
        _same(ids_g, dist_g, oi, od, ("long", k, ef))


def test_build_ivf_hnsw_index_end_to_end():
    # ivf_smoke_tests.jl:108-133: the convenience builder, single / batch /
    # vector-of-vectors queries; here also against the oracle on the GPU's own
    # k-means partition, and the recall against brute force
    _banner("build_ivf_hnsw_index 6000 x 32, nlist 16")
    rng = np.random.default_rng(4)
    X = rng.standard_normal((6000, 32)).astype(np.float32)
    data = np.asfortranarray(X.T)
    idx = I.build_ivf_hnsw_index(data, 16, routing_k=4, hnsw_M=8, hnsw_ef_construction=50,
                                 hnsw_ef_search=24, rng=3)
    assert isinstance(idx, I.MultiLevelIndex) and isinstance(idx.merge_strategy, I.DisjointMerge)
    q = rng.standard_normal(32).astype(np.float32)
    single = MA.query(idx, data, q, 5)
    assert len(single) == 5
    Q = rng.standard_normal((200, 32)).astype(np.float32)
    batch = MA.query(idx, data, np.asfortranarray(Q.T), 10)
    assert len(batch) == 200 and all(len(r) == 10 for r in batch)
    vv = MA.query(idx, data, [Q[0], Q[1], Q[2]], 10)
    assert [r[0].id for r in vv] == [r[0].id for r in batch[:3]]
    # oracle on the same partition, levels and routing
    nbr, deg, entry, maxl, level = I.hnsw_graph(idx)
    ids_all = np.concatenate(idx.id_mappings)
    sizes = idx.list_sizes
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    g = O.ivf_hnsw_build(X, off, ids_all, idx.n_buckets, 8, 50, level=level)
    assert np.array_equal(deg, g.deg) and np.array_equal(entry, g.entry)
    C = np.ascontiguousarray(idx.transform.centroids.T)
    route = O.route_distances(Q, C)
    ids_g, dist_g = I.query_arrays(idx, data, np.asfortranarray(Q.T), 10)
    oi, od = O.ivf_hnsw_search(X, off, ids_all, idx.n_buckets, g, Q, route, 10, 4, 24,
                               merge="disjoint")
    _same(ids_g, dist_g, oi, od, "end-to-end")
    truth, _ = O.bf_knn(X, Q, 10)
    rec = O.recall_at_k(ids_g, truth)
    print(f"recall@10 = {rec:.3f}")
    assert rec > 0.5


def test_argument_errors():
    X, C, assign, Q = _data(1, 500, 8, 4)
    with pytest.raises(MA.ArgumentError):
        I.HnswTerminal.from_params(dict(M=0))
    with pytest.raises(MA.ArgumentError):
        I.HnswTerminal.from_params(dict(neighbor_policy="unknown"))
    term = I.HnswTerminal(M=64)           # above the GPU build's maximum
    with pytest.raises(MA.ArgumentError):
        _index(X, C, assign, 2, term, 0, I.DisjointMerge())
    idx = _index(X, C, assign, 2, I.HnswTerminal(M=4, ef_construction=8), 0, I.DisjointMerge())
    data = np.asfortranarray(X.T)
    with pytest.raises(MA.ArgumentError):
        I.query_arrays(idx, data, np.asfortranarray(Q.T), 10, ef_search=0)
    with pytest.raises(MA.DimensionMismatch):
        I.query_arrays(idx, data, np.zeros((7, 3), np.float32), 10)
    ids, dist = I.query_arrays(idx, data, np.asfortranarray(Q.T[:, :0]), 10)
    assert ids.shape == (0, 10)


def test_installed_graphs_search_like_the_oracle():
    # graphs built elsewhere (here: the oracle's CPU build, the role the
    # reference's own HNSWIndex children play behind the Julia shim) installed
    # with mann_ivf_hnsw_set_graph: the GPU search over them equals the
    # oracle's, with the levels given and inferred; the export returns them
    _banner("IVF-HNSW over installed oracle graphs, 4000 x 20, 10 lists")
    X, C, assign, Q = _data(17, 4000, 20, 10)
    off, ids, nb = O.partition(assign)
    g = O.ivf_hnsw_build(X, off, ids, nb, 10, 40, seed=123)
    route = O.route_distances(Q, C)
    data = np.asfortranarray(X.T)
    for with_level in (True, False):
        idx = I.build_index_from_assignment(data, np.asfortranarray(C.T), assign,
                                            I.TopKRouting(3), merge_strategy=I.DisjointMerge())
        I.set_hnsw_graph(idx, g.nbr, g.deg, g.entry, g.max_layer,
                         level=g.level if with_level else None, ef_search=20)
        nbr, deg, entry, maxl, level = I.hnsw_graph(idx)
        assert np.array_equal(deg, g.deg) and np.array_equal(entry, g.entry)
        for layer in range(deg.shape[0]):
            live = np.arange(g.nbr.shape[2])[None, :] < deg[layer][:, None]
            assert np.array_equal(np.where(live, nbr[layer], 0), np.where(live, g.nbr[layer], 0))
        for po in (False, True):
            ids_g, dist_g = I.query_arrays(idx, data, np.asfortranarray(Q.T), 10, probe_order=po)
            oi, od = O.ivf_hnsw_search(X, off, ids, nb, g, Q, route, 10, 3, 20, merge="disjoint",
                                       probe_order=po)
            _same(ids_g, dist_g, oi, od, ("installed", with_level, po))


def test_installed_graphs_are_validated():
    _banner("IVF-HNSW graph validation")
    X, C, assign, Q = _data(2, 600, 8, 3)
    off, ids, nb = O.partition(assign)
    g = O.ivf_hnsw_build(X, off, ids, nb, 4, 16, seed=1)
    data = np.asfortranarray(X.T)
    idx = I.build_index_from_assignment(data, np.asfortranarray(C.T), assign, I.TopKRouting(2))
    bad = g.nbr.copy()
    bad[0, 0, 0] = off[1] - off[0] + 5          # an id past its list
    with pytest.raises(MA.ArgumentError):
        I.set_hnsw_graph(idx, bad, g.deg, g.entry, g.max_layer, level=g.level)
    badd = g.deg.copy()
    badd[0, 3] = 9                                # a degree above M
    with pytest.raises(MA.ArgumentError):
        I.set_hnsw_graph(idx, g.nbr, badd, g.entry, g.max_layer, level=g.level)
    if g.nbr.shape[0] > 1:
        lv = g.level.copy()
        r = int(np.nonzero(g.deg[1])[0][0])
        lv[r] = 0                                 # a row with neighbours above its level
        with pytest.raises(MA.ArgumentError):
            I.set_hnsw_graph(idx, g.nbr, g.deg, g.entry, g.max_layer, level=lv)
    I.set_hnsw_graph(idx, g.nbr, g.deg, g.entry, g.max_layer, level=g.level)


def test_graphs_at_a_realistic_list_size():
    # 64 lists of about 1,000 rows in 64 dimensions with the builder's defaults
    # (M 16, ef_construction 200): every graph equal to the oracle's, and the
    # searches at ef 64
    _banner("IVF-HNSW 64,000 x 64, 64 lists, M 16, ef_c 200")
    X, C, assign, Q = _data(29, 64_000, 64, 64, nq=128)
    term = I.HnswTerminal(M=16, ef_construction=200, ef_search=64)
    idx = _index(X, C, assign, 5, term, 77, I.DisjointMerge())
    off, ids, nb, g = _check_graph(idx, X, assign, term, 77)
    route = O.route_distances(Q, C)
    ids_g, dist_g = I.query_arrays(idx, np.asfortranarray(X.T), np.asfortranarray(Q.T), 10)
    oi, od = O.ivf_hnsw_search(X, off, ids, nb, g, Q, route, 10, 5, 64, merge="disjoint")
    _same(ids_g, dist_g, oi, od, "realistic")
