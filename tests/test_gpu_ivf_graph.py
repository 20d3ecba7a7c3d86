"""The batched MultiLevelIndex search captured in a HIP graph, and the IVF
autotuner.

A captured search is replayed several times with new query contents written
into its query buffer and other searches run eagerly between the replays;
every replay must equal the eager search of the same queries and the oracle
(query.jl:279-399, merge.jl:79-124).  Small cases send queries through the
refinement pass and, for 100 queries tied on identical rows, through more
exact-fallback work than the graph itself covers (the launcher's eager tail).
The autotuner's choice must leave the results exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import manifoldann_amd as M
from manifoldann_amd import datasets
from manifoldann_amd import ivf as I
from oracle import oracle as O


def _banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _index(n=40_000, d=64, nlist=128, seed=4, iters=6):
    X = datasets.zipf_anisotropic_mixture(n, d, components=40, seed=seed)
    t = I.KMeansTransform(k=nlist, max_iters=iters)
    I.fit(t, X.T, rng=seed)
    a = I.take_pending_assignments(t)
    C = np.ascontiguousarray(np.asarray(t.centroids).T)
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(8))
    return X, C, a, index


def _oracle(X, C, a, Q, k, nprobe, R=None):
    off, lids, nb = O.partition(a)
    R = O.route_distances(Q, C) if R is None else R
    return O.ivf_search(X, off, lids, nb, Q, R, k, nprobe)


def _same(ids, dist, oi, od, tag):
    ids = ids.cpu().numpy() if hasattr(ids, "cpu") else ids
    dist = dist.cpu().numpy() if hasattr(dist, "cpu") else dist
    assert (ids == oi).all(), (tag, np.argwhere(ids != oi)[:5])
    assert (dist.view(np.uint32) == od.view(np.uint32)).all(), tag


def test_ivf_graph_replays_match_eager_and_oracle():
    _banner("ivf graph replays 40000x64 nlist 128, 600 queries, k=10, nprobe 8")
    X, C, a, index = _index()
    rows = torch.empty((600, 64), dtype=torch.float32, device="cuda")
    rows.copy_(torch.from_numpy(datasets.mixture_queries(600, 64, components=40, seed=4,
                                                         query_seed=1)))
    g = I.IVFSearchGraph(index, rows.t(), 10, nprobe=8)
    other = torch.from_numpy(datasets.mixture_queries(200, 64, components=40, seed=4,
                                                      query_seed=77)).cuda()
    for rep in range(4):
        Qh = datasets.mixture_queries(600, 64, components=40, seed=4, query_seed=20 + rep)
        rows.copy_(torch.from_numpy(Qh))
        ids, dist = g.launch()
        gi, gd = ids.cpu().numpy().copy(), dist.cpu().numpy().copy()
        # eager work between replays on the same thread
        I.query_arrays(index, X.T, other.t(), 7, nprobe=5)
        torch.cuda.synchronize()
        ei, ed = I.query_arrays(index, X.T, Qh.T, 10, nprobe=8)
        _same(gi, gd, ei, ed, ("eager", rep))
        oi, od = _oracle(X, C, a, Qh, 10, 8)
        _same(gi, gd, oi, od, ("oracle", rep))


def test_ivf_graph_with_supplied_routes():
    _banner("ivf graph with device routing distances")
    X, C, a, index = _index(n=20_000, d=32, nlist=64, seed=6, iters=4)
    Qh = datasets.mixture_queries(200, 32, components=40, seed=6, query_seed=3)
    rows = torch.from_numpy(Qh).cuda()
    R = torch.from_numpy(np.random.default_rng(3).random((200, 64)).astype(np.float32)).cuda()
    g = I.IVFSearchGraph(index, rows.t(), 10, nprobe=6, route=R)
    for rep in range(2):
        Rh = np.random.default_rng(10 + rep).random((200, 64)).astype(np.float32)
        R.copy_(torch.from_numpy(Rh))
        ids, dist = g.launch()
        oi, od = _oracle(X, C, a, Qh, 10, 6, R=Rh)
        _same(ids, dist, oi, od, ("routes", rep))


def test_ivf_graph_refinement_pass():
    # the shell case of the eager refinement test: the first-list threshold
    # admits the whole shell, the refinement pass runs inside the graph
    _banner("ivf graph refinement pass")
    rng = np.random.default_rng(21)
    d = 16
    shell = rng.standard_normal((2000, d))
    shell /= np.linalg.norm(shell, axis=1, keepdims=True)
    near = 0.5 * np.eye(d)[0] + 0.01 * rng.standard_normal((10, d))
    far = 5.0 * np.eye(d)[0] + 0.3 * rng.standard_normal((490, d))
    X = np.concatenate([shell, near, far]).astype(np.float32)
    a = np.concatenate([np.ones(2000, np.int64), np.full(500, 2, np.int64)])
    C = np.stack([X[:2000].mean(0), X[2000:].mean(0)]).astype(np.float32)
    Qh = (0.01 * rng.standard_normal((8, d))).astype(np.float32)
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(2))
    rows = torch.from_numpy(Qh).cuda()
    g = I.IVFSearchGraph(index, rows.t(), 10)
    for rep in range(2):
        ids, dist = g.launch()
        st = M.last_stats()
        assert st[2] >= 1, st
        oi, od = _oracle(X, C, a, Qh, 10, 2)
        _same(ids, dist, oi, od, ("refine", rep))


def test_ivf_graph_fallback_beyond_captured_slots():
    # 3000 identical rows: every query ties and takes the exact fallback, 100
    # queries against the graph's 64 fallback slots
    _banner("ivf graph with 100 fallback queries")
    rng = np.random.default_rng(22)
    d = 24
    X = np.concatenate([np.ones((3000, d)), rng.standard_normal((1500, d)) + 4.0]).astype(np.float32)
    a = np.concatenate([np.ones(3000, np.int64), np.full(1500, 2, np.int64)])
    C = np.stack([X[:3000].mean(0), X[3000:].mean(0)]).astype(np.float32)
    Qh = np.ones((100, d), np.float32)
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(2))
    rows = torch.from_numpy(Qh).cuda()
    g = I.IVFSearchGraph(index, rows.t(), 10)
    oi, od = _oracle(X, C, a, Qh, 10, 2)
    for rep in range(2):
        ids, dist = g.launch()
        st = M.last_stats()
        assert st[3] == 100, st
        _same(ids, dist, oi, od, ("fallback", rep))


def test_ivf_graph_argument_checks():
    _banner("ivf graph argument checks")
    X, C, a, index = _index(n=5000, d=16, nlist=16, seed=8, iters=3)
    rows = torch.from_numpy(X[:50].copy()).cuda()
    with pytest.raises(M.ArgumentError):
        I.IVFSearchGraph(index, torch.from_numpy(X[:50].T.copy()).cuda(), 5)  # not rows.t()
    flat = I.ivf_flat_from_assignment(X.T, index.transform, a, nprobe=3)
    with pytest.raises(M.ArgumentError):
        I.IVFSearchGraph(flat, rows.t(), 5)
    R = torch.zeros((16, 50), dtype=torch.float32, device="cuda").t()  # 50 x 16, not contiguous
    with pytest.raises(M.ArgumentError):
        I.IVFSearchGraph(index, rows.t(), 5, route=R)
    index.merge_strategy = I.DisjointMerge()
    with pytest.raises(M.ArgumentError):
        I.IVFSearchGraph(index, rows.t(), 5)


def test_ivf_autotune_keeps_results_exact():
    _banner("ivf autotune 40000x64, 1000 queries, nprobe 8")
    X, C, a, index = _index()
    Qh = datasets.mixture_queries(1000, 64, components=40, seed=4, query_seed=5)
    try:
        ch = I.autotune(index, Qh.T, 10, nprobe=8)
        print("ivf choice", ch, flush=True)
        assert ch["trials"] >= 10 and ch["ms"] > 0
        ids, dist = I.query_arrays(index, X.T, Qh.T, 10, nprobe=8)
        oi, od = _oracle(X, C, a, Qh, 10, 8)
        _same(ids, dist, oi, od, "tuned")
    finally:
        M.lib().mann_autotune_clear()
