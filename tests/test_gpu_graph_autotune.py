"""The exact search captured in a HIP graph, and the autotuner.

A captured search is replayed several times with new query contents written
into its query buffer and other searches run eagerly between the replays;
every replay must equal the eager search of the same queries (and the
oracle).  A batch of identical points sends every query to the exact fallback,
more than the graph's own fallback covers, so the eager tail after the replay
is exercised too.  The autotuner's choice must leave the results exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import manifoldann_amd as M
from manifoldann_amd import datasets
from oracle import oracle as O


def _banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _same(ids, dist, oi, od, tag):
    ids = ids.cpu().numpy() if hasattr(ids, "cpu") else ids
    dist = dist.cpu().numpy() if hasattr(dist, "cpu") else dist
    assert ids.shape == oi.shape, tag
    assert (ids == oi).all(), (tag, np.argwhere(ids != oi)[:5])
    assert (dist.view(np.uint32) == od.view(np.uint32)).all(), tag


def test_graph_replays_match_eager_and_oracle():
    _banner("graph replays 20000x96, 500 queries, k=12")
    X = datasets.gaussian(20_000, 96, seed=21)
    index = M.build_index(M.BruteForceIndex, X.T)
    ds = M.Dataset(X.T)
    rows = torch.empty((500, 96), dtype=torch.float32, device="cuda")
    rows.copy_(torch.from_numpy(datasets.gaussian(500, 96, seed=22)))
    g = M.SearchGraph(index, ds, rows.t(), 12)
    other = torch.from_numpy(datasets.gaussian(300, 96, seed=99)).cuda()
    for rep in range(4):
        Qh = datasets.gaussian(500, 96, seed=30 + rep)
        rows.copy_(torch.from_numpy(Qh))
        ids, dist = g.launch()
        gi, gd = ids.cpu().numpy().copy(), dist.cpu().numpy().copy()
        # eager work between replays: another search on the same thread
        M.query_arrays(index, ds, other.t(), 7)
        torch.cuda.synchronize()
        ei, ed = M.query_arrays(index, ds, Qh.T, 12)
        _same(gi, gd, ei, ed, ("eager", rep))
        oi, od = O.bf_knn(X, Qh, 12)
        _same(gi, gd, oi, od, ("oracle", rep))


def test_graph_fallback_beyond_captured_slots():
    # identical rows: every query ties everywhere and takes the exact fallback
    _banner("graph with 100 fallback queries")
    Y = np.ones((3000, 32), np.float32)
    Y[2999] = 2.0
    index = M.build_index(M.BruteForceIndex, Y.T)
    ds = M.Dataset(Y.T)
    rows = torch.ones((100, 32), dtype=torch.float32, device="cuda")
    g = M.SearchGraph(index, ds, rows.t(), 5)
    for rep in range(2):
        ids, dist = g.launch()
        st = M.last_stats()
        assert st[3] == 100, st
        assert ids.cpu().numpy().tolist() == [[1, 2, 3, 4, 5]] * 100
        assert (dist.cpu().numpy() == 0).all()


def test_graph_cosine_and_argument_checks():
    _banner("graph cosine 8000x64 k=10")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((8000, 64)).astype(np.float32)
    index = M.build_index(M.BruteForceIndex, X.T, distance=M.CosineDist())
    ds = M.Dataset(X.T, M.CosineDist())
    Qh = rng.standard_normal((256, 64)).astype(np.float32)
    rows = torch.from_numpy(Qh).cuda()
    g = M.SearchGraph(index, ds, rows.t(), 10)
    ids, dist = g.launch()
    oi, od = O.bf_knn(X, Qh, 10, O.COSINE)
    _same(ids, dist, oi, od, "cosine")
    with pytest.raises(M.ArgumentError):
        M.SearchGraph(index, ds, torch.from_numpy(Qh.T.copy()).cuda(), 10)  # rows not contiguous
    with pytest.raises(M.ArgumentError):
        M.SearchGraph(index, ds, rows.t(), 300)


def test_autotune_keeps_results_exact():
    _banner("autotune 30000x128 queries and self-join")
    X = datasets.gaussian(30_000, 128, seed=41)
    Q = datasets.gaussian(2000, 128, seed=42)
    index = M.build_index(M.BruteForceIndex, X.T)
    ds = M.Dataset(X.T)
    try:
        ch = M.autotune(index, ds, Q.T, 10)
        print("query choice", ch, flush=True)
        assert ch["trials"] >= 10 and ch["ms"] > 0
        ids, dist = M.query_arrays(index, ds, Q.T, 10)
        oi, od = O.bf_knn(X, Q, 10)
        _same(ids, dist, oi, od, "tuned queries")
        Xs = X[:8000]
        dss = M.Dataset(Xs.T)
        idx_s = M.build_index(M.BruteForceIndex, Xs.T)
        chg = M.autotune(idx_s, dss, None, 15)
        print("graph choice", chg, flush=True)
        a, _ = M.build_knn_graph(idx_s, dss, k=15, return_arrays=True)
        assert (a == O.knn_graph(Xs, 15)).all()
    finally:
        M.lib().mann_autotune_clear()
