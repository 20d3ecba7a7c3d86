"""The benchmark harness's wrapper contract (wrappers/base.py:8-200,
manifoldann.py:15-287, 740-824, driven as benchmark.py:741-828 drives it) over
the GPU library: prepare_data / fit / prepare_queries / query_batch_raw /
finalize_batch_ids return the oracle's neighbours as 0-based id lists, and the
single-query path agrees with the batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from manifoldann_amd import harness as H
from manifoldann_amd import ivf as I
from oracle import oracle as O


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _run_contract(algo, X, Q, k):
    """The harness's sequence (benchmark.py:741-828): untimed preparation,
    fit, a warm query, then the timed raw query and the untimed finalise."""
    prepared_train = algo.prepare_data(X)
    algo.fit(prepared_train)
    prepared_test = algo.prepare_queries(Q)
    algo.finalize_batch_ids(algo.query_batch_raw(prepared_test, k))
    raw = algo.query_batch_raw(prepared_test, k)
    return algo.finalize_batch_ids(raw)


@pytest.mark.parametrize("metric", ["euclidean", "angular"])
def test_bruteforce_wrapper_returns_oracle_ids(metric):
    _banner(f"harness BruteForce {metric}")
    rng = np.random.default_rng(11)
    X = rng.standard_normal((3000, 48)).astype(np.float32)
    Q = rng.standard_normal((200, 48)).astype(np.float32)
    algo = H.ManifoldANN_BruteForce(metric)
    preds = _run_contract(algo, X, Q, 10)
    code = O.COSINE if metric == "angular" else O.EUCLIDEAN
    oi, _ = O.bf_knn(X, Q, 10, code)
    truth = (oi - 1).tolist()
    assert preds == truth
    assert H.compute_recall_batch(preds, np.asarray(truth), 10) == 1.0
    for i in (0, 17, 199):
        assert algo.query(Q[i], 10) == truth[i]
    # a numpy batch converts inside query_batch_raw, as the reference's does
    assert algo.finalize_batch_ids(algo.query_batch_raw(Q[:5], 10)) == truth[:5]
    # k above n: a short list, no padding ids
    small = H.ManifoldANN_BruteForce(metric)
    got = _run_contract(small, X[:7], Q[:3], 10)
    assert all(len(r) == 7 for r in got)


@pytest.mark.parametrize("metric", ["euclidean", "angular"])
def test_ivf_flat_wrapper_matches_heap_oracle(metric):
    _banner(f"harness IVF-Flat {metric}")
    rng = np.random.default_rng(12)
    X = rng.standard_normal((4000, 32)).astype(np.float32)
    Q = rng.standard_normal((150, 32)).astype(np.float32)
    algo = H.ManifoldANN_IVFFlat(metric, nlist=16, nprobe=3)
    assert algo.get_name() == "ManifoldANN-IVF-Flat"
    code = O.COSINE if metric == "angular" else O.EUCLIDEAN
    for nprobe in (3, 7):
        if nprobe != 3:
            algo.set_query_params(nprobe=nprobe)
            preds = algo.finalize_batch_ids(algo.query_batch_raw(algo.prepare_queries(Q), 10))
        else:
            preds = _run_contract(algo, X, Q, 10)
        index = algo._index
        route = I.route_distances(index, Q.T)
        lists = index.lists
        off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
        ids = np.concatenate(lists).astype(np.int64)
        oi, _od, cnt = O.ivf_flat_query(X, off, ids, Q, route, 10, nprobe, metric=code)
        truth = [(oi[q, :cnt[q]] - 1).tolist() for q in range(Q.shape[0])]
        assert preds == truth, nprobe
        # the single-query path (probe order) gives the same ids here
        assert algo.query(Q[5], 10) == truth[5]


def test_ivf_hnsw_wrapper_matches_the_oracle():
    # manifoldann.py:649-737 through the harness contract: the ids equal the
    # oracle's IVF-HNSW search over the index's own lists, levels and routing
    _banner("harness IVF-HNSW euclidean")
    rng = np.random.default_rng(6)
    X = rng.standard_normal((4000, 32)).astype(np.float32)
    Q = rng.standard_normal((150, 32)).astype(np.float32)
    algo = H.ManifoldANN_IVFHNSW("euclidean", nlist=12, routing_k=3, M=8, ef_construction=40,
                                 ef_search=20)
    preds = _run_contract(algo, X, Q, 10)
    idx = algo._index
    nbr, deg, entry, maxl, level = I.hnsw_graph(idx)
    ids_all = np.concatenate(idx.id_mappings)
    off = np.concatenate([[0], np.cumsum(idx.list_sizes)]).astype(np.int64)
    g = O.ivf_hnsw_build(X, off, ids_all, idx.n_buckets, 8, 40, level=level)
    C = np.ascontiguousarray(idx.transform.centroids.T)
    oi, _ = O.ivf_hnsw_search(X, off, ids_all, idx.n_buckets, g, Q, O.route_distances(Q, C), 10,
                              3, 20, merge="disjoint")
    assert preds == [[int(i) - 1 for i in row if i > 0] for row in oi]
    # the single-query path agrees with the batch on untied rows
    assert algo.query(Q[0], 10) == preds[0]
