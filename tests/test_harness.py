"""Host logic of the harness wrappers (no GPU): metric validation, names,
id conversion and the harness's recall (benchmarking/utils/evaluation.py:7-52)."""
import numpy as np
import pytest

from manifoldann_amd import harness as H


def test_metric_validation_and_names():
    with pytest.raises(ValueError):
        H.ManifoldANN_BruteForce("cosine")
    bf = H.ManifoldANN_BruteForce("euclidean")
    assert bf.get_name() == "ManifoldANN-BruteForce" and str(bf) == "ManifoldANN-BruteForce()"
    iv = H.ManifoldANN_IVFFlat("angular", nlist=50, nprobe=7)
    assert str(iv) == ("ManifoldANN-IVF-Flat(nlist=50, nprobe=7, "
                       "kmeans_init=kmeans_plus_plus, kmeans_max_iters=25)")
    iv.set_query_params(nprobe=12, unrelated=3)
    assert iv._nprobe == 12
    assert type(iv._get_distance_function()).__name__ == "CosineDist"
    assert type(bf._get_distance_function()).__name__ == "Euclidean"


def test_ids_to_lists_shifts_and_drops_padding():
    ids = np.array([[3, 1, 2], [5, 0, 0], [0, 0, 0]], np.int64)
    assert H.ids_to_lists(ids) == [[2, 0, 1], [4], []]


def test_recall_matches_the_harness_definition():
    truth = np.array([[0, 1, 2, 3], [4, 4, 5, 6]])
    # duplicates in the truth count once (set semantics), k defaults to len(pred)
    assert H.compute_recall([0, 1, 9, 8], truth[0]) == 0.5
    assert H.compute_recall([4, 5], truth[1], k=2) == 1.0
    assert H.compute_recall([1], np.array([]), k=1) == 0.0
    assert H.compute_recall_batch([], truth, 4) == 0.0
    assert H.compute_recall_batch([[0, 1, 2, 3], [7, 7, 7, 7]], truth, 4) == pytest.approx(0.5)


def test_ivf_hnsw_wrapper_names_and_params():
    h = H.ManifoldANN_IVFHNSW("euclidean", nlist=32, routing_k=4, M=12, ef_construction=80,
                              ef_search=40)
    assert H.ManifoldANN_IVFHNSW.get_name() == "ManifoldANN-IVF-HNSW"
    assert str(h) == ("ManifoldANN-IVF-HNSW(nlist=32, routing_k=4, M=12, ef_construction=80, "
                      "ef_search=40, neighbor_policy=diversified, kmeans_init=kmeans_plus_plus)")
    h.set_query_params(ef_search=90)
    assert h._ef_search == 90
