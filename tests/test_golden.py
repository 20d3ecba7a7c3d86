"""The oracle reproduces the committed golden fixtures (tests/golden), and the
GPU path reproduces them too (gpu-marked)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


def test_oracle_bf_fixture():
    ids, dist = O.bf_knn(G["bf_base"], G["bf_q"], 5)
    assert (ids == G["bf_ids"]).all() and (dist == G["bf_dist"]).all()


def test_oracle_graph_fixture():
    g = O.knn_graph(G["graph_x"], 6)
    assert (g == G["graph_adj"]).all()
    lists, _ = O.symmetrize(g)
    assert (np.concatenate(lists) == G["graph_sym_ids"]).all()


def test_oracle_lloyd_and_ivf_fixture():
    C, a, it = O.lloyd(G["graph_x"], G["lloyd_c0"], 10)
    assert (C == G["lloyd_c"]).all() and (a == G["lloyd_assign"]).all()
    assert it == G["lloyd_iters"][0]
    off, lids, nb = O.partition(a)
    for npb in (1, 3, 8):
        oi, od = O.ivf_search(G["graph_x"], off, lids, nb, G["ivf_q"], G["ivf_route"], 7, npb)
        assert (oi == G[f"ivf_ids_np{npb}"]).all()


@pytest.mark.gpu
def test_gpu_reproduces_fixtures():
    import manifoldann_amd as M
    from manifoldann_amd import ivf as I
    base = G["bf_base"]
    idx = M.build_index(M.BruteForceIndex, base.T)
    ids, dist = M.query_arrays(idx, base.T, G["bf_q"].T, 5)
    assert (ids == G["bf_ids"]).all() and (dist == G["bf_dist"]).all()
    X = G["graph_x"]
    gi = M.build_index(M.BruteForceIndex, X.T)
    adj, _ = M.build_knn_graph(gi, X.T, k=6, return_arrays=True)
    assert (adj == G["graph_adj"]).all()
    t = I.KMeansTransform(k=8, max_iters=10)
    t.init_centroids = G["lloyd_c0"].T
    I.fit(t, X.T)
    assert (t.centroids.T == G["lloyd_c"]).all()
    index = I.build_index_from_assignment(X.T, t.centroids, t.pending_assignments, I.TopKRouting(3))
    for npb in (1, 3, 8):
        oi, _ = I.query_arrays(index, X.T, G["ivf_q"].T, 7, nprobe=npb)
        assert (oi == G[f"ivf_ids_np{npb}"]).all()
