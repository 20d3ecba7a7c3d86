"""Writes tests/golden/golden.npz: small known-answer and seeded fixtures of the
search path, produced by the CPU oracle and cross-checked against an FP64
brute force (ids) when written.  The known answers restate the reference's
own tests (file:line in the keys' comments in test_golden.py)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def main():
    out = {}
    rng = np.random.default_rng(20260504)
    base = rng.standard_normal((200, 16)).astype(np.float32)
    Q = rng.standard_normal((20, 16)).astype(np.float32)
    ids, dist = O.bf_knn(base, Q, 5)
    ids64, d64 = O.fp64_knn(base, Q, 6)
    assert O.fp32_rounding_gap_ambiguous(d64, 5, 16) == 0 and (ids == ids64[:, :5]).all()
    out.update(bf_base=base, bf_q=Q, bf_ids=ids, bf_dist=dist)

    X = rng.standard_normal((1000, 16)).astype(np.float32)
    g = O.knn_graph(X, 6)
    lists, _ = O.symmetrize(g)
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])])
    out.update(graph_x=X, graph_adj=g, graph_sym_off=off, graph_sym_ids=np.concatenate(lists))

    C0 = X[:8].copy()
    C, a, it = O.lloyd(X, C0, 10)
    out.update(lloyd_c0=C0, lloyd_c=C, lloyd_assign=a, lloyd_iters=np.array([it]))

    off, lids, nb = O.partition(a)
    Qi = rng.standard_normal((25, 16)).astype(np.float32)
    R = O.route_distances(Qi, C)
    for npb in (1, 3, 8):
        oi, od = O.ivf_search(X, off, lids, nb, Qi, R, 7, npb)
        out[f"ivf_ids_np{npb}"] = oi
        out[f"ivf_dist_np{npb}"] = od
    out.update(ivf_q=Qi, ivf_route=R)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"))


if __name__ == "__main__":
    main()
