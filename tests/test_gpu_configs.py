"""Oracle parity at the sizes of the measured IVF workloads.

* IVF via KMeansTransform: 1,000,000 x 256 Zipf(1.1) mixture of 1,000
  anisotropic Gaussians, nlist = 1,024, 20 Lloyd iterations fitted on the GPU
  (kmeans++ seeding), k = 10.  The final assignment equals the oracle's E-step
  on the fitted centroids for every point, and the batched search returns the
  oracle's ids and distance bits at nprobe 1, 16 and 64 on 256 queries.
* Learned probe router on that index: MLP 256 -> 512 -> 1,024.  The gradient
  of one batch agrees with the FP64 oracle (bf16 operand rounding emulated),
  and the routed searches (fixed nprobe and probability mass) return the
  oracle's ids for the router's own routing distances.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import datasets
from manifoldann_amd import ivf as I
from manifoldann_amd import router as RT
from oracle import oracle as O
from oracle import router as RO

N, D, NLIST = 1_000_000, 256, 1024


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def big_ivf():
    _banner("build 1M x 256 IVF: data, GPU k-means (20 iterations), lists")
    X = datasets.zipf_anisotropic_mixture(N, D, components=1000, seed=3)
    t = I.KMeansTransform(k=NLIST, max_iters=20)
    cfg = I.TransformedConfig(t, I.TopKRouting(16), I.TerminalConfig(M.BruteForceIndex))
    index = I.build_index(M.MultiLevelIndex, X.T, cfg, rng=20260)
    C = np.ascontiguousarray(index.transform.centroids.T)        # (nlist, d)
    # the GPU's lists as an assignment: bucket of every point
    sizes = index.list_sizes
    ids = np.concatenate([np.asarray(m) for m in index.id_mappings])
    assign = np.zeros(N, np.int64)
    nz = np.flatnonzero(sizes > 0)
    assign[ids - 1] = np.repeat(nz + 1, sizes[nz])
    print(f"fit done: {index.transform.n_iters} iterations, {len(nz)} non-empty lists", flush=True)
    return X, C, assign, index


def test_kmeans_final_assignment_matches_oracle(big_ivf):
    # lloyd! ends with an E-step on the final centroids (lloyd.jl:130-131):
    # every point's bucket is the oracle's strict-'<' argmin over them
    _banner("final assignment of 1M points vs oracle E-step")
    X, C, assign, _ = big_ivf
    oa, _ = O.kmeans_assign(X, C)
    bad = np.flatnonzero(oa != assign)
    assert bad.size == 0, (bad.size, bad[:5])


@pytest.mark.parametrize("nprobe", [1, 16, 64])
def test_ivf_1m_search_matches_oracle(big_ivf, nprobe):
    _banner(f"1M x 256 IVF search, 256 queries, k 10, nprobe {nprobe}")
    X, C, assign, index = big_ivf
    Q = datasets.mixture_queries(256, D, components=1000, seed=3, query_seed=4242)
    ids, dist = I.query_arrays(index, X.T, Q.T, 10, nprobe=nprobe)
    off, lids, nb = O.partition(assign)
    R = O.route_distances(Q, C)
    oi, od = O.ivf_search(X, off, lids, nb, Q, R, 10, nprobe)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()
    st = M.last_stats()
    print(f"stats {st}", flush=True)


def test_router_256_512_1024_gradient_matches_oracle(big_ivf):
    _banner("router 256 -> 512 -> 1024 gradient vs FP64 oracle")
    X, C, assign, index = big_ivf
    d, h, c, B = D, 512, NLIST, 256
    Xb = datasets.mixture_queries(B, D, components=1000, seed=3, query_seed=515)
    labels = RT.neighbour_list_labels(index, X.T, Xb.T, k=10)
    r = RT.LearnedProbeRouter(d, c, hidden=h, seed=11)
    p = r.get_params()
    loss, g = r.grad(Xb.T, labels)
    oloss, og = RO.loss_and_grad(p.astype(np.float64), Xb, labels, d, h, c, bf16=True)
    assert abs(loss - oloss) <= 1e-4 * (1 + abs(oloss))
    err = np.abs(g - og).max() / (np.abs(og).max() + 1e-12)
    print(f"loss {loss:.6f} oracle {oloss:.6f} max rel grad err {err:.2e}", flush=True)
    assert err < 2e-3
    # per block (W1, b1, W2, b2), relative to the block's own scale
    sizes = [h * d, h, c * h, c]
    o = 0
    for s in sizes:
        e = np.abs(g[o:o + s] - og[o:o + s]).max() / (np.abs(og[o:o + s]).max() + 1e-12)
        assert e < 5e-3, (s, e)
        o += s


def test_router_routed_1m_search_matches_oracle(big_ivf):
    _banner("router-routed 1M search vs oracle (fixed nprobe and mass)")
    X, C, assign, index = big_ivf
    r = RT.LearnedProbeRouter(D, NLIST, hidden=512, seed=12)
    Xt = datasets.mixture_queries(4096, D, components=1000, seed=3, query_seed=616)
    labels = RT.neighbour_list_labels(index, X.T, Xt.T, k=10)
    r.set_prior(index, r.fit_prior(index, Xt.T, labels))
    r.train(Xt.T, labels, epochs=1, batch=512, lr=1e-3)
    Q = datasets.mixture_queries(256, D, components=1000, seed=3, query_seed=717)
    route = -r.scores(Q.T)
    off, lids, nb = O.partition(assign)
    ids, dist = RT.query_arrays_routed(index, r, Q.T, 10, 12)
    oi, od = O.ivf_search(X, off, lids, nb, Q, route, 10, 12)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()
    ids, dist = RT.query_arrays_routed(index, r, Q.T, 10, 32, mass=0.9)
    chosen = RT.probe_mass(index, route, 32, 0.9)
    for q in range(Q.shape[0]):
        lists = [int(b) - 1 for b in chosen[q] if b > 0]
        rq = np.full((1, NLIST), np.inf, np.float32)
        rq[0, lists] = np.arange(len(lists), dtype=np.float32)
        oi, od = O.ivf_search(X, off, lids, nb, Q[q:q + 1], rq, 10, len(lists))
        assert (ids[q] == oi[0]).all(), q
        assert (dist[q].view(np.uint32) == od[0].view(np.uint32)).all(), q
