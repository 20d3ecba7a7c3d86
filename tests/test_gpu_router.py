"""GPU learned router: gradients against the FP64 oracle (bf16 operand rounding
emulated), training reduces the loss, and the routed IVF search returns the
oracle's ids for the router's own routing distances."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from manifoldann_amd import datasets
from manifoldann_amd import ivf as I
from manifoldann_amd import router as RT
from oracle import oracle as O
from oracle import router as RO


def test_router_gradient_matches_oracle():
    print("\n[gpu] router gradient", flush=True)
    rng = np.random.default_rng(0)
    d, h, c, B = 48, 64, 40, 96
    r = RT.LearnedProbeRouter(d, c, hidden=h, seed=3)
    p = r.get_params()
    X = rng.standard_normal((B, d)).astype(np.float32)
    labels = rng.integers(0, c + 1, size=(B, 4))
    labels[:, 0] = rng.integers(1, c + 1, size=B)
    loss, g = r.grad(X.T, labels)
    oloss, og = RO.loss_and_grad(p.astype(np.float64), X, labels, d, h, c, bf16=True)
    # tolerance: FP32 accumulation order and fast exp/log vs FP64 on identical
    # bf16-rounded operands
    assert abs(loss - oloss) <= 1e-4 * (1 + abs(oloss))
    err = np.abs(g - og).max() / (np.abs(og).max() + 1e-12)
    print(f"max rel grad err {err:.2e}", flush=True)
    assert err < 2e-3
    # and against the unrounded FP64 gradient: bf16-level agreement
    _, og64 = RO.loss_and_grad(p.astype(np.float64), X, labels, d, h, c, bf16=False)
    assert np.abs(g - og64).max() / np.abs(og64).max() < 5e-2


def test_router_scores_match_oracle():
    rng = np.random.default_rng(1)
    d, h, c = 32, 48, 24
    r = RT.LearnedProbeRouter(d, c, hidden=h, seed=5)
    X = rng.standard_normal((70, d)).astype(np.float32)
    s = r.scores(X.T)
    os_ = RO.scores(r.get_params().astype(np.float64), X, d, h, c)
    assert np.abs(s - os_).max() <= 1e-4 * (1 + np.abs(os_).max())


def test_router_training_learns_labels():
    print("\n[gpu] router training", flush=True)
    rng = np.random.default_rng(2)
    d, c, n = 32, 16, 8192
    centres = rng.standard_normal((c, d)).astype(np.float32) * 3
    lab = rng.integers(0, c, size=n)
    X = centres[lab] + rng.standard_normal((n, d)).astype(np.float32)
    labels = np.stack([lab + 1, np.zeros(n, np.int64)], 1)
    r = RT.LearnedProbeRouter(d, c, hidden=64, seed=7)
    l0, _ = r.grad(X[:1024].T, labels[:1024])
    loss = r.train(X.T, labels, epochs=4, batch=512, lr=3e-3)
    acc = (r.scores(X.T).argmax(1) == lab).mean()
    print(f"loss {l0:.3f} -> {loss:.3f}, acc {acc:.3f}", flush=True)
    assert loss < 0.5 * l0 and acc > 0.95


def test_routed_ivf_search_matches_oracle():
    print("\n[gpu] routed ivf search", flush=True)
    X = datasets.zipf_anisotropic_mixture(20_000, 32, components=30, seed=8)
    C0 = X[np.random.default_rng(8).choice(20_000, 64, replace=False)]
    oc, oa, _ = O.lloyd(X, C0, 5)
    index = I.build_index_from_assignment(X.T, oc.T, oa, I.TopKRouting(6))
    r = RT.LearnedProbeRouter(32, 64, hidden=64, seed=9)
    Q = datasets.mixture_queries(300, 32, components=30, seed=8, query_seed=12)
    ids, dist = RT.query_arrays_routed(index, r, Q.T, 10, 6)
    route = -r.scores(Q.T)
    off, lids, nb = O.partition(oa)
    oi, od = O.ivf_search(X, off, lids, nb, Q, route, 10, 6)
    assert (ids == oi).all()
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()


def _prior_index(d, c, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3000, d)).astype(np.float32)
    C = rng.standard_normal((c, d)).astype(np.float32)
    a, _ = O.kmeans_assign(X, C)
    return C, I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(4))


def test_router_prior_gradient_and_scores():
    print("\n[gpu] router centroid prior", flush=True)
    rng = np.random.default_rng(4)
    d, h, c, B, gamma = 48, 64, 40, 96, 0.37
    C, index = _prior_index(d, c, 4)
    r = RT.LearnedProbeRouter(d, c, hidden=h, seed=3)
    r.set_prior(index, gamma)
    p = r.get_params().astype(np.float64)
    X = rng.standard_normal((B, d)).astype(np.float32)
    labels = rng.integers(0, c + 1, size=(B, 3))
    labels[:, 0] = rng.integers(1, c + 1, size=B)
    prior = RO.prior_logits(X, C, gamma)
    loss, g = r.grad(X.T, labels)
    oloss, og = RO.loss_and_grad(p, X, labels, d, h, c, bf16=True, prior=prior)
    assert abs(loss - oloss) <= 1e-4 * (1 + abs(oloss))
    err = np.abs(g - og).max() / (np.abs(og).max() + 1e-12)
    print(f"max rel grad err {err:.2e}", flush=True)
    assert err < 2e-3
    s = r.scores(X.T)
    os_ = RO.scores(p, X, d, h, c, prior=prior)
    assert np.abs(s - os_).max() <= 1e-4 * (1 + np.abs(os_).max())
    # prior alone (zero MLP): the ranking is the centroid ranking
    r.zero_output_layer()
    s0 = r.scores(X.T)
    R = O.route_distances(X, C)
    assert (np.argsort(-s0, axis=1, kind="stable")[:, :5] ==
            np.argsort(R, axis=1, kind="stable")[:, :5]).mean() > 0.99
    # removing the prior restores the plain MLP
    r.set_prior(None, 0.0)
    assert np.abs(r.scores(X.T)).max() == 0.0


def test_router_prior_training_gains_over_centroids():
    # lists of an anisotropic mixture: the trained correction must lower the
    # held-out cross-entropy of the centroid prior alone (what the network adds)
    print("\n[gpu] router training over the prior", flush=True)
    X = datasets.zipf_anisotropic_mixture(30_000, 32, components=40, seed=9)
    C0 = X[np.random.default_rng(9).choice(30_000, 48, replace=False)]
    oc, oa, _ = O.lloyd(X, C0, 6)
    index = I.build_index_from_assignment(X.T, oc.T, oa, I.TopKRouting(4))
    Tq = datasets.mixture_queries(12_000, 32, components=40, seed=9, query_seed=5)
    labels = RT.neighbour_list_labels(index, X.T, Tq.T, k=10)
    tr, te = slice(0, 10_000), slice(10_000, 12_000)
    r = RT.LearnedProbeRouter(32, 48, hidden=64, seed=2)
    gamma = r.fit_prior(index, Tq[tr].T, labels[tr])
    r.zero_output_layer()
    l_prior, _ = r.grad(Tq[te].T, labels[te])
    r.train(Tq[tr].T, labels[tr], epochs=6, batch=500, lr=2e-3, seed=3)
    l_trained, _ = r.grad(Tq[te].T, labels[te])
    print(f"gamma {gamma:.4g}: held-out loss prior {l_prior:.4f} -> trained {l_trained:.4f}",
          flush=True)
    assert gamma > 0
    assert l_trained < l_prior
