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


def _mass_reference(route, list_sizes, nprobe_max, mass):
    """FP64 restatement of probability-mass probing: lists in ascending route
    order (ties to the lower bucket), softmax of -route, stop once the mass is
    reached or nprobe_max non-empty lists are taken."""
    out = []
    avail = min(len(list_sizes), np.asarray(route).shape[1])
    for r in np.asarray(route, np.float64)[:, :avail]:
        order = np.lexsort((np.arange(r.size), r))
        p = np.exp(r.min() - r)
        p /= p.sum()
        cum, taken, cums = 0.0, [], []
        for b in order:
            if cum >= mass or len(taken) >= nprobe_max:
                break
            cum += p[b]
            if list_sizes[b] > 0:
                taken.append(b + 1)
                cums.append(cum)
        out.append((taken, cums))
    return out


def test_probability_mass_probing_matches_definition():
    print("\n[gpu] probability-mass probing", flush=True)
    X, C, a = _small_ivf_data()
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(4))
    rng = np.random.default_rng(12)
    route = (3.0 * rng.standard_normal((300, C.shape[0]))).astype(np.float32)
    for mass in (0.5, 0.9, 0.99):
        got = RT.probe_mass(index, route, 16, mass)
        ref = _mass_reference(route, index.list_sizes, 16, mass)
        for q in range(route.shape[0]):
            g = [int(b) for b in got[q] if b > 0]
            taken, cums = ref[q]
            # identical list order; counts may differ only where the FP32
            # running sum sits within 1e-5 of the mass
            n = min(len(g), len(taken))
            assert g[:n] == taken[:n]
            if len(g) != len(taken):
                edge = cums[min(len(g), len(taken)) - 1]
                assert abs(edge - mass) < 1e-5, (q, g, taken)


def _small_ivf_data():
    X = datasets.zipf_anisotropic_mixture(20_000, 32, components=30, seed=8)
    C0 = X[np.random.default_rng(8).choice(20_000, 64, replace=False)]
    oc, oa, _ = O.lloyd(X, C0, 5)
    return X, oc, oa


def test_mass_routed_ivf_search_matches_oracle():
    print("\n[gpu] mass-routed ivf search", flush=True)
    X, C, a = _small_ivf_data()
    index = I.build_index_from_assignment(X.T, C.T, a, I.TopKRouting(6))
    r = RT.LearnedProbeRouter(32, 64, hidden=64, seed=9)
    r.set_prior(index, r.fit_prior(index, *_prior_training(X, index)))
    Q = datasets.mixture_queries(200, 32, components=30, seed=8, query_seed=13)
    ids, dist = RT.query_arrays_routed(index, r, Q.T, 10, 12, mass=0.9)
    probes_total = RT.last_probe_count()
    route = -r.scores(Q.T)
    chosen = RT.probe_mass(index, route, 12, 0.9)
    assert probes_total == int((chosen > 0).sum())
    off, lids, nb = O.partition(a)
    for q in range(Q.shape[0]):
        lists = [int(b) - 1 for b in chosen[q] if b > 0]
        rq = np.full((1, C.shape[0]), np.inf, np.float32)
        rq[0, lists] = np.arange(len(lists), dtype=np.float32)
        oi, od = O.ivf_search(X, off, lids, nb, Q[q:q + 1], rq, 10, len(lists))
        assert (ids[q] == oi[0]).all(), q
        assert (dist[q].view(np.uint32) == od[0].view(np.uint32)).all(), q


def _prior_training(X, index):
    Tq = datasets.mixture_queries(2000, 32, components=30, seed=8, query_seed=14)
    labels = RT.neighbour_list_labels(index, X.T, Tq.T, k=10)
    return Tq.T, labels
