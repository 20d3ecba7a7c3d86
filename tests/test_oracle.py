"""Pins the CPU oracle: the reference's own known answers, closed forms, an FP64
brute force and exhaustive tiny cases.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


# --- per-pair arithmetic -----------------------------------------------------

def test_a16_order_matches_written_out_tree(orc):
    # Hand-evaluated A16 order on d=37 (two full groups + a ragged tail),
    # independent of the C code: accumulate acc[i & 15] then the 8/4/2/1 tree.
    rng = np.random.default_rng(3)
    x = rng.standard_normal(37).astype(np.float32)
    q = rng.standard_normal(37).astype(np.float32)
    acc = [np.float32(0)] * 16
    for i in range(37):
        t = np.float32(x[i] - q[i])
        # fmaf(t, t, acc): exact product + acc rounded once == float64 sum rounded
        # when |t*t| fits in 48 bits of mantissa (it does: t is f32)
        acc[i & 15] = np.float32(np.float64(t) * np.float64(t) + np.float64(acc[i & 15]))
    for s in (8, 4, 2, 1):
        for j in range(s):
            acc[j] = np.float32(acc[j] + acc[j + s])
    assert orc.sumsq_diff(x, q) == acc[0]


def test_metrics_closed_forms(orc):
    x = np.array([1.0, 2.0, 3.0], np.float32)
    y = np.array([1.5, 2.5, -1.0], np.float32)
    # bruteforce_tests.jl:4-13: d^2 ~ dsq (atol 1e-5), Float32 results
    d = orc.metric(O.EUCLIDEAN, x, y)
    dsq = orc.metric(O.SQEUCLIDEAN, x, y)
    assert d >= 0 and dsq >= 0
    assert abs(float(d) ** 2 - float(dsq)) < 1e-5
    assert dsq == np.float32(16.5)  # 0.25 + 0.25 + 16 exactly
    # cosine: orthogonal -> 1, parallel -> 0, antiparallel -> 2
    e1 = np.array([1, 0, 0], np.float32)
    e2 = np.array([0, 3, 0], np.float32)
    assert orc.metric(O.COSINE, e1, e2) == 1.0
    assert orc.metric(O.COSINE, e1, 2 * e1) == 0.0
    assert orc.metric(O.COSINE, e1, -e1) == 2.0


# --- BruteForceIndex -----------------------------------------------------------

def test_bruteforce_known_answer(orc):
    # bruteforce_tests.jl:22-28: data [0 1 2; 0 0 0], q=[0.25, 0], k=2 -> [1, 2]
    base = np.array([[0, 0], [1, 0], [2, 0]], np.float32)
    ids, dist = orc.bf_knn(base, np.array([[0.25, 0.0]], np.float32), 2)
    assert ids.tolist() == [[1, 2]]
    assert dist.tolist() == [[0.25, 0.75]]


def test_bruteforce_k_clamp_and_padding(orc):
    base = np.array([[0, 0], [1, 0], [2, 0]], np.float32)
    ids, dist = orc.bf_knn(base, np.array([[5.0, 0.0]], np.float32), 5)
    assert ids.tolist() == [[3, 2, 1, 0, 0]]
    assert np.isinf(dist[0, 3:]).all()


def test_bruteforce_ties_lower_index(orc):
    # all points equidistant -> ids ascending
    base = np.array([[1, 0], [0, 1], [-1, 0], [0, -1]], np.float32)
    ids, _ = orc.bf_knn(base, np.zeros((1, 2), np.float32), 4)
    assert ids.tolist() == [[1, 2, 3, 4]]


@pytest.mark.parametrize("metric", [O.EUCLIDEAN, O.SQEUCLIDEAN, O.COSINE])
def test_bruteforce_vs_fp64(orc, metric):
    rng = np.random.default_rng(11)
    base = rng.standard_normal((500, 24)).astype(np.float32)
    Q = rng.standard_normal((40, 24)).astype(np.float32)
    k = 7
    ids, dist = orc.bf_knn(base, Q, k, metric)
    ids64, d64 = O.fp64_knn(base, Q, k + 1, metric)
    # identical unless the FP64 gap at the boundary is within FP32 rounding
    amb = O.fp32_rounding_gap_ambiguous(d64, k, 24)
    assert amb == 0
    assert (ids == ids64[:, :k]).all()
    np.testing.assert_allclose(dist, d64[:, :k], rtol=2e-6, atol=2e-6)


def test_batch_equals_single(orc):
    # abstract_batch_query_tests.jl:9-61 (16 x 1000 randn, 80 queries, k=5)
    rng = np.random.default_rng(5)
    base = rng.standard_normal((1000, 16)).astype(np.float32)
    Q = rng.standard_normal((80, 16)).astype(np.float32)
    ids, dist = orc.bf_knn(base, Q, 5)
    for i in range(0, 80, 7):
        i1, d1 = orc.bf_knn(base, Q[i:i + 1], 5)
        assert (i1[0] == ids[i]).all() and (d1[0] == dist[i]).all()


# --- kNN graph -------------------------------------------------------------------

def test_knn_graph_known_answer(orc):
    # graph_tests.jl:4-48: data = reshape(1:12, 3, 4), collinear equally spaced
    data = np.arange(1, 13, dtype=np.float32).reshape(4, 3)
    g = orc.knn_graph(data, 2)
    assert g.tolist() == [[2, 3], [1, 3], [2, 4], [3, 2]]
    gs = orc.knn_graph(data, 3, include_self=True)
    for i, row in enumerate(gs.tolist()):
        assert (i + 1) in row and len(row) == 3
    with pytest.raises(ValueError):
        orc.knn_graph(data, 0)
    with pytest.raises(ValueError):
        orc.knn_graph(data, 4)


def test_symmetrize_union_sorted(orc):
    adj = np.array([[2, 3], [1, 3], [2, 4], [3, 2]], np.int64)
    lists, _ = orc.symmetrize(adj)
    assert [l.tolist() for l in lists] == [[2, 3], [1, 3, 4], [1, 2, 4], [2, 3]]


# --- k-means ---------------------------------------------------------------------

def test_lloyd_empty_cluster_known_answer(orc):
    # test_kmeans_rng_and_empty_clusters.jl:14-51: 5 points on a line, three
    # centroids at 1e9, max_iters=1 -> the three repaired centroids are
    # distinct data points (farthest first, marked -Inf).
    X = np.array([[0, 0], [1, 0], [2, 0], [3, 0], [4, 0]], np.float32)
    C0 = np.array([[0, 0], [1e9, 1e9], [1e9, 1e9], [1e9, 1e9]], np.float32)
    C, assign, iters = orc.lloyd(X, C0, 1)
    assert iters == 1
    assert C.tolist() == [[2, 0], [4, 0], [3, 0], [2, 0]]
    reps = [tuple(r) for r in C[1:].tolist()]
    assert len(set(reps)) == 3
    assert all(r in [tuple(x) for x in X.tolist()] for r in reps)


def test_kmeans_assign_vs_fp64_and_self_distance(orc):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((400, 12)).astype(np.float32)
    C = X[:9].copy()
    a, md = orc.kmeans_assign(X, C)
    # test_kmeans_optimizations.jl:107-122: clamp keeps self distance >= 0, ~0
    assert (md[:9] >= 0).all() and (md[:9] < 0.01).all()
    D64 = np.sqrt(((X[:, None, :].astype(np.float64) - C[None].astype(np.float64)) ** 2).sum(-1))
    best = D64.argmin(1) + 1
    srt = np.sort(D64, 1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-4 * (1 + srt[:, 1])
    assert (a[clear] == best[clear]).all()
    # test_kmeans_optimizations.jl:9-67: norms form ~ direct norm (rtol 1e-5, plus
    # the absolute cancellation floor of the norms form near 0)
    np.testing.assert_allclose(md, D64.min(1), rtol=1e-5, atol=1e-3)


def test_lloyd_textbook_tiny(orc):
    # Two well-separated blobs; Lloyd from two seed points converges to the blob
    # means (FP64 closed form) and assigns every point to its own blob.
    rng = np.random.default_rng(9)
    A = rng.standard_normal((50, 3)) * 0.1
    B = rng.standard_normal((60, 3)) * 0.1 + 10
    X = np.vstack([A, B]).astype(np.float32)
    C, assign, iters = orc.lloyd(X, X[[0, 50]], 20)
    assert (assign[:50] == 1).all() and (assign[50:] == 2).all()
    np.testing.assert_allclose(C[0], X[:50].astype(np.float64).mean(0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(C[1], X[50:].astype(np.float64).mean(0), rtol=1e-6, atol=1e-6)
    assert iters < 20  # converged (movement < tol) before the cap


def test_kmeanspp_properties(orc):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3000, 8)).astype(np.float32)
    C, idx = orc.kmeanspp(X, 16, seed=1234)
    assert len(set(idx.tolist())) == 16            # D^2 sampling never repeats a centre
    assert (C == X[idx - 1]).all()
    C2, idx2 = orc.kmeanspp(X, 16, seed=1234)
    assert (idx == idx2).all()                      # same seed -> same centres
    _, idx3 = orc.kmeanspp(X, 16, seed=99)
    assert (idx != idx3).any()
    u0 = O.uniform_stream(1234, 1)[0]
    assert idx[0] == int(u0 * 3000) + 1             # first centre = floor(u0 * n)


def test_splitmix64_known_answers(orc):
    # The published splitmix64 reference outputs for state 0 (Vigna's
    # splitmix64.c): the stream every seeded draw of the library comes from.
    assert O.splitmix64_stream(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                         0x06C45D188009454F]
    # uniform = top 53 bits / 2^53
    assert O.uniform_stream(0, 1)[0] == (0xE220A8397B1DCDAF >> 11) / 2.0**53


def test_kmeanspp_closed_form_d2_sampling(orc):
    # init.jl:56-102 on three points of a line at 0, 1 and 3 with the first
    # centre at 0: the squared distances are [0, 1, 9], the inclusive
    # cumulative sum [0, 1, 10], so round 1 picks the first j whose cumsum
    # reaches u * 10.
    X = np.array([[0.0], [1.0], [3.0]], np.float32)

    def pick(u1):
        _, idx = O.kmeanspp_draws(X, 2, [0.0, u1])
        assert idx[0] == 1
        return int(idx[1])

    # u in (0.1, 1): threshold in (1, 10] -> point 3.  A D-weighted CDF
    # ([0, 1, 4] / 4) would pick point 2 for u in (0.1, 0.25).
    assert pick(0.2) == 3
    assert pick(0.12) == 3
    assert pick(0.99) == 3
    # threshold exactly on the cumulative value 1 (0.1 * 10 == 1.0 in FP64):
    # searchsortedfirst takes cumsum[j] >= threshold -> point 2 (a strict '>'
    # search would take point 3)
    assert pick(0.1) == 2
    # below the boundary: point 2.  An exclusive cumulative sum ([0, 0, 1])
    # would take point 3 here.
    assert pick(0.05) == 2
    # u = 0: threshold 0 is reached by cumsum[1] = 0 at the first centre itself
    # (searchsortedfirst of 0 over [0, 1, 10] is 1)
    assert pick(0.0) == 1
    # first centre: rand(rng, 1:n) as floor(u_0 n) + 1
    for u0, want in [(0.0, 1), (0.34, 2), (0.7, 3), (0.999999, 3)]:
        _, idx = O.kmeanspp_draws(X, 1, [u0])
        assert idx[0] == want
    # round 2 after centres 0 and 3: squared distances [0, 1, 0], cumsum
    # [0, 1, 1]: every u in (0, 1) picks point 2
    for u2 in (1e-9, 0.5, 0.999):
        _, idx = O.kmeanspp_draws(X, 3, [0.0, 0.5, u2])
        assert idx.tolist() == [1, 3, 2]


def test_kmeanspp_sampling_frequencies(orc):
    # D^2 weighting as a distribution: with the first centre at 0 on points
    # {0, 1, 3, 7}, squared distances [0, 1, 9, 49] -> P = [0, 1, 9, 49] / 59.
    # 20,000 evenly spaced draws reproduce the CDF to within 1e-3.
    X = np.array([[0.0], [1.0], [3.0], [7.0]], np.float32)
    m = 20000
    counts = np.zeros(5, np.int64)
    for i in range(m):
        _, idx = O.kmeanspp_draws(X, 2, [0.0, (i + 0.5) / m])
        counts[idx[1]] += 1
    p = counts[1:] / m
    np.testing.assert_allclose(p, np.array([0, 1, 9, 49]) / 59.0, atol=1e-3)


def test_kmeanspp_blocked_cumsum_matches_fp64(orc):
    # Larger input: every pick of the seeded stream equals the first j whose
    # FP64 cumulative sum of the running minimum squared distance reaches
    # u_i * total, recomputed here in FP64 from the points themselves (the
    # oracle's FP32 norms form differs from it by rounding only; points are
    # integers so both are exact).
    rng = np.random.default_rng(5)
    X = rng.integers(-20, 21, size=(2500, 3)).astype(np.float32)
    k = 12
    u = O.uniform_stream(77, k)
    _, idx = O.kmeanspp_draws(X, k, u)
    X64 = X.astype(np.float64)
    first = int(u[0] * len(X))
    assert idx[0] == first + 1
    mins = np.full(len(X), np.inf)
    chosen = [first]
    for i in range(1, k):
        c = X64[chosen[-1]]
        mins = np.minimum(mins, ((X64 - c) ** 2).sum(1))
        cs = np.cumsum(mins)
        j = int(np.searchsorted(cs, u[i] * cs[-1], side="left"))
        j = min(j, len(X) - 1)
        assert idx[i] == j + 1
        chosen.append(j)
    # the seeded entry point draws the same stream in the same order
    _, idx_seed = orc.kmeanspp(X, k, seed=77)
    assert (idx_seed == idx).all()


# --- routing / partition / merge --------------------------------------------------------

def test_topk_routing_known_answers(orc):
    # routing_tests.jl:8-57
    d = np.array([0.9, 0.8, 0.7, 0.6, 0.5, 0.1, 0.2], np.float32)
    assert orc.select_topk(d, 5, 3).tolist() == [5, 4, 3]
    assert orc.select_topk(np.array([0.4, 0.3, 0.2, 0.1], np.float32), 4, 10).tolist() == [4, 3, 2, 1]
    assert orc.select_topk(np.array([0.1, 0.2], np.float32), 0, 3).tolist() == []
    with pytest.raises(ValueError):
        orc.select_topk(np.zeros(0, np.float32), 3, 5)


def test_partition_stable(orc):
    # utils_tests.jl:6-26: partitions cover all points, ids ascending per bucket
    a = np.array([2, 1, 2, 3, 1, 2], np.int64)
    off, ids, nb = orc.partition(a)
    assert nb == 3
    assert off.tolist() == [0, 2, 5, 6]
    assert ids.tolist() == [2, 5, 1, 3, 6, 4]


def test_simple_merge_docstring_example(orc):
    # merge.jl:60-66
    ids = np.array([10, 20, 30, 40, 50])
    dist = np.array([0.5, 0.8, 0.3, 1.2, 0.6], np.float32)
    oi, od = orc.simple_merge(ids, dist, 3)
    assert oi.tolist() == [30, 10, 50]
    # stable ties keep list order; duplicates are dropped
    oi, _ = orc.simple_merge(np.array([7, 3, 7, 1]), np.array([1, 1, 1, 0], np.float32), 3)
    assert oi.tolist() == [1, 7, 3]


def test_disjoint_merge_heap(orc):
    ids = np.array([10, 20, 30, 40, 50])
    dist = np.array([0.5, 0.8, 0.3, 1.2, 0.6], np.float32)
    oi, od = orc.disjoint_merge(ids, dist, 3)
    assert oi.tolist() == [30, 10, 50]
    # strict '<' eviction: an equal-distance newcomer does not evict
    oi, _ = orc.disjoint_merge(np.array([5, 4, 9]), np.array([1, 1, 1], np.float32), 2)
    assert oi.tolist() == [4, 5]


def test_ivf_exhaustive_equals_bruteforce(orc):
    # query_behavior_tests.jl:63-96 / distance_tests.jl:13-32: probing every
    # bucket returns the brute-force neighbours.
    rng = np.random.default_rng(21)
    X = rng.standard_normal((600, 10)).astype(np.float32)
    Q = rng.standard_normal((30, 10)).astype(np.float32)
    C = X[:8].copy()
    a, _ = orc.kmeans_assign(X, C)
    off, ids, nb = orc.partition(a)
    route = orc.route_distances(Q, C)
    for merge in ("simple", "disjoint"):
        oi, od = orc.ivf_search(X, off, ids, nb, Q, route, 6, 8, merge=merge)
        bi, bd = orc.bf_knn(X, Q, 6)
        assert (oi == bi).all()
        assert (od == bd).all()


def test_ivf_nprobe1_scans_nearest_list_only(orc):
    rng = np.random.default_rng(22)
    X = rng.standard_normal((300, 6)).astype(np.float32)
    Q = rng.standard_normal((10, 6)).astype(np.float32)
    C = X[:5].copy()
    a, _ = orc.kmeans_assign(X, C)
    off, ids, nb = orc.partition(a)
    route = orc.route_distances(Q, C)
    oi, _ = orc.ivf_search(X, off, ids, nb, Q, route, 4, 1)
    for qi in range(10):
        b = int(np.lexsort((np.arange(route.shape[1]), route[qi]))[0])
        members = set(ids[off[b]:off[b + 1]].tolist())
        got = [g for g in oi[qi].tolist() if g]
        assert set(got) <= members
        sub = np.array(sorted(members)) - 1
        si, _ = orc.bf_knn(X[sub], Q[qi:qi + 1], 4)
        assert got == [int(sub[s - 1] + 1) for s in si[0] if s]


def test_recall_closed_form(orc):
    found = np.array([[1, 2, 3], [4, 5, 6]])
    truth = np.array([[3, 2, 9], [7, 8, 9]])
    assert orc.recall_at_k(found, truth) == pytest.approx((2 / 3 + 0) / 2)
