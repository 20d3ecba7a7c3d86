"""FP8 (OCP E4M3) first pass of the exact search against the oracle.

The first pass changes, the answer does not: with `first_pass="fp8"` the
returned ids and distance bits equal the oracle's FP32 brute force (and the
bf16 first pass's), on unit vectors (cosine), on centred Euclidean data with a
dimension that is not a multiple of 128, on the self-join of the kNN graph
(whose rows carry folded norms), and on duplicated points, whose ties push
queries through the second pass and the exact fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import datasets
from oracle import oracle as O


def _banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _same(ids, dist, oi, od, tag):
    assert ids.shape == oi.shape, tag
    bad = np.argwhere(ids != oi)
    assert bad.size == 0, (tag, bad[:5])
    assert (dist.view(np.uint32) == od.view(np.uint32)).all(), tag


def _both(index, ds, Q, k):
    """(fp8 ids, dist, stats, emitted), then the bf16 pass on the same rows."""
    ds.set_first_pass("fp8")
    i8, d8 = M.query_arrays(index, ds, Q.T, k)
    st8, em8 = M.last_stats(), M.last_emitted()
    ds.set_first_pass("bf16")
    i16, d16 = M.query_arrays(index, ds, Q.T, k)
    st16, em16 = M.last_stats(), M.last_emitted()
    print(f"fp8 stats {st8} emitted/query {em8 / len(Q):.1f}; "
          f"bf16 stats {st16} emitted/query {em16 / len(Q):.1f}", flush=True)
    return i8, d8, i16, d16, st8


def test_fp8_cosine_unit_mixture_200k_k100():
    # 200,000 x 256 unit vectors (100 components of ~2,000 rows), k = 100
    _banner("fp8 cosine 200000x256 k=100")
    X = datasets.unit_mixture(200_000, 256, components=100, seed=5)
    Q = datasets.unit_mixture(64, 256, components=100, seed=77)
    index = M.build_index(M.BruteForceIndex, X.T, distance=M.CosineDist())
    ds = M.Dataset(X.T, M.CosineDist())
    i8, d8, i16, d16, st8 = _both(index, ds, Q, 100)
    oi, od = O.bf_knn(X, Q, 100, O.COSINE)
    _same(i8, d8, oi, od, "fp8")
    _same(i16, d16, oi, od, "bf16")
    assert st8[3] == 0, st8  # no query needed the exact fallback


@pytest.mark.parametrize("k", [1, 10, 100])
def test_fp8_euclidean_odd_dimension(k):
    # d = 200: FP8 rows padded to 256 bytes; centred rows (not normalised)
    _banner(f"fp8 euclidean 30000x200 k={k}")
    X = datasets.zipf_anisotropic_mixture(30_000, 200, components=60, seed=8)
    Q = datasets.mixture_queries(97, 200, components=60, seed=8, query_seed=9)
    index = M.build_index(M.BruteForceIndex, X.T)
    ds = M.Dataset(X.T)
    i8, d8, i16, d16, _ = _both(index, ds, Q, k)
    oi, od = O.bf_knn(X, Q, k)
    _same(i8, d8, oi, od, "fp8")
    _same(i16, d16, oi, od, "bf16")


def test_fp8_sqeuclidean_small_dimension():
    _banner("fp8 sqeuclidean 5000x37 k=9")
    rng = np.random.default_rng(12)
    X = rng.standard_normal((5000, 37)).astype(np.float32)
    Q = rng.standard_normal((200, 37)).astype(np.float32)
    index = M.build_index(M.BruteForceIndex, X.T, distance=M.SqEuclidean())
    ds = M.Dataset(X.T, M.SqEuclidean(), first_pass="fp8")
    ids, dist = M.query_arrays(index, ds, Q.T, 9)
    oi, od = O.bf_knn(X, Q, 9, O.SQEUCLIDEAN)
    _same(ids, dist, oi, od, "fp8")


def test_fp8_knn_graph_self_join_folded_rows():
    # d = 100: bf16 rows of 128 columns carry folded norms past column d, which
    # the bf16 re-score of the FP8 survivors must not sum
    _banner("fp8 knn graph 12000x100 k=15")
    X = datasets.clipped_mixture(12_000, 100, clusters=10, seed=4)
    index = M.build_index(M.BruteForceIndex, X.T)
    ds = M.Dataset(X.T, first_pass="fp8")
    a8, d8 = M.build_knn_graph(index, ds, k=15, return_arrays=True)
    print(f"fp8 stats {M.last_stats()}", flush=True)
    ds.set_first_pass("bf16")
    a16, d16 = M.build_knn_graph(index, ds, k=15, return_arrays=True)
    oi = O.knn_graph(X, 15)
    assert (a8 == oi).all(), np.argwhere(a8 != oi)[:5]
    assert (a16 == oi).all()
    assert (d8.view(np.uint32) == d16.view(np.uint32)).all()


def test_fp8_duplicates_reach_second_pass_and_fallback():
    # every base point stored 8 times: ties at the k-th place
    _banner("fp8 duplicated points 8x1500x64 k=20")
    rng = np.random.default_rng(3)
    base = rng.standard_normal((1500, 64)).astype(np.float32)
    X = np.repeat(base, 8, axis=0)[rng.permutation(12_000)]
    Q = (base[rng.integers(0, 1500, 128)] + 0.01 * rng.standard_normal((128, 64))).astype(
        np.float32)
    index = M.build_index(M.BruteForceIndex, X.T)
    ds = M.Dataset(X.T, first_pass="fp8")
    for k in (20, 200):
        ids, dist = M.query_arrays(index, ds, Q.T, k)
        print(f"k={k} stats {M.last_stats()}", flush=True)
        oi, od = O.bf_knn(X, Q, k)
        _same(ids, dist, oi, od, ("dup", k))
    # all rows identical: every query falls back to the exact scan
    Y = np.ones((3000, 32), np.float32)
    ds1 = M.Dataset(Y.T, first_pass="fp8")
    ids, dist = M.query_arrays(M.build_index(M.BruteForceIndex, Y.T), ds1, Y[:3].T, 5)
    assert ids.tolist() == [[1, 2, 3, 4, 5]] * 3
    assert (dist == 0).all()


def test_first_pass_argument_checked():
    X = np.zeros((4, 10), np.float32)
    ds = M.Dataset(X)
    with pytest.raises(M.ArgumentError):
        ds.set_first_pass("int4")
