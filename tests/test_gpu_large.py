"""GPU parity of the exact path on the large-search configuration's shapes:
a Float16 base (upcast exactly to Float32), cosine, k=100, and the query-chunked
threshold estimation.  Ids and distances bit-identical to the oracle's FP32
brute force on the upcast data."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import datasets
from manifoldann_amd._lib import check, lib
from oracle import oracle as O


def _banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def test_f16_cosine_k100_bit_exact():
    # unit-normalised mixture stored as Float16, cosine, k = 100
    _banner("f16 cosine 200000x256 k=100")
    X16 = datasets.unit_mixture(200_000, 256, components=5000, seed=5).astype(np.float16)
    Q = datasets.unit_mixture(64, 256, components=5000, seed=6).astype(np.float16)
    index = M.build_index(M.BruteForceIndex, X16.T, distance=M.cosine_distance)
    ids, dist = M.query_arrays(index, X16.T, Q.T, 100)
    st = M.last_stats()
    print(f"stats={st}", flush=True)
    oi, od = O.bf_knn(X16.astype(np.float32), Q.astype(np.float32), 100, metric=O.COSINE)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()


def test_chunked_threshold_sample_bit_exact():
    # a 65,536-row sample and 5,000 queries: the dense sample block is computed
    # in query chunks (2^28 values each)
    _banner("chunked threshold sample 100000x32, 5000 queries")
    ids, dist, X, Q = _search_with_sample(65536, 100_000, 5_000, 16, seed=7)
    oi, od = O.bf_knn(X, Q, 16)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()


def _search_with_sample(sample_rows, n, nq, k, seed):
    L = lib()
    old = ctypes.c_int64()
    check(L.mann_tune_get(b"sample_rows", ctypes.byref(old)))
    check(L.mann_tune_set(b"sample_rows", sample_rows))
    try:
        X = datasets.gaussian(n, 32, seed=seed)
        Q = datasets.gaussian(nq, 32, seed=seed + 1)
        ds = M.Dataset(X.T)
        index = M.build_index(M.BruteForceIndex, X.T)
        ids, dist = M.query_arrays(index, ds, Q.T, k)
    finally:
        check(L.mann_tune_set(b"sample_rows", old.value))
    return ids, dist, X, Q


def test_block_minimum_threshold_sample_bit_exact():
    # a 65,536-row sample and k = 2: the threshold rank (about 20) lies far
    # below the 2,048 blocks of 32 sample rows, so only the block minima are
    # stored and ranked
    _banner("block-minimum threshold sample 100000x32, 3000 queries, k=2")
    ids, dist, X, Q = _search_with_sample(65536, 100_000, 3_000, 2, seed=11)
    st = M.last_stats()
    print(f"stats={st}", flush=True)
    oi, od = O.bf_knn(X, Q, 2)
    assert (ids == oi).all(), np.argwhere(ids != oi)[:5]
    assert (dist.view(np.uint32) == od.view(np.uint32)).all()
