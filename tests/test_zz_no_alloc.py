"""mann_bf_knn_noalloc: the allocation-free exact search gives mann_bf_knn's
results bit for bit, on two calls in a row (the second smaller, so it runs on
the scratch the first one left behind), for L2 and cosine, and rejects a batch
whose scratch would not fit with an ArgumentError instead of allocating."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import manifoldann_amd as M
from manifoldann_amd import datasets
from manifoldann_amd._lib import check, lib


def _run(fn, ds, Qd, k, metric):
    import torch
    nq, d = Qd.shape
    ids = torch.zeros((nq, k), dtype=torch.int64, device="cuda")
    dist = torch.zeros((nq, k), dtype=torch.float32, device="cuda")
    rc = fn(ds.handle, Qd.data_ptr(), d, nq, d, k, metric, ids.data_ptr(), dist.data_ptr())
    return rc, ids.cpu().numpy(), dist.cpu().numpy()


@pytest.mark.parametrize("metric", [M.Euclidean(), M.CosineDist()])
def test_noalloc_matches_bf_knn_twice(metric):
    torch = pytest.importorskip("torch")
    L = lib()
    print(f"\n[case] noalloc vs bf_knn, metric {metric.code}", flush=True)
    X = datasets.clipped_mixture(20_000, 200, seed=5)
    Xd = torch.from_numpy(X).cuda()
    ds = M.Dataset(Xd.t(), metric)
    for nq, k, seed in ((1500, 16, 1), (333, 7, 2)):
        Q = datasets.clipped_mixture(nq, 200, seed=10 + seed)
        Qd = torch.from_numpy(Q).cuda()
        rc0, i0, d0 = _run(L.mann_bf_knn, ds, Qd, k, metric.code)
        check(rc0)
        rc1, i1, d1 = _run(L.mann_bf_knn_noalloc, ds, Qd, k, metric.code)
        check(rc1)
        print(f"nq {nq} k {k}: stats {M.last_stats()}", flush=True)
        assert (i1 == i0).all(), np.argwhere(i1 != i0)[:5]
        assert (d1.view(np.uint32) == d0.view(np.uint32)).all()


def test_noalloc_rejects_oversized_batches_and_host_buffers():
    torch = pytest.importorskip("torch")
    L = lib()
    X = datasets.gaussian(5000, 64, seed=3)
    ds = M.Dataset(torch.from_numpy(X).cuda().t())
    # 20,000 queries need more than the 64 MiB scratch (about 8 KiB each)
    Qd = torch.from_numpy(datasets.gaussian(20_000, 64, seed=4)).cuda()
    rc, _, _ = _run(L.mann_bf_knn_noalloc, ds, Qd, 10, M.EUCLIDEAN)
    assert rc == -2
    assert b"split the query batch" in L.mann_last_error()
    # host queries would need a staging allocation
    Qh = datasets.gaussian(10, 64, seed=5)
    ids = torch.zeros((10, 3), dtype=torch.int64, device="cuda")
    dist = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    rc = L.mann_bf_knn_noalloc(ds.handle, Qh.ctypes.data, 64, 10, 64, 3, M.EUCLIDEAN,
                               ids.data_ptr(), dist.data_ptr())
    assert rc == -2
    # k above the selection's capacity takes the all-distances sort, which
    # needs scratch that grows with the dataset
    Qd = torch.from_numpy(datasets.gaussian(10, 64, seed=9)).cuda()
    rc, _, _ = _run(L.mann_bf_knn_noalloc, ds, Qd, 300, M.EUCLIDEAN)
    assert rc == -2
    assert b"allocation-free search's maximum" in L.mann_last_error()
    # the search still works afterwards
    Qd = torch.from_numpy(datasets.gaussian(100, 64, seed=6)).cuda()
    rc0, i0, d0 = _run(L.mann_bf_knn, ds, Qd, 10, M.EUCLIDEAN)
    rc1, i1, d1 = _run(L.mann_bf_knn_noalloc, ds, Qd, 10, M.EUCLIDEAN)
    assert rc0 == 0 and rc1 == 0 and (i0 == i1).all()
    assert (d0.view(np.uint32) == d1.view(np.uint32)).all()


def test_noalloc_leaves_device_memory_unchanged():
    # after the one-time set-up, repeated calls allocate nothing: the free
    # device memory the runtime reports does not move
    torch = pytest.importorskip("torch")
    L = lib()
    X = datasets.gaussian(8000, 96, seed=7)
    ds = M.Dataset(torch.from_numpy(X).cuda().t())
    Qd = torch.from_numpy(datasets.gaussian(700, 96, seed=8)).cuda()
    ids = torch.zeros((700, 12), dtype=torch.int64, device="cuda")
    dist = torch.zeros((700, 12), dtype=torch.float32, device="cuda")
    check(L.mann_bf_knn_noalloc(ds.handle, Qd.data_ptr(), 96, 700, 96, 12, M.EUCLIDEAN,
                                ids.data_ptr(), dist.data_ptr()))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(5):
        check(L.mann_bf_knn_noalloc(ds.handle, Qd.data_ptr(), 96, 700, 96, 12, M.EUCLIDEAN,
                                    ids.data_ptr(), dist.data_ptr()))
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
