"""Benchmark of the MI355X search core (one JSON line on stdout).

Default workload (N=1): the all-points kNN graph self-join on the 60,000 x 784
clipped 10-cluster mixture (k=20, L2) -- the exact-kNN configuration whose
throughput is reported as TFLOP/s = 2 n^2 d / t.  One step is one full
build_knn_graph over data already resident in HBM (device dataset prepared
once, outside the timed region, like the reference harness's prepare_*).

Also reported: the dominant kernel's roofline (HIP-event timing of the
bf16-MFMA scan kernel inside this run) and the CPU oracle timed on this
host's cores on a bounded query sample.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

METRIC = "QPS at recall@10≥0.95, learned vs centroid IVF routing; exact kNN TFLOP/s"
MFMA_BF16_PEAK_TFLOPS = 2500.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=60_000)
    ap.add_argument("--d", type=int, default=784)
    ap.add_argument("--k", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the sharded design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import manifoldann_amd as M
    from manifoldann_amd import datasets
    from manifoldann_amd._lib import check, lib

    n, d, k = args.n, args.d, args.k
    print(f"[bench] generating {n}x{d} clipped mixture", file=sys.stderr, flush=True)
    X = datasets.clipped_mixture(n, d, seed=2)
    Xd = torch.from_numpy(X).cuda()
    ds = M.Dataset(Xd.t())
    ids = torch.zeros((n, k), dtype=torch.int64, device="cuda")
    dist = torch.zeros((n, k), dtype=torch.float32, device="cuda")
    L = lib()

    def step():
        check(L.mann_knn_graph(ds.handle, k, 0, M.EUCLIDEAN, ids.data_ptr(), dist.data_ptr()))

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup {i + 1}/{args.warmup}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    stats = M.last_stats()
    ms = 1e3 * t / args.steps
    flop = 2.0 * n * n * d
    value = flop / (t / args.steps) / 1e12
    print(f"[bench] timed {args.steps} steps: {ms:.3f} ms/step", file=sys.stderr, flush=True)

    # dominant kernel duration from HIP events on its own stream (untimed step)
    check(L.mann_profile_enable(1))
    check(L.mann_profile_reset())
    step()
    kernels = {}
    for name in ("scan_filter", "scan_dense", "tau_sample", "select_rerank", "scan_filter_pass2",
                 "select_rerank_pass2", "exact_fallback"):
        c = ctypes.c_int64()
        tot = ctypes.c_double()
        check(L.mann_profile_query(name.encode(), ctypes.byref(c), ctypes.byref(tot)))
        if c.value:
            kernels[name] = round(tot.value / c.value, 4)
    check(L.mann_profile_enable(0))
    scan_ms = kernels.get("scan_filter", float("nan"))
    achieved = flop / (scan_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4),
                "traffic": None, "kernel": "scan_filter (bf16 MFMA distance + threshold emit)",
                "algorithmic_flop_per_launch": flop}

    cpu = None
    if not args.skip_cpu:
        from oracle import oracle as O
        cores = os.cpu_count() or 1
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            pass
        nsq = 256
        rows = np.random.default_rng(1).choice(n, nsq, replace=False)
        O.bf_knn(X, X[rows[:8]], k + 1, nthreads=cores)  # warm
        t1 = time.perf_counter()
        O.bf_knn(X, X[rows], k + 1, nthreads=cores)
        tc = time.perf_counter() - t1
        cpu = {"value": round(2.0 * nsq * n * d / tc / 1e12, 6), "unit": "TFLOP/s",
               "cores": cores, "kind": "port",
               "sample": f"{nsq} self-join queries (k=21) against all {n} points, "
                         f"{tc:.2f} s, oracle per-pair A16 FP32 scan + partial sort"}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 first pass (fp32 accumulate) + fp32 exact re-rank",
        "data": "synthetic (clipped 10-cluster Gaussian mixture, sigma 0.15, Philox seed 2)",
        "config": {"workload": f"kNN graph self-join n={n} d={d} k={k} L2 (exact ids)",
                   "n": n, "d": d, "k": k},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms": kernels,
        "search_stats": {"reranked": stats[0], "overflow": stats[1], "second_pass": stats[2],
                         "exact_fallback": stats[3]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
