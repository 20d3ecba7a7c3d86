"""Benchmark of the MI355X search core (one JSON line on stdout).

Headline workload (`value`): the all-points kNN graph self-join on the
60,000 x 784 clipped 10-cluster mixture (k=20, L2, exact ids), reported as
TFLOP/s = 2 n^2 d / t.  One step is one full build_knn_graph over data already
resident in HBM (the device dataset is prepared once, outside the timed
region, like the reference harness's prepare_*).

Also reported (key "large_cosine"): exact cosine kNN over a 10,000,000 x 256
Float16 base (unit 5,000-component mixture), 100,000 queries, k=100, with
both first passes (bf16 MFMA, and FP8 E4M3 MFMA re-scored on bf16) timed on the
same dataset, their first-pass candidates per query, the scan's fraction of its
MFMA peak, and a sample of queries checked against the CPU oracle.

Also reported (key "ivf"): the IVF workload -- 1,000,000 x 256 Zipf mixture,
nlist 1,024, 20 Lloyd iterations, 10,000 queries, k=10 -- as QPS at
recall@10 >= 0.95 for centroid TopKRouting and for the learned probe router
(QPS = queries / wall time of one batched search call, routing included,
median of 5 after a warm-up, as benchmarking/benchmark.py:799-828).

The dominant kernel's roofline uses HIP-event timing of its launches in this
run; the CPU oracle is timed on this host's cores on a bounded sample.
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
MFMA_FP8_PEAK_TFLOPS = 5000.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


PMC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           "scan_filter_knn_graph_60000x784_pmc.txt")


def measured_traffic():
    """HBM bytes per launch of the FILTER scan from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def kernel_avg_ms(L, check, names):
    out = {}
    for name in names:
        c = ctypes.c_int64()
        tot = ctypes.c_double()
        check(L.mann_profile_query(name.encode(), ctypes.byref(c), ctypes.byref(tot)))
        if c.value:
            out[name] = round(tot.value / c.value, 4)
    return out


def fp64_metric(np, X, Q, ids, metric):
    """FP64 distances of queries Q (m, d) to rows ids (m, j) (1-based) of X."""
    X64 = X.astype(np.float64)
    out = np.empty(ids.shape)
    for s0 in range(0, ids.shape[0], 512):
        s1 = min(ids.shape[0], s0 + 512)
        R = X64[ids[s0:s1] - 1]
        q = Q[s0:s1].astype(np.float64)[:, None, :]
        if metric == 2:   # CosineDist
            dot = (R * q).sum(-1)
            out[s0:s1] = np.maximum(
                1.0 - dot / (np.linalg.norm(R, axis=-1) * np.linalg.norm(q, axis=-1)), 0.0)
        else:
            out[s0:s1] = np.sqrt(((R - q) ** 2).sum(-1))
    return out


def precision_report(np, X, Q, k, gpu_ids, gpu_dist, o_ids, o_dist, metric):
    """Parity and FP64 error of one configuration's GPU results on the queries
    Q: ids and distance bits against the oracle (o_ids / o_dist hold the
    oracle's top-(k+1)), the GPU distances' error against FP64, and the number
    of queries whose FP64 k-th and (k+1)-th distances lie within the FP32
    rounding bound (where FP32 and FP64 may order the neighbours differently)."""
    d = X.shape[1]
    u = 2.0 ** -24
    d64 = fp64_metric(np, X, Q, gpu_ids, metric)
    err = np.abs(gpu_dist.astype(np.float64) - d64)
    rel = err / np.maximum(d64, 1e-30)
    o64 = np.sort(fp64_metric(np, X, Q, o_ids[:, :k + 1], metric), axis=1)
    gap = o64[:, k] - o64[:, k - 1]
    bound = (2 * d + 16) * u if metric == 2 else (d + 4) * u * o64[:, k]
    return {"queries": int(Q.shape[0]),
            "ids_equal_oracle": bool((gpu_ids == o_ids[:, :k]).all()),
            "dist_bits_equal_oracle": bool((gpu_dist.view(np.uint32) ==
                                            o_dist[:, :k].view(np.uint32)).all()),
            "max_abs_err_vs_fp64": float(err.max()), "max_rel_err_vs_fp64": float(rel.max()),
            "fp32_ambiguous_queries": int((gap <= bound).sum())}


def bench_config1(args, M, datasets, L, check, np, torch):
    """Plain BruteForceIndex exact kNN, L2: base 10,000 x 128 N(0, 1), 1,000
    queries, k=10 -- the CPU-runnable parity configuration."""
    from oracle import oracle as O
    X = datasets.gaussian(10_000, 128, seed=1)
    Q = datasets.gaussian(1_000, 128, seed=2)
    Xd = torch.from_numpy(X).cuda()
    Qd = torch.from_numpy(Q).cuda()
    ds = M.Dataset(Xd.t())
    ids = torch.zeros((1000, 10), dtype=torch.int64, device="cuda")
    dist = torch.zeros((1000, 10), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # (the library runs on its own stream)

    def run():
        check(L.mann_bf_knn(ds.handle, Qd.data_ptr(), 128, 1000, 128, 10, M.EUCLIDEAN,
                            ids.data_ptr(), dist.data_ptr()))

    qps = qps_of(run, 1000, args.reps, torch)
    oi, od = O.bf_knn(X, Q, 11, nthreads=host_cores())
    rep = precision_report(np, X, Q, 10, ids.cpu().numpy(), dist.cpu().numpy(), oi, od, 0)
    # the same search captured in a HIP graph (one launch per batch)
    g = M.SearchGraph(M.BruteForceIndex(128, 10_000), ds, Qd.t(), 10)
    qps_g = qps_of(g.launch, 1000, args.reps, torch)
    gi, gd = g.launch()
    same = bool((gi == ids).all().item() and (gd == dist).all().item())
    # k beyond the certified selection's capacity (any k, as bruteforce.jl:46-71):
    # every exact distance sorted per query; all 1,000 queries against the oracle
    kl = 1000
    idl = torch.zeros((1000, kl), dtype=torch.int64, device="cuda")
    dil = torch.zeros((1000, kl), dtype=torch.float32, device="cuda")

    def run_large():
        check(L.mann_bf_knn(ds.handle, Qd.data_ptr(), 128, 1000, 128, kl, M.EUCLIDEAN,
                            idl.data_ptr(), dil.data_ptr()))

    qps_l = qps_of(run_large, 1000, args.reps, torch)
    li, ld = O.bf_knn(X, Q, kl, nthreads=host_cores())
    large_k = {"k": kl, "qps": round(qps_l, 1),
               "path": "every exact FP32 distance + per-query segmented radix sort",
               "ids_equal_oracle": bool((idl.cpu().numpy() == li).all()),
               "dist_bits_equal_oracle": bool((dil.cpu().numpy().view(np.uint32)
                                               == ld.view(np.uint32)).all())}
    log(f"exact k={kl}: {qps_l:.0f} QPS, oracle equal {large_k['ids_equal_oracle']}")
    rep.update({"workload": "exact kNN L2: base 10000x128 N(0,1), 1000 queries, k=10",
                "qps": round(qps, 1), "qps_hip_graph": round(qps_g, 1),
                "hip_graph_same_result": same, "large_k": large_k})
    del g, ds, Xd, Qd, idl, dil
    return rep


def bench_knn_graph(args, M, datasets, L, check, np, torch):
    n, d, k = args.n, args.d, args.k
    log(f"kNN graph: generating {n}x{d} clipped mixture")
    X = datasets.clipped_mixture(n, d, seed=2)
    Xd = torch.from_numpy(X).cuda()
    ds = M.Dataset(Xd.t())
    ids = torch.zeros((n, k), dtype=torch.int64, device="cuda")
    dist = torch.zeros((n, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def step():
        check(L.mann_knn_graph(ds.handle, k, 0, M.EUCLIDEAN, ids.data_ptr(), dist.data_ptr()))

    # the autotuner picks this shape's knobs (untimed, like the harness's prepare step)
    tuned = None
    if not args.no_autotune:
        tuned = M.autotune(M.BruteForceIndex(d, n), ds, None, k)
        log(f"kNN graph: autotuned {tuned}")
    for i in range(args.warmup):
        step()
    log(f"kNN graph: {args.warmup} warmup steps done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    stats = M.last_stats()
    flop = 2.0 * n * n * d
    log(f"kNN graph: {1e3 * t:.3f} ms/step")
    check(L.mann_profile_enable(1))
    check(L.mann_profile_reset())
    step()
    kernels = kernel_avg_ms(L, check, ["scan_filter", "scan_dense", "tau_sample", "select_rerank",
                                       "scan_filter_pass2", "select_rerank_pass2",
                                       "exact_fallback"])
    check(L.mann_profile_enable(0))
    scan_ms = kernels.get("scan_filter", float("nan"))
    achieved = flop / (scan_ms * 1e-3) / 1e12
    traffic = measured_traffic()
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "traffic_source": ("rocprofv3 FETCH_SIZE+WRITE_SIZE, " + os.path.relpath(
                PMC_SUMMARY, os.path.dirname(os.path.abspath(__file__)))) if traffic else None,
            "kernel": "scan_grid_kernel FILTER (bf16 MFMA distances + threshold emit)",
            "algorithmic_flop_per_launch": flop, "avg_launch_ms": scan_ms}
    if args.copy_calibration:
        # a plain device copy of known size: the reference reading for the
        # FETCH_SIZE / WRITE_SIZE counters of the --pmc runs
        nb = 1 << 30
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nb} bytes read and written")
        del src, dst
    cpu = None
    if not args.skip_cpu:
        from oracle import oracle as O
        cores = host_cores()
        nsq = args.cpu_queries
        rows = np.random.default_rng(1).choice(n, nsq, replace=False)
        O.bf_knn(X, X[rows[:8]], k + 1, nthreads=cores)
        t1 = time.perf_counter()
        O.bf_knn(X, X[rows], k + 1, nthreads=cores)
        tc = time.perf_counter() - t1
        oi, od = O.bf_knn(X, X[rows], k + 2, nthreads=cores)
        gi, gd = ids.cpu().numpy()[rows], dist.cpu().numpy()[rows]
        # the oracle's self-join rows without the point itself (by id)
        keep_i = np.zeros((nsq, k + 1), np.int64)
        keep_d = np.zeros((nsq, k + 1), np.float32)
        for t_, r in enumerate(rows):
            m_ = oi[t_] != r + 1
            keep_i[t_] = oi[t_][m_][:k + 1]
            keep_d[t_] = od[t_][m_][:k + 1]
        precision = precision_report(np, X, X[rows], k, gi, gd, keep_i, keep_d, 0)
        cpu = {"value": round(2.0 * nsq * n * d / tc / 1e12, 6), "unit": "TFLOP/s",
               "cores": cores, "kind": "port",
               "sample": f"{nsq} self-join queries (k=21) against all {n} points in {tc:.2f} s "
                         f"(oracle per-pair A16 FP32 scan + partial sort, OpenMP)"}
    res = {"ms_per_step": 1e3 * t, "value": flop / t / 1e12, "kernels_ms": kernels,
           "autotune": tuned,
           "precision": precision if not args.skip_cpu else None,
           "stats": {"reranked": stats[0], "overflow": stats[1], "second_pass": stats[2],
                     "exact_fallback": stats[3]}}
    del ds, Xd
    torch.cuda.empty_cache()
    return res, roof, cpu


def qps_of(fn, nq, reps, torch):
    times = []
    fn()
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    times.sort()
    return nq / times[len(times) // 2]


def bench_ivf(args, M, datasets, L, check, np, torch):
    from manifoldann_amd import ivf as I
    from manifoldann_amd import router as RT
    n, d, nlist, nq = args.ivf_n, 256, args.nlist, args.ivf_nq
    t_all = time.perf_counter()
    log(f"IVF: generating {n}x{d} Zipf mixture")
    X = datasets.zipf_anisotropic_mixture(n, d, components=1000, seed=3)
    Xd = torch.from_numpy(X).cuda()
    ds = M.Dataset(Xd.t())
    t0 = time.perf_counter()
    t = I.KMeansTransform(k=nlist, max_iters=args.iters, tol=1e-12)
    I.fit(t, ds, rng=7)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    log(f"IVF: k-means fit {fit_s:.2f} s ({t.n_iters} iterations)")
    index = I.build_index_from_assignment(ds, t.centroids, t.pending_assignments,
                                          I.TopKRouting(1))
    Q = datasets.mixture_queries(nq, d, components=1000, seed=3, query_seed=33)
    Qd = torch.from_numpy(Q).cuda()
    bf = M.BruteForceIndex(d, n)
    gt, _ = M.query_arrays(bf, ds, Qd.t(), 10)
    gt = gt.cpu().numpy()
    ids = torch.zeros((nq, 10), dtype=torch.int64, device="cuda")
    dist = torch.zeros((nq, 10), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def recall(found):
        f = found.cpu().numpy()
        return float(np.mean([len(set(a.tolist()) & set(b.tolist())) / 10.0
                              for a, b in zip(f, gt)]))

    if args.ivf_search_only:
        # only the centroid search at one nprobe, three times, for counter runs
        for _ in range(3):
            check(L.mann_ivf_search(index.handle, Qd.data_ptr(), d, nq, d, 10,
                                    args.ivf_search_only, 0, None, ids.data_ptr(),
                                    dist.data_ptr()))
        torch.cuda.synchronize()
        return {"workload": f"IVF n={n} d={d} nlist={nlist} nq={nq} k=10",
                "search_only_nprobe": args.ivf_search_only, "recall": round(recall(ids), 4)}
    sweep_c = []
    # every probe count from 8 to 20: the centroid operating point at recall
    # 0.95 is found on a fine grid, not between powers of two
    for npb in (1, 2, 4, *range(8, 21), 24, 32, 48, 64):
        def run(npb=npb):
            check(L.mann_ivf_search(index.handle, Qd.data_ptr(), d, nq, d, 10, npb, 0, None,
                                    ids.data_ptr(), dist.data_ptr()))
        qps = qps_of(run, nq, args.reps, torch)
        rc = recall(ids)
        sweep_c.append({"nprobe": npb, "recall": round(rc, 4), "qps": round(qps, 1)})
        log(f"IVF centroid nprobe={npb}: recall {rc:.4f} qps {qps:.0f}")

    # IVFFlatIndex over the same fit (ivf_flat.jl:30-168, default nprobe 10):
    # one list per centroid and the bounded-heap merge
    flat = I.ivf_flat_from_assignment(ds, t, t.pending_assignments, nprobe=10)
    sweep_f = []
    for npb in (10, 16):
        def run_f(npb=npb):
            check(L.mann_ivf_search(flat.handle, Qd.data_ptr(), d, nq, d, 10, npb, 0, None,
                                    ids.data_ptr(), dist.data_ptr()))
        qps = qps_of(run_f, nq, args.reps, torch)
        rc = recall(ids)
        rep = ctypes.c_int64()
        check(L.mann_tune_get(b"last_heap_replays", ctypes.byref(rep)))
        sweep_f.append({"nprobe": npb, "recall": round(rc, 4), "qps": round(qps, 1),
                        "heap_replays": int(rep.value)})
        log(f"IVF-Flat nprobe={npb}: recall {rc:.4f} qps {qps:.0f} heap replays {rep.value}")
    del flat

    # learned router: labels from the exact kNN of training queries
    t1 = time.perf_counter()
    Tq = datasets.mixture_queries(args.train_q, d, components=1000, seed=3, query_seed=77)
    Tqd = torch.from_numpy(Tq).cuda()
    labels = RT.neighbour_list_labels(index, ds, Tqd.t(), k=10)
    r = RT.LearnedProbeRouter(d, nlist, hidden=512, seed=11)
    gamma = None
    if not args.no_prior:
        m = min(16384, args.train_q)
        gamma = r.fit_prior(index, Tqd[:m].t(), labels[:m])
        r.zero_output_layer()
        log(f"IVF: centroid prior scale {gamma:.4g}")
    loss = r.train(Tqd.t(), labels, epochs=args.epochs, batch=1024, lr=args.lr, seed=5)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t1
    log(f"IVF: router labels+training {train_s:.2f} s, final loss {loss:.3f}")
    sweep_l = []
    for npb in (1, 2, 4, 8, 16, 32, 64):
        def run(npb=npb):
            check(L.mann_ivf_search_routed(index.handle, r.handle, Qd.data_ptr(), d, nq, d, 10,
                                           npb, 0, ids.data_ptr(), dist.data_ptr()))
        qps = qps_of(run, nq, args.reps, torch)
        rc = recall(ids)
        sweep_l.append({"nprobe": npb, "recall": round(rc, 4), "qps": round(qps, 1)})
        log(f"IVF learned nprobe={npb}: recall {rc:.4f} qps {qps:.0f}")

    # probability-mass probing: lists in router order until their softmax mass
    # reaches the target (at most 64), for the trained router and, as the
    # ablation, for its centroid prior alone (zero MLP)
    def mass_sweep(router, name):
        out = []
        for mass in (0.6, 0.8, 0.9, 0.95, 0.97, 0.98, 0.99, 0.995):
            def run(mass=mass):
                check(L.mann_ivf_search_routed_mass(index.handle, router.handle, Qd.data_ptr(), d,
                                                    nq, d, 10, 64, mass, 0, ids.data_ptr(),
                                                    dist.data_ptr()))
            qps = qps_of(run, nq, args.reps, torch)
            rc = recall(ids)
            probes = RT.last_probe_count() / nq
            out.append({"mass": mass, "recall": round(rc, 4), "qps": round(qps, 1),
                        "mean_probes": round(probes, 2)})
            log(f"IVF {name} mass={mass}: recall {rc:.4f} qps {qps:.0f} probes {probes:.2f}")
        return out

    sweep_m = mass_sweep(r, "learned")
    sweep_p = None
    if gamma is not None:
        # the baseline of probability-mass probing without the network: the
        # softmax of the scaled centroid distances alone (the router's prior
        # with a zero network), i.e. centroid routing with mass probing
        r0 = RT.LearnedProbeRouter(d, nlist, hidden=512, seed=11)
        r0.set_params(np.zeros(r0.n_params, np.float32))
        r0.set_prior(index, gamma)
        sweep_p = mass_sweep(r0, "centroid+mass")
        del r0

    def at95(sw):
        ok = [s for s in (sw or []) if s["recall"] >= 0.95]
        return max(ok, key=lambda s: s["qps"]) if ok else None

    bc, bl_fixed, bl_mass = at95(sweep_c), at95(sweep_l), at95(sweep_m)
    cands = [s for s in (bl_fixed, bl_mass) if s]
    bl = max(cands, key=lambda s: s["qps"]) if cands else None
    # the best centroid point over fixed nprobe and centroid+mass probing
    bcm = at95(sweep_p)
    bc_best = max([s for s in (bc, bcm) if s], key=lambda s: s["qps"]) if (bc or bcm) else None
    # kernel timing of the list scan at the centroid operating point
    kern = {}
    ivf_stats = None
    ivf_roof = None
    if bc:
        check(L.mann_profile_enable(1))
        check(L.mann_profile_reset())
        check(L.mann_ivf_search(index.handle, Qd.data_ptr(), d, nq, d, 10, bc["nprobe"], 0, None,
                                ids.data_ptr(), dist.data_ptr()))
        kern = kernel_avg_ms(L, check, ["prep_queries", "ivf_route", "ivf_probe",
                                        "ivf_route_scan", "ivf_route_select", "ivf_plan",
                                        "ivf_residual", "ivf_tau_scan", "ivf_tau", "ivf_scan",
                                        "ivf_select", "ivf_exact_fallback"])
        ivf_stats = M.last_stats()
        check(L.mann_profile_enable(0))
        # HBM roofline of the list scan: its algorithmic bytes are the bf16
        # residual rows (and norms) of every probed list, read once, plus one
        # residual query row per probed (list, query) entry
        route = I.route_distances(index, Qd.t())
        index.routing = I.TopKRouting(bc["nprobe"])
        probes = I.probe_lists(index, route)
        index.routing = I.TopKRouting(1)
        used = np.unique(probes[probes > 0]) - 1
        kp = (d + 31) // 32 * 32
        rows = int(np.asarray(index.list_sizes)[used].sum())
        entries = int((probes > 0).sum())
        alg = rows * (kp * 2 + 4) + entries * (kp * 2 + 8)
        # and its MFMA roofline: 2 d FLOP per (query, row of a probed list)
        sizes = np.asarray(index.list_sizes)
        scanned = int(sizes[probes[probes > 0] - 1].sum())
        flop = 2.0 * d * scanned
        scan_ms = kern.get("ivf_scan")
        if scan_ms:
            secs = scan_ms * 1e-3
            gbs = alg / secs / 1e9
            tfs = flop / secs / 1e12
            f_hbm, f_mfma = gbs / HBM_PEAK_GBS, tfs / MFMA_BF16_PEAK_TFLOPS
            # the bound is the roof nearer to the measured time:
            # max(bytes / BW, FLOP / peak) (the search roofline of the IVF workload)
            if f_mfma >= f_hbm:
                ivf_roof = {"bound": "mfma", "achieved": round(tfs, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": round(f_mfma, 4)}
            else:
                ivf_roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(f_hbm, 4)}
            ivf_roof.update({"kernel": "scan_items_kernel GATHER (IVF list scan)",
                             "hbm_gbs": round(gbs, 1), "hbm_frac": round(f_hbm, 4),
                             "mfma_tflops": round(tfs, 1), "mfma_frac": round(f_mfma, 4),
                             "algorithmic_bytes_per_launch": alg,
                             "algorithmic_flop_per_launch": flop,
                             "scanned_pairs": scanned,
                             "probed_list_rows": rows, "probed_entries": entries,
                             "avg_launch_ms": scan_ms, "nprobe": bc["nprobe"],
                             "traffic_source": "profiles/ivf_search_nprobe16_pmc.md "
                                               "(FETCH_SIZE + WRITE_SIZE at nprobe 16)"})
    # and at the learned router's operating point (probability mass)
    kern_l = {}
    if bl_mass:
        check(L.mann_profile_enable(1))
        check(L.mann_profile_reset())
        check(L.mann_ivf_search_routed_mass(index.handle, r.handle, Qd.data_ptr(), d, nq, d, 10,
                                            64, bl_mass["mass"], 0, ids.data_ptr(),
                                            dist.data_ptr()))
        kern_l = kernel_avg_ms(L, check, ["prep_queries", "router_forward", "ivf_probe",
                                          "ivf_plan", "ivf_residual", "ivf_tau_scan", "ivf_tau",
                                          "ivf_scan", "ivf_select", "ivf_exact_fallback"])
        check(L.mann_profile_enable(0))
    # the centroid operating point again, with the autotuner's knobs for its
    # shape and as a HIP-graph replay (after every sweep, so that the sweeps
    # above all ran on the same default knobs)
    tuned = None
    if bc and not args.no_autotune:
        npb = bc["nprobe"]
        choice = I.autotune(index, Qd.t(), 10, nprobe=npb)

        def run_t():
            check(L.mann_ivf_search(index.handle, Qd.data_ptr(), d, nq, d, 10, npb, 0, None,
                                    ids.data_ptr(), dist.data_ptr()))
        qps_t = qps_of(run_t, nq, args.reps, torch)
        ref_i, ref_d = ids.clone(), dist.clone()
        g = I.IVFSearchGraph(index, Qd.t(), 10, nprobe=npb)
        qps_g = qps_of(g.launch, nq, args.reps, torch)
        gi, gd = g.launch()
        same = bool((gi == ref_i).all().item() and (gd == ref_d).all().item())
        tuned = {"nprobe": npb, "recall": round(recall(ref_i), 4), "autotune": choice,
                 "qps": round(qps_t, 1), "qps_hip_graph": round(qps_g, 1),
                 "hip_graph_same_result": same}
        log(f"IVF centroid nprobe={npb} tuned: qps {qps_t:.0f}, HIP graph {qps_g:.0f} "
            f"(same result {same}), choice {choice}")
        del g
        check(L.mann_autotune_clear())
    cpu = None
    if not args.skip_cpu and bc:
        from oracle import oracle as O
        cores = host_cores()
        m = min(args.cpu_ivf_queries, nq)
        off = np.concatenate([[0], np.cumsum(index.list_sizes)])
        lids = np.zeros(n, np.int64)
        check(L.mann_ivf_list_ids(index.handle, lids.ctypes.data))
        Cc = np.ascontiguousarray(t.centroids.T)
        O.route_distances(Q[:8], Cc, nthreads=cores)
        t2 = time.perf_counter()
        R = O.route_distances(Q[:m], Cc, nthreads=cores)
        O.ivf_search(X, off, lids, index.n_buckets, Q[:m], R, 10, bc["nprobe"], nthreads=cores)
        tc = time.perf_counter() - t2
        cpu = {"value": round(m / tc, 1), "unit": "QPS", "cores": cores, "kind": "port",
               "sample": f"{m} queries at nprobe={bc['nprobe']} (oracle routing + per-list "
                         f"scans + merge, OpenMP) in {tc:.2f} s"}
    hn = None if args.skip_hnsw else bench_ivf_hnsw(args, M, I, L, check, np, torch, index, X,
                                                     Q, Qd, ids, dist, recall)
    out = {"workload": f"IVF n={n} d={d} nlist={nlist} iters={t.n_iters} nq={nq} k=10",
           "kmeans_fit_s": round(fit_s, 3), "router_train_s": round(train_s, 3),
           "router_final_loss": round(loss, 4), "train_queries": args.train_q,
           "router_prior_scale": gamma, "router_epochs": args.epochs,
           "centroid": sweep_c, "ivf_flat": sweep_f, "learned": sweep_l, "learned_mass": sweep_m,
           "centroid_mass": sweep_p,
           "centroid_at_0.95": bc, "centroid_mass_at_0.95": bcm, "centroid_best_at_0.95": bc_best,
           "centroid_at_0.95_tuned": tuned,
           "learned_at_0.95": bl,
           "learned_fixed_at_0.95": bl_fixed, "learned_mass_at_0.95": bl_mass,
           # learned router against the best centroid point (fixed nprobe on the
           # fine grid, or centroid distances with mass probing)
           "learned_over_centroid_qps": (round(bl["qps"] / bc_best["qps"], 3)
                                         if bc_best and bl else None),
           "learned_over_centroid_fixed_qps": (round(bl["qps"] / bc["qps"], 3)
                                               if bc and bl else None),
           "kernels_ms": kern, "kernels_ms_learned_mass": kern_l, "search_stats": ivf_stats,
           "roofline": ivf_roof,
           "cpu_baseline": cpu,
           "ivf_hnsw": hn,
           "total_s": round(time.perf_counter() - t_all, 1)}
    return out


def bench_ivf_hnsw(args, M, I, L, check, np, torch, index, X, Q, Qd, ids, dist, recall):
    """IVF-HNSW (build_ivf_hnsw_index's defaults: M 16, ef_construction 200,
    ef_search 64, diversified neighbours, DisjointMerge) on the IVF workload's
    k-means lists: build time, recall@10 and QPS over (nprobe, ef_search), the
    kernel times, 16 queries against the oracle and the oracle's CPU search over
    the same graphs."""
    n, d = X.shape
    nq = Q.shape[0]
    term = I.HnswTerminal()
    t0 = time.perf_counter()
    I._attach_hnsw(index.handle, term, 5)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"IVF-HNSW: graphs built in {build_s:.2f} s")

    def search(npb, ef, q=Qd, m=nq, oi=ids, od=dist):
        check(L.mann_ivf_hnsw_search(index.handle, q.data_ptr(), d, m, d, 10, npb, ef, None, 1, 0,
                                     oi.data_ptr(), od.data_ptr()))

    sweep = []
    for npb, ef in ((5, 64), (8, 64), (16, 64), (16, 128), (24, 128), (32, 128), (32, 256)):
        qps = qps_of(lambda npb=npb, ef=ef: search(npb, ef), nq, args.reps, torch)
        rc = recall(ids)
        sweep.append({"nprobe": npb, "ef_search": ef, "recall": round(rc, 4),
                      "qps": round(qps, 1)})
        log(f"IVF-HNSW nprobe={npb} ef={ef}: recall {rc:.4f} qps {qps:.0f}")
    ok = [s_ for s_ in sweep if s_["recall"] >= 0.95]
    best = max(ok, key=lambda s_: s_["qps"]) if ok else None
    check(L.mann_profile_enable(1))
    check(L.mann_profile_reset())
    search(5, 64)
    kern = kernel_avg_ms(L, check, ["prep_queries", "ivf_route", "ivf_probe", "hnsw_search",
                                    "hnsw_merge"])
    check(L.mann_profile_enable(0))
    # the walk's algorithmic bytes: one FP32 row (d floats) per exact distance,
    # divided by the search kernel's HIP-event time; it is bound by the latency
    # of its dependent steps, which this fraction of the HBM roof shows
    ndist = ctypes.c_int64()
    check(L.mann_tune_get(b"last_hnsw_distances", ctypes.byref(ndist)))
    roof = None
    if kern.get("hnsw_search"):
        gbs = ndist.value * d * 4 / (kern["hnsw_search"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4),
                "kernel": "hnsw_search_kernel (per-(query, list) graph walks, nprobe 5, ef 64)",
                "distances_per_query": round(ndist.value / nq, 1),
                "algorithmic_bytes_per_launch": ndist.value * d * 4,
                "avg_launch_ms": kern["hnsw_search"],
                "note": "latency bound: each step waits for the neighbour list, then the rows"}
    # the oracle over the GPU's graphs: 16 queries checked, a sample timed
    from oracle import oracle as O
    nbr, deg, entry, maxl, level = I.hnsw_graph(index)
    g = O.HnswGraphs(nbr, deg, entry, maxl, level, term.M)
    off = np.concatenate([[0], np.cumsum(index.list_sizes)]).astype(np.int64)
    lids = np.zeros(n, np.int64)
    check(L.mann_ivf_list_ids(index.handle, lids.ctypes.data))
    Cc = np.ascontiguousarray(index.transform.centroids.T)
    R16 = O.route_distances(Q[:16], Cc)
    oi, od = O.ivf_hnsw_search(X, off, lids, index.n_buckets, g, Q[:16], R16, 10, 5, 64,
                               merge="disjoint")
    search(5, 64)
    gi, gd = ids[:16].cpu().numpy(), dist[:16].cpu().numpy()
    parity = {"queries": 16, "ids_equal_oracle": bool(np.array_equal(gi, oi)),
              "dist_bits_equal_oracle": bool(np.array_equal(gd.view(np.uint32),
                                                            od.view(np.uint32)))}
    cpu = None
    build_cpu = None
    if not args.skip_cpu:
        cores = host_cores()
        # the oracle's serial insertion over the same lists and levels, timed on
        # the host's cores, and every graph compared with the GPU's
        t3 = time.perf_counter()
        go = O.ivf_hnsw_build(X, off, lids, index.n_buckets, term.M, term.ef_construction,
                              term.policy_code, level=level, nthreads=cores)
        tb = time.perf_counter() - t3
        live = np.arange(term.M + 1)[None, None, :] < deg[:, :, None]
        same = bool(np.array_equal(deg, go.deg) and np.array_equal(entry, go.entry) and
                    np.array_equal(np.where(live, nbr, -1), np.where(live, go.nbr, -1)))
        del go, live
        build_cpu = {"value_s": round(tb, 2), "cores": cores, "kind": "port",
                     "graphs_equal_gpu": same,
                     "sample": f"all {index.n_buckets} lists (oracle serial insertion per list, "
                               f"lists in parallel, OpenMP)"}
        log(f"IVF-HNSW: oracle build {tb:.1f} s on {cores} cores, graphs equal {same}")
        m = min(2000, nq)
        t2 = time.perf_counter()
        R = O.route_distances(Q[:m], Cc, nthreads=cores)
        O.ivf_hnsw_search(X, off, lids, index.n_buckets, g, Q[:m], R, 10, 5, 64,
                          merge="disjoint", nthreads=cores)
        tc = time.perf_counter() - t2
        cpu = {"value": round(m / tc, 1), "unit": "QPS", "cores": cores, "kind": "port",
               "sample": f"{m} queries at nprobe=5, ef_search=64 (oracle routing + per-list "
                         f"HNSW searches over the same graphs + DisjointMerge, OpenMP) in "
                         f"{tc:.2f} s"}
    del nbr, deg, g
    return {"workload": f"IVF-HNSW n={n} d={d} nlist={index.nlist} M={term.M} "
                        f"ef_construction={term.ef_construction} nq={nq} k=10",
            "build_s": round(build_s, 2), "max_list": int(np.max(index.list_sizes)),
            "sweep": sweep, "best_at_0.95": best, "kernels_ms": kern, "roofline": roof,
            "precision": parity,
            "cpu_baseline": cpu, "build_cpu_baseline": build_cpu}


def bench_large_cosine(args, M, L, check, np, torch):
    """Large cosine search: n unit-normalised rows of a 5,000-component mixture
    (noise 0.1), stored as Float16, queries from the same mixture, k=100."""
    from oracle import oracle as O
    n, d, nq, k = args.large_n, 256, args.large_nq, 100
    t_all = time.perf_counter()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(55)
    C = torch.randn(5000, d, generator=gen, device="cuda")
    C /= C.norm(dim=1, keepdim=True)

    def draw(count, seed):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        out = torch.empty((count, d), dtype=torch.float16, device="cuda")
        lab = torch.randint(0, 5000, (count,), generator=g, device="cuda")
        step = 1 << 20
        for s in range(0, count, step):
            e = min(count, s + step)
            v = C[lab[s:e]] + 0.1 * torch.randn((e - s, d), generator=g, device="cuda")
            out[s:e] = (v / v.norm(dim=1, keepdim=True)).to(torch.float16)
        return out

    log(f"large: generating {n}x{d} fp16 unit mixture on the GPU")
    X16 = draw(n, 5)
    Q32 = draw(nq, 6).to(torch.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ds = M.Dataset(X16.t(), M.cosine_distance)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    log(f"large: dataset prepared in {prep_s:.2f} s")
    ids = torch.zeros((nq, k), dtype=torch.int64, device="cuda")
    dist = torch.zeros((nq, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def run():
        check(L.mann_bf_knn(ds.handle, Q32.data_ptr(), d, nq, d, k, M.COSINE, ids.data_ptr(),
                            dist.data_ptr()))

    flop = 2.0 * nq * n * d
    # exactness on a sample of queries against the FP32 oracle on the upcast data
    m = args.large_check
    rows = np.random.default_rng(3).choice(nq, m, replace=False)
    Xh = X16.cpu().numpy().astype(np.float32)
    Qh = Q32[torch.from_numpy(rows).cuda()].cpu().numpy()
    log(f"large: {m} oracle queries on the CPU")
    oi, od = O.bf_knn(Xh, Qh, k + 1, metric=O.COSINE, nthreads=host_cores())

    def measure(kind):
        t1 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        log(f"large[{kind}]: first search {time.perf_counter() - t1:.2f} s, "
            f"stats {M.last_stats()}")
        ts = []
        for _ in range(args.large_reps):
            t1 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
            log(f"large[{kind}]: search {ts[-1]:.3f} s")
        t = sorted(ts)[len(ts) // 2]
        stats = M.last_stats()
        emitted = M.last_emitted()
        check(L.mann_profile_enable(1))
        check(L.mann_profile_reset())
        run()
        kern = kernel_avg_ms(L, check, ["prep_queries", "prep_queries_fp8", "scan_filter",
                                        "scan_dense", "tau_sample", "rescore_bf16",
                                        "select_rerank", "scan_filter_pass2",
                                        "rescore_bf16_pass2", "select_rerank_pass2",
                                        "exact_fallback"])
        check(L.mann_profile_enable(0))
        scan_ms = kern.get("scan_filter", float("nan"))
        peak = MFMA_FP8_PEAK_TFLOPS if kind == "fp8" else MFMA_BF16_PEAK_TFLOPS
        gi = ids.cpu().numpy()[rows]
        gd = dist.cpu().numpy()[rows]
        return {"ms": round(1e3 * t, 2), "qps": round(nq / t, 1),
                "tflops": round(flop / t / 1e12, 2),
                "scan_filter_tflops": round(flop / (scan_ms * 1e-3) / 1e12, 2),
                "scan_filter_frac_peak": round(flop / (scan_ms * 1e-3) / 1e12 / peak, 4),
                "peak_tflops": peak,
                "first_pass_candidates_per_query": round(emitted / nq, 1),
                "reranked_per_query": round(stats[0] / nq, 1),
                "kernels_ms": kern,
                "search_stats": {"reranked": stats[0], "overflow": stats[1],
                                 "second_pass": stats[2], "exact_fallback": stats[3]},
                "precision": precision_report(np, Xh, Qh, k, gi, gd, oi, od, 2)}

    res = {"bf16": measure("bf16")}
    t0 = time.perf_counter()
    ds.set_first_pass("fp8")
    torch.cuda.synchronize()
    fp8_prep_s = time.perf_counter() - t0
    log(f"large: FP8 copy prepared in {fp8_prep_s:.2f} s")
    res["fp8"] = measure("fp8")
    res["fp8"]["fp8_copy_prep_s"] = round(fp8_prep_s, 3)
    best = min(res, key=lambda kk: res[kk]["ms"])
    out = {"workload": f"cosine exact kNN: base {n}x{d} Float16 (unit 5,000-component mixture, "
                       f"noise 0.1), {nq} queries, k={k}",
           "first_pass": best,
           "ms": res[best]["ms"], "qps": res[best]["qps"], "tflops": res[best]["tflops"],
           "fp8_over_bf16": round(res["bf16"]["ms"] / res["fp8"]["ms"], 3),
           "dataset_prep_s": round(prep_s, 3),
           "precision": res[best]["precision"],
           "by_first_pass": res,
           "total_s": round(time.perf_counter() - t_all, 1)}
    del ds, X16, Q32, ids, dist
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the kNN graph with the default knobs")
    ap.add_argument("--skip-ivf", action="store_true")
    ap.add_argument("--skip-hnsw", action="store_true",
                    help="skip the IVF-HNSW key of the IVF workload")
    ap.add_argument("--ivf-search-only", type=int, default=0, metavar="NPROBE",
                    help="IVF: build the index, then only three centroid searches at NPROBE "
                         "(for counter collection)")
    ap.add_argument("--n", type=int, default=60_000)
    ap.add_argument("--d", type=int, default=784)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--ivf-n", type=int, default=1_000_000)
    ap.add_argument("--nlist", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--ivf-nq", type=int, default=10_000)
    ap.add_argument("--train-q", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--cpu-queries", type=int, default=16384,
                    help="self-join queries timed on the CPU oracle (bounded sample)")
    ap.add_argument("--cpu-ivf-queries", type=int, default=4000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration under rocprofv3)")
    ap.add_argument("--skip-large", action="store_true",
                    help="skip the 10M x 256 Float16 cosine search (k=100)")
    ap.add_argument("--large-n", type=int, default=10_000_000)
    ap.add_argument("--large-nq", type=int, default=100_000)
    ap.add_argument("--large-reps", type=int, default=3)
    ap.add_argument("--large-check", type=int, default=16,
                    help="queries checked against the CPU oracle")
    ap.add_argument("--no-prior", action="store_true",
                    help="train the router without the centroid prior")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the search path runs on one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding is a design, not an implementation)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    torch.cuda.init()
    import manifoldann_amd as M
    from manifoldann_amd import datasets
    from manifoldann_amd._lib import check, lib
    L = lib()

    config1 = None if args.skip_cpu else bench_config1(args, M, datasets, L, check, np, torch)
    res, roof, cpu = bench_knn_graph(args, M, datasets, L, check, np, torch)
    ivf = None if args.skip_ivf else bench_ivf(args, M, datasets, L, check, np, torch)
    large = None if args.skip_large else bench_large_cosine(args, M, L, check, np, torch)
    out = {
        "metric": METRIC,
        "value": round(res["value"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate) first pass + fp32 exact re-rank",
        "data": "synthetic (Philox-seeded Gaussian mixtures; no datasets downloaded)",
        "config": {"workload": f"kNN graph self-join n={args.n} d={args.d} k={args.k} L2, "
                               f"exact ids", "n": args.n, "d": args.d, "k": args.k},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms": res["kernels_ms"],
        "autotune": res["autotune"],
        "search_stats": res["stats"],
        "precision": {"exact_knn_10000x128": config1, "knn_graph_60000x784": res["precision"],
                      "large_cosine_10m_256": large["precision"] if large else None},
        "ivf": ivf,
        "large_cosine": large,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
