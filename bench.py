"""Benchmark of the all-GPU ForceAtlas2 iteration on one MI355X.

Workload (the configuration the headline metric is quoted on): an R-MAT
graph of 2^21 nodes and 16,000,000 distinct undirected edges (a=0.57,
b=c=0.19, d=0.05, self-loops and duplicates dropped), unweighted, start
positions uniform in [-10 sqrt(n), 10 sqrt(n))^2, Barnes-Hut theta = 1.2,
linlog and dissuade_hubs on, scaling 2, gravity 1. A step is one FA2
iteration of the whole graph: codes, sort, tree, repulsion, attraction,
speed, displacement. The graph and its state are resident in HBM before the
timed region starts.

Before the timed region the autotuner times its candidate launch shapes
on the graph and keeps the fastest. The timed iterations are replayed from
the captured HIP graph, as a layout runs. The same number of iterations is
then launched one kernel at a time with HIP events on the batch's stream
between the phases: their times give the phase breakdown, the roofline of
the dominant kernel (the Barnes-Hut traversal, FP32 VALU) and the iteration
roofline (per phase, the larger of minimum HBM bytes over 8 TB/s and FLOP
over the FP32 peak, summed, over the measured iteration time), with each
phase's HBM rate.

Printed: one JSON line with node-iters/s, the rooflines and the CPU
baseline (the C restatement of the reference CPU layout, Gephi Barnes-Hut
in FP64 with OpenMP, on a bounded sample of iterations of the same graph).
Beside it, in the same line:
  * "allpairs": the exact all-pairs repulsion kernel of the offload path on
    a 50,000-node preferential-attachment graph, FP32, its FLOP/s against
    the FP32 peak (12 FLOP per ordered pair);
  * "offload": the same size of graph laid out through the repulsion
    offload (the host runs the iteration), exact and Barnes-Hut, each beside
    the oracle's CPU layout with the same algorithm, and through the
    all-GPU path;
  * "batched": 512 stochastic-block-model graphs (200..20000 nodes,
    log-uniform; 2..20 equal blocks; mean in-block degree 8, cross-block 1),
    200 iterations each, laid out as one batch and one at a time;
  * "serving": a Poisson stream of layout requests (preferential-attachment
    or block-model graphs of 1k..1M nodes, log-uniform; mean degree 4..16;
    100..1000 iterations; linlog, prevent_overlap with node sizes in [1, 3],
    5% pinned nodes and double precision mixed in; 200 requests) served
    by the layout server, against the same requests solved one at a time;
    p50/p99 latency and layouts/s, at the offered rate and, for the first
    40 requests, at 3 requests/s (below the server's capacity); and the
    ceiling of the speedup that the size classes allow (each class at the
    rate of one batch of all its requests). Both start from prepared
    incidence lists and include the upload and the download of every
    graph.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The serving leg runs 8 streams at once on the hardware queues HIP gives the
# process (GPU_MAX_HW_QUEUES, 4 by default), which the bench leaves as the
# environment sets them. With 16 queues the server's batch rebuilds waited
# less (lane time 12 -> 5.6 s on the 200-request stream), but the served rate
# did not rise (17.2 against 17.6 layouts/s), and a layout made after the
# server or after a repulsion-offload session in the same process ran up to
# 2.7 times slower: the attraction's stream, forked beside the traversal,
# then shared a hardware queue with the batch's own.

METRIC = ("FA2 node-iters/s at 2M nodes; served layouts/s + p99 vs "
          "one-at-a-time plain")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibration_copy(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime that
    libfa2hip.so itself uses: a dispatch of known size for scaling the HBM
    byte counters of a profiler run (FETCH_SIZE, WRITE_SIZE)."""
    import ctypes
    import forceatlas2_rc_amd  # noqa: F401  (loads the library and its runtime)
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("could not allocate the calibration buffers")
    if hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():  # device to device
        raise RuntimeError("the calibration copy failed")
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy of {nbytes} bytes done")


def sort_yardstick(keys, engine_ms, reps=50):
    """rocPRIM's radix_sort_pairs (libfa2yard.so, built beside the library
    and never part of it) on the same 2^21 (Morton code, node) pairs as the
    engine's sort, all 32 bits: a stated yardstick for the engine's sort
    phase, which also keeps each graph of a batch apart."""
    import ctypes
    path = os.path.join(ROOT, "forceatlas2_rc_amd", "libfa2yard.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    f = lib.fa2yard_radix_sort_pairs
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    keys = np.ascontiguousarray(keys, np.uint32)
    ms, ok = ctypes.c_double(), ctypes.c_int()
    if f(0, int(keys.size), keys.ctypes.data, reps, ctypes.byref(ms), ctypes.byref(ok)):
        return None
    return dict(rocprim_radix_sort_pairs_ms=ms.value, stable_order=bool(ok.value),
                engine_sort_phase_ms=engine_ms, pairs=int(keys.size), reps=reps)


def host_threads():
    """Threads for the CPU baseline: every core this process may run on,
    within the OpenMP thread count the host grants it (OMP_NUM_THREADS, when
    set). Returns the count and what it was taken from."""
    cpus = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = cpus
    omp = os.environ.get("OMP_NUM_THREADS", "")
    omp = int(omp) if omp.isdigit() and int(omp) > 0 else None
    threads = min(allowed, omp) if omp else allowed
    return threads, dict(os_cpu_count=cpus, sched_affinity=allowed, omp_num_threads=omp)


def cpu_baseline(n, e, pos, args):
    """The oracle's CPU layout (Gephi BH, FP64, OpenMP) on the same graph:
    one warm-up iteration, then args.cpu_iters timed ones."""
    from oracle import oracle as O
    threads, host = host_threads()
    kw = dict(scaling=2.0, gravity=1.0, theta=1.2, linlog=True, outbound=True,
              barnes_hut=True, threads=threads)
    p = pos.copy()
    p = O.layout(n, e, p, 1, **kw)
    t0 = time.perf_counter()
    O.layout(n, e, p, args.cpu_iters, **kw)
    dt = time.perf_counter() - t0
    # the host preprocessing of C_forceatlas2_large (mass, incidence lists)
    t1 = time.perf_counter()
    O.large_prepare(n, e, None, True)
    prep = time.perf_counter() - t1
    return dict(value=n * args.cpu_iters / dt, unit="node-iters/s",
                cores=threads, kind="port", host=host,
                sample=(f"{args.cpu_iters} iterations (after 1 warm-up) of the "
                        f"same {n}-node graph, Gephi Barnes-Hut theta=1.2, FP64, "
                        f"{threads} OpenMP threads, {dt:.1f} s"),
                prepare_s=prep)


# Minimum HBM bytes per node of each phase of one FP32 iteration (positions,
# masses, keys, tree records and forces each read or written once), and per
# incidence for the attraction's list entry and position gather: 456n + 16I
# in total. FP64 position accumulators add 16 bytes per node to the move.
STAGE_BYTES_PER_NODE = dict(codes=8 + 16,            # bbox, Morton codes
                            sort=4 * 16,              # 4 LSD passes
                            tree=32 + 32 + 144 + 40,  # payload gather, Karras, summaries, finalize
                            repulsion=28,             # streaming part of the traversal
                            attraction=32,            # per node; + 16 per incidence
                            move=24 + 36)             # swing, apply


def iteration_roofline(n, n_inc, rep_flop, stages, t_iter, flop_peak):
    """t_roof = sum over phases of max(bytes / HBM peak, FLOP / FLOP peak),
    and the HBM rate each phase achieves on its minimum bytes."""
    out, t_roof = {}, 0.0
    for k, per_node in STAGE_BYTES_PER_NODE.items():
        nbytes = per_node * n + (16 * n_inc if k == "attraction" else 0)
        flop = rep_flop if k == "repulsion" else 0
        t_min = max(nbytes / (HBM_PEAK_GBS * 1e9), flop / (flop_peak * 1e12))
        t_roof += t_min
        t = stages[k]
        out[k] = dict(bytes=nbytes, ms=t * 1e3, gbs=nbytes / t / 1e9,
                      hbm_frac=nbytes / t / 1e9 / HBM_PEAK_GBS, roof_ms=t_min * 1e3)
    return dict(t_roof_ms=t_roof * 1e3, t_iter_ms=t_iter * 1e3, frac=t_roof / t_iter,
                bytes_per_iteration=sum(v["bytes"] for v in out.values()),
                hbm_peak_gbs=HBM_PEAK_GBS, stages=out)


def allpairs_leg(args):
    """The all-pairs repulsion of the offload path (fa2_gpu_repulse_exact) on
    a 50,000-node preferential-attachment graph (m = 3), FP32: the kernel's
    time from HIP events on the session's stream, 12 FLOP per ordered pair."""
    import ctypes
    from forceatlas2_rc_amd import graphs
    from forceatlas2_rc_amd._lib import last_error, lib
    n, e = graphs.barabasi_albert(args.allpairs_nodes, 3, seed=args.seed + 6)
    mass = 1.0 + np.bincount(e.ravel(), minlength=n).astype(np.float64)
    rng = np.random.default_rng(args.seed + 7)
    r = 10.0 * np.sqrt(n)
    x = rng.uniform(-r, r, n)
    y = rng.uniform(-r, r, n)
    dx = np.zeros(n)
    dy = np.zeros(n)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    g = lib.fa2_gpu_open(0, 0, n, 0, 0, 1)
    if not g:
        raise RuntimeError(last_error())
    ms = []
    try:
        for it in range(args.allpairs_calls + 2):
            if lib.fa2_gpu_repulse_exact(g, ptr(x), ptr(y), ptr(mass), None, 2.0, ptr(dx),
                                         ptr(dy)) != 0:
                raise RuntimeError(last_error())
            t = ctypes.c_double()
            if lib.fa2_gpu_kernel_time(g, ctypes.byref(t)) != 0:
                raise RuntimeError(last_error())
            if it >= 2:
                ms.append(t.value)
    finally:
        lib.fa2_gpu_close(g)
    k_ms = float(np.mean(ms))
    flop = 12.0 * n * (n - 1)
    achieved = flop / (k_ms * 1e-3) / 1e12
    log(f"all-pairs: {n} nodes, {k_ms:.3f} ms per kernel, {achieved:.1f} TFLOP/s")
    return dict(workload=(f"all-pairs repulsion (fa2_gpu_repulse_exact), preferential attachment "
                          f"n={n}, m=3, FP32"), nodes=n, pairs=n * (n - 1), calls=len(ms),
                kernel="k_exact", kernel_ms=k_ms, flop_per_launch=flop, achieved=achieved,
                peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=achieved / FP32_PEAK_TFLOPS,
                bound="valu")


def offload_leg(args):
    """The C2 workload through the repulsion offload (forceatlas2_gpu: the
    host runs the iteration, the GPU the repulsion): a 50,000-node
    preferential-attachment graph (m = 3), weights uniform in [0.5, 2)
    rounded to float32, linlog off, single precision. Exact all-pairs
    repulsion and Barnes-Hut (theta 1.2, the Gephi quadtree built on the host
    each iteration, the traversal on the GPU), each against the oracle's CPU
    layout with the same algorithm in FP64 on the host's cores (the offload's
    host part uses as many threads); and the same graph through the all-GPU
    path. Bounded samples: --offload-iters
    iterations on the GPU, 2 (exact) and 10 (Barnes-Hut) on the CPU, each
    after one warm-up iteration."""
    import forceatlas2_rc_amd as F
    from forceatlas2_rc_amd import Batch, graphs
    from oracle import oracle as O
    n, e = graphs.barabasi_albert(args.allpairs_nodes, 3, seed=args.seed + 8)
    rng = np.random.default_rng(args.seed + 9)
    w = rng.uniform(0.5, 2.0, len(e)).astype(np.float32).astype(np.float64)
    r = 10.0 * np.sqrt(n)
    pos = rng.uniform(-r, r, size=(n, 2))
    K = args.offload_iters
    threads, _ = host_threads()
    out = {}
    for name, bh in (("exact", False), ("barnes_hut", True)):
        F.layout_forceatlas2_gpu(n, e, 2, weights=w, init=pos, barnes_hut=bh, theta=1.2,
                                 threads=threads)
        t0 = time.perf_counter()
        _, tm = F.layout_forceatlas2_gpu(n, e, K, weights=w, init=pos, barnes_hut=bh,
                                         theta=1.2, threads=threads, timings=True)
        dt = time.perf_counter() - t0
        out[name] = dict(ms_per_iteration=dt / K * 1e3, node_iters_per_s=n * K / dt,
                         host_threads=threads,
                         phase_ms_per_iteration={k: v / K * 1e3 for k, v in tm.items()})
    for name, bh, k in (("exact", False, 2), ("barnes_hut", True, 10)):
        kw = dict(scaling=2.0, gravity=1.0, theta=1.2, barnes_hut=bh, w=w, threads=threads)
        p1 = O.layout(n, e, pos, 1, **kw)
        t0 = time.perf_counter()
        O.layout(n, e, p1, k, **kw)
        dt = time.perf_counter() - t0
        out[name]["cpu_node_iters_per_s"] = n * k / dt
        out[name]["cpu_sample"] = f"{k} iterations after 1 warm-up, FP64, {threads} threads"
        out[name]["gpu_over_cpu"] = out[name]["node_iters_per_s"] / out[name]["cpu_node_iters_per_s"]
    g = dict(n=n, edges=e, weights=w, init=pos, theta=1.2, scaling=2.0, gravity=1.0)
    with Batch([g], autotune=False) as b:
        b.run(2)
        b.sync()
        t0 = time.perf_counter()
        b.run(K)
        b.sync()
        dt = time.perf_counter() - t0
    out["all_gpu_barnes_hut"] = dict(ms_per_iteration=dt / K * 1e3, node_iters_per_s=n * K / dt)
    log("offload:", json.dumps({k: round(v["ms_per_iteration"], 3) for k, v in out.items()}))
    return dict(workload=(f"repulsion offload (forceatlas2_gpu), preferential attachment n={n}, "
                          "m=3, weights U(0.5, 2) as float32, linlog off, FP32; exact and "
                          "Barnes-Hut theta=1.2, and the all-GPU path on the same graph"),
                nodes=n, edges=int(len(e)), iterations=K, **out)


def batched_leg(args):
    """512 SBM graphs, 200 iterations: one batch against one at a time."""
    from forceatlas2_rc_amd import Batch, workloads
    t0 = time.time()
    gs = workloads.sbm_batch(args.batch_graphs, args.batch_iters, args.seed)
    n_tot = sum(g["n"] for g in gs)
    m_tot = sum(len(g["frm"]) for g in gs)
    log(f"batched: {len(gs)} graphs, {n_tot} nodes, {m_tot} edges ({time.time() - t0:.1f} s)")
    it = args.batch_iters
    # warm-up: code objects, allocator
    with Batch(gs[:2]) as b:
        b.run(2)
        b.sync()
    t0 = time.perf_counter()
    b = Batch(gs)
    t_up = time.perf_counter()
    t_open = t_up - t0
    b.run(it)
    b.sync()
    t_run = time.perf_counter()
    res_b = [b.positions(g) for g in range(len(gs))]
    shapes = b.tuning  # chosen by the autotuner at the first run (timed above)
    b.close()
    t_batch = time.perf_counter() - t0
    log(f"batched: {t_batch:.2f} s ({(t_run - t_up) / it * 1e3:.3f} ms per iteration)")
    t0 = time.perf_counter()
    same = 0
    if args.no_batch_alone:  # a profiler's run: the batch only
        gs_alone = []
    else:
        gs_alone = gs
    for g, gr in enumerate(gs_alone):
        with Batch([gr]) as b1:
            b1.run(it)
            b1.sync()
            same += int(np.array_equal(b1.positions(0), res_b[g]))
    t_one = time.perf_counter() - t0 if gs_alone else float("nan")
    log(f"one at a time: {t_one:.2f} s")
    return dict(workload=(f"{len(gs)} SBM graphs, 200..20000 nodes log-uniform, 2..20 blocks, "
                          f"in-block degree 8, cross-block 1, {it} iterations each, BH theta=1.2"),
                graphs=len(gs), nodes=n_tot, edges=m_tot, iterations=it,
                batched_s=t_batch, open_s=t_open,
                batched_ms_per_iteration=(t_run - t_up) / it * 1e3,
                one_at_a_time_s=t_one, layouts_per_s_batched=len(gs) / t_batch,
                layouts_per_s_one_at_a_time=len(gs) / t_one, speedup=t_one / t_batch,
                node_iters_per_s_batched=n_tot * it / t_batch, launch_shapes=shapes,
                identical_results=f"{same}/{len(gs_alone)}")


def serving_requests(args):
    from forceatlas2_rc_amd import workloads
    return workloads.serving_requests(args.serve_requests, args.serve_rate,
                                      args.serve_max_nodes, args.seed)


def serve_stream(reqs, arrive, args):
    """Submit the requests at their arrival times; per-request latency from
    arrival to result."""
    from forceatlas2_rc_amd import Server
    with Server(max_batches=args.serve_batches, chunk=args.serve_chunk) as srv:
        ids = []
        t0 = time.perf_counter()
        for (g, prec, _), ta in zip(reqs, arrive):
            wait = t0 + ta - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(g, prec))
        out = [srv.result(j) for j in ids]
        span = time.perf_counter() - t0
        stats = srv.stats()
    lat = np.array([t["total"] for _, t in out])
    log(f"served at {len(reqs) / arrive[-1] if arrive[-1] > 0 else 0:.1f} requests/s offered: "
        f"{span:.2f} s, p50 {np.percentile(lat, 50):.2f} s, p99 {np.percentile(lat, 99):.2f} s, "
        f"{stats}")
    return out, span, lat, stats


def queue_latency(arrive, svc):
    """Latencies of the same arrivals served by one solver in turn."""
    fin, lat = 0.0, []
    for ta, s_ in zip(arrive, svc):
        fin = max(fin, ta) + s_
        lat.append(fin - ta)
    return np.array(lat)


SIZE_CLASSES = ((1_000, 10_000), (10_000, 100_000), (100_000, 1 << 30))


def class_ceiling(reqs, svc, iters=40):
    """The speedup no scheduler of these requests can pass while each size
    class runs at the rate of its own best batch: per (size class,
    precision), the node-iterations per second of one batch holding all of
    that class's requests (timed over `iters` iterations after 2 warm-up
    ones), against the same requests' rate one at a time (their measured
    service times). Returns the per-class rates and the ceiling."""
    from forceatlas2_rc_amd import Batch
    out, t_best = [], 0.0
    for lo, hi in SIZE_CLASSES:
        for prec in ("single", "double"):
            sel = [k for k, (g, p, _) in enumerate(reqs) if lo <= g["n"] < hi and p == prec]
            if not sel:
                continue
            work = sum(reqs[k][0]["n"] * reqs[k][0]["iterations"] for k in sel)
            alone_s = float(sum(svc[k] for k in sel))
            gs = [dict(reqs[k][0], iterations=1 << 30) for k in sel]
            with Batch(gs, precision=prec, autotune=False) as b:
                b.run(2)
                b.sync()
                t0 = time.perf_counter()
                b.run(iters)
                b.sync()
                dt = time.perf_counter() - t0
            rate = sum(g["n"] for g in gs) * iters / dt
            t_best += work / rate
            out.append(dict(nodes=[lo, hi], precision=prec, requests=len(sel),
                            node_iterations=work, batched_node_iters_per_s=rate,
                            one_at_a_time_node_iters_per_s=work / alone_s))
    return out, float(sum(svc)) / t_best


def serving_leg(args):
    """A Poisson stream of layout requests through the server, against the
    same requests solved one at a time: at the offered rate (above capacity,
    so the stream saturates the server) and, with the first requests, at a
    rate below capacity; then the size-class ceiling of the speedup."""
    from forceatlas2_rc_amd import Batch
    t0 = time.time()
    reqs, arrive = serving_requests(args)
    n_tot = sum(g["n"] for g, _, _ in reqs)
    work = sum(g["n"] * g["iterations"] for g, _, _ in reqs)
    log(f"serving: {len(reqs)} requests, {n_tot} nodes, {work:.3g} node-iterations "
        f"({time.time() - t0:.1f} s)")
    with Batch([reqs[0][0]]) as b:  # warm-up
        b.run(2)
        b.sync()
    out, span, lat, stats = serve_stream(reqs, arrive, args)
    n_sub = min(len(reqs), args.serve_sub_requests)
    arrive_sub = arrive[:n_sub] * (args.serve_rate / args.serve_sub_rate)
    out_sub, span_sub, lat_sub, stats_sub = serve_stream(reqs[:n_sub], arrive_sub, args)
    # one at a time, in arrival order, as fa2_full_run lays out one graph
    # (default launch shapes: no autotuning, which the server does not do
    # either)
    svc, same = [], 0
    for k, (g, prec, _) in enumerate(reqs):
        t1 = time.perf_counter()
        with Batch([g], precision=prec, autotune=False) as b:
            b.run(g["iterations"])
            b.sync()
            p = b.positions(0)
        svc.append(time.perf_counter() - t1)
        same += int(np.array_equal(p, out[k][0]) and
                    (k >= n_sub or np.array_equal(p, out_sub[k][0])))
        if k % 50 == 49:
            log(f"one at a time: {k + 1} requests")
    svc = np.array(svc)
    lat1 = queue_latency(arrive, svc)
    lat1_sub = queue_latency(arrive_sub, svc[:n_sub])
    log(f"one at a time: {svc.sum():.2f} s of service")
    classes, ceiling = class_ceiling(reqs, svc)
    log(f"size-class ceiling of the speedup: {ceiling:.2f}")
    pct = lambda a, q: float(np.percentile(a, q))
    return dict(workload=(f"Poisson stream, {args.serve_rate:g} requests/s, {len(reqs)} "
                          f"requests; preferential-attachment or SBM graphs of 1000.."
                          f"{args.serve_max_nodes} nodes (log-uniform), mean degree 4..16, "
                          "100..1000 iterations; linlog, prevent_overlap (sizes 1..3), 5% "
                          "pinned, dissuade_hubs and double precision mixed in"),
                requests=len(reqs), nodes=n_tot, node_iterations=work,
                layouts_per_s=len(reqs) / span, p50_s=pct(lat, 50), p99_s=pct(lat, 99),
                node_iters_per_s=work / span,
                one_at_a_time=dict(layouts_per_s=len(reqs) / svc.sum(),
                                   p50_s=pct(lat1, 50), p99_s=pct(lat1, 99),
                                   service_s=float(svc.sum()),
                                   node_iters_per_s=work / svc.sum()),
                speedup=(len(reqs) / span) / (len(reqs) / svc.sum()),
                size_class_ceiling=dict(speedup=ceiling, classes=classes),
                batches=stats["batches"], refills=stats["refills"],
                refill_s=stats["refill_s"], uncaptured_batches=stats["uncaptured"],
                hw_queues=os.environ.get("GPU_MAX_HW_QUEUES"),
                max_batches_running=stats["peak_running"],
                below_capacity=dict(rate=args.serve_sub_rate, requests=n_sub, span_s=span_sub,
                                    layouts_per_s=n_sub / span_sub,
                                    p50_s=pct(lat_sub, 50), p99_s=pct(lat_sub, 99),
                                    one_at_a_time_p50_s=pct(lat1_sub, 50),
                                    one_at_a_time_p99_s=pct(lat1_sub, 99),
                                    batches=stats_sub["batches"], refills=stats_sub["refills"],
                                    max_batches_running=stats_sub["peak_running"]),
                identical_results=f"{same}/{len(reqs)}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--layout-iters", type=int, default=300,
                    help="continue the timed layout to this many iterations, timed by chunks")
    ap.add_argument("--scale", type=int, default=21)
    ap.add_argument("--edges", type=int, default=16_000_000)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--precision", default="single")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-allpairs", action="store_true")
    ap.add_argument("--allpairs-nodes", type=int, default=50_000)
    ap.add_argument("--allpairs-calls", type=int, default=10)
    ap.add_argument("--no-offload", action="store_true")
    ap.add_argument("--offload-iters", type=int, default=20)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-graphs", type=int, default=512)
    ap.add_argument("--batch-iters", type=int, default=200)
    ap.add_argument("--no-batch-alone", action="store_true",
                    help="skip laying out each graph of the batch alone")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-yardstick", action="store_true",
                    help="skip timing rocPRIM's radix sort on the C3 codes")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-sub-requests", type=int, default=40,
                    help="requests of the stream replayed below capacity")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-sub-rate", type=float, default=3.0)
    ap.add_argument("--serve-max-nodes", type=int, default=1_000_000)
    ap.add_argument("--serve-batches", type=int, default=8)
    ap.add_argument("--serve-chunk", type=int, default=25)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first copy this many bytes device to device (a dispatch of known "
                         "size for a profiler's HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU "
              "path (each graph of a batch lives on one device)",
              file=sys.stderr)
        sys.exit(2)

    from forceatlas2_rc_amd import Batch, workloads
    from forceatlas2_rc_amd.api import prepare_large

    if args.copy_bytes:
        calibration_copy(args.copy_bytes)
    t0 = time.time()
    n, e, pos = workloads.rmat_large(args.scale, args.edges, args.seed)
    log(f"graph n={n} m={len(e)} ({time.time() - t0:.1f} s)")
    t0 = time.perf_counter()
    mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1],
                                          np.ones(len(e)), True)
    t_prep = time.perf_counter() - t0
    log(f"host preprocessing {t_prep:.2f} s, incidences={int(off[-1])}")
    graph = dict(n=n, frm=e[:, 0], to=e[:, 1], mass=mass, inc_off=off,
                 inc_node=node, inc_coef=coef, init=pos, theta=1.2,
                 linlog=True, dissuade_hubs=True, scaling=2.0, gravity=1.0)
    raw = {k: v for k, v in graph.items() if k not in ("mass", "inc_off", "inc_node",
                                                         "inc_coef")}
    # preprocessing on the device (edge upload included), then prepared lists
    small = dict(n=64, frm=np.arange(63, dtype=np.int32), to=np.arange(1, 64, dtype=np.int32),
                 init=pos[:64])
    with Batch([small], device=0, precision=args.precision, ingest="device"):  # warm-up
        pass
    t0 = time.perf_counter()
    b0 = Batch([raw], device=0, precision=args.precision, ingest="device")
    t_dev = time.perf_counter() - t0
    b0.close()
    # the timed region: iterations replayed from the captured HIP graph
    t0 = time.perf_counter()
    b = Batch([graph], device=0, precision=args.precision, autotune=False)
    t_up = time.perf_counter() - t0
    log(f"uploaded prepared lists in {t_up:.2f} s; device preprocessing and upload {t_dev:.2f} s")
    # launch shapes: the autotuner times every candidate on this graph
    shapes, shape_ms = b.tune(force=True)
    log(f"autotuner: {shapes}, candidate ms {np.round(shape_ms, 4).tolist()}")
    b.run(args.warmup)
    b.sync()
    log("warm-up done")
    t0 = time.perf_counter()
    b.run(args.steps)
    b.sync()
    dt = time.perf_counter() - t0
    pos_replay = b.positions(0)
    ms = dt / args.steps * 1e3
    value = n * args.steps / dt
    log(f"timed: {args.steps} iterations, {ms:.4f} ms per iteration")
    # the rest of the layout (to --layout-iters iterations in all), timed by
    # chunks: how the cost of an iteration moves as the layout converges
    progression, done_it = [], args.warmup + args.steps
    while done_it < args.layout_iters:
        k = min(25, args.layout_iters - done_it)
        t1 = time.perf_counter()
        b.run(k)
        b.sync()
        progression.append(dict(iterations=[done_it + 1, done_it + k],
                                ms_per_iteration=(time.perf_counter() - t1) / k * 1e3))
        done_it += k
    b.close()
    if progression:
        log("layout progression (ms per iteration):",
            [round(p_["ms_per_iteration"], 3) for p_ in progression])

    # the same iterations launched one kernel at a time with HIP events
    # between the phases: phase times and the traversal's duration
    b = Batch([graph], device=0, precision=args.precision, profile=True, autotune=False)
    b.set_tuning(**shapes)
    b.run(args.warmup)
    b.sync()
    base = b.stage_times()
    t0 = time.perf_counter()
    b.run(args.steps)
    b.sync()
    dt_prof = time.perf_counter() - t0
    st = b.stage_times()
    stages = {k: (st[k] - base[k]) / args.steps for k in st if k != "total"}
    log("per-iteration device time by phase (ms):",
        {k: round(v * 1e3, 4) for k, v in stages.items()})
    work = b.count_work()
    pos_out = b.positions(0)
    finite = bool(np.all(np.isfinite(pos_out)))
    yard = None
    if not args.no_yardstick:
        tr = b.tree(0)
        keys_in = np.empty(n, np.uint32)  # the codes in node order, as the sort takes them
        keys_in[tr["order"].astype(np.int64)] = tr["keys"]
        yard = sort_yardstick(keys_in, stages["sort"] * 1e3)
        log("sort yardstick:", yard)
    b.close()

    A, O_, P, steps = work
    flops = 6 * (A + O_) + 7 * A + 12 * P
    t_rep = stages["repulsion"]
    dbl = args.precision == "double"
    peak = FP64_PEAK_TFLOPS if dbl else FP32_PEAK_TFLOPS
    achieved = flops / t_rep / 1e12
    n_inc = int(off[-1])
    iter_roof = iteration_roofline(n, n_inc, flops, stages, ms * 1e-3, peak)
    log("iteration roofline:", json.dumps(iter_roof))
    traffic = None
    # the counter-measured bytes of the single-precision traversal at this
    # workload's default size (none are recorded for other runs)
    tpath = os.path.join(ROOT, "profiles", "repulse_rmat21_traffic.json")
    if (args.precision == "single" and args.scale == 21 and args.edges == 16_000_000
            and os.path.exists(tpath)):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof = dict(kernel="k_repulse (Barnes-Hut traversal)", bound="valu",
                achieved=achieved, peak=peak, unit="TFLOP/s",
                frac=achieved / peak, traffic=traffic,
                flop_per_launch=flops, launch_ms=t_rep * 1e3,
                work=dict(accepted=A, opened=O_, pairs=P, wave_steps=steps,
                          lane_steps_per_wave_step=(A + O_) / max(1, steps)))
    out = dict(metric=METRIC, value=value, unit="node-iters/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=ms,
               higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64" if dbl else "fp32",
               data="synthetic",
               config=dict(workload=(f"R-MAT scale {args.scale}, {len(e)} "
                                     "undirected edges, all-GPU FA2, BH "
                                     "theta=1.2, linlog, dissuade_hubs"),
                           nodes=n, edges=int(len(e)), precision=args.precision,
                           timed_path="HIP graph replay (two iterations per graph)",
                           launch_shapes=dict(chosen=shapes, candidate_ms=dict(
                               repulse_threads_64_to_1024=shape_ms[:5].tolist(),
                               repulse_xcd_run_0_8_32_128_512=shape_ms[5:10].tolist(),
                               attract_lane_unroll_1_2_4_8=shape_ms[10:14].tolist(),
                               attract_16lane_unroll_1_2_4_8=shape_ms[14:18].tolist(),
                               attract_wave_unroll_1_2_4_8=shape_ms[18:22].tolist(),
                               attract_fork_0_1_2_3=shape_ms[22:26].tolist(),
                               beside_attraction_threads_chosen_256=shape_ms[26:28].tolist())),
                           stage_ms={k: v * 1e3 for k, v in stages.items()},
                           profiled_ms_per_step=dt_prof / args.steps * 1e3,
                           replay_equals_eager=bool(np.array_equal(pos_replay, pos_out)),
                           layout_progression=progression,
                           positions_finite=finite,
                           sort_yardstick=yard,
                           ingest=dict(host_numpy_prepare_s=t_prep, upload_prepared_s=t_up,
                                       device_prepare_with_upload_s=t_dev)),
               roofline=roof, iteration_roofline=iter_roof)
    if not args.no_cpu:
        log("cpu baseline...")
        out["cpu_baseline"] = cpu_baseline(n, e, pos, args)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    del graph, mass, off, node, coef
    if not args.no_allpairs:
        out["allpairs"] = allpairs_leg(args)
    if not args.no_offload:
        out["offload"] = offload_leg(args)
    if not args.no_batched:
        out["batched"] = batched_leg(args)
    if not args.no_serving:
        out["serving"] = serving_leg(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
