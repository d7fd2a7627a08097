"""Benchmark of the all-GPU ForceAtlas2 iteration on one MI355X.

Workload (the configuration the headline metric is quoted on): an R-MAT
graph of 2^21 nodes and 16,000,000 distinct undirected edges (a=0.57,
b=c=0.19, d=0.05, self-loops and duplicates dropped), unweighted, start
positions uniform in [-10 sqrt(n), 10 sqrt(n))^2, Barnes-Hut theta = 1.2,
linlog and dissuade_hubs on, scaling 2, gravity 1. A step is one FA2
iteration of the whole graph: codes, sort, tree, repulsion, attraction,
speed, displacement. The graph and its state are resident in HBM before the
timed region starts.

Printed: one JSON line with node-iters/s, the roofline of the dominant
kernel (timed with HIP events on the batch's stream over the timed region)
and the CPU baseline (the C restatement of the reference CPU layout,
Gephi Barnes-Hut in FP64 with OpenMP, on a bounded sample of iterations of
the same graph).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("FA2 node-iters/s at 2M nodes; served layouts/s + p99 vs "
          "one-at-a-time plain")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_graph(scale, n_edges, seed):
    from forceatlas2_rc_amd import graphs
    cache = f"/tmp/fa2_rmat_{scale}_{n_edges}_{seed}.npy"
    if os.path.exists(cache):
        e = np.load(cache)
        return 1 << scale, e
    n, e = graphs.rmat(scale, n_edges, seed=seed)
    try:
        np.save(cache, e)
    except OSError:
        pass
    return n, e


def cpu_baseline(n, e, pos, args):
    """The oracle's CPU layout (Gephi BH, FP64, OpenMP) on the same graph:
    one warm-up iteration, then args.cpu_iters timed ones."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    kw = dict(scaling=2.0, gravity=1.0, theta=1.2, linlog=True, outbound=True,
              barnes_hut=True, threads=threads)
    p = pos.copy()
    p = O.layout(n, e, p, 1, **kw)
    t0 = time.perf_counter()
    O.layout(n, e, p, args.cpu_iters, **kw)
    dt = time.perf_counter() - t0
    return dict(value=n * args.cpu_iters / dt, unit="node-iters/s",
                cores=threads, kind="port",
                sample=(f"{args.cpu_iters} iterations (after 1 warm-up) of the "
                        f"same {n}-node graph, Gephi Barnes-Hut theta=1.2, FP64, "
                        f"{threads} OpenMP threads, {dt:.1f} s"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=21)
    ap.add_argument("--edges", type=int, default=16_000_000)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--precision", default="single")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many "
                         "bytes (a known-size reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU "
              "path (each graph of a batch lives on one device)",
              file=sys.stderr)
        sys.exit(2)

    from forceatlas2_rc_amd import Batch
    from forceatlas2_rc_amd.api import prepare_large

    if args.copy_bytes:
        import torch
        src = torch.empty(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    t0 = time.time()
    n, e = make_graph(args.scale, args.edges, args.seed)
    rng = np.random.default_rng(args.seed)
    r = 10.0 * np.sqrt(n)
    pos = rng.uniform(-r, r, size=(n, 2))
    mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1],
                                          np.ones(len(e)), True)
    log(f"graph n={n} m={len(e)} incidences={int(off[-1])} "
        f"({time.time() - t0:.1f} s)")
    graph = dict(n=n, frm=e[:, 0], to=e[:, 1], mass=mass, inc_off=off,
                 inc_node=node, inc_coef=coef, init=pos, theta=1.2,
                 linlog=True, dissuade_hubs=True, scaling=2.0, gravity=1.0)
    b = Batch([graph], device=0, precision=args.precision, profile=True)
    log("uploaded")
    b.run(args.warmup)
    b.sync()
    base = b.stage_times()
    log("warm-up done")

    t0 = time.perf_counter()
    b.run(args.steps)
    b.sync()
    dt = time.perf_counter() - t0
    st = b.stage_times()
    stages = {k: (st[k] - base[k]) / args.steps for k in st if k != "total"}
    log("per-iteration device time by phase (ms):",
        {k: round(v * 1e3, 4) for k, v in stages.items()})
    work = b.count_work()
    pos_out = b.positions(0)
    finite = bool(np.all(np.isfinite(pos_out)))
    b.close()

    ms = dt / args.steps * 1e3
    value = n * args.steps / dt
    A, O_, P = work
    flops = 6 * (A + O_) + 7 * A + 12 * P
    t_rep = stages["repulsion"]
    dbl = args.precision == "double"
    peak = FP64_PEAK_TFLOPS if dbl else FP32_PEAK_TFLOPS
    achieved = flops / t_rep / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "repulse_rmat21_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof = dict(kernel="k_repulse (Barnes-Hut traversal)", bound="valu",
                achieved=achieved, peak=peak, unit="TFLOP/s",
                frac=achieved / peak, traffic=traffic,
                flop_per_launch=flops, launch_ms=t_rep * 1e3,
                work=dict(accepted=A, opened=O_, pairs=P))
    out = dict(metric=METRIC, value=value, unit="node-iters/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=ms,
               higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64" if dbl else "fp32",
               data="synthetic",
               config=dict(workload=(f"R-MAT scale {args.scale}, {len(e)} "
                                     "undirected edges, all-GPU FA2, BH "
                                     "theta=1.2, linlog, dissuade_hubs"),
                           nodes=n, edges=int(len(e)), precision=args.precision,
                           stage_ms={k: v * 1e3 for k, v in stages.items()},
                           positions_finite=finite),
               roofline=roof)
    if not args.no_cpu:
        log("cpu baseline...")
        out["cpu_baseline"] = cpu_baseline(n, e, pos, args)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
