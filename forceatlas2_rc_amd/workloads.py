"""The benchmark workloads, seeded, shared by the benchmark and the tests.

* ``rmat_large``: one R-MAT graph of 2^21 nodes and 16,000,000 distinct
  undirected edges (a=0.57, b=c=0.19, d=0.05), unweighted, start positions
  uniform in [-10 sqrt(n), 10 sqrt(n))^2, laid out with Barnes-Hut
  theta=1.2, linlog and dissuade_hubs.
* ``sbm_batch``: stochastic-block-model graphs of 200..20000 nodes
  (log-uniform), 2..20 equal blocks, mean in-block degree 8 and cross-block
  degree 1, theta=1.2, to be laid out together.
* ``serving_requests``: a Poisson stream of layout requests, preferential
  attachment or block-model graphs of 1000..max_nodes nodes (log-uniform),
  mean degree 4..16, 100..1000 iterations, with linlog, prevent_overlap
  (node sizes uniform in [1, 3]), 5% pinned nodes, dissuade_hubs and double
  precision mixed in.

Every graph dict carries its prepared mass and incidence lists
(``prepare_large``), as a caller that prepares them once would.
"""
import os

import numpy as np

from . import graphs
from .api import prepare_large

__all__ = ["rmat_large", "sbm_batch", "serving_requests"]


def rmat_large(scale=21, n_edges=16_000_000, seed=2026, cache_dir="/tmp"):
    """(n, edges, positions) of the R-MAT workload. The edge list is cached
    under cache_dir (it takes about 20 s to draw at full size)."""
    path = None if cache_dir is None else os.path.join(
        cache_dir, f"fa2_rmat_{scale}_{n_edges}_{seed}.npy")
    e = None
    if path and os.path.exists(path):
        try:
            e = np.load(path)
        except (OSError, ValueError):
            e = None
    if e is None:
        _, e = graphs.rmat(scale, n_edges, seed=seed)
        if path:
            try:
                np.save(path, e)
            except OSError:
                pass
    n = 1 << scale
    rng = np.random.default_rng(seed)
    r = 10.0 * np.sqrt(n)
    pos = rng.uniform(-r, r, size=(n, 2))
    return n, e, pos


def sbm_batch(count=512, iterations=200, seed=2026):
    """Graph dicts of the batch workload (prepared lists, theta 1.2)."""
    rng = np.random.default_rng(seed + 4)
    out = []
    for g in range(count):
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(20000)))))
        k = int(rng.integers(2, 21))
        n, e = graphs.sbm(n, k, deg_in=8.0, deg_out=1.0, seed=seed + 1000 + g)
        r = 10.0 * np.sqrt(n)
        pos = rng.uniform(-r, r, size=(n, 2))
        mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1], np.ones(len(e)), False)
        out.append(dict(n=n, frm=e[:, 0], to=e[:, 1], mass=mass, inc_off=off, inc_node=node,
                        inc_coef=coef, init=pos, theta=1.2, iterations=iterations))
    return out


def serving_requests(count=32, rate=50.0, max_nodes=1_000_000, seed=2026,
                     iterations=(100, 1000)):
    """(requests, arrival times): each request is (graph dict, precision,
    kind); arrivals are a Poisson process of `rate` per second starting at 0.
    `iterations` is the range of iteration counts (inclusive)."""
    rng = np.random.default_rng(seed + 5)
    reqs = []
    for q in range(count):
        n = int(round(np.exp(rng.uniform(np.log(1000), np.log(max_nodes)))))
        deg = rng.uniform(4.0, 16.0)
        if rng.random() < 0.5:
            n, e = graphs.preferential_attachment(n, max(1, int(round(deg / 2))),
                                                  seed=seed + 2000 + q)
            kind = "pa"
        else:
            n, e = graphs.sbm(n, int(rng.integers(2, 21)), deg_in=deg - 1.0, deg_out=1.0,
                              seed=seed + 2000 + q)
            kind = "sbm"
        r = 10.0 * np.sqrt(n)
        pos = rng.uniform(-r, r, size=(n, 2))
        g = dict(n=n, frm=e[:, 0], to=e[:, 1], init=pos, theta=1.2,
                 iterations=int(rng.integers(iterations[0], iterations[1] + 1)))
        if rng.random() < 0.5:
            g["linlog"] = True
        if rng.random() < 0.3:
            g["adjust"] = True
            g["size"] = rng.uniform(1.0, 3.0, n)
        if rng.random() < 0.5:
            g["fixed"] = rng.random(n) < 0.05
        outbound = rng.random() < 0.3
        g["dissuade_hubs"] = outbound
        mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1], np.ones(len(e)), outbound)
        g.update(mass=mass, inc_off=off, inc_node=node, inc_coef=coef)
        prec = "double" if rng.random() < 0.25 else "single"
        reqs.append((g, prec, kind))
    gaps = rng.exponential(1.0 / rate, len(reqs))
    gaps[0] = 0.0
    return reqs, np.cumsum(gaps)
