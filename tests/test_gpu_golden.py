"""The GPU paths against the golden fixtures of tests/golden (written by the
oracle, tests/golden/make_fixtures.py).

Gates: the large path's Morton keys, permutation and Karras tree bit-exact
on the duplicate-key case; the all-pairs offload kernel within 1e-5
normwise of the exact FP64 forces in single precision and 1e-12 in double
(C2, 50,000 nodes); the offload layout of C1 in double within 1e-9 of the
CPU after 10 iterations (test-gpu.R:34-58)."""
import ctypes
import os

import numpy as np
import pytest

import forceatlas2_rc_amd as F
from forceatlas2_rc_amd import Batch
from forceatlas2_rc_amd._lib import last_error, lib

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_duplicate_keys_tree():
    f = load("dupkeys_700.npz")
    pos = f["pos"].astype(np.float64)
    n = len(pos)
    ring = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    with Batch([dict(n=n, edges=ring, init=pos, iterations=1)]) as b:
        b.run(1)
        b.sync()
        t = b.tree(0)
    for k in ("keys", "order", "child", "range", "parent"):
        assert np.array_equal(t[k].astype(np.int64), f[k].astype(np.int64)), k


@pytest.mark.parametrize("use_double,tol", [(0, 1e-5), (1, 1e-12)])
def test_c2_exact_forces(use_double, tol):
    f = load("c2_exact_4096.npz")
    pos, mass, idx, force = f["pos"], f["mass"].astype(np.float64), f["idx"], f["force"]
    n = len(pos)
    x = np.ascontiguousarray(pos[:, 0], np.float64)
    y = np.ascontiguousarray(pos[:, 1], np.float64)
    dx, dy = np.zeros(n), np.zeros(n)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    g = lib.fa2_gpu_open(0, use_double, n, 0, 0, 1)
    assert g, last_error()
    try:
        assert lib.fa2_gpu_repulse_exact(g, ptr(x), ptr(y), ptr(mass), None, 2.0, ptr(dx),
                                         ptr(dy)) == 0, last_error()
        t = ctypes.c_double()
        assert lib.fa2_gpu_kernel_time(g, ctypes.byref(t)) == 0
        assert t.value > 0
    finally:
        lib.fa2_gpu_close(g)
    got = np.stack([dx[idx], dy[idx]], 1)
    assert normwise(got, force) <= tol


def test_c1_islands_offload_double():
    f = load("c1_islands.npz")
    n = len(f["init"])
    got = F.layout_forceatlas2_gpu(n, f["edges"], iterations=10, init=f["init"],
                                   scaling_ratio=2.0, barnes_hut=False, precision="double")
    assert np.abs(got - f["pos10"]).mean() / np.abs(f["pos10"]).mean() < 1e-9


def edge_length_ratio(xy, edges):
    """Mean edge length over the layout's standard deviation: the scale-free
    measure of a final layout that the long runs are compared by."""
    d = np.hypot(*(xy[edges[:, 0]] - xy[edges[:, 1]]).T)
    return d.mean() / np.std(xy, ddof=1)


@pytest.mark.parametrize("path", ["offload_single", "offload_double", "large_single",
                                  "large_double"])
def test_c1_islands_1000_iterations_quality(path):
    """The full 1000-iteration layout of the islands graph (5 x 30 nodes,
    exact repulsion, scaling 2, gravity 1), which diverges chaotically from
    the CPU's after tens of iterations: its edge-length ratio must be within
    2% of the golden CPU layout's (FP64, exact). The large path uses
    Barnes-Hut theta = 1.2 as the reference's forceatlas2_large does; the
    offload uses the exact all-pairs kernel."""
    f = load("c1_islands.npz")
    n, e = len(f["init"]), f["edges"]
    prec = "double" if path.endswith("double") else "single"
    if path.startswith("offload"):
        got = F.layout_forceatlas2_gpu(n, e, iterations=1000, init=f["init"],
                                       scaling_ratio=2.0, barnes_hut=False, precision=prec)
    else:
        got = F.layout_forceatlas2_large(n, e, iterations=1000, init=f["init"],
                                         scaling_ratio=2.0, precision=prec)
    assert np.all(np.isfinite(got))
    want = edge_length_ratio(f["pos1000"], e)
    assert edge_length_ratio(got, e) == pytest.approx(want, rel=0.02)


def test_rmat_300_iterations_quality():
    """R-MAT scale 14 (120,000 edges), linlog and dissuade_hubs, 300
    iterations on the all-GPU path against the CPU layout (Gephi Barnes-Hut,
    theta 1.2, FP64): edge-length ratio within 2%, spread within 5%."""
    from forceatlas2_rc_amd import graphs
    from oracle import oracle as O
    n, e = graphs.rmat(14, 120_000, seed=41)
    r = 10 * np.sqrt(n)
    pos = graphs.uniform_positions(n, -r, r, 42)
    got = F.layout_forceatlas2_large(n, e, 300, init=pos, linlog=True, dissuade_hubs=True)
    cpu = O.layout(n, e, pos, 300, linlog=True, outbound=True, barnes_hut=True, theta=1.2,
                   threads=8)
    assert np.all(np.isfinite(got))
    assert edge_length_ratio(got, e) == pytest.approx(edge_length_ratio(cpu, e), rel=0.02)
    assert np.std(got, ddof=1) == pytest.approx(np.std(cpu, ddof=1), rel=0.05)
