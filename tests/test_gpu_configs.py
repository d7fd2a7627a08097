"""The benchmark workloads themselves, against the oracle.

* R-MAT 2^21 nodes / 16,000,000 edges (linlog, dissuade_hubs, theta 1.2):
  Morton codes, sorted permutation and Karras topology bit-exact against the
  reference large path restated on the CPU on the same float32 positions
  (pipeline.cl:105-270); one iteration's forces within 1e-5 normwise of the
  FP64 oracle, over the graph and within each attraction class; iterations
  replayed from the captured HIP graph (with the autotuner's shapes) equal to
  kernel-by-kernel launches bit for bit; finite positions. In double
  precision the tree is bit-exact and the forces within 1e-12 of the FP64
  oracle.
* 512 stochastic-block-model graphs in one batch: each graph's layout alone
  equals its layout in the batch bit for bit, and the smallest, median and
  largest graphs' first-iteration forces in the batch are within 1e-5 of the
  oracle's.
* a served stream of block-model and preferential-attachment graphs of up to
  200,000 nodes with linlog, prevent_overlap (sizes 1..3), 5% pinned nodes,
  dissuade_hubs and double precision mixed in: each result equals the
  layout alone bit for bit; for a subset, the layout alone's first-iteration
  force is within 1e-5 (single) or 1e-12 (double) of the oracle's.
"""
import numpy as np
import pytest

from forceatlas2_rc_amd import Batch, Server, workloads
from forceatlas2_rc_amd.api import prepare_large
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = [(1, 16), (17, 256), (257, 1024), (1025, 1 << 30)]


def normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def rmat21():
    n, e, pos = workloads.rmat_large()
    mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1], np.ones(len(e)), True)
    g = dict(n=n, frm=e[:, 0], to=e[:, 1], mass=mass, inc_off=off, inc_node=node,
             inc_coef=coef, init=pos, theta=1.2, linlog=True, dissuade_hubs=True)
    return n, e, pos, g, np.diff(off)


def test_rmat21_tree_and_forces(rmat21):
    n, e, pos, g, deg = rmat21
    assert n == 1 << 21 and len(e) == 16_000_000
    with Batch([g], autotune=False) as b:
        b.run(1)
        b.sync()
        tree = b.tree(0)
        f = b.forces(0)
        sp = b.speed(0)
    p32 = pos.astype(np.float32)
    ref = O.large_tree(p32[:, 0], p32[:, 1], "f")
    for k in ("keys", "order", "child", "range", "parent"):
        assert np.array_equal(tree[k].astype(np.int64), ref[k].astype(np.int64)), k
    del tree, ref
    _, tr = O.large_run(n, e, pos, 1, mode="m", theta=1.2, linlog=True, outbound=True,
                        trace=True)
    fr = tr["force"][0]
    assert normwise(f, fr) <= 1e-5
    for lo, hi in CLASSES:
        sel = (deg >= lo) & (deg <= hi)
        assert sel.any()
        assert normwise(f[sel], fr[sel]) <= 1e-5, (lo, hi)
    iso = deg == 0  # half the nodes have no edges: repulsion and gravity only
    assert iso.sum() > n // 4 and normwise(f[iso], fr[iso]) <= 1e-5
    assert sp == pytest.approx(tr["speed"][0], rel=1e-5)


def test_rmat21_double_tree_and_forces(rmat21):
    """The same workload in double precision: the Morton codes of the FP64
    positions, the permutation and the topology bit-exact, and one
    iteration's forces within 1e-12 normwise of the FP64 oracle (mode d),
    over the graph and within each attraction class."""
    n, e, pos, g, deg = rmat21
    with Batch([g], precision="double", autotune=False) as b:
        b.run(1)
        b.sync()
        tree = b.tree(0)
        f = b.forces(0)
    ref = O.large_tree(pos[:, 0], pos[:, 1], "d")
    for k in ("keys", "order", "child", "range", "parent"):
        assert np.array_equal(tree[k].astype(np.int64), ref[k].astype(np.int64)), k
    del tree, ref
    _, tr = O.large_run(n, e, pos, 1, mode="d", theta=1.2, linlog=True, outbound=True,
                        trace=True)
    fr = tr["force"][0]
    assert normwise(f, fr) <= 1e-12
    for lo, hi in CLASSES:
        sel = (deg >= lo) & (deg <= hi)
        assert normwise(f[sel], fr[sel]) <= 1e-12, (lo, hi)


def test_rmat21_replay_equals_eager(rmat21):
    *_, g, _ = rmat21
    with Batch([g]) as replay:  # tuned shapes, iterations replayed from the HIP graph
        replay.run(6)
        replay.sync()
        shapes = replay.tuning
        a = replay.positions(0)
    with Batch([g], profile=True, autotune=False) as eager:  # one kernel at a time
        eager.set_tuning(**shapes)
        eager.run(6)
        eager.sync()
        b = eager.positions(0)
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b)


def test_sbm_batch_of_512():
    gs = workloads.sbm_batch(512, iterations=200)
    sizes = np.array([gr["n"] for gr in gs])
    picks = [int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]),
             int(np.argmax(sizes))]
    with Batch(gs) as b:
        b.run(1)
        b.sync()
        f1 = {k: b.forces(k) for k in picks}
        b.run(2)
        b.sync()
        together = [b.positions(k) for k in range(len(gs))]
    for k in picks:
        gr = gs[k]
        e = np.stack([gr["frm"], gr["to"]], 1)
        _, tr = O.large_run(gr["n"], e, gr["init"], 1, mode="m", theta=1.2, trace=True)
        assert normwise(f1[k], tr["force"][0]) <= 1e-5, (k, gr["n"])
    for k, gr in enumerate(gs):
        with Batch([gr], autotune=False) as one:
            one.run(3)
            one.sync()
            assert np.array_equal(one.positions(0), together[k]), k


def _oracle_kw(g, precision):
    kw = dict(theta=g.get("theta", 1.2), linlog=g.get("linlog", False),
              outbound=g.get("dissuade_hubs", False), adjust=g.get("adjust", False))
    if "size" in g:
        s = np.asarray(g["size"], np.float64)
        kw["size"] = s.astype(np.float32).astype(np.float64) if precision == "single" else s
    if g.get("fixed") is not None:
        kw["fixed"] = np.asarray(g["fixed"], np.int32)
    return kw


def test_served_mix_against_alone_and_oracle():
    reqs, arrive = workloads.serving_requests(16, rate=200.0, max_nodes=200_000, seed=77,
                                              iterations=(8, 40))
    kinds = dict(linlog=0, adjust=0, fixed=0, dissuade_hubs=0, double=0, large=0)
    for g, prec, _ in reqs:
        kinds["linlog"] += bool(g.get("linlog"))
        kinds["adjust"] += bool(g.get("adjust"))
        kinds["fixed"] += g.get("fixed") is not None
        kinds["dissuade_hubs"] += bool(g.get("dissuade_hubs"))
        kinds["double"] += prec == "double"
        kinds["large"] += g["n"] > 100_000
    assert min(kinds.values()) >= 1, kinds
    import time
    with Server(max_batches=4, chunk=10) as srv:
        t0 = time.perf_counter()
        ids = []
        for (g, prec, _), ta in zip(reqs, arrive):
            wait = t0 + ta - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(g, prec))
        out = [srv.result(j)[0] for j in ids]
        assert srv.stats()["admitted"] == len(reqs)
    checked = 0
    for k, ((g, prec, _), pos) in enumerate(zip(reqs, out)):
        with Batch([g], precision=prec, autotune=False) as one:
            one.run(g["iterations"])
            one.sync()
            assert np.array_equal(one.positions(0), pos), k
        if k % 3 == 0:  # the first iteration of a subset against the oracle
            with Batch([g], precision=prec, autotune=False) as one:
                one.run(1)
                one.sync()
                f = one.forces(0)
            e = np.stack([g["frm"], g["to"]], 1)
            mode, tol = ("m", 1e-5) if prec == "single" else ("d", 1e-12)
            _, tr = O.large_run(g["n"], e, g["init"], 1, mode=mode, trace=True,
                                **_oracle_kw(g, prec))
            assert normwise(f, tr["force"][0]) <= tol, (k, g["n"], prec)
            checked += 1
    assert checked >= 5
