"""Parity of the all-GPU path (fa2_full_run / fa2_batch_*) with the oracle.

Gates (written in each test):
  * Morton codes, sorted permutation, Karras topology: bit-exact against
    the reference large path restated on the CPU, on identical float32
    positions.
  * forces of one iteration from identical state: normwise
    max|dF| / max|F_ref| <= 1e-5 in single precision against the FP64
    oracle evaluated on the same float32 positions, <= 1e-12 in double.
  * trajectories: step by step over 10 iterations, then the quality ratio
    (mean edge length / sd of the layout) within 2% of the CPU reference.
Relational tests follow tests/testthat/test-large.R of the reference.
"""
import numpy as np
import pytest

import forceatlas2_rc_amd as F
from forceatlas2_rc_amd import Batch, graphs
from oracle import oracle as O
from oracle.ref_numpy import update_speed

pytestmark = pytest.mark.gpu


def normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def mean_rel(a, b):
    return np.abs(a - b).mean() / np.abs(b).mean()


def pa(n, m=2, seed=42):
    return graphs.barabasi_albert(n, m, seed=seed)


def start(n, seed=7, r=None):
    r = 10 * np.sqrt(n) if r is None else r
    return graphs.uniform_positions(n, -r, r, seed)


def one_step(n, e, pos, precision="single", **kw):
    g = dict(n=n, edges=e, init=pos, iterations=1, **kw)
    with Batch([g], precision=precision) as b:
        b.run(1)
        b.sync()
        return b.tree(0), b.forces(0), b.speed(0), b.positions(0)


def oracle_kw(kw):
    out = dict(theta=kw.get("theta", 1.2), linlog=kw.get("linlog", False),
               outbound=kw.get("dissuade_hubs", False),
               adjust=kw.get("adjust", False), strong=kw.get("strong", False),
               scaling=kw.get("scaling", 2.0), gravity=kw.get("gravity", 1.0),
               jt=kw.get("jitter_tolerance", 1.0))
    if "size" in kw:
        out["size"] = kw["size"]
    if "fixed" in kw:
        out["fixed"] = kw["fixed"]
    return out


@pytest.mark.parametrize("n,dup", [(2, 0), (3, 0), (700, 30), (3000, 0), (9000, 300)])
def test_tree_is_bit_exact(n, dup):
    """Codes, permutation and topology against the reference large path
    (pipeline.cl:105-270) on the same float32 positions; coincident nodes
    force equal Morton codes (test-large.R:121-129)."""
    n, e = pa(n) if n > 2 else (n, np.array([[0, 1]]))
    pos = start(n)
    if dup:
        pos[:dup] = pos[np.arange(dup) % 3]
    t, *_ = one_step(n, e, pos)
    p32 = pos.astype(np.float32)
    ref = O.large_tree(p32[:, 0], p32[:, 1], "f")
    for k in ("keys", "order", "child", "range", "parent"):
        assert np.array_equal(t[k].astype(np.int64), ref[k].astype(np.int64)), k


def test_tree_is_bit_exact_in_double():
    n, e = pa(2000)
    pos = start(n)
    t, *_ = one_step(n, e, pos, precision="double")
    ref = O.large_tree(pos[:, 0], pos[:, 1], "d")
    for k in ("keys", "order", "child", "range", "parent"):
        assert np.array_equal(t[k].astype(np.int64), ref[k].astype(np.int64)), k


def dfs_skip(child, n):
    """The node that a left-first depth-first walk of the reference tree
    visits after each region's subtree (-1 after the last)."""
    pre, stack = [], [0]
    while stack:
        r = stack.pop()
        pre.append(r)
        if r < n - 1:
            stack.append(child[2 * r + 1])
            stack.append(child[2 * r])
    where = {r: i for i, r in enumerate(pre)}
    size = np.ones(2 * n - 1, np.int64)
    for r in reversed(pre):
        if r < n - 1:
            size[r] = 1 + size[child[2 * r]] + size[child[2 * r + 1]]
    out = np.empty(n - 1, np.int64)
    for r in range(n - 1):
        j = where[r] + size[r]
        out[r] = pre[j] if j < len(pre) else -1
    return out


@pytest.mark.parametrize("precision,n,mode", [("single", 3000, "f"), ("single", 9000, "f"),
                                              ("single", 2048, "f"), ("single", 70000, "f"),
                                              ("double", 3000, "d"), ("double", 1500, "d"),
                                              ("double", 1024, "d"), ("double", 70000, "d")])
def test_region_records(precision, n, mode):
    """Every region's traversal record (pipeline.cl:276-323) against the
    oracle on the same positions: masses, centres and size^2 bit-exact
    (w * w + h * h is rounded term by term, as the reference does); skip links are
    the depth-first successors in the oracle's tree. The blocks of the first
    summary level hold 1024 nodes in single precision and 512 in double, and
    the units above them 16 and 256 blocks: 1024-9000 nodes need the second
    level (1024 in double is two blocks), 70000 nodes the third."""
    n, e = pa(n)
    pos = start(n)
    if precision == "single":
        pos = pos.astype(np.float32).astype(np.float64)
    mass = 1.0 + np.bincount(e[:, 0], minlength=n) + np.bincount(e[:, 1], minlength=n)
    g = dict(n=n, edges=e, init=pos, iterations=1)
    with Batch([g], precision=precision) as b:
        b.run(1)
        b.sync()
        rec, skip = b.regions(0)
    ref = O.large_regions(pos[:, 0], pos[:, 1], mass, mode)
    assert np.array_equal(rec[:, 2], ref[:, 2])
    assert np.array_equal(rec[:, :2], ref[:, :2])
    assert np.array_equal(rec[:, 3], ref[:, 3])
    tree = O.large_tree(pos[:, 0], pos[:, 1], mode)
    assert np.array_equal(skip, dfs_skip(tree["child"], n))


OPTS = [
    dict(),
    dict(linlog=True, dissuade_hubs=True),
    dict(strong=True, gravity=0.3),
    dict(adjust=True, size="rand"),
    dict(adjust=True, size="rand", linlog=True, dissuade_hubs=True),
    dict(fixed="alt", scaling=5.0, gravity=3.0),
]


def _opts(o, n, single=True):
    o = dict(o)
    if o.get("size") == "rand":
        # the device holds sizes as float32: the oracle gets the same values
        size = np.random.default_rng(1).uniform(0.5, 4, n)
        o["size"] = size.astype(np.float32).astype(np.float64) if single else size
    if o.get("fixed") == "alt":
        o["fixed"] = np.arange(n) % 2 == 0
    return o


@pytest.mark.parametrize("opts", OPTS, ids=lambda o: ",".join(o) or "default")
@pytest.mark.parametrize("theta", [0.0, 1.2])
def test_one_iteration_forces_single(opts, theta):
    """Single precision against the FP64 oracle on the same float32
    positions: normwise <= 1e-5."""
    n, e = pa(4000, 3, seed=5)
    pos = start(n, seed=6)
    kw = _opts(opts, n)
    kw["theta"] = theta
    _, f, sp, _ = one_step(n, e, pos, **kw)
    _, tr = O.large_run(n, e, pos, 1, mode="m", trace=True, **oracle_kw(kw))
    assert normwise(f, tr["force"][0]) <= 1e-5
    assert sp == pytest.approx(tr["speed"][0], rel=1e-5)


@pytest.mark.parametrize("opts", OPTS[:4], ids=lambda o: ",".join(o) or "default")
def test_one_iteration_forces_double(opts):
    n, e = pa(3000, 3, seed=5)
    pos = start(n, seed=6)
    kw = _opts(opts, n, single=False)
    _, f, sp, _ = one_step(n, e, pos, precision="double", **kw)
    _, tr = O.large_run(n, e, pos, 1, mode="d", trace=True, **oracle_kw(kw))
    assert normwise(f, tr["force"][0]) <= 1e-12
    assert sp == pytest.approx(tr["speed"][0], rel=1e-12)


def fp64_move(f_old, f, mass, speed, eff, jt=1.0):
    """Swing, traction, speed and displacement of one iteration in FP64
    (pipeline.cl:438-506 with fa2_update_speed, src/forceatlas2.c:445-485),
    no adjust, nothing fixed."""
    n = len(mass)
    sw = mass * np.hypot(f_old[:, 0] - f[:, 0], f_old[:, 1] - f[:, 1])
    tr = 0.5 * mass * np.hypot(f_old[:, 0] + f[:, 0], f_old[:, 1] + f[:, 1])
    sp, _ = update_speed(n, sw.sum(), tr.sum(), jt, speed, eff)
    return f * (sp / (1.0 + np.sqrt(sp * sw)))[:, None]


def move_kappa(f_old, f, mass, speed, eff, trials=8):
    """Largest normwise displacement change per normwise force change over
    random perturbations of the force by up to 1e-7 max|F| per component."""
    rng = np.random.default_rng(1)
    d0 = fp64_move(f_old, f, mass, speed, eff)
    bound = 1e-7 * np.abs(f).max()
    kappa = 0.0
    for _ in range(trials):
        pf = f + rng.uniform(-bound, bound, f.shape)
        kappa = max(kappa, normwise(fp64_move(f_old, pf, mass, speed, eff), d0) / normwise(pf, f))
    return kappa


@pytest.mark.parametrize("n0,seed", [(3000, 8), (5000, 11)])
def test_ten_iterations_step_by_step(n0, seed):
    """The first 10 iterations, step by step from identical state: the
    oracle's state before step t (positions, previous force, speed and
    efficiency) is loaded into the device, one iteration runs, and:
      * its force is within 1e-5 normwise of the oracle's;
      * its displacement is the FP64 move of that force, within 1e-9;
      * its displacement is within 1e-5 normwise of the oracle's, or, where
        the move is ill-conditioned, within kappa times the force error,
        kappa being the largest displacement change per unit force change
        seen over random force perturbations of 1e-7 max|F|. The factor
        speed / (1 + sqrt(speed * swing)) has a sensitivity growing like
        1 / sqrt(swing), and a node whose force barely changed has
        swing = m |F_old - F| near zero: there a 5e-8 error in the force
        moves the displacement by 1e-5 and more, whatever computes it
        (kappa reaches 1e4 on these graphs)."""
    n, e = pa(n0, 3, seed=seed)
    pos = start(n, seed=seed + 1)
    mass = 1.0 + np.bincount(e[:, 0], minlength=n) + np.bincount(e[:, 1], minlength=n)
    _, tr = O.large_run(n, e, pos, 10, mode="m", theta=1.2, linlog=True,
                        outbound=True, trace=True)
    g = dict(n=n, edges=e, init=pos, theta=1.2, linlog=True, dissuade_hubs=True)
    with Batch([g]) as b:
        for t in range(10):
            p0 = pos if t == 0 else tr["pos"][t - 1]
            f_old = np.zeros((n, 2)) if t == 0 else tr["force"][t - 1]
            sp, ef = (1.0, 1.0) if t == 0 else tr["speed"][t - 1][2:]
            if t > 0:
                b.set_state(0, p0, f_old, sp, ef)
            b.run(1)
            b.sync()
            f = b.forces(0)
            assert normwise(f, tr["force"][t]) <= 1e-5, t
            disp = b.positions(0) - p0
            assert normwise(disp, fp64_move(f_old, f, mass, sp, ef)) <= 1e-9, t
            ref = tr["pos"][t] - p0
            assert normwise(disp, ref) <= max(1e-5, move_kappa(f_old, tr["force"][t], mass, sp, ef)
                                              * normwise(f, tr["force"][t])), t
            assert b.speed(0)[2:] == pytest.approx(tr["speed"][t][2:], rel=1e-5)


def test_double_without_approximation_is_the_exact_layout():
    """test-large.R:10-38: theta = 0 in double equals the exact CPU layout
    within 1e-8 after 5 iterations, with coincident nodes."""
    n, e = pa(700, 2)
    pos = start(n, r=50)
    pos[:30] = pos[np.repeat(np.arange(3), 10)]
    rng = np.random.default_rng(3)
    w = rng.uniform(0.2, 3, len(e))
    size = rng.uniform(0.5, 4, n)
    for kw in (dict(), dict(scaling_ratio=5, gravity=3),
               dict(linlog=True, strong_gravity=True),
               dict(dissuade_hubs=True, gravity=0),
               dict(weights=w, edge_weight_influence=0.5, jitter_tolerance=0.3),
               dict(prevent_overlap=True, node_size=size),
               dict(prevent_overlap=True, node_size=size, linlog=True,
                    dissuade_hubs=True, weights=w),
               dict(fixed=np.arange(n) % 2 == 0)):
        got = F.layout_forceatlas2_large(n, e, 5, init=pos, theta=0.0,
                                         precision="double", **kw)
        okw = dict(scaling=kw.get("scaling_ratio", 2.0),
                   gravity=kw.get("gravity", 1.0),
                   strong=kw.get("strong_gravity", False),
                   linlog=kw.get("linlog", False),
                   outbound=kw.get("dissuade_hubs", False),
                   adjust=kw.get("prevent_overlap", False),
                   jt=kw.get("jitter_tolerance", 1.0), fixed=kw.get("fixed"),
                   size=kw.get("node_size", 1.0))
        if "weights" in kw:
            okw["w"] = F.prepare_weights(len(e), kw["weights"],
                                         kw.get("edge_weight_influence", 1.0))
        want = O.layout(n, e, pos, 5, **okw)
        assert mean_rel(got, want) < 1e-8, kw.keys()


def test_theta_sweep():
    """test-large.R:40-57."""
    n, e = pa(3000)
    pos = start(n, r=50)
    exact = O.layout(n, e, pos, 1)

    def err(theta, precision="double"):
        xy = F.layout_forceatlas2_large(n, e, 1, init=pos, theta=theta,
                                        precision=precision)
        return np.abs(xy - exact).max() / np.ptp(exact)

    assert err(0.1) < 1e-4
    assert err(0.1) < err(0.5) < err(1.2) < 0.02
    assert err(1.2, "single") < 0.02
    assert err(0.0, "single") < 1e-3


def test_records_beyond_4_gib_in_double():
    """Traversal records past 4 GiB: in double precision a record takes 64
    bytes, so a batch of 2^25 + 3000 nodes holds more than 2^26 records and
    the traversal indexes them in 64 bits. The small graph after the large
    one walks records beyond 4 GiB; it must give its layout alone bit for
    bit, and the large one (isolated nodes, uniform) finite positions."""
    n_a = 1 << 25
    rng = np.random.default_rng(31)
    big = dict(n=n_a, edges=np.zeros((0, 2), np.int32), init=rng.uniform(-1e4, 1e4, (n_a, 2)),
               iterations=2)
    n_b, e_b = pa(3000, 3, seed=32)
    small = dict(n=n_b, edges=e_b, init=start(n_b, seed=33), linlog=True, iterations=2)
    with Batch([big, small], precision="double", autotune=False) as b:
        b.run(2)
        b.sync()
        pos_b = b.positions(1)
        assert np.all(np.isfinite(b.positions(0)))
    with Batch([small], precision="double", autotune=False) as one:
        one.run(2)
        one.sync()
        assert np.array_equal(one.positions(0), pos_b)


def test_reproducible_and_same_alone_as_in_a_batch():
    """test-large.R:59-71, and batch composition invariance."""
    n, e = pa(5000)
    pos = start(n)
    a = F.layout_forceatlas2_large(n, e, 30, init=pos)
    b = F.layout_forceatlas2_large(n, e, 30, init=pos)
    assert np.array_equal(a, b)
    gs = []
    for i, m in enumerate([300, 5000, 2049, 777]):
        m, ee = pa(m, 2, seed=20 + i)
        gs.append(dict(n=m, edges=ee, init=start(m, seed=i), iterations=25,
                       linlog=bool(i % 2), theta=1.0 + 0.1 * i))
    with Batch(gs) as bb:
        bb.run(25)
        bb.sync()
        together = [bb.positions(i) for i in range(len(gs))]
    for i, g in enumerate(gs):
        with Batch([g]) as one:
            one.run(25)
            one.sync()
            assert np.array_equal(one.positions(0), together[i])


def test_graphs_without_fp64_sums_are_unchanged_beside_those_with():
    """A batch holding a graph at theta 0 or with node sizes runs the
    traversal kernel that carries FP64 force sums; the graphs that keep
    FP32 sums (theta > 0, no sizes) give the same layout bit for bit there
    as alone, and so do the others. The theta-0 graph's forces are within
    1e-5 normwise of the oracle (its lanes sum every other node)."""
    specs = [(1500, dict(linlog=True)), (900, dict(theta=0.0)),
             (1200, dict(adjust=True, size=1.5)), (2100, dict(theta=0.8))]
    gs = []
    for i, (m, kw) in enumerate(specs):
        m, ee = pa(m, 2, seed=40 + i)
        gs.append(dict(n=m, edges=ee, init=start(m, seed=50 + i), iterations=12, **kw))
    with Batch(gs, autotune=False) as bb:
        bb.run(12)
        bb.sync()
        together = [bb.positions(i) for i in range(len(gs))]
    for i, g in enumerate(gs):
        with Batch([g], autotune=False) as one:
            one.run(12)
            one.sync()
            assert np.array_equal(one.positions(0), together[i]), specs[i]
    g = gs[1]
    with Batch([dict(g, iterations=1)], autotune=False) as one:
        one.run(1)
        one.sync()
        f = one.forces(0)
    _, tr = O.large_run(g["n"], g["edges"], g["init"], 1, mode="m", theta=0.0, trace=True)
    ref = tr["force"][0]
    assert np.abs(f - ref).max() / np.abs(ref).max() <= 1e-5


def test_layout_quality_matches_cpu():
    """test-large.R:73-94 (5%) and the 2% edge-length/stress gate: islands
    4 x 100, 300 iterations."""
    n, e = graphs.islands(4, 100, 0.1, 2, seed=1)
    pos = start(n, r=50)

    def edge_length(xy):
        d = np.hypot(*(xy[e[:, 0]] - xy[e[:, 1]]).T)
        return d.mean() / np.std(xy, ddof=1)

    cpu = O.layout(n, e, pos, 300, barnes_hut=True)
    gpu = F.layout_forceatlas2_large(n, e, 300, init=pos)
    assert np.all(np.isfinite(gpu))
    assert edge_length(gpu) == pytest.approx(edge_length(cpu), rel=0.02)
    assert np.std(gpu, ddof=1) == pytest.approx(np.std(cpu, ddof=1), rel=0.05)


def test_cliques_separate():
    e = [(i, j) for b in (0, 10) for i in range(b, b + 10) for j in range(i + 1, b + 10)]
    e.append((0, 10))
    xy = F.layout_forceatlas2_large(20, np.array(e), 500, rng=1)
    d = np.hypot(*(xy[:, None] - xy[None]).transpose(2, 0, 1))
    assert np.all(np.isfinite(xy))
    assert d[:10, :10].max() < d[1:10, 11:20].min()


def test_degenerate_graphs():
    """test-large.R:96-133."""
    assert F.layout_forceatlas2_large(0, np.zeros((0, 2), int), 10).shape == (0, 2)
    assert F.layout_forceatlas2_large(1, np.zeros((0, 2), int), 10, rng=1).shape == (1, 2)
    for n in range(2, 13):
        ring = np.array([(i, (i + 1) % n) for i in range(n)])
        xy = F.layout_forceatlas2_large(n, ring, 50, rng=n)
        assert np.all(np.isfinite(xy))
        d = np.hypot(*(xy[:, None] - xy[None]).transpose(2, 0, 1))
        assert d[np.triu_indices(n, 1)].min() > 0
    assert np.all(np.isfinite(F.layout_forceatlas2_large(300, np.zeros((0, 2), int), 100, rng=1)))
    g = np.array([[0, 0], [0, 1], [0, 1], [1, 2]])
    assert np.all(np.isfinite(F.layout_forceatlas2_large(5, g, 100, rng=1)))
    n, e = pa(500)
    zero = np.zeros((n, 2))
    assert np.array_equal(F.layout_forceatlas2_large(n, e, 20, init=zero), zero)
    line = np.stack([np.arange(1, n + 1), np.zeros(n)], 1).astype(float)
    assert np.all(np.isfinite(F.layout_forceatlas2_large(n, e, 20, init=line)))
    pos = start(n, r=50)
    pos[:300] = 1.0
    assert np.all(np.isfinite(F.layout_forceatlas2_large(n, e, 20, init=pos)))
    assert np.allclose(F.layout_forceatlas2_large(n, e, 0, init=pos), pos)


def test_errors():
    n, e = pa(60)
    with pytest.raises(ValueError, match="device"):
        F.layout_forceatlas2_large(n, e, 5, device=len(F.gpu_devices()) + 1)
    with pytest.raises(ValueError, match="theta"):
        F.layout_forceatlas2_large(n, e, 5, theta=-1)


def test_profile_times_every_phase():
    n, e = pa(3000)
    xy, t = F.layout_forceatlas2_large(n, e, 5, rng=1, timings=True)
    assert t["total"] > 0
    for k in ("codes", "sort", "tree", "repulsion", "attraction", "move"):
        assert t[k] > 0, k


@pytest.mark.parametrize("precision", ["single", "double"])
def test_uniform_coefficients_read_once_per_node(precision):
    """A graph whose nodes each carry one coefficient (unweighted, without
    dissuade_hubs, whose coefficient varies along a list with the other end
    of each edge) reads it once per node. Against the per-incidence reads,
    forced by changing one coefficient of one node p (the graph is then not
    uniform), the first iteration's force of every other node is
    bit-identical: its own terms did not change. Degrees span every
    attraction class, linlog on; the device's preprocessing gives the same
    forces as the host's."""
    from forceatlas2_rc_amd.api import prepare_large
    n, e = graphs.rmat(13, 60_000, seed=21)
    pos = start(n, seed=22, r=10 * np.sqrt(n))
    mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1], np.ones(len(e)), False)
    deg = np.diff(off)
    p = int(np.argmax(deg))  # a hub: its incidences are work items
    assert deg[p] > 1024 and (deg > 0).sum() > 100 and ((deg > 16) & (deg <= 256)).any()
    coef2 = coef.copy()
    coef2[off[p] + 1] *= 1.5
    base = dict(n=n, frm=e[:, 0], to=e[:, 1], mass=mass, inc_off=off, inc_node=node,
                init=pos, theta=1.2, linlog=True)
    forces = []
    for c in (coef, coef2):
        with Batch([dict(base, inc_coef=c)], precision=precision, autotune=False) as b:
            b.run(1)
            b.sync()
            forces.append(b.forces(0))
    others = np.arange(n) != p
    assert np.array_equal(forces[0][others], forces[1][others])
    assert not np.array_equal(forces[0][p], forces[1][p])
    with Batch([dict(n=n, edges=e, init=pos, theta=1.2, linlog=True)],
               precision=precision, autotune=False, ingest="device") as b:
        b.run(1)
        b.sync()
        assert np.array_equal(b.forces(0), forces[0])
