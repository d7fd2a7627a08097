"""Parity on power-law graphs whose hubs reach every attraction class.

The attraction (pipeline.cl:392-429) runs one loop per node over its whole
incidence list, whatever its degree; the reference's tests have no degree
cap (tests/testthat/test-large.R:10-57). Here a node's list goes to one of
four paths by length: one lane (1..16 incidences), 16 lanes (17..256), one
wave per work item of 1024 incidences (more than 256), and the items' partial
sums added in item order (more than 1024). R-MAT graphs put hubs of
thousands of incidences in the last two.

Gates (against the oracle on the same positions):
  * one iteration, single precision: normwise max|dF| / max|F_ref| <= 1e-5
    against the FP64 model of the single-precision path (mode "m"), over the
    whole graph and within each degree class separately, so that small
    nodes are not hidden behind the hubs' large forces;
  * one iteration, double precision: <= 1e-12 against mode "d", over the
    graph and per class;
  * ten iterations step by step from the oracle's state, as in
    test_gpu_large.py.
Settings: linlog with dissuade_hubs (coefficients read per incidence),
unweighted without dissuade_hubs (one coefficient per node), edge weights,
prevent_overlap with node sizes; theta 1.2 on 2^16 nodes, theta 0 (exact)
on 2^14 nodes.
"""
import numpy as np
import pytest

from forceatlas2_rc_amd import Batch, graphs
from forceatlas2_rc_amd.api import prepare_large
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# list lengths that separate the attraction paths
CLASSES = [(1, 16), (17, 256), (257, 1024), (1025, 1 << 30)]


def normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def class_errors(f, ref, deg):
    out = []
    for lo, hi in CLASSES:
        sel = (deg >= lo) & (deg <= hi)
        assert sel.any(), (lo, hi)
        out.append(normwise(f[sel], ref[sel]))
    return out


_CACHE = {}


def rmat_case(scale, n_edges, seed):
    key = (scale, n_edges, seed)
    if key not in _CACHE:
        n, e = graphs.rmat(scale, n_edges, seed=seed)
        r = 10 * np.sqrt(n)
        pos = graphs.uniform_positions(n, -r, r, seed + 1)
        deg = np.bincount(e[:, 0], minlength=n) + np.bincount(e[:, 1], minlength=n)
        _CACHE[key] = (n, e, pos, deg)
    return _CACHE[key]


SETTINGS = {
    "linlog,dissuade_hubs": dict(linlog=True, dissuade_hubs=True),
    "unweighted": dict(),
    "weights,dissuade_hubs": dict(weights="rand", dissuade_hubs=True),
    "adjust,linlog": dict(adjust=True, size="rand", linlog=True),
}


def _setting(name, n, m, single):
    kw = dict(SETTINGS[name])
    rng = np.random.default_rng(17)
    okw = dict(linlog=kw.get("linlog", False), outbound=kw.get("dissuade_hubs", False),
               adjust=kw.get("adjust", False))
    if kw.get("weights") == "rand":
        w = rng.uniform(0.5, 2.0, m).astype(np.float32).astype(np.float64)
        kw["weights"] = w
        okw["w"] = w
    if kw.get("size") == "rand":
        s = rng.uniform(0.5, 3.0, n)
        # the device holds sizes as float32 in single precision
        s = s.astype(np.float32).astype(np.float64) if single else s
        kw["size"] = s
        okw["size"] = s
    return kw, okw


def _one_step(n, e, pos, kw, theta, precision):
    g = dict(n=n, edges=e, init=pos, iterations=1, theta=theta, **kw)
    with Batch([g], precision=precision, autotune=False) as b:
        b.run(1)
        b.sync()
        return b.forces(0), b.speed(0)


@pytest.mark.parametrize("setting", list(SETTINGS))
@pytest.mark.parametrize("precision", ["single", "double"])
def test_hub_forces_theta_1_2(setting, precision):
    n, e, pos, deg = rmat_case(16, 400_000, 3)
    assert deg.max() > 4 * 1024  # several work items per hub
    kw, okw = _setting(setting, n, len(e), precision == "single")
    f, sp = _one_step(n, e, pos, kw, 1.2, precision)
    mode, tol = ("m", 1e-5) if precision == "single" else ("d", 1e-12)
    _, tr = O.large_run(n, e, pos, 1, mode=mode, theta=1.2, trace=True, **okw)
    ref = tr["force"][0]
    assert normwise(f, ref) <= tol
    errs = class_errors(f, ref, deg)
    assert max(errs) <= tol, errs
    assert sp == pytest.approx(tr["speed"][0], rel=tol)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_hub_forces_exact_repulsion(precision):
    """theta = 0: nothing is approximated, so the repulsion is the exact
    all-pairs sum (test-large.R:10-38), beside the hubs' attraction."""
    n, e, pos, deg = rmat_case(14, 120_000, 5)
    assert deg.max() > 1024
    kw, okw = _setting("linlog,dissuade_hubs", n, len(e), precision == "single")
    f, _ = _one_step(n, e, pos, kw, 0.0, precision)
    mode, tol = ("m", 1e-5) if precision == "single" else ("d", 1e-12)
    _, tr = O.large_run(n, e, pos, 1, mode=mode, theta=0.0, trace=True, **okw)
    ref = tr["force"][0]
    assert normwise(f, ref) <= tol
    assert max(class_errors(f, ref, deg)) <= tol


def test_hub_attraction_alone():
    """The attraction itself, without the repulsion that dominates the
    force: with scaling 0 and gravity 0 the force is the edges' pull alone,
    compared per class with the oracle's."""
    n, e, pos, deg = rmat_case(16, 400_000, 3)
    for precision, mode, tol in (("single", "m", 1e-5), ("double", "d", 1e-12)):
        kw = dict(linlog=True, dissuade_hubs=True, scaling=0.0, gravity=0.0)
        f, _ = _one_step(n, e, pos, kw, 1.2, precision)
        _, tr = O.large_run(n, e, pos, 1, mode=mode, theta=1.2, linlog=True, outbound=True,
                            scaling=0.0, gravity=0.0, trace=True)
        ref = tr["force"][0]
        # an independent FP64 sum over the incidence lists (pipeline.cl:412-427)
        mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1], np.ones(len(e)), True)
        p = pos.astype(np.float32).astype(np.float64) if precision == "single" else pos
        src = np.repeat(np.arange(n), np.diff(off))
        xd = p[src] - p[node]
        d = np.hypot(xd[:, 0], xd[:, 1])
        fac = np.where(d > 0, np.log1p(d) / np.where(d > 0, d, 1.0), 1.0)
        term = -(coef * fac)[:, None] * xd
        direct = np.zeros((n, 2))
        np.add.at(direct, src, term)
        assert normwise(ref, direct) <= 1e-12
        assert normwise(f, ref) <= tol, precision
        assert max(class_errors(f, ref, deg)) <= tol, precision


def test_hub_ten_iterations_step_by_step():
    """The first 10 iterations from the oracle's state (positions, previous
    force, speed, efficiency), one at a time: force within 1e-5 normwise and
    per class, speed and efficiency within 1e-5."""
    n, e, pos, deg = rmat_case(15, 200_000, 7)
    assert deg.max() > 1024
    _, tr = O.large_run(n, e, pos, 10, mode="m", theta=1.2, linlog=True, outbound=True,
                        trace=True)
    g = dict(n=n, edges=e, init=pos, theta=1.2, linlog=True, dissuade_hubs=True)
    with Batch([g], autotune=False) as b:
        for t in range(10):
            if t > 0:
                b.set_state(0, tr["pos"][t - 1], tr["force"][t - 1], *tr["speed"][t - 1][2:])
            b.run(1)
            b.sync()
            f = b.forces(0)
            ref = tr["force"][t]
            assert normwise(f, ref) <= 1e-5, t
            assert max(class_errors(f, ref, deg)) <= 1e-5, t
            assert b.speed(0)[2:] == pytest.approx(tr["speed"][t][2:], rel=1e-5), t
