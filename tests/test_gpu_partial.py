"""Parity of the repulsion offload (fa2_gpu_* of include/fa2.h and the host
driver fa2_layout_run) with the oracle, following
tests/testthat/test-gpu.R of the reference.

Gates: one repulsion evaluation normwise <= 1e-5 in single precision
against the FP64 oracle on the same float32 inputs, <= 1e-12 in double;
whole layouts in double within 1e-9 of the CPU layout after 10 iterations
(test-gpu.R:34-58)."""
import ctypes

import numpy as np
import pytest

import forceatlas2_rc_amd as F
from forceatlas2_rc_amd import _lib, graphs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
lib = _lib.lib


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def mean_rel(a, b):
    return np.abs(a - b).mean() / np.abs(b).mean()


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def repulse_exact(x, y, m, s, scaling, use_double):
    n = len(x)
    g = lib.fa2_gpu_open(0, use_double, n, s is not None, 0, 1)
    assert g, _lib.last_error()
    dx, dy = np.full(n, 0.5), np.full(n, -0.25)  # the forces are ADDED
    try:
        assert lib.fa2_gpu_repulse_exact(g, P(x), P(y), P(m), P(s), scaling,
                                         P(dx), P(dy)) == 0, _lib.last_error()
        t = np.zeros(3)
        lib.fa2_gpu_times(g, P(t))
        assert t[1] > 0
    finally:
        lib.fa2_gpu_close(g)
    return np.stack([dx - 0.5, dy + 0.25], 1)


@pytest.mark.parametrize("n", [1, 2, 100, 3000, 50000])
@pytest.mark.parametrize("adjust", [False, True])
def test_exact_repulsion_single(n, adjust):
    rng = np.random.default_rng(n)
    r = 10 * np.sqrt(n)
    x, y = f32(rng.uniform(-r, r, n)), f32(rng.uniform(-r, r, n))
    m = 1.0 + rng.integers(0, 30, n)
    s = f32(rng.uniform(0.5, 3, n)) if adjust else None
    got = repulse_exact(x, y, m, s, 2.0, 0)
    want = O.repulse_exact(x, y, m, 2.0, size=s)
    if n == 1:
        assert np.all(got == 0)
        return
    assert normwise(got, want) <= 1e-5


@pytest.mark.parametrize("adjust", [False, True])
def test_exact_repulsion_double(adjust):
    n = 4000
    rng = np.random.default_rng(1)
    x, y = rng.uniform(-500, 500, n), rng.uniform(-500, 500, n)
    m = 1.0 + rng.integers(0, 30, n)
    s = rng.uniform(0.5, 3, n) if adjust else None
    got = repulse_exact(x, y, m, s, 2.0, 1)
    want = O.repulse_exact(x, y, m, 2.0, size=s)
    assert normwise(got, want) <= 1e-12


def _bh_case(n, adjust, use_double):
    rng = np.random.default_rng(n + 7)
    r = 10 * np.sqrt(n)
    x, y = rng.uniform(-r, r, n), rng.uniform(-r, r, n)
    if not use_double:
        x, y = f32(x), f32(y)
    x[: n // 20] = x[0]  # coincident nodes
    y[: n // 20] = y[0]
    m = 1.0 + rng.integers(0, 30, n)
    s = f32(rng.uniform(0.5, 3, n)) if adjust else None
    return x, y, m, s


def _bh_gpu(x, y, m, s, use_double, theta):
    from tests.test_abi import host_tree
    n = len(x)
    t, keep = host_tree(x, y, m, s)
    g = lib.fa2_gpu_open(0, use_double, n, int(s is not None), 1, 1)
    assert g, _lib.last_error()
    dx, dy = np.zeros(n), np.zeros(n)
    try:
        assert lib.fa2_gpu_repulse_bh(g, ctypes.byref(t), 2.0, theta, P(dx),
                                      P(dy)) == 0, _lib.last_error()
    finally:
        lib.fa2_gpu_close(g)
    return np.stack([dx, dy], 1)


@pytest.mark.parametrize("adjust", [False, True])
@pytest.mark.parametrize("theta", [0.0, 1.2])
def test_bh_repulsion_host_tree(adjust, theta):
    """The offload's Barnes-Hut over the host Gephi tree. Single precision
    against the FP64 model of the single-precision kernel (float32 records
    and acceptance test, FP64 terms: the same regions approximated):
    normwise <= 1e-5 over all nodes and over the half with the smallest
    forces; against the FP64 oracle (whose decisions may differ where the
    test is borderline) within 1e-4. Double precision against the FP64
    oracle: <= 1e-12."""
    x, y, m, s = _bh_case(20000, adjust, 0)
    got = _bh_gpu(x, y, m, s, 0, theta)
    model = O.repulse_bh_single(x, y, m, 2.0, theta, size=s)
    assert normwise(got, model) <= 1e-5
    small = np.abs(model).max(1) <= np.median(np.abs(model).max(1))
    assert normwise(got[small], model[small]) <= 1e-5
    assert normwise(got, O.repulse_bh(x, y, m, 2.0, theta, size=s)) <= 1e-4
    x, y, m, s = _bh_case(5000, adjust, 1)
    got = _bh_gpu(x, y, m, s, 1, theta)
    assert normwise(got, O.repulse_bh(x, y, m, 2.0, theta, size=s)) <= 1e-12


def test_bh_offload_takes_the_float32_decisions():
    """At a theta where the float32 and the FP64 acceptance tests disagree
    for some (node, region) pair, the single-precision kernel follows the
    float32 decision: on the nodes where the two decisions give forces more
    than 1e-4 apart, the kernel is within 1e-5 of the float32-decision
    model; over all nodes, within 1e-5 normwise."""
    from tests.test_oracle import _borderline_theta, _walk_single_numpy
    x, y, m, _ = _bh_case(3000, False, 0)
    t = O.gephi_tree(x, y, m)
    theta = _borderline_theta(t)
    got = _bh_gpu(x, y, m, None, 0, theta)
    model = O.repulse_bh_single(x, y, m, 2.0, theta)
    other = _walk_single_numpy(t, theta, 2.0, fp64_decisions=True)
    scale = np.abs(model).max(1)
    flip = np.abs(model - other).max(1) > 1e-4 * scale
    assert flip.any()
    assert np.all(np.abs(got - model).max(1)[flip] <= 1e-5 * scale[flip])
    assert normwise(got, model) <= 1e-5


@pytest.mark.parametrize("bh", [False, True])
def test_double_layout_equals_cpu(bh):
    """test-gpu.R:34-58: 10 iterations, 1e-9, over the option sets."""
    n, e = graphs.barabasi_albert(400, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    rng = np.random.default_rng(3)
    w = rng.uniform(0.2, 3, len(e))
    size = rng.uniform(0.5, 4, n)
    sets = [dict(), dict(theta=0.5, scaling_ratio=5, gravity=3),
            dict(linlog=True, dissuade_hubs=True, strong_gravity=True),
            dict(weights=w, edge_weight_influence=0.5),
            dict(prevent_overlap=True, node_size=size),
            dict(fixed=np.arange(n) % 2 == 0)]
    for kw in sets:
        got = F.layout_forceatlas2_gpu(n, e, 10, init=pos, barnes_hut=bh,
                                       precision="double", **kw)
        okw = dict(scaling=kw.get("scaling_ratio", 2.0),
                   gravity=kw.get("gravity", 1.0), theta=kw.get("theta", 1.2),
                   strong=kw.get("strong_gravity", False),
                   linlog=kw.get("linlog", False),
                   outbound=kw.get("dissuade_hubs", False),
                   adjust=kw.get("prevent_overlap", False),
                   size=kw.get("node_size", 1.0), fixed=kw.get("fixed"),
                   barnes_hut=bh)
        if "weights" in kw:
            okw["w"] = F.prepare_weights(len(e), w, 0.5)
        want = O.layout(n, e, pos, 10, **okw)
        assert mean_rel(got, want) < 1e-9, list(kw)


@pytest.mark.parametrize("bh", [False, True])
def test_layout_does_not_depend_on_host_threads(bh):
    """test-algorithm.R:88-103 for the offload: with threads > 1 the host
    runs the tree build and the attraction, swing and displacement loops on
    several threads (the attraction node by node over incidence lists in
    edge order); the layout is bit-identical to one thread's, over the
    option sets, with self-loops and a repeated edge."""
    n, e = graphs.barabasi_albert(3000, 3, seed=44)
    e = np.concatenate([e, [[5, 5], [7, 9], [7, 9]]])
    pos = graphs.uniform_positions(n, -200, 200, 45)
    rng = np.random.default_rng(46)
    w = rng.uniform(0.2, 3, len(e))
    size = rng.uniform(0.5, 4, n)
    sets = [dict(), dict(linlog=True, dissuade_hubs=True, weights=w),
            dict(prevent_overlap=True, node_size=size, strong_gravity=True),
            dict(fixed=np.arange(n) % 3 == 0, gravity=2.0)]
    for kw in sets:
        for precision in ("single", "double"):
            one = F.layout_forceatlas2_gpu(n, e, 8, init=pos, barnes_hut=bh, threads=1,
                                           precision=precision, **kw)
            many = F.layout_forceatlas2_gpu(n, e, 8, init=pos, barnes_hut=bh, threads=6,
                                            precision=precision, **kw)
            assert np.array_equal(one, many), (list(kw), precision)


def test_single_layout_close_to_cpu():
    """test-gpu.R:60-79."""
    n, e = graphs.barabasi_albert(400, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    for bh in (False, True):
        cpu = O.layout(n, e, pos, 2, barnes_hut=bh)
        gpu = F.layout_forceatlas2_gpu(n, e, 2, init=pos, barnes_hut=bh)
        assert np.abs(gpu - cpu).max() / np.ptp(cpu) < 1e-3
        assert not np.array_equal(gpu, cpu)
    edges = [(i, j) for b in (0, 10) for i in range(b, b + 10) for j in range(i + 1, b + 10)]
    edges.append((0, 10))
    xy = F.layout_forceatlas2_gpu(20, np.array(edges), 500, rng=1)
    d = np.hypot(*(xy[:, None] - xy[None]).transpose(2, 0, 1))
    assert np.all(np.isfinite(xy))
    assert d[:10, :10].max() < d[1:10, 11:20].min()


def test_degenerate_graphs():
    """test-gpu.R:81-102."""
    assert F.layout_forceatlas2_gpu(0, np.zeros((0, 2), int), 10).shape == (0, 2)
    assert np.all(np.isfinite(F.layout_forceatlas2_gpu(1, np.zeros((0, 2), int), 10, rng=1)))
    n, e = graphs.barabasi_albert(50, 2, seed=42)
    zero = np.zeros((n, 2))
    pos = graphs.uniform_positions(n, -50, 50, 7)
    pos[:20] = 1.0
    for bh in (False, True):
        assert np.array_equal(F.layout_forceatlas2_gpu(n, e, 50, init=zero, barnes_hut=bh), zero)
        assert np.all(np.isfinite(F.layout_forceatlas2_gpu(n, e, 50, init=pos, barnes_hut=bh)))


def test_devices_and_errors():
    """test-gpu.R:12-32."""
    d = F.gpu_devices()
    assert len(d) >= 1 and d[0]["type"] == "GPU" and d[0]["double"]
    assert set(d[0]) == {"name", "vendor", "platform", "version", "type",
                         "memory_gb", "double"}
    n, e = graphs.barabasi_albert(20, 2, seed=1)
    with pytest.raises(ValueError, match="device"):
        F.layout_forceatlas2_gpu(n, e, device=len(d) + 1)
    with pytest.raises(ValueError, match="device"):
        F.layout_forceatlas2_gpu(n, e, device=0)
    assert lib.fa2_gpu_open(len(d), 0, 10, 0, 0, 1) is None
    assert "device" in _lib.last_error()
    assert lib.fa2_gpu_open(0, 0, (1 << 27) + 1, 0, 0, 1) is None
