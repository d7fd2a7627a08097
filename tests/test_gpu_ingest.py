"""Preprocessing of the all-GPU layout on the device (fa2_batch_open_edges):
mass = 1 + degree, incidence lists in edge order without self-loops, and
the attraction coefficients of C_forceatlas2_large
(src/forceatlas2.c:866-902).

Gates: the device's mass, row offsets and incidence nodes equal the
oracle's restatement of the reference preprocessing (orc_large_prepare) and
the host preprocessing (api.prepare_large) exactly; the coefficients exactly
in double precision and as their float32 rounding in single; and layouts
from either preprocessing are bit-identical. The cases reach every path of
the device sort (one to three passes of 8-bit digits; tile counts scanned by
a thread, a wave and a workgroup) and of the block scan of the row counts.
"""
import numpy as np
import pytest

from forceatlas2_rc_amd import Batch, GPUError, Server, graphs
from forceatlas2_rc_amd.api import prepare_large
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _messy(n, m, seed):
    """A graph with self-loops, repeated edges, isolated nodes and weights."""
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n - n // 10, size=(m, 2)).astype(np.int32)  # last tenth isolated
    e[: m // 50, 1] = e[: m // 50, 0]  # self-loops
    e = np.concatenate([e, e[: m // 20]])  # repeated edges
    w = rng.uniform(0.2, 3.0, len(e))
    return e, w


def _graph(n, e, w, outbound, seed, **kw):
    pos = graphs.uniform_positions(n, -10 * np.sqrt(n), 10 * np.sqrt(n), seed)
    return dict(n=n, edges=e, weights=w, init=pos, dissuade_hubs=outbound, **kw)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_device_preprocessing_equals_host(precision):
    cases = [(3000, 9000, True), (777, 2000, False), (50, 0, True), (20000, 60000, True),
             (1200, 5000, False)]
    gs, refs = [], []
    for k, (n, m, outbound) in enumerate(cases):
        if m:
            e, w = _messy(n, m, k)
        else:
            e, w = np.zeros((0, 2), np.int32), np.zeros(0)
        gs.append(_graph(n, e, w, outbound, k))
        refs.append(O.large_prepare(n, e, w, outbound))
        for a, c in zip(refs[-1], prepare_large(n, e[:, 0], e[:, 1], w, outbound)):
            assert np.array_equal(a, c)
    with Batch(gs, precision=precision, ingest="device") as b:
        for g, (mass, off, node, coef) in enumerate(refs):
            dm, doff, dnode, dcoef = b.csr(g)
            assert np.array_equal(dm, mass), g
            assert np.array_equal(doff, off), g
            assert np.array_equal(dnode, node), g
            want = coef if precision == "double" else coef.astype(np.float32).astype(np.float64)
            assert np.array_equal(dcoef, want), g


def test_device_preprocessing_at_scale():
    """R-MAT scale 17 (hubs of thousands of incidences, many isolated
    nodes), with dissuade_hubs."""
    n, e = graphs.rmat(17, 1_000_000, seed=3)
    mass, off, node, coef = O.large_prepare(n, e, None, True)
    g = _graph(n, e, None, True, 5)
    with Batch([g], precision="double", ingest="device") as b:
        dm, doff, dnode, dcoef = b.csr(0)
    assert np.array_equal(dm, mass)
    assert np.array_equal(doff, off)
    assert np.array_equal(dnode, node)
    assert np.array_equal(dcoef, coef)


def test_layouts_equal_with_either_preprocessing():
    gs = []
    for k, (n, m) in enumerate([(2500, 8000), (600, 1500)]):
        e, w = _messy(n, m, 10 + k)
        gs.append(_graph(n, e, w, k == 0, 20 + k, linlog=True))
    out = {}
    for ingest in ("host", "device"):
        with Batch(gs, ingest=ingest) as b:
            b.run(15)
            b.sync()
            out[ingest] = [b.positions(g) for g in range(len(gs))]
    for a, c in zip(out["host"], out["device"]):
        assert np.array_equal(a, c)


def test_server_with_device_preprocessing():
    reqs = []
    for k, (n, m) in enumerate([(1800, 6000), (400, 900), (5000, 15000)]):
        e, w = _messy(n, m, 30 + k)
        reqs.append(_graph(n, e, w, k != 1, 40 + k, iterations=12 + 7 * k))
    with Server(max_batches=2, chunk=5) as srv:
        ids = [srv.submit(g, ingest="device") for g in reqs]
        ids.append(srv.submit(reqs[0], ingest="host"))
        res = [srv.result(j)[0] for j in ids]
    for g, pos in zip(reqs + [reqs[0]], res):
        with Batch([g]) as b:
            b.run(g["iterations"])
            b.sync()
            assert np.array_equal(pos, b.positions(0))


def test_device_preprocessing_errors():
    """Node ids out of range: caught on the host for an edge array, and by
    the device's own check for raw from/to arrays."""
    e = np.array([[0, 1], [1, 5]], np.int32)
    with pytest.raises(ValueError, match="does not exist"):
        Batch([_graph(4, e, None, False, 1)], ingest="device")
    for bad in ([0, 1], [1, 5]), ([0, -1], [1, 2]):
        g = _graph(4, None, None, False, 1)
        g["frm"] = np.array(bad[0], np.int32)
        g["to"] = np.array(bad[1], np.int32)
        with pytest.raises(GPUError, match="does not exist"):
            Batch([g], ingest="device")
