"""Failure paths of the all-GPU layout, through the C ABI.

The reference's fa2_full_run (src/gpu_full.c:164-532) stops with status 1
and a message in two cases: the caller's interrupt callback returns non-zero
(src/gpu_full.c:404-411, polled once per 10^6 node and incidence visits;
src/forceatlas2.c:927 passes R's interrupt check), and the swing or traction
totals are not numbers (src/gpu_full.c:486-491). Here the callback is also
polled while the host waits for the device, and invalid numbers are seen
once per run: a graph whose totals stop being numbers stops moving (status
2 on the device), and its request fails alone in a batch or in the server.
"""
import ctypes

import numpy as np
import pytest

from forceatlas2_rc_amd import Batch, GPUError, Server, graphs
from forceatlas2_rc_amd._lib import FA2_FULL_NTIMES, INTERRUPT_FN, FullParams, last_error, lib
from forceatlas2_rc_amd.api import _fill_params

pytestmark = pytest.mark.gpu


def _graph(n, seed, **kw):
    n, e = graphs.barabasi_albert(n, 3, seed=seed)
    pos = graphs.uniform_positions(n, -10 * np.sqrt(n), 10 * np.sqrt(n), seed + 1)
    return dict(n=n, edges=e, init=pos, **kw)


def _full_run(g, callback=None, precision="single"):
    """fa2_full_run on graph dict g: (status, message, positions)."""
    p = FullParams()
    keep, xy = _fill_params(p, g, 0, precision, False)
    if callback is not None:
        p.interrupted = INTERRUPT_FN(callback)
    times = np.zeros(FA2_FULL_NTIMES)
    st = lib.fa2_full_run(ctypes.byref(p), xy.ctypes.data_as(ctypes.c_void_p),
                          times.ctypes.data_as(ctypes.c_void_p))
    del keep
    return st, (last_error() if st else ""), xy.reshape(2, -1).T.copy()


@pytest.mark.parametrize("precision", ["single", "double"])
def test_interrupt_stops_the_layout(precision):
    g = _graph(3000, 1, iterations=4000, linlog=True)
    calls = []

    def stop_on_third():
        calls.append(1)
        return int(len(calls) >= 3)

    st, msg, _ = _full_run(g, stop_on_third, precision)
    assert st == 1
    assert msg == "the layout was interrupted"
    assert len(calls) == 3


def test_callback_that_never_interrupts_changes_nothing():
    g = _graph(2000, 2, iterations=401, dissuade_hubs=True)
    calls = []

    def never():
        calls.append(1)
        return 0

    st0, _, ref = _full_run(g)
    st1, _, got = _full_run(g, never)
    assert st0 == st1 == 0
    assert len(calls) >= 3  # 401 iterations of 14,000 visits: the engine polls 5 times
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_invalid_numbers_stop_the_layout(precision):
    g = _graph(2000, 3, iterations=20)
    g["init"] = g["init"].copy()
    g["init"][17] = np.nan
    st, msg, _ = _full_run(g, precision=precision)
    assert st == 1
    assert msg == "the computation on the GPU produced invalid numbers"
    # the same through a callback-driven run
    st, msg, _ = _full_run(g, lambda: 0, precision)
    assert (st, msg) == (1, "the computation on the GPU produced invalid numbers")


def test_invalid_graph_fails_alone_in_a_batch():
    """In a batch the graph with a NaN position gets status 2 and stops; the
    other graphs give their layouts alone bit for bit."""
    good = _graph(1500, 4, iterations=12, linlog=True)
    bad = _graph(900, 5, iterations=12)
    bad["init"] = bad["init"].copy()
    bad["init"][3, 1] = np.inf
    with Batch([good, bad, good]) as b:
        b.run(12)
        with pytest.raises(GPUError, match="invalid numbers"):
            b.sync()
        done, status = b.progress()
        assert status.tolist() == [0, 2, 0]
        assert done[0] == done[2] == 12 and done[1] < 12
        together = b.positions(0), b.positions(2)
    with Batch([good]) as one:
        one.run(12)
        one.sync()
        alone = one.positions(0)
    assert np.array_equal(together[0], alone)
    assert np.array_equal(together[1], alone)


def test_invalid_request_fails_alone_in_the_server():
    reqs = [_graph(2500, 6, iterations=30), _graph(800, 7, iterations=25),
            _graph(1200, 8, iterations=40, adjust=True, size=1.5)]
    reqs[1]["init"] = reqs[1]["init"].copy()
    reqs[1]["init"][0, 0] = np.nan
    with Server(max_batches=2, chunk=10) as srv:
        ids = [srv.submit(g) for g in reqs]
        ok0 = srv.result(ids[0])[0]
        with pytest.raises(GPUError, match="invalid numbers"):
            srv.result(ids[1])
        assert srv.status(ids[1]) == "failed"
        ok2 = srv.result(ids[2])[0]
    for g, pos in ((reqs[0], ok0), (reqs[2], ok2)):
        with Batch([g]) as one:
            one.run(g["iterations"])
            one.sync()
            assert np.array_equal(one.positions(0), pos)
