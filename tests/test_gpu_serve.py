"""Batched iterations replayed from a HIP graph, and the layout server
(fa2_serve_* of include/fa2_hip.h).

Gates: a graph's positions are bit-identical
  * whether its iterations are launched one kernel at a time or replayed
    from the captured HIP graph, over several replays with eager work and
    state changes in between;
  * alone (a batch of one) and inside the server's mixed batches, whatever
    else is running.
"""
import numpy as np
import pytest

from forceatlas2_rc_amd import Batch, GPUError, Server, graphs, serve_footprint

pytestmark = pytest.mark.gpu


def _graph(n, seed, m=3, **kw):
    n, e = graphs.barabasi_albert(n, m, seed=seed)
    pos = graphs.uniform_positions(n, -10 * np.sqrt(n), 10 * np.sqrt(n), seed + 1)
    return dict(n=n, edges=e, init=pos, **kw)


def _alone(g, precision="single"):
    with Batch([g], precision=precision) as b:
        b.run(g["iterations"])
        b.sync()
        return b.positions(0)


def test_graph_replay_equals_eager_launches():
    """profile=True launches every kernel; profile=False replays a captured
    pair of iterations. Checked after odd and even counts (the replay starts
    from an aligned force buffer), after set_state and across graphs."""
    gs = [_graph(3000, 1, linlog=True, dissuade_hubs=True), _graph(700, 2, adjust=True,
                                                                   size=2.0)]
    with Batch(gs, profile=True) as eager, Batch(gs, profile=False) as replay:
        for steps in (3, 4, 1, 6, 5):
            eager.run(steps)
            replay.run(steps)
            eager.sync()
            replay.sync()
            for g in range(2):
                assert np.array_equal(eager.positions(g), replay.positions(g)), steps
        p = eager.positions(0) * 0.5
        eager.set_state(0, p)
        replay.set_state(0, p)
        eager.run(8)
        replay.run(8)
        eager.sync()
        replay.sync()
        for g in range(2):
            assert np.array_equal(eager.positions(g), replay.positions(g))


def test_server_results_equal_each_layout_alone():
    """Requests of uneven sizes, settings and iteration counts, submitted
    together so that several share a batch and several batches run at once:
    each result equals the layout of that graph alone, bit for bit."""
    reqs = [_graph(2000, 11, iterations=40),
            _graph(500, 12, iterations=75, linlog=True),
            _graph(9000, 13, iterations=30, dissuade_hubs=True),
            _graph(300, 14, iterations=120, adjust=True, size=1.5),
            _graph(4100, 15, iterations=55, strong=True, gravity=0.5),
            _graph(1200, 16, iterations=25, fixed=np.arange(1200) % 20 == 0)]
    with Server(max_batches=3, chunk=10) as srv:
        ids = []
        for k, g in enumerate(reqs):
            ids.append(srv.submit(g))
            if k == 2:  # the first batches are under way when the rest arrive
                srv.result(ids[0])
        out = [srv.result(j) for j in ids]
        st = srv.stats()
    assert st["admitted"] == len(reqs)
    assert st["batches"] >= 2
    for g, (pos, t) in zip(reqs, out):
        assert np.array_equal(pos, _alone(g))
        assert t["total"] >= t["running"] > 0


def test_server_mixed_precision():
    """Single and double requests interleaved: each batch holds one
    precision, and each result equals the layout alone in its precision."""
    reqs = [(_graph(800, 21, iterations=20), "double"),
            (_graph(2500, 22, iterations=35, linlog=True), "single"),
            (_graph(1500, 23, iterations=28, adjust=True, size=2.0), "double"),
            (_graph(600, 24, iterations=16), "single")]
    with Server(max_batches=2, chunk=8) as srv:
        ids = [srv.submit(g, prec) for g, prec in reqs]
        out = [srv.result(j)[0] for j in ids]
        assert srv.stats()["batches"] >= 2
    for (g, prec), pos in zip(reqs, out):
        assert np.array_equal(pos, _alone(g, prec))


def test_server_admission_by_footprint():
    """A budget that holds one request at a time runs them one after the
    other; a request larger than the budget fails alone."""
    a = _graph(3000, 31, iterations=15)
    b = _graph(3000, 32, iterations=15)
    one = serve_footprint(3000, 2 * len(a["edges"]))
    with Server(budget_bytes=1.5 * one, max_batches=4, chunk=5) as srv:
        ia, ib = srv.submit(a), srv.submit(b)
        (pa, ta), (pb, tb) = srv.result(ia), srv.result(ib)
        assert srv.stats()["peak_running"] == 1
    assert np.array_equal(pa, _alone(a))
    assert np.array_equal(pb, _alone(b))
    with Server(budget_bytes=0.5 * one) as srv:
        j = srv.submit(a)
        with pytest.raises(GPUError, match="budget"):
            srv.result(j)
        assert srv.status(j) == "failed"


def test_server_argument_errors():
    from forceatlas2_rc_amd._lib import lib
    with pytest.raises(GPUError):
        Server(max_batches=0)
    with pytest.raises(GPUError):
        Server(device=99)
    with Server() as srv:
        assert lib.fa2_serve_status(srv.h, 999) == -1
        assert lib.fa2_serve_wait(srv.h, 999, None, None) == 1
        with pytest.raises(GPUError):
            srv.submit(_graph(100, 41, iterations=0))


def test_waiting_requests_join_a_running_batch():
    """Continuous batching: with one stream, requests submitted while a
    batch runs join it at a chunk boundary once one of its requests has
    finished (the batch is rebuilt with the remaining request's state);
    requests of the other precision wait for the stream. Every result equals
    the layout alone bit for bit, the continuing one included."""
    long_ = _graph(3000, 51, iterations=400, linlog=True)
    short = _graph(800, 52, iterations=10)
    late = [_graph(1500, 53, iterations=35, dissuade_hubs=True),
            _graph(600, 54, iterations=60, adjust=True, size=1.5),
            _graph(2200, 55, iterations=500)]
    other = _graph(900, 56, iterations=20)
    with Server(max_batches=1, chunk=5) as srv:
        i_long, i_short = srv.submit(long_), srv.submit(short)
        srv.result(i_short)  # the batch runs on; a slot is free
        ids = [srv.submit(g) for g in late]
        i_other = srv.submit(other, "double")
        res = [srv.result(j)[0] for j in [i_long] + ids]
        res_other = srv.result(i_other)[0]
        st = srv.stats()
    assert st["refills"] >= 1
    for g, pos in zip([long_] + late, res):
        assert np.array_equal(pos, _alone(g))
    assert np.array_equal(res_other, _alone(other, "double"))


def test_requests_from_edge_lists_join_a_running_batch():
    """Continuous batching of requests whose incidence lists the device
    builds from their edge lists (fa2_serve_submit_edges): rebuilt batches
    prepare their newcomers on the device beside the resumed requests, and
    every result equals the layout alone bit for bit."""
    long_ = _graph(2500, 61, iterations=300, linlog=True, dissuade_hubs=True)
    short = _graph(700, 62, iterations=10)
    late = [_graph(1800, 63, iterations=40), _graph(900, 64, iterations=70, adjust=True, size=1.2)]
    with Server(max_batches=1, chunk=5) as srv:
        i_long, i_short = srv.submit(long_, ingest="device"), srv.submit(short, ingest="device")
        srv.result(i_short)
        ids = [srv.submit(g, ingest="device") for g in late]
        res = [srv.result(j)[0] for j in [i_long] + ids]
        st = srv.stats()
    assert st["refills"] >= 1
    for g, pos in zip([long_] + late, res):
        assert np.array_equal(pos, _alone(g))
