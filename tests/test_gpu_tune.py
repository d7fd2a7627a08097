"""The autotuner of a batch's launch shapes (fa2_batch_tune,
fa2_batch_set_tuning of include/fa2_hip.h).

Gates: every candidate shape (the attraction's fork point included) gives
positions bit-identical to the default shapes, with the iterations replayed
from the captured HIP graph (which is
recaptured after each change) and launched kernel by kernel; tuning leaves
the layout state untouched; the tuner keeps a shape other than the default
only when it timed it at least 2% faster;
a large server batch (default shapes, no second stream) gives each
graph's layout alone.
"""
import numpy as np
import pytest

from forceatlas2_rc_amd import Batch, GPUError, Server, graphs

pytestmark = pytest.mark.gpu

SHAPES = [((1, 1, 1), 64, 0, 0), ((2, 2, 2), 128, 2, 1), ((8, 8, 8), 256, 32, 3),
          ((4, 1, 8), 64, 8, 2), ((8, 4, 2), 128, 1, 0), ((4, 4, 4), 256, 4096, 1),
          ((2, 4, 8), 512, 8, 3), ((4, 8, 1), 1024, 1, 2)]


def _graphs():
    """Every attraction class occupied: R-MAT hubs of thousands of
    incidences (wave items and hub partials), preferential attachment
    (16-lane class), a sparse block model (one lane per node); linlog,
    dissuade_hubs and node sizes mixed in."""
    n, e = graphs.rmat(14, 120_000, seed=5)
    r = 10 * np.sqrt(n)
    a = dict(n=n, edges=e, init=graphs.uniform_positions(n, -r, r, 6), linlog=True,
             dissuade_hubs=True)
    n2, e2 = graphs.barabasi_albert(6000, 12, seed=7)
    b = dict(n=n2, edges=e2, init=graphs.uniform_positions(n2, -800, 800, 8))
    n3, e3 = graphs.sbm(3000, 6, deg_in=4.0, deg_out=1.0, seed=9)
    c = dict(n=n3, edges=e3, init=graphs.uniform_positions(n3, -500, 500, 10), adjust=True,
             size=np.random.default_rng(11).uniform(1.0, 3.0, n3))
    return [a, b, c]


def _layout(gs, iters, shape=None, profile=False, tune=False, precision="single"):
    with Batch(gs, profile=profile, autotune=False, precision=precision) as b:
        if shape is not None:
            b.set_tuning(*shape)
        if tune:
            b.tune(force=True)
        b.run(iters)
        b.sync()
        return [b.positions(g) for g in range(len(gs))]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_every_shape_gives_the_same_layout(precision):
    gs = _graphs()
    ref = _layout(gs, 7, precision=precision)
    for shape in SHAPES:
        for profile in (False, True):
            out = _layout(gs, 7, shape, profile, precision=precision)
            for g in range(len(gs)):
                assert np.array_equal(out[g], ref[g]), (shape, profile, g)


def test_tuning_keeps_the_state_and_a_faster_shape():
    gs = _graphs()
    ref = _layout(gs, 6)
    with Batch(gs, autotune=False) as b:
        b.run(3)
        chosen, times = b.tune(force=True)
        assert times is not None and np.all(times[:10] > 0)
        assert np.all(times[10:22] > 0)  # every attraction class is occupied
        assert np.all(times[22:26] > 0)
        # the fastest candidate replaces the default (256 threads, XCD runs
        # of 32 blocks, unroll 4, attraction forked after the tree) when it
        # was timed at least 2% faster
        def kept(cands, values, chosen_value, default):
            t = dict(zip(cands, values))
            fastest = min(cands, key=lambda c: t[c])
            want = fastest if t[fastest] < 0.98 * t[default] else default
            assert chosen_value == want
        # a size other than 256 also has to win beside the attraction
        threads = dict(zip((64, 128, 256, 512, 1024), times[:5]))
        alone = min(threads, key=lambda c: threads[c])
        if threads[alone] < 0.98 * threads[256] and alone != 256:
            assert times[26] > 0 and times[27] > 0
            want = alone if times[26] < 0.98 * times[27] else 256
        else:
            assert times[26] == 0 and times[27] == 0
            want = 256
        assert chosen["repulse_threads"] == want
        kept((0, 8, 32, 128, 512), times[5:10], chosen["repulse_xcd_run"], 32)
        for c in range(3):
            kept((1, 2, 4, 8), times[10 + 4 * c:14 + 4 * c], chosen["attract_unroll"][c], 4)
        kept((0, 1, 2, 3), times[22:26], chosen["attract_fork"], 2)
        assert b.tuning == chosen
        b.run(3)
        b.sync()
        for g in range(len(gs)):
            assert np.array_equal(b.positions(g), ref[g])
        # a second batch of the same size class reuses the choice untimed
        with Batch(gs, autotune=False) as b2:
            again, t2 = b2.tune(force=False)
            assert t2 is None and again == chosen
    assert np.array_equal(_layout(gs, 6, tune=True)[0], ref[0])


def test_tuning_errors():
    gs = _graphs()[1:2]
    with Batch(gs, autotune=False) as b:
        for bad in (((3, 4, 4), 256, 0), ((4, 4, 4), 96, 0), ((4, 4, 4), 2048, 0), ((4, 0, 4), 64, 0),
                    ((4, 4, 4), 256, -1), ((4, 4, 4), 256, 5000), ((4, 4, 4), 256, 32, -1),
                    ((4, 4, 4), 256, 32, 4)):
            with pytest.raises(GPUError):
                b.set_tuning(*bad)
        assert b.tuning == dict(attract_unroll=(4, 4, 4), repulse_threads=256, repulse_xcd_run=32,
                                attract_fork=2)


def test_server_large_batch_equals_each_graph_alone():
    reqs = []
    for k, n in enumerate((70_000, 5_000)):  # the first alone is above the size
        n, e = graphs.barabasi_albert(n, 4, seed=60 + k)
        r = 10 * np.sqrt(n)
        reqs.append(dict(n=n, edges=e, init=graphs.uniform_positions(n, -r, r, 70 + k),
                         iterations=9 + 4 * k, linlog=k == 1))
    with Server(max_batches=1, chunk=5) as srv:
        ids = [srv.submit(g) for g in reqs]
        out = [srv.result(j)[0] for j in ids]
    for g, pos in zip(reqs, out):
        with Batch([g], autotune=False) as b:
            b.run(g["iterations"])
            b.sync()
            assert np.array_equal(pos, b.positions(0))
