"""Pins of the CPU oracle (oracle/fa2_oracle.c) before it is trusted.

Each check is independent of the oracle's own code: the numpy restatement
of the reference's R test oracle (tests/testthat/helper-reference.R:3-100),
closed forms, conservation laws, brute force on tiny inputs, library
routines (numpy's stable argsort), and the relational known-answer tests of
the reference suite (tests/testthat/test-algorithm.R, test-large.R)."""
import numpy as np
import pytest

from forceatlas2_rc_amd import graphs
from oracle import oracle as O
from oracle import ref_numpy as RN


def normwise(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def mean_rel(a, b):
    """testthat's expect_equal tolerance: mean |a - b| / mean |b|."""
    return np.abs(a - b).mean() / np.abs(b).mean()


@pytest.fixture(scope="module")
def pa40():
    n, e = graphs.barabasi_albert(40, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    return n, e, pos


# test-algorithm.R:1-17 and :19-61
OPTION_SETS = [
    dict(),
    dict(strong=True, gravity=0.3),
    dict(linlog=True, scaling=3.0),
    dict(outbound=True, gravity=0.0),
    dict(w="sqrt", jt=0.2),
    dict(adjust=True, size="rand"),
    dict(adjust=True, size="rand", linlog=True, outbound=True, w="rand"),
    dict(fixed=5),
]


def _expand(opts, n, m):
    o = dict(opts)
    rng = np.random.default_rng(3)
    w = rng.uniform(0.2, 3, m)
    size = rng.uniform(0.5, 4, n)
    if o.get("w") == "sqrt":
        o["w"] = np.sqrt(w)
    elif o.get("w") == "rand":
        o["w"] = w
    if o.get("size") == "rand":
        o["size"] = size
    if "fixed" in o:
        o["fixed"] = np.arange(n) < o["fixed"]
    return o


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda o: ",".join(o) or "default")
def test_c_layout_matches_r_transcription(pa40, opts):
    n, e, pos = pa40
    o = _expand(opts, n, len(e))
    scaling = o.pop("scaling", 10.0)
    a = O.layout(n, e, pos, 10, scaling=scaling, **o)
    kw = dict(o)
    jt = kw.pop("jt", 1.0)
    size = kw.pop("size", 1.0)
    b = RN.reference_fa2(n, e, pos, 10, scaling, size=size, jt_user=jt, **kw)
    assert mean_rel(a, b) < 1e-9
    if opts.get("fixed"):
        f = o["fixed"]
        assert np.array_equal(a[f], pos[f])


def test_c_layout_long_run_close_to_r_transcription(pa40):
    n, e, pos = pa40
    a = O.layout(n, e, pos, 100, scaling=10.0)
    b = RN.reference_fa2(n, e, pos, 100, 10.0)
    assert mean_rel(a, b) < 1e-3


def test_update_speed_table():
    """Every branch of fa2_update_speed against the numpy restatement."""
    cases = [
        (100, 0.0, 5.0, 1.0, 1.0, 1.0),        # nothing swings
        (100, 10.0, 2.0, 1.0, 1.0, 1.0),       # erratic: sw > 2 tr
        (100, 10.0, 2.0, 1.0, 1.0, 0.04),      # erratic, efficiency floor
        (100, 1.0, 100.0, 1.0, 1.0, 1.0),      # calm: efficiency rises
        (100, 1.0, 100.0, 1.0, 2000.0, 1.0),   # calm but already fast
        (10 ** 6, 3.0, 5.0, 0.3, 7.0, 0.5),    # jitter clamped low
        (4, 1.0, 1e9, 2.0, 0.5, 1.0),          # jitter clamped at 10
        (1000, 50.0, 40.0, 1.0, 3.0, 0.2),     # sw > jt * tr
    ]
    for n, sw, tr, jt, sp, ef in cases:
        a = O.update_speed(n, sw, tr, jt, sp, ef)
        b = RN.update_speed(n, sw, tr, jt, sp, ef)
        assert a == pytest.approx(b, rel=1e-15, abs=0)


def test_closed_forms():
    # two nodes: |F| = scaling m1 m2 / d, along the line, opposite
    x = np.array([0.0, 3.0])
    y = np.array([0.0, 4.0])
    m = np.array([2.0, 5.0])
    f = O.repulse_exact(x, y, m, 1.5)
    assert np.hypot(*f[0]) == pytest.approx(1.5 * 10 / 5, rel=1e-15)
    assert np.allclose(f[0], -f[1], rtol=0, atol=1e-15)
    assert np.allclose(f[0] / np.hypot(*f[0]), [-0.6, -0.8])
    # overlap with sizes: 100 * scaling * m1 m2 along the line
    f = O.repulse_exact(x, y, m, 1.5, size=np.array([3.0, 3.0]))
    assert np.hypot(*f[0]) == pytest.approx(100 * 1.5 * 10 * 5, rel=1e-15)
    # one iteration of two unconnected nodes = repulsion + gravity only
    pos = np.array([[1.0, 2.0], [-3.0, 0.5]])
    _, _, tf, _ = O.layout(2, np.zeros((0, 2), int), pos, 1, scaling=2.0,
                           gravity=0.7, trace=True)
    rep = O.repulse_exact(pos[:, 0], pos[:, 1], np.ones(2), 2.0)
    grav = -0.7 * pos / np.hypot(pos[:, 0], pos[:, 1])[:, None]
    assert np.allclose(tf[0], rep + grav, rtol=1e-14, atol=0)
    _, _, tf, _ = O.layout(2, np.zeros((0, 2), int), pos, 1, scaling=2.0,
                           gravity=0.7, strong=True, trace=True)
    assert np.allclose(tf[0], rep - 0.7 * pos, rtol=1e-14, atol=0)
    # linlog attraction of one edge: c log(1 + d) along the edge
    pos = np.array([[0.0, 0.0], [1e6, 0.0]])
    _, _, tf, _ = O.layout(2, np.array([[0, 1]]), pos, 1, scaling=1e-30,
                           gravity=0.0, linlog=True, trace=True)
    assert tf[0][0, 0] == pytest.approx(np.log1p(1e6), rel=1e-12)
    assert tf[0][1, 0] == pytest.approx(-np.log1p(1e6), rel=1e-12)


def test_conservation():
    n, e = graphs.barabasi_albert(200, 3, seed=1)
    pos = graphs.uniform_positions(n, -100, 100, 2)
    m = 1.0 + np.bincount(e.ravel(), minlength=n)
    f = O.repulse_exact(pos[:, 0], pos[:, 1], m, 2.0)
    assert np.abs(f.sum(0)).max() < 1e-10 * np.abs(f).sum()
    # attraction without outbound: sum zero (gravity off, repulsion tiny)
    _, _, tf, _ = O.layout(n, e, pos, 1, scaling=1e-30, gravity=0.0,
                           trace=True)
    assert np.abs(tf[0].sum(0)).max() < 1e-10 * np.abs(tf[0]).sum()


def test_barnes_hut_converges_to_exact():
    """test-algorithm.R:63-86."""
    n, e = graphs.barabasi_albert(300, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    run = lambda it, **kw: O.layout(n, e, pos, it, **kw)
    assert mean_rel(run(5, barnes_hut=True, theta=0.0), run(5)) < 1e-9
    assert mean_rel(run(3, barnes_hut=True, theta=0.0, adjust=True, size=2.0),
                    run(3, adjust=True, size=2.0)) < 1e-9
    exact = run(1)
    err = lambda th: np.abs(run(1, barnes_hut=True, theta=th) - exact).max() \
        / np.ptp(exact)
    assert err(0.1) < 1e-4
    assert err(0.1) < err(0.5) < err(1.2) < 0.02


def test_threads_do_not_change_results():
    """test-algorithm.R:88-103."""
    n, e = graphs.barabasi_albert(500, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    for bh in (False, True):
        a = O.layout(n, e, pos, 40, barnes_hut=bh, threads=1)
        b = O.layout(n, e, pos, 40, barnes_hut=bh, threads=4)
        assert np.array_equal(a, b)
    a = O.layout(n, e, pos, 40, adjust=True, threads=1)
    b = O.layout(n, e, pos, 40, adjust=True, threads=4)
    assert np.array_equal(a, b)


def test_degenerate_inputs():
    """test-algorithm.R:162-194."""
    one = O.layout(1, np.zeros((0, 2), int), np.array([[3.0, 4.0]]), 10)
    assert np.all(np.isfinite(one))
    assert np.all(np.isfinite(O.layout(20, np.zeros((0, 2), int),
                                       graphs.uniform_positions(20, -5, 5, 1), 500)))
    e = np.array([[0, 0], [0, 1], [0, 1], [1, 2]])
    for bh in (False, True):
        assert np.all(np.isfinite(O.layout(5, e, graphs.uniform_positions(5, -5, 5, 2),
                                           200, barnes_hut=bh)))
    n, e = graphs.barabasi_albert(50, 2, seed=42)
    zero = np.zeros((50, 2))
    for bh in (False, True):
        assert np.array_equal(O.layout(n, e, zero, 200, barnes_hut=bh), zero)
    pos = graphs.uniform_positions(50, -50, 50, 7)
    pos[:20] = 1.0
    assert mean_rel(O.layout(n, e, pos, 5, barnes_hut=True, theta=0.0),
                    O.layout(n, e, pos, 5)) < 1e-9
    assert np.array_equal(O.layout(n, e, pos, 0), pos)


# --- large-path stages ----------------------------------------------------

def test_spread_bits_brute_force():
    rng = np.random.default_rng(0)
    for v in list(range(70)) + list(rng.integers(0, 1 << 16, 500)) + [0xFFFF]:
        want = sum(((int(v) >> b) & 1) << (2 * b) for b in range(16))
        assert O.lib.orc_spread16(int(v)) == want


def test_morton_hand_cases_and_float32_restatement():
    x = np.array([0.0, 10.0, 10.0, 0.0, 5.0], np.float32)
    y = np.array([0.0, 10.0, 0.0, 10.0, 2.5], np.float32)
    k = O.morton(x, y, "f")
    assert k[0] == 0 and k[1] == 0xFFFFFFFF
    assert k[2] == 0x55555555 and k[3] == 0xAAAAAAAA
    # zero extent: every key 0
    assert np.all(O.morton(np.ones(7, np.float32), np.ones(7, np.float32), "f") == 0)
    # numpy float32 arithmetic, IEEE round to nearest, as written in the
    # reference kernel (pipeline.cl:122-128)
    rng = np.random.default_rng(1)
    for _ in range(5):
        x = rng.normal(0, 1e3, 3000).astype(np.float32)
        y = rng.normal(0, 10, 3000).astype(np.float32)
        mnx, mny = x.min(), y.min()
        ext = max(x.max() - mnx, y.max() - mny)
        sc = np.float32(65535) / ext
        ux = np.clip((x - mnx) * sc, np.float32(0), np.float32(65535)).astype(np.uint32)
        uy = np.clip((y - mny) * sc, np.float32(0), np.float32(65535)).astype(np.uint32)
        spread = lambda v: sum(((v >> b) & 1) << (2 * b) for b in range(16))
        want = spread(ux.astype(np.uint64)) | (spread(uy.astype(np.uint64)) << 1)
        assert np.array_equal(O.morton(x, y, "f"), want.astype(np.uint32))


def test_radix_sort_is_stable_argsort():
    rng = np.random.default_rng(2)
    for n in (1, 2, 17, 5000):
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = keys[0]  # ties
        ks, perm = O.radix_sort(keys)
        want = np.argsort(keys, kind="stable")
        assert np.array_equal(perm, want)
        assert np.array_equal(ks, keys[want])


def _radix_tree_brute(ext):
    """Top-down binary radix tree of sorted distinct 64-bit extended keys:
    a range splits where its highest differing bit changes."""
    n = len(ext)
    nodes = {}

    def build(lo, hi):
        if lo == hi:
            return ("leaf", lo)
        bit = (ext[lo] ^ ext[hi]).bit_length() - 1
        s = lo
        while (ext[s + 1] >> bit) & 1 == 0:
            s += 1
        nodes[(lo, hi)] = s
        build(lo, s)
        build(s + 1, hi)

    build(0, n - 1)
    return nodes


def test_karras_equals_brute_force_radix_tree():
    rng = np.random.default_rng(4)
    for n in (2, 3, 5, 9, 33, 200):
        for dup in (False, True):
            k = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
            if dup:
                k[: n // 2] = k[0]
            k = np.sort(k)
            child, rng_, parent = O.karras(k)
            # equal keys are told apart by their position (pipeline.cl:215)
            ext = [(int(v) << 32) | i for i, v in enumerate(k)]
            want = _radix_tree_brute(ext)
            got = {}
            for r in range(n - 1):
                lo, hi = rng_[2 * r], rng_[2 * r + 1]
                left = child[2 * r]
                split = left - (n - 1) if left >= n - 1 else rng_[2 * left + 1]
                got[(lo, hi)] = split
            assert got == want
            assert rng_[0] == 0 and rng_[1] == n - 1 and parent[0] == -1
            for r in range(n - 1):
                for c in child[2 * r: 2 * r + 2]:
                    assert parent[c] == r


def test_gephi_tree_invariants():
    n, e = graphs.barabasi_albert(800, 2, seed=3)
    pos = graphs.uniform_positions(n, -100, 100, 5)
    pos[:40] = pos[0]  # coincident nodes become single-node leaves
    m = 1.0 + np.bincount(e.ravel(), minlength=n)
    t = O.gephi_tree(pos[:, 0], pos[:, 1], m)
    rd, ri = t["reg_d"], t["reg_i"]
    assert sorted(t["perm"]) == list(range(n))
    assert rd[0, 0] == pytest.approx(m.sum())
    for r in range(t["nreg"]):
        start, count, child, nchild, skip = ri[r]
        idx = t["perm"][start:start + count]
        if count > 1 and nchild == 0 or count > 1:
            cx = (pos[idx, 0] * m[idx]).sum() / m[idx].sum()
            assert rd[r, 1] == pytest.approx(cx, rel=1e-9, abs=1e-9)
            far = np.hypot(pos[idx, 0] - rd[r, 1], pos[idx, 1] - rd[r, 2]).max()
            assert rd[r, 3] == pytest.approx(2 * far, rel=1e-9, abs=1e-12)
        if nchild:
            kids = ri[child: child + nchild]
            assert kids[:, 1].sum() == count
            assert kids[0, 0] == start
            assert kids[-1, 4] == child + nchild - 2 or nchild == 1


@pytest.mark.parametrize("opts", [
    dict(), dict(scaling=5.0, gravity=3.0), dict(linlog=True, strong=True),
    dict(outbound=True, gravity=0.0), dict(w="rand", jt=0.3),
    dict(adjust=True, size="rand"),
    dict(adjust=True, size="rand", linlog=True, outbound=True, w="rand"),
    dict(fixed="alt"),
], ids=lambda o: ",".join(o) or "default")
def test_large_path_without_approximation_is_exact(opts):
    """test-large.R:10-38: theta = 0 in double equals the exact CPU layout,
    with coincident nodes so that equal Morton codes occur."""
    n, e = graphs.barabasi_albert(700, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    pos[:30] = pos[np.repeat(np.arange(3), 10)]
    rng = np.random.default_rng(3)
    o = dict(opts)
    if o.get("w") == "rand":
        o["w"] = rng.uniform(0.2, 3, len(e))
    if o.get("size") == "rand":
        o["size"] = rng.uniform(0.5, 4, n)
    if o.get("fixed") == "alt":
        o["fixed"] = (np.arange(n) % 2 == 0)
    scaling = o.pop("scaling", 2.0)
    a = O.large_run(n, e, pos, 5, mode="d", theta=0.0, scaling=scaling, **o)
    b = O.layout(n, e, pos, 5, scaling=scaling, **o)
    assert mean_rel(a, b) < 1e-8


def test_large_path_theta_sweep():
    """test-large.R:40-57."""
    n, e = graphs.barabasi_albert(3000, 2, seed=42)
    pos = graphs.uniform_positions(n, -50, 50, 7)
    exact = O.layout(n, e, pos, 1)
    err = lambda th, mode="d": np.abs(
        O.large_run(n, e, pos, 1, mode=mode, theta=th) - exact).max() / np.ptp(exact)
    assert err(0.1) < 1e-4
    assert err(0.1) < err(0.5) < err(1.2) < 0.02
    assert err(1.2, "f") < 0.02
    assert err(0.0, "f") < 1e-3
    assert err(0.0, "m") < 1e-5


def test_large_prepare_matches_product_preprocessing():
    from forceatlas2_rc_amd.api import prepare_large
    n, e = graphs.barabasi_albert(500, 3, seed=9)
    e = np.concatenate([e, [[3, 3], [4, 5], [4, 5]]])  # self-loop, multi-edge
    w = np.random.default_rng(1).uniform(0.5, 2, len(e))
    for outbound in (False, True):
        a = O.large_prepare(n, e, w, outbound)
        b = prepare_large(n, e[:, 0], e[:, 1], w, outbound)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_large_regions_equal_sums_over_each_region():
    """Region records (pipeline.cl:276-323) against sums taken directly over
    the nodes each region covers in the oracle's own tree: mass, weighted
    centre, size^2 = 2 (w^2 + h^2) of the bounding box; coincident nodes keep
    the centre inside the box (modes d and m; m sums in float32, so within
    float32 roundings of the sums)."""
    rng = np.random.default_rng(5)
    n = 1500
    x = rng.uniform(-300, 300, n)
    y = rng.uniform(-80, 80, n)
    x[:40] = x[np.arange(40) % 4]
    y[:40] = y[np.arange(40) % 4]
    m = rng.integers(1, 30, n).astype(np.float64)
    for mode in ("d", "m"):
        px = x.astype(np.float32).astype(np.float64) if mode == "m" else x
        py = y.astype(np.float32).astype(np.float64) if mode == "m" else y
        rec = O.large_regions(x, y, m, mode)
        t = O.large_tree(x, y, mode if mode == "d" else "f")
        order = t["order"].astype(np.int64)
        rg = t["range"].reshape(-1, 2)
        for r in range(n - 1):
            idx = order[rg[r, 0]:rg[r, 1] + 1]
            M = m[idx].sum()
            cx = (m[idx] * px[idx]).sum() / M
            cy = (m[idx] * py[idx]).sum() / M
            w = px[idx].max() - px[idx].min()
            h = py[idx].max() - py[idx].min()
            tol = 1e-12 if mode == "d" else 1e-5
            assert rec[r, 2] == M
            scale = max(np.abs(px[idx]).max(), np.abs(py[idx]).max())
            assert abs(rec[r, 0] - cx) <= tol * scale
            assert abs(rec[r, 1] - cy) <= tol * scale
            assert px[idx].min() <= rec[r, 0] <= px[idx].max()
            assert py[idx].min() <= rec[r, 1] <= py[idx].max()
            assert rec[r, 3] == pytest.approx(2 * (w * w + h * h), rel=tol, abs=1e-9)


CLASSES = [(0, 0), (1, 16), (17, 256), (257, 1024), (1025, 1 << 30)]


def test_single_model_takes_the_reference_single_decisions():
    """Mode m (the FP64 model of the single-precision path) decides which
    regions to approximate in float32 from float32 records, as the
    reference's "single" path (mode f) does: its forces differ from f's by
    float32 rounding only, within 1e-5 normwise in every degree class, on
    an R-MAT graph with hubs and on four block-model graphs. The FP64
    decisions of mode d on the same positions now and then take a region
    that f and m open (or the reverse), which moves that node's force by the
    region's approximation error: on one of the block-model graphs this is
    4e-5 of the largest force, so a comparison with d would fail a 1e-5
    gate. At theta = 0 no region is accepted and m equals d bit for bit."""
    kw = dict(theta=1.2, trace=True)
    n, e = graphs.rmat(13, 60_000, seed=21)
    pos = graphs.uniform_positions(n, -900, 900, 22).astype(np.float32).astype(np.float64)
    deg = np.bincount(e[:, 0], minlength=n) + np.bincount(e[:, 1], minlength=n)
    fm = O.large_run(n, e, pos, 1, mode="m", linlog=True, outbound=True, **kw)[1]["force"][0]
    ff = O.large_run(n, e, pos, 1, mode="f", linlog=True, outbound=True, **kw)[1]["force"][0]
    for lo, hi in CLASSES:
        sel = (deg >= lo) & (deg <= hi)
        assert normwise(ff[sel], fm[sel]) <= 1e-5, (lo, hi)
    worst_d = 0.0
    for seed in range(4):
        n, e = graphs.sbm(20000, 8, seed=seed)
        pos = graphs.uniform_positions(n, -1400, 1400, seed + 5)
        pos = pos.astype(np.float32).astype(np.float64)
        fm = O.large_run(n, e, pos, 1, mode="m", **kw)[1]["force"][0]
        ff = O.large_run(n, e, pos, 1, mode="f", **kw)[1]["force"][0]
        fd = O.large_run(n, e, pos, 1, mode="d", **kw)[1]["force"][0]
        assert normwise(ff, fm) <= 1e-5, seed
        worst_d = max(worst_d, normwise(fd, fm))
    assert worst_d > 1e-5
    kw["theta"] = 0.0
    sub = e[(e < 500).all(1)]
    am = O.large_run(500, sub, pos[:500], 2, mode="m", **kw)[1]["force"]
    ad = O.large_run(500, sub, pos[:500], 2, mode="d", **kw)[1]["force"]
    assert np.array_equal(am[0], ad[0])


def _walk_single_numpy(t, theta, scaling, fp64_decisions=False):
    """A second restatement of the single-precision offload model, over the
    oracle's own Gephi tree: float32 records and acceptance test (numpy
    float32 scalars, each operation rounded), FP64 terms; the children of a
    region from the last to the first (src/gpu.c:82-123). With
    fp64_decisions the test is taken in FP64 on the FP64 records instead."""
    f32 = np.float32
    n = len(t["px"])
    rd, ri = t["reg_d"], t["reg_i"]
    px, py, pm = (a.astype(f32) for a in (t["px"], t["py"], t["pm"]))
    th2 = f32(theta * theta)
    out = np.zeros((n, 2))
    for k in range(n):
        xf, yf = px[k], py[k]
        fx = fy = 0.0
        stack = [0]
        while stack:
            r = stack.pop()
            start, count, child, nchild = ri[r, :4]
            if count > 1:
                cx, cy = f32(rd[r, 1]), f32(rd[r, 2])
                if fp64_decisions:
                    ux, uy = float(xf) - rd[r, 1], float(yf) - rd[r, 2]
                    accept = (ux * ux + uy * uy) * (theta * theta) > rd[r, 3] * rd[r, 3]
                else:
                    txd, tyd = f32(xf - cx), f32(yf - cy)
                    d2 = f32(f32(txd * txd) + f32(tyd * tyd))
                    accept = f32(d2 * th2) > f32(rd[r, 3] * rd[r, 3])
                if accept:
                    if fp64_decisions:
                        xd, yd = ux, uy
                    else:
                        xd, yd = float(xf) - float(cx), float(yf) - float(cy)
                    g = scaling * (float(pm[k]) * float(f32(rd[r, 0]))) / (xd * xd + yd * yd)
                    fx += xd * g
                    fy += yd * g
                    continue
                if nchild > 0:
                    stack.extend(range(child, child + nchild))
                    continue
            for j in range(start, start + count):
                if j == k:
                    continue
                xd, yd = float(xf) - float(px[j]), float(yf) - float(py[j])
                d2 = xd * xd + yd * yd
                g = scaling * (float(pm[k]) * float(pm[j])) / d2 if d2 > 0 else 0.0
                fx += xd * g
                fy += yd * g
        out[t["perm"][k]] = (fx, fy)
    return out


def _borderline_theta(t):
    """A theta at which the float32 and the FP64 acceptance tests disagree
    for some node of the tree and some child of the root (which every node
    tests): found by stepping theta across the FP64 threshold."""
    f32 = np.float32
    rd, ri = t["reg_d"], t["reg_i"]
    child, nchild = ri[0, 2], ri[0, 3]
    px, py = t["px"].astype(f32), t["py"].astype(f32)
    for r in range(child, child + nchild):
        if ri[r, 1] < 2:
            continue
        cx, cy, s2 = f32(rd[r, 1]), f32(rd[r, 2]), rd[r, 3] * rd[r, 3]
        for k in np.argsort(np.abs(px.astype(float) - rd[r, 1]))[:200]:
            ux, uy = float(px[k]) - rd[r, 1], float(py[k]) - rd[r, 2]
            th = np.sqrt(s2 / (ux * ux + uy * uy))
            for _ in range(64):
                txd, tyd = f32(px[k] - cx), f32(py[k] - cy)
                a = f32(f32(f32(txd * txd) + f32(tyd * tyd)) * f32(th * th)) > f32(s2)
                if a != ((ux * ux + uy * uy) * (th * th) > s2):
                    return th
                th = np.nextafter(th, np.inf)
    raise AssertionError("no borderline theta found")


def test_single_offload_model():
    """orc_repulse_bh_single (the FP64 model of the single-precision offload)
    against the numpy restatement above on the same tree, within FP64
    rounding, at theta 1.2 and at a theta where float32 and FP64 decisions
    disagree for one (node, region) pair: there the FP64-decision variant
    differs by far more than rounding, so a model that decided in FP64
    would fail. theta = 0 is the exact all-pairs sum on the float32
    inputs."""
    rng = np.random.default_rng(8)
    n = 1500
    x, y = rng.uniform(-300, 300, n), rng.uniform(-300, 300, n)
    x[:60], y[:60] = x[np.arange(60) % 6], y[np.arange(60) % 6]  # coincident nodes
    m = 1.0 + rng.integers(0, 20, n)
    t = O.gephi_tree(x, y, m)
    for theta in (1.2, _borderline_theta(t)):
        got = O.repulse_bh_single(x, y, m, 2.0, theta)
        want = _walk_single_numpy(t, theta, 2.0)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), theta
    other = _walk_single_numpy(t, theta, 2.0, fp64_decisions=True)
    node_err = np.abs(got - other).max(1) / np.abs(got).max(1)
    assert node_err.max() > 1e-4
    xf = x.astype(np.float32).astype(np.float64)
    yf = y.astype(np.float32).astype(np.float64)
    exact = O.repulse_exact(xf, yf, m, 2.0)
    assert normwise(O.repulse_bh_single(x, y, m, 2.0, 0.0), exact) <= 1e-12
