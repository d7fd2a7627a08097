"""The oracle against the golden fixtures of tests/golden (written by
tests/golden/make_fixtures.py), and the fixtures against independent
restatements where one is cheap: the numpy restatement of the reference's
own R test oracle for C1, numpy's stable argsort for the sorted
permutation, and a direct numpy sum for sampled all-pairs forces."""
import os

import numpy as np

from oracle import oracle as O
from oracle.ref_numpy import reference_fa2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def test_c1_islands_layout():
    """Exact FP64 CPU layout of the islands graph after 1, 10 and 1000
    iterations (the C1 configuration)."""
    f = load("c1_islands.npz")
    e = f["edges"]
    n = len(f["init"])
    out, tp, _, _ = O.layout(n, e, f["init"], 1000, scaling=2.0, gravity=1.0, trace=True)
    for it in (1, 10, 1000):
        assert np.allclose(tp[it - 1], f[f"pos{it}"], rtol=1e-12, atol=1e-12), it
    assert np.array_equal(out, tp[999])
    ref = reference_fa2(n, e, f["init"], 10, 2.0)
    assert np.abs(ref - f["pos10"]).mean() / np.abs(f["pos10"]).mean() < 1e-9


def test_duplicate_keys_tree():
    f = load("dupkeys_700.npz")
    pos = f["pos"]
    t = O.large_tree(pos[:, 0], pos[:, 1], "f")
    for k in ("keys", "order", "child", "range", "parent"):
        assert np.array_equal(t[k], f[k]), k
    key_unsorted = O.morton(pos[:, 0], pos[:, 1], "f")
    assert np.array_equal(np.argsort(key_unsorted, kind="stable"), f["order"])


def test_c2_exact_forces():
    """Exact FP64 repulsion of 4,096 sampled nodes of the 50,000-node graph:
    the oracle matches the fixture, and a direct numpy sum matches it on 32
    of them (k = scaling m_i m_j / d^2 for d^2 > 0, 0 otherwise)."""
    f = load("c2_exact_4096.npz")
    pos, mass, idx, force = f["pos"], f["mass"].astype(np.float64), f["idx"], f["force"]
    x, y = pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64)
    got = O.repulse_exact(x, y, mass, 2.0, threads=os.cpu_count() or 1)[idx]
    assert np.allclose(got, force, rtol=1e-12, atol=1e-9 * np.abs(force).max())
    for i, fi in zip(idx[:32], force[:32]):
        xd, yd = x[i] - x, y[i] - y
        d2 = xd * xd + yd * yd
        k = np.where(d2 > 0, 2.0 * mass[i] * mass / np.where(d2 > 0, d2, 1.0), 0.0)
        direct = np.array([np.sum(xd * k), np.sum(yd * k)])
        assert np.allclose(direct, fi, rtol=1e-9, atol=1e-9 * np.abs(fi).max())
