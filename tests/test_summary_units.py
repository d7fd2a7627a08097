"""CPU check of the summary levels of the engine's tree (engine.hip:
k_karras, k_aggregate_tile, k_aggregate_unit), restated in Python over the
oracle's Karras tree (reference numbering, pipeline.cl:226-270).

A region belongs to the smallest unit of sorted positions that holds its
whole range: a block of B positions (level 1), else a unit of B * 16^(L-1)
positions (level L). Gates, on trees of uniform, clustered, duplicate-heavy
and geometrically spaced positions:
  * no unit lists more than 960 regions (the bound behind kUnitCap = 1024);
  * a listed region's children are leaves, regions of lower levels, or
    regions listed in the same unit: never a higher level or another unit;
  * the climb of k_aggregate_unit (children from outside counted as arrived,
    a region with two of them merged first, the second arrival merging the
    parent) merges every region exactly once, after both of its children,
    level after level.
"""
import numpy as np
import pytest

from oracle import oracle as O

FAN = 16


def _tree(x, y):
    t = O.large_tree(np.asarray(x, np.float32), np.asarray(y, np.float32), "f")
    n = len(x)
    return n, t["child"].reshape(n - 1, 2), t["range"].reshape(n - 1, 2)


def _levels(n, rng, B):
    """Level and unit of each region (level 1: inside a block)."""
    first, last = rng[:, 0].astype(np.int64), rng[:, 1].astype(np.int64)
    level = np.ones(len(first), np.int64)
    unit = first // B
    size = B
    crossing = first // size != last // size
    L = 1
    while crossing.any():
        L += 1
        size *= FAN
        level[crossing] = L
        unit[crossing] = first[crossing] // size
        crossing = crossing & (first // size != last // size)
    return level, unit


def _check(n, child, rng, B):
    level, unit = _levels(n, rng, B)
    counts = {}
    for r in np.nonzero(level >= 2)[0]:
        counts[(level[r], unit[r])] = counts.get((level[r], unit[r]), 0) + 1
    assert max(counts.values(), default=0) <= 15 * 64
    parent = np.full(n - 1, -1)
    for r in range(n - 1):
        for c in child[r]:
            if c < n - 1:
                parent[c] = r
                assert level[c] <= level[r]
                if level[c] == level[r] and level[r] >= 2:
                    assert unit[c] == unit[r]
    # the climb, level after level; the block level completes its regions
    done = level == 1
    merged = np.zeros(n - 1, np.int64)
    for L in range(2, int(level.max()) + 1):
        for key in [k for k in counts if k[0] == L]:
            members = np.nonzero((level == L) & (unit == key[1]))[0]
            listed = set(members.tolist())
            vis = {r: sum(1 for c in child[r] if c >= n - 1 or c not in listed)
                   for r in members}
            starts = [r for r in members if vis[r] == 2]
            for r in starts:
                merged[r] += 1
                assert all(c >= n - 1 or done[c] for c in child[r])
                done[r] = True
                while True:
                    p = parent[r]
                    if p < 0 or p not in listed:
                        break
                    vis[p] += 1
                    if vis[p] != 2:
                        break
                    assert all(c >= n - 1 or done[c] for c in child[p])
                    merged[p] += 1
                    done[p] = True
                    r = p
    assert done.all()
    assert np.all(merged[level >= 2] == 1)
    return max(counts.values(), default=0), int(level.max())


def _cases():
    rng = np.random.default_rng(3)
    n = 200_000
    yield "uniform", rng.uniform(-1e3, 1e3, n), rng.uniform(-1e3, 1e3, n)
    c = rng.integers(0, 40, n)
    scale = 10.0 ** rng.uniform(-3, 2, 40)
    cx, cy = rng.uniform(-1e3, 1e3, 40), rng.uniform(-1e3, 1e3, 40)
    yield ("clustered", cx[c] + rng.normal(0, 1, n) * scale[c],
           cy[c] + rng.normal(0, 1, n) * scale[c])
    pts = rng.uniform(0, 100, (60, 2))
    d = pts[rng.integers(0, 60, 120_000)]
    yield "duplicates", d[:, 0], d[:, 1]
    k = np.arange(70_000)
    yield "geometric", 2.0 ** -(k % 40) + 1e-3 * (k // 40), 2.0 ** -((k * 7) % 40)


@pytest.mark.parametrize("B", [1024, 512])
@pytest.mark.parametrize("case", ["uniform", "clustered", "duplicates", "geometric"])
def test_summary_units(case, B):
    x, y = next((a, b) for name, a, b in _cases() if name == case)
    n, child, rng = _tree(x, y)
    most, top = _check(n, child, rng, B)
    assert top >= 2  # the graphs span several blocks
    assert most <= 1024
