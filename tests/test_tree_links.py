"""CPU check of the index arithmetic of the engine's tree (engine.hip:
k_karras, the skip links written by k_aggregate, the stackless walk of
k_repulse), restated in Python over a batch of graphs stored side by side.

The stackless walk must visit, for every node, exactly the regions the
reference's stack traversal visits (inst/opencl/pipeline.cl:344-370, left
child first), in the same order, and must end. The topology is compared with
the oracle's Karras tree (reference numbering)."""
import numpy as np
import pytest

from oracle import oracle as O

END = -(1 << 31)


def clz32(v):
    return 32 - int(v).bit_length()


def kdelta(ks, n, a, b):
    if b < 0 or b >= n:
        return -1
    x, y = int(ks[a]), int(ks[b])
    if x == y:
        return 32 + clz32(a ^ b)
    return clz32(x ^ y)


def engine_tree(keys_per_graph):
    """k_karras + skip links for graphs laid side by side."""
    ns = [len(k) for k in keys_per_graph]
    G = len(ns)
    noff = np.concatenate([[0], np.cumsum(ns)])[:-1]
    ioff = noff - np.arange(G)
    N = sum(ns)
    NI = N - G
    child = [None] * NI
    tl = [None] * NI
    parent_leaf = [None] * N
    parent_int = [None] * NI
    escape = [None] * NI
    for g in range(G):
        n, base, ks = ns[g], int(noff[g]), keys_per_graph[g]
        for i in range(n - 1):
            t = int(ioff[g]) + i
            d = 1 if kdelta(ks, n, i, i + 1) > kdelta(ks, n, i, i - 1) else -1
            dmin = kdelta(ks, n, i, i - d)
            lmax = 2
            while kdelta(ks, n, i, i + lmax * d) > dmin:
                lmax *= 2
            l = 0
            s = lmax // 2
            while s >= 1:
                if kdelta(ks, n, i, i + (l + s) * d) > dmin:
                    l += s
                s //= 2
            j = i + l * d
            dnode = kdelta(ks, n, i, j)
            s, st = 0, l
            while True:
                st = (st + 1) >> 1
                if kdelta(ks, n, i, i + (s + st) * d) > dnode:
                    s += st
                if st <= 1:
                    break
            split = i + s * d + min(d, 0)
            first, last = min(i, j), max(i, j)
            S = base + split
            left = ~S if first == split else S - g
            right = ~(S + 1) if last == split + 1 else S + 1 - g
            child[t] = (left, right)
            for c in (left, right):
                if c >= 0:
                    parent_int[c] = t
                else:
                    parent_leaf[~c] = t
            if i == 0:
                parent_int[t] = -1
            escape[S - g] = right
            tl[t] = [left, 0, base + first, base + last]
    for g in range(G):
        gend = int(noff[g]) + ns[g]
        for i in range(ns[g] - 1):
            t = int(ioff[g]) + i
            last = tl[t][3]
            tl[t][1] = END if last == gend - 1 else escape[last - g]
    return dict(ns=ns, noff=noff, ioff=ioff, child=child, tl=tl,
                escape=escape, parent_leaf=parent_leaf, parent_int=parent_int)


def stackless_visits(T, g, k, accept):
    """Regions and leaves k_repulse visits for sorted node k of graph g;
    accept(r) says whether region r (global id) is far enough."""
    n, base = T["ns"][g], int(T["noff"][g])
    gend = base + n
    cur = int(T["ioff"][g])
    out = []
    for _ in range(2 * n):
        if cur == END:
            break
        if cur >= 0:
            left, skip, first, last = T["tl"][cur]
            out.append(("r", cur - int(T["ioff"][g])))
            if accept(cur):
                cur = skip
                continue
            if last - first >= 8:
                cur = left
                continue
            out.append(("b", first - base, last - base))
            cur = skip
        else:
            p = ~cur
            # k_aggregate_tile writes a leaf's record only when the leaf's
            # parent holds more than 8 nodes: a walk never reaches another
            par = T["tl"][T["parent_leaf"][p]]
            assert par[3] - par[2] >= 8, "a leaf record under a small region"
            out.append(("l", p - base))
            cur = END if p == gend - 1 else T["escape"][p - g]
    assert cur == END, "the walk did not end"
    return out


def stack_visits(n, child, rng_, k, accept_ref):
    """The reference traversal (pipeline.cl:344-370) in reference numbering."""
    out = []
    stack = [0]
    while stack:
        r = stack.pop()
        if r >= n - 1:
            out.append(("l", r - (n - 1)))
            continue
        out.append(("r", r))
        if accept_ref(r):
            continue
        first, last = rng_[2 * r], rng_[2 * r + 1]
        if last - first >= 8:
            stack.append(child[2 * r + 1])
            stack.append(child[2 * r])
            continue
        out.append(("b", first, last))
    return out


def _keys(rng, n, dup):
    k = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    if dup:
        k[: n // 2] = k[0]  # many equal keys: ties broken by position
        k[n // 2: n // 2 + 7] = 12345
    return np.sort(k, kind="stable")


@pytest.mark.parametrize("dup", [False, True])
def test_topology_and_walk(dup):
    rng = np.random.default_rng(5 + dup)
    sizes = [2, 3, 9, 40, 257]
    keys = [_keys(rng, n, dup) for n in sizes]
    T = engine_tree(keys)
    for g, n in enumerate(sizes):
        child, rng_, parent = O.karras(keys[g])
        io, no = int(T["ioff"][g]), int(T["noff"][g])
        ref = lambda r: r - io if r >= 0 else n - 1 + (~r - no)
        for i in range(n - 1):
            c = T["child"][io + i]
            assert (ref(c[0]), ref(c[1])) == (child[2 * i], child[2 * i + 1])
            assert T["tl"][io + i][2] - no == rng_[2 * i]
            assert T["tl"][io + i][3] - no == rng_[2 * i + 1]
        # walks: accept a pseudo-random half of the regions
        acc = rng.random(max(1, n - 1)) < 0.4
        for k in range(0, n, max(1, n // 17)):
            a = stackless_visits(T, g, k, lambda r: bool(acc[r - io]))
            b = stack_visits(n, child, rng_, k, lambda r: bool(acc[r]))
            assert a == b
        # nothing accepted: every node is reached exactly once
        a = stackless_visits(T, g, 0, lambda r: False)
        seen = []
        for v in a:
            if v[0] == "l":
                seen.append(v[1])
            elif v[0] == "b":
                seen.extend(range(v[1], v[2] + 1))
        assert sorted(seen) == list(range(n))


def wave_visits(T, g, lanes, accept):
    """The wave walk of k_repulse: the union of the lanes' visits in
    depth-first order, each lane taking part only where its own traversal
    goes. Returns the sequence each lane took part in."""
    n, base = T["ns"][g], int(T["noff"][g])
    gend = base + n
    io = int(T["ioff"][g])
    done = {k: -(1 << 31) + 1 for k in lanes}
    seq = {k: [] for k in lanes}
    c = io
    steps = 0
    while c != END:
        steps += 1
        assert steps <= 2 * n
        if c >= 0:
            left, skip, first, last = T["tl"][c]
            act = [k for k in lanes if first > done[k]]
            acc = [k for k in act if accept(k, c)]
            opn = [k for k in act if k not in acc]
            for k in act:
                seq[k].append(("r", c - io))
            for k in acc:
                done[k] = last
            if not opn:
                c = skip
                continue
            if last - first >= 8:
                c = left
                continue
            for k in opn:
                seq[k].append(("b", first - base, last - base))
                done[k] = last
            c = skip
        else:
            p = ~c
            par = T["tl"][T["parent_leaf"][p]]
            assert par[3] - par[2] >= 8, "a leaf record under a small region"
            for k in lanes:
                if p > done[k]:
                    seq[k].append(("l", p - base))
                    done[k] = p
            c = END if p == gend - 1 else T["escape"][p - g]
    return seq


@pytest.mark.parametrize("dup", [False, True])
def test_wave_walk_gives_each_lane_its_own_traversal(dup):
    rng = np.random.default_rng(11 + dup)
    sizes = [3, 70, 300]
    keys = [_keys(rng, n, dup) for n in sizes]
    T = engine_tree(keys)
    for g, n in enumerate(sizes):
        child, rng_, parent = O.karras(keys[g])
        io = int(T["ioff"][g])
        # each lane accepts a region with its own pseudo-random rule,
        # correlated with its position like a real distance test
        salt = rng.integers(0, 1 << 30, n)
        acc = lambda k, r: ((int(salt[k]) // 7 + int(r) * 2654435761) >> 3) % 3 == 0 \
            and abs(rng_[2 * r] - k) > 4
        lanes = list(range(min(n, 64)))
        got = wave_visits(T, g, lanes, lambda k, c: acc(k, c - io))
        for k in lanes:
            want = stack_visits(n, child, rng_, k, lambda r: acc(k, r))
            assert got[k] == want
