"""Seeded synthetic graphs of the benchmark configurations.

All generators use numpy's PCG64 so that the same seed gives the same graph
on every machine. Edges are returned as an (m, 2) int32 array, 0-based, in
generation order; undirected graphs list each edge once.
"""
import numpy as np

__all__ = ["islands", "barabasi_albert", "rmat", "sbm", "uniform_positions"]


def uniform_positions(n, lo, hi, seed):
    """(n, 2) positions uniform in [lo, hi)^2."""
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(n, 2))


def islands(n_islands=5, island_size=30, p_in=0.3, n_between=2, seed=0):
    """Islands graph: G(island_size, p_in) inside each island, plus
    n_between random edges between every pair of islands (the shape of
    igraph::sample_islands)."""
    rng = np.random.default_rng(seed)
    edges = []
    iu, ju = np.triu_indices(island_size, k=1)
    for k in range(n_islands):
        keep = rng.random(iu.size) < p_in
        off = k * island_size
        edges.append(np.stack([iu[keep] + off, ju[keep] + off], 1))
    for a in range(n_islands):
        for b in range(a + 1, n_islands):
            u = rng.integers(0, island_size, n_between) + a * island_size
            v = rng.integers(0, island_size, n_between) + b * island_size
            edges.append(np.stack([u, v], 1))
    return n_islands * island_size, np.concatenate(edges).astype(np.int32)


def barabasi_albert(n, m=3, seed=0):
    """Preferential attachment: node t attaches to min(m, t) distinct earlier
    nodes chosen with probability proportional to degree + 1."""
    rng = np.random.default_rng(seed)
    # urn of node ids: one ticket per node (the +1) and one per edge end
    urn = np.empty(n + 2 * m * n, np.int64)
    size = 0
    edges = np.empty((max(0, (n - 1)) * m, 2), np.int64)
    ne = 0
    urn[size] = 0
    size += 1
    draws = rng.random(8 * m * n + 64)
    di = 0
    for t in range(1, n):
        k = min(m, t)
        chosen = []
        while len(chosen) < k:
            if di >= draws.size:
                draws = rng.random(8 * m * n)
                di = 0
            c = int(urn[int(draws[di] * size)])
            di += 1
            if c not in chosen:
                chosen.append(c)
        for c in chosen:
            edges[ne] = (t, c)
            ne += 1
            urn[size] = c
            size += 1
        urn[size:size + k] = t
        size += k
        urn[size] = t
        size += 1
    return n, edges[:ne].astype(np.int32)


def preferential_attachment(n, m=3, seed=0):
    """Preferential attachment for large n, vectorised by generations: the
    nodes of a generation (a quarter of the nodes so far) each draw m
    earlier nodes with probability proportional to degree + 1 as it was
    when the generation began; repeated draws of one node collapse into one
    edge. Starts from a clique of m + 1 nodes."""
    rng = np.random.default_rng(seed)
    k0 = min(n, m + 1)
    a, b = np.tril_indices(k0, -1)
    edges = [np.stack([a, b], 1)]
    urn = np.concatenate([a, b, np.arange(k0)])
    t = k0
    while t < n:
        c = min(n - t, max(64, t // 4))
        new = np.arange(t, t + c)
        picks = urn[rng.integers(0, urn.size, size=(c, m))]
        e = np.stack([np.repeat(new, m), picks.ravel()], 1)
        key = e[:, 0].astype(np.int64) * n + e[:, 1]
        _, idx = np.unique(key, return_index=True)
        e = e[np.sort(idx)]
        edges.append(e)
        urn = np.concatenate([urn, e[:, 0], e[:, 1], new])
        t += c
    return n, np.concatenate(edges).astype(np.int32)


def rmat(scale=21, n_edges=16_000_000, a=0.57, b=0.19, c=0.19, seed=0,
         chunk=1 << 22):
    """R-MAT graph on 2^scale nodes: edges drawn quadrant by quadrant with
    probabilities (a, b, c, d), made undirected (u < v), self-loops and
    duplicates dropped, until n_edges distinct edges exist. The kept edges
    are the first occurrences, in generation order."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    keys = []
    have = 0
    seen = np.empty(0, np.int64)
    first_pos = np.empty(0, np.int64)
    total = 0
    pa, pab, pabc = a, a + b, a + b + c
    while have < n_edges:
        k = chunk
        u = np.zeros(k, np.int64)
        v = np.zeros(k, np.int64)
        for lvl in range(scale):
            r = rng.random(k)
            bit_u = (r >= pab).astype(np.int64)            # quadrants c, d
            bit_v = ((r >= pa) & (r < pab)) | (r >= pabc)  # quadrants b, d
            u |= bit_u << lvl
            v |= bit_v.astype(np.int64) << lvl
        lo = np.minimum(u, v)
        hi = np.maximum(u, v)
        ok = lo != hi
        key = (lo[ok] << scale) | hi[ok]
        pos = np.nonzero(ok)[0] + total
        total += k
        allk = np.concatenate([seen, key])
        allp = np.concatenate([first_pos, pos])
        uk, idx = np.unique(allk, return_index=True)
        seen = uk
        first_pos = allp[idx]
        have = uk.size
    order = np.argsort(first_pos, kind="stable")[:n_edges]
    key = seen[order]
    edges = np.stack([key >> scale, key & (n - 1)], 1).astype(np.int32)
    return n, edges


def sbm(n, n_blocks, deg_in=8.0, deg_out=1.0, seed=0):
    """Stochastic block model with n_blocks equal blocks: p_in gives a mean
    in-block degree deg_in, p_out a mean cross-block degree deg_out."""
    rng = np.random.default_rng(seed)
    bsz = np.full(n_blocks, n // n_blocks)
    bsz[: n % n_blocks] += 1
    start = np.concatenate([[0], np.cumsum(bsz)])
    edges = []
    for i in range(n_blocks):
        for j in range(i, n_blocks):
            if i == j:
                b = int(bsz[i])
                pairs = b * (b - 1) // 2
                p = min(1.0, deg_in / max(1, b - 1))
            else:
                pairs = int(bsz[i]) * int(bsz[j])
                p = min(1.0, deg_out / max(1, n - int(bsz[i])))
            cnt = rng.binomial(pairs, p) if pairs > 0 else 0
            if cnt == 0:
                continue
            u = rng.integers(0, bsz[i], cnt) + start[i]
            v = rng.integers(0, bsz[j], cnt) + start[j]
            keep = u != v
            edges.append(np.stack([u[keep], v[keep]], 1))
    if not edges:
        return n, np.zeros((0, 2), np.int32)
    e = np.concatenate(edges)
    lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
    key = lo.astype(np.int64) * n + hi
    _, idx = np.unique(key, return_index=True)
    idx.sort()
    return n, e[idx].astype(np.int32)
