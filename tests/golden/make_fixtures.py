"""Writes the golden fixtures of this directory with the oracle (the C
restatement of the reference in oracle/). Run from the repository root after
`make -C oracle`:

    python tests/golden/make_fixtures.py

The reference ships no golden vectors: its tests draw inputs from R's RNG and
igraph at run time (tests/testthat/helper-reference.R:102-110). These
fixtures hold inputs and outputs of the configurations the reference's tests
and benchmarks use, from this repository's seeded generators, so that later
changes to the oracle or to the GPU path are checked against stored data:

  c1_islands.npz   islands 5 x 30 (within-island G(30, 0.3), 2 edges between
                   every pair of islands), start U[0, 100)^2, exact FP64 CPU
                   layout (C_forceatlas2, scaling 2, gravity 1): positions
                   after 1, 10 and 1000 iterations.
  dupkeys_700.npz  700 float32 positions, the first 60 on 3 points: Morton
                   keys, the sorted permutation and the Karras tree of the
                   large path (pipeline.cl:105-270), in the reference
                   numbering.
  c2_exact_4096.npz  50,000 nodes (preferential attachment, m = 3) at
                   float32 positions U[-10 sqrt(n), 10 sqrt(n))^2 with their
                   masses, and the exact FP64 repulsion (scaling 2) of 4,096
                   sampled nodes from all others.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from forceatlas2_rc_amd import graphs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def c1():
    n, e = graphs.islands(5, 30, 0.3, 2, seed=11)
    pos = graphs.uniform_positions(n, 0.0, 100.0, 12)
    # one run of 1000 iterations (a new call would restart the speed)
    _, tp, _, _ = O.layout(n, e, pos, 1000, scaling=2.0, gravity=1.0, trace=True)
    np.savez_compressed(os.path.join(HERE, "c1_islands.npz"), edges=e, init=pos, pos1=tp[0],
                        pos10=tp[9], pos1000=tp[999])


def dupkeys():
    rng = np.random.default_rng(21)
    n = 700
    pos = rng.uniform(-300.0, 300.0, size=(n, 2)).astype(np.float32)
    pos[:60] = pos[np.arange(60) % 3]
    t = O.large_tree(pos[:, 0], pos[:, 1], "f")
    np.savez_compressed(os.path.join(HERE, "dupkeys_700.npz"), pos=pos, keys=t["keys"],
                        order=t["order"], child=t["child"], range=t["range"],
                        parent=t["parent"])


def c2():
    n, e = graphs.barabasi_albert(50_000, 3, seed=31)
    mass = 1.0 + np.bincount(e.ravel(), minlength=n).astype(np.float64)
    r = 10.0 * np.sqrt(n)
    pos = graphs.uniform_positions(n, -r, r, 32).astype(np.float32)
    x, y = pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64)
    f = O.repulse_exact(x, y, mass, 2.0, threads=os.cpu_count() or 1)
    idx = np.sort(np.random.default_rng(33).choice(n, 4096, replace=False)).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "c2_exact_4096.npz"), pos=pos,
                        mass=mass.astype(np.uint16), idx=idx, force=f[idx])


if __name__ == "__main__":
    c1()
    dupkeys()
    c2()
