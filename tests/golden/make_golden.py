"""Writes the golden fixtures of the self-join (tests/golden/*.npz).

Each fixture holds the input points (float64), k, eps and the canonical
neighbour table (CSR in original point IDs, rows ascending) computed by the
CPU oracle; the script refuses to write a fixture unless brute force with
the same predicate gives the identical table.  The reference ships no
vectors of its own, so these are generated here.

Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "self-join-hpbdc2019_amd"))

import oracle as O  # noqa: E402
from sjamd import datasets as D  # noqa: E402


def lattice(side, k):
    xs = np.arange(side, dtype=np.float64)
    return np.array(np.meshgrid(*([xs] * k), indexing="ij")).reshape(k, -1).T.copy()


def boundary_pairs():
    """Pairs at distance exactly eps, one ulp outside, and pairs whose FMA
    and plain accumulations fall on different sides of the threshold."""
    eps = 0.5
    T = O.eps_threshold(eps)
    pts = [[0.0, 0.0], [0.5, 0.0], [0.0, 0.5], [0.3, 0.4],   # d = 0.5 exactly
           [10.0, 10.0], [10.5 + 2 ** -49, 10.0],              # just outside
           [20.0, 20.0], [20.3, 20.4]]                         # 0.3^2+0.4^2 rounding
    rng = np.random.default_rng(11)
    flips = 0
    base = np.array([40.0, 40.0])
    while flips < 24:
        ang = rng.uniform(0, 2 * math.pi)
        r = eps * (1 + rng.uniform(-4e-16, 4e-16))
        p = base + rng.uniform(0, 50, 2)
        q = p + r * np.array([math.cos(ang), math.sin(ang)])
        d_fma = O.dist2(p, q)
        df = q - p
        d_plain = df[0] * df[0] + df[1] * df[1]
        if (d_fma <= T) != (d_plain <= T):
            pts.append(list(p))
            pts.append(list(q))
            base = base + 100.0
            flips += 1
    return np.array(pts, dtype=np.float64), eps


def fixtures():
    out = {}
    out["lattice2d_eps1p5"] = (lattice(30, 2), 2, 1.5)
    out["lattice3d_eps_sqrt2"] = (lattice(10, 3), 3, math.sqrt(2.0))
    out["lattice3d_eps_below_sqrt2"] = (lattice(10, 3), 3, float(np.nextafter(math.sqrt(2.0), 0)))
    bp, beps = boundary_pairs()
    out["boundary_fma_pairs"] = (bp, 2, beps)
    out["uniform2d_5k"] = (D.uniform(5000, 2, 101, 0.0, 50.0), 2, 0.9)
    out["exp2d_5k"] = (D.exponential_clipped(5000, 2, 102), 2, 3e-3)
    out["uniform3d_4k"] = (D.uniform(4000, 3, 103), 3, 0.08)
    out["uniform6d_3k"] = (D.uniform(3000, 6, 104), 6, 0.35)
    out["clusters64d_k6_1500"] = (D.clusters(1500, 64, 4), 6, 4.5)
    out["identical_points"] = (np.full((40, 2), 3.25), 2, 0.125)
    # coordinates on exact multiples of eps from the grid origin
    g = np.arange(0, 6, dtype=np.float64) * 0.25
    out["cell_edges_2d"] = (np.array(np.meshgrid(g, g, indexing="ij")).reshape(2, -1).T.copy(), 2, 0.25)
    return out


def main():
    for name, spec in fixtures().items():
        pts, k, eps = spec
        off, ids = O.join(pts, k, eps, O.LID, sort_rows=True)
        boff, bids = O.brute(pts, eps)
        if not (np.array_equal(off, boff) and np.array_equal(ids, bids)):
            raise SystemExit(f"{name}: grid oracle and brute force disagree; fixture not written")
        for form in (O.FULL, O.UNICOMP, O.BUNICOMP):
            f_off, f_ids = O.join(pts, k, eps, form, sort_rows=True)
            assert np.array_equal(f_off, off) and np.array_equal(f_ids, ids), (name, form)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, points=pts, k=np.int64(k), eps=np.float64(eps),
                            offsets=off, ids=ids)
        print(f"{name}: n={pts.shape[0]} d={pts.shape[1]} k={k} eps={float(eps).hex()} |S|={int(off[-1])}")


if __name__ == "__main__":
    main()
