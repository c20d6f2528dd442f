"""The oracle reproduces every committed golden fixture (written from the
oracle and cross-checked by brute force when generated), so a later change
of the oracle's arithmetic or ID order shows up here."""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(GOLDEN) >= 10


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_reproduces_fixture(path):
    f = load(path)
    for form in (O.FULL, O.UNICOMP, O.BUNICOMP, O.LID):
        off, ids = O.join(f["points"], int(f["k"]), float(f["eps"]), form, sort_rows=True)
        np.testing.assert_array_equal(off, f["offsets"])
        np.testing.assert_array_equal(ids, f["ids"])


def test_boundary_fixture_decides_with_fma():
    f = load(os.path.join(os.path.dirname(__file__), "golden", "boundary_fma_pairs.npz"))
    pts, off, ids = f["points"], f["offsets"].astype(np.int64), f["ids"]
    T = O.eps_threshold(float(f["eps"]))
    # pairs (2i, 2i+1) from index 8 on were chosen where the fma and plain
    # accumulations straddle T; the table follows the fma one
    for a in range(8, pts.shape[0], 2):
        b = a + 1
        fma_hit = O.dist2(pts[a], pts[b]) <= T
        row = ids[off[a]:off[a + 1]]
        assert (b in row) == fma_hit
    # exact-eps pairs are hits, one ulp beyond is not
    assert 1 in ids[off[0]:off[1]] and 2 in ids[off[0]:off[1]] and 3 in ids[off[0]:off[1]]
    assert 5 not in ids[off[4]:off[5]]
