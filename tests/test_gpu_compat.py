"""The reference-named C++ interface (libsjamd_compat.so) driven like the
reference program: import, bin pre-sort, grid, batched join, neighbour
table, for every search mode; compared with the oracle."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "self-join-hpbdc2019_amd", "sjamd", "sjamd_selfjoin")


def lattice_points(side, seed):
    # distinct integer parts: the bin pre-sort has no ties, so its order is
    # the same for std::sort and a stable sort
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    pts = np.stack([i.ravel(), j.ravel()], 1).astype(np.float64) + rng.uniform(0, 0.5, (side * side, 2))
    return pts[rng.permutation(len(pts))]


@pytest.mark.parametrize("mode", [3, 4, 5, 6, 7, 9, 10])
def test_reference_flow_all_modes(mode, tmp_path):
    pts = lattice_points(60, 5)
    csv = tmp_path / "pts.csv"
    csv.write_text("\n".join(f"{x!r},{y!r}" for x, y in pts.tolist()) + "\n")
    dump = tmp_path / "table.txt"
    print(f"[compat] mode={mode}", flush=True)
    sys.stdout.flush()
    out = subprocess.run([DRIVER, str(csv), "1.2", "2", str(mode), "2", str(dump)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    sorted_pts = O.sort_in_nd_bins_stable(pts)
    off, ids = O.join(sorted_pts, 2, 1.2, O.LID, sort_rows=True)
    lines = dump.read_text().splitlines()
    assert len(lines) == len(pts)
    for p, line in enumerate(lines):
        got = [int(v) for v in line.split(":")[1].split()]
        assert got == ids[int(off[p]):int(off[p + 1])].tolist(), f"point {p}"
    total = int([l for l in out.stdout.splitlines() if l.startswith("totalNeighbors:")][0].split()[1])
    assert total == int(off[-1])
