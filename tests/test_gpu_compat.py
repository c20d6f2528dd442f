"""The reference-named C++ interface (libsjamd_compat.so) driven like the
reference program (main.cu:36-226): import, bin pre-sort, grid, batched
join, neighbour table, the per-mode stats line and the neighbour-table
listing, for every search mode; compared with the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "self-join-hpbdc2019_amd", "sjamd", "sjamd_selfjoin")

STATS_FILE = {3: "gpu_global_stats.txt", 4: "gpu_unicomp_stats.txt", 5: "gpu_b-unicomp_stats.txt",
              6: "gpu_linID-unicomp_stats.txt", 7: "gpu_sorting-cells_stats.txt",
              8: "gpu_dynamic-threads-fixed_stats.txt", 9: "gpu_working-queue_stats.txt",
              10: "gpu_working-queue_lid-unicomp_stats.txt"}


def lattice_points(side, seed):
    # distinct integer parts: the bin pre-sort has no ties, so its order is
    # the same for std::sort and a stable sort
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    pts = np.stack([i.ravel(), j.ravel()], 1).astype(np.float64) + rng.uniform(0, 0.5, (side * side, 2))
    return pts[rng.permutation(len(pts))]


def reference_listing(off, ids):
    """The text main.cu:212-222 prints for a table whose rows are sorted:
    every row's IDs except the last, each followed by a comma."""
    out = []
    for p in range(off.size - 1):
        row = ids[int(off[p]):int(off[p + 1])]
        out.append(f"\npoint id: {p}, neighbors: nb neighbor {row.size}\n")
        out.append("".join(f"{v}," for v in row[:-1].tolist()))
    out.append("\n\n\n")
    return "".join(out)


@pytest.mark.parametrize("mode", [3, 4, 5, 6, 7, 8, 9, 10])
def test_reference_flow_all_modes(mode, tmp_path):
    pts = lattice_points(60, 5)
    csv = tmp_path / "pts.csv"
    csv.write_text("\n".join(f"{x!r},{y!r}" for x, y in pts.tolist()) + "\n")
    dump = tmp_path / "table.txt"
    print(f"[compat] mode={mode}", flush=True)
    sys.stdout.flush()
    out = subprocess.run([DRIVER, str(csv), "1.2", "2", str(mode), "2", str(dump)],
                         capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    sorted_pts = O.sort_in_nd_bins_stable(pts)
    off, ids = O.join(sorted_pts, 2, 1.2, O.LID, sort_rows=True)
    assert dump.read_text() == reference_listing(off, ids)
    total = int([l for l in out.stdout.splitlines() if l.startswith("totalNeighbors:")][0].split()[1])
    assert total == int(off[-1])
    # main.cu:203-206: time, dataset, epsilon, |S|, build parameters
    line = (tmp_path / STATS_FILE[mode]).read_text().splitlines()
    assert len(line) == 1
    f = line[0].split(", ")
    assert f[1] == str(csv) and f[2] == "1.2" and int(f[3]) == int(off[-1])
    assert f[4:] == ["GPUNUMDIM/NUMINDEXEDDIM/DTYPE/BLOCKSIZE/THperPOINT/SORT: 2", "2", "double", "256",
                     "1", "1"]
    float(f[0])
