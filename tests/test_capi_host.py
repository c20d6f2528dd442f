"""C ABI checks that need no GPU: the library loads, exports every entry
point include/sjamd.h declares, and its host-only functions (dataset
import, bin pre-sort, eps threshold) agree with the oracle."""
import os
import re
import tempfile

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sjamd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"SJ_API\s+[\w\s\*]+?\b(sj_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "sj_join" in names and "sj_grid_build" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    import sjamd

    L = sjamd.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(sjamd.EXPORTS) == declared()


def test_eps_threshold_matches_oracle():
    import sjamd

    rng = np.random.default_rng(1)
    for eps in [0.5, 1.0, 2.0, 1.5e-4, 4.5] + list(rng.uniform(1e-5, 1e3, 100)):
        assert sjamd.eps_threshold(eps) == O.eps_threshold(eps)


def test_import_csv_matches_oracle():
    import sjamd

    rng = np.random.default_rng(2)
    vals = rng.uniform(-100, 100, (101, 3))
    lines = [",".join(repr(float(v)) for v in row) for row in vals]
    text = "\n".join(lines) + "\n1.5, 2e3 ,bad,7\n"
    with tempfile.NamedTemporaryFile("w", suffix=".csv", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        a = sjamd.import_csv(path, 3)
        b = O.import_csv(path, 3)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a[:101], vals)
    finally:
        os.unlink(path)


def test_import_csv_missing_file_is_io_error():
    import sjamd

    with pytest.raises(sjamd.SJError) as e:
        sjamd.import_csv("/nonexistent/points.csv", 2)
    assert e.value.status == 6


def test_sort_in_nd_bins_stable_matches_oracle():
    import sjamd

    rng = np.random.default_rng(3)
    pts = rng.uniform(-20, 20, (700, 3))
    np.testing.assert_array_equal(sjamd.sort_in_nd_bins(pts, stable=True),
                                  O.sort_in_nd_bins_stable(pts))


def test_sort_in_nd_bins_reference_order_is_a_valid_order():
    import sjamd

    rng = np.random.default_rng(4)
    pts = rng.uniform(-20, 20, (700, 2))
    out = sjamd.sort_in_nd_bins(pts, stable=False)
    key = np.trunc(out).astype(np.int64)
    assert all(tuple(key[i]) <= tuple(key[i + 1]) for i in range(len(key) - 1))
    np.testing.assert_array_equal(np.sort(out.view("u8"), axis=None), np.sort(pts.view("u8"), axis=None))
    # distinct keys: identical to the stable order
    pts2 = np.arange(300, dtype=np.float64)[::-1].reshape(150, 2) * 3.0
    np.testing.assert_array_equal(sjamd.sort_in_nd_bins(pts2), sjamd.sort_in_nd_bins(pts2, stable=True))


def test_sort_in_nd_bins_rejects_out_of_int_range():
    import sjamd

    with pytest.raises(sjamd.SJError) as e:
        sjamd.sort_in_nd_bins(np.array([[1e12, 0.0]]))
    assert e.value.status == 4
