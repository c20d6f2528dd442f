"""GPU parity of the HIP grid build and join against the CPU oracle.

Every comparison is exact: grid arrays bit for bit, neighbour tables as
sets per row (rows sorted on both sides)."""
import sys

import numpy as np
import pytest

import oracle as O
from sjamd import datasets as D

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _canon(off, ids):
    return O.csr_rows_sorted(np.asarray(off, dtype=np.uint64), np.asarray(ids, dtype=np.int32))


def _assert_same(gpu, ref):
    goff, gids = _canon(*gpu)
    roff, rids = ref
    assert goff.shape == roff.shape
    assert int(goff[-1]) == int(roff[-1]), f"|S| gpu {int(goff[-1])} oracle {int(roff[-1])}"
    np.testing.assert_array_equal(goff, roff)
    np.testing.assert_array_equal(gids, rids)


CASES = {
    "uniform2d_100k": lambda: (D.make_points("uniform2d"), 2, 0.5),
    "exp2d_200k": lambda: (D.exponential_clipped(200_000, 2, 2), 2, 5.0e-4),
    "uniform6d_50k": lambda: (D.uniform(50_000, 6, 3), 6, 0.23),
    "uniform3d_60k": lambda: (D.uniform(60_000, 3, 11), 3, 0.03),
    "uniform4d_40k": lambda: (D.uniform(40_000, 4, 12), 4, 0.07),
    "clusters64d_k6_8k": lambda: (D.clusters(8_000, 64, 4), 6, 4.5),
    "uniform8d_k5_20k": lambda: (D.uniform(20_000, 8, 13), 5, 0.35),
    # d = k = 5 and 7: the tile walks built for k = 5..7 (four-wave FP64 count)
    "uniform5d_20k": lambda: (D.uniform(20_000, 5, 14), 5, 0.18),
    "uniform7d_20k": lambda: (D.uniform(20_000, 7, 15), 7, 0.30),
}

_cache = {}


def _case(name):
    if name not in _cache:
        pts, k, eps = CASES[name]()
        g = O.OracleGrid(pts, k, eps)
        _cache[name] = (pts, k, eps, g, g.join(O.LID, sort_rows=True))
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_grid_arrays_match_oracle(ctx, name):
    pts, k, eps, og, _ = _case(name)
    _say(f"[grid] {name} n={pts.shape[0]} d={pts.shape[1]} k={k}")
    g = ctx.grid_build(pts, k, eps)
    mn, mx, nc, tot = g.dims()
    np.testing.assert_array_equal(mn, og.minArr)
    np.testing.assert_array_equal(mx, og.maxArr)
    np.testing.assert_array_equal(nc, og.nCells)
    assert tot == og.totalCells
    assert g.nNonEmpty == og.nNonEmpty
    a = g.arrays()
    np.testing.assert_array_equal(a["cellLinear"], og.cellLinear)
    np.testing.assert_array_equal(a["indexMin"], og.indexMin)
    np.testing.assert_array_equal(a["indexMax"], og.indexMax)
    np.testing.assert_array_equal(a["indexLookup"], og.indexLookup)
    np.testing.assert_array_equal(a["cellCoord"], og.cellCoord)
    mask, offs = g.nd_mask()
    np.testing.assert_array_equal(mask, og.ndMask)
    np.testing.assert_array_equal(offs, og.ndMaskOffsets)
    g.free()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("form", ["full", "unicomp", "bunicomp", "lid"])
@pytest.mark.parametrize("walk", ["tile", "point"])
def test_join_fp64_matches_oracle(ctx, name, form, walk):
    pts, k, eps, og, ref = _case(name)
    _say(f"[join fp64] {name} form={form} walk={walk}")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form=form, per_point=(walk == "point"))
    _assert_same(r.csr(), ref)
    st = r.stats()
    assert st["total_pairs"] == int(ref[0][-1])
    r.free()
    g.free()


@pytest.mark.parametrize("name", ["uniform2d_100k", "exp2d_200k", "uniform6d_50k", "clusters64d_k6_8k",
                                  "uniform5d_20k", "uniform7d_20k"])
@pytest.mark.parametrize("form", ["bunicomp", "lid"])
def test_join_fp32_recheck_exact(ctx, name, form):
    import sjamd

    pts, k, eps, og, ref = _case(name)
    _say(f"[join fp32] {name} form={form}")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form=form, precision=sjamd.FP32_RECHECK)
    _assert_same(r.csr(), ref)
    st = r.stats()
    assert st["fp32_misclassified"] <= st["fp32_band_pairs"]
    r.free()
    g.free()


@pytest.mark.parametrize("name", ["uniform2d_100k", "exp2d_200k", "uniform6d_50k"])
@pytest.mark.parametrize("form", ["full", "lid", "bunicomp"])
def test_join_workqueue(ctx, name, form):
    import sjamd

    pts, k, eps, og, ref = _case(name)
    _say(f"[join queue] {name} form={form}")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form=form, schedule=sjamd.WORKQUEUE)
    _assert_same(r.csr(), ref)
    r.free()
    g.free()


def test_device_views(ctx):
    """cell-order table + row IDs, and the original-order CSR built from it"""
    from sjamd import download

    pts, k, eps, og, ref = _case("uniform2d_100k")
    _say("[device views] uniform2d_100k")
    r = ctx.self_join(pts, k, eps, form="lid")
    n, tot = pts.shape[0], r.total
    o, i, rid = r.device_table()
    off = download(o, n + 1, np.uint64)
    ids = download(i, tot, np.int32)
    rows = download(rid, n, np.uint32).astype(np.int64)
    np.testing.assert_array_equal(np.sort(rows), np.arange(n))
    for rr in [0, 5, n // 2, n - 1]:
        p = rows[rr]
        got = np.sort(ids[int(off[rr]):int(off[rr + 1])])
        np.testing.assert_array_equal(got, ref[1][int(ref[0][p]):int(ref[0][p + 1])])
    o2, i2 = r.device_csr()
    _assert_same((download(o2, n + 1, np.uint64), download(i2, tot, np.int32)), ref)


def test_sorted_rows_device(ctx):
    pts, k, eps, og, ref = _case("exp2d_200k")
    _say("[sorted rows] exp2d_200k")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form="bunicomp", sort_rows=True)
    off, ids = r.csr()
    np.testing.assert_array_equal(off, ref[0])
    np.testing.assert_array_equal(ids, ref[1])


@pytest.mark.parametrize("form", ["full", "unicomp", "lid"])
def test_batched_rounds(ctx, form):
    pts, k, eps, og, ref = _case("uniform6d_50k")
    total = int(ref[0][-1])
    cap = total // 5 + 1
    _say(f"[batched] uniform6d_50k form={form} capacity={cap}")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form=form, capacity=cap)
    assert r.stats()["num_batches"] >= 4
    _assert_same(r.csr(), ref)
    tab = r.neighbor_table()
    # every row through the reference-style lookup table
    for p in [0, 17, 4999, pts.shape[0] - 1]:
        pid, lo, hi, b = tab[p]
        data = r.batch_data(int(b))
        got = np.sort(data[lo:hi + 1])
        want = ref[1][int(ref[0][p]):int(ref[0][p + 1])]
        np.testing.assert_array_equal(got, want)


def test_grid_key_width_from_last_build(ctx):
    """A grid of the same shape and eps as the context's previous one is
    sorted with that grid's key width before its own parameters reach the
    host; a grid that needs more bits (or 64-bit keys) is sorted again.
    Each grid and join below must equal the oracle's whatever the guess."""
    rng = np.random.default_rng(21)
    base = rng.random((3000, 3))
    eps = 0.05
    # ~14-bit keys, then 64-bit keys (totalCells > 2^32), then ~21 bits
    # (a wider guess than needed), then the first set again
    for scale in (1.0, 100.0, 6.0, 1.0):
        pts = np.ascontiguousarray(base * scale)
        og = O.OracleGrid(pts, 3, eps)
        _say(f"[grid guess] scale={scale} totalCells={og.totalCells}")
        g = ctx.grid_build(pts, 3, eps)
        mn, mx, nc, tot = g.dims()
        assert tot == og.totalCells
        np.testing.assert_array_equal(nc, og.nCells)
        a = g.arrays()
        np.testing.assert_array_equal(a["cellLinear"], og.cellLinear)
        np.testing.assert_array_equal(a["indexLookup"], og.indexLookup)
        np.testing.assert_array_equal(a["indexMin"], og.indexMin)
        r = g.join(form="lid")
        _assert_same(r.csr(), og.join(O.LID, sort_rows=True))
        r.free()
        g.free()


@pytest.mark.parametrize("case", ["clusters2d_sparse", "clusters6d_sparse"])
def test_sparse_grids_general_kernels(ctx, case):
    """Grids too sparse for the dense lookup table (totalCells above 2^26
    and above 16 n): the 2-D and 6-D joins then run the kernels with the
    binary-search lookups rather than the dense-grid variants, and must
    equal the oracle in every form and both precisions."""
    import sjamd

    if case == "clusters2d_sparse":
        pts, k, eps = D.clusters(20_000, 2, 16, n_clusters=20, box=1e5, sigma=2.0), 2, 1.0
    else:
        pts, k, eps = D.clusters(20_000, 6, 17, n_clusters=20, box=100.0, sigma=0.3), 6, 0.5
    ref = O.join(pts, k, eps, O.LID, sort_rows=True)
    g = ctx.grid_build(pts, k, eps)
    n = pts.shape[0]
    tc = g.dims()[3]
    assert tc > (1 << 26) and tc > 16 * n and not g.info.dense_lookup
    for form in ("full", "unicomp", "bunicomp", "lid"):
        for prec in (sjamd.FP64, sjamd.FP32_RECHECK):
            _say(f"[sparse] {case} form={form} prec={prec}")
            r = g.join(form=form, precision=prec)
            _assert_same(r.csr(), ref)
            r.free()
    g.free()


def test_edge_single_point(ctx):
    _say("[edge] single point")
    r = ctx.self_join(np.array([[0.25, 0.75]]), 2, 0.1)
    off, ids = r.csr()
    assert list(off) == [0, 1] and list(ids) == [0]


def test_edge_identical_points(ctx):
    _say("[edge] identical points")
    pts = np.full((300, 3), 1.5)
    ref = O.join(pts, 3, 0.01, O.LID)
    for form in ["full", "unicomp", "bunicomp", "lid"]:
        r = ctx.self_join(pts, 3, 0.01, form=form)
        assert r.total == 300 * 300
        _assert_same(r.csr(), ref)


def test_edge_lattice_boundaries(ctx):
    _say("[edge] integer lattice, eps on the boundary")
    xs = np.arange(30, dtype=np.float64)
    pts = np.stack(np.meshgrid(xs, xs, indexing="ij"), -1).reshape(-1, 2)
    for eps, per_interior in [(1.0, 5), (1.5, 9), (2.0, 13)]:
        ref = O.join(pts, 2, eps, O.LID)
        r = ctx.self_join(pts, 2, eps, form="bunicomp")
        _assert_same(r.csr(), ref)
        off = ref[0].astype(np.int64)
        cnt = np.diff(off).reshape(30, 30)
        assert (cnt[3:-3, 3:-3] == per_interior).all()
