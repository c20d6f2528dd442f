"""Pins of the CPU oracle against facts it does not compute itself:
exact rational arithmetic, closed-form lattices, hand-built grids, numpy
restatements and brute force.  A plausible slip in the oracle (no FMA,
wrong eps threshold, transposed ID order, wrong stencil claim) fails one of
these."""
import math
import os
import tempfile
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from sjamd import datasets as D


# ------------------------------------------------------------- predicate
def _fma_exact(a, b, c):
    # correctly rounded a*b+c: exact rational, then one rounding
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _dist2_exact_fma(p, q):
    acc = 0.0
    for a, b in zip(p, q):
        df = float(Fraction(b) - Fraction(a))  # correctly rounded subtraction
        acc = _fma_exact(df, df, acc)
    return acc


def test_eps_threshold_is_largest_sqrt_preimage():
    rng = np.random.default_rng(0)
    eps_list = [0.5, 1.0, 1.5, 2.0, math.sqrt(2.0), 1.596e-4, 0.1207, 4.5, 1e-300, 1e150]
    eps_list += list(rng.uniform(1e-6, 1e3, 200))
    for eps in eps_list:
        t = O.eps_threshold(eps)
        assert math.sqrt(t) <= eps
        assert math.sqrt(np.nextafter(t, np.inf)) > eps


def test_eps_threshold_known_values():
    # sqrt(1 + 2^-52) = 1 + 2^-53 - ... rounds to 1, sqrt(1 + 2^-51) rounds
    # to 1 + 2^-52: the threshold of eps=1 is one ulp above 1
    assert O.eps_threshold(1.0) == 1.0 + 2.0 ** -52
    assert O.eps_threshold(2.0) == 4.0 + 2.0 ** -50
    # d2 == 2 is a hit for eps = fl(sqrt 2) and a miss one ulp below
    assert 2.0 <= O.eps_threshold(math.sqrt(2.0))
    assert 2.0 > O.eps_threshold(np.nextafter(math.sqrt(2.0), 0))


@pytest.mark.parametrize("d", [1, 2, 3, 6, 64])
def test_dist2_equals_exact_fma_chain(d):
    rng = np.random.default_rng(d)
    for _ in range(300):
        p = rng.normal(size=d) * rng.choice([1e-3, 1.0, 1e3])
        q = p + rng.normal(size=d) * rng.choice([1e-5, 1e-2, 1.0])
        assert O.dist2(p, q) == _dist2_exact_fma(p, q)
        assert O.dist2(p, q) == O.dist2(q, p)


def test_dist2_uses_fma_not_plain_sum():
    # search pairs where fma accumulation and round-after-multiply differ
    rng = np.random.default_rng(5)
    found = 0
    for _ in range(20000):
        p = rng.uniform(0, 1, 3)
        q = p + rng.uniform(-1e-4, 1e-4, 3)
        plain = 0.0
        for a, b in zip(p, q):
            df = b - a
            plain = plain + df * df
        ex = _dist2_exact_fma(p, q)
        if plain != ex:
            assert O.dist2(p, q) == ex
            found += 1
    assert found > 0


# ------------------------------------------------------------------ grid
def test_hand_built_grid():
    pts = np.array([[0, 0], [0.5, 0.5], [2, 0], [2.5, 1.5], [0, 3], [1.9, 2.9]], dtype=np.float64)
    g = O.OracleGrid(pts, 2, 1.0)
    np.testing.assert_array_equal(g.minArr, [-1.0, -1.0])
    np.testing.assert_array_equal(g.maxArr, [3.5, 4.0])
    np.testing.assert_array_equal(g.nCells, [5, 5])
    assert g.totalCells == 25
    np.testing.assert_array_equal(g.cellLinear, [6, 8, 13, 17, 21])
    np.testing.assert_array_equal(g.indexLookup, [0, 1, 2, 3, 5, 4])
    np.testing.assert_array_equal(g.indexMin, [0, 2, 3, 4, 5])
    np.testing.assert_array_equal(g.indexMax, [1, 2, 3, 4, 5])
    np.testing.assert_array_equal(g.ndMask, [1, 2, 3, 1, 2, 3, 4])
    np.testing.assert_array_equal(g.ndMaskOffsets, [0, 2, 3, 6])
    off, ids = g.join(O.LID)
    np.testing.assert_array_equal(off, [0, 2, 4, 5, 6, 7, 8])
    np.testing.assert_array_equal(ids, [0, 1, 0, 1, 2, 3, 4, 5])


def _numpy_grid(pts, k, eps):
    mn = pts[:, :k].min(0) - eps
    mx = pts[:, :k].max(0) + eps
    nc = np.ceil((mx - mn) / eps).astype(np.uint64)
    coords = ((pts[:, :k] - mn) / eps).astype(np.uint64)
    mult = np.concatenate([[1], np.cumprod(nc)[:-1]]).astype(np.uint64)
    lin = (coords * mult).sum(1).astype(np.uint64)
    order = np.argsort(lin, kind="stable")
    uniq, first, counts = np.unique(lin, return_index=False, return_counts=True, return_inverse=False), None, None
    uniq, counts = np.unique(lin, return_counts=True)
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    masks = [np.unique(coords[:, j]) for j in range(k)]
    return mn, mx, nc, int(np.prod(nc)), uniq, order, start, start + counts - 1, masks


@pytest.mark.parametrize("name,k,eps", [
    ("u2", 2, 0.5), ("e2", 2, 2e-3), ("u6", 6, 0.3), ("c64", 6, 4.5), ("u3k2", 2, 0.05)])
def test_grid_matches_numpy_restatement(name, k, eps):
    pts = {
        "u2": lambda: D.uniform(20_000, 2, 1, 0, 100),
        "e2": lambda: D.exponential_clipped(20_000, 2, 2),
        "u6": lambda: D.uniform(20_000, 6, 3),
        "c64": lambda: D.clusters(3_000, 64, 4),
        "u3k2": lambda: D.uniform(20_000, 3, 9),
    }[name]()
    g = O.OracleGrid(pts, k, eps)
    mn, mx, nc, tot, uniq, order, imin, imax, masks = _numpy_grid(pts, k, eps)
    np.testing.assert_array_equal(g.minArr, mn)
    np.testing.assert_array_equal(g.maxArr, mx)
    np.testing.assert_array_equal(g.nCells, nc)
    assert g.totalCells == tot
    np.testing.assert_array_equal(g.cellLinear, uniq)
    np.testing.assert_array_equal(g.indexLookup, order)
    np.testing.assert_array_equal(g.indexMin, imin)
    np.testing.assert_array_equal(g.indexMax, imax)
    np.testing.assert_array_equal(g.ndMask, np.concatenate(masks))


def test_linear_id_dim0_fastest():
    nc = np.array([7, 5, 3], dtype=np.uint32)
    lib = O.lib()
    import ctypes as C
    idx = np.array([2, 3, 1], dtype=np.uint32)
    v = lib.sjo_linear_id(idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                          nc.ctypes.data_as(C.POINTER(C.c_uint32)), 3)
    assert v == 2 + 3 * 7 + 1 * 35


# --------------------------------------------------------------- stencils
def _interior_cells(k, side):
    rng = np.arange(2, side - 2)
    return np.array(np.meshgrid(*([rng] * k), indexing="ij")).reshape(k, -1).T


def _offsets(k):
    o = np.array(np.meshgrid(*([[-1, 0, 1]] * k), indexing="ij")).reshape(k, -1).T
    return o[np.any(o != 0, axis=1)]


@pytest.mark.parametrize("form,k", [
    (O.UNICOMP, 2), (O.UNICOMP, 3), (O.UNICOMP, 4), (O.BUNICOMP, 2), (O.BUNICOMP, 3),
    (O.BUNICOMP, 4), (O.BUNICOMP, 5), (O.LID, 2), (O.LID, 3), (O.LID, 4)])
def test_claims_exactly_once(form, k):
    side = 6 if k <= 3 else 5
    nc = np.full(k, side, dtype=np.uint32)
    offs = _offsets(k)
    rng = np.random.default_rng(k)
    cells = _interior_cells(k, side)
    if len(cells) > 40:
        cells = cells[rng.choice(len(cells), 40, replace=False)]
    per_cell = []
    for c in cells:
        n_claim = 0
        for o in offs:
            a = O.claims(form, k, c, o, nc)
            b = O.claims(form, k, c + o, -o, nc)
            assert a != b, f"pair {c}->{c + o} claimed {int(a) + int(b)} times"
            n_claim += a
        per_cell.append(n_claim)
    if form in (O.BUNICOMP, O.LID):
        assert all(x == (3 ** k - 1) // 2 for x in per_cell)
    else:
        assert all(0 <= x <= 3 ** k - 1 for x in per_cell)


def test_unicomp_visit_counts_2d_and_3d():
    # exhaustive enumeration: a Unicomp cell visits between 0 and 3^k-1
    # of its neighbour cells
    for k in (2, 3):
        nc = np.full(k, 8, dtype=np.uint32)
        counts = set()
        for c in _interior_cells(k, 8):
            counts.add(sum(O.claims(O.UNICOMP, k, c, o, nc) for o in _offsets(k)))
        assert min(counts) == 0 and max(counts) == 3 ** k - 1


# ------------------------------------------------------------------ joins
@pytest.mark.parametrize("k,eps,per_interior", [(2, 1.0, 5), (2, 1.5, 9), (2, 2.0, 13),
                                                (3, 1.0, 7), (3, math.sqrt(2.0), 19)])
def test_lattice_closed_form(k, eps, per_interior):
    side = 12
    xs = np.arange(side, dtype=np.float64)
    pts = np.array(np.meshgrid(*([xs] * k), indexing="ij")).reshape(k, -1).T.copy()
    for form in (O.FULL, O.UNICOMP, O.BUNICOMP, O.LID):
        off, ids = O.join(pts, k, eps, form)
        cnt = np.diff(off.astype(np.int64)).reshape([side] * k)
        inner = cnt[tuple([slice(3, -3)] * k)]
        assert (inner == per_interior).all()


def test_lattice_eps_just_below_sqrt2():
    xs = np.arange(10, dtype=np.float64)
    pts = np.array(np.meshgrid(xs, xs, xs, indexing="ij")).reshape(3, -1).T.copy()
    off, _ = O.join(pts, 3, np.nextafter(math.sqrt(2.0), 0), O.LID)
    cnt = np.diff(off.astype(np.int64)).reshape(10, 10, 10)
    assert (cnt[3:-3, 3:-3, 3:-3] == 7).all()


@pytest.mark.parametrize("d,k,n,eps,kind", [
    (2, 2, 4000, 1.5, "u100"), (2, 2, 4000, 0.01, "exp"), (3, 3, 3000, 0.08, "u1"),
    (6, 6, 2000, 0.35, "u1"), (8, 4, 2000, 0.5, "u1"), (64, 6, 1500, 4.5, "clusters")])
def test_grid_join_equals_brute_force(d, k, n, eps, kind):
    if kind == "u100":
        pts = D.uniform(n, d, 21, 0, 100)
    elif kind == "exp":
        pts = D.exponential_clipped(n, d, 22)
    elif kind == "clusters":
        pts = D.clusters(n, d, 4)
    else:
        pts = D.uniform(n, d, 23)
    boff, bids = O.brute(pts, eps)
    for form in (O.FULL, O.UNICOMP, O.BUNICOMP, O.LID):
        off, ids = O.join(pts, k, eps, form)
        np.testing.assert_array_equal(off, boff)
        np.testing.assert_array_equal(ids, bids)


def test_forms_agree_and_are_symmetric():
    pts = D.exponential_clipped(30_000, 2, 2)
    g = O.OracleGrid(pts, 2, 1.2e-3)
    ref = g.join(O.FULL)
    for form in (O.UNICOMP, O.BUNICOMP, O.LID):
        off, ids = g.join(form)
        np.testing.assert_array_equal(off, ref[0])
        np.testing.assert_array_equal(ids, ref[1])
    off, ids = ref
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))
    fwd = set(zip(rows.tolist(), ids.tolist()))
    assert all((b, a) in fwd for a, b in list(fwd)[:200000])
    assert all((i, i) in fwd for i in range(len(off) - 1))


# ---------------------------------------------------------- workload etc.
def test_workload_matches_numpy():
    pts = D.uniform(5_000, 2, 31, 0, 10)
    g = O.OracleGrid(pts, 2, 0.25)
    coords = g.cellCoord.astype(np.int64)
    cells = {tuple(c.tolist()): i for i, c in enumerate(coords)}
    size = g.indexMax - g.indexMin + 1
    w_full, w_lid = g.workload(False), g.workload(True)
    for i, c in enumerate(coords.tolist()):
        sf = sl = 0
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                j = cells.get((c[0] + dx, c[1] + dy))
                if j is None:
                    continue
                sf += size[j]
                if g.cellLinear[j] >= g.cellLinear[i]:
                    sl += size[j]
        assert w_full[i] == sf and w_lid[i] == sl
    order = g.workload_order(False)
    ws = w_full[order]
    assert (np.diff(ws) <= 0).all()
    ties = np.diff(ws) == 0
    assert (np.diff(order.astype(np.int64))[ties] > 0).all()


def test_batch_policy_known_answers():
    # 1% sample of |S|~0.88M (100k uniform, eps 0.5) -> 3 batches
    p = O.batch_policy(8800)
    assert p["estimated"] == 880000
    assert p["num_batches"] == 3
    assert p["buffer_size"] == int(1.5 * int(880000 * 1.1 / 3))
    # |S|~128M (2M exponential, 64 per point) -> 40M buffers, 4 batches
    p = O.batch_policy(1_280_000)
    assert p["num_batches"] == 4
    assert p["buffer_size"] == 60_000_000
    # empty result stays well-defined
    p = O.batch_policy(0)
    assert p["num_batches"] == 1


def test_sample_count_matches_brute():
    pts = D.uniform(3_000, 2, 41, 0, 10)
    g = O.OracleGrid(pts, 2, 0.4)
    off, _ = O.brute(pts, 0.4)
    cnt = np.diff(off.astype(np.int64))
    assert g.sample_count(100) == int(cnt[::100][:30].sum())


@pytest.mark.parametrize("lid", [False, True])
def test_sample_count_workqueue_matches_brute(lid):
    # the sample is the first n/stride points of the workload-ordered list;
    # the order comes from a numpy count of each cell's 3x3 stencil and a
    # stable sort, the hits from brute force over all points
    pts = D.exponential_clipped(4_000, 2, 43, 40.0)
    eps = 0.01
    g = O.OracleGrid(pts, 2, eps)
    coords = g.cellCoord.astype(np.int64)
    cells = {tuple(c.tolist()): i for i, c in enumerate(coords)}
    size = (g.indexMax - g.indexMin + 1).astype(np.int64)
    w = np.zeros(g.nNonEmpty, dtype=np.int64)
    for i, c in enumerate(coords.tolist()):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                j = cells.get((c[0] + dx, c[1] + dy))
                if j is not None and (not lid or g.cellLinear[j] >= g.cellLinear[i]):
                    w[i] += size[j]
    order = np.argsort(-w, kind="stable")
    plist = np.concatenate([g.indexLookup[g.indexMin[c]:g.indexMax[c] + 1] for c in order])
    off, _ = O.brute(pts, eps)
    cnt = np.diff(off.astype(np.int64))
    for stride in (100, 37):
        nsamp = int(pts.shape[0] * (1.0 / stride))
        assert g.sample_count_workqueue(stride, lid) == int(cnt[plist[:nsamp]].sum())


def test_neighbor_table_run_length():
    keys = np.array([0, 0, 2, 2, 2, 3], dtype=np.int32)
    vals = np.array([0, 1, 2, 3, 4, 3], dtype=np.int32)
    tab = O.neighbor_table(keys, vals, 5, batch=7)
    np.testing.assert_array_equal(tab[0], [0, 0, 1, 7])
    np.testing.assert_array_equal(tab[2], [2, 2, 4, 7])
    np.testing.assert_array_equal(tab[3], [3, 5, 5, 7])
    assert tab[1][3] == -1 and tab[4][3] == -1


# --------------------------------------------------------------- import
def test_import_csv_semantics():
    text = "1.5,2\n3e1, 4.25\n-5 ,6\n7,8x,9\n10,11,12"
    with tempfile.NamedTemporaryFile("w", suffix=".csv", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        pts = O.import_csv(path, 2)
        # fields: "1.5" "2\n3e1" " 4.25\n-5 " "6\n7" "8x" "9\n10" "11" "12"
        want = [1.5, 2.0, 30.0, 4.25, -5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 11.0, 12.0]
        np.testing.assert_array_equal(pts.ravel(), want)
    finally:
        os.unlink(path)


def test_sort_in_nd_bins_stable():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-5, 5, (500, 3))
    out = O.sort_in_nd_bins_stable(pts)
    key = np.trunc(pts).astype(np.int64)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    np.testing.assert_array_equal(out, pts[order])


# ---------------------------------------------------------------- datasets
def test_splitmix64_known_answer():
    z = D.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_datasets_deterministic():
    a = D.make_points("uniform2d", n=1000)
    b = D.make_points("uniform2d", n=1000)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= 0 and a.max() < 100
    e = D.make_points("exp2d", n=1000)
    assert e.min() >= 0 and e.max() <= 1.0
