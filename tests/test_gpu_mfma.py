"""Matrix-core FP32 path (16 <= d <= 64, FP32 + FP64 band recheck) against
the FP64 grid oracle: the pair sets must be identical, row by row."""
import sys

import numpy as np
import pytest

import oracle as O
from sjamd import datasets as D

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _assert_same(gpu, ref):
    goff, gids = O.csr_rows_sorted(np.asarray(gpu[0], dtype=np.uint64), np.asarray(gpu[1], dtype=np.int32))
    roff, rids = ref
    assert int(goff[-1]) == int(roff[-1]), f"|S| gpu {int(goff[-1])} oracle {int(roff[-1])}"
    np.testing.assert_array_equal(goff, roff)
    np.testing.assert_array_equal(gids, rids)


def _lattice16():
    # coordinates in {0,1}^16: the squared distance is the Hamming
    # distance, so many pairs sit exactly on eps^2 = 4 and land in the
    # recheck band
    rng = np.random.default_rng(7)
    return rng.integers(0, 2, size=(6000, 16)).astype(np.float64), 3, 2.0


CASES = {
    "clusters16d_k3": lambda: (D.clusters(12_000, 16, 21, n_clusters=40), 3, 2.1),
    "clusters24d_k2": lambda: (D.clusters(10_000, 24, 22, n_clusters=20), 2, 2.5),
    "clusters32d_k6": lambda: (D.clusters(10_000, 32, 23, n_clusters=15), 6, 3.0),
    "clusters64d_k6": lambda: (D.clusters(12_000, 64, 4, n_clusters=3), 6, 4.5),
    "uniform20d_k4": lambda: (D.uniform(8_000, 20, 24), 4, 1.0),
    "lattice16d_k3": _lattice16,
}

_cache = {}


def _case(name):
    if name not in _cache:
        pts, k, eps = CASES[name]()
        ref = O.join(pts, k, eps, O.LID)
        _cache[name] = (pts, k, eps, ref)
    return _cache[name]


def test_cases_have_neighbours():
    """the fixtures are not trivially self-pairs only"""
    for name in CASES:
        pts, k, eps, ref = _case(name)
        assert int(ref[0][-1]) > 2 * pts.shape[0], name


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("form", ["full", "unicomp", "bunicomp", "lid"])
def test_mfma_matches_fp64_oracle(ctx, name, form):
    import sjamd

    pts, k, eps, ref = _case(name)
    _say(f"[mfma] {name} n={pts.shape[0]} d={pts.shape[1]} k={k} form={form}")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form=form, precision=sjamd.FP32_RECHECK)
    _assert_same(r.csr(), ref)
    st = r.stats()
    assert st["fp32_misclassified"] <= st["fp32_band_pairs"]
    if name == "lattice16d_k3":
        assert st["fp32_band_pairs"] > 0
    r.free()
    g.free()


@pytest.mark.parametrize("name", ["clusters16d_k3", "clusters64d_k6"])
def test_mfma_workqueue_and_batches(ctx, name):
    import sjamd

    pts, k, eps, ref = _case(name)
    total = int(ref[0][-1])
    _say(f"[mfma queue+batches] {name}")
    g = ctx.grid_build(pts, k, eps)
    r = g.join(form="lid", precision=sjamd.FP32_RECHECK, schedule=sjamd.WORKQUEUE)
    _assert_same(r.csr(), ref)
    r.free()
    r = g.join(form="bunicomp", precision=sjamd.FP32_RECHECK, capacity=total // 4 + 1)
    assert r.stats()["num_batches"] >= 4
    _assert_same(r.csr(), ref)
    r.free()
    g.free()


def test_mfma_fp32_and_fp64_forms_agree(ctx):
    import sjamd

    pts, k, eps, ref = _case("clusters64d_k6")
    _say("[mfma vs fp64] clusters64d_k6")
    a = ctx.self_join(pts, k, eps, form="lid", precision=sjamd.FP64)
    b = ctx.self_join(pts, k, eps, form="lid", precision=sjamd.FP32_RECHECK)
    assert a.total == b.total == int(ref[0][-1])
    _assert_same(b.csr(), (O.csr_rows_sorted(*a.csr())))


def test_same_grid_fp32_walks_in_turn(ctx):
    """one grid joined in FP32 per lane (rows of d floats) and then on the
    matrix cores (rows padded to 16): the cached FP32 copy grows"""
    import sjamd

    pts, k, eps, ref = _case("clusters16d_k3")
    _say("[mfma] same grid, per-point then matrix-core FP32")
    g = ctx.grid_build(pts, k, eps)
    for per_point in (True, False, True):
        r = g.join(form="lid", precision=sjamd.FP32_RECHECK, per_point=per_point)
        _assert_same(r.csr(), ref)
        r.free()
    g.free()
