"""GPU parity of the workload kernels (W1/W2), the batch estimators (B1),
the key/value view and the degenerate inputs of the single join, against
the CPU oracle.

Reference: sortByWorkLoad / sortByWorkLoadLidUnicomp kernel.cu:33-250 and
the descending order GPU.cu:1695-1698; callGPUBatchEst GPU.cu:854-1065 and
callGPUBatchEstWorkQueue GPU.cu:1072-1284; thrust::sort_by_key
GPU.cu:2326-2349; argument handling main.cu:46-71."""
import math
import sys

import numpy as np
import pytest

import oracle as O
import sjamd
from sjamd import datasets as D

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    "uniform2d_30k": lambda: (D.uniform(30_000, 2, 21, 0.0, 50.0), 2, 0.5),
    "exp2d_100k": lambda: (D.exponential_clipped(100_000, 2, 22), 2, 6.0e-4),
    "uniform3d_40k": lambda: (D.uniform(40_000, 3, 23), 3, 0.04),
    "uniform6d_30k": lambda: (D.uniform(30_000, 6, 24), 6, 0.25),
    "clusters16d_k4_5k": lambda: (D.clusters(5_000, 16, 25), 4, 2.0),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("lid", [False, True])
def test_workload_and_order_match_oracle(ctx, name, lid):
    pts, k, eps = CASES[name]()
    _say(f"[workload] {name} lid={lid}")
    og = O.OracleGrid(pts, k, eps)
    g = ctx.grid_build(pts, k, eps)
    w, order = g.workload(lid=lid)
    np.testing.assert_array_equal(w, og.workload(lid))
    np.testing.assert_array_equal(order, og.workload_order(lid))
    g.free()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("stride", [100, 7])
def test_batch_estimate_matches_oracle(ctx, name, stride):
    pts, k, eps = CASES[name]()
    _say(f"[estimate] {name} stride={stride}")
    og = O.OracleGrid(pts, k, eps)
    g = ctx.grid_build(pts, k, eps)
    e = g.batch_estimate(stride)
    cnt = og.sample_count(stride)
    assert e["sample_count"] == cnt
    pol = O.batch_policy(cnt, stride)
    assert e["estimated"] == pol["estimated"]
    assert e["estimated_with_alpha"] == pol["estimated_with_alpha"]
    assert e["num_batches"] == pol["num_batches"]
    assert e["buffer_size"] == pol["buffer_size"]
    for wq in ("full", "lid"):
        e = g.batch_estimate(stride, workqueue=wq)
        cnt = og.sample_count_workqueue(stride, wq == "lid")
        assert e["sample_count"] == cnt, wq
        pol = O.batch_policy(cnt, stride)
        assert (e["estimated_with_alpha"], e["num_batches"], e["buffer_size"]) == \
            (pol["estimated_with_alpha"], pol["num_batches"], pol["buffer_size"]), wq
    g.free()


@pytest.mark.parametrize("capacity_div", [0, 5])
def test_key_value_pairs(ctx, capacity_div):
    """sj_result_copy_pairs: keys ascending, each key's values the row."""
    pts, k, eps = CASES["uniform3d_40k"]()
    off, ids = O.join(pts, k, eps, O.LID, sort_rows=True)
    total = int(off[-1])
    cap = total // capacity_div + 1 if capacity_div else 0
    _say(f"[pairs] uniform3d_40k capacity={cap}")
    r = ctx.self_join(pts, k, eps, form="unicomp", capacity=cap, sort_rows=True)
    if capacity_div:
        assert r.stats()["num_batches"] >= 4
    keys, vals = r.pairs()
    assert keys.size == total
    np.testing.assert_array_equal(keys, np.repeat(np.arange(pts.shape[0], dtype=np.int32),
                                                  np.diff(off.astype(np.int64))))
    np.testing.assert_array_equal(vals, ids)
    # the reference builds its table from exactly this view (GPU.cu:2595-2637)
    tab = O.neighbor_table(keys, vals, pts.shape[0])
    np.testing.assert_array_equal(tab[:, 1], off[:-1].astype(np.int64))
    np.testing.assert_array_equal(tab[:, 2], off[1:].astype(np.int64) - 1)


def _status(fn):
    with pytest.raises(sjamd.SJError) as e:
        fn()
    return e.value.status


def test_degenerate_inputs(ctx):
    _say("[degenerate] single join")
    pts = D.uniform(1000, 2, 26)
    assert _status(lambda: ctx.self_join(np.zeros((0, 2)), 2, 0.1)) == 1  # N = 0
    assert _status(lambda: ctx.self_join(pts, 2, 0.0)) == 1               # eps = 0
    assert _status(lambda: ctx.self_join(pts, 2, -0.5)) == 1              # eps < 0
    assert _status(lambda: ctx.self_join(pts, 2, math.nan)) == 1          # eps NaN
    assert _status(lambda: ctx.self_join(pts, 2, math.inf)) == 1          # eps inf
    assert _status(lambda: ctx.self_join(pts, 3, 0.1)) == 1               # k > d
    bad = pts.copy()
    bad[417, 1] = math.nan
    assert _status(lambda: ctx.self_join(bad, 2, 0.1)) == 1               # NaN coordinate
    bad[417, 1] = math.inf
    assert _status(lambda: ctx.self_join(bad, 2, 0.1)) == 1               # inf coordinate
    # nCells = 2e15 per dim: totalCells overflows 2^64
    wide = np.array([[0.0, 0.0], [1.0e15, 1.0e15]])
    assert _status(lambda: ctx.self_join(wide, 2, 0.5)) == 4
    # the context still works after every rejected call
    r = ctx.self_join(pts, 2, 0.05)
    off, ids = r.csr()
    roff, rids = O.join(pts, 2, 0.05, O.LID, sort_rows=True)
    np.testing.assert_array_equal(off, roff)
    np.testing.assert_array_equal(O.csr_rows_sorted(off, ids)[1], rids)


def test_context_close_frees_live_handles():
    """Closing a context first frees the grids and results still alive on
    it; their later collection does not touch the destroyed context."""
    import gc

    _say("[lifetime] close with live handles")
    c = sjamd.Context(0)
    pts = D.uniform(5_000, 2, 27)
    g = c.grid_build(pts, 2, 0.01)
    r = c.self_join(pts, 2, 0.01)
    r2 = g.join(form="lid")
    c.close()
    assert not g.h and not r.h and not r2.h
    del g, r, r2
    gc.collect()


@pytest.mark.parametrize("name", ["uniform2d_30k", "exp2d_100k", "uniform6d_30k", "clusters16d_k4_5k"])
def test_threads_per_point_schedule(ctx, name):
    """Reference mode 8 (kernel.cu:2072-2188): 16/8/2 lanes per point over
    the workload order, full stencil, FP64; one window and >= 4 windows."""
    pts, k, eps = CASES[name]()
    _say(f"[threads per point] {name}")
    off, ids = O.join(pts, k, eps, O.LID, sort_rows=True)
    for cap in (0, int(off[-1]) // 4 + 1):
        r = ctx.self_join(pts, k, eps, form="full", schedule=sjamd.THREADS_PER_POINT, capacity=cap)
        if cap:
            assert r.stats()["num_batches"] >= 4
        goff, gids = r.csr()
        np.testing.assert_array_equal(goff, off)
        np.testing.assert_array_equal(O.csr_rows_sorted(goff, gids)[1], ids)
        r.free()
    # the schedule is defined for the full stencil in FP64 only
    assert _status(lambda: ctx.self_join(pts, k, eps, form="lid", schedule=sjamd.THREADS_PER_POINT)) == 5
    assert _status(lambda: ctx.self_join(pts, k, eps, form="full", schedule=sjamd.THREADS_PER_POINT,
                                         precision=sjamd.FP32_RECHECK)) == 5
