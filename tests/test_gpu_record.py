"""The fill that replays the count pass's hit record must give the table of
the fill that evaluates every distance again, and the oracle's, for every
form, precision and schedule of the LDS tile walk; a record that does not
fit makes that join evaluate again and the next join on the context
replay."""
import sys

import numpy as np
import pytest

import oracle as O
from sjamd import datasets as D

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(gpu, ref, tag):
    goff, gids = O.csr_rows_sorted(np.asarray(gpu[0], dtype=np.uint64), np.asarray(gpu[1], dtype=np.int32))
    roff, rids = ref
    assert int(goff[-1]) == int(roff[-1]), f"{tag}: |S| gpu {int(goff[-1])} oracle {int(roff[-1])}"
    np.testing.assert_array_equal(goff, roff, err_msg=tag)
    np.testing.assert_array_equal(gids, rids, err_msg=tag)


CASES = {
    "exp2d_200k": lambda: (D.exponential_clipped(200_000, 2, 2), 2, 5.0e-4),
    "uniform1d_30k": lambda: (D.uniform(30_000, 1, 21), 1, 2.0e-4),
    "uniform3d_60k": lambda: (D.uniform(60_000, 3, 11), 3, 0.03),
    "uniform6d_50k": lambda: (D.uniform(50_000, 6, 3), 6, 0.23),
    "uniform8d_12k": lambda: (D.uniform(12_000, 8, 14), 8, 0.3),
}

_cache = {}


def _case(name):
    if name not in _cache:
        pts, k, eps = CASES[name]()
        _cache[name] = (pts, k, eps, O.join(pts, k, eps, O.LID))
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("form", ["full", "unicomp", "bunicomp", "lid"])
def test_replayed_fill_matches(ctx, name, form):
    import sjamd

    pts, k, eps, ref = _case(name)
    for prec in (sjamd.FP64, sjamd.FP32_RECHECK):
        for sched in (sjamd.STRIDED, sjamd.WORKQUEUE):
            tag = f"{name} {form} prec={prec} sched={sched}"
            _say(f"[record] {tag}")
            g = ctx.grid_build(pts, k, eps)
            r0 = g.join(form=form, precision=prec, schedule=sched, hit_record=False)
            assert r0.stats()["fill_from_record"] == 0
            _same(r0.csr(), ref, tag + " evaluated")
            r0.free()
            r1 = g.join(form=form, precision=prec, schedule=sched)
            r2 = g.join(form=form, precision=prec, schedule=sched)
            assert r2.stats()["fill_from_record"] == 1, tag
            _same(r1.csr(), ref, tag + " first")
            _same(r2.csr(), ref, tag + " replayed")
            r1.free()
            r2.free()
            g.free()


def test_record_overflow_then_replay():
    """6-D uniform, 300k points: the record outgrows the initial space of a
    fresh context, so the first join evaluates again and the second
    replays"""
    import sjamd

    pts = D.uniform(300_000, 6, 17)
    eps = 0.2
    _say("[record] overflow uniform6d_300k")
    ref = O.join(pts, 6, eps, O.LID)
    c = sjamd.Context(0)
    try:
        g = c.grid_build(pts, 6, eps)
        r1 = g.join(form="lid")
        first = r1.stats()["fill_from_record"]
        _same(r1.csr(), ref, "first")
        r1.free()
        r2 = g.join(form="lid")
        assert r2.stats()["fill_from_record"] == 1
        _same(r2.csr(), ref, "second")
        r2.free()
        g.free()
        assert first == 0, "the record was expected to outgrow the initial space"
    finally:
        c.close()


def test_unusable_record_in_ahead_fill():
    """A context that has seen a larger result fills ahead of the host; when
    the record then outgrows its space (6-D uniform, 300k points), the
    replay is skipped on the device and the host evaluates the fill in a
    second round, into the same buffer"""
    import sjamd

    pts = D.uniform(300_000, 6, 17)
    eps = 0.2
    _say("[record] ahead fill, unusable record, uniform6d_300k")
    ref = O.join(pts, 6, eps, O.LID)
    c = sjamd.Context(0)
    try:
        ra = c.self_join(D.uniform(60_000, 2, 23), 2, 0.08, form="lid")
        big = ra.total
        ra.free()
        assert big > int(ref[0][-1]), "the earlier result must be the larger"
        g = c.grid_build(pts, 6, eps)
        r = g.join(form="lid")
        st = r.stats()
        _same(r.csr(), ref, "ahead, unusable record")
        r.free()
        g.free()
        assert st["fill_from_record"] == 0, "the record was expected to outgrow the initial space"
        assert st["overflow_retries"] == 0
    finally:
        c.close()


def test_record_with_batches(ctx):
    """result rounds: every round's rows equal the oracle's"""
    pts, k, eps, ref = _case("uniform6d_50k")
    _say("[record] batches uniform6d_50k")
    g = ctx.grid_build(pts, k, eps)
    for form in ("full", "lid", "bunicomp"):
        r = g.join(form=form, capacity=int(ref[0][-1]) // 5 + 1)
        assert r.stats()["num_batches"] >= 4
        _same(r.csr(), ref, f"batches {form}")
        r.free()
    g.free()


MFMA_CASES = {
    "clusters16d_k4_6k": lambda: (D.clusters(6_000, 16, 21), 4, 2.4),
    "uniform20d_k3_4k": lambda: (D.uniform(4_000, 20, 22), 3, 1.05),
}


@pytest.mark.parametrize("name", list(MFMA_CASES))
@pytest.mark.parametrize("form", ["full", "bunicomp", "lid"])
def test_matrix_core_replay_matches(name, form):
    """FP32+recheck on the matrix cores (d > 8): the second join on a fresh
    context replays the count's chunk record; whole and batched results
    equal the oracle"""
    import sjamd

    if name not in _cache:
        pts, k, eps = MFMA_CASES[name]()
        _cache[name] = (pts, k, eps, O.join(pts, k, eps, O.LID))
    pts, k, eps, ref = _cache[name]
    assert int(ref[0][-1]) > pts.shape[0]
    _say(f"[record mfma] {name} {form}")
    c = sjamd.Context(0)
    try:
        g = c.grid_build(pts, k, eps)
        for sched in (sjamd.STRIDED, sjamd.WORKQUEUE):
            r1 = g.join(form=form, precision=sjamd.FP32_RECHECK, schedule=sched)
            _same(r1.csr(), ref, f"{name} {form} sched={sched} first")
            r1.free()
            r2 = g.join(form=form, precision=sjamd.FP32_RECHECK, schedule=sched)
            assert r2.stats()["fill_from_record"] == 1
            _same(r2.csr(), ref, f"{name} {form} sched={sched} replayed")
            r2.free()
        r3 = g.join(form=form, precision=sjamd.FP32_RECHECK, capacity=int(ref[0][-1]) // 4 + 1)
        assert r3.stats()["num_batches"] >= 3 and r3.stats()["fill_from_record"] == 1
        _same(r3.csr(), ref, f"{name} {form} batches")
        r3.free()
        g.free()
    finally:
        c.close()


def test_batches_without_record(ctx):
    """result rounds whose fill evaluates distances again over the count
    pass's tiles (hit record off), every form and both precisions"""
    import sjamd

    pts, k, eps, ref = _case("uniform6d_50k")
    _say("[record] batches without record uniform6d_50k")
    g = ctx.grid_build(pts, k, eps)
    for form in ("full", "unicomp", "bunicomp", "lid"):
        for prec in (sjamd.FP64, sjamd.FP32_RECHECK):
            r = g.join(form=form, precision=prec, capacity=int(ref[0][-1]) // 5 + 1, hit_record=False)
            st = r.stats()
            assert st["num_batches"] >= 4 and st["fill_from_record"] == 0
            _same(r.csr(), ref, f"batches without record {form} prec={prec}")
            r.free()
    g.free()
