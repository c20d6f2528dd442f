"""Many joins in one graph-recorded persistent launch per pass: every
instance's neighbour table must equal the single-join result and the CPU
oracle, on the first run, on graph replays separated by eager work, after
re-tuning and after the autotuner."""
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


def _batch():
    """(points, k, eps): dims 1..8 with d == k (tile path), d != k and
    d = 20 (per-point path), uniform and exponential data, uneven sizes."""
    spec = [
        (2, 2, "u", 4_000, 0.02), (3, 3, "e", 9_000, 0.004), (4, 4, "u", 3_000, 0.12),
        (6, 6, "u", 12_000, 0.25), (8, 8, "e", 5_000, 0.02), (2, 2, "e", 30_000, 0.001),
        (1, 1, "u", 2_000, 0.0005), (5, 5, "u", 7_000, 0.2), (7, 7, "u", 2_500, 0.35),
        (5, 3, "u", 6_000, 0.18), (20, 4, "u", 3_000, 1.1), (3, 3, "u", 40_000, 0.025),
    ]
    out = []
    for i, (d, k, kind, n, eps) in enumerate(spec):
        pts = D.uniform(n, d, 900 + i) if kind == "u" else D.exponential_clipped(n, d, 900 + i)
        out.append((pts, k, eps))
    return out


_cache = {}


def _refs():
    if "refs" not in _cache:
        b = _batch()
        _cache["batch"] = b
        _cache["refs"] = [O.join(p, k, e, O.LID) for p, k, e in b]
    return _cache["batch"], _cache["refs"]


def test_batch_is_meaningful():
    b, refs = _refs()
    for (p, k, e), r in zip(b, refs):
        assert int(r[0][-1]) > p.shape[0]  # more than the self-pairs


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("schedule", [0, 1])
def test_many_matches_oracle(ctx, precision, schedule):
    b, refs = _refs()
    _say(f"[many] precision={precision} schedule={schedule} instances={len(b)}")
    m = ctx.many(b, form="bunicomp", precision=precision, schedule=schedule)
    for i, ref in enumerate(refs):
        _same(m.csr(i), ref, f"instance {i}")
    st = m.stats()
    assert st["total_pairs"] == sum(int(r[0][-1]) for r in refs)
    if precision == 1:
        assert st["fp32_misclassified"] <= st["fp32_band_pairs"]
    m.free()


def test_many_graph_replays_with_eager_work_between(ctx):
    """replays of the recorded graph, each after unrelated eager joins and
    synchronisations, give the eager (single-join) result every time"""
    import torch

    b, refs = _refs()
    _say("[many] replays")
    eager = []
    for p, k, e in b:
        r = ctx.self_join(p, k, e, form="lid")
        eager.append(O.csr_rows_sorted(*r.csr()))
        r.free()
    m = ctx.many(b, form="lid", precision=1, schedule=1)
    for rep in range(4):
        r = ctx.self_join(b[rep][0], b[rep][1], b[rep][2], form="full")  # eager work in between
        r.free()
        torch.cuda.synchronize()
        m.run()
        for i in range(len(b)):
            _same(m.csr(i), eager[i], f"replay {rep} instance {i}")
    m.free()


def test_many_tuning_and_autotune(ctx):
    import sjamd

    b, refs = _refs()
    _say("[many] tuning")
    m = ctx.many(b, form="lid", precision=1, schedule=0)
    for form, sched, bpc, ch in [("full", 1, 4, 2), ("unicomp", 1, 16, 4), ("bunicomp", 0, 32, 1)]:
        m.set_tuning(form, sched, bpc, ch)
        assert m.tuning()["form"] == sjamd.FORMS[form]
        for i in [0, 5, 10, 11]:
            _same(m.csr(i), refs[i], f"{form}/{sched}/{bpc}/{ch} instance {i}")
    best = m.autotune(tune_form=True, reps=2)
    assert best == m.tuning()
    assert best["schedule"] in (0, 1) and best["blocks_per_cu"] in (4, 8, 16, 32)
    for i, ref in enumerate(refs):
        _same(m.csr(i), ref, f"autotuned instance {i}")
    t = m.times()
    assert t["ms_run"] > 0 and t["ms_plan"] >= t["ms_grid"] > 0
    m.free()


def test_many_rejects_bad_input(ctx):
    import sjamd

    b, _ = _refs()
    with pytest.raises(sjamd.SJError):
        ctx.many([(b[0][0], 0, 0.1)])
    with pytest.raises(sjamd.SJError):
        ctx.many([(b[0][0], 2, -1.0)])
