"""The allocation-free entry point (sj_self_join_static) through the C ABI
against sj_self_join's device table, to the same bar as every join test:
offsets, row IDs and (canonical, sorted) neighbour rows identical.  Two
calls in a row, the second with a smaller input, so a stale scratch would
show."""
import ctypes as C
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


def _lib():
    L = sjamd.lib()
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int32
    L.sj_self_join_static.argtypes = [vp, vp, u64, i32, i32, C.c_double, C.POINTER(sjamd.JoinParams), vp, vp,
                                      u64, vp, C.POINTER(u64)]
    L.sj_static_need.argtypes = [u64, i32, i32, C.POINTER(sjamd.JoinParams)]
    L.sj_static_need.restype = u64
    L.sj_static_scratch_bytes.restype = u64
    return L


def _static_join(ctx, x, k, eps, cap, **kw):
    import torch

    L = _lib()
    n, d = x.shape
    p = sjamd.make_params(**kw)
    off = torch.empty(n + 1, dtype=torch.int64, device=x.device)
    ids = torch.empty(max(cap, 1), dtype=torch.int32, device=x.device)
    rows = torch.empty(n, dtype=torch.int32, device=x.device)
    tot = C.c_uint64()
    torch.cuda.synchronize()
    st = L.sj_self_join_static(ctx.h, C.c_void_p(x.data_ptr()), n, d, k, float(eps), C.byref(p),
                               C.c_void_p(off.data_ptr()), C.c_void_p(ids.data_ptr()), cap,
                               C.c_void_p(rows.data_ptr()), C.byref(tot))
    return st, tot.value, off.cpu().numpy(), ids.cpu().numpy(), rows.cpu().numpy()


def _reference_table(ctx, x, k, eps, **kw):
    r = ctx.self_join(x, k, eps, **kw)
    n = x.shape[0]
    o, i, rid = r.device_table()
    off = sjamd.download(o, n + 1, np.uint64)
    ids = sjamd.download(i, r.total, np.int32)
    rows = sjamd.download(rid, n, np.uint32)
    r.free()
    return off, ids, rows


CASES = [
    ("exp2d_300k", lambda: D.exponential_clipped(300_000, 2, 2), 2, 1.2e-3,
     dict(form="bunicomp", schedule=sjamd.WORKQUEUE, sort_rows=True)),
    ("uniform3d_80k", lambda: D.uniform(80_000, 3, 7), 3, 0.02,
     dict(form="lid", precision=sjamd.FP32_RECHECK, sort_rows=True)),
    ("clusters16d_k4_20k", lambda: D.clusters(20_000, 16, 9), 4, 2.0,
     dict(form="unicomp", precision=sjamd.FP32_RECHECK, sort_rows=True)),
]


@pytest.mark.parametrize("name,make,k,eps,kw", CASES, ids=[c[0] for c in CASES])
def test_static_matches_allocating_entry(ctx, name, make, k, eps, kw):
    import torch

    pts = make()
    for step, n in enumerate([pts.shape[0], pts.shape[0] // 3]):
        x = ctx.to_device(pts[:n])
        _say(f"[no-alloc] {name} call {step + 1}: n={n}")
        roff, rids, rrows = _reference_table(ctx, x, k, eps, **kw)
        total = int(roff[-1])
        st, tot, off, ids, rows = _static_join(ctx, x, k, eps, total + 16, **kw)
        assert st == 0, sjamd.lib().sj_last_error()
        assert tot == total
        np.testing.assert_array_equal(off.astype(np.uint64), roff)
        np.testing.assert_array_equal(rows.astype(np.uint32), rrows)
        np.testing.assert_array_equal(ids[:total], rids)
        del x
        torch.cuda.synchronize()


def test_static_against_oracle_and_limits(ctx):
    import torch

    L = _lib()
    pts = D.uniform(50_000, 2, 31, 0.0, 20.0)
    x = ctx.to_device(pts)
    _say("[no-alloc] oracle parity, capacity and size limits")
    roff, rids = O.join(pts, 2, 0.05, O.LID, sort_rows=True)
    total = int(roff[-1])
    st, tot, off, ids, rows = _static_join(ctx, x, 2, 0.05, total, form="lid", sort_rows=True)
    assert st == 0 and tot == total
    # cell-order rows back to original order
    cnt = np.diff(off.astype(np.int64))
    out_off = np.zeros(pts.shape[0] + 1, dtype=np.int64)
    out_off[1:][rows] = cnt
    out_off = np.cumsum(out_off)
    out = np.empty(total, dtype=np.int32)
    for r in range(pts.shape[0]):
        out[out_off[rows[r]]:out_off[rows[r]] + cnt[r]] = ids[off[r]:off[r + 1]]
    np.testing.assert_array_equal(out_off.astype(np.uint64), roff)
    np.testing.assert_array_equal(out, rids)
    # a capacity one short: SJ_ERR_OVERFLOW with |S| reported
    st, tot, *_ = _static_join(ctx, x, 2, 0.05, total - 1, form="lid")
    assert st == 4 and tot == total
    # an input above the scratch: rejected before any device work
    p = sjamd.make_params(form="lid")
    assert L.sj_static_need(50_000, 2, 2, C.byref(p)) < L.sj_static_scratch_bytes()
    big = 4_000_000
    assert L.sj_static_need(big, 2, 2, C.byref(p)) > L.sj_static_scratch_bytes()
    dummy = torch.empty(1, dtype=torch.float64, device=x.device)
    tot = C.c_uint64()
    st = L.sj_self_join_static(ctx.h, C.c_void_p(dummy.data_ptr()), big, 2, 2, 0.05, C.byref(p),
                               C.c_void_p(dummy.data_ptr()), C.c_void_p(dummy.data_ptr()), 0,
                               C.c_void_p(dummy.data_ptr()), C.byref(tot))
    assert st == 4
    assert b"static scratch" in sjamd.lib().sj_last_error()
    # the context still serves the allocating entry points afterwards
    r = ctx.self_join(x, 2, 0.05, form="lid")
    assert r.total == total
    r.free()
