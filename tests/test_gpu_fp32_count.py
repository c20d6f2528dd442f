"""sj_stats.fp32_misclassified against a host count.

Fixture: 2-D points on the 2^-10 lattice (coordinates below 2^14), so every
difference of coordinates, and every coordinate minus a tile centre, is
exact in FP64 and FP32.  The VALU tile walk's FP32 squared distance (FMA
chain of the squared FP32 differences, dimension order) then depends only
on the pair, and the host reproduces it bit for bit.  Pairs are placed at
integer offsets (dx, dy) with dx^2 + dy^2 within 2^7 of 2^26 lattice units
around eps = 8 (T = 64 + 2^-46), where FP32 rounds many of them across the
threshold.  With the FULL stencil every ordered pair is evaluated once, so
the count is the number of ordered pairs whose plain FP32 decision
(s32 <= fl32(T)) differs from the FP64 predicate."""
import numpy as np
import pytest

import oracle as O
import sjamd
from fp32emu import F32, fma32

pytestmark = pytest.mark.gpu

EPS = 8.0
UNIT = 2.0 ** -10


def fixture_points():
    target = 1 << 26
    vecs = []
    for dx in range(0, 8200):
        lo = target - 128 - dx * dx
        hi = target + 128 - dx * dx
        if hi < 0:
            break
        for dy in range(int(np.sqrt(max(lo, 0))) - 1, int(np.sqrt(hi)) + 2):
            if dy >= 0 and lo <= dy * dy <= hi:
                vecs.append((dx, dy))
    vecs = np.array(sorted(set(vecs)), dtype=np.int64)
    pts = []
    for g, (dx, dy) in enumerate(vecs.tolist()):
        ax, ay = 24 * 1024 * (g % 24), 24 * 1024 * (g // 24)  # groups 24 eps apart
        pts.append((ax, ay))
        pts.append((ax + dx, ay + dy))
    return np.array(pts, dtype=np.float64) * UNIT, len(vecs)


def host_misclassified(pts):
    T = O.eps_threshold(EPS)
    t32 = F32(T)
    n = pts.shape[0]
    i, j = np.nonzero(np.ones((n, n), dtype=bool))
    near = np.abs(pts[i] - pts[j]).max(1) <= 2 * EPS
    i, j = i[near], j[near]
    f = (pts[j].astype(F32) - pts[i].astype(F32)).astype(F32)  # exact
    s32 = (f[:, 0] * f[:, 0]).astype(F32)
    s32 = fma32(f[:, 1], f[:, 1], s32)
    d64 = O.dist2_batch(pts[i], pts[j])
    return int(((s32 <= t32) != (d64 <= T)).sum())


def test_fp32_misclassified_matches_host_count(ctx):
    pts, nvec = fixture_points()
    print(f"[fp32 count] {pts.shape[0]} points, {nvec} offsets", flush=True)
    want = host_misclassified(pts)
    assert want > 20  # the fixture does exercise the FP32 rounding
    off, ids = O.join(pts, 2, EPS, O.LID, sort_rows=True)
    r = ctx.self_join(pts, 2, EPS, form="full", precision=sjamd.FP32_RECHECK, sort_rows=True)
    st = r.stats()
    goff, gids = r.csr()
    np.testing.assert_array_equal(goff, off)
    np.testing.assert_array_equal(gids, ids)
    assert st["fp32_misclassified"] == want
    assert st["fp32_band_pairs"] >= want
    r.free()
