"""Host checks of the FP32 precision form's error bands (no GPU).

The join decides a candidate in FP32 and re-decides in FP64 only inside a
band; it is exact only if the band covers every FP32 error.  These tests
emulate, in numpy float32 arithmetic with correctly rounded FMAs, the exact
recipes the kernels run, over pairs placed within a few ulps of the
threshold, and assert the implications the band promises against the
canonical FP64 predicate (the oracle's FMA chain, kernel.cu:341-349):
  VALU walk:    d2_64 <= T  =>  s32 <= hi,   d2_64 > T  =>  s32 >= lo
  matrix cores: d2 + M <= tlo  =>  d2_64 <= T,   d2 - M > thi  =>  d2_64 > T
A band that dropped the staging error (the dominant term at these
radii) fails them: on these pairs the observed error reaches 26-38% of the
VALU band's half-width."""
import numpy as np
import pytest

import oracle as O
import sjamd
from fp32emu import F32, fma32


def near_threshold_pairs(rng, m, d, T, R, centre):
    """Pairs (p, q) with every coordinate within R of the centre and
    |p - q|^2 within a few ulps of T."""
    r = np.sqrt(T)
    dirs = rng.normal(size=(m, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rel = rng.uniform(-4e-15, 4e-15, size=(m, 1))
    p = centre + rng.uniform(-(R - r), R - r, size=(m, d))
    q = p + dirs * (r * (1.0 + rel))
    return p, q


CASES_VALU = [(d, eps, m) for d in (2, 3, 6, 8) for eps, m in ((1.5962e-4, 4.0), (0.125, 6.0), (3.7, 64.0))]


@pytest.mark.parametrize("d,eps,mult", CASES_VALU)
def test_valu_band_covers_fp32_error(d, eps, mult):
    rng = np.random.default_rng(1000 + d)
    T = O.eps_threshold(eps)
    R = mult * eps * (1.0 + 2.0 ** -20)  # the tile walk's bound of |x - centre|
    usable, lo, hi, t32 = sjamd.fp32_band(T, R, d)
    assert usable
    assert lo < T < hi and t32 == F32(T)
    centre = rng.uniform(0.0, 1.0, size=d)
    p, q = near_threshold_pairs(rng, 200_000, d, T, R, centre)
    d64 = O.dist2_batch(p, q)
    # stored coordinates y = fl32(fl64(x - c)), s32 = FMA chain of the
    # FP32 differences in dimension order (d2_fp32 in the join kernel)
    yp = (p - centre).astype(F32)
    yq = (q - centre).astype(F32)
    f = (yq - yp).astype(F32)
    s32 = (f[:, 0] * f[:, 0]).astype(F32)
    for j in range(1, d):
        s32 = fma32(f[:, j], f[:, j], s32)
    hit = d64 <= T
    assert hit.any() and (~hit).any()
    assert (s32[hit] <= F32(hi)).all()
    assert (s32[~hit] >= F32(lo)).all()
    # the band is the error bound at the threshold: |s32 - d2_64| <= M
    M = max(hi - T, T - lo)
    assert np.abs(s32.astype(np.float64) - d64).max() <= M
    # and it is not vacuous: a few ulps of T wide, far below T
    assert M < 1e-4 * T


def test_valu_band_unusable_when_too_wide():
    T = O.eps_threshold(1e-6)
    usable, *_ = sjamd.fp32_band(T, 1e6, 2)
    assert not usable


def _dot_orders(A, B):
    """float32 dot products of rows in three summation orders (the matrix
    core's order is not specified): sequential FMA, pairwise tree of
    rounded products, reversed FMA."""
    m, n = A.shape
    seq = np.zeros(m, dtype=F32)
    for t in range(n):
        seq = fma32(A[:, t], B[:, t], seq)
    prod = (A * B).astype(F32)
    while prod.shape[1] > 1:
        if prod.shape[1] % 2:
            prod = np.concatenate([prod, np.zeros((m, 1), dtype=F32)], axis=1)
        prod = (prod[:, 0::2] + prod[:, 1::2]).astype(F32)
    rev = np.zeros(m, dtype=F32)
    for t in reversed(range(n)):
        rev = fma32(A[:, t], B[:, t], rev)
    return [seq, prod[:, 0], rev]


def _norm_chain(V):
    nn = np.zeros(V.shape[0], dtype=F32)
    for t in range(V.shape[1]):
        nn = fma32(V[:, t], V[:, t], nn)
    return nn


CASES_MFMA = [(d, eps, src32) for d in (6, 16, 64) for eps in (0.4, 4.5) for src32 in (False, True)]


@pytest.mark.parametrize("d,eps,src32", CASES_MFMA)
def test_mfma_band_decisions_are_exact(d, eps, src32):
    rng = np.random.default_rng(2000 + d)
    T = O.eps_threshold(eps)
    span = 3.0 * eps
    origin = np.full(d, -0.5)
    centre = origin + rng.uniform(0.5, 8.0, size=d)
    p, q = near_threshold_pairs(rng, 60_000, d, T, span, centre)
    tile0 = centre  # the tile's first point
    if src32:
        # FP32 copy relative to the grid origin, rows within R of it
        P32 = (p - origin).astype(F32)
        Q32 = (q - origin).astype(F32)
        C32 = (tile0 - origin).astype(F32)
        R = float(np.sqrt(((np.abs(np.concatenate([p, q, tile0[None]])) - origin).max(0) ** 2).sum()))
        R = max(R, float(np.linalg.norm(np.abs(np.concatenate([p, q]) - origin).max(0))))
        R *= 1.0 + 2.0 ** -20
        A = (P32 - C32).astype(F32)
        B = (Q32 - C32).astype(F32)
    else:
        R = -1.0
        A = (p - tile0).astype(F32)
        B = (q - tile0).astype(F32)
    DP, alpha, beta, tlo, thi = sjamd.mfma_band(T, R, d)
    assert DP >= d
    d64 = O.dist2_batch(p, q)
    sq = (_norm_chain(A) + _norm_chain(B)).astype(F32)
    M = fma32(np.full_like(sq, alpha), sq, F32(beta))
    hit = d64 <= T
    assert hit.any() and (~hit).any()
    for dot in _dot_orders(A, B):
        d2 = fma32(np.full_like(sq, -2.0), dot, sq)
        certain_hit = (d2 + M).astype(F32) <= F32(tlo)
        certain_miss = ~((d2 - M).astype(F32) <= F32(thi))
        assert hit[certain_hit].all()
        assert (~hit[certain_miss]).all()
    # the band leaves most far pairs decided: pairs at 0.5 T and 2 T
    far_q = p + (q - p) * np.sqrt(0.5)
    A2 = A if not src32 else A
    B2 = ((far_q - origin).astype(F32) - C32).astype(F32) if src32 else (far_q - tile0).astype(F32)
    sq2 = (_norm_chain(A2) + _norm_chain(B2)).astype(F32)
    d2 = fma32(np.full_like(sq2, -2.0), _dot_orders(A2, B2)[0], sq2)
    M2 = fma32(np.full_like(sq2, alpha), sq2, F32(beta))
    assert ((d2 + M2).astype(F32) <= F32(tlo)).mean() > 0.99
