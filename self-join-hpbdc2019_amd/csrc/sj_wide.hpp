// Device side of the plain FP64 traversal for points with d != k or d > 8 (up to 64 dims):
// the canonical predicate of kernel.cu:341-349 (acc = fma(df, df, acc) in
// dimension order, acc <= T) evaluated from LDS-staged candidates.
//
// A wave takes one row-aligned tile of up to 64 query points (one per
// lane).  For every claimed neighbour row the union of the tile's
// candidate ranges is staged in LDS, 32 candidates at a time, as FP64 rows.
// Each lane keeps one accumulator per staged candidate and walks the
// dimensions in order, two at a time; every candidate coordinate is one
// LDS broadcast read shared by all lanes.  The per-pair accumulation order
// is the canonical one, so the pair set is bit-exact without any recheck.
// Hits of a chunk form a 32-bit mask per lane; own hits are written one
// source lane at a time as contiguous row writes, mirrored hits go through
// LDS counters per candidate as in the low-dimensional tile walk.
#pragma once

#include "sj_traverse.hpp"

namespace sj {


constexpr int WIDE_CH = 32;  // staged candidates per chunk
constexpr int WIDE_MAXD = 64;

struct WideLds {
    double x[WIDE_CH * WIDE_MAXD];  // chunk rows, stride d
    int32_t id[WIDE_CH];
    uint32_t cnt[WIDE_CH];
    unsigned long long base[WIDE_CH];
};

template <int PASS>
__device__ __forceinline__ void process_tile_wide(const JoinArgs &a, uint32_t tile, WideLds &L, LaneStats &st) {
    const unsigned lane = __lane_id();
    const int k = a.k, d = a.d;
    const uint32_t ts0 = a.tileStart[tile], ts1 = a.tileStart[tile + 1];
    const uint64_t s64 = (uint64_t)ts0 + lane;
    const bool valid = s64 < ts1;
    const uint32_t s = valid ? (uint32_t)s64 : 0u;
    const uint64_t vmask = __ballot(valid);
    if (!vmask) return;
    const int lastL = 63 - __builtin_clzll(vmask);
    const uint32_t sEff = valid ? s : ts0 + (uint32_t)lastL;
    uint32_t c[DEV_MAXK];
    const uint32_t cellA = a.cellOf[sEff];
#pragma unroll
    for (int j = 0; j < DEV_MAXK; j++) c[j] = j < k ? a.cellCoord[(uint64_t)cellA * k + j] : 0u;
    const uint64_t linA = a.cellLinear[cellA];
    const uint64_t rowBase = linA - c[0];
    const uint64_t rb0 = readlane_u64(rowBase, 0);
    const bool badTile = __ballot(valid && rowBase != rb0) != 0ull;
    const uint32_t xa = readlane_u32(c[0], 0);
    const uint32_t xb = readlane_u32(c[0], lastL);
    const uint32_t nC0 = a.nCells[0];
    const uint32_t x = c[0];
    const double *qp = a.pts + (uint64_t)sEff * d;

    const bool ownW = valid && s >= a.wBegin && s < a.wEnd;
    const int32_t myId = valid ? (int32_t)a.origId[s] : 0;
    uint32_t own = 0;
    uint64_t ownPos = 0, ownLim = 0;
    if (PASS == PASS_FILL && ownW) {
        ownPos = a.offS[s] - a.outBase;
        ownLim = ownPos + a.ownCnt[s];
    }
    uint64_t cand = 0, err = 0;
    const bool anyMirror = a.form != SJ_FORM_FULL;

    int o[DEV_MAXK];
#pragma unroll
    for (int j = 0; j < DEV_MAXK; j++) o[j] = j < k ? -1 : 0;
    o[0] = 0;
    int nrow = 1;
    for (int j = 1; j < k; j++) nrow *= 3;
    if (badTile) {
        nrow = 0;
        if (lane == 0) err++;
    }
    for (int it = 0; it < nrow; ++it) {
        bool inb = true, originRow = true;
        int64_t dl = 0;
#pragma unroll
        for (int j = 1; j < DEV_MAXK; j++) {
            if (j < k) {
                originRow = originRow && o[j] == 0;
                const int64_t v = (int64_t)c[j] + o[j];
                inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
                dl += (int64_t)o[j] * a.mult[j];
            }
        }
        if (inb) {
            const uint64_t rb = (uint64_t)((int64_t)rowBase + dl);
            bool useL = false, useC = false, useR = false;
            if (valid) {
                int oo[DEV_MAXK];
#pragma unroll
                for (int j = 0; j < DEV_MAXK; j++) oo[j] = o[j];
                oo[0] = -1;
                useL = x >= 1 && claims_rt(a.form, k, c, oo, linA, rb + x - 1);
                oo[0] = 1;
                useR = x + 1 < nC0 && claims_rt(a.form, k, c, oo, linA, rb + x + 1);
                oo[0] = 0;
                useC = originRow || claims_rt(a.form, k, c, oo, linA, rb + x);
            }
            if (__ballot(useL || useC || useR)) {
                const uint32_t ux0 = xa >= 1 ? xa - 1 : 0u;
                const uint32_t ux1 = xb + 1 < nC0 ? xb + 1 : nC0 - 1;
                const uint32_t u0 = pos_lb(a, rb + ux0);
                const uint32_t u1 = pos_lb(a, rb + ux1 + 1);
                uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;
                if (valid) {
                    e1 = pos_lb(a, rb + x);
                    e2 = pos_lb(a, rb + x + 1);
                    e0 = x >= 1 ? pos_lb(a, rb + x - 1) : e1;
                    e3 = x + 1 < nC0 ? pos_lb(a, rb + x + 2) : e2;
                }
                if (PASS == PASS_COUNT && valid)
                    cand += (useL ? e1 - e0 : 0) + (useC ? e2 - e1 : 0) + (useR ? e3 - e2 : 0);
                const bool mirC = anyMirror && !originRow;
                for (uint32_t cb = u0; cb < u1; cb += WIDE_CH) {
                    const uint32_t nv = min((uint32_t)WIDE_CH, u1 - cb);
                    // stage: the chunk's rows are contiguous in the sorted copy
                    const double *src = a.pts + (uint64_t)cb * d;
                    for (uint32_t t = lane; t < nv * (uint32_t)d; t += 64) L.x[t] = src[t];
                    if (lane < WIDE_CH) {
                        L.id[lane] = lane < nv ? (int32_t)a.origId[cb + lane] : -1;
                        L.cnt[lane] = 0u;
                    }
                    wave_lds_sync();
                    // claimed candidates of this lane in the chunk
                    uint32_t claim = 0;
                    if (valid) {
                        if (useL) claim |= (uint32_t)range_bits(cb, e0, e1);
                        if (useC) claim |= (uint32_t)range_bits(cb, e1, e2);
                        if (useR) claim |= (uint32_t)range_bits(cb, e2, e3);
                    }
                    claim &= nv >= 32 ? 0xFFFFFFFFu : ((1u << nv) - 1u);
                    uint32_t hit = 0;
                    if (__ballot(claim != 0u)) {
                        // canonical sums, dimension order, two dims per step
                        double acc[WIDE_CH];
#pragma unroll
                        for (int cc = 0; cc < WIDE_CH; cc++) acc[cc] = 0.0;
                        // fma(f, f, +0) = fl(f*f): the first step equals the
                        // canonical acc = df*df
                        int j = 0;
                        for (; j + 1 < d; j += 2) {
                            const double q0 = qp[j], q1 = qp[j + 1];
#pragma unroll
                            for (int cc = 0; cc < WIDE_CH; cc++) {
                                const double f0 = L.x[cc * d + j] - q0, f1 = L.x[cc * d + j + 1] - q1;
                                acc[cc] = fma(f0, f0, acc[cc]);
                                acc[cc] = fma(f1, f1, acc[cc]);
                            }
                        }
                        if (j < d) {
                            const double q0 = qp[j];
#pragma unroll
                            for (int cc = 0; cc < WIDE_CH; cc++) {
                                const double f0 = L.x[cc * d + j] - q0;
                                acc[cc] = fma(f0, f0, acc[cc]);
                            }
                        }
#pragma unroll
                        for (int cc = 0; cc < WIDE_CH; cc++) hit |= (acc[cc] <= a.T ? 1u : 0u) << cc;
                        hit &= claim;
                    }
                    uint32_t mirBits = 0;
                    if (anyMirror) {
                        uint32_t elig = (uint32_t)range_bits(cb, e0, e1) | (uint32_t)range_bits(cb, e2, e3);
                        if (mirC) elig |= (uint32_t)range_bits(cb, e1, e2);
                        mirBits = hit & elig & (uint32_t)range_bits(cb, a.wBegin, a.wEnd);
                    }
                    emit_chunk32<PASS>(a, cb, nv, hit, mirBits, ownW, myId, ownPos, ownLim, own, L.id, L.cnt,
                                       L.base, anyMirror, err);
                    wave_lds_sync();
                }
            }
        }
        bool carry = true;
#pragma unroll
        for (int j = 1; j < DEV_MAXK; j++) {
            if (j < k && carry) {
                if (o[j] == 1) {
                    o[j] = -1;
                } else {
                    o[j]++;
                    carry = false;
                }
            }
        }
    }
    if (PASS == PASS_COUNT) {
        if (ownW) a.ownCnt[s] = badTile ? 0u : own;
    } else {
        if (ownW && !badTile && own != (ownLim - ownPos)) err++;
    }
    st.cand += cand;
    st.err += err;
}

}  // namespace sj
