// Device side of the FP32 norm-expansion traversal on the matrix cores (the precision form of
// the join, reference predicate kernel.cu:341-349), for 1 <= d <= 64.
//
// A wave takes one row-aligned tile of up to 64 query points.  Queries and
// candidates are centred on the tile's first point: in FP64 and rounded to
// FP32 once (the rounding error is relative to the small centred vectors),
// or, when coordinates are within 1024 eps of the grid origin, in FP32
// from the FP32 copy (rows of DP floats: half the staged bytes and float4
// staging, an origin-relative bound).  For every claimed neighbour row the union of the tile's
// candidate ranges is staged in LDS, 32 candidates at a time, and
// v_mfma_f32_32x32x2f32 forms the 64 x 32 dot products (dims padded with
// zeros to DP = 4, 8, 16, 32 or 64).  With s = |q|^2 + |c|^2 the squared
// distance is d2 = s - 2 q.c, and the bound M = alpha*s + beta of
// mfma_consts splits the pairs:
//   d2 + M <= tlo  -> hit, certain;
//   d2 - M >  thi  -> miss, certain;
//   otherwise      -> decided by the canonical FP64 predicate.
// The pair set is therefore the FP64 one whatever the rounding of the
// matrix cores.  The compares come out as wave ballots, two queries per
// output register; a select per ballot regroups them so that each lane
// holds its own query's 32-bit masks, and the chunk is emitted like the other
// traversals (claims, recheck, own and mirrored hits per lane).
//
// Output layout of v_mfma_f32_32x32x2f32 used here: register r of lane l
// holds C[i][j] with j = l % 32 and i = 8*(r/4) + 4*(l/32) + r%4; operand A
// element (i, kk) sits in lane i + 32*kk, operand B element (kk, j) in lane
// j + 32*kk.  Lanes 0-31 carry dims [0, DP/2) and lanes 32-63 dims
// [DP/2, DP), one dim pair per instruction.
#pragma once

#include "sj_traverse.hpp"

namespace sj {

typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int DP>
struct MfmaLds {
    float x[32][DP + 4];  // staged candidates, centred; +4 floats keep b128 reads conflict-free
    float nc[32];         // their squared norms (1e30 for empty slots)
    int32_t id[32];       // their original IDs
    uint32_t cnt[32];     // mirrored hits per candidate
    unsigned long long base[32];
    double ctr[DP];       // tile centre, FP64
    float ctr32[DP];      // tile centre, FP32 copy
    float nq[64];         // squared norms of the tile's queries
};

// RP (fill only): take each chunk's hit masks from the count pass's hit
// record (one directory entry per executed 32-candidate chunk) instead of
// the matrix cores; chunks without hits are skipped.
template <int DP, int PASS, bool RP = false>
__device__ __forceinline__ void process_tile_mfma(const JoinArgs &a, uint32_t tile, MfmaLds<DP> &L,
                                                  LaneStats &st, TileRec &R) {
    static_assert(!RP || PASS == PASS_FILL, "only the fill replays a record");
    constexpr int HD = DP / 2;  // dims per lane half
    const unsigned lane = __lane_id();
    const unsigned jl = lane & 31u, g = lane >> 5;
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
    // tiles are row-aligned by construction; anything else is reported
    const bool badTile = __ballot(valid && rowBase != rb0) != 0ull;
    const uint32_t xa = readlane_u32(c[0], 0);
    const uint32_t xb = readlane_u32(c[0], lastL);
    const uint32_t nC0 = a.nCells[0];
    const uint32_t x = c[0];
    const double *P = a.pts;
    const float *P32 = a.pts32;
    const bool src32 = a.mSrc32 != 0;  // stage from the FP32 copy (origin-relative bound)
    // hit record: entries [tileDir[tile], tileDir[tile+1]), one per chunk
    const bool recOn = a.recDir != nullptr;
    bool rec = false;
    uint64_t dbase = 0, dend = 0, piece_it = 0;
    if (PASS == PASS_COUNT && recOn) {
        dbase = readlane_u64(a.tileDir[tile], 0);
        dend = readlane_u64(a.tileDir[tile + 1], 0);
        rec = dend <= a.dirCap && !badTile;
        if (!rec && lane == 0) atomicAdd(&a.recCtr[REC_BAD], 1ull);
    }
    if constexpr (RP) dbase = readlane_u64(a.tileDir[tile], 0);

    // operand A: queries 32h + jl, dims [g*HD, (g+1)*HD), centred in FP64
    float A0[HD], A1[HD];
    if constexpr (!RP) {
    // tile centre: the first point (the FP32 copy has rows of DP floats,
    // zero padded)
    if (src32) {
        for (int j = (int)lane; j < DP; j += 64) L.ctr32[j] = P32[(uint64_t)ts0 * DP + j];
    } else {
        for (int j = (int)lane; j < d; j += 64) L.ctr[j] = P[(uint64_t)ts0 * d + j];
    }
    wave_lds_sync();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint64_t q = (uint64_t)ts0 + 32u * h + jl;
        const bool qv = q < ts1;
        float *A = h == 0 ? A0 : A1;
        float nq = 0.0f;
#pragma unroll
        for (int t = 0; t < HD; t++) {
            const int j = (int)g * HD + t;
            float v = 0.0f;
            if (qv && j < d) v = src32 ? P32[q * DP + j] - L.ctr32[j] : (float)(P[q * d + j] - L.ctr[j]);
            A[t] = v;
            nq = fmaf(v, v, nq);
        }
        nq += __shfl_xor(nq, 32, 64);
        if (g == 0) L.nq[32 * h + jl] = nq;
    }
    }  // operand A

    const bool ownW = valid && s >= a.wBegin && s < a.wEnd;
    const int32_t myId = valid ? (int32_t)a.origId[s] : 0;
    uint32_t own = 0;
    uint64_t ownPos = 0, ownLim = 0;
    if (PASS == PASS_FILL && ownW) {
        ownPos = a.offS[s] - a.outBase;
        ownLim = ownPos + a.ownCnt[s];
    }
    uint64_t cand = 0, band = 0, mis = 0, err = 0;
    const bool anyMirror = a.form != SJ_FORM_FULL;
    // a fill writes only rows [wBegin, wEnd): chunks without such candidates
    // matter only to tiles that own rows there
    const bool anyOwnW = __ballot(ownW) != 0ull;
    // staging: element t of a chunk (row t / d, dim t % d) goes to lane t % 64
    const uint32_t cstep = 64u / (uint32_t)d, jstep = 64u % (uint32_t)d;
    const uint32_t c0l = lane / (uint32_t)d, j0l = lane % (uint32_t)d;

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
                for (uint32_t cb = u0; cb < u1; cb += 32) {
                    const uint32_t nv = min(32u, u1 - cb);
                    if (PASS == PASS_FILL && !anyOwnW && (cb + nv <= a.wBegin || cb >= a.wEnd)) {
                        piece_it++;  // nothing of this chunk lands in the rows written
                        continue;
                    }
                    if constexpr (RP) {
                        const uint64_t *e = a.recDir + REC_DW * (dbase + piece_it);
                        const uint64_t pres = e[0];
                        piece_it++;
                        if (!pres) continue;  // no hits in the chunk
                        const uint32_t hit =
                            (pres >> lane) & 1ull ? (uint32_t)a.recMask[e[1] + popc_below64(pres)] : 0u;
                        if (lane < 32) {
                            L.id[lane] = lane < nv ? (int32_t)a.origId[cb + lane] : -1;
                            L.cnt[lane] = 0u;
                        }
                        wave_lds_sync();
                        uint32_t mirBits = 0;
                        if (anyMirror) {
                            uint32_t elig = (uint32_t)range_bits(cb, e0, e1) | (uint32_t)range_bits(cb, e2, e3);
                            if (mirC) elig |= (uint32_t)range_bits(cb, e1, e2);
                            mirBits = hit & elig & (uint32_t)range_bits(cb, a.wBegin, a.wEnd);
                        }
                        emit_chunk32<PASS>(a, cb, nv, hit, mirBits, ownW, myId, ownPos, ownLim, own, L.id, L.cnt,
                                           L.base, anyMirror, err);
                        wave_lds_sync();
                        continue;
                    }
                    // stage: the chunk's rows are contiguous in the sorted copy
                    if (src32) {
                        // float4 f of the chunk goes to lane f % 64; the DP/4
                        // lanes of a candidate are consecutive, so its norm is
                        // a shuffle reduction
                        for (uint32_t f = lane; f < 8u * DP; f += 64) {
                            const uint32_t cj = f / (DP / 4), part = f % (DP / 4);
                            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                            if (cj < nv) {
                                const float4 w =
                                    *reinterpret_cast<const float4 *>(P32 + (uint64_t)(cb + cj) * DP + part * 4);
                                const float4 cc = *reinterpret_cast<const float4 *>(&L.ctr32[part * 4]);
                                v = make_float4(w.x - cc.x, w.y - cc.y, w.z - cc.z, w.w - cc.w);
                            }
                            *reinterpret_cast<float4 *>(&L.x[cj][part * 4]) = v;
                            float nn = v.x * v.x;
                            nn = fmaf(v.y, v.y, nn);
                            nn = fmaf(v.z, v.z, nn);
                            nn = fmaf(v.w, v.w, nn);
#pragma unroll
                            for (int off = DP / 8; off > 0; off >>= 1) nn += __shfl_xor(nn, off, 64);
                            if (part == 0) L.nc[cj] = cj < nv ? nn : 1e30f;
                        }
                        if (lane < 32) {
                            L.id[lane] = lane < nv ? (int32_t)a.origId[cb + lane] : -1;
                            L.cnt[lane] = 0u;
                        }
                        wave_lds_sync();
                    } else {
                        const double *src = P + (uint64_t)cb * d;
                        uint32_t cc = c0l, jj = j0l;
                        for (uint32_t t = lane; t < nv * (uint32_t)d; t += 64) {
                            L.x[cc][jj] = (float)(src[t] - L.ctr[jj]);
                            cc += cstep;
                            jj += jstep;
                            if (jj >= (uint32_t)d) {
                                jj -= (uint32_t)d;
                                cc++;
                            }
                        }
                        if (lane < 32) {
                            L.id[lane] = lane < nv ? (int32_t)a.origId[cb + lane] : -1;
                            L.cnt[lane] = 0u;
                        }
                        wave_lds_sync();
                        if (lane < 32) {
                            float nn = 0.0f;
#pragma unroll
                            for (int t = 0; t < DP; t++) nn = fmaf(L.x[lane][t], L.x[lane][t], nn);
                            L.nc[lane] = lane < nv ? nn : 1e30f;
                        }
                        wave_lds_sync();
                    }
                    // dot products: operand B (candidate jl, dims [g*HD, (g+1)*HD))
                    const float ncv = L.nc[jl];
                    f32x16 acc0 = {}, acc1 = {};
                    if constexpr (HD % 4 == 0) {
#pragma unroll
                        for (int t = 0; t < HD; t += 4) {
                            const float4 w = *reinterpret_cast<const float4 *>(&L.x[jl][g * HD + t]);
                            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[t], w.x, acc0, 0, 0, 0);
                            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[t], w.x, acc1, 0, 0, 0);
                            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[t + 1], w.y, acc0, 0, 0, 0);
                            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[t + 1], w.y, acc1, 0, 0, 0);
                            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[t + 2], w.z, acc0, 0, 0, 0);
                            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[t + 2], w.z, acc1, 0, 0, 0);
                            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[t + 3], w.w, acc0, 0, 0, 0);
                            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[t + 3], w.w, acc1, 0, 0, 0);
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < HD; t++) {
                            const float w = L.x[jl][g * HD + t];
                            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[t], w, acc0, 0, 0, 0);
                            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[t], w, acc1, 0, 0, 0);
                        }
                    }
                    // decisions as ballots, regrouped so that lane q holds
                    // the masks of query q over the 32 candidates: ballot
                    // (h, r) carries queries q0 (low word) and q0 + 4 (high
                    // word) with q0 = 32h + 8(r/4) + r%4, whose bit 2 is 0
                    uint32_t myM = 0, myS = 0, myL = 0;
                    bool anyM = false;
                    const unsigned laneQ = lane & ~4u;
                    const bool laneHi = (lane & 4u) != 0u;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
#pragma unroll
                        for (int r = 0; r < 16; r++) {
                            const int i = 8 * (r / 4) + 4 * (int)g + (r % 4);
                            const float sq = L.nq[32 * h + i] + ncv;
                            const float dot = h == 0 ? acc0[r] : acc1[r];
                            const float d2 = fmaf(-2.0f, dot, sq);
                            const float M = fmaf(a.mAlpha, sq, a.mBeta);
                            const uint64_t bm = __ballot(d2 - M <= a.mThi);
                            const uint64_t bs = __ballot(d2 + M <= a.mTlo);
                            const bool mine = laneQ == (unsigned)(32 * h + 8 * (r / 4) + (r % 4));
                            myM = mine ? (laneHi ? (uint32_t)(bm >> 32) : (uint32_t)bm) : myM;
                            myS = mine ? (laneHi ? (uint32_t)(bs >> 32) : (uint32_t)bs) : myS;
                            if (PASS == PASS_COUNT) {
                                const uint64_t bl = __ballot(d2 <= a.mT32);
                                myL = mine ? (laneHi ? (uint32_t)(bl >> 32) : (uint32_t)bl) : myL;
                            }
                            anyM = anyM || bm != 0ull;
                        }
                    }
                    uint32_t hitRec = 0;  // the chunk's final hits, for the record
                    if (anyM) {
                        // claimed candidates of this lane's query in the chunk
                        uint32_t claim = 0;
                        if (valid) {
                            if (useL) claim |= (uint32_t)range_bits(cb, e0, e1);
                            if (useC) claim |= (uint32_t)range_bits(cb, e1, e2);
                            if (useR) claim |= (uint32_t)range_bits(cb, e2, e3);
                        }
                        uint32_t hit = myS & claim;
                        const uint32_t bnd = myM & ~myS & claim;
                        if (__ballot(bnd != 0u)) {
                            // canonical FP64 predicate for the band
                            uint32_t exm = 0, b = bnd;
                            const double *pq = P + (uint64_t)s * d;
                            while (b) {
                                const int bi = __builtin_ctz(b);
                                b &= b - 1u;
                                if (d2_fp64<0>(pq, P + (uint64_t)(cb + bi) * d, d) <= a.T) exm |= 1u << bi;
                            }
                            hit |= exm;
                            if (PASS == PASS_COUNT) {
                                band += (uint32_t)__popc(bnd);
                                mis += (uint32_t)__popc((myL & bnd) ^ exm);
                            }
                        }
                        uint32_t mirBits = 0;
                        if (anyMirror) {
                            uint32_t elig = (uint32_t)range_bits(cb, e0, e1) | (uint32_t)range_bits(cb, e2, e3);
                            if (mirC) elig |= (uint32_t)range_bits(cb, e1, e2);
                            mirBits = hit & elig & (uint32_t)range_bits(cb, a.wBegin, a.wEnd);
                        }
                        emit_chunk32<PASS>(a, cb, nv, hit, mirBits, ownW, myId, ownPos, ownLim, own, L.id, L.cnt,
                                           L.base, anyMirror, err);
                        hitRec = hit;
                    }
                    if (PASS == PASS_COUNT && recOn) record_piece(a, R, rec, dbase, dend, piece_it, (uint64_t)hitRec);
                    piece_it++;
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
    st.band += band;
    st.mis += mis;
    st.err += err;
}

template <int DP>
__device__ __forceinline__ void mfma_lds_clear(MfmaLds<DP> &L) {
    // pad dims of the staged rows are never written again: zero them once
    float *xs = &L.x[0][0];
    for (int t = (int)__lane_id(); t < 32 * (DP + 4); t += 64) xs[t] = 0.0f;
    wave_lds_sync();
}

}  // namespace sj
