// FP32 norm-expansion traversal on the matrix cores, for d >= 16 (the
// precision form of the join, reference predicate kernel.cu:341-349).
//
// A wave takes one row-aligned tile of up to 64 query points.  Queries and
// candidates are centred on the tile's first point; for every claimed
// neighbour row the union of the tile's candidate ranges is staged in LDS,
// 32 candidates at a time, and v_mfma_f32_32x32x2f32 forms the 64 x 32 dot
// products.  With s = |q|^2 + |c|^2 the squared distance is
// d2 = s - 2 q.c, and a per-pair error bound M = alpha*s + beta (derived in
// mfma_consts) splits the pairs:
//   d2 + M <= tlo  -> hit, certain;
//   d2 - M >  thi  -> miss, certain;
//   otherwise      -> decided by the canonical FP64 predicate.
// The pair set is therefore the FP64 one whatever the rounding of the
// matrix cores.  Hits are rare at these dimensions, so claim masks, the
// FP64 recheck and the emission run only on non-zero hit masks, which
// come out of the compares as wave ballots.
//
// Output layout of v_mfma_f32_32x32x2f32 used here: register r of lane l
// holds C[i][j] with j = l % 32 and i = 8*(r/4) + 4*(l/32) + r%4; operand A
// element (i, kk) sits in lane i + 32*kk, operand B element (kk, j) in lane
// j + 32*kk.  Lanes 0-31 carry dims [0, DP/2) and lanes 32-63 dims
// [DP/2, DP) of the padded rows, one dim pair per instruction.
#include <cmath>
#include <cstring>

#include "sj_traverse.hpp"

namespace sj {

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t range_bits32(uint32_t start, uint32_t A, uint32_t B) {
    return (uint32_t)range_bits(start, A, B);  // start..start+31 only: low word
}

template <int DP>
struct MfmaLds {
    float x[32][DP + 4];  // staged candidates, centred; +4 floats keep b128 reads conflict-free
    float nc[32];         // their squared norms (1e30 for empty slots)
    int32_t id[32];       // their original IDs
    float ctr[DP];        // tile centre
    float nq[64];         // squared norms of the tile's queries
    uint64_t mk[32][3];   // non-empty (maybe, sure, FP32) masks of the 32 output registers
};

template <int DP, int PASS>
__device__ __forceinline__ void process_tile_mfma(const JoinArgs &a, const MfmaConsts &mc, uint32_t tile,
                                                  MfmaLds<DP> &L, LaneStats &st) {
    constexpr int HD = DP / 2;  // dims per lane half
    const unsigned lane = __lane_id();
    const unsigned jl = lane & 31u, g = lane >> 5;
    const int k = a.k;
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

    // tile centre: FP32 row of the first point
    const float *P = a.pts32;
    if (lane < DP / 4)
        *reinterpret_cast<float4 *>(&L.ctr[lane * 4]) =
            *reinterpret_cast<const float4 *>(P + (uint64_t)ts0 * DP + lane * 4);
    wave_lds_sync();

    // operand A: queries 32h + jl, dims [g*HD, (g+1)*HD), centred
    float A0[HD], A1[HD];
    float nq0 = 0.0f, nq1 = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint64_t q = (uint64_t)ts0 + 32u * h + jl;
        const bool qv = q < ts1;
        float *A = h == 0 ? A0 : A1;
        float nq = 0.0f;
#pragma unroll
        for (int t = 0; t < HD; t += 4) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (qv) {
                const float4 w = *reinterpret_cast<const float4 *>(P + q * DP + g * HD + t);
                const float4 cc = *reinterpret_cast<const float4 *>(&L.ctr[g * HD + t]);
                v = make_float4(w.x - cc.x, w.y - cc.y, w.z - cc.z, w.w - cc.w);
            }
            A[t] = v.x;
            A[t + 1] = v.y;
            A[t + 2] = v.z;
            A[t + 3] = v.w;
            nq = fmaf(v.x, v.x, nq);
            nq = fmaf(v.y, v.y, nq);
            nq = fmaf(v.z, v.z, nq);
            nq = fmaf(v.w, v.w, nq);
        }
        nq += __shfl_xor(nq, 32, 64);
        if (h == 0) nq0 = nq; else nq1 = nq;
    }
    // query norms by row, read back in the output layout (register r of
    // this lane is row i = 8*(r/4) + 4*g + r%4 of each 32-row half)
    if (g == 0) {
        L.nq[jl] = nq0;
        L.nq[32 + jl] = nq1;
    }

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
                const uint32_t useBits = (useL ? 1u : 0u) | (useC ? 2u : 0u) | (useR ? 4u : 0u);
                const bool mirC = anyMirror && !originRow;
                for (uint32_t cb = u0; cb < u1; cb += 32) {
                    const uint32_t nv = min(32u, u1 - cb);
                    // stage 32 candidates: float4 f of the chunk goes to lane f % 64
#pragma unroll
                    for (int mm = 0; mm < DP / 8; mm++) {
                        const uint32_t f = lane + 64u * mm;
                        const uint32_t cj = f / (DP / 4), part = f % (DP / 4);
                        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (cj < nv) {
                            const float4 w =
                                *reinterpret_cast<const float4 *>(P + (uint64_t)(cb + cj) * DP + part * 4);
                            const float4 cc = *reinterpret_cast<const float4 *>(&L.ctr[part * 4]);
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
                    if (lane < 32) L.id[lane] = lane < nv ? (int32_t)a.origId[cb + lane] : -1;
                    wave_lds_sync();
                    // operand B (candidate jl, dims [g*HD, (g+1)*HD)) streamed
                    // from LDS four dims at a time
                    const float ncv = L.nc[jl];
                    f32x16 acc0 = {}, acc1 = {};
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
                    // phase 1: ballots of the 32 output registers; non-empty
                    // masks are parked in LDS, one summary bit each
                    uint32_t summary = 0;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
#pragma unroll
                        for (int r = 0; r < 16; r++) {
                            const float sq = L.nq[32 * h + 8 * (r / 4) + 4 * (int)g + (r % 4)] + ncv;
                            const float dot = h == 0 ? acc0[r] : acc1[r];
                            const float d2 = fmaf(-2.0f, dot, sq);
                            const float M = fmaf(mc.alpha, sq, mc.beta);
                            const uint64_t maybe = __ballot(d2 - M <= mc.thi);
                            const uint64_t sure = __ballot(d2 + M <= mc.tlo);
                            const uint64_t low = __ballot(d2 <= mc.t32);
                            if (maybe) {
                                const int idx = 16 * h + r;
                                if (lane == 0) {
                                    L.mk[idx][0] = maybe;
                                    L.mk[idx][1] = sure;
                                    L.mk[idx][2] = low;
                                }
                                summary |= 1u << idx;
                            }
                        }
                    }
                    if (summary) wave_lds_sync();
                    // phase 2: claims, FP64 recheck of the band, emission
                    while (summary) {
                        const int idx = __builtin_ctz(summary);
                        summary &= summary - 1u;
                        const int h = idx >> 4, r = idx & 15;
                        const uint64_t maybe = __builtin_amdgcn_readfirstlane((int)(L.mk[idx][0] & 0xffffffffu)) |
                                               ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(L.mk[idx][0] >> 32)) << 32);
                        const uint64_t sure = (uint32_t)__builtin_amdgcn_readfirstlane((int)(L.mk[idx][1] & 0xffffffffu)) |
                                              ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(L.mk[idx][1] >> 32)) << 32);
                        const uint64_t low = (uint32_t)__builtin_amdgcn_readfirstlane((int)(L.mk[idx][2] & 0xffffffffu)) |
                                             ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(L.mk[idx][2] >> 32)) << 32);
                        for (int gg = 0; gg < 2; gg++) {
                            uint32_t bm = (uint32_t)(maybe >> (32 * gg));
                            if (!bm) continue;
                            const int qi = 32 * h + 8 * (r >> 2) + 4 * gg + (r & 3);
                            if (!((vmask >> qi) & 1ull)) continue;
                            uint32_t hit = (uint32_t)(sure >> (32 * gg));
                            const uint32_t q0 = readlane_u32(e0, qi), q1 = readlane_u32(e1, qi);
                            const uint32_t q2 = readlane_u32(e2, qi), q3 = readlane_u32(e3, qi);
                            const uint32_t ub = readlane_u32(useBits, qi);
                            uint32_t claim = 0;
                            if (ub & 1u) claim |= range_bits32(cb, q0, q1);
                            if (ub & 2u) claim |= range_bits32(cb, q1, q2);
                            if (ub & 4u) claim |= range_bits32(cb, q2, q3);
                            bm &= claim;
                            hit &= claim;
                            const uint32_t bnd = bm & ~hit;
                            if (bnd) {
                                // canonical FP64 predicate for the band
                                const bool mine = g == (unsigned)gg && ((bnd >> jl) & 1u);
                                bool ex = false;
                                if (mine) {
                                    const double *pq = a.pts + (uint64_t)(ts0 + qi) * a.d;
                                    const double *pc = a.pts + (uint64_t)(cb + jl) * a.d;
                                    ex = d2_fp64<0>(pq, pc, a.d) <= a.T;
                                }
                                const uint32_t exm = (uint32_t)(__ballot(mine && ex) >> (32 * gg));
                                hit |= exm;
                                if (PASS == PASS_COUNT && lane == 0) {
                                    band += (uint32_t)__popc(bnd);
                                    const uint32_t f32 = (uint32_t)(low >> (32 * gg)) & bnd;
                                    mis += (uint32_t)__popc(f32 ^ exm);
                                }
                            }
                            if (!hit) continue;
                            const uint32_t pc = (uint32_t)__popc(hit);
                            const bool qOwn = readlane_u32(ownW ? 1u : 0u, qi) != 0u;
                            if (PASS == PASS_FILL && qOwn) {
                                const uint64_t w0 = readlane_u64(ownPos + own, qi);
                                const uint64_t lim = readlane_u64(ownLim, qi);
                                if (w0 + pc <= lim) {
                                    if (lane < 32 && ((hit >> lane) & 1u))
                                        a.ids[w0 + __popc(hit & ((1u << lane) - 1u))] = L.id[lane];
                                } else if (lane == 0) {
                                    err++;
                                }
                            }
                            if (qOwn && lane == (unsigned)qi) own += pc;
                            if (anyMirror) {
                                uint32_t elig = range_bits32(cb, q0, q1) | range_bits32(cb, q2, q3);
                                if (mirC) elig |= range_bits32(cb, q1, q2);
                                const uint32_t mir = hit & elig & range_bits32(cb, a.wBegin, a.wEnd);
                                if (lane < 32 && ((mir >> lane) & 1u)) {
                                    const uint32_t t = cb + lane;
                                    if (PASS == PASS_COUNT) {
                                        atomicAdd(&a.mirCnt[t], 1u);
                                    } else {
                                        const int32_t qid = (int32_t)readlane_u32((uint32_t)myId, qi);
                                        const unsigned long long wp = atomicAdd(&a.mirCur[t], 1ull);
                                        if (wp < a.offS[t + 1] - a.outBase) a.ids[wp] = qid;
                                        else err++;
                                    }
                                }
                            }
                        }
                    }
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

template <int DP, int PASS>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(2))) void k_join_mfma(JoinArgs a,
                                                                                          MfmaConsts mc) {
    __shared__ MfmaLds<DP> lds[WPB];
    const unsigned wave = threadIdx.x >> 6;
    const unsigned lane = __lane_id();
    MfmaLds<DP> &L = lds[wave];
    const uint32_t nTiles = *a.nTilesDev;
    LaneStats st;
    if (a.tileOrder) {
        const uint32_t ns = min(gridDim.x, (uint32_t)QSLICES);
        const uint32_t sl = blockIdx.x % ns;
        const uint32_t nbl = (gridDim.x - sl + ns - 1) / ns;
        uint64_t i = (uint64_t)(blockIdx.x / ns) * WPB + wave;
        for (;;) {
            const uint64_t q = sl + i * ns;
            if (q >= nTiles) break;
            uint32_t nx = 0;
            if (lane == 0) nx = atomicAdd(a.qctr + sl * QSTRIDE, 1u);
            process_tile_mfma<DP, PASS>(a, mc, a.tileOrder[q], L, st);
            i = (uint64_t)nbl * WPB + readlane_u32(nx, 0);
        }
    } else {
        for (uint32_t tile = blockIdx.x * WPB + wave; tile < nTiles; tile += gridDim.x * WPB)
            process_tile_mfma<DP, PASS>(a, mc, tile, L, st);
    }
    flush_stats(a, st);
}

// ---------------------------------------------------------------- host
int mfma_padded_dims(int d) {
    if (d < MFMA_MIN_D || d > MFMA_MAX_D) return 0;
    return d <= 16 ? 16 : (d <= 32 ? 32 : 64);
}

// Error bound of the matrix-core squared distance against the canonical
// FP64 one, as alpha*s + beta with s = |q|^2 + |c|^2 of the centred FP32
// vectors (u = 2^-24, R = bound on |x - origin|_2 over all points):
//  * FP32 copy and centring: each vector is off by at most
//    u*(1.0001 R + |v|), so the distance by E <= u*(2.0002 R + sqrt(2s));
//    with d <= sqrt(2s) and sqrt(2s) <= (2s+1)/2 this moves d^2 by at most
//    u*(4 R s + 2 R + 4.004 s) + E^2, E^2 <= u^2*(8 R^2 + 8.02 s);
//  * norms, dot product (DP terms, any order) and s - 2 q.c in FP32:
//    (2*gamma + 3u)*s with gamma = 1.01*DP*u;
//  * the canonical FP64 sum is within rho = 1.01*(d+2)*2^-53 of the exact
//    value relatively: it enters through tlo = T/(1+rho) and
//    thi = T/(1-rho), each widened by 4u for the FP32 compare;
//  * 5% on alpha and beta covers their own FP32 evaluation and s's
//    rounding; DP*1e-30 covers denormal flushing in the matrix core.
MfmaConsts mfma_consts(double T, double Rcoord, int d, int DP, bool *usable) {
    const double u = std::ldexp(1.0, -24);
    const double R = Rcoord * std::sqrt((double)d) * (1.0 + std::ldexp(1.0, -20));
    const double gam = 1.01 * DP * u;
    double alpha = (2.0 * gam + 3.0 * u) + u * (4.0001 * R + 4.004) + u * u * 8.02;
    double beta = u * 2.0001 * R + u * u * 8.0 * R * R;
    alpha *= 1.05;
    beta = beta * 1.05 + DP * 1e-30;
    const double rho = 1.01 * (d + 2) * std::ldexp(1.0, -53);
    double tlo = T / (1.0 + rho) * (1.0 - 4.0 * u);
    double thi = T / (1.0 - rho) * (1.0 + 4.0 * u);
    MfmaConsts m;
    m.alpha = (float)alpha;
    if ((double)m.alpha < alpha) m.alpha = std::nextafter(m.alpha, INFINITY);
    m.beta = (float)beta;
    if ((double)m.beta < beta) m.beta = std::nextafter(m.beta, INFINITY);
    m.tlo = (float)tlo;
    if ((double)m.tlo > tlo) m.tlo = std::nextafter(m.tlo, -INFINITY);
    m.thi = (float)thi;
    if ((double)m.thi < thi) m.thi = std::nextafter(m.thi, INFINITY);
    m.t32 = (float)T;
    m.dp = DP;
    *usable = std::isfinite(R) && R < 1e15 && std::isfinite(thi) && thi < 1e30 && T > 0.0 &&
              std::isfinite(m.tlo) && std::isfinite(m.thi);
    return m;
}

using MfmaKernel = void (*)(JoinArgs, MfmaConsts);
static MfmaKernel pick_mfma(int DP, int pass) {
    switch (DP) {
    case 16: return pass == PASS_COUNT ? k_join_mfma<16, PASS_COUNT> : k_join_mfma<16, PASS_FILL>;
    case 32: return pass == PASS_COUNT ? k_join_mfma<32, PASS_COUNT> : k_join_mfma<32, PASS_FILL>;
    case 64: return pass == PASS_COUNT ? k_join_mfma<64, PASS_COUNT> : k_join_mfma<64, PASS_FILL>;
    }
    return nullptr;
}

sj_status launch_join_mfma(const JoinArgs &a, const MfmaConsts &mc, int pass, int cu_count, hipStream_t s) {
    MfmaKernel kfn = pick_mfma(mc.dp, pass);
    if (!kfn) {
        set_error("no matrix-core kernel for padded dims %d", mc.dp);
        return SJ_ERR_UNSUPPORTED;
    }
    const uint64_t len = (uint64_t)a.qEnd - a.qBegin;
    uint64_t blocks = std::min<uint64_t>((uint64_t)cu_count * 8, (len + JB - 1) / JB * 4);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(kfn, dim3((unsigned)blocks), dim3(JB), 0, s, a, mc);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

}  // namespace sj
