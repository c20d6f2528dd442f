// Cell-pair traversal of the grid join (reference kernel.cu:949-1538,
// 2204-2531) producing the neighbour table, plus workload ordering
// (kernel.cu:33-250, GPU.cu:1660-1734), the atomic work queue
// (kernel.cu:2202-2236) and the batch estimator (kernel.cu:602-942).
//
// Work unit: a tile of 64 consecutive points in cell order, one wave.
// When the tile's cells lie in one cell row (same coordinates in dims
// 1..k-1), every neighbour row of the stencil is a contiguous range of
// sorted positions: the wave stages that range in LDS once and each lane
// walks its own three-cell sub-range (x-1, x, x+1) from LDS.  Tiles that
// straddle a row fall back to a per-lane walk over neighbour cells.
//
// Two passes over the same traversal:
//   count: hits each query point owns (own row) and, for the Unicomp
//          forms, mirrored hits given to the candidate (LDS counters per
//          staged candidate, one global atomic per candidate and chunk);
//   fill:  own hits leave one source lane at a time: the lane's hit mask is
//          broadcast and each thread stores one hit, so a lane's hits form
//          one contiguous row write; for mirrored hits each lane keeps a
//          bitmask over the staged chunk, the wave reserves one range per
//          candidate with one atomic, and the hits are placed through an
//          LDS cursor per candidate.
// Rows are laid out in cell order, so a tile writes one contiguous block.
// Exact predicate: acc = sum fma((x_q - x_p)^2) in FP64, dimension order;
// hit <=> acc <= T with T = max{x : sqrt_rn(x) <= eps}.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "sj_common.hpp"
#include "sj_mfma.hpp"
#include "sj_traverse.hpp"
#include "sj_wide.hpp"

namespace sj {

template <int D>
__device__ __forceinline__ float d2_fp32(const float *p, const float *q, int d) {
    float acc = 0.0f;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            const float df = q[j] - p[j];
            acc = j == 0 ? df * df : fmaf(df, df, acc);
        }
    } else {
        for (int j = 0; j < d; j++) {
            const float df = q[j] - p[j];
            acc = j == 0 ? df * df : fmaf(df, df, acc);
        }
    }
    return acc;
}

// Decision for one candidate.  FP32 mode: outside the margin band the FP32
// value decides, inside it the FP64 predicate does.
template <int D, int PASS, int PREC>
__device__ __forceinline__ bool decide(const JoinArgs &a, const double *p, const float *p32,
                                       const float *q32, uint32_t t, uint64_t &band, uint64_t &mis) {
    const int d = D > 0 ? D : a.d;
    if constexpr (PREC == 1) {
        const float s32 = d2_fp32<D>(p32, q32, d);
        if (s32 < a.lo32) return true;
        if (s32 > a.hi32) return false;
        const bool hit = d2_fp64<D>(p, a.pts + (uint64_t)t * d, d) <= a.T;
        if (PASS == PASS_COUNT) {
            band++;
            mis += (hit != (s32 <= a.T32));
        }
        return hit;
    } else {
        return d2_fp64<D>(p, a.pts + (uint64_t)t * d, d) <= a.T;
    }
}

// ------------------------------------------------------- per-lane fallback
// One query point, neighbour cells one at a time straight from memory.
template <int K, int D, int PASS, int PREC, bool DN = false>
__device__ __forceinline__ void process_point(const JoinArgs &a, uint32_t s, LaneStats &st) {
    constexpr int DR = D > 0 ? D : 1;
    const int d = D > 0 ? D : a.d;
    const unsigned lane = __lane_id();
    const uint32_t cellA = a.cellOf[s];
    uint32_t c[K];
#pragma unroll
    for (int j = 0; j < K; j++) c[j] = a.cellCoord[(uint64_t)cellA * K + j];
    const uint64_t linA = a.cellLinear[cellA];
    double pr[DR];
    float p32r[DR];
    const double *p = a.pts + (uint64_t)s * d;
    const float *p32 = PREC == 1 ? a.pts32 + (uint64_t)s * d : nullptr;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; j++) pr[j] = p[j];
        p = pr;
        if constexpr (PREC == 1) {
#pragma unroll
            for (int j = 0; j < D; j++) p32r[j] = p32[j];
            p32 = p32r;
        }
    }
    const bool ownW = s >= a.wBegin && s < a.wEnd;
    const int32_t myId = (int32_t)a.origId[s];
    uint32_t own = 0;
    uint64_t ownPos = 0, ownLim = 0;
    if (PASS == PASS_FILL && ownW) {
        ownPos = a.offS[s] - a.outBase;
        ownLim = ownPos + a.ownCnt[s];
    }
    uint64_t cand = 0, band = 0, mis = 0, err = 0;
    const bool anyMirror = a.form != SJ_FORM_FULL;

    auto visit = [&](uint32_t b, bool mirror) __attribute__((always_inline)) {
        const uint32_t t0 = a.cellStart[b], t1 = a.cellStart[b + 1];
        if (PASS == PASS_COUNT) cand += t1 - t0;
        for (uint32_t t = t0; t < t1; ++t) {
            const float *q32 = PREC == 1 ? a.pts32 + (uint64_t)t * d : nullptr;
            const bool hit = decide<D, PASS, PREC>(a, p, p32, q32, t, band, mis);
            if (hit && ownW) {
                if (PASS == PASS_FILL) {
                    const uint64_t w = ownPos + own;
                    if (w < ownLim) a.ids[w] = (int32_t)a.origId[t];
                    else err++;
                }
                own++;
            }
            if (mirror) {
                const bool mh = hit && t >= a.wBegin && t < a.wEnd;
                const uint64_t m = __ballot(mh);
                if (m) {
                    const int first = __builtin_ctzll(m);
                    const uint32_t tl = __shfl(t, first, 64);
                    const uint64_t same = __ballot(mh && t == tl);
                    if (PASS == PASS_COUNT) {
                        if ((int)lane == first) atomicAdd(&a.mirCnt[tl], (uint32_t)__popcll(same));
                        if (mh && t != tl) atomicAdd(&a.mirCnt[t], 1u);
                    } else {
                        unsigned long long base = 0;
                        if ((int)lane == first)
                            base = atomicAdd(&a.mirCur[tl], (unsigned long long)__popcll(same));
                        base = __shfl(base, first, 64);
                        if (mh) {
                            uint64_t w;
                            if (t == tl) w = base + popc_below64(same);
                            else w = atomicAdd(&a.mirCur[t], 1ull);
                            if (w < a.offS[t + 1] - a.outBase) a.ids[w] = myId;
                            else err++;
                        }
                    }
                }
            }
        }
    };

    visit(cellA, false);
    int o[K];
#pragma unroll
    for (int j = 0; j < K; j++) o[j] = -1;
    int noff = 1;
#pragma unroll
    for (int j = 0; j < K; j++) noff *= 3;
    for (int it = 0; it < noff; ++it) {
        bool zero = true, inb = true;
        int64_t dl = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            zero = zero && o[j] == 0;
            const int64_t v = (int64_t)c[j] + o[j];
            inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
            dl += (int64_t)o[j] * a.mult[j];
        }
        if (!zero && inb) {
            const uint64_t linB = (uint64_t)((int64_t)linA + dl);
            if (claims<K>(a.form, c, o, linA, linB)) {
                const int64_t b = find_cell_t<DN>(a, linB);
                if (b >= 0) visit((uint32_t)b, anyMirror);
            }
        }
        bool carry = true;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (carry) {
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
        if (ownW) a.ownCnt[s] = own;
    } else {
        if (ownW && own != (ownLim - ownPos)) err++;
    }
    st.cand += cand;
    st.band += band;
    st.mis += mis;
    st.err += err;
}

// ------------------------------------------------------------ tile path
template <int D, int PREC>
struct TileCfg {
    static constexpr int DD = D > 0 ? D : 1;
    static constexpr int CH = D <= 4 ? 256 : 128;  // staged candidates per chunk
    static constexpr int MW = CH / 64;                              // mask words per lane (<= 4)
};

// Staged rows are padded by 64: a lane reads the 64 rows after its piece
// start, masking the rows beyond its span.
template <int D, int PREC>
struct TileLds {
    using C = TileCfg<D, PREC>;
    double x64[PREC == 1 ? 1 : C::CH + 64][C::DD];
    float x32[PREC == 1 ? C::CH + 64 : 1][C::DD];
    int32_t id[C::CH];
    uint32_t cnt[C::CH];
    unsigned long long base[C::CH];
};

// Own hits of one piece staged in output order by the replaying fill (at
// most OWN_STAGE), then stored 64 consecutive entries at a time.
constexpr uint32_t OWN_STAGE = 384;

// LDS of a count pass: the staged coordinates and the mirrored-hit
// counters (the candidate IDs and reservation slots belong to the fills).
template <int D, int PREC>
struct TileLdsCount {
    using C = TileCfg<D, PREC>;
    double x64[PREC == 1 ? 1 : C::CH + 64][C::DD];
    float x32[PREC == 1 ? C::CH + 64 : 1][C::DD];
    int32_t id[1];
    uint32_t cnt[C::CH];
    unsigned long long base[1];
};

// LDS of the fill that replays a hit record: candidate IDs and mirrored-hit
// counters, chunked like the count pass of precision PREC, and the own-hit
// stage: entry i is neighbour ID sid[i] for position wc[l] + t of lane l's
// row, with sdst[i] = l * 64 + t.
template <int D, int PREC>
struct TileLdsIds {
    using C = TileCfg<D, PREC>;
    int32_t id[C::CH];
    uint32_t cnt[C::CH];
    unsigned long long base[C::CH];
    unsigned long long wc[64];
    int32_t sid[OWN_STAGE];
    uint32_t sdst[OWN_STAGE];
};

// Upper bound of the pieces a tile walk executes: per neighbour row in the
// grid, ceil(staged candidates / cw).  LDS tile walk: cw = 64 (chunks are
// whole multiples of 64 candidates, a chunk runs at most ceil(chunk / 64)
// pieces); matrix-core walk: cw = 32, one piece per 32-candidate chunk.
template <int K>
__device__ __forceinline__ uint32_t tile_piece_bound(const JoinArgs &a, const uint32_t *c, uint64_t rowBase,
                                                     uint32_t xa, uint32_t xb, uint32_t cw) {
    const uint32_t nC0 = a.nCells[0];
    const uint32_t ux0 = xa >= 1 ? xa - 1 : 0u;
    const uint32_t ux1 = xb + 1 < nC0 ? xb + 1 : nC0 - 1;
    int o[K];
#pragma unroll
    for (int j = 0; j < K; j++) o[j] = -1;
    int nrow = 1;
#pragma unroll
    for (int j = 1; j < K; j++) nrow *= 3;
    uint32_t bound = 0;
    for (int it = 0; it < nrow; ++it) {
        bool inb = true;
        int64_t dl = 0;
#pragma unroll
        for (int j = 1; j < K; j++) {
            const int64_t v = (int64_t)c[j] + o[j];
            inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
            dl += (int64_t)o[j] * a.mult[j];
        }
        if (inb) {
            const uint64_t rb = (uint64_t)((int64_t)rowBase + dl);
            bound += (pos_lb(a, rb + ux1 + 1) - pos_lb(a, rb + ux0) + cw - 1u) / cw;
        }
        bool carry = true;
#pragma unroll
        for (int j = 1; j < K; j++) {
            if (carry) {
                if (o[j] == 1) {
                    o[j] = -1;
                } else {
                    o[j]++;
                    carry = false;
                }
            }
        }
    }
    return bound;
}

// One tile of <= 64 points of one cell row against its 3^(k-1) neighbour
// rows.  RP (fill only): take the hit masks from the count pass's record
// instead of evaluating distances; L is then a TileLdsIds.
template <int K, int D, int PASS, int PREC, bool RP, class LDS, bool DN = false>
__device__ __forceinline__ void process_tile(const JoinArgs &a, uint32_t tile, LDS &L, LaneStats &st,
                                             TileRec &R) {
    using Cfg = TileCfg<D, PREC>;
    constexpr int CH = Cfg::CH;
    constexpr int DD = Cfg::DD;
    static_assert(!RP || PASS == PASS_FILL, "only the fill replays a record");
    const unsigned lane = __lane_id();
    const uint32_t ts0 = a.tileStart[tile], ts1 = a.tileStart[tile + 1];
    const uint64_t s64 = (uint64_t)ts0 + lane;
    const bool valid = s64 < ts1;
    const uint32_t s = valid ? (uint32_t)s64 : 0u;
    const uint64_t vmask = __ballot(valid);
    if (!vmask) return;
    const int firstL = __builtin_ctzll(vmask);
    const int lastL = 63 - __builtin_clzll(vmask);

    // lanes past the end take the last point's cell so that row-level
    // control flow stays uniform across the wave; they never emit
    const uint32_t sEff = valid ? s : ts0 + (uint32_t)lastL;
    uint32_t c[K];
    const uint32_t cellA = a.cellOf[sEff];
#pragma unroll
    for (int j = 0; j < K; j++) c[j] = a.cellCoord[(uint64_t)cellA * K + j];
    const uint64_t linA = a.cellLinear[cellA];
    const uint64_t rowBase = linA - c[0];
    const uint64_t rb0 = __shfl(rowBase, firstL, 64);
    if (__ballot(valid && rowBase == rb0) != vmask) {
        // the tile crosses a cell row: per-lane walk (FP64: the per-lane
        // FP32 walk needs the global FP32 copy, which tiles do not make)
        // a record cannot describe it: the count marks the record unusable
        if (valid) process_point<K, D, PASS, 0, DN>(a, s, st);
        if (PASS == PASS_COUNT && lane == (unsigned)firstL) {
            st.fallback++;
            if (a.recDir) atomicAdd(&a.recCtr[REC_BAD], 1ull);
        }
        return;
    }
    const uint32_t xa = __shfl(c[0], firstL, 64);
    const uint32_t xb = __shfl(c[0], lastL, 64);
    const uint32_t nC0 = a.nCells[0];
    const uint32_t x = c[0];

    // query point in registers (FP32 form: the FP32 copy only; the FP64
    // coordinates are read again for the rare band recheck)
    double p[PREC == 1 ? 1 : DD];
    float p32[DD];
#pragma unroll
    for (int j = 0; j < DD; j++) {
        if constexpr (PREC != 1) p[j] = 0.0;
        p32[j] = 0.0f;
    }
    // FP32 form: coordinates centred on the tile's first point in FP64 and
    // rounded once; every staged coordinate is within R of that centre
    // (R from the tile's span of cells in dim 0), so the band is set per tile
    double ctr[DD];
    float lo32 = 0.0f, hi32 = 0.0f, t32 = 0.0f;
    if constexpr (PREC == 1 && !RP) {
#pragma unroll
        for (int j = 0; j < DD; j++) ctr[j] = uniform_f64(a.pts[(uint64_t)ts0 * DD + j]);
        const double Rt = (double)max(xb - xa + 3u, 4u) * a.cellW * (1.0 + ldexp(1.0, -20));
        const Fp32Margin m = fp32_margin(a.T, Rt, D);
        lo32 = m.usable ? m.lo : -INFINITY;
        hi32 = m.usable ? m.hi : INFINITY;
        t32 = m.t32;
    }
    if (valid && !RP) {
        if constexpr (PREC == 1) {
#pragma unroll
            for (int j = 0; j < DD; j++) p32[j] = (float)(a.pts[(uint64_t)s * DD + j] - ctr[j]);
        } else {
#pragma unroll
            for (int j = 0; j < DD; j++) p[j] = a.pts[(uint64_t)s * DD + j];
        }
    }
    // hit record: entries [tileDir[tile], tileDir[tile+1]) hold the tile's
    // pieces (a bound of them, laid out by a scan before the count)
    const bool recOn = a.recDir != nullptr;
    bool rec = false;
    uint64_t dbase = 0, dend = 0;
    if (PASS == PASS_COUNT && recOn) {
        dbase = readlane_u64(a.tileDir[tile], 0);
        dend = readlane_u64(a.tileDir[tile + 1], 0);
        rec = dend <= a.dirCap;
        if (!rec && lane == 0) atomicAdd(&a.recCtr[REC_BAD], 1ull);
    }
    if constexpr (RP) dbase = readlane_u64(a.tileDir[tile], 0);
    uint64_t piece_it = 0;  // pieces executed so far: the directory index
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

    // union of the lanes' three-cell windows in dim 0
    const uint32_t ux0 = xa >= 1 ? xa - 1 : 0u;
    const uint32_t ux1 = xb + 1 < nC0 ? xb + 1 : nC0 - 1;
    int nrow = 1;
#pragma unroll
    for (int j = 1; j < K; j++) nrow *= 3;

    // candidates [c0, c0 + m) into LDS rows [0, m)
    auto stage = [&](uint32_t c0, uint32_t m) __attribute__((always_inline)) {
        for (uint32_t j = lane; j < m; j += 64) {
            const uint64_t t = (uint64_t)c0 + j;
            if constexpr (RP) {
            } else if constexpr (PREC == 1) {
#pragma unroll
                for (int q = 0; q < DD; q++) L.x32[j][q] = (float)(a.pts[t * DD + q] - ctr[q]);
            } else {
#pragma unroll
                for (int q = 0; q < DD; q++) L.x64[j][q] = a.pts[t * DD + q];
            }
            if (PASS == PASS_FILL) L.id[j] = (int32_t)a.origId[t];  // emission only
            L.cnt[j] = 0u;
        }
    };

    // One neighbour row at offsets o[1..K-1] (o[0] unused) with cell-row
    // base rb; the caller may pass the union's bounds [pu0, pu1) (given).
    // Every lane shares the row coordinates, so control flow is uniform.
    auto row = [&](const int (&o)[K], uint64_t rb, bool originRow, bool given, uint32_t pu0,
                   uint32_t pu1) __attribute__((always_inline)) {
        // claimed dim-0 offsets of this lane's cell in this row
        bool useL = false, useC = false, useR = false;
        // every lane claims every cell of the row: a lane's span has no
        // unclaimed gap (uniform)
        const bool allClaimed = given && !originRow && a.form != SJ_FORM_BUNICOMP;
        if (valid && allClaimed) {
            // a row the lookup kept off the origin row: under FULL, LID and
            // Unicomp every lane claims every cell of it
            useL = x >= 1;
            useC = true;
            useR = x + 1 < nC0;
        } else if (valid) {
            int oo[K];
#pragma unroll
            for (int j = 0; j < K; j++) oo[j] = o[j];
            oo[0] = -1;
            useL = x >= 1 && claims<K>(a.form, c, oo, linA, rb + x - 1);
            oo[0] = 1;
            useR = x + 1 < nC0 && claims<K>(a.form, c, oo, linA, rb + x + 1);
            oo[0] = 0;
            useC = originRow || claims<K>(a.form, c, oo, linA, rb + x);
        }
        if (__ballot(useL || useC || useR)) {
            const uint32_t u0 = given ? pu0 : pos_lb_t<DN>(a, rb + ux0);
            const uint32_t u1 = given ? pu1 : pos_lb_t<DN>(a, rb + ux1 + 1);
            uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;
            // given: the first chunk is staged while the window bounds load
            // (both requests in flight together); the lookups run on every
            // lane, clamped to the row, and past-the-end lanes drop them
            bool staged0 = false;
            if (given) {
                const uint32_t xl = x >= 1 ? x - 1 : x;
                const uint32_t xr = x + 1 < nC0 ? x + 2 : x + 1;
                const uint32_t f0 = pos_lb_t<DN>(a, rb + xl), f1 = pos_lb_t<DN>(a, rb + x);
                const uint32_t f2 = pos_lb_t<DN>(a, rb + x + 1), f3 = pos_lb_t<DN>(a, rb + xr);
                if constexpr (!RP) {
                    const uint32_t m0 = min((uint32_t)CH, u1 - u0);
                    if (!(PASS == PASS_FILL && !anyOwnW && (u0 + m0 <= a.wBegin || u0 >= a.wEnd))) {
                        stage(u0, m0);
                        staged0 = true;
                    }
                }
                e0 = valid ? f0 : 0u;
                e1 = valid ? f1 : 0u;
                e2 = valid ? f2 : 0u;
                e3 = valid ? f3 : 0u;
            } else if (valid) {
                e1 = pos_lb_t<DN>(a, rb + x);
                e2 = pos_lb_t<DN>(a, rb + x + 1);
                e0 = x >= 1 ? pos_lb_t<DN>(a, rb + x - 1) : e1;
                e3 = x + 1 < nC0 ? pos_lb_t<DN>(a, rb + x + 2) : e2;
            }
            // this lane's span: from the first to the last claimed cell;
            // unclaimed cells inside it are masked per candidate
            uint32_t lo = useL ? e0 : (useC ? e1 : e2);
            uint32_t hi = useR ? e3 : (useC ? e2 : e1);
            if (!(useL || useC || useR)) lo = hi = e0;
            if (PASS == PASS_COUNT && valid)
                cand += (useL ? e1 - e0 : 0) + (useC ? e2 - e1 : 0) + (useR ? e3 - e2 : 0);
            const bool mirC = anyMirror && !originRow;
            for (uint32_t c0 = u0; c0 < u1; c0 += CH) {
                const uint32_t m = min((uint32_t)CH, u1 - c0);
                const uint32_t clo = max(lo, c0);
                const uint32_t chi = min(hi, c0 + m);
                const uint32_t len = chi > clo ? chi - clo : 0u;
                static_assert(CH <= 256, "lane spans need at most 9 bits");
                const uint32_t maxLen = wave_max_small<CH == 256 ? 9 : 8>(len);
                // replay: the chunk's recorded masks, requested before the
                // staging so that both loads are in flight together
                uint64_t rm[Cfg::MW];
#pragma unroll
                for (int w = 0; w < Cfg::MW; w++) rm[w] = 0ull;
                const uint32_t np = (maxLen + 63u) / 64u;  // pieces of the chunk
                const uint64_t chunkIt = piece_it;
                if (PASS == PASS_FILL && !anyOwnW && (c0 + m <= a.wBegin || c0 >= a.wEnd)) {
                    piece_it += np;  // nothing of this chunk lands in the rows written
                    continue;
                }
                if constexpr (RP) {
                    uint64_t anyHit = 0ull;
#pragma unroll
                    for (int w = 0; w < Cfg::MW; w++) {
                        if ((uint32_t)w < np) {
                            const uint64_t *e = a.recDir + REC_DW * (dbase + piece_it + (uint64_t)w);
                            const uint64_t pres = e[0];
                            anyHit |= pres;
                            if ((pres >> lane) & 1ull) rm[w] = a.recMask[e[1] + popc_below64(pres)];
                        }
                    }
                    // no hits in the chunk: nothing to stage or emit
                    if (!anyHit) {
                        piece_it += np;
                        continue;
                    }
                    // mirrored hits: each candidate's range comes from the
                    // reservation the count pass recorded, packed as
                    // (row position << 7) | hits
                    const uint64_t mo = anyMirror ? a.recDir[REC_DW * (dbase + piece_it) + 2] : 0ull;
                    if (mo) {
                        const uint64_t *mr = a.recMask + (mo - 1);
                        uint32_t pre = Cfg::MW;
#pragma unroll
                        for (int w = 0; w < Cfg::MW; w++) {
                            const uint64_t hv = mr[w];
                            if ((hv >> lane) & 1ull) {
                                const uint64_t rv = mr[pre + popc_below64(hv)];
                                const uint32_t j = lane + 64u * (uint32_t)w;
                                const uint64_t row = (uint64_t)c0 + j;
                                unsigned long long bs = NO_SLOT;
                                if (row >= a.wBegin && row < a.wEnd) {
                                    const uint64_t v = rv & 255ull;
                                    const uint64_t b0 = a.offS[row] - a.outBase + a.ownCnt[row] + (rv >> 8);
                                    if (b0 + v <= a.offS[row + 1] - a.outBase) bs = (b0 << 7) | v;
                                    else err++;
                                }
                                L.base[j] = bs;
                            }
                            pre += (uint32_t)__popcll(hv);
                        }
                    }
                }
                if (!(staged0 && c0 == u0)) stage(c0, m);
                wave_lds_sync();
                // mirrored hits of the chunk, one word per 64-candidate piece
                uint64_t mk0 = 0ull, mk1 = 0ull, mk2 = 0ull, mk3 = 0ull;
                bool anyMir = false;
                uint32_t piece = 0;
                for (uint32_t base = 0; base < maxLen; base += 64, piece++, piece_it++) {
                    const uint32_t n = min(64u, maxLen - base);
                    const uint32_t pstart = clo + base;
                    uint64_t mask = 0ull, maybe = 0ull;
                    if constexpr (RP) {
#pragma unroll
                        for (int w = 0; w < Cfg::MW; w++)
                            if ((uint32_t)w == piece) mask = rm[w];
                    } else {
                    // claimed candidates inside this lane's span, as bits
                    // of the piece; lanes without work read row 0
                    uint64_t okm = 0ull;
                    uint32_t r0 = 0u;
                    if (base < len) {
                        if (allClaimed) {
                            okm = range_bits(pstart, pstart, clo + len);
                        } else {
                            if (useL) okm |= range_bits(pstart, e0, e1);
                            if (useC) okm |= range_bits(pstart, e1, e2);
                            if (useR) okm |= range_bits(pstart, e2, e3);
                            okm &= range_bits(pstart, pstart, clo + len);
                        }
                        r0 = pstart - c0;
                    }
                    // uniform trip count, US candidates per step, one bit
                    // per candidate; masks are applied per piece
                    constexpr int US = D <= 3 ? 8 : 4;
                    for (uint32_t i = 0; i < n; i += US) {
                        uint32_t hb = 0u, mb = 0u;
#pragma unroll
                        for (int u = 0; u < US; u++) {
                            const uint32_t row = r0 + i + (uint32_t)u;
                            if constexpr (PREC == 1) {
                                float q[DD];
#pragma unroll
                                for (int z = 0; z < DD; z++) q[z] = L.x32[row][z];
                                const float s32 = d2_fp32<D>(p32, q, DD);
                                hb |= (s32 < lo32 ? 1u : 0u) << u;
                                mb |= (s32 <= hi32 ? 1u : 0u) << u;
                            } else {
                                double q[DD];
#pragma unroll
                                for (int z = 0; z < DD; z++) q[z] = L.x64[row][z];
                                hb |= (d2_fp64<D>(p, q, DD) <= a.T ? 1u : 0u) << u;
                            }
                        }
                        mask |= (uint64_t)hb << i;
                        if constexpr (PREC == 1) maybe |= (uint64_t)mb << i;
                    }
                    mask &= okm;
                    if constexpr (PREC == 1) {
                        // candidates in [T - M, T + M]: canonical FP64 predicate
                        const uint64_t bnd = maybe & ~mask & okm;
                        if (__ballot(bnd != 0ull)) {
                            uint64_t exm = 0ull, bb = bnd;
                            while (bb) {
                                const int bi = __builtin_ctzll(bb);
                                bb &= bb - 1ull;
                                const uint64_t t = (uint64_t)pstart + (uint32_t)bi;
                                const bool ex = d2_fp64<D>(a.pts + (uint64_t)s * DD, a.pts + t * DD, DD) <= a.T;
                                if (ex) exm |= 1ull << bi;
                                if (PASS == PASS_COUNT) {
                                    // the plain FP32 decision, for the count of
                                    // pairs it would have misclassified
                                    float q[DD];
#pragma unroll
                                    for (int z = 0; z < DD; z++) q[z] = L.x32[r0 + (uint32_t)bi][z];
                                    mis += (d2_fp32<D>(p32, q, DD) <= t32) != ex ? 1u : 0u;
                                }
                            }
                            mask |= exm;
                            if (PASS == PASS_COUNT) band += (uint64_t)__popcll(bnd);
                        }
                    }
                    }  // distance evaluation
                    if (PASS == PASS_COUNT && recOn) record_piece(a, R, rec, dbase, dend, piece_it, mask);
                    // emission: mirrored hits (claimed non-own cells inside
                    // the rows written) and own hits of the piece
                    uint64_t mirBits = 0ull;
                    if (anyMirror) {
                        // every hit is eligible when the lane claimed all cells
                        // off the origin row
                        uint64_t elig = ~0ull;
                        if (!allClaimed) {
                            elig = range_bits(pstart, e0, e1) | range_bits(pstart, e2, e3);
                            if (mirC) elig |= range_bits(pstart, e1, e2);
                        }
                        mirBits = mask & elig & range_bits(pstart, a.wBegin, a.wEnd);
                    }
                    const uint32_t myCnt = (uint32_t)__popcll(mask);
                    uint64_t ownBits = 0ull;
                    if (PASS == PASS_FILL && ownW) {
                        // more hits than the row has room for: count and
                        // fill disagree; drop the piece and report it
                        if (ownPos + own + myCnt <= ownLim) ownBits = mask;
                        else err++;
                    }
                    const uint64_t wcur = ownPos + own;
                    const uint32_t jb = pstart - c0;  // LDS row of the piece start (lanes with bits)
                    // mirrored hits, lane by lane: the count and the
                    // evaluating fill tally them per candidate; the
                    // replayed fill places them in the recorded ranges
                    if (anyMirror) {
                        uint64_t mb = mirBits;
                        while (mb) {
                            const int bi = __builtin_ctzll(mb);
                            mb &= mb - 1ull;
                            const uint32_t jj = min(jb + (uint32_t)bi, (uint32_t)(CH - 1));
                            if constexpr (RP) {
                                const uint32_t k2 = atomicAdd(&L.cnt[jj], 1u);
                                const unsigned long long bs = L.base[jj];
                                if (bs != NO_SLOT && k2 < (uint32_t)(bs & 127ull)) a.ids[(bs >> 7) + k2] = myId;
                                else if (bs != NO_SLOT) err++;
                            } else {
                                atomicAdd(&L.cnt[jj], 1u);
                            }
                        }
                    }
                    // own hits: lane by lane when no lane has more than 24
                    // (measured best from 8 to 24 at 2-D to 6-D), else
                    // one source lane at a time, its mask broadcast so that
                    // each thread takes one candidate bit and the hits
                    // leave as one contiguous row write
                    const uint32_t cntOwn = PASS == PASS_FILL ? (uint32_t)__popcll(ownBits) : 0u;
                    const uint32_t mp = PASS == PASS_FILL ? wave_max_small<7>(cntOwn) : 0u;
                    bool staged = false;
                    if constexpr (RP) {
                        // replay: the piece's own hits (when few per lane and
                        // in total) go through the LDS stage in output order,
                        // so that each store covers a few row segments
                        // instead of one position in each of 64 rows
                        if (mp <= 24u) {
                            uint32_t pre = 0, tot = 0;
#pragma unroll
                            for (int b = 0; b < 7; b++) {
                                const uint64_t bm = __ballot((cntOwn >> b) & 1u);
                                pre += popc_below64(bm) << b;
                                tot += (uint32_t)__popcll(bm) << b;
                            }
                            if (tot <= OWN_STAGE) {
                                if (cntOwn) L.wc[lane] = wcur;
                                uint64_t ob = ownBits;
                                uint32_t q = pre, t = lane * 64u;
                                while (ob) {
                                    const int bi = __builtin_ctzll(ob);
                                    ob &= ob - 1ull;
                                    L.sid[q] = L.id[min(jb + (uint32_t)bi, (uint32_t)(CH - 1))];
                                    L.sdst[q] = t;
                                    q++;
                                    t++;
                                }
                                wave_lds_sync();
                                for (uint32_t i = lane; i < tot; i += 64) {
                                    const uint32_t pk = L.sdst[i];
                                    a.ids[L.wc[pk >> 6] + (pk & 63u)] = L.sid[i];
                                }
                                wave_lds_sync();
                                staged = true;
                            }
                        }
                    }
                    if (staged) {
                    } else if (PASS == PASS_FILL && mp <= 24u) {
                        uint64_t ob = ownBits;
                        uint64_t wp = wcur;
                        while (ob) {
                            const int bi = __builtin_ctzll(ob);
                            ob &= ob - 1ull;
                            a.ids[wp++] = L.id[min(jb + (uint32_t)bi, (uint32_t)(CH - 1))];
                        }
                    } else if (PASS == PASS_FILL) {
                        uint64_t todo = __ballot(ownBits != 0ull);
                        // four source lanes per round so that their LDS reads overlap
                        while (todo) {
                            int l[4];
                            bool lv[4];
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                lv[u] = todo != 0ull;
                                l[u] = lv[u] ? __builtin_ctzll(todo) : 0;
                                todo &= todo - 1ull;
                            }
                            uint64_t oB[4];
                            uint32_t j[4];
                            int32_t idv[4];
                            uint64_t w[4];
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                oB[u] = lv[u] ? readlane_u64(ownBits, l[u]) : 0ull;
                                j[u] = min(readlane_u32(pstart, l[u]) + lane - c0, (uint32_t)(CH - 1));
                            }
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                idv[u] = L.id[j[u]];
                                w[u] = readlane_u64(wcur, l[u]) + popc_below64(oB[u]);
                            }
#pragma unroll
                            for (int u = 0; u < 4; u++)
                                if ((oB[u] >> lane) & 1ull) a.ids[w[u]] = idv[u];
                        }
                    }
                    if (ownW) own += myCnt;
                    if (PASS == PASS_FILL && mirBits) {
                        anyMir = true;
                        if (piece == 0) mk0 = mirBits;
                        else if (piece == 1) mk1 = mirBits;
                        else if (piece == 2) mk2 = mirBits;
                        else mk3 = mirBits;
                    }
                }
                wave_lds_sync();
                if (anyMirror) {
                    if (PASS == PASS_COUNT) {
                        // reserve the chunk's mirrored hits of each candidate;
                        // a record keeps each reservation for the fill
                        uint64_t hv[Cfg::MW], rv[Cfg::MW];
                        uint32_t nzm = Cfg::MW;
#pragma unroll
                        for (int w = 0; w < Cfg::MW; w++) {
                            const uint32_t j = lane + 64u * (uint32_t)w;
                            const uint32_t v = j < m ? L.cnt[j] : 0u;
                            uint32_t b = 0;
                            if (v) b = atomicAdd(&a.mirCnt[c0 + j], v);
                            hv[w] = __ballot(v != 0u);
                            rv[w] = ((uint64_t)b << 8) | v;
                            nzm += (uint32_t)__popcll(hv[w]);
                        }
                        if (recOn && rec && np > 0 && nzm > (uint32_t)Cfg::MW) {
                            if (rec_room(a, R, nzm)) {
                                const uint64_t off = R.blk + R.used;
                                uint32_t pre = Cfg::MW;
#pragma unroll
                                for (int w = 0; w < Cfg::MW; w++) {
                                    if (lane == (unsigned)w) a.recMask[off + w] = hv[w];
                                    if ((hv[w] >> lane) & 1ull) a.recMask[off + pre + popc_below64(hv[w])] = rv[w];
                                    pre += (uint32_t)__popcll(hv[w]);
                                }
                                R.used += nzm;
                                if (lane == 0) a.recDir[REC_DW * (dbase + chunkIt) + 2] = off + 1;
                            }
                        }
                        R.want += nzm > (uint32_t)Cfg::MW ? nzm : 0u;
                    } else if constexpr (!RP) {
                        // reserve one range per candidate, then place; a
                        // reservation past the candidate's row end is an
                        // error and its hits are dropped
                        for (uint32_t j = lane; j < m; j += 64) {
                            const uint32_t v = L.cnt[j];
                            if (v) {
                                unsigned long long b = atomicAdd(&a.mirCur[c0 + j], (unsigned long long)v);
                                if (b + v > a.offS[c0 + j + 1] - a.outBase) {
                                    err++;
                                    b = NO_SLOT;
                                }
                                L.base[j] = b;
                            }
                            L.cnt[j] = 0u;
                        }
                        wave_lds_sync();
                        // place the mirrored hits, one source lane at a time
                        if (__ballot(anyMir)) {
#pragma unroll
                            for (int w = 0; w < Cfg::MW; w++) {
                                const uint64_t mb = w == 0 ? mk0 : (w == 1 ? mk1 : (w == 2 ? mk2 : mk3));
                                uint64_t todo = __ballot(mb != 0ull);
                                while (todo) {
                                    int l[4];
                                    bool lv[4];
#pragma unroll
                                    for (int u = 0; u < 4; u++) {
                                        lv[u] = todo != 0ull;
                                        l[u] = lv[u] ? __builtin_ctzll(todo) : 0;
                                        todo &= todo - 1ull;
                                    }
                                    bool hb[4];
                                    uint32_t j[4];
                                    int32_t oid[4];
#pragma unroll
                                    for (int u = 0; u < 4; u++) {
                                        const uint64_t bits = lv[u] ? readlane_u64(mb, l[u]) : 0ull;
                                        hb[u] = (bits >> lane) & 1ull;
                                        j[u] = min(readlane_u32(clo, l[u]) + (uint32_t)w * 64u + lane - c0,
                                                   (uint32_t)(CH - 1));
                                        oid[u] = (int32_t)readlane_u32((uint32_t)myId, l[u]);
                                    }
                                    uint32_t k2[4];
                                    unsigned long long bb[4];
#pragma unroll
                                    for (int u = 0; u < 4; u++) {
                                        k2[u] = hb[u] ? atomicAdd(&L.cnt[j[u]], 1u) : 0u;
                                        bb[u] = L.base[j[u]];
                                    }
#pragma unroll
                                    for (int u = 0; u < 4; u++)
                                        if (hb[u] && bb[u] != NO_SLOT) a.ids[bb[u] + k2[u]] = oid[u];
                                }
                            }
                        }
                    }
                }
                wave_lds_sync();
            }
        }
    };

    if constexpr (K < 4) {
        // neighbour rows one after another: odometer over dims 1..K-1
        int o[K];
#pragma unroll
        for (int j = 0; j < K; j++) o[j] = -1;
        o[0] = 0;
        for (int it = 0; it < nrow; ++it) {
            bool inb = true, originRow = true;
            int64_t dl = 0;
#pragma unroll
            for (int j = 1; j < K; j++) {
                originRow = originRow && o[j] == 0;
                const int64_t v = (int64_t)c[j] + o[j];
                inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
                dl += (int64_t)o[j] * a.mult[j];
            }
            if (inb) row(o, (uint64_t)((int64_t)rowBase + dl), originRow, false, 0u, 0u);
            bool carry = true;
#pragma unroll
            for (int j = 1; j < K; j++) {
                if (carry) {
                    if (o[j] == 1) {
                        o[j] = -1;
                    } else {
                        o[j]++;
                        carry = false;
                    }
                }
            }
        }
    } else {
        // 3^(K-1) neighbour rows, 64 at a time: lane l looks up row b0 + l
        // (its bounds and whether it can hold claimed candidates), then the
        // wave visits the rows that can, in the same order as the odometer
        // (o[1] fastest).  Rows left out are outside the grid, empty, or
        // claimed by no lane: LID claims a row with every lane and cell iff
        // its offset dl > 0 (|dl| >= nCells[0] off the origin row), Unicomp
        // off the origin row by the parity of c[j] of its highest non-zero
        // offset, the same for every lane.
        for (int b0 = 0; b0 < nrow; b0 += 64) {
            const int my = b0 + (int)lane;
            int od[K];
            od[0] = 0;
            bool inb = my < nrow, org = true;
            int64_t dl = 0;
            int t = my;
#pragma unroll
            for (int j = 1; j < K; j++) {
                od[j] = t % 3 - 1;
                t /= 3;
                org = org && od[j] == 0;
                const int64_t v = (int64_t)c[j] + od[j];
                inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
                dl += (int64_t)od[j] * a.mult[j];
            }
            uint32_t r0 = 0u, r1 = 0u;
            bool act = false;
            if (inb) {
                const uint64_t rbl = (uint64_t)((int64_t)rowBase + dl);
                bool may = true;
                if (a.form == SJ_FORM_LID) may = org || dl > 0;
                else if (a.form == SJ_FORM_UNICOMP && !org) may = claims<K>(a.form, c, od, linA, rbl);
                if (may) {
                    r0 = pos_lb_t<DN>(a, rbl + ux0);
                    r1 = pos_lb_t<DN>(a, rbl + ux1 + 1);
                    act = r1 > r0;
                }
            }
            // the visited rows' offsets (2 bits per dim) and bases come from
            // the lane that looked the row up
            uint32_t opk = 0;
#pragma unroll
            for (int j = 1; j < K; j++) opk |= (uint32_t)(od[j] + 1) << (2 * (j - 1));
            const uint64_t rbl = (uint64_t)((int64_t)rowBase + dl);
            uint64_t am = __ballot(act);
            while (am) {
                const int bl = __builtin_ctzll(am);
                am &= am - 1ull;
                const uint32_t ok = readlane_u32(opk, bl);
                int o[K];
                o[0] = 0;
                bool originRow = true;
#pragma unroll
                for (int j = 1; j < K; j++) {
                    o[j] = (int)((ok >> (2 * (j - 1))) & 3u) - 1;
                    originRow = originRow && o[j] == 0;
                }
                row(o, readlane_u64(rbl, bl), originRow, true, readlane_u32(r0, bl), readlane_u32(r1, bl));
            }
        }
    }
    if (PASS == PASS_COUNT) {
        if (ownW) a.ownCnt[s] = own;
    } else {
        if (ownW && own != (ownLim - ownPos)) err++;
    }
    st.cand += cand;
    st.band += band;
    st.mis += mis;
    st.err += err;
}

template <int K, int D, int PASS, int PREC, bool RP, class LDS, bool DN = false>
__device__ __forceinline__ void join_body(const JoinArgs &a, LDS *lds) {
    if (PASS == PASS_FILL && !fill_gate(a)) return;
    const unsigned wave = threadIdx.x >> 6;
    const unsigned lane = __lane_id();
    LDS &L = lds[wave];
    const uint32_t nTiles = *a.nTilesDev;
    LaneStats st;
    TileRec R;
    auto run = [&](uint32_t tile) __attribute__((always_inline)) {
        if constexpr (RP) {
            process_tile<K, D, PASS, PREC, true, LDS, DN>(a, tile, L, st, R);
        } else if (D > 0 && a.useTiles) {
            process_tile<K, D, PASS, PREC, false, LDS, DN>(a, tile, L, st, R);
        } else {
            const uint64_t s = (uint64_t)a.tileStart[tile] + lane;
            if (s < a.tileStart[tile + 1]) process_point<K, D, PASS, PREC, DN>(a, (uint32_t)s, st);
        }
    };
    if (a.tileOrder) {
        // work queue in QSLICES interleaved slices with one counter each:
        // blocks are dealt to the XCDs round-robin, so a slice is served by
        // the waves of one XCD.  A wave's first tile is static; the next
        // grab is issued before the current tile runs, hiding its latency.
        const uint32_t ns = min(gridDim.x, (uint32_t)QSLICES);
        const uint32_t sl = blockIdx.x % ns;
        const uint32_t nbl = (gridDim.x - sl + ns - 1) / ns;  // blocks of this slice
        uint64_t i = (uint64_t)(blockIdx.x / ns) * WPB + wave;
        for (;;) {
            const uint64_t q = queue_entry(i, sl, ns);
            if (q >= nTiles) break;
            uint32_t nx = 0;
            if (lane == 0) nx = atomicAdd(a.qctr + sl * QSTRIDE, 1u);
            run(a.tileOrder[q]);
            i = (uint64_t)nbl * WPB + readlane_u32(nx, 0);
        }
    } else {
        // persistent grid-stride over the tiles in cell order
        for (uint32_t tile = blockIdx.x * WPB + wave; tile < nTiles; tile += gridDim.x * WPB) run(tile);
    }
    flush_stats(a, st);
    // mask words this wave's pieces needed, plus its last partial block,
    // on the second word of the wave's counter slice
    if (PASS == PASS_COUNT && a.recDir && lane == 0) {
        const uint32_t sl = (blockIdx.x * WPB + wave) % REC_NSLICE;
        atomicAdd(&a.recCtr[REC_SLICE0 + sl * REC_STRIDE + 1], (unsigned long long)(R.want + REC_BLK));
    }
}

template <int K, int D, int PASS, int PREC>
__global__ __launch_bounds__(JB) void k_join(JoinArgs a) {
    __shared__ TileLds<D, PREC> lds[WPB];
    join_body<K, D, PASS, PREC, false>(a, lds);
}

// k <= 2: four waves per SIMD (at most 128 VGPRs); measured faster at
// 2-D than the wider allocation the compiler picks by itself
template <int K, int D, int PASS, int PREC, bool DN = false>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(4))) void k_join_o4(JoinArgs a) {
    __shared__ TileLds<D, PREC> lds[WPB];
    join_body<K, D, PASS, PREC, false, TileLds<D, PREC>, DN>(a, lds);
}

// 5 <= k <= 7: three waves per SIMD (at most 168 VGPRs), which the LDS
// of these tiles still allows; measured faster at 6-D than the two waves
// the compiler's own allocation leaves
template <int K, int D, int PASS, int PREC, bool DN = false>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(3))) void k_join_o3(JoinArgs a) {
    __shared__ TileLds<D, PREC> lds[WPB];
    join_body<K, D, PASS, PREC, false, TileLds<D, PREC>, DN>(a, lds);
}

// 5 <= k <= 7, FP64 count: four waves per SIMD, which its registers (at
// most 128) and the count-only LDS allow
template <int K, int D, bool DN = false>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(4))) void k_join_c4(JoinArgs a) {
    __shared__ TileLdsCount<D, 0> lds[WPB];
    join_body<K, D, PASS_COUNT, 0, false, TileLdsCount<D, 0>, DN>(a, lds);
}

// Fill that replays the hit record of a count pass of precision PREC
template <int K, int D, int PREC, bool DN = false>
__global__ __launch_bounds__(JB) void k_join_rp(JoinArgs a) {
    __shared__ TileLdsIds<D, PREC> lds[WPB];
    join_body<K, D, PASS_FILL, PREC, true, TileLdsIds<D, PREC>, DN>(a, lds);
}

// k <= 2: five waves per SIMD (at most 96 VGPRs); measured faster at 2-D,
// slower at 6-D
template <int K, int D, int PREC, bool DN = false>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(5))) void k_join_rp_o5(JoinArgs a) {
    __shared__ TileLdsIds<D, PREC> lds[WPB];
    join_body<K, D, PASS_FILL, PREC, true, TileLdsIds<D, PREC>, DN>(a, lds);
}

// ------------------------------------------------------------ many joins
// One persistent launch over the tiles of many independent joins: work
// item w is tile workTile[w] of instance workInst[w], whose arguments are
// args[instance].  FP32 instances run the FP32 LDS tile walk when
// d == k <= 8 and the matrix-core walk otherwise; FP64 ones the LDS tile
// walk when d == k <= 8 and the wide walk otherwise.  Statistics go to the
// instance's counters.
template <int A, int B>
struct cmax {
    static constexpr int v = A > B ? A : B;
};
template <int D>
struct LdsMax {
    static constexpr int v =
        cmax<cmax<cmax<(int)sizeof(TileLds<D, 0>), (int)sizeof(TileLds<D, 1>)>::v,
                  cmax<(int)sizeof(TileLdsIds<D, 0>), (int)sizeof(TileLdsIds<D, 1>)>::v>::v,
             LdsMax<D - 1>::v>::v;
};
template <>
struct LdsMax<0> {
    static constexpr int v = cmax<(int)sizeof(WideLds), (int)sizeof(MfmaLds<64>)>::v;
};
constexpr int MANY_LDS = (LdsMax<DEV_MAXK>::v + 15) / 16 * 16;
// FP32 instances none of which uses the matrix cores: the FP32 LDS tile
// walks only
template <int D>
struct LdsMax32 {
    static constexpr int v =
        cmax<cmax<(int)sizeof(TileLds<D, 1>), (int)sizeof(TileLdsIds<D, 1>)>::v, LdsMax32<D - 1>::v>::v;
};
template <>
struct LdsMax32<0> {
    static constexpr int v = 0;
};
constexpr int MANY_LDS32 = (LdsMax32<DEV_MAXK>::v + 15) / 16 * 16;

template <int K, int PASS, int PREC, bool RP>
__device__ __forceinline__ void many_tile(const JoinArgs &a, uint32_t tile, unsigned char *raw, LaneStats &st,
                                          TileRec &R) {
    if constexpr (RP) {
        // the LDS tile walks replay their record; the wide walk keeps none
        if (PREC == 1) process_tile<K, K, PASS, 1, true>(a, tile, *reinterpret_cast<TileLdsIds<K, 1> *>(raw), st, R);
        else if (a.d == K) process_tile<K, K, PASS, 0, true>(a, tile, *reinterpret_cast<TileLdsIds<K, 0> *>(raw), st, R);
        else process_tile_wide<PASS>(a, tile, *reinterpret_cast<WideLds *>(raw), st);
    } else {
        if (PREC == 1) process_tile<K, K, PASS, 1, false>(a, tile, *reinterpret_cast<TileLds<K, 1> *>(raw), st, R);
        else if (a.d == K) process_tile<K, K, PASS, 0, false>(a, tile, *reinterpret_cast<TileLds<K, 0> *>(raw), st, R);
        else process_tile_wide<PASS>(a, tile, *reinterpret_cast<WideLds *>(raw), st);
    }
}

template <int DP, int PASS, bool RP>
__device__ __forceinline__ void many_mfma(const JoinArgs &a, uint32_t tile, unsigned char *raw, int &cleared,
                                          LaneStats &st, TileRec &R) {
    // the pad columns [d, DP) of the staged rows must be zero: clear when
    // (DP, d) changes or another walk used the LDS
    MfmaLds<DP> &L = *reinterpret_cast<MfmaLds<DP> *>(raw);
    const int key = DP * 128 + a.d;
    if (cleared != key) {
        mfma_lds_clear(L);
        cleared = key;
    }
    process_tile_mfma<DP, PASS, RP>(a, tile, L, st, R);
}

// RP (fill): the instances' LDS tile and matrix-core walks replay the hit
// record of the count kernel; a record marked unusable makes every item an
// error instead.  MF = 2: FP32 instances without matrix-core walks, whose
// smaller LDS and register needs allow three waves per SIMD.
template <int PASS, int MF, bool RP>
__device__ __forceinline__ void many_body(const JoinArgs *__restrict__ args, const uint32_t *__restrict__ workInst,
                                          const uint32_t *__restrict__ workTile, uint32_t W, uint32_t *qctr,
                                          uint32_t chunk, int useQueue, unsigned char *rawBase, int ldsBytes) {
    const unsigned wave = threadIdx.x >> 6;
    const unsigned lane = __lane_id();
    unsigned char *raw = rawBase + (size_t)wave * ldsBytes;
    LaneStats st;
    TileRec R;  // the wave's block of record words, across its items
    uint32_t cur = 0xFFFFFFFFu;
    int cleared = 0;  // (DP, d) whose zero pad columns the LDS holds, 0: none
    auto run = [&](uint32_t w) __attribute__((always_inline)) {
        const uint32_t inst = (uint32_t)__builtin_amdgcn_readfirstlane((int)workInst[w]);
        const uint32_t tile = (uint32_t)__builtin_amdgcn_readfirstlane((int)workTile[w]);
        if (inst != cur) {
            if (cur != 0xFFFFFFFFu) flush_stats(args[cur], st);
            st = LaneStats();
            cur = inst;
        }
        const JoinArgs &a = args[inst];
        if (RP && a.recCtr[REC_BAD]) {
            if (lane == 0) st.err++;
            return;
        }
        if (MF == 1 && a.mDP > 0) {
            switch (a.mDP) {
            case 4: many_mfma<4, PASS, RP>(a, tile, raw, cleared, st, R); break;
            case 8: many_mfma<8, PASS, RP>(a, tile, raw, cleared, st, R); break;
            case 16: many_mfma<16, PASS, RP>(a, tile, raw, cleared, st, R); break;
            case 32: many_mfma<32, PASS, RP>(a, tile, raw, cleared, st, R); break;
            case 64: many_mfma<64, PASS, RP>(a, tile, raw, cleared, st, R); break;
            }
            return;
        }
        cleared = 0;
        switch (a.k) {
        case 1: many_tile<1, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 2: many_tile<2, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 3: many_tile<3, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 4: many_tile<4, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 5: many_tile<5, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 6: many_tile<6, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 7: many_tile<7, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        case 8: many_tile<8, PASS, (MF ? 1 : 0), RP>(a, tile, raw, st, R); break;
        }
    };
    if (useQueue) {
        // chunks of `chunk` consecutive items: the wave's first chunk is
        // static, later ones come from the counter; the next grab is issued
        // before the current chunk runs
        const uint64_t nw = (uint64_t)gridDim.x * WPB;
        uint64_t m = (uint64_t)blockIdx.x * WPB + wave;
        for (;;) {
            const uint64_t b0 = m * chunk;
            if (b0 >= W) break;
            uint32_t nx = 0;
            if (lane == 0) nx = atomicAdd(qctr, 1u);
            const uint64_t b1 = min<uint64_t>(b0 + chunk, W);
            for (uint64_t w = b0; w < b1; w++) run((uint32_t)w);
            m = nw + readlane_u32(nx, 0);
        }
    } else {
        for (uint32_t w = blockIdx.x * WPB + wave; w < W; w += gridDim.x * WPB) run(w);
    }
    if (cur != 0xFFFFFFFFu) {
        flush_stats(args[cur], st);
        // record words this wave's items needed (the record is shared)
        if (PASS == PASS_COUNT && args[cur].recDir && lane == 0) {
            const uint32_t sl = (blockIdx.x * WPB + wave) % REC_NSLICE;
            atomicAdd(&args[cur].recCtr[REC_SLICE0 + sl * REC_STRIDE + 1], (unsigned long long)(R.want + REC_BLK));
        }
    }
}

template <int PASS, int MF, bool RP>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(2))) void k_join_many(const JoinArgs *__restrict__ args,
                                                  const uint32_t *__restrict__ workInst,
                                                  const uint32_t *__restrict__ workTile, uint32_t W,
                                                  uint32_t *qctr, uint32_t chunk, int useQueue) {
    __shared__ __align__(16) unsigned char lds[WPB][MANY_LDS];
    many_body<PASS, MF, RP>(args, workInst, workTile, W, qctr, chunk, useQueue, &lds[0][0], MANY_LDS);
}

template <int PASS, bool RP>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(3))) void k_join_many32(const JoinArgs *__restrict__ args,
                                                    const uint32_t *__restrict__ workInst,
                                                    const uint32_t *__restrict__ workTile, uint32_t W,
                                                    uint32_t *qctr, uint32_t chunk, int useQueue) {
    __shared__ __align__(16) unsigned char lds[WPB][MANY_LDS32];
    many_body<PASS, 2, RP>(args, workInst, workTile, W, qctr, chunk, useQueue, &lds[0][0], MANY_LDS32);
}

// Tiles: runs of at most 64 consecutive points of one cell row within
// [q0, q1).  A head starts every row and every 64 points inside a row.
__global__ __launch_bounds__(JB) void k_tile_heads(JoinArgs a, uint32_t q0, uint32_t q1, uint32_t *head) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    const uint64_t s = (uint64_t)q0 + i;
    if (s >= q1) return;
    const uint32_t cell = a.cellOf[s];
    const uint64_t row = a.cellLinear[cell] - a.cellCoord[(uint64_t)cell * a.k];
    const uint32_t rs = max(pos_lb(a, row), q0);
    head[i] = ((s - rs) & 63u) == 0 ? 1u : 0u;
}

// z: counters of the join that follows, cleared here (each when set)
__global__ __launch_bounds__(JB) void k_tile_compact(const uint32_t *head, const uint32_t *ex, uint32_t q0,
                                                     uint32_t q1, uint32_t *tileStart, uint32_t *nTiles,
                                                     TileZero z) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    const uint64_t len = q1 - q0;
    if (z.classes && i < 2 * 256) z.classes[i] = 0u;
    if (z.mir && i < z.n) z.mir[i] = 0u;
    if (z.stat && i < 8) z.stat[i] = 0ull;
    if (z.qctr && i < (uint64_t)QSLICES * QSTRIDE) z.qctr[i] = 0u;
    if (i >= len) return;
    if (head[i]) tileStart[ex[i]] = q0 + (uint32_t)i;
    if (i == len - 1) {
        tileStart[ex[len]] = q1;
        *nTiles = ex[len];
    }
}

// ------------------------------------------------------------ workload
// Per non-empty cell: sizes of the non-empty cells of its 3^k stencil, own
// included; lid counts only cells with linear ID >= own
// (sortByWorkLoad / sortByWorkLoadLidUnicomp, kernel.cu:33-250).
template <int K>
__global__ __launch_bounds__(JB) void k_workload(JoinArgs a, int lid, uint32_t *w) {
    const uint64_t cell = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (cell >= a.nne) return;
    uint32_t c[K];
#pragma unroll
    for (int j = 0; j < K; j++) c[j] = a.cellCoord[cell * K + j];
    const uint64_t linA = a.cellLinear[cell];
    uint32_t sum = a.cellStart[cell + 1] - a.cellStart[cell];
    int o[K];
#pragma unroll
    for (int j = 0; j < K; j++) o[j] = -1;
    int noff = 1;
#pragma unroll
    for (int j = 0; j < K; j++) noff *= 3;
    for (int it = 0; it < noff; ++it) {
        bool zero = true, inb = true;
        int64_t dl = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            zero = zero && o[j] == 0;
            const int64_t v = (int64_t)c[j] + o[j];
            inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
            dl += (int64_t)o[j] * a.mult[j];
        }
        if (!zero && inb) {
            const uint64_t linB = (uint64_t)((int64_t)linA + dl);
            if (!lid || linB >= linA) {
                const int64_t b = find_cell(a, linB);
                if (b >= 0) sum += a.cellStart[b + 1] - a.cellStart[b];
            }
        }
        bool carry = true;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (carry) {
                if (o[j] == 1) {
                    o[j] = -1;
                } else {
                    o[j]++;
                    carry = false;
                }
            }
        }
    }
    w[cell] = sum;
}

__global__ void k_desc_keys(const uint32_t *w, uint64_t n, uint32_t *key, uint32_t *val) {
    uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (i >= n) return;
    key[i] = 0xFFFFFFFFu - w[i];
    val[i] = (uint32_t)i;
}

// Descending 8-bit log-scale bucket of a cost: 8 steps per octave.
__device__ __forceinline__ uint32_t cost_key(uint32_t v) {
    if (v == 0) return 255u;
    const int e = 31 - __builtin_clz(v);
    const uint32_t m = e >= 3 ? (v >> (e - 3)) & 7u : (v << (3 - e)) & 7u;
    return 255u - min(255u, 8u * (uint32_t)e + m);
}

// Work-queue classes: the 256 cost buckets.  bucket[0..QB) counts the tiles
// of each class, bucket[QB..2 QB) are the placement cursors; both start at
// zero (k_tile_compact clears them).
constexpr uint32_t QB = 256;
static_assert(JB == (int)QB, "one thread per class");

// Tile cost: the candidates the tile stages, i.e. the union of its
// neighbour cell ranges over the 3^(k-1) neighbour rows, as a class
// (cost_key) per tile, counted per class in bucket[].
template <int K>
__global__ __launch_bounds__(JB) void k_tile_cost(JoinArgs a, uint32_t bound, uint32_t *key, uint32_t *bucket) {
    __shared__ uint32_t hist[QB];
    hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t t = blockIdx.x * JB + threadIdx.x;
    uint32_t cost = 0;
    if (t < bound && t < *a.nTilesDev) {
        const uint32_t s0 = a.tileStart[t], s1 = a.tileStart[t + 1] - 1u;
        const uint32_t cA = a.cellOf[s0], cB = a.cellOf[s1];
        uint32_t c[K];
#pragma unroll
        for (int j = 0; j < K; j++) c[j] = a.cellCoord[(uint64_t)cA * K + j];
        const uint32_t xb = a.cellCoord[(uint64_t)cB * K];
        const uint64_t rowBase = a.cellLinear[cA] - c[0];
        const uint32_t nC0 = a.nCells[0];
        const uint32_t ux0 = c[0] >= 1 ? c[0] - 1 : 0u;
        const uint32_t ux1 = xb + 1 < nC0 ? xb + 1 : nC0 - 1;
        int o[K];
#pragma unroll
        for (int j = 0; j < K; j++) o[j] = -1;
        int nrow = 1;
#pragma unroll
        for (int j = 1; j < K; j++) nrow *= 3;
        for (int it = 0; it < nrow; ++it) {
            bool inb = true;
            int64_t dl = 0;
#pragma unroll
            for (int j = 1; j < K; j++) {
                const int64_t v = (int64_t)c[j] + o[j];
                inb = inb && v >= 0 && v < (int64_t)a.nCells[j];
                dl += (int64_t)o[j] * a.mult[j];
            }
            if (inb) {
                const uint64_t rb = (uint64_t)((int64_t)rowBase + dl);
                cost += pos_lb(a, rb + ux1 + 1) - pos_lb(a, rb + ux0);
            }
            bool carry = true;
#pragma unroll
            for (int j = 1; j < K; j++) {
                if (carry) {
                    if (o[j] == 1) {
                        o[j] = -1;
                    } else {
                        o[j]++;
                        carry = false;
                    }
                }
            }
        }
        const uint32_t kk = cost_key(cost);
        key[t] = kk;
        atomicAdd(&hist[kk], 1u);
    }
    __syncthreads();
    const uint32_t h = hist[threadIdx.x];
    if (h) atomicAdd(&bucket[threadIdx.x], h);
}

// The queue: tiles by class, heaviest first.  Each block takes the class
// offsets from the counts (a 256-entry scan of its own), ranks its tiles
// stably inside each class (wave64 ballots over the 8 key bits, then the
// waves in order) and reserves one run per class with one atomic, so a
// class holds runs of tiles in cell order (neighbouring tiles stay close
// in the queue).  Positions stay below nTiles: the runs of a class add up
// to its count.
__global__ __launch_bounds__(JB) void k_tile_place(const uint32_t *nTilesDev, const uint32_t *key, uint32_t *bucket,
                                                   uint32_t *order) {
    __shared__ uint32_t base[QB];                // the block's first position in each class
    __shared__ uint32_t wcnt[JB / 64][QB];       // tiles of each class per wave, then offsets
    __shared__ uint32_t wsum[JB / 64];
    const unsigned lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint32_t nT = *nTilesDev;
    // class offsets: exclusive scan of the counts
    const uint32_t x = bucket[threadIdx.x];
    uint32_t inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(inc, off, 64);
        if (lane >= (unsigned)off) inc += u;
    }
    if (lane == 63) wsum[wave] = inc;
#pragma unroll
    for (int w = 0; w < JB / 64; w++) wcnt[w][threadIdx.x] = 0u;
    __syncthreads();
    uint32_t before = 0;
    for (unsigned w = 0; w < wave; w++) before += wsum[w];
    const uint32_t classBase = before + inc - x;
    // stable rank of each tile among the wave's tiles of its class
    const uint32_t t = blockIdx.x * JB + threadIdx.x;
    const bool valid = t < nT;
    const uint32_t kk = valid ? key[t] : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (kk >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    if (valid && rank == 0) wcnt[wave][kk] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
        // class threadIdx.x: the waves' offsets in order, then one run for the block
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < JB / 64; w++) {
            const uint32_t c = wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = run;
            run += c;
        }
        base[threadIdx.x] = run ? classBase + atomicAdd(&bucket[QB + threadIdx.x], run) : 0u;
    }
    __syncthreads();
    if (valid) order[base[kk] + wcnt[wave][kk] + rank] = t;
}

// Pieces a tile walk may execute per tile (tile_piece_bound); zero
// for entries from nTiles to bound.  Scanned, they lay out the hit record.
template <int K>
// zeroCtr (when set): the hit-record counters, cleared here
__global__ __launch_bounds__(JB) void k_tile_bound(JoinArgs a, uint32_t bound, uint32_t cw, uint32_t *out,
                                                   unsigned long long *zeroCtr) {
    const uint32_t t = blockIdx.x * JB + threadIdx.x;
    if (zeroCtr && t < REC_NCTR) zeroCtr[t] = 0ull;
    if (t >= bound) return;
    uint32_t v = 0;
    if (t < *a.nTilesDev) {
        const uint32_t cA = a.cellOf[a.tileStart[t]], cB = a.cellOf[a.tileStart[t + 1] - 1u];
        uint32_t c[K];
#pragma unroll
        for (int j = 0; j < K; j++) c[j] = a.cellCoord[(uint64_t)cA * K + j];
        const uint64_t rowBase = a.cellLinear[cA] - c[0];
        v = tile_piece_bound<K>(a, c, rowBase, c[0], a.cellCoord[(uint64_t)cB * K], cw);
    }
    out[t] = v;
}

// ----------------------------------------------------- small row kernels
// The host's view after the count pass, gathered for one copy: out[0] =
// |S|, out[1..8] = statistics, out[16] = record unusable, out[18] =
// directory entries the record needed (the last two when recording).
// The count's error counter is cleared once read, for the fill's errors.
// The live statistics sit at out[SC_STAT..] of the same buffer, so one copy
// of it after the fill also carries the fill's errors and overflow flag.
constexpr int SC_STAT = 24, SC_WORDS = 32;
struct CountScalars {
    const uint64_t *total;
    unsigned long long *stat;
    const unsigned long long *recCtr;  // null: not recording
    const uint64_t *dirTotal;          // null: not recording
    uint64_t *out;                     // null: nothing to gather
};

__device__ __forceinline__ void gather_count_scalars(const CountScalars &c, unsigned t) {
    if (t == 0) c.out[0] = *c.total;
    if (t >= 1 && t <= 8) c.out[t] = c.stat[t - 1];
    if (t == 1 + STAT_ERR) c.stat[STAT_ERR] = 0ull;
    if (t == 16) c.out[16] = c.recCtr ? c.recCtr[REC_BAD] : 0ull;
    if (t == 18) c.out[18] = c.dirTotal ? *c.dirTotal : 0ull;
}

__global__ void k_count_scalars(CountScalars c) { gather_count_scalars(c, threadIdx.x); }

// mirrored-hit cursors of rows [w0, w1): start right after the own hits.
// recBad (when set): only when the hit record is unusable (the fill then
// evaluates distances again and takes its mirrored slots from the cursors)
// qctr (when set): the work-queue counters at zero for the fill that follows.
// sc.out (when set): block 0 also gathers the count pass's scalars.
__global__ void k_mir_cursors(const uint64_t *offS, const uint32_t *own, uint32_t w0, uint32_t w1,
                              uint64_t outBase, unsigned long long *mirCur, const unsigned long long *recBad,
                              uint32_t *qctr, CountScalars sc) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (sc.out && blockIdx.x == 0) gather_count_scalars(sc, threadIdx.x);
    if (qctr && i < (uint64_t)QSLICES * QSTRIDE) qctr[i] = 0u;
    uint64_t s = (uint64_t)w0 + i;
    if (s >= w1) return;
    if (recBad && *recBad == 0ull) return;
    mirCur[s] = offS[s] - outBase + own[s];
}

// FP32 copy of the coordinates relative to the grid origin, rows of
// `stride` floats (zero padded), and the largest |x - origin| (one atomic
// per block)
__global__ __launch_bounds__(JB) void k_to_fp32(const double *pts, uint64_t n, int d, int k, int stride,
                                                const GridParams *gp, float *out, unsigned int *rmaxBits) {
    __shared__ float wmax[WPB];
    const uint64_t s = (uint64_t)blockIdx.x * JB + threadIdx.x;
    float r = 0.0f;
    if (s < n) {
        for (int j = 0; j < d; j++) {
            const double o = j < k ? gp->minArr[j] : 0.0;
            const double v = pts[s * d + j] - o;
            out[s * stride + j] = (float)v;
            r = fmaxf(r, (float)fabs(v));
        }
        for (int j = d; j < stride; j++) out[s * stride + j] = 0.0f;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r = fmaxf(r, __shfl_xor(r, off, 64));
    if (__lane_id() == 0) wmax[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = wmax[0];
        for (int w = 1; w < WPB; w++) m = fmaxf(m, wmax[w]);
        atomicMax(rmaxBits, __float_as_uint(m));
    }
}

__global__ void k_sample_positions(const uint32_t *sortedIds, uint64_t n, uint64_t stride, uint64_t nsamp,
                                   uint32_t *posOf, uint32_t *list, int phase) {
    uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (phase == 0) {
        if (i < n) posOf[sortedIds[i]] = (uint32_t)i;
    } else {
        if (i < nsamp) list[i] = posOf[i * stride];
    }
}

// estimator: every stride-th original point, full stencil, counted per point
template <int K, int D>
__global__ __launch_bounds__(JB) void k_sample_count(JoinArgs a, const uint32_t *list, uint64_t m) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    LaneStats st;
    if (i < m) process_point<K, D, PASS_COUNT, 0>(a, list[i], st);
    flush_stats(a, st);
}

__global__ void k_sum_gather(const uint32_t *v, const uint32_t *idx, uint64_t m, unsigned long long *out) {
    uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    unsigned long long x = i < m ? v[idx[i]] : 0ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (__lane_id() == 0 && x) atomicAdd(out, x);
}

// ---------------------------------------------------------------- host
static unsigned nblk(uint64_t n) {
    uint64_t g = (n + JB - 1) / JB;
    return (unsigned)(g ? g : 1);
}

using JoinKernel = void (*)(JoinArgs);
using SampleKernel = void (*)(JoinArgs, const uint32_t *, uint64_t);

template <int K, int D>
static JoinKernel pick_pass(int pass, int prec) {
    if constexpr (K <= 2 && D > 0) {
        if (pass == PASS_COUNT) return prec ? k_join_o4<K, D, PASS_COUNT, 1> : k_join_o4<K, D, PASS_COUNT, 0>;
        return prec ? k_join_o4<K, D, PASS_FILL, 1> : k_join_o4<K, D, PASS_FILL, 0>;
    } else if constexpr (K >= 5 && K <= 7 && D > 0) {
        if (pass == PASS_COUNT) return prec ? k_join_o3<K, D, PASS_COUNT, 1> : k_join_c4<K, D>;
        return prec ? k_join_o3<K, D, PASS_FILL, 1> : k_join_o3<K, D, PASS_FILL, 0>;
    } else {
        if (pass == PASS_COUNT) return prec ? k_join<K, D, PASS_COUNT, 1> : k_join<K, D, PASS_COUNT, 0>;
        return prec ? k_join<K, D, PASS_FILL, 1> : k_join<K, D, PASS_FILL, 0>;
    }
}

template <int K>
static JoinKernel pick_d(int d, int pass, int prec) {
    if (d == K) return pick_pass<K, K>(pass, prec);
    return pick_pass<K, 0>(pass, prec);
}

static JoinKernel pick_kernel(int k, int d, int pass, int prec) {
    switch (k) {
    case 1: return pick_d<1>(d, pass, prec);
    case 2: return pick_d<2>(d, pass, prec);
    case 3: return pick_d<3>(d, pass, prec);
    case 4: return pick_d<4>(d, pass, prec);
    case 5: return pick_d<5>(d, pass, prec);
    case 6: return pick_d<6>(d, pass, prec);
    case 7: return pick_d<7>(d, pass, prec);
    case 8: return pick_d<8>(d, pass, prec);
    }
    return nullptr;
}

template <int K>
static SampleKernel pick_sample_d(int d) {
    return d == K ? k_sample_count<K, K> : k_sample_count<K, 0>;
}

static SampleKernel pick_sample(int k, int d) {
    switch (k) {
    case 1: return pick_sample_d<1>(d);
    case 2: return pick_sample_d<2>(d);
    case 3: return pick_sample_d<3>(d);
    case 4: return pick_sample_d<4>(d);
    case 5: return pick_sample_d<5>(d);
    case 6: return pick_sample_d<6>(d);
    case 7: return pick_sample_d<7>(d);
    case 8: return pick_sample_d<8>(d);
    }
    return nullptr;
}

using WorkKernel = void (*)(JoinArgs, int, uint32_t *);
static WorkKernel pick_workload(int k) {
    switch (k) {
    case 1: return k_workload<1>;
    case 2: return k_workload<2>;
    case 3: return k_workload<3>;
    case 4: return k_workload<4>;
    case 5: return k_workload<5>;
    case 6: return k_workload<6>;
    case 7: return k_workload<7>;
    case 8: return k_workload<8>;
    }
    return nullptr;
}

using TileCostKernel = void (*)(JoinArgs, uint32_t, uint32_t *, uint32_t *);
static TileCostKernel pick_tile_cost(int k) {
    switch (k) {
    case 1: return k_tile_cost<1>;
    case 2: return k_tile_cost<2>;
    case 3: return k_tile_cost<3>;
    case 4: return k_tile_cost<4>;
    case 5: return k_tile_cost<5>;
    case 6: return k_tile_cost<6>;
    case 7: return k_tile_cost<7>;
    case 8: return k_tile_cost<8>;
    }
    return nullptr;
}

using TileBoundKernel = void (*)(JoinArgs, uint32_t, uint32_t, uint32_t *, unsigned long long *);
static TileBoundKernel pick_tile_bound(int k) {
    switch (k) {
    case 1: return k_tile_bound<1>;
    case 2: return k_tile_bound<2>;
    case 3: return k_tile_bound<3>;
    case 4: return k_tile_bound<4>;
    case 5: return k_tile_bound<5>;
    case 6: return k_tile_bound<6>;
    case 7: return k_tile_bound<7>;
    case 8: return k_tile_bound<8>;
    }
    return nullptr;
}

double eps_threshold(double eps) {
    double t = eps * eps;
    if (!(eps >= 0.0) || std::isinf(eps)) return NAN;
    if (std::sqrt(t) <= eps) {
        for (;;) {
            double nx = std::nextafter(t, INFINITY);
            if (std::sqrt(nx) <= eps) t = nx; else break;
        }
    } else {
        while (!(std::sqrt(t) <= eps)) t = std::nextafter(t, 0.0);
    }
    return t;
}

// Margin M of the FP32 distance: |s32 - T| > M decides the pair exactly.
// Inputs: R bounds |x - origin| over all coordinates, D dims.  Derivation:
// stored y = fl32(fl64(x - o)), |y - (x-o)| <= R(u32 + 2u64) =: E0/2;
// f_j = fl32(y_q - y_p): |f_j - a_j| <= E0 + u32(|a| + E0) =: E;
// FMA sum of D squares: |s32 - sum f^2| <= gamma_D sum f^2;
// |sum f^2 - sum a^2| <= E(2 sqrt(D) |a| + D E);
// FP64 canonical: |d2_64 - |a|^2| <= rho |a|^2, rho = (D+3) u64.
// d2_64 <= T  =>  s32 <= hi(T/(1-rho));  d2_64 > T  =>  s32 >= lo(T/(1+rho)).
// Device time between two events on a stream.  The events are the
// context's (reused by every join) or, when none are given, its own.
struct EvTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    bool used = false, own = false;
    explicit EvTimer(hipStream_t st, hipEvent_t ea = nullptr, hipEvent_t eb = nullptr) : a(ea), b(eb), s(st) {
        if (!a || !b) {
            own = true;
            a = b = nullptr;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
        }
    }
    ~EvTimer() {
        if (!own) return;
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start() {
        (void)hipEventRecord(a, s);
        used = true;
    }
    void stop() { (void)hipEventRecord(b, s); }
    // start where prev stopped, with no GPU work between: its end event is
    // this one's start (one marker fewer in the stream).  Events not owned.
    void start_at(const EvTimer &prev) {
        a = prev.b;
        used = true;
    }
    double ms() {
        if (!used) return 0.0;
        float f = 0;
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&f, a, b);
        return f;
    }
};

JoinArgs base_args(sj_grid_s *g, int form, double T) {
    JoinArgs a;
    memset(&a, 0, sizeof a);
    a.pts = g->sortedPts.p;
    a.origId = g->sortedIds.p;
    a.cellOf = g->cellOf.p;
    a.cellStart = g->cellStart.p;
    a.cellLinear = g->cellLinear.p;
    a.cellCoord = g->cellCoord.p;
    a.posLB = g->dense ? g->posLB.p : nullptr;
    a.nne = g->nNonEmpty;
    for (int j = 0; j < DEV_MAXK; j++) {
        a.nCells[j] = j < g->k ? g->gp.nCells[j] : 1;
        a.mult[j] = j < g->k ? (int64_t)g->gp.mult[j] : 0;
    }
    a.d = g->d;
    a.k = g->k;
    a.form = form;
    a.useTiles = 1;
    a.T = T;
    a.cellW = g->eps;
    a.qBegin = 0;
    a.qEnd = (uint32_t)g->n;
    a.wBegin = 0;
    a.wEnd = (uint32_t)g->n;
    return a;
}

static sj_status cell_workload(sj_grid_s *g, int lid, DBuf<uint32_t> &w) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t nne = g->nNonEmpty;
    JoinArgs a = base_args(g, SJ_FORM_FULL, 0.0);
    SJ_TRY(w.alloc(pool, nne));
    hipLaunchKernelGGL(pick_workload(g->k), dim3(nblk(nne)), dim3(JB), 0, s, a, lid, w.p);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// Row-aligned tiles over [q0, q1); the count stays on the device.
sj_status build_tiles(sj_grid_s *g, const JoinArgs &a, uint32_t q0, uint32_t q1, DBuf<uint32_t> &tileStart,
                             DBuf<uint32_t> &nTiles, const TileZero &zero) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t len = q1 - q0;
    DBuf<uint32_t> head, ex;
    SJ_TRY(head.alloc(pool, len));
    SJ_TRY(ex.alloc(pool, len + 1));
    SJ_TRY(tileStart.alloc(pool, len + 1));
    SJ_TRY(nTiles.alloc(pool, 1));
    if (len == 0) {
        SJ_TRY(fill_u32(s, nTiles.p, 1, 0u));
        return SJ_OK;
    }
    hipLaunchKernelGGL(k_tile_heads, dim3(nblk(len)), dim3(JB), 0, s, a, q0, q1, head.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u32(pool, s, head.p, len, ex.p));
    const uint64_t nz = std::max<uint64_t>(len, std::max<uint64_t>(2 * QB, std::max<uint64_t>(zero.n, 256)));
    hipLaunchKernelGGL(k_tile_compact, dim3(nblk(nz)), dim3(JB), 0, s, head.p, ex.p, q0, q1, tileStart.p, nTiles.p,
                       zero);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// Work queue: tiles by descending staged-candidate bucket.  The tile count
// stays on the device: the sort covers an upper bound (one tile per 64
// points plus one per non-empty row, at most one per non-empty cell).
static sj_status build_tile_queue(sj_grid_s *g, const JoinArgs &a, uint64_t len, DBuf<uint32_t> &order,
                                  uint32_t *classes) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t bound64 = (len + 63) / 64 + std::min<uint64_t>(len, g->nNonEmpty);
    if (bound64 > 0xFFFFFFFFull) {
        set_error("work queue: %llu tiles exceed 32-bit indices", (unsigned long long)bound64);
        return SJ_ERR_OVERFLOW;
    }
    const uint32_t bound = (uint32_t)bound64;
    DBuf<uint32_t> key;
    SJ_TRY(key.alloc(pool, bound));
    SJ_TRY(order.alloc(pool, bound));
    if (bound == 0) return SJ_OK;
    hipLaunchKernelGGL(pick_tile_cost(g->k), dim3(nblk(bound)), dim3(JB), 0, s, a, bound, key.p, classes);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_place, dim3(nblk(bound)), dim3(JB), 0, s, a.nTilesDev, key.p, classes, order.p);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// Workload per cell (w) and the cells by descending workload, ties by
// ascending cell index (order), both on the device.
static sj_status workload_order_device(sj_grid_s *g, int lid, DBuf<uint32_t> &w, DBuf<uint32_t> &order) {
    if (g->k > DEV_MAXK) {
        set_error("workload: k=%d above the device stencil limit %d", g->k, DEV_MAXK);
        return SJ_ERR_UNSUPPORTED;
    }
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t nne = g->nNonEmpty;
    DBuf<uint32_t> key, key2, val2;
    SJ_TRY(cell_workload(g, lid, w));
    SJ_TRY(key.alloc(pool, nne));
    SJ_TRY(key2.alloc(pool, nne));
    SJ_TRY(order.alloc(pool, nne));
    SJ_TRY(val2.alloc(pool, nne));
    hipLaunchKernelGGL(k_desc_keys, dim3(nblk(nne)), dim3(JB), 0, s, w.p, nne, key.p, order.p);
    SJ_LAUNCH_CHECK();
    uint32_t *kp = key.p, *ka = key2.p, *vp = order.p, *va = val2.p;
    SJ_TRY(radix_sort_pairs_u32(pool, s, &kp, &ka, &vp, &va, nne, 32));
    if (vp != order.p) std::swap(order, val2);
    return SJ_OK;
}

sj_status grid_workload(sj_grid_s *g, int lid, int64_t *nb, uint32_t *order) {
    const uint64_t nne = g->nNonEmpty;
    DBuf<uint32_t> w, ord;
    SJ_TRY(workload_order_device(g, lid, w, ord));
    std::vector<uint32_t> hw(nne);
    SJ_TRY_HIP(hipStreamSynchronize(g->ctx->stream));
    SJ_TRY_HIP(hipMemcpy(hw.data(), w.p, nne * 4, hipMemcpyDeviceToHost));
    if (nb)
        for (uint64_t i = 0; i < nne; i++) nb[i] = hw[i];
    if (order) SJ_TRY_HIP(hipMemcpy(order, ord.p, nne * 4, hipMemcpyDeviceToHost));
    return SJ_OK;
}

template <int K>
static JoinKernel pick_replay_k(int prec) {
    if constexpr (K <= 2) return prec ? k_join_rp_o5<K, K, 1> : k_join_rp_o5<K, K, 0>;
    else return prec ? k_join_rp<K, K, 1> : k_join_rp<K, K, 0>;
}

// k = d = 2 or 6 on a dense grid: kernels without the binary-search
// lookups (the 2-D and 6-D shapes; other k use the general kernels)
static JoinKernel pick_dense(int k, int pass, int prec, bool replay) {
    if (k == 2) {
        if (replay) return prec ? k_join_rp_o5<2, 2, 1, true> : k_join_rp_o5<2, 2, 0, true>;
        if (pass == PASS_COUNT)
            return prec ? k_join_o4<2, 2, PASS_COUNT, 1, true> : k_join_o4<2, 2, PASS_COUNT, 0, true>;
        return prec ? k_join_o4<2, 2, PASS_FILL, 1, true> : k_join_o4<2, 2, PASS_FILL, 0, true>;
    }
    if (replay) return prec ? k_join_rp<6, 6, 1, true> : k_join_rp<6, 6, 0, true>;
    if (pass == PASS_COUNT) return prec ? k_join_o3<6, 6, PASS_COUNT, 1, true> : k_join_c4<6, 6, true>;
    return prec ? k_join_o3<6, 6, PASS_FILL, 1, true> : k_join_o3<6, 6, PASS_FILL, 0, true>;
}

static JoinKernel pick_replay(int k, int prec) {
    switch (k) {
    case 1: return pick_replay_k<1>(prec);
    case 2: return pick_replay_k<2>(prec);
    case 3: return pick_replay_k<3>(prec);
    case 4: return pick_replay_k<4>(prec);
    case 5: return pick_replay_k<5>(prec);
    case 6: return pick_replay_k<6>(prec);
    case 7: return pick_replay_k<7>(prec);
    case 8: return pick_replay_k<8>(prec);
    }
    return nullptr;
}

// Walks whose count pass keeps a hit record the fill can replay: the LDS
// tile walk (d = k <= 8) and the matrix-core walk (FP32, mDP > 0).
static bool record_walk(const sj_grid_s *g, const JoinArgs &a) {
    return a.useTiles && ((g->d == g->k && g->d <= DEV_MAXK) || a.mDP > 0);
}

sj_status tpp_point_order(sj_grid_s *g, const uint32_t *corder, DBuf<uint32_t> &order);
sj_status launch_join_tpp(const JoinArgs &a, int k, int pass, uint64_t n, hipStream_t s);

// replay: the fill replays the count pass's hit record (record_walk grids)
static sj_status launch_join(sj_grid_s *g, const JoinArgs &a, int pass, int prec, bool replay = false) {
    hipStream_t s = g->ctx->stream;
    if (a.tppOrder) return launch_join_tpp(a, g->k, pass, g->n, s);
    if (prec == SJ_PREC_FP32_RECHECK && a.mDP > 0) return launch_join_mfma(a, pass, g->ctx->cu_count, s, replay);
    if (prec == SJ_PREC_FP64 && a.useTiles && wide_supported(g->d, g->k))
        return launch_join_wide(a, pass, g->ctx->cu_count, s);
    JoinKernel kfn = replay ? pick_replay(g->k, prec) : pick_kernel(g->k, g->d, pass, prec);
    if (a.posLB && g->k == g->d && (g->k == 2 || g->k == 6)) kfn = pick_dense(g->k, pass, prec, replay);
    if (!kfn) {
        set_error("no traversal kernel for k=%d", g->k);
        return SJ_ERR_UNSUPPORTED;
    }
    // persistent waves: enough to fill every CU, at most one per tile bound
    const uint64_t len = (uint64_t)a.qEnd - a.qBegin;
    const uint64_t maxTiles = len;  // one point per tile at worst
    uint64_t blocks = std::min<uint64_t>((uint64_t)g->ctx->cu_count * 16, (maxTiles + WPB - 1) / WPB);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(kfn, dim3((unsigned)blocks), dim3(JB), 0, s, a);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

sj_status launch_join_many(const JoinArgs *d_args, const uint32_t *workInst, const uint32_t *workTile,
                           uint32_t W, uint32_t *qctr, uint32_t chunk, int useQueue, int pass, int mfma,
                           unsigned blocks, hipStream_t s, bool replay) {
    if (W == 0) return SJ_OK;
    if (chunk == 0) chunk = 1;
    if (blocks == 0) blocks = 1;
    using ManyKernel = void (*)(const JoinArgs *, const uint32_t *, const uint32_t *, uint32_t, uint32_t *, uint32_t,
                                int);
    ManyKernel kfn;
    if (mfma == 2)  // FP32, no matrix-core instance
        kfn = pass == PASS_COUNT ? k_join_many32<PASS_COUNT, false>
                                 : (replay ? k_join_many32<PASS_FILL, true> : k_join_many32<PASS_FILL, false>);
    else if (pass == PASS_COUNT) kfn = mfma ? k_join_many<PASS_COUNT, 1, false> : k_join_many<PASS_COUNT, 0, false>;
    else if (replay) kfn = mfma ? k_join_many<PASS_FILL, 1, true> : k_join_many<PASS_FILL, 0, true>;
    else kfn = mfma ? k_join_many<PASS_FILL, 1, false> : k_join_many<PASS_FILL, 0, false>;
    hipLaunchKernelGGL(kfn, dim3(blocks), dim3(JB), 0, s, d_args, workInst, workTile, W, qctr, chunk, useQueue);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

sj_status launch_tile_bound(const JoinArgs &a, int k, uint32_t bound, uint32_t cw, uint32_t *out, hipStream_t s) {
    TileBoundKernel kfn = pick_tile_bound(k);
    if (!kfn) {
        set_error("tile bound: k=%d above the device stencil limit", k);
        return SJ_ERR_UNSUPPORTED;
    }
    if (bound == 0) return SJ_OK;
    hipLaunchKernelGGL(kfn, dim3((bound + JB - 1) / JB), dim3(JB), 0, s, a, bound, cw, out, nullptr);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

sj_status launch_tile_cost(const JoinArgs &a, int k, uint32_t bound, uint32_t *key, uint32_t *val, hipStream_t s) {
    TileCostKernel kfn = pick_tile_cost(k);
    if (!kfn) {
        set_error("tile cost: k=%d above the device stencil limit", k);
        return SJ_ERR_UNSUPPORTED;
    }
    if (bound == 0) return SJ_OK;
    hipLaunchKernelGGL(kfn, dim3((bound + JB - 1) / JB), dim3(JB), 0, s, a, bound, key, val);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// Sizes of the cells in workload order (the scan input of k_order_sample).
__global__ void k_order_sizes(const uint32_t *order, const uint32_t *cellStart, uint64_t nne, uint32_t *size) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (i >= nne) return;
    const uint32_t c = order[i];
    size[i] = cellStart[c + 1] - cellStart[c];
}

// The first nsamp points of the cells taken in workload order (the points
// of a cell in ascending ID): list[r] = sorted position of the r-th point.
__global__ void k_order_sample(const uint32_t *order, const uint32_t *cellStart, const uint64_t *rankBase,
                               uint64_t nne, uint64_t nsamp, uint32_t *list) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (i >= nne) return;
    const uint64_t r0 = rankBase[i];
    if (r0 >= nsamp) return;
    const uint32_t c = order[i];
    const uint32_t p0 = cellStart[c], p1 = cellStart[c + 1];
    for (uint32_t p = p0; p < p1 && r0 + (p - p0) < nsamp; p++) list[r0 + (p - p0)] = p;
}

static sj_status sample_count_list(sj_grid_s *g, const uint32_t *list, uint64_t nsamp, uint64_t *count);

// Work-queue estimate (callGPUBatchEstWorkQueue, GPU.cu:1072-1284;
// kernelNDGridIndexWorkQueueBatchEstimatorOLD, kernel.cu:785-942): the
// first n/stride points of the workload-ordered point list (full or LID
// workload, as modes 9 and 10 order it, GPU.cu:1685-1692), full stencil.
sj_status batch_sample_count_wq(sj_grid_s *g, int lid, uint64_t stride, uint64_t *count) {
    if (g->k > DEV_MAXK) {
        set_error("estimator: k=%d above the device stencil limit %d", g->k, DEV_MAXK);
        return SJ_ERR_UNSUPPORTED;
    }
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t n = g->n, nne = g->nNonEmpty;
    if (stride == 0) stride = 1;
    const uint64_t nsamp = (uint64_t)((double)n * (1.0 / (double)stride));
    *count = 0;
    if (nsamp == 0) return SJ_OK;
    DBuf<uint32_t> w, order, size, list;
    DBuf<uint64_t> rank;
    SJ_TRY(workload_order_device(g, lid, w, order));
    SJ_TRY(size.alloc(pool, nne));
    SJ_TRY(rank.alloc(pool, nne + 1));
    SJ_TRY(list.alloc(pool, nsamp));
    hipLaunchKernelGGL(k_order_sizes, dim3(nblk(nne)), dim3(JB), 0, s, order.p, g->cellStart.p, nne, size.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u64(pool, s, size.p, nne, rank.p));
    hipLaunchKernelGGL(k_order_sample, dim3(nblk(nne)), dim3(JB), 0, s, order.p, g->cellStart.p, rank.p, nne, nsamp,
                       list.p);
    SJ_LAUNCH_CHECK();
    return sample_count_list(g, list.p, nsamp, count);
}

// Sampled estimate of |S| (every stride-th original point, full stencil).
sj_status batch_sample_count(sj_grid_s *g, uint64_t stride, uint64_t *count) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t n = g->n;
    if (g->k > DEV_MAXK) {
        set_error("estimator: k=%d above the device stencil limit %d", g->k, DEV_MAXK);
        return SJ_ERR_UNSUPPORTED;
    }
    if (stride == 0) stride = 1;
    const uint64_t nsamp = (uint64_t)((double)n * (1.0 / (double)stride));
    *count = 0;
    if (nsamp == 0) return SJ_OK;
    DBuf<uint32_t> posOf, list;
    SJ_TRY(posOf.alloc(pool, n));
    SJ_TRY(list.alloc(pool, nsamp));
    hipLaunchKernelGGL(k_sample_positions, dim3(nblk(n)), dim3(JB), 0, s, g->sortedIds.p, n, stride, nsamp,
                       posOf.p, list.p, 0);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sample_positions, dim3(nblk(nsamp)), dim3(JB), 0, s, g->sortedIds.p, n, stride,
                       nsamp, posOf.p, list.p, 1);
    SJ_LAUNCH_CHECK();
    return sample_count_list(g, list.p, nsamp, count);
}

// Full-stencil hits of the sorted positions list[0..nsamp), summed.
static sj_status sample_count_list(sj_grid_s *g, const uint32_t *list, uint64_t nsamp, uint64_t *count) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    DBuf<uint32_t> own;
    DBuf<unsigned long long> stat;
    SJ_TRY(own.alloc(pool, g->n));
    SJ_TRY(stat.alloc(pool, 8));
    SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(stat.p), 8, 0ull));
    JoinArgs a = base_args(g, SJ_FORM_FULL, eps_threshold(g->eps));
    a.ownCnt = own.p;
    a.mirCnt = own.p;  // unused by FULL
    a.stat = stat.p;
    hipLaunchKernelGGL(pick_sample(g->k, g->d), dim3(nblk(nsamp)), dim3(JB), 0, s, a, list, nsamp);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_gather, dim3(nblk(nsamp)), dim3(JB), 0, s, own.p, list, nsamp, stat.p + 5);
    SJ_LAUNCH_CHECK();
    uint64_t *pin = g->ctx->pinned;
    SJ_TRY_HIP(hipMemcpyAsync(pin, stat.p + 5, 8, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipStreamSynchronize(s));
    *count = pin[0];
    return SJ_OK;
}

sj_status sort_rows_device(sj_context ctx, const uint64_t *offsets, int32_t *ids, uint64_t n, uint64_t base,
                           uint64_t idsCap = 0);

// Decision constants of the precision form.  The FP32 LDS tile walk
// (d == k <= 8) centres each tile in FP64 and sets its band per tile: it
// needs nothing here.  The others use an FP32 copy relative to the grid
// origin and its largest coordinate (one host sync): the matrix cores
// (d > 8 or d != k, d <= 64) stage from that copy when the coordinates lie
// within 1024 eps of the origin and centre each tile in FP64 otherwise; the
// per-lane FP32 walk (no tiles) uses the global margin of fp32_margin and
// falls back to FP64 (*prec) when that margin would not be useful.
bool use_mfma(const sj_grid_s *g, int useTiles) {
    return useTiles && mfma_padded_dims(g->d) && (g->d > DEV_MAXK || g->d != g->k);
}

sj_status prepare_fp32(sj_grid_s *g, JoinArgs &a, int *prec) {
    sj_context ctx = g->ctx;
    Pool *pool = &ctx->pool;
    hipStream_t s = ctx->stream;
    const uint64_t n = g->n;
    a.mDP = 0;
    a.mSrc32 = 0;
    const bool mf = use_mfma(g, a.useTiles);
    if (!mf && a.useTiles && g->d == g->k && g->d <= DEV_MAXK) return SJ_OK;  // tiles centre in FP64
    DBuf<unsigned int> rbits;
    SJ_TRY(rbits.alloc(pool, 1));
    SJ_TRY(fill_u32(s, rbits.p, 1, 0u));
    // rows padded to the matrix-core operand width for the matrix cores
    const int stride = mf ? mfma_padded_dims(g->d) : g->d;
    if (!g->sortedPts32.p || g->sortedPts32.n < n * (uint64_t)stride)
        SJ_TRY(g->sortedPts32.alloc(pool, n * (uint64_t)stride));
    hipLaunchKernelGGL(k_to_fp32, dim3(nblk(n)), dim3(JB), 0, s, g->sortedPts.p, n, g->d, g->k, stride,
                       g->d_gp.p, g->sortedPts32.p, rbits.p);
    SJ_LAUNCH_CHECK();
    uint64_t *pin = ctx->pinned;
    SJ_TRY_HIP(hipMemcpyAsync(pin, rbits.p, 4, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipStreamSynchronize(s));
    const uint32_t hb = (uint32_t)(pin[0] & 0xffffffffu);
    float rf;
    memcpy(&rf, &hb, 4);
    // the FP32 rounding of |v| may be below |v|: widen by one FP32 ulp
    const double R = (double)std::nextafter(rf, INFINITY);
    if (mf) {
        // FP32 copy for the matrix cores when the coordinates lie within
        // 1024 eps of the origin; FP64 centring otherwise
        const double Re = R * std::sqrt((double)g->d) * (1.0 + std::ldexp(1.0, -20));
        const bool src32 = Re <= 1024.0 * std::sqrt(a.T);
        mfma_consts(a, a.T, src32 ? Re : -1.0);
        if (src32) a.pts32 = g->sortedPts32.p;
        return SJ_OK;
    }
    Fp32Margin m = fp32_margin(a.T, R, g->d);
    if (m.usable) {
        a.pts32 = g->sortedPts32.p;
        a.lo32 = m.lo;
        a.hi32 = m.hi;
        a.T32 = m.t32;
    } else {
        *prec = SJ_PREC_FP64;  // band would cover the whole range
    }
    return SJ_OK;
}

sj_status join(sj_grid_s *g, const sj_join_params *P, sj_result_s *r) {
    const auto t_wall0 = std::chrono::steady_clock::now();
    if (!g || !P || !r) {
        set_error("join: null argument");
        return SJ_ERR_INVALID;
    }
    if (P->form < 0 || P->form > 3 || P->precision < 0 || P->precision > 1 || P->schedule < 0 ||
        P->schedule > 2) {
        set_error("join: invalid form/precision/schedule");
        return SJ_ERR_INVALID;
    }
    if (P->schedule == SJ_SCHED_THREADS_PER_POINT &&
        (P->form != SJ_FORM_FULL || P->precision != SJ_PREC_FP64)) {
        set_error("join: the threads-per-point schedule runs the full stencil in FP64 (reference mode 8)");
        return SJ_ERR_UNSUPPORTED;
    }
    if (P->block_size != 0 && P->block_size != JB) {
        set_error("join: block_size must be 0 or %d", JB);
        return SJ_ERR_UNSUPPORTED;
    }
    if (g->k > DEV_MAXK) {
        set_error("join: k=%d above the device stencil limit %d", g->k, DEV_MAXK);
        return SJ_ERR_UNSUPPORTED;
    }
    sj_context ctx = g->ctx;
    Pool *pool = &ctx->pool;
    hipStream_t s = ctx->stream;
    const uint64_t n = g->n;
    r->ctx = ctx;
    r->n = n;
    memset(&r->stats, 0, sizeof r->stats);
    const double T = eps_threshold(g->eps);
    JoinArgs a = base_args(g, P->form, T);
    if (P->cells_per_tile < 0) a.useTiles = 0;  // per-point walk only (diagnostics)
    EvTimer tw(s, ctx->tev[0], ctx->tev[1]), tc(s, ctx->tev[2], ctx->tev[3]), ts(s, ctx->tev[4], ctx->tev[5]),
        tf(s, ctx->tev[6], ctx->tev[7]), tr(s, ctx->tev[8], ctx->tev[9]);

    // precision form: FP32 copy relative to the grid origin + margin
    int prec = P->precision;
    if (prec == SJ_PREC_FP32_RECHECK) SJ_TRY(prepare_fp32(g, a, &prec));

    // threads per point: the points in workload order, no tiles
    DBuf<uint32_t> tppCellW, tppCellOrder, tppOrder;
    if (P->schedule == SJ_SCHED_THREADS_PER_POINT) {
        tw.start();
        SJ_TRY(workload_order_device(g, 0, tppCellW, tppCellOrder));
        SJ_TRY(tpp_point_order(g, tppCellOrder.p, tppOrder));
        tw.stop();
        a.tppOrder = tppOrder.p;
        a.useTiles = 0;
    }

    // row-aligned tiles; work queue of tiles in descending workload
    // the count pass's counters (own and mirrored hits per row, statistics,
    // queue counters), cleared by the tile compaction
    DBuf<uint32_t> own, mir;
    DBuf<uint64_t> scal;  // the host's scalars, the live statistics at SC_STAT
    SJ_TRY(own.alloc(pool, n));
    SJ_TRY(mir.alloc(pool, n));
    SJ_TRY(scal.alloc(pool, SC_WORDS));
    unsigned long long *const stat = reinterpret_cast<unsigned long long *>(scal.p) + SC_STAT;
    DBuf<uint32_t> tileStart, nTilesDev, tileOrder, qctr, qclass;
    const bool queue = P->schedule == SJ_SCHED_WORKQUEUE;
    if (queue) {
        SJ_TRY(qclass.alloc(pool, 2 * QB));
        SJ_TRY(qctr.alloc(pool, QSLICES * QSTRIDE));
    }
    const TileZero tz{queue ? qclass.p : nullptr, mir.p, n, stat, queue ? qctr.p : nullptr};
    SJ_TRY(build_tiles(g, a, 0u, (uint32_t)n, tileStart, nTilesDev, tz));
    a.tileStart = tileStart.p;
    a.nTilesDev = nTilesDev.p;
    // count pass (the hit record's directory, then the work queue, then
    // the count)
    a.ownCnt = own.p;
    a.mirCnt = mir.p;
    a.stat = stat;
    // hit record of the count pass for the fill, in space that persists
    // on the context; without the space the fill evaluates distances again
    DBuf<uint64_t> recDir, recMask, tileDir;
    DBuf<uint32_t> tileBound;
    DBuf<unsigned long long> recCtr;
    bool recording = P->hit_record >= 0 && record_walk(g, a) && n > 0;
    const uint64_t tBound = (n + 63) / 64 + std::min<uint64_t>(n, g->nNonEmpty);
    if (recording) {
        const uint64_t dirCap = std::max<uint64_t>(ctx->recDirCap, std::max<uint64_t>(1ull << 20, n));
        const uint64_t maskCap = std::max<uint64_t>(ctx->recMaskCap, std::max<uint64_t>(1ull << 23, 16 * n));
        if (tBound > 0xFFFFFFFFull || recDir.alloc(pool, (uint64_t)REC_DW * dirCap) != SJ_OK ||
            recMask.alloc(pool, maskCap) != SJ_OK || tileDir.alloc(pool, tBound + 1) != SJ_OK ||
            tileBound.alloc(pool, tBound) != SJ_OK || recCtr.alloc(pool, REC_NCTR) != SJ_OK) {
            recording = false;  // no room: the fill evaluates distances again
            (void)hipGetLastError();
        } else {
            const uint32_t cw = a.mDP > 0 ? 32u : 64u;  // candidates per recorded piece
            // the record's counters cleared by the same launch
            hipLaunchKernelGGL(pick_tile_bound(g->k), dim3(nblk(std::max<uint64_t>(tBound, REC_NCTR))), dim3(JB), 0,
                               s, a, (uint32_t)tBound, cw, tileBound.p, recCtr.p);
            SJ_LAUNCH_CHECK();
            SJ_TRY(scan_exclusive_u32_to_u64(pool, s, tileBound.p, tBound, tileDir.p));
            a.recDir = recDir.p;
            a.recMask = recMask.p;
            a.tileDir = tileDir.p;
            a.recCtr = recCtr.p;
            a.dirCap = dirCap;
            a.maskCap = maskCap;
        }
    }
    if (queue) {
        tw.start();
        SJ_TRY(build_tile_queue(g, a, n, tileOrder, qclass.p));
        tw.stop();
        a.tileOrder = tileOrder.p;
        a.qctr = qctr.p;
        tc.start_at(tw);  // nothing runs between the queue and the count
    } else {
        tc.start();
    }
    SJ_TRY(launch_join(g, a, PASS_COUNT, prec));
    tc.stop();

    // row offsets in cell order (the caller's array when it gave one)
    ts.start_at(tc);  // nothing runs between the count and the scan
    if (r->extOffsets) {
        r->d_offsets.p = r->extOffsets;
        r->d_offsets.n = n + 1;
        r->d_offsets.pool = nullptr;
    } else {
        SJ_TRY(r->d_offsets.alloc(pool, n + 1));
    }
    SJ_TRY(scan_exclusive_sum2_u32_to_u64(pool, s, own.p, mir.p, n, r->d_offsets.p));
    ts.stop();
    // |S|, the statistics and the record's state, gathered for one readback.
    // When the context has seen results of about this size (idsHint) and no
    // batches are asked for, the fill goes ahead at once into a buffer of
    // that capacity, gated on the device (fill_gate: the record's state, the
    // result fitting); the host reads everything once at the end and fills
    // again only if the result outgrew the buffer.
    uint64_t *pin = ctx->pinned;
    const CountScalars csc{r->d_offsets.p + n, stat, recording ? recCtr.p : nullptr,
                           recording ? tileDir.p + tBound : nullptr, scal.p};
    const bool ahead = P->capacity == 0 && !r->extIds && ctx->idsHint > 0;
    // ahead: the fill's first cursor kernel gathers them
    if (!ahead) {
        hipLaunchKernelGGL(k_count_scalars, dim3(1), dim3(64), 0, s, csc);
        SJ_LAUNCH_CHECK();
    }
    bool replay = false;
    // the count pass's scalars (sc: k_count_scalars layout) into the result
    auto absorb = [&](const uint64_t *sc) -> sj_status {
        if (recording) {
            replay = sc[16 + REC_BAD] == 0;
            if (!replay) {
                // grow to the demand seen (at least doubling the mask space, whose
                // slices may fill unevenly), within a quarter of the free memory
                size_t freeB = 0, totalB = 0;
                if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) freeB = 0;
                const uint64_t wantDir = std::max<uint64_t>(sc[18], a.dirCap);
                std::vector<unsigned long long> hc(REC_NCTR);
                SJ_TRY_HIP(hipMemcpy(hc.data(), recCtr.p, REC_NCTR * 8, hipMemcpyDeviceToHost));
                uint64_t want = 0;
                for (uint32_t sl = 0; sl < REC_NSLICE; sl++) want += hc[REC_SLICE0 + sl * REC_STRIDE + 1];
                const uint64_t wantMask = std::max<uint64_t>(want + want / 4 + (uint64_t)REC_NSLICE * REC_BLK * 4,
                                                             sc[18] > a.dirCap ? a.maskCap : 2 * a.maskCap);
                if ((wantDir * 8 * REC_DW + wantMask * 8) <= freeB / 4) {
                    ctx->recDirCap = std::max(ctx->recDirCap, wantDir);
                    ctx->recMaskCap = std::max(ctx->recMaskCap, wantMask);
                }
            }
        }
        r->total = sc[0];
        r->stats.total_pairs = sc[0];
        r->stats.candidates = sc[1 + STAT_CAND];
        r->stats.fp32_band_pairs = sc[1 + STAT_BAND];
        r->stats.fp32_misclassified = sc[1 + STAT_MIS];
        if (sc[1 + STAT_ERR]) {
            set_error("join: count pass met %llu inconsistent tiles", (unsigned long long)sc[1 + STAT_ERR]);
            return SJ_ERR_STATE;
        }
        return SJ_OK;
    };
    if (!ahead) {
        SJ_TRY_HIP(hipMemcpyAsync(pin, scal.p, 19 * 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        SJ_TRY(absorb(pin));
    }
    r->stats.overflow_retries = 0;
    // row IDs: a grid made for this call alone hands its sorted IDs over at
    // the end; otherwise they are copied (into the caller's array when it
    // gave one)
    if (r->extRowIds) {
        SJ_TRY_HIP(hipMemcpyAsync(r->extRowIds, g->sortedIds.p, n * 4, hipMemcpyDeviceToDevice, s));
        r->d_rowIds.p = r->extRowIds;
        r->d_rowIds.n = n;
        r->d_rowIds.pool = nullptr;
    } else if (!g->transient) {
        SJ_TRY(r->d_rowIds.alloc(pool, n));
        SJ_TRY_HIP(hipMemcpyAsync(r->d_rowIds.p, g->sortedIds.p, n * 4, hipMemcpyDeviceToDevice, s));
    }

    DBuf<unsigned long long> mirCur;
    SJ_TRY(mirCur.alloc(pool, n));
    a.mirCur = mirCur.p;
    a.offS = r->d_offsets.p;

    // the fill of the whole result into ids[cap]: the replayed LDS tile walk
    // places mirrored hits at the recorded reservations, every other fill
    // takes them from the cursors.  known: the record's state is known on the
    // host (replay); otherwise, when recording, only the replay is launched,
    // gated on the record being usable (the caller evaluates in a second
    // round when it was not).  gather: the cursor kernel also gathers the
    // count pass's scalars
    auto launch_fill = [&](uint64_t cap, bool known, bool gather) -> sj_status {
        a.ids = r->d_ids.p;
        a.outBase = 0;
        a.idsCap = cap;
        const bool gate = !known && recording;
        uint32_t *qz = a.tileOrder ? qctr.p : nullptr;  // queue counters, zeroed by the cursor kernel
        const unsigned nbq = (unsigned)std::max<uint64_t>(nblk(n), nblk(QSLICES * QSTRIDE));
        CountScalars gsc = csc;
        if (!gather) gsc.out = nullptr;
        if (!gate) {
            if (!(replay && a.mDP == 0)) {
                hipLaunchKernelGGL(k_mir_cursors, dim3(nbq), dim3(JB), 0, s, r->d_offsets.p, own.p, 0u,
                                   (uint32_t)n, 0ull, mirCur.p, nullptr, qz, gsc);
                SJ_LAUNCH_CHECK();
            } else {
                if (gather) {
                    hipLaunchKernelGGL(k_count_scalars, dim3(1), dim3(64), 0, s, csc);
                    SJ_LAUNCH_CHECK();
                }
                if (qz) SJ_TRY(fill_u32(s, qz, QSLICES * QSTRIDE, 0u));
            }
            SJ_TRY(launch_join(g, a, PASS_FILL, prec, replay));
        } else {
            hipLaunchKernelGGL(k_mir_cursors, dim3(nbq), dim3(JB), 0, s, r->d_offsets.p, own.p, 0u,
                               (uint32_t)n, 0ull, mirCur.p, a.mDP == 0 ? recCtr.p + REC_BAD : nullptr, qz, gsc);
            SJ_LAUNCH_CHECK();
            // replay when the record is usable; an unusable one (the
            // record's space fell short, seen on the host afterwards) is
            // evaluated in a second round
            JoinArgs ar = a;
            ar.recGate = recCtr.p + REC_BAD;
            ar.recGateMode = 1;
            SJ_TRY(launch_join(g, ar, PASS_FILL, prec, true));
        }
        return SJ_OK;
    };

    const bool whole = ahead || P->capacity == 0 || r->total <= P->capacity;
    double ms_fill = 0.0;
    if (whole) {
        bool known = !ahead;
        uint64_t cap = r->total;
        if (r->extIds) {
            if (r->total > r->extIdsCap) {
                set_error("join: %llu result pairs exceed the caller's capacity of %llu",
                          (unsigned long long)r->total, (unsigned long long)r->extIdsCap);
                return SJ_ERR_OVERFLOW;
            }
            r->d_ids.p = r->extIds;
            r->d_ids.n = r->total;
            r->d_ids.pool = nullptr;
        } else {
            if (ahead) cap = ctx->idsHint;
            SJ_TRY(r->d_ids.alloc(pool, cap));
        }
        for (int round = 0; round < 2; round++) {
            // from the first round's start to the last round's end; ahead,
            // nothing runs between the scan and the fill
            if (round == 0) {
                if (ahead)
                    tf.start_at(ts);
                else
                    tf.start();
            }
            SJ_TRY(launch_fill(known ? 0 : cap, known, round == 0 && ahead));
            tf.stop();
            if (P->sort_rows) {
                tr.start();
                SJ_TRY(sort_rows_device(ctx, r->d_offsets.p, r->d_ids.p, n, 0, known ? 0 : cap));
                tr.stop();
            }
            // the count's scalars and the fill's errors and overflow flag
            SJ_TRY_HIP(hipMemcpyAsync(pin, scal.p, SC_WORDS * 8, hipMemcpyDeviceToHost, s));
            SJ_TRY_HIP(hipStreamSynchronize(s));
            if (known) break;
            SJ_TRY(absorb(pin));
            // the gated replay skips an unusable record before its capacity
            // check, so the size is checked here too
            const bool outgrew = pin[SC_STAT + STAT_OVF] != 0 || r->total > cap;
            if (!outgrew && (replay || !recording)) break;
            // fill again: into a buffer of the result's size when it outgrew
            // this one, evaluating distances when the record was unusable
            known = true;
            if (outgrew) {
                r->stats.overflow_retries = 1;
                SJ_TRY(r->d_ids.alloc(pool, r->total));
            }
            SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(stat) + STAT_ERR, 1, 0ull));
        }
        if (pin[SC_STAT + STAT_ERR]) {
            set_error("join: fill pass disagreed with the count pass (%llu writes)",
                      (unsigned long long)pin[SC_STAT + STAT_ERR]);
            return SJ_ERR_STATE;
        }
        r->stats.fill_from_record = replay ? 1 : 0;
        r->on_device = true;
        r->stats.num_batches = 1;
        ms_fill = tf.ms();
        r->stats.ms_sort_rows = tr.ms();
    } else {
        // Batches: windows of rows in cell order, each within `capacity`
        // pairs.  Host copies land in pinned memory from the context's pool;
        // batch b's download on the copy stream overlaps the fill of batch
        // b+1, the two device buffers taken in turn.
        Pool *hp = &ctx->hpool;
        const uint64_t nne = g->nNonEmpty;
        DBuf<uint64_t> hOff, hLin;
        DBuf<uint32_t> hCellOf, hIds, hStart, hTile;
        SJ_TRY(hOff.alloc(hp, n + 1));
        SJ_TRY(hLin.alloc(hp, nne));
        SJ_TRY(hCellOf.alloc(hp, n));
        SJ_TRY(hIds.alloc(hp, n));
        SJ_TRY(hStart.alloc(hp, nne + 1));
        SJ_TRY_HIP(hipMemcpyAsync(hOff.p, r->d_offsets.p, (n + 1) * 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hLin.p, g->cellLinear.p, nne * 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hCellOf.p, g->cellOf.p, n * 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hIds.p, g->sortedIds.p, n * 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hStart.p, g->cellStart.p, (nne + 1) * 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(pin, nTilesDev.p, 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        // the count pass's tiles: a window's query range is a run of them
        const uint32_t nT = (uint32_t)(pin[0] & 0xffffffffu);
        SJ_TRY(hTile.alloc(hp, (uint64_t)nT + 1));
        SJ_TRY_HIP(hipMemcpyAsync(hTile.p, tileStart.p, (nT + 1) * 4ull, hipMemcpyDeviceToHost, s));
        std::vector<std::pair<uint32_t, uint32_t>> win;
        uint64_t w0 = 0;
        while (w0 < n) {
            uint64_t w1 = w0;
            while (w1 < n && hOff.p[w1 + 1] - hOff.p[w0] <= P->capacity) w1++;
            if (w1 == w0) {
                set_error("join: capacity %llu is below one neighbour row (%llu pairs)",
                          (unsigned long long)P->capacity, (unsigned long long)(hOff.p[w0 + 1] - hOff.p[w0]));
                return SJ_ERR_OVERFLOW;
            }
            win.emplace_back((uint32_t)w0, (uint32_t)w1);
            w0 = w1;
        }
        uint64_t Mmax = 0;
        for (int j = 0; j < g->k; j++) Mmax += g->gp.mult[j];
        if (!ctx->copyStream) SJ_TRY_HIP(hipStreamCreateWithFlags(&ctx->copyStream, hipStreamNonBlocking));
        hipStream_t cs = ctx->copyStream;
        struct Ev {
            hipEvent_t e = nullptr;
            Ev() { (void)hipEventCreateWithFlags(&e, hipEventDisableTiming); }
            ~Ev() {
                if (e) (void)hipEventDestroy(e);
            }
        } filled[2], copied[2];
        DBuf<int32_t> dbuf[2];
        DBuf<uint32_t> bCount[2];
        // declared after the buffers, so destroyed before them: every exit
        // from here, error paths included, waits for the downloads in flight
        // before the device buffers go back to the pool
        struct CopyDrain {
            hipStream_t s, cs;
            ~CopyDrain() {
                (void)hipStreamSynchronize(s);
                (void)hipStreamSynchronize(cs);
            }
        } drain{s, cs};
        for (int j = 0; j < 2; j++) {
            SJ_TRY(dbuf[j].alloc(pool, P->capacity));
            SJ_TRY(bCount[j].alloc(pool, 1));
        }
        std::vector<std::unique_ptr<EvTimer>> tfs;
        SJ_TRY_HIP(hipStreamSynchronize(s));  // hTile
        r->batches.resize(win.size());
        for (size_t bi = 0; bi < win.size(); bi++) {
            const int sl = (int)(bi & 1);
            const uint32_t s0 = win[bi].first, s1 = win[bi].second;
            // query range: the window's cells, widened to every cell that may
            // claim one of them (adjacent cells lie within Mmax linear IDs;
            // LID cells claim only larger linear IDs)
            uint64_t q0 = s0, q1 = s1;
            if (P->form != SJ_FORM_FULL) {
                const uint64_t linLo = hLin.p[hCellOf.p[s0]], linHi = hLin.p[hCellOf.p[s1 - 1]];
                const uint64_t lo = linLo > Mmax ? linLo - Mmax : 0;
                const uint64_t hi = P->form == SJ_FORM_LID ? linHi : linHi + Mmax;
                const uint64_t c0 = std::lower_bound(hLin.p, hLin.p + nne, lo) - hLin.p;
                const uint64_t c1 = std::upper_bound(hLin.p, hLin.p + nne, hi) - hLin.p;
                q0 = hStart.p[c0];
                q1 = hStart.p[c1];
            }
            // tiles [t0, t1) of the count pass cover [q0, q1); their points
            // outside the window emit nothing into it
            const uint32_t t0 = (uint32_t)(std::upper_bound(hTile.p, hTile.p + nT + 1, (uint32_t)q0) - hTile.p) - 1u;
            const uint32_t t1 = (uint32_t)(std::lower_bound(hTile.p, hTile.p + nT + 1, (uint32_t)q1) - hTile.p);
            const uint64_t base = hOff.p[s0];
            hipLaunchKernelGGL(k_mir_cursors, dim3(nblk(s1 - s0)), dim3(JB), 0, s, r->d_offsets.p, own.p, s0, s1,
                               (unsigned long long)base, mirCur.p, nullptr, nullptr, CountScalars{});
            SJ_LAUNCH_CHECK();
            // the buffer's previous download must be done before it is refilled
            if (bi >= 2) SJ_TRY_HIP(hipStreamWaitEvent(s, copied[sl].e, 0));
            SJ_TRY(fill_u32(s, bCount[sl].p, 1, t1 - t0));
            JoinArgs ab = a;
            ab.tileOrder = nullptr;  // batches walk their tiles in order
            ab.tileStart = tileStart.p + t0;
            ab.tileDir = a.tileDir ? a.tileDir + t0 : nullptr;
            ab.nTilesDev = bCount[sl].p;
            ab.qBegin = hTile.p[t0];
            ab.qEnd = hTile.p[t1];
            ab.wBegin = s0;
            ab.wEnd = s1;
            ab.outBase = base;
            ab.ids = dbuf[sl].p;
            tfs.emplace_back(new EvTimer(s));
            tfs.back()->start();
            SJ_TRY(launch_join(g, ab, PASS_FILL, prec, replay));
            tfs.back()->stop();
            if (P->sort_rows) SJ_TRY(sort_rows_device(ctx, r->d_offsets.p + s0, dbuf[sl].p, s1 - s0, base));
            SJ_TRY_HIP(hipEventRecord(filled[sl].e, s));
            const uint64_t cnt = hOff.p[s1] - base;
            sj_batch_host &B = r->batches[bi];
            SJ_TRY(B.ids.alloc(hp, cnt));
            SJ_TRY(B.rowIds.alloc(hp, (size_t)(s1 - s0)));
            SJ_TRY(B.rowOffsets.alloc(hp, (size_t)(s1 - s0) + 1));
            std::copy(hIds.p + s0, hIds.p + s1, B.rowIds.p);
            for (uint32_t q = s0; q <= s1; q++) B.rowOffsets.p[q - s0] = hOff.p[q] - base;
            SJ_TRY_HIP(hipStreamWaitEvent(cs, filled[sl].e, 0));
            SJ_TRY_HIP(hipMemcpyAsync(B.ids.p, dbuf[sl].p, cnt * 4, hipMemcpyDeviceToHost, cs));
            SJ_TRY_HIP(hipEventRecord(copied[sl].e, cs));
        }
        SJ_TRY_HIP(hipStreamSynchronize(cs));
        for (auto &t : tfs) ms_fill += t->ms();
        SJ_TRY_HIP(hipMemcpyAsync(pin, stat + STAT_ERR, 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        if (pin[0]) {
            set_error("join: fill pass disagreed with the count pass (%llu writes)",
                      (unsigned long long)pin[0]);
            return SJ_ERR_STATE;
        }
        r->on_device = false;
        r->d_ids.reset();
        r->stats.num_batches = win.size();
        r->stats.fill_from_record = replay ? 1 : 0;
    }
    if (g->transient && !r->extRowIds) r->d_rowIds = std::move(g->sortedIds);
    // the next join on this context fills ahead into a buffer of about this size
    if (!r->extIds) ctx->idsHint = std::max<uint64_t>(ctx->idsHint, r->total + r->total / 16 + 4096);
    r->stats.ms_workload = tw.ms();
    r->stats.ms_count = tc.ms();
    r->stats.ms_scan = ts.ms();
    r->stats.ms_fill = ms_fill;
    r->stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();
    return SJ_OK;
}

// ------------------------------------------------------- canonical rows
// Rows sorted ascending, one wave per row: rows up to 1024 entries by a
// bitonic sort in LDS, longer rows by a serial shell sort (rare).
constexpr int SR_CAP = 1024;

__global__ __launch_bounds__(256) void k_sort_rows(const uint64_t *offsets, int32_t *ids, uint64_t n,
                                                   uint64_t base, uint64_t idsCap) {
    __shared__ int32_t buf[4][SR_CAP];
    const unsigned wave = threadIdx.x >> 6, lane = __lane_id();
    const uint64_t row = (uint64_t)blockIdx.x * 4 + wave;
    if (row >= n) return;
    if (idsCap && offsets[n] - base > idsCap) return;  // the fill wrote nothing
    const uint64_t b = offsets[row] - base, e = offsets[row + 1] - base;
    const uint64_t len = e - b;
    if (len < 2) return;
    if (len <= SR_CAP) {
        uint32_t p2 = 2;
        while (p2 < len) p2 <<= 1;
        for (uint32_t i = lane; i < p2; i += 64) buf[wave][i] = i < len ? ids[b + i] : INT32_MAX;
        wave_lds_sync();
        for (uint32_t kk = 2; kk <= p2; kk <<= 1) {
            for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
                for (uint32_t i = lane; i < p2; i += 64) {
                    const uint32_t l = i ^ jj;
                    if (l > i) {
                        const int32_t x = buf[wave][i], y = buf[wave][l];
                        const bool up = (i & kk) == 0;
                        if ((x > y) == up) {
                            buf[wave][i] = y;
                            buf[wave][l] = x;
                        }
                    }
                }
                wave_lds_sync();
            }
        }
        for (uint32_t i = lane; i < len; i += 64) ids[b + i] = buf[wave][i];
    } else if (lane == 0) {
        int32_t *v = ids + b;
        for (uint64_t gap = len / 2; gap > 0; gap /= 2)
            for (uint64_t i = gap; i < len; i++) {
                int32_t x = v[i];
                uint64_t j = i;
                for (; j >= gap && v[j - gap] > x; j -= gap) v[j] = v[j - gap];
                v[j] = x;
            }
    }
}

sj_status sort_rows_device(sj_context ctx, const uint64_t *offsets, int32_t *ids, uint64_t n, uint64_t base,
                           uint64_t idsCap) {
    if (!n) return SJ_OK;
    hipLaunchKernelGGL(k_sort_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, offsets, ids, n,
                       base, idsCap);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// Original-ID-order CSR from the cell-order table (on demand).
__global__ void k_row_lengths_orig(const uint64_t *off, const uint32_t *rowIds, uint64_t n, uint32_t *len) {
    uint64_t r = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (r >= n) return;
    len[rowIds[r]] = (uint32_t)(off[r + 1] - off[r]);
}

__global__ __launch_bounds__(256) void k_copy_rows_orig(const uint64_t *off, const int32_t *ids,
                                                        const uint32_t *rowIds, uint64_t n,
                                                        const uint64_t *offO, int32_t *idsO) {
    const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint64_t b = off[r], e = off[r + 1];
    const uint64_t d = offO[rowIds[r]];
    for (uint64_t i = b + __lane_id(); i < e; i += 64) idsO[d + (i - b)] = ids[i];
}

sj_status result_to_orig(sj_result_s *r) {
    if (r->d_offsetsOrig.p) return SJ_OK;
    Pool *pool = &r->ctx->pool;
    hipStream_t s = r->ctx->stream;
    const uint64_t n = r->n;
    DBuf<uint32_t> len;
    SJ_TRY(len.alloc(pool, n));
    SJ_TRY(r->d_offsetsOrig.alloc(pool, n + 1));
    SJ_TRY(r->d_idsOrig.alloc(pool, r->total));
    hipLaunchKernelGGL(k_row_lengths_orig, dim3(nblk(n)), dim3(JB), 0, s, r->d_offsets.p, r->d_rowIds.p, n,
                       len.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u64(pool, s, len.p, n, r->d_offsetsOrig.p));
    hipLaunchKernelGGL(k_copy_rows_orig, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, r->d_offsets.p,
                       r->d_ids.p, r->d_rowIds.p, n, r->d_offsetsOrig.p, r->d_idsOrig.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY_HIP(hipStreamSynchronize(s));
    return SJ_OK;
}

}  // namespace sj
