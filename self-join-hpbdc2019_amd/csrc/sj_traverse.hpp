// Device-side definitions shared by the traversal kernels: launch
// constants, the argument block, lane/wave helpers, cell lookups, the
// stencil claim rules and the canonical FP64 predicate.
#pragma once

#include "sj_common.hpp"

namespace sj {

constexpr int JB = 256;         // traversal block size (4 waves)
constexpr int WPB = JB / 64;    // waves per block
constexpr int DEV_MAXK = 8;     // indexed dims handled by the device stencil
constexpr int QSLICES = 8;      // work-queue slices (one counter each)
constexpr int QGROUP = 8;       // consecutive queue entries a slice takes together
constexpr int QSTRIDE = 32;     // u32 words between slice counters (128 B)
enum { PASS_COUNT = 0, PASS_FILL = 1 };
enum { STAT_CAND = 0, STAT_BAND = 1, STAT_MIS = 2, STAT_ERR = 3, STAT_FALLBACK = 4, STAT_OVF = 6 };

struct JoinArgs {
    const double *pts;        // n x d, cell order
    const float *pts32;       // n x d, (x - origin) in FP32, or null
    const uint32_t *origId;   // original ID of each sorted position
    const uint32_t *cellOf;
    const uint32_t *cellStart;
    const uint64_t *cellLinear;
    const uint32_t *cellCoord;
    const uint32_t *posLB;    // dense grids: first position with linear ID >= L
    uint64_t nne;
    uint32_t nCells[DEV_MAXK];
    int64_t mult[DEV_MAXK];
    int d, k, form, useTiles;
    double T;
    float lo32, hi32, T32, pad1;
    double cellW;  // grid cell width (eps)
    // matrix-core FP32 form (mDP > 0): error bound mAlpha * s + mBeta,
    // certain hit at or below mTlo, certain miss above mThi
    float mAlpha, mBeta, mTlo, mThi, mT32;
    int mDP, mSrc32;  // mSrc32: stage from pts32 (FP32 copy, rows of mDP floats)
    // query range (sorted positions) and tiles of 64 over it
    uint32_t qBegin, qEnd;
    const uint32_t *tileStart;  // tile t = sorted positions [tileStart[t], tileStart[t+1])
    const uint32_t *nTilesDev;  // number of tiles (device)
    const uint32_t *tileOrder;  // work queue order of tiles, or null
    uint32_t *qctr;
    const uint32_t *tppOrder;   // threads-per-point schedule: points in workload order, or null
    // fill launched before the host read the count pass's results: the
    // result must fit idsCap IDs (0: no check), and recGate (when set)
    // runs the pass only if the hit record is usable (recGateMode 1) or
    // unusable (recGateMode 2)
    uint64_t idsCap;
    const unsigned long long *recGate;
    int recGateMode, pad2;
    // rows written by this launch (sorted positions)
    uint32_t wBegin, wEnd;
    // count pass
    uint32_t *ownCnt, *mirCnt;
    unsigned long long *stat;
    // fill pass: row r of the window starts at offS[r] - outBase
    const uint64_t *offS;
    uint64_t outBase;
    unsigned long long *mirCur;
    int32_t *ids;
    // hit record (LDS tile walk): the count pass stores, per executed
    // 64-candidate piece, a directory entry {lanes with hits, first mask
    // word} and the hit masks of those lanes back to back; the fill replays
    // them instead of evaluating distances again.  recDir null: no record.
    uint64_t *recDir;              // REC_DW words per entry
    uint64_t *recMask;
    const uint64_t *tileDir;       // tile t: entries [tileDir[t], tileDir[t+1])
    unsigned long long *recCtr;    // REC_* counters
    uint64_t dirCap, maskCap;      // entries, words
};

// Hit-record counters: record unusable (any cause), then REC_NSLICE
// slices REC_STRIDE words apart, each {blocks taken, mask words wanted}.  A
// wave takes blocks of REC_BLK mask words from slice (wave id mod
// REC_NSLICE); block b of slice c is global block b * REC_NSLICE + c.
enum { REC_BAD = 0, REC_SLICE0 = 32 };
constexpr uint32_t REC_BLK = 512;
// Directory entry: {lanes with hits, first mask word, 1 + first word of the
// chunk's mirror reservations (first piece of a chunk only; 0: none), 0}
constexpr uint32_t REC_DW = 4;
constexpr uint32_t REC_NSLICE = 64;
constexpr uint32_t REC_STRIDE = 32;
constexpr uint32_t REC_NCTR = REC_SLICE0 + REC_NSLICE * REC_STRIDE;

// Per-wave state of the hit record: the wave's current block of mask words.
struct TileRec {
    uint64_t blk = 0;
    uint32_t used = REC_BLK;
    bool ok = true;
    uint64_t want = 0;
};

__device__ __forceinline__ uint32_t popc_below64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Per-lane statistics, reduced across the wave once per kernel: per-lane
// atomics on one address would serialise in L2.
struct LaneStats {
    uint64_t cand = 0, band = 0, mis = 0, err = 0, fallback = 0;
};

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Bits i in [0,64) with start + i in [A, B).  Branch-free: both ends
// clamped to [0, 64] with unsigned differences, then one run of w bits.
__device__ __forceinline__ uint64_t range_bits(uint32_t start, uint32_t A, uint32_t B) {
    const uint32_t lo = min(A > start ? A - start : 0u, 64u);
    const uint32_t hi = min(B > start ? B - start : 0u, 64u);
    const uint32_t w = hi > lo ? hi - lo : 0u;
    return w ? (~0ull >> (64u - w)) << lo : 0ull;  // w > 0: lo < 64
}

// Largest v over the wave's active lanes, for v < 2^BITS: the bits of the
// result from the top down, each kept when some lane reaches it (ballots
// only, no cross-lane data movement).
template <int BITS>
__device__ __forceinline__ uint32_t wave_max_small(uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (int b = BITS - 1; b >= 0; b--) {
        const uint32_t t = r | (1u << b);
        if (__ballot(v >= t)) r = t;
    }
    return r;
}

// Value of v in lane l (l wave-uniform).
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
    return ((uint64_t)readlane_u32((uint32_t)(v >> 32), l) << 32) | readlane_u32((uint32_t)v, l);
}

// v as the value of the wave's first active lane, held in scalar registers
__device__ __forceinline__ double uniform_f64(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

constexpr unsigned long long NO_SLOT = ~0ull;

// Room for nz more record words in the wave's current block, taking a new
// block from the wave's counter slice when needed; false when the record
// space is exhausted (the record is then unusable).
__device__ __forceinline__ bool rec_room(const JoinArgs &a, TileRec &R, uint32_t nz) {
    if (R.ok && R.used + nz > REC_BLK) {
        const uint32_t sl = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % REC_NSLICE;
        uint64_t b = 0;
        if (__lane_id() == 0) b = atomicAdd(&a.recCtr[REC_SLICE0 + sl * REC_STRIDE], 1ull);
        R.blk = (readlane_u64(b, 0) * REC_NSLICE + sl) * REC_BLK;
        R.used = 0;
        R.ok = R.blk + REC_BLK <= a.maskCap;
        if (!R.ok && __lane_id() == 0) atomicAdd(&a.recCtr[REC_BAD], 1ull);
    }
    return R.ok;
}

// Count pass, one executed piece it of a tile whose entries are
// [dbase, dend): directory entry {lanes with hits, first mask word, 0, 0}
// and the masks of the lanes with hits, back to back.
__device__ __forceinline__ void record_piece(const JoinArgs &a, TileRec &R, bool &rec, uint64_t dbase,
                                             uint64_t dend, uint64_t it, uint64_t mask) {
    const unsigned lane = __lane_id();
    const uint64_t pres = __ballot(mask != 0ull);
    const uint32_t nz = (uint32_t)__popcll(pres);
    R.want += nz;
    if (rec && dbase + it >= dend) {
        rec = false;
        if (lane == 0) atomicAdd(&a.recCtr[REC_BAD], 1ull);
    }
    if (rec && rec_room(a, R, nz)) {
        const uint64_t off = R.blk + R.used;
        if (mask) a.recMask[off + popc_below64(pres)] = mask;
        R.used += nz;
        if (lane == 0) {
            uint64_t *e = a.recDir + REC_DW * (dbase + it);
            e[0] = pres;
            e[1] = off;
            e[2] = 0ull;
            e[3] = 0ull;
        }
    }
}

// Entry i of queue slice sl of ns: slices take groups of QGROUP
// consecutive entries in turn, so neighbouring tiles (in cell order inside
// a cost bucket) stay on the XCD that serves the slice.
__device__ __forceinline__ uint64_t queue_entry(uint64_t i, uint32_t sl, uint32_t ns) {
    return (sl + (i / QGROUP) * ns) * QGROUP + i % QGROUP;
}

__device__ __forceinline__ void flush_stats(const JoinArgs &a, const LaneStats &st) {
    const uint64_t c = wave_sum_u64(st.cand), b = wave_sum_u64(st.band), m = wave_sum_u64(st.mis);
    const uint64_t e = wave_sum_u64(st.err), f = wave_sum_u64(st.fallback);
    if (__lane_id() == 0) {
        if (c) atomicAdd(&a.stat[STAT_CAND], (unsigned long long)c);
        if (b) atomicAdd(&a.stat[STAT_BAND], (unsigned long long)b);
        if (m) atomicAdd(&a.stat[STAT_MIS], (unsigned long long)m);
        if (e) atomicAdd(&a.stat[STAT_ERR], (unsigned long long)e);
        if (f) atomicAdd(&a.stat[STAT_FALLBACK], (unsigned long long)f);
    }
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int64_t find_cell(const JoinArgs &a, uint64_t lin) {
    if (a.posLB) {
        const uint32_t p0 = a.posLB[lin];
        return p0 < a.posLB[lin + 1] ? (int64_t)a.cellOf[p0] : -1;
    }
    uint64_t lo = 0, hi = a.nne;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (a.cellLinear[mid] < lin) lo = mid + 1; else hi = mid;
    }
    return (lo < a.nne && a.cellLinear[lo] == lin) ? (int64_t)lo : -1;
}

// Whether a fill launched ahead of the host runs: its result fits the ids
// buffer (else STAT_OVF is raised and nothing is written) and the hit
// record is in the state the pass was launched for.  Uniform per launch.
__device__ __forceinline__ bool fill_gate(const JoinArgs &a) {
    if (a.recGate) {
        const bool bad = *a.recGate != 0ull;
        if (bad == (a.recGateMode == 1)) return false;
    }
    if (a.idsCap && a.offS[a.wEnd] - a.outBase > a.idsCap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.stat[STAT_OVF], 1ull);
        return false;
    }
    return true;
}

// first sorted position whose cell has linear ID >= lin
__device__ __forceinline__ uint32_t pos_lb(const JoinArgs &a, uint64_t lin);

// DN: the grid has the dense table (kernels built for it alone carry no
// binary search)
template <bool DN>
__device__ __forceinline__ uint32_t pos_lb_t(const JoinArgs &a, uint64_t lin) {
    if constexpr (DN) return a.posLB[lin];
    else return pos_lb(a, lin);
}
template <bool DN>
__device__ __forceinline__ int64_t find_cell_t(const JoinArgs &a, uint64_t lin) {
    if constexpr (DN) {
        const uint32_t p0 = a.posLB[lin];
        return p0 < a.posLB[lin + 1] ? (int64_t)a.cellOf[p0] : -1;
    } else {
        return find_cell(a, lin);
    }
}

__device__ __forceinline__ uint32_t pos_lb(const JoinArgs &a, uint64_t lin) {
    if (a.posLB) return a.posLB[lin];
    uint64_t lo = 0, hi = a.nne;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (a.cellLinear[mid] < lin) lo = mid + 1; else hi = mid;
    }
    return a.cellStart[lo];
}

// Stencil claim rules: FULL every neighbour (kernel.cu:1076-1093); UNICOMP
// stamp (stamploops.h): the highest differing dimension's coordinate is
// odd; B-UNICOMP checkerboard (kernel.cu:1346-1392; other k use the
// balanced rule [off >lex 0] xor ([sum off even] and parity(sum c)));
// LID: neighbour linear ID larger (kernel.cu:1513-1536).
template <int K>
__device__ __forceinline__ bool claims(int form, const uint32_t (&c)[K], const int (&o)[K],
                                       uint64_t linA, uint64_t linB) {
    if (form == SJ_FORM_FULL) return true;
    if (form == SJ_FORM_LID) return linB > linA;
    if (form == SJ_FORM_UNICOMP) {
        bool found = false, odd = false;
#pragma unroll
        for (int j = K - 1; j >= 0; j--) {
            if (!found && o[j] != 0) {
                odd = (c[j] & 1u) != 0;
                found = true;
            }
        }
        return odd;
    }
    uint32_t sc = 0;
#pragma unroll
    for (int j = 0; j < K; j++) sc += c[j];
    const bool odd = sc & 1u;
    if constexpr (K == 2) {
        if (odd) return o[1] >= 0 && o[0] != 0;
        return o[1] > 0 || (o[1] < 0 && o[0] == 0);
    } else if constexpr (K == 3) {
        if (odd)
            return (o[0] != 0 && o[1] >= 0 && o[2] >= 0) || (o[0] == 0 && o[1] >= 0 && o[2] > 0) ||
                   (o[1] > 0 && o[2] < 0);
        return (o[1] > 0 && o[2] >= 0) || (o[0] == 0 && o[1] < 0 && o[2] >= 0) ||
               (o[0] != 0 && o[1] > 0 && o[2] < 0) || (o[1] == 0 && o[2] > 0);
    } else {
        int so = 0;
        bool pos = false, found = false;
#pragma unroll
        for (int j = K - 1; j >= 0; j--) {
            so += o[j];
            if (!found && o[j] != 0) {
                pos = o[j] > 0;
                found = true;
            }
        }
        const bool even = (so & 1) == 0;
        return pos != (even && odd);
    }
}

// Runtime-k form of the claim rules of claims<K>.
__device__ __forceinline__ bool claims_rt(int form, int k, const uint32_t (&c)[DEV_MAXK],
                                          const int (&o)[DEV_MAXK], uint64_t linA, uint64_t linB) {
    if (form == SJ_FORM_FULL) return true;
    if (form == SJ_FORM_LID) return linB > linA;
    if (form == SJ_FORM_UNICOMP) {
        bool found = false, odd = false;
#pragma unroll
        for (int j = DEV_MAXK - 1; j >= 0; j--) {
            if (j < k && !found && o[j] != 0) {
                odd = (c[j] & 1u) != 0;
                found = true;
            }
        }
        return odd;
    }
    uint32_t sc = 0;
#pragma unroll
    for (int j = 0; j < DEV_MAXK; j++)
        if (j < k) sc += c[j];
    const bool odd = sc & 1u;
    if (k == 2) {
        if (odd) return o[1] >= 0 && o[0] != 0;
        return o[1] > 0 || (o[1] < 0 && o[0] == 0);
    }
    if (k == 3) {
        if (odd)
            return (o[0] != 0 && o[1] >= 0 && o[2] >= 0) || (o[0] == 0 && o[1] >= 0 && o[2] > 0) ||
                   (o[1] > 0 && o[2] < 0);
        return (o[1] > 0 && o[2] >= 0) || (o[0] == 0 && o[1] < 0 && o[2] >= 0) ||
               (o[0] != 0 && o[1] > 0 && o[2] < 0) || (o[1] == 0 && o[2] > 0);
    }
    int so = 0;
    bool pos = false, found = false;
#pragma unroll
    for (int j = DEV_MAXK - 1; j >= 0; j--) {
        if (j < k) {
            so += o[j];
            if (!found && o[j] != 0) {
                pos = o[j] > 0;
                found = true;
            }
        }
    }
    const bool even = (so & 1) == 0;
    return pos != (even && odd);
}

// Band of the FP32 direct-difference distance (VALU walks).  Stored
// y = fl32(fl64(x - o)) with every |x_j - o_j| <= R, so
// |y_j - (x_j - o_j)| <= E0 = 2R(u32 + 2u64); a difference of such values is
// off by at most E = E0 + u32(|a| + E0).  The FMA sum of D squares is within
// gamma_D of the sum of squared differences, which is within
// E(2 sqrt(D)|a| + D E) of |a|^2, and the canonical FP64 sum is within
// (D+3) u64 |a|^2 of it.  Hence the canonical value is <= T if the FP32 sum
// is < lo and > T if it is > hi; between lo and hi the FP64 predicate
// decides.  usable is false when the band would exceed T/4.
struct Fp32Margin {
    bool usable;
    float lo, hi, t32;
};

__host__ __device__ inline Fp32Margin fp32_margin(double T, double R, int D) {
    const double u32 = ldexp(1.0, -24), u64 = ldexp(1.0, -53);
    const double Rh = R * (1.0 + ldexp(1.0, -20)) + ldexp(1.0, -126);
    const double E0 = 2.0 * Rh * (u32 + 2.0 * u64);
    const double gam = D * u32 / (1.0 - D * u32);
    const double rho = (D + 3) * u64 * 1.01;
    const double sD = sqrt((double)D);
    const double tiny = D * ldexp(1.0, -126) * 8.0;  // possible FP32 denormal flushes
    const double Ahi = T / (1.0 - rho);
    const double nhi = sqrt(Ahi);
    const double Ehi = E0 + u32 * (nhi + E0);
    const double up = (1.0 + gam) * (Ahi + Ehi * (2.0 * sD * nhi + D * Ehi)) + tiny;
    const double Alo = T / (1.0 + rho);
    const double nlo = sqrt(Alo);
    const double Elo = E0 + u32 * (nlo + E0);
    const double dn = (1.0 - gam) * (Alo - 2.0 * Elo * sD * nlo) - tiny;
    const double M = fmax(up - T, T - dn) * 1.05;
    Fp32Margin m;
    m.usable = __builtin_isfinite(M) && M < 0.25 * T && (T + M) < 1e37 && R < 1e30;
    const double lo = T - M, hi = T + M;
    float flo = (float)lo, fhi = (float)hi;
    if ((double)flo > lo) flo = nextafterf(flo, -INFINITY);
    if ((double)fhi < hi) fhi = nextafterf(fhi, INFINITY);
    m.lo = flo;
    m.hi = fhi;
    m.t32 = (float)T;
    return m;
}

// FP64 canonical predicate: dimension order, explicit fma.
template <int D>
__device__ __forceinline__ double d2_fp64(const double *p, const double *q, int d) {
    double acc = 0.0;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            const double df = q[j] - p[j];
            acc = j == 0 ? df * df : fma(df, df, acc);
        }
    } else {
        for (int j = 0; j < d; j++) {
            const double df = q[j] - p[j];
            acc = j == 0 ? df * df : fma(df, df, acc);
        }
    }
    return acc;
}

// ------------------------------------------------- matrix-core path
constexpr int MFMA_MAX_D = 64;  // FP32+recheck joins with d <= MFMA_MAX_D run on the matrix cores

// Padded dims of the matrix-core operands (4, 8, 16, 32 or 64), 0 above
// MFMA_MAX_D.
int mfma_padded_dims(int d);
// Decision constants of the matrix-core walk: with s = |q|^2 + |c|^2 and
// M = alpha s + beta, d2 + M <= tlo is a certain hit and d2 - M > thi a
// certain miss of the canonical FP64 predicate (derivation in sj_mfma.hip).
struct MfmaBand {
    float alpha, beta, tlo, thi, t32;
    int DP;
};
MfmaBand mfma_band(double T, double R, int d);
// Fills the matrix-core fields of a (mAlpha .. mSrc32) for threshold T;
// R >= 0: staging from the FP32 copy whose coordinates are within R of the
// origin (Euclidean), R < 0: staging centred in FP64.
void mfma_consts(JoinArgs &a, double T, double R);
// replay: the fill replays the count pass's hit record
sj_status launch_join_mfma(const JoinArgs &a, int pass, int cu_count, hipStream_t s, bool replay = false);

// Own and mirrored hits of one staged chunk of at most 32 candidates at
// sorted positions [cb, cb + nv).  hit: this lane's hits (bit c =
// candidate cb + c); mirBits: those whose mirrored pair goes to the
// candidate's row.  Lid, Lcnt, Lbase: the chunk's IDs, mirrored-hit
// counters (zero on entry) and reserved bases in LDS.  Own hits leave one
// source lane at a time, four lanes per round, as contiguous row writes;
// mirrored hits are counted with one atomic per candidate (count pass) or
// reserved with one atomic per candidate and placed through the LDS
// counter (fill pass).
template <int PASS>
__device__ __forceinline__ void emit_chunk32(const JoinArgs &a, uint32_t cb, uint32_t nv, uint32_t hit,
                                             uint32_t mirBits, bool ownW, int32_t myId, uint64_t ownPos,
                                             uint64_t ownLim, uint32_t &own, const int32_t *Lid, uint32_t *Lcnt,
                                             unsigned long long *Lbase, bool anyMirror, uint64_t &err) {
    const unsigned lane = __lane_id();
    const uint32_t myCnt = (uint32_t)__popc(hit);
    uint32_t ownBits = 0;
    if (PASS == PASS_FILL && ownW) {
        // more hits than the row has room for: count and fill disagree
        if (ownPos + own + myCnt <= ownLim) ownBits = hit;
        else err++;
    }
    const uint64_t wcur = ownPos + own;
    const uint32_t lc = min(lane, 31u);
    uint64_t todo = __ballot(ownBits != 0u || mirBits != 0u);
    while (todo) {
        int l[4];
        bool lv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            lv[u] = todo != 0ull;
            l[u] = lv[u] ? __builtin_ctzll(todo) : 0;
            todo &= todo - 1ull;
        }
        uint32_t oB[4], mB[4];
        uint64_t w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            oB[u] = lv[u] ? readlane_u32(ownBits, l[u]) : 0u;
            mB[u] = lv[u] ? readlane_u32(mirBits, l[u]) : 0u;
            w[u] = readlane_u64(wcur, l[u]) + (uint32_t)__popc(oB[u] & ((1u << lc) - 1u));
        }
        const int32_t idv = Lid[lc];
        if (lane < 32) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (PASS == PASS_FILL && ((oB[u] >> lane) & 1u)) a.ids[w[u]] = idv;
                if ((mB[u] >> lane) & 1u) atomicAdd(&Lcnt[lane], 1u);
            }
        }
    }
    if (ownW) own += myCnt;
    if (!anyMirror) return;
    wave_lds_sync();
    if (PASS == PASS_COUNT) {
        if (lane < nv && Lcnt[lane]) atomicAdd(&a.mirCnt[cb + lane], Lcnt[lane]);
        return;
    }
    if (lane < nv) {
        const uint32_t v = Lcnt[lane];
        if (v) {
            unsigned long long b = atomicAdd(&a.mirCur[cb + lane], (unsigned long long)v);
            if (b + v > a.offS[cb + lane + 1] - a.outBase) {
                err++;
                b = NO_SLOT;
            }
            Lbase[lane] = b;
        }
        Lcnt[lane] = 0u;
    }
    wave_lds_sync();
    uint64_t todo2 = __ballot(mirBits != 0u);
    while (todo2) {
        const int l = __builtin_ctzll(todo2);
        todo2 &= todo2 - 1ull;
        const uint32_t mB = readlane_u32(mirBits, l);
        const int32_t qid = (int32_t)readlane_u32((uint32_t)myId, l);
        if (lane < 32 && ((mB >> lane) & 1u)) {
            const uint32_t k2 = atomicAdd(&Lcnt[lane], 1u);
            const unsigned long long b = Lbase[lane];
            if (b != NO_SLOT) a.ids[b + k2] = qid;
        }
    }
}

// ------------------------------------------------- host helpers (sj_join.hip)
double eps_threshold(double eps);
JoinArgs base_args(sj_grid_s *g, int form, double T);
// Counters the tile compaction clears for the join that follows (each when
// set): the work queue's 2 * 256 class counters, n mirrored-hit counts, 8
// statistics, the queue slice counters.
struct TileZero {
    uint32_t *classes = nullptr;
    uint32_t *mir = nullptr;
    uint64_t n = 0;
    unsigned long long *stat = nullptr;
    uint32_t *qctr = nullptr;
};
sj_status build_tiles(sj_grid_s *g, const JoinArgs &a, uint32_t q0, uint32_t q1, DBuf<uint32_t> &tileStart,
                      DBuf<uint32_t> &nTiles, const TileZero &zero = TileZero{});
sj_status prepare_fp32(sj_grid_s *g, JoinArgs &a, int *prec);

// One persistent launch over (instance, tile) work items of many joins;
// mfma: every instance runs the matrix-core FP32 walk, else the FP64 walks.
// replay (fill): the instances' LDS tile and matrix-core walks replay the
// hit record of the count kernel
sj_status launch_join_many(const JoinArgs *d_args, const uint32_t *workInst, const uint32_t *workTile,
                           uint32_t W, uint32_t *qctr, uint32_t chunk, int useQueue, int pass, int mfma,
                           unsigned blocks, hipStream_t s, bool replay = false);
// Per-tile bound of the pieces a walk executes (cw candidates per piece)
// for the first `bound` tiles of a.
sj_status launch_tile_bound(const JoinArgs &a, int k, uint32_t bound, uint32_t cw, uint32_t *out, hipStream_t s);

// Work-queue keys of tiles [0, bound) of one join (descending 8-bit
// log-scale bucket of the staged candidates; val[t] = t).
sj_status launch_tile_cost(const JoinArgs &a, int k, uint32_t bound, uint32_t *key, uint32_t *val, hipStream_t s);

// Plain FP64 traversal from LDS-staged candidates for d != k or d > 8, d <= 64.
bool wide_supported(int d, int k);
sj_status launch_join_wide(const JoinArgs &a, int pass, int cu_count, hipStream_t s);

}  // namespace sj
