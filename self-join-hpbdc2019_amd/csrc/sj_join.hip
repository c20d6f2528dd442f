// Cell-pair traversal of the grid join (reference kernel.cu:949-1538,
// 2204-2531) producing the neighbour table as CSR, plus workload ordering
// (kernel.cu:33-250, GPU.cu:1660-1734), the atomic work queue
// (kernel.cu:2202-2236) and the batch estimator (kernel.cu:602-942).
//
// Two passes over the same traversal:
//   count: per query point the hits it owns (own row) and, for the
//          Unicomp forms, the mirrored hits it gives to the other point;
//   fill:  own hits go to a lane-private cursor in the point's row, the
//          mirrored hits are wave-aggregated: lanes hitting the same
//          candidate take one atomic on that candidate's row cursor and
//          place their IDs with a ballot popcount prefix.
// Exact predicate: acc = sum fma((x_q - x_p)^2) (FP64, dimension order),
// hit <=> acc <= T with T = max{x : sqrt_rn(x) <= eps}.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "sj_common.hpp"

namespace sj {

constexpr int JB = 256;          // traversal block size
constexpr int DEV_MAXK = 8;      // indexed dims handled by the device stencil
enum { PASS_COUNT = 0, PASS_FILL = 1 };
enum { STAT_CAND = 0, STAT_BAND = 1, STAT_MIS = 2, STAT_ERR = 3 };

struct JoinArgs {
    const double *pts;        // n x d, cell order
    const float *pts32;       // n x d, (x - origin) in FP32, or null
    const uint32_t *origId;   // original ID of each sorted position
    const uint32_t *cellOf;
    const uint32_t *cellStart;
    const uint64_t *cellLinear;
    const uint32_t *cellCoord;
    const int32_t *dense;
    uint64_t nne;
    uint32_t nCells[DEV_MAXK];
    int64_t mult[DEV_MAXK];
    int d, k, form, pad0;
    double T;
    float lo32, hi32, T32, pad1;
    // query items: strided range, or the work queue when order != null
    uint32_t qBegin, qEnd;
    const uint32_t *order;
    uint32_t nItems, chunk;
    uint32_t *qctr;
    // rows written by this launch (sorted positions)
    uint32_t wBegin, wEnd;
    // count pass
    uint32_t *ownCnt, *mirCnt;
    unsigned long long *stat;
    // fill pass
    const uint64_t *ownBase;
    unsigned long long *mirCur;
    const uint64_t *rowEnd;
    int32_t *ids;
};

__device__ __forceinline__ uint32_t popc_below64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int64_t find_cell(const JoinArgs &a, uint64_t lin) {
    if (a.dense) return a.dense[lin];
    uint64_t lo = 0, hi = a.nne;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (a.cellLinear[mid] < lin) lo = mid + 1; else hi = mid;
    }
    return (lo < a.nne && a.cellLinear[lo] == lin) ? (int64_t)lo : -1;
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
    // B-UNICOMP
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

template <int D, int PREC>
struct QueryPoint {
    double x[D > 0 ? D : 1];
    float y[D > 0 ? D : 1];
};

// FP64 canonical predicate, dimension order, explicit fma.
template <int D>
__device__ __forceinline__ double dist2_fp64(const JoinArgs &a, const double *p, uint32_t t,
                                             uint32_t s) {
    const int d = D > 0 ? D : a.d;
    const double *q = a.pts + (uint64_t)t * d;
    double acc = 0.0;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            const double df = q[j] - p[j];
            acc = j == 0 ? df * df : fma(df, df, acc);
        }
    } else {
        const double *pp = a.pts + (uint64_t)s * d;
        for (int j = 0; j < d; j++) {
            const double df = q[j] - pp[j];
            acc = j == 0 ? df * df : fma(df, df, acc);
        }
    }
    return acc;
}

template <int D>
__device__ __forceinline__ float dist2_fp32(const JoinArgs &a, const float *p, uint32_t t, uint32_t s) {
    const int d = D > 0 ? D : a.d;
    const float *q = a.pts32 + (uint64_t)t * d;
    float acc = 0.0f;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            const float df = q[j] - p[j];
            acc = j == 0 ? df * df : fmaf(df, df, acc);
        }
    } else {
        const float *pp = a.pts32 + (uint64_t)s * d;
        for (int j = 0; j < d; j++) {
            const float df = q[j] - pp[j];
            acc = j == 0 ? df * df : fmaf(df, df, acc);
        }
    }
    return acc;
}

template <int K, int D, int PASS, int PREC>
__device__ void process_point(const JoinArgs &a, uint32_t s) {
    constexpr int DR = D > 0 ? D : 1;
    const unsigned lane = __lane_id();
    const uint32_t cellA = a.cellOf[s];
    uint32_t c[K];
#pragma unroll
    for (int j = 0; j < K; j++) c[j] = a.cellCoord[(uint64_t)cellA * K + j];
    const uint64_t linA = a.cellLinear[cellA];
    double p[DR];
    float p32[DR];
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; j++) p[j] = a.pts[(uint64_t)s * D + j];
        if constexpr (PREC == 1) {
#pragma unroll
            for (int j = 0; j < D; j++) p32[j] = a.pts32[(uint64_t)s * D + j];
        }
    }
    const bool ownW = s >= a.wBegin && s < a.wEnd;
    const int32_t myId = (int32_t)a.origId[s];
    uint32_t own = 0;
    uint64_t ownPos = 0, ownLim = 0;
    if (PASS == PASS_FILL && ownW) {
        ownPos = a.ownBase[s];
        ownLim = ownPos + a.ownCnt[s];
    }
    uint64_t cand = 0, band = 0, mis = 0, err = 0;
    const bool anyMirror = a.form != SJ_FORM_FULL;

    auto visit = [&](uint32_t b, bool mirror) {
        const uint32_t t0 = a.cellStart[b], t1 = a.cellStart[b + 1];
        if (PASS == PASS_COUNT) cand += t1 - t0;
        for (uint32_t t = t0; t < t1; ++t) {
            bool hit;
            if constexpr (PREC == 1) {
                const float s32 = dist2_fp32<D>(a, p32, t, s);
                if (s32 < a.lo32) {
                    hit = true;
                } else if (s32 > a.hi32) {
                    hit = false;
                } else {
                    hit = dist2_fp64<D>(a, p, t, s) <= a.T;
                    if (PASS == PASS_COUNT) {
                        band++;
                        mis += (hit != (s32 <= a.T32));
                    }
                }
            } else {
                hit = dist2_fp64<D>(a, p, t, s) <= a.T;
            }
            if (hit && ownW) {
                if (PASS == PASS_COUNT) {
                    own++;
                } else {
                    const uint64_t w = ownPos + own;
                    if (w < ownLim) a.ids[w] = (int32_t)a.origId[t];
                    else err++;
                    own++;
                }
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
                            if (w < a.rowEnd[t]) a.ids[w] = myId;
                            else err++;
                        }
                    }
                }
            }
        }
    };

    // origin cell: every ordered pair, self included (kernel.cu:445-496)
    visit(cellA, false);

    // neighbour cells by odometer over {-1,0,1}^K
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
                const int64_t b = find_cell(a, linB);
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
        if (cand) atomicAdd(&a.stat[STAT_CAND], (unsigned long long)cand);
        if (band) atomicAdd(&a.stat[STAT_BAND], (unsigned long long)band);
        if (mis) atomicAdd(&a.stat[STAT_MIS], (unsigned long long)mis);
    } else {
        if (ownW && own != (ownLim - ownPos)) err++;
        if (err) atomicAdd(&a.stat[STAT_ERR], (unsigned long long)err);
    }
}

template <int K, int D, int PASS, int PREC>
__global__ __launch_bounds__(JB) void k_join(JoinArgs a) {
    if (a.order) {
        const unsigned lane = __lane_id();
        for (;;) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(a.qctr, a.chunk);
            base = __shfl(base, 0, 64);
            if (base >= a.nItems) break;
            const uint32_t end = min(base + a.chunk, a.nItems);
            for (uint32_t it = base + lane; it < end; it += 64)
                process_point<K, D, PASS, PREC>(a, a.order[it]);
        }
    } else {
        const uint64_t s = (uint64_t)a.qBegin + (uint64_t)blockIdx.x * JB + threadIdx.x;
        if (s < a.qEnd) process_point<K, D, PASS, PREC>(a, (uint32_t)s);
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

__global__ void k_order_sizes(const uint32_t *orderCell, const uint32_t *cellStart, uint64_t nne,
                              uint32_t *size, uint32_t *rank) {
    uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (i >= nne) return;
    const uint32_t c = orderCell[i];
    size[i] = cellStart[c + 1] - cellStart[c];
    rank[c] = (uint32_t)i;
}

__global__ void k_queue_points(const uint32_t *cellOf, const uint32_t *cellStart, const uint32_t *rank,
                               const uint32_t *pos, uint64_t n, uint32_t *queue) {
    uint64_t s = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = cellOf[s];
    queue[pos[rank[c]] + ((uint32_t)s - cellStart[c])] = (uint32_t)s;
}

// ----------------------------------------------------- small row kernels
__global__ void k_totals_to_orig(const uint32_t *own, const uint32_t *mir, const uint32_t *origId,
                                 uint64_t n, uint32_t *cntOrig) {
    uint64_t s = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (s >= n) return;
    cntOrig[origId[s]] = own[s] + mir[s];
}

__global__ void k_totals_sorted(const uint32_t *own, const uint32_t *mir, uint64_t n, uint32_t *tot) {
    uint64_t s = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (s >= n) return;
    tot[s] = own[s] + mir[s];
}

// whole result: rows in original ID order
__global__ void k_row_bases_orig(const uint64_t *offsets, const uint32_t *origId, const uint32_t *own,
                                 uint64_t n, uint64_t *ownBase, unsigned long long *mirCur,
                                 uint64_t *rowEnd) {
    uint64_t s = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (s >= n) return;
    const uint32_t id = origId[s];
    const uint64_t b = offsets[id];
    ownBase[s] = b;
    mirCur[s] = b + own[s];
    rowEnd[s] = offsets[id + 1];
}

// batch: rows [w0, w1) in sorted order, relative to offS[w0]
__global__ void k_row_bases_window(const uint64_t *offS, const uint32_t *own, uint32_t w0, uint32_t w1,
                                   uint64_t *ownBase, unsigned long long *mirCur, uint64_t *rowEnd) {
    uint64_t s = (uint64_t)w0 + (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (s >= w1) return;
    const uint64_t b = offS[s] - offS[w0];
    ownBase[s] = b;
    mirCur[s] = b + own[s];
    rowEnd[s] = offS[s + 1] - offS[w0];
}

__global__ void k_to_fp32(const double *pts, uint64_t n, int d, int k, const GridParams *gp, float *out,
                          unsigned int *rmaxBits) {
    uint64_t e = (uint64_t)blockIdx.x * JB + threadIdx.x;
    float r = 0.0f;
    if (e < n * (uint64_t)d) {
        const int j = (int)(e % (uint64_t)d);
        const double o = j < k ? gp->minArr[j] : 0.0;
        const double v = pts[e] - o;
        out[e] = (float)v;
        r = (float)fabs(v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r = fmaxf(r, __shfl_xor(r, off, 64));
    if (__lane_id() == 0) atomicMax(rmaxBits, __float_as_uint(r));
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

template <int K, int D>
static JoinKernel pick_pass(int pass, int prec) {
    if (pass == PASS_COUNT) return prec ? k_join<K, D, PASS_COUNT, 1> : k_join<K, D, PASS_COUNT, 0>;
    return prec ? k_join<K, D, PASS_FILL, 1> : k_join<K, D, PASS_FILL, 0>;
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
struct Fp32Margin {
    bool usable;
    float lo, hi, t32;
};

static Fp32Margin fp32_margin(double T, double R, int D) {
    const double u32 = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    const double Rh = R * (1.0 + std::ldexp(1.0, -20)) + std::ldexp(1.0, -126);
    const double E0 = 2.0 * Rh * (u32 + 2.0 * u64);
    const double gam = D * u32 / (1.0 - D * u32);
    const double rho = (D + 3) * u64 * 1.01;
    const double sD = std::sqrt((double)D);
    const double tiny = D * std::ldexp(1.0, -126) * 8.0;  // possible FP32 denormal flushes
    const double Ahi = T / (1.0 - rho);
    const double nhi = std::sqrt(Ahi);
    const double Ehi = E0 + u32 * (nhi + E0);
    const double up = (1.0 + gam) * (Ahi + Ehi * (2.0 * sD * nhi + D * Ehi)) + tiny;
    const double Alo = T / (1.0 + rho);
    const double nlo = std::sqrt(Alo);
    const double Elo = E0 + u32 * (nlo + E0);
    const double dn = (1.0 - gam) * (Alo - 2.0 * Elo * sD * nlo) - tiny;
    double M = std::max(up - T, T - dn) * 1.05;
    Fp32Margin m;
    m.usable = std::isfinite(M) && M < 0.25 * T && (T + M) < 1e37 && R < 1e30;
    double lo = T - M, hi = T + M;
    float flo = (float)lo, fhi = (float)hi;
    if ((double)flo > lo) flo = std::nextafter(flo, -INFINITY);
    if ((double)fhi < hi) fhi = std::nextafter(fhi, INFINITY);
    m.lo = flo;
    m.hi = fhi;
    m.t32 = (float)T;
    return m;
}

struct EvTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EvTimer(hipStream_t st) : s(st) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~EvTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start() { (void)hipEventRecord(a, s); }
    void stop() { (void)hipEventRecord(b, s); }
    double ms() {
        float f = 0;
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&f, a, b);
        return f;
    }
};

static JoinArgs base_args(sj_grid_s *g, int form, double T) {
    JoinArgs a;
    memset(&a, 0, sizeof a);
    a.pts = g->sortedPts.p;
    a.origId = g->sortedIds.p;
    a.cellOf = g->cellOf.p;
    a.cellStart = g->cellStart.p;
    a.cellLinear = g->cellLinear.p;
    a.cellCoord = g->cellCoord.p;
    a.dense = g->dense ? g->denseCell.p : nullptr;
    a.nne = g->nNonEmpty;
    for (int j = 0; j < DEV_MAXK; j++) {
        a.nCells[j] = j < g->k ? g->gp.nCells[j] : 1;
        a.mult[j] = j < g->k ? (int64_t)g->gp.mult[j] : 0;
    }
    a.d = g->d;
    a.k = g->k;
    a.form = form;
    a.T = T;
    a.qBegin = 0;
    a.qEnd = (uint32_t)g->n;
    a.wBegin = 0;
    a.wEnd = (uint32_t)g->n;
    return a;
}

// Cells in descending workload, flattened to a queue of sorted positions.
static sj_status build_queue(sj_grid_s *g, int lid, DBuf<uint32_t> &queue, DBuf<uint32_t> *orderOut,
                             DBuf<uint32_t> *workOut) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t nne = g->nNonEmpty, n = g->n;
    JoinArgs a = base_args(g, SJ_FORM_FULL, 0.0);
    DBuf<uint32_t> w, key, key2, val, val2;
    SJ_TRY(w.alloc(pool, nne));
    SJ_TRY(key.alloc(pool, nne));
    SJ_TRY(key2.alloc(pool, nne));
    SJ_TRY(val.alloc(pool, nne));
    SJ_TRY(val2.alloc(pool, nne));
    hipLaunchKernelGGL(pick_workload(g->k), dim3(nblk(nne)), dim3(JB), 0, s, a, lid, w.p);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_desc_keys, dim3(nblk(nne)), dim3(JB), 0, s, w.p, nne, key.p, val.p);
    SJ_LAUNCH_CHECK();
    uint32_t *kp = key.p, *ka = key2.p, *vp = val.p, *va = val2.p;
    SJ_TRY(radix_sort_pairs_u32(pool, s, &kp, &ka, &vp, &va, nne, 32));
    DBuf<uint32_t> size, rank, pos;
    SJ_TRY(size.alloc(pool, nne));
    SJ_TRY(rank.alloc(pool, nne));
    SJ_TRY(pos.alloc(pool, nne + 1));
    hipLaunchKernelGGL(k_order_sizes, dim3(nblk(nne)), dim3(JB), 0, s, vp, g->cellStart.p, nne, size.p,
                       rank.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u32(pool, s, size.p, nne, pos.p));
    SJ_TRY(queue.alloc(pool, n));
    hipLaunchKernelGGL(k_queue_points, dim3(nblk(n)), dim3(JB), 0, s, g->cellOf.p, g->cellStart.p, rank.p,
                       pos.p, n, queue.p);
    SJ_LAUNCH_CHECK();
    if (orderOut) {
        SJ_TRY(orderOut->alloc(pool, nne));
        SJ_TRY_HIP(hipMemcpyAsync(orderOut->p, vp, nne * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    if (workOut) *workOut = std::move(w);
    return SJ_OK;
}

sj_status grid_workload(sj_grid_s *g, int lid, int64_t *nb, uint32_t *order) {
    if (g->k > DEV_MAXK) {
        set_error("workload: k=%d above the device stencil limit %d", g->k, DEV_MAXK);
        return SJ_ERR_UNSUPPORTED;
    }
    DBuf<uint32_t> queue, ord, w;
    SJ_TRY(build_queue(g, lid, queue, &ord, &w));
    std::vector<uint32_t> hw(g->nNonEmpty);
    SJ_TRY_HIP(hipStreamSynchronize(g->ctx->stream));
    SJ_TRY_HIP(hipMemcpy(hw.data(), w.p, hw.size() * 4, hipMemcpyDeviceToHost));
    if (nb)
        for (uint64_t i = 0; i < hw.size(); i++) nb[i] = hw[i];
    if (order) SJ_TRY_HIP(hipMemcpy(order, ord.p, g->nNonEmpty * 4, hipMemcpyDeviceToHost));
    return SJ_OK;
}

static sj_status launch_join(sj_grid_s *g, const JoinArgs &a, int pass, int prec, uint32_t nItems) {
    hipStream_t s = g->ctx->stream;
    JoinKernel kfn = pick_kernel(g->k, g->d, pass, prec);
    if (!kfn) {
        set_error("no traversal kernel for k=%d", g->k);
        return SJ_ERR_UNSUPPORTED;
    }
    unsigned blocks;
    if (a.order) {
        blocks = (unsigned)std::min<uint64_t>((uint64_t)g->ctx->cu_count * 8, nblk(nItems));
    } else {
        blocks = nblk((uint64_t)a.qEnd - a.qBegin);
    }
    hipLaunchKernelGGL(kfn, dim3(blocks), dim3(JB), 0, s, a);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// Sampled estimate of |S| (every stride-th original point, full stencil).
sj_status batch_sample_count(sj_grid_s *g, uint64_t stride, uint64_t *count) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t n = g->n;
    if (stride == 0) stride = 1;
    const uint64_t nsamp = (uint64_t)((double)n * (1.0 / (double)stride));
    *count = 0;
    if (nsamp == 0) return SJ_OK;
    DBuf<uint32_t> posOf, list, own, qctr;
    DBuf<unsigned long long> stat;
    SJ_TRY(posOf.alloc(pool, n));
    SJ_TRY(list.alloc(pool, nsamp));
    SJ_TRY(own.alloc(pool, n));
    SJ_TRY(qctr.alloc(pool, 1));
    SJ_TRY(stat.alloc(pool, 8));
    hipLaunchKernelGGL(k_sample_positions, dim3(nblk(n)), dim3(JB), 0, s, g->sortedIds.p, n, stride, nsamp,
                       posOf.p, list.p, 0);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sample_positions, dim3(nblk(nsamp)), dim3(JB), 0, s, g->sortedIds.p, n, stride,
                       nsamp, posOf.p, list.p, 1);
    SJ_LAUNCH_CHECK();
    SJ_TRY(fill_u32(s, qctr.p, 1, 0u));
    SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(stat.p), 8, 0ull));
    JoinArgs a = base_args(g, SJ_FORM_FULL, eps_threshold(g->eps));
    a.order = list.p;
    a.nItems = (uint32_t)nsamp;
    a.chunk = 64;
    a.qctr = qctr.p;
    a.ownCnt = own.p;
    a.mirCnt = own.p;  // unused by FULL
    a.stat = stat.p;
    SJ_TRY(launch_join(g, a, PASS_COUNT, 0, (uint32_t)nsamp));
    SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(stat.p) + 4, 1, 0ull));
    hipLaunchKernelGGL(k_sum_gather, dim3(nblk(nsamp)), dim3(JB), 0, s, own.p, list.p, nsamp, stat.p + 4);
    SJ_LAUNCH_CHECK();
    unsigned long long h = 0;
    SJ_TRY_HIP(hipMemcpyAsync(&h, stat.p + 4, 8, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipStreamSynchronize(s));
    *count = h;
    return SJ_OK;
}

struct JoinPlan {
    int form, prec, sched, sort_rows;
    uint64_t capacity;
    uint32_t chunk;
};

sj_status sort_rows_device(sj_context ctx, const uint64_t *offsets, int32_t *ids, uint64_t n);

sj_status join(sj_grid_s *g, const sj_join_params *P, sj_result_s *r) {
    const auto t_wall0 = std::chrono::steady_clock::now();
    if (!g || !P || !r) {
        set_error("join: null argument");
        return SJ_ERR_INVALID;
    }
    if (P->form < 0 || P->form > 3 || P->precision < 0 || P->precision > 1 || P->schedule < 0 ||
        P->schedule > 1) {
        set_error("join: invalid form/precision/schedule");
        return SJ_ERR_INVALID;
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
    EvTimer tw(s), tc(s), ts(s), tf(s), tr(s);

    // precision form: FP32 copy relative to the grid origin + margin
    int prec = P->precision;
    if (prec == SJ_PREC_FP32_RECHECK) {
        DBuf<unsigned int> rbits;
        SJ_TRY(rbits.alloc(pool, 1));
        SJ_TRY(fill_u32(s, rbits.p, 1, 0u));
        if (!g->sortedPts32.p) SJ_TRY(g->sortedPts32.alloc(pool, n * (uint64_t)g->d));
        hipLaunchKernelGGL(k_to_fp32, dim3(nblk(n * (uint64_t)g->d)), dim3(JB), 0, s, g->sortedPts.p, n, g->d,
                           g->k, g->d_gp.p, g->sortedPts32.p, rbits.p);
        SJ_LAUNCH_CHECK();
        uint32_t hb = 0;
        SJ_TRY_HIP(hipMemcpyAsync(&hb, rbits.p, 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        float rf;
        memcpy(&rf, &hb, 4);
        // the FP32 rounding of |v| may be below |v|: widen by one FP32 ulp
        const double R = (double)std::nextafter(rf, INFINITY);
        Fp32Margin m = fp32_margin(T, R, g->d);
        if (m.usable) {
            a.pts32 = g->sortedPts32.p;
            a.lo32 = m.lo;
            a.hi32 = m.hi;
            a.T32 = m.t32;
        } else {
            prec = SJ_PREC_FP64;  // band would cover the whole range
        }
    }

    // scheduling
    DBuf<uint32_t> queue, qctr;
    if (P->schedule == SJ_SCHED_WORKQUEUE) {
        tw.start();
        SJ_TRY(build_queue(g, P->form == SJ_FORM_LID ? 1 : 0, queue, nullptr, nullptr));
        tw.stop();
        SJ_TRY(qctr.alloc(pool, 1));
        a.order = queue.p;
        a.nItems = (uint32_t)n;
        a.chunk = P->queue_chunk > 0 ? (uint32_t)P->queue_chunk : 64u;
        a.qctr = qctr.p;
    }

    // count pass
    DBuf<uint32_t> own, mir;
    DBuf<unsigned long long> stat;
    SJ_TRY(own.alloc(pool, n));
    SJ_TRY(mir.alloc(pool, n));
    SJ_TRY(stat.alloc(pool, 8));
    SJ_TRY(fill_u32(s, mir.p, n, 0u));
    SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(stat.p), 8, 0ull));
    if (a.order) SJ_TRY(fill_u32(s, qctr.p, 1, 0u));
    a.ownCnt = own.p;
    a.mirCnt = mir.p;
    a.stat = stat.p;
    tc.start();
    SJ_TRY(launch_join(g, a, PASS_COUNT, prec, (uint32_t)n));
    tc.stop();

    // offsets in original ID order (the neighbour table product)
    ts.start();
    DBuf<uint32_t> cntOrig;
    SJ_TRY(cntOrig.alloc(pool, n));
    SJ_TRY(r->d_offsets.alloc(pool, n + 1));
    hipLaunchKernelGGL(k_totals_to_orig, dim3(nblk(n)), dim3(JB), 0, s, own.p, mir.p, g->sortedIds.p, n,
                       cntOrig.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u64(pool, s, cntOrig.p, n, r->d_offsets.p));
    ts.stop();
    uint64_t *pin = ctx->pinned;
    SJ_TRY_HIP(hipMemcpyAsync(pin, r->d_offsets.p + n, 8, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipMemcpyAsync(pin + 1, stat.p, 3 * 8, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipStreamSynchronize(s));
    const uint64_t total = pin[0];
    r->total = total;
    r->stats.total_pairs = total;
    r->stats.candidates = pin[1 + STAT_CAND];
    r->stats.fp32_band_pairs = pin[1 + STAT_BAND];
    r->stats.fp32_misclassified = pin[1 + STAT_MIS];
    SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(stat.p) + STAT_ERR, 1, 0ull));

    DBuf<uint64_t> ownBase, rowEnd;
    DBuf<unsigned long long> mirCur;
    SJ_TRY(ownBase.alloc(pool, n));
    SJ_TRY(rowEnd.alloc(pool, n));
    SJ_TRY(mirCur.alloc(pool, n));
    a.ownBase = ownBase.p;
    a.rowEnd = rowEnd.p;
    a.mirCur = mirCur.p;

    const bool whole = P->capacity == 0 || total <= P->capacity;
    double ms_fill = 0.0;
    if (whole) {
        SJ_TRY(r->d_ids.alloc(pool, total));
        hipLaunchKernelGGL(k_row_bases_orig, dim3(nblk(n)), dim3(JB), 0, s, r->d_offsets.p, g->sortedIds.p,
                           own.p, n, ownBase.p, mirCur.p, rowEnd.p);
        SJ_LAUNCH_CHECK();
        a.ids = r->d_ids.p;
        if (a.order) SJ_TRY(fill_u32(s, qctr.p, 1, 0u));
        tf.start();
        SJ_TRY(launch_join(g, a, PASS_FILL, prec, (uint32_t)n));
        tf.stop();
        if (P->sort_rows) {
            tr.start();
            SJ_TRY(sort_rows_device(ctx, r->d_offsets.p, r->d_ids.p, n));
            tr.stop();
        }
        SJ_TRY_HIP(hipMemcpyAsync(pin, stat.p + STAT_ERR, 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        if (pin[0]) {
            set_error("join: fill pass disagreed with the count pass (%llu writes)",
                      (unsigned long long)pin[0]);
            return SJ_ERR_STATE;
        }
        r->on_device = true;
        r->stats.num_batches = 1;
        ms_fill = tf.ms();
        r->stats.ms_sort_rows = P->sort_rows ? tr.ms() : 0.0;
    } else {
        // Batches of rows in sorted order, each within `capacity` pairs.
        DBuf<uint32_t> tot;
        DBuf<uint64_t> offS;
        SJ_TRY(tot.alloc(pool, n));
        SJ_TRY(offS.alloc(pool, n + 1));
        hipLaunchKernelGGL(k_totals_sorted, dim3(nblk(n)), dim3(JB), 0, s, own.p, mir.p, n, tot.p);
        SJ_LAUNCH_CHECK();
        SJ_TRY(scan_exclusive_u32_to_u64(pool, s, tot.p, n, offS.p));
        std::vector<uint64_t> hOff(n + 1);
        std::vector<uint64_t> hLin(g->nNonEmpty);
        std::vector<uint32_t> hCellOf(n), hIds(n);
        SJ_TRY_HIP(hipMemcpyAsync(hOff.data(), offS.p, (n + 1) * 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hLin.data(), g->cellLinear.p, g->nNonEmpty * 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hCellOf.data(), g->cellOf.p, n * 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipMemcpyAsync(hIds.data(), g->sortedIds.p, n * 4, hipMemcpyDeviceToHost, s));
        std::vector<uint32_t> hStart(g->nNonEmpty + 1);
        SJ_TRY_HIP(hipMemcpyAsync(hStart.data(), g->cellStart.p, (g->nNonEmpty + 1) * 4, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        // plan windows
        std::vector<std::pair<uint32_t, uint32_t>> win;
        uint64_t w0 = 0;
        while (w0 < n) {
            uint64_t w1 = w0;
            while (w1 < n && hOff[w1 + 1] - hOff[w0] <= P->capacity) w1++;
            if (w1 == w0) {
                set_error("join: capacity %llu is below one neighbour row (%llu pairs)",
                          (unsigned long long)P->capacity, (unsigned long long)(hOff[w0 + 1] - hOff[w0]));
                return SJ_ERR_OVERFLOW;
            }
            win.emplace_back((uint32_t)w0, (uint32_t)w1);
            w0 = w1;
        }
        uint64_t Mmax = 0;
        for (int j = 0; j < g->k; j++) Mmax += g->gp.mult[j];
        DBuf<int32_t> dbuf;
        SJ_TRY(dbuf.alloc(pool, P->capacity));
        a.ids = dbuf.p;
        r->batches.resize(win.size());
        for (size_t bi = 0; bi < win.size(); bi++) {
            const uint32_t s0 = win[bi].first, s1 = win[bi].second;
            // query cells: the window's cells, widened to every cell that may
            // claim one of them (adjacent cells lie within Mmax linear IDs)
            uint64_t q0 = s0, q1 = s1;
            if (P->form != SJ_FORM_FULL) {
                const uint64_t linLo = hLin[hCellOf[s0]], linHi = hLin[hCellOf[s1 - 1]];
                const uint64_t lo = linLo > Mmax ? linLo - Mmax : 0;
                const uint64_t hi = linHi + Mmax;
                const uint64_t c0 = std::lower_bound(hLin.begin(), hLin.end(), lo) - hLin.begin();
                const uint64_t c1 = std::upper_bound(hLin.begin(), hLin.end(), hi) - hLin.begin();
                q0 = hStart[c0];
                q1 = hStart[c1];
            }
            hipLaunchKernelGGL(k_row_bases_window, dim3(nblk(s1 - s0)), dim3(JB), 0, s, offS.p, own.p, s0, s1,
                               ownBase.p, mirCur.p, rowEnd.p);
            SJ_LAUNCH_CHECK();
            JoinArgs ab = a;
            ab.order = nullptr;  // batches use the strided range of their cells
            ab.qBegin = (uint32_t)q0;
            ab.qEnd = (uint32_t)q1;
            ab.wBegin = s0;
            ab.wEnd = s1;
            tf.start();
            SJ_TRY(launch_join(g, ab, PASS_FILL, prec, 0));
            tf.stop();
            const uint64_t cnt = hOff[s1] - hOff[s0];
            sj_batch_host &B = r->batches[bi];
            B.ids.resize(cnt);
            B.rowIds.assign(hIds.begin() + s0, hIds.begin() + s1);
            B.rowOffsets.resize((size_t)(s1 - s0) + 1);
            for (uint32_t q = s0; q <= s1; q++) B.rowOffsets[q - s0] = hOff[q] - hOff[s0];
            SJ_TRY_HIP(hipMemcpyAsync(B.ids.data(), dbuf.p, cnt * 4, hipMemcpyDeviceToHost, s));
            SJ_TRY_HIP(hipStreamSynchronize(s));
            ms_fill += tf.ms();
            if (P->sort_rows) {
                for (uint32_t q = 0; q < s1 - s0; q++)
                    std::sort(B.ids.begin() + B.rowOffsets[q], B.ids.begin() + B.rowOffsets[q + 1]);
            }
        }
        SJ_TRY_HIP(hipMemcpyAsync(pin, stat.p + STAT_ERR, 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        if (pin[0]) {
            set_error("join: fill pass disagreed with the count pass (%llu writes)",
                      (unsigned long long)pin[0]);
            return SJ_ERR_STATE;
        }
        r->on_device = false;
        r->stats.num_batches = win.size();
    }
    r->stats.ms_workload = P->schedule == SJ_SCHED_WORKQUEUE ? tw.ms() : 0.0;
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

__global__ __launch_bounds__(256) void k_sort_rows(const uint64_t *offsets, int32_t *ids, uint64_t n) {
    __shared__ int32_t buf[4][SR_CAP];
    const unsigned wave = threadIdx.x >> 6, lane = __lane_id();
    const uint64_t row = (uint64_t)blockIdx.x * 4 + wave;
    if (row >= n) return;
    const uint64_t b = offsets[row], e = offsets[row + 1];
    const uint64_t len = e - b;
    if (len < 2) return;
    if (len <= SR_CAP) {
        uint32_t p2 = 2;
        while (p2 < len) p2 <<= 1;
        for (uint32_t i = lane; i < p2; i += 64) buf[wave][i] = i < len ? ids[b + i] : INT32_MAX;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
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
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
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

sj_status sort_rows_device(sj_context ctx, const uint64_t *offsets, int32_t *ids, uint64_t n) {
    if (!n) return SJ_OK;
    hipLaunchKernelGGL(k_sort_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, offsets, ids, n);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

}  // namespace sj
