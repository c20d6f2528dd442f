// Device grid construction (reference main.cu:240-503, built on the host
// there).  Steps, all on the context stream:
//   1. per-dimension min/max + finiteness check        (k_minmax)
//   2. grid parameters: min-eps, max+eps, nCells, ID multipliers
//                                                       (k_grid_params)
//   3. per-point cell coordinates -> linear cell ID    (k_cell_keys)
//   4. stable LSD radix sort of (linearID, pointID)    (points stay in
//      ascending ID inside a cell, as main.cu:291-313 produces)
//   5. head flags + exclusive scan -> compacted non-empty cell table
//   6. coordinates gathered into cell order, per-cell coordinates and the
//      O(1) dense lookup table when the grid is small enough.
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstring>

#include "sj_common.hpp"

namespace sj {

constexpr int GB = 256;

// Per-block min/max of the first k coordinates and a check that all d are
// finite, into partial[block * (2k + 1) ..]: k minima, k maxima, then 1.0
// when a coordinate was not finite.  DT = d when d <= 8 (the row is loaded
// in full, four rows in flight per thread), DT = 0 for any d.
template <int DT>
__global__ __launch_bounds__(GB) void k_minmax(const double *pts, uint64_t n, int d, int k,
                                               double *partial) {
    constexpr int KM = DT > 0 ? (DT < SJ_MAXK ? DT : SJ_MAXK) : SJ_MAXK;
    constexpr int U = DT > 0 ? 4 : 1;
    double mn[KM], mx[KM];
#pragma unroll
    for (int j = 0; j < KM; j++) {
        mn[j] = INFINITY;
        mx[j] = -INFINITY;
    }
    bool bad = false;
    const uint64_t stride = (uint64_t)gridDim.x * GB;
    for (uint64_t i0 = (uint64_t)blockIdx.x * GB + threadIdx.x; i0 < n; i0 += stride * U) {
        if constexpr (DT > 0) {
            double v[U][DT];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t i = i0 + (uint64_t)u * stride;
#pragma unroll
                for (int j = 0; j < DT; j++) v[u][j] = i < n ? pts[i * DT + j] : pts[i0 * DT + j];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
#pragma unroll
                for (int j = 0; j < DT; j++) {
                    bad |= !isfinite(v[u][j]);
                    if (j < k) {
                        mn[j] = fmin(mn[j], v[u][j]);
                        mx[j] = fmax(mx[j], v[u][j]);
                    }
                }
            }
        } else {
            const double *p = pts + i0 * (uint64_t)d;
            for (int j = 0; j < d; j++) bad |= !isfinite(p[j]);
#pragma unroll
            for (int j = 0; j < KM; j++) {
                if (j < k) {
                    const double x = p[j];
                    mn[j] = fmin(mn[j], x);
                    mx[j] = fmax(mx[j], x);
                }
            }
        }
    }
    __shared__ uint32_t sbad;
    if (threadIdx.x == 0) sbad = 0u;
    __syncthreads();
    if (bad) sbad = 1u;
    __shared__ double smn[GB / 64][SJ_MAXK], smx[GB / 64][SJ_MAXK];
    const unsigned lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < KM; j++) {
        double a = mn[j], b = mx[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a = fmin(a, __shfl_xor(a, off, 64));
            b = fmax(b, __shfl_xor(b, off, 64));
        }
        if (lane == 0) {
            smn[wave][j] = a;
            smx[wave][j] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)k) {
        int j = threadIdx.x;
        double a = INFINITY, b = -INFINITY;
        for (int w = 0; w < GB / 64; w++) {
            a = fmin(a, smn[w][j]);
            b = fmax(b, smx[w][j]);
        }
        partial[(uint64_t)blockIdx.x * (2 * k + 1) + j] = a;
        partial[(uint64_t)blockIdx.x * (2 * k + 1) + k + j] = b;
    }
    if (threadIdx.x == 0) partial[(uint64_t)blockIdx.x * (2 * k + 1) + 2 * k] = sbad ? 1.0 : 0.0;
}

using MinMaxKernel = void (*)(const double *, uint64_t, int, int, double *);
static MinMaxKernel pick_minmax(int d) {
    switch (d) {
    case 1: return k_minmax<1>;
    case 2: return k_minmax<2>;
    case 3: return k_minmax<3>;
    case 4: return k_minmax<4>;
    case 5: return k_minmax<5>;
    case 6: return k_minmax<6>;
    case 7: return k_minmax<7>;
    case 8: return k_minmax<8>;
    }
    return k_minmax<0>;
}

// main.cu:454-501: min/max, then minArr -= eps, maxArr += eps,
// nCells = ceil((max-min)/eps), totalCells = prod nCells (dim 0 fastest).
// One block: the block reduces the per-block partials of each dimension.
// The error flags start here: flags[0] = GP_NONFINITE when a coordinate was
// not finite, 0 otherwise (later kernels add their bits).
__global__ __launch_bounds__(GB) void k_grid_params(const double *partial, int nb, int k, double eps,
                                                    GridParams *gp, uint32_t *flags) {
    __shared__ double smn[GB / 64][SJ_MAXK], smx[GB / 64][SJ_MAXK];
    __shared__ double fmn[SJ_MAXK], fmx[SJ_MAXK];
    __shared__ uint32_t sbad;
    const unsigned lane = __lane_id(), wave = threadIdx.x >> 6;
    const int rs = 2 * k + 1;
    if (threadIdx.x == 0) sbad = 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += GB)
        if (partial[(uint64_t)i * rs + 2 * k] != 0.0) sbad = 1u;
    for (int j = 0; j < k; j++) {
        double a = INFINITY, b = -INFINITY;
        for (int i = threadIdx.x; i < nb; i += GB) {
            a = fmin(a, partial[(uint64_t)i * rs + j]);
            b = fmax(b, partial[(uint64_t)i * rs + k + j]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a = fmin(a, __shfl_xor(a, off, 64));
            b = fmax(b, __shfl_xor(b, off, 64));
        }
        if (lane == 0) {
            smn[wave][j] = a;
            smx[wave][j] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)k) {
        const int j = threadIdx.x;
        double a = INFINITY, b = -INFINITY;
        for (int w = 0; w < GB / 64; w++) {
            a = fmin(a, smn[w][j]);
            b = fmax(b, smx[w][j]);
        }
        fmn[j] = a;
        fmx[j] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint32_t f = sbad ? (uint32_t)GP_NONFINITE : 0u;
    *flags = f;  // the parameters' own overflow bits below go to gp->flags
    uint64_t tot = 1, mul = 1;
    for (int j = 0; j < k; j++) {
        const double a = fmn[j] - eps;
        const double b = fmx[j] + eps;
        gp->minArr[j] = a;
        gp->maxArr[j] = b;
        const double c = ceil((b - a) / eps);
        uint32_t nc = 0;
        if (!(c < 4294967296.0) || !(c >= 1.0))
            f |= GP_OVERFLOW;
        else
            nc = (uint32_t)c;
        gp->nCells[j] = nc;
        gp->mult[j] = mul;
        if (__umul64hi(tot, (uint64_t)nc) != 0) f |= GP_OVERFLOW;
        tot = tot * (uint64_t)nc;
        mul = tot;
    }
    for (int j = k; j < SJ_MAXK; j++) {
        gp->minArr[j] = 0;
        gp->maxArr[j] = 0;
        gp->nCells[j] = 1;
        gp->mult[j] = 0;
    }
    gp->totalCells = tot;
    gp->flags = f;
    gp->pad = 0;
}

// Cell coordinate (unsigned)((x - minArr)/eps): FP64 subtract, correctly
// rounded divide, truncation (main.cu:263,294; kernel.cu:1457).  One block
// per sort tile of RADIX_TILE points: the keys and values, and the tile's
// count of each lowest 8-bit key digit for the sort's first pass
// (hist0[digit * ntiles + tile]).
constexpr int KB = 1024;  // threads of the key kernel: four keys each per sort tile
static_assert(RADIX_TILE % KB == 0 && KB >= 256, "sort tile");
template <class KT>
__global__ __launch_bounds__(KB) void k_cell_keys(const double *pts, uint64_t n, int d, int k,
                                                  double eps, const GridParams *gp, KT *keys,
                                                  uint32_t *vals, uint32_t *flags, uint32_t *hist0,
                                                  uint32_t ntiles) {
    __shared__ uint32_t bins[256];
    if (threadIdx.x < 256) bins[threadIdx.x] = 0u;
    __syncthreads();
    bool bad = false;
    const uint64_t base = (uint64_t)blockIdx.x * RADIX_TILE;
#pragma unroll
    for (uint32_t r = 0; r < RADIX_TILE / KB; r++) {
        const uint64_t i = base + (uint64_t)r * KB + threadIdx.x;
        if (i < n) {
            const double *p = pts + i * (uint64_t)d;
            uint64_t lin = 0;
            for (int j = 0; j < k; j++) {
                double v = (p[j] - gp->minArr[j]) / eps;
                uint32_t c = 0;
                if (!(v >= 0.0) || !(v < (double)gp->nCells[j]))
                    bad = true;
                else
                    c = (uint32_t)v;
                lin += (uint64_t)c * gp->mult[j];
            }
            keys[i] = (KT)lin;
            vals[i] = (uint32_t)i;
            atomicAdd(&bins[(uint32_t)lin & 255u], 1u);
        }
    }
    if (bad) atomicOr(flags, (uint32_t)GP_OVERFLOW);
    __syncthreads();
    if (threadIdx.x < 256) hist0[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = bins[threadIdx.x];
}

template <class KT>
__global__ __launch_bounds__(GB) void k_compact_cells(const KT *keys, const uint32_t *ex, uint64_t n,
                                                      uint32_t *cellOf, uint32_t *cellStart,
                                                      uint64_t *cellLinear, uint32_t *nne) {
    uint64_t i = (uint64_t)blockIdx.x * GB + threadIdx.x;
    if (i >= n) return;
    const bool head = i == 0 || keys[i] != keys[i - 1];
    uint32_t c = ex[i] + (head ? 1u : 0u) - 1u;
    cellOf[i] = c;
    if (head) {
        cellStart[c] = (uint32_t)i;
        cellLinear[c] = (uint64_t)keys[i];
    }
    if (i == n - 1) {
        cellStart[c + 1] = (uint32_t)n;
        *nne = c + 1;
    }
}

// one thread per (point, coordinate): row s of the output is row ids[s]
__global__ __launch_bounds__(GB) void k_gather_points(const double *pts, const uint32_t *ids,
                                                      uint64_t n, int d, double *out) {
    const uint64_t e0 = (uint64_t)blockIdx.x * GB + threadIdx.x;
    const uint64_t total = n * (uint64_t)d;
    if (e0 >= total) return;
    // 32-bit division when the element count allows it
    const uint64_t row = total <= 0xFFFFFFFFull ? (uint64_t)((uint32_t)e0 / (uint32_t)d) : e0 / (uint64_t)d;
    const int j = (int)(e0 - row * d);
    out[e0] = pts[(uint64_t)ids[row] * d + j];
}

__global__ __launch_bounds__(GB) void k_cell_coords(const double *spts, const uint32_t *cellStart,
                                                    uint64_t nne, int d, int k, double eps,
                                                    const GridParams *gp, uint32_t *coord) {
    uint64_t c = (uint64_t)blockIdx.x * GB + threadIdx.x;
    if (c >= nne) return;
    const double *p = spts + (uint64_t)cellStart[c] * d;
    for (int j = 0; j < k; j++) coord[c * k + j] = (uint32_t)((p[j] - gp->minArr[j]) / eps);
}

// posLB[L], L in [0, tc]: the first sorted position whose cell has linear
// ID >= L, i.e. cellStart of the first non-empty cell at or after L (n past
// the last).  A cell L is non-empty iff posLB[L] < posLB[L + 1].  Thread t
// fills the run [t * POSLB_RUN, (t + 1) * POSLB_RUN); its wave covers 64
// runs from Lw.  The loads are kept off long dependent chains:
//   - the wave finds lo, its first cell at or after Lw, with a 64-way
//     search (each round the 64 lanes probe evenly spaced cells);
//   - the lanes load the 64 cells from lo, one each, and a lane counts
//     those below its own first ID with a 6-step search over the lanes'
//     registers; a lane with all 64 below (a dense stretch) searches the
//     rest in memory;
//   - the at most POSLB_RUN cells inside the lane's run are loaded
//     independently, and each entry takes the start of the first of them
//     at or after its ID; runs with no cell in them (the usual case in
//     sparse stretches) take the next cell's start without more loads.
constexpr uint32_t POSLB_RUN = 16;
__global__ __launch_bounds__(GB) void k_pos_lb(const uint64_t *cellLinear, const uint32_t *cellStart,
                                               uint64_t nne, uint64_t tc, uint32_t *posLB) {
    const uint64_t t = (uint64_t)blockIdx.x * GB + threadIdx.x;
    const unsigned lane = __lane_id();
    const uint64_t Lw = (t - lane) * POSLB_RUN;  // the wave's lowest ID
    if (Lw > tc) return;                          // wave-uniform
    // first index with cellLinear >= Lw, within [lo, hi]
    uint64_t lo = 0, hi = nne;
    while (lo < hi) {
        const uint64_t step = (hi - lo + 63) / 64;
        const uint64_t pos = lo + step * lane;
        const bool less = pos < hi && cellLinear[pos] < Lw;
        const uint32_t c = (uint32_t)__popcll(__ballot(less));  // probes below Lw: a prefix
        if (c == 0) {
            hi = lo;
        } else {
            const uint64_t last = lo + step * (c - 1);  // below Lw
            const uint64_t next = lo + step * c;        // at or above Lw when < hi
            lo = last + 1;
            hi = next < hi ? next : hi;
        }
    }
    // every lane active here: the shuffles read all 64 lanes
    const uint64_t L0 = t * POSLB_RUN;
    const uint64_t cj = lo + lane < nne ? cellLinear[lo + lane] : ~0ull;
    uint32_t c = 0;  // cells lo .. lo+63 below L0 (they are sorted)
#pragma unroll
    for (uint32_t step = 32; step; step >>= 1) {
        const uint64_t v = __shfl(cj, (int)(c + step - 1), 64);
        if (v < L0) c += step;
    }
    if (__shfl(cj, 63, 64) < L0) c = 64;  // c was 63: the 64th is below too
    uint64_t a = lo + c;  // first index with cellLinear >= L0 when c < 64
    if (L0 > tc) return;
    if (c == 64) {
        // at most L0 - Lw cells lie in [Lw, L0): the first at or after L0 is
        // within lo + (L0 - Lw)
        uint64_t b = lo + (L0 - Lw) < nne ? lo + (L0 - Lw) : nne;
        while (a < b) {
            const uint64_t mid = (a + b) >> 1;
            if (cellLinear[mid] < L0) a = mid + 1; else b = mid;
        }
    }
    // a <= nne; cellStart holds nne + 1 entries
    const uint64_t L1 = L0 + POSLB_RUN < tc + 1 ? L0 + POSLB_RUN : tc + 1;
    uint32_t v[POSLB_RUN];
    const uint64_t c0 = a < nne ? cellLinear[a] : ~0ull;
    const uint32_t s0 = cellStart[a];
    if (c0 >= L0 + POSLB_RUN) {
#pragma unroll
        for (uint32_t q = 0; q < POSLB_RUN; q++) v[q] = s0;
    } else {
        uint64_t cl[POSLB_RUN];
        uint32_t cs[POSLB_RUN + 1];
        cl[0] = c0;
        cs[0] = s0;
#pragma unroll
        for (uint32_t q = 1; q < POSLB_RUN; q++) cl[q] = a + q < nne ? cellLinear[a + q] : ~0ull;
#pragma unroll
        for (uint32_t q = 1; q <= POSLB_RUN; q++) cs[q] = a + q <= nne ? cellStart[a + q] : 0u;
#pragma unroll
        for (uint32_t q = 0; q < POSLB_RUN; q++) {
            uint32_t x = cs[0];
#pragma unroll
            for (uint32_t j = 0; j < POSLB_RUN; j++)
                if (cl[j] < L0 + q) x = cs[j + 1];
            v[q] = x;
        }
    }
    if (L1 - L0 == POSLB_RUN) {
        uint4 *o = reinterpret_cast<uint4 *>(posLB + L0);
#pragma unroll
        for (uint32_t q = 0; q < POSLB_RUN / 4; q++) o[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (uint32_t q = 0; q < POSLB_RUN; q++)
            if (L0 + q < L1) posLB[L0 + q] = v[q];
    }
}

static unsigned blocks_for(uint64_t n, unsigned b = GB) {
    uint64_t g = (n + b - 1) / b;
    return (unsigned)(g ? g : 1);
}

static int bits_for(uint64_t maxv) {
    int b = 0;
    while (b < 64 && (maxv >> b)) b++;
    return b;
}

// pinned words 16.. hold the grid parameters read back by the host
static GridParams *pinned_params(sj_context ctx) {
    static_assert(16 * 8 + sizeof(GridParams) <= 64 * 8, "pinned readback space");
    return reinterpret_cast<GridParams *>(ctx->pinned + 16);
}

template <class KT>
static sj_status sort_and_compact(sj_grid_s *g, const double *d_points, uint32_t *d_flags, int bits) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t n = g->n;
    DBuf<KT> keys, keys2;
    DBuf<uint32_t> vals, vals2;
    SJ_TRY(keys.alloc(pool, n));
    SJ_TRY(keys2.alloc(pool, n));
    SJ_TRY(vals.alloc(pool, n));
    SJ_TRY(vals2.alloc(pool, n));
    const uint32_t nt = radix_tiles(n);
    DBuf<uint32_t> hist0;  // the sort's first histogram, counted with the keys
    SJ_TRY(hist0.alloc(pool, (uint64_t)256 * nt));
    hipLaunchKernelGGL((k_cell_keys<KT>), dim3(nt), dim3(KB), 0, s, d_points, n, g->d, g->k, g->eps, g->d_gp.p,
                       keys.p, vals.p, d_flags, hist0.p, nt);
    SJ_LAUNCH_CHECK();
    KT *kp = keys.p, *ka = keys2.p;
    uint32_t *vp = vals.p, *va = vals2.p;
    sj_status st;
    if constexpr (sizeof(KT) == 8)
        st = radix_sort_pairs_u64(pool, s, &kp, &ka, &vp, &va, n, bits, hist0.p);
    else
        st = radix_sort_pairs_u32(pool, s, &kp, &ka, &vp, &va, n, bits, hist0.p);
    SJ_TRY(st);

    // cell index of each sorted position: the scan of the run heads
    DBuf<uint32_t> ex;
    SJ_TRY(ex.alloc(pool, n + 1));
    SJ_TRY(scan_exclusive_heads(pool, s, kp, n, ex.p));
    SJ_TRY(g->cellOf.alloc(pool, n));
    SJ_TRY(g->cellStart.alloc(pool, n + 1));
    SJ_TRY(g->cellLinear.alloc(pool, n));
    hipLaunchKernelGGL((k_compact_cells<KT>), dim3(blocks_for(n)), dim3(GB), 0, s, kp, ex.p, n, g->cellOf.p,
                       g->cellStart.p, g->cellLinear.p, d_flags + 1);
    SJ_LAUNCH_CHECK();
    // sorted point IDs (indexLookupArr): the sort's value buffer itself
    g->sortedIds = std::move(vp == vals.p ? vals : vals2);
    // the parameters with the error flags and nNonEmpty (d_flags[1], written
    // by the compaction) in one copy, read by grid_build once the gather
    // behind them is enqueued
    SJ_TRY_HIP(hipMemcpyAsync(pinned_params(g->ctx), g->d_gp.p, sizeof(GridParams), hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipEventRecord(g->ctx->ev, s));
    return SJ_OK;
}

sj_status grid_build(sj_context ctx, const double *d_points, uint64_t n, int d, int k, double eps,
                     sj_grid_s *g) {
    if (!ctx || !d_points || n == 0 || d < 1 || d > SJ_MAXD || k < 1 || k > SJ_MAXK || k > d) {
        set_error("grid_build: invalid arguments (n=%llu d=%d k=%d)", (unsigned long long)n, d, k);
        return SJ_ERR_INVALID;
    }
    if (!(eps > 0.0) || std::isinf(eps)) {
        set_error("grid_build: eps must be finite and > 0");
        return SJ_ERR_INVALID;
    }
    if (n >= 0x7fffffffull) {
        set_error("grid_build: at most 2^31-2 points (int32 point IDs)");
        return SJ_ERR_OVERFLOW;
    }
    Pool *pool = &ctx->pool;
    hipStream_t s = ctx->stream;
    g->ctx = ctx;
    g->n = n;
    g->d = d;
    g->k = k;
    g->eps = eps;

    const unsigned nbmm = (unsigned)std::min<uint64_t>(1024, blocks_for(n));
    DBuf<double> partial;
    SJ_TRY(partial.alloc(pool, (uint64_t)nbmm * (2 * k + 1)));
    SJ_TRY(g->d_gp.alloc(pool, 1));
    // error flags (set first by k_grid_params), then the non-empty cell count
    uint32_t *flags = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(g->d_gp.p) + offsetof(GridParams, err));
    static_assert(offsetof(GridParams, nne) == offsetof(GridParams, err) + 4, "flags, then the count");
    hipLaunchKernelGGL(pick_minmax(d), dim3(nbmm), dim3(GB), 0, s, d_points, n, d, k, partial.p);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_grid_params, dim3(1), dim3(GB), 0, s, partial.p, (int)nbmm, k, eps, g->d_gp.p, flags);
    SJ_LAUNCH_CHECK();
    GridParams *pgp = pinned_params(ctx);
    auto check_params = [&]() -> sj_status {
        g->gp = *pgp;
        if (g->gp.flags & GP_NONFINITE) {
            set_error("grid_build: non-finite coordinate");
            return SJ_ERR_INVALID;
        }
        if (g->gp.flags & GP_OVERFLOW) {
            set_error("grid_build: cell count overflows (nCells >= 2^32 or totalCells >= 2^64)");
            return SJ_ERR_OVERFLOW;
        }
        return SJ_OK;
    };
    // Same shape and eps as the context's last grid: sort with its key
    // width now and check the parameters when the host next waits (the
    // keys are computed from the device parameters either way; a wider
    // grid than guessed is sorted again).  Otherwise wait for them first.
    const auto &h = ctx->gridHint;
    const bool guess = h.valid && h.n == n && h.d == d && h.k == k && h.eps == eps;
    int bits = h.bits;
    bool wide = h.wide;
    if (!guess) {
        SJ_TRY_HIP(hipMemcpyAsync(pgp, g->d_gp.p, sizeof(GridParams), hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        SJ_TRY(check_params());
        bits = bits_for(g->gp.totalCells ? g->gp.totalCells - 1 : 0);
        wide = g->gp.totalCells > 0xFFFFFFFFull;
    }
    for (int attempt = 0;; attempt++) {
        if (wide)
            SJ_TRY(sort_and_compact<uint64_t>(g, d_points, flags, bits));
        else
            SJ_TRY(sort_and_compact<uint32_t>(g, d_points, flags, bits));
        // the coordinates in cell order need no host value: enqueued before
        // the host waits for the flags and the non-empty cell count
        SJ_TRY(g->sortedPts.alloc(pool, n * (uint64_t)d));
        hipLaunchKernelGGL(k_gather_points, dim3(blocks_for(n * (uint64_t)d)), dim3(GB), 0, s, d_points,
                           g->sortedIds.p, n, d, g->sortedPts.p);
        SJ_LAUNCH_CHECK();
        SJ_TRY_HIP(hipEventSynchronize(ctx->ev));
        if (guess && attempt == 0) {
            SJ_TRY(check_params());  // copied with the flags before the event
            const int need = bits_for(g->gp.totalCells ? g->gp.totalCells - 1 : 0);
            const bool needWide = g->gp.totalCells > 0xFFFFFFFFull;
            if (need > bits || (needWide && !wide)) {
                bits = need;
                wide = needWide;
                continue;
            }
        }
        break;
    }
    g->gp = *pgp;  // the last copy: parameters, flags and the non-empty cell count
    if (g->gp.err & GP_OVERFLOW) {
        set_error("cell coordinate outside the grid (overflow)");
        return SJ_ERR_OVERFLOW;
    }
    g->nNonEmpty = g->gp.nne;
    ctx->gridHint.valid = true;
    ctx->gridHint.wide = g->gp.totalCells > 0xFFFFFFFFull;
    ctx->gridHint.d = d;
    ctx->gridHint.k = k;
    ctx->gridHint.bits = bits_for(g->gp.totalCells ? g->gp.totalCells - 1 : 0);
    ctx->gridHint.n = n;
    ctx->gridHint.eps = eps;
    const uint64_t nne = g->nNonEmpty;
    SJ_TRY(g->cellCoord.alloc(pool, nne * (uint64_t)k));
    hipLaunchKernelGGL(k_cell_coords, dim3(blocks_for(nne)), dim3(GB), 0, s, g->sortedPts.p,
                       g->cellStart.p, nne, d, k, eps, g->d_gp.p, g->cellCoord.p);
    SJ_LAUNCH_CHECK();

    // O(1) lookups when the linear-ID table is at most 64M entries or 16x n:
    // posLB gives both a cell's sorted range and whether it is non-empty
    const uint64_t tc = g->gp.totalCells;
    g->dense = tc <= (1ull << 26) || tc <= 16ull * n;
    if (g->dense && pool->arena_active() && (tc + 1) * 4 + 256 + g->reserveAfter > pool->arena_free())
        g->dense = false;  // a static scratch too small for the table: binary search
    if (g->dense) {
        SJ_TRY(g->posLB.alloc(pool, tc + 1));
        hipLaunchKernelGGL(k_pos_lb, dim3(blocks_for(tc / POSLB_RUN + 1)), dim3(GB), 0, s, g->cellLinear.p,
                           g->cellStart.p, nne, tc, g->posLB.p);
        SJ_LAUNCH_CHECK();
    }
    return SJ_OK;
}

// Per-dimension occupied-coordinate mask (main.cu:381-401), host side.
sj_status grid_nd_mask_host(sj_grid_s *g, std::vector<uint32_t> &mask, std::vector<uint32_t> &offsets) {
    const uint64_t nne = g->nNonEmpty;
    const int k = g->k;
    std::vector<uint32_t> coord(nne * k);
    SJ_TRY_HIP(hipMemcpy(coord.data(), g->cellCoord.p, coord.size() * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    mask.clear();
    offsets.assign(2 * k, 0);
    for (int j = 0; j < k; j++) {
        std::vector<uint8_t> occ(g->gp.nCells[j], 0);
        for (uint64_t c = 0; c < nne; c++) occ[coord[c * k + j]] = 1;
        offsets[2 * j] = (uint32_t)mask.size();
        for (uint32_t v = 0; v < g->gp.nCells[j]; v++)
            if (occ[v]) mask.push_back(v);
        offsets[2 * j + 1] = (uint32_t)mask.size() - 1;
    }
    return SJ_OK;
}

}  // namespace sj
