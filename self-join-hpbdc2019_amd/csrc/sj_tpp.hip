// Threads-per-point schedule: reference search mode 8
// (kernelNDGridIndexGlobalSortedCellsDynamicThreadsV3, kernel.cu:2072-2188,
// launched with NB_THREAD_THIRD1/2/3 = 16, 8 and 2 threads per point by
// thirds of the batches, GPU.cu:2239-2263, params.h:13-15).
//
// Query points are taken in workload order (cells by descending full
// stencil workload, ties by cell index; the points of a cell in ID order,
// GPU.cu:1697-1716).  The heaviest third of that order runs with 16 lanes
// per point, the middle third with 8, the last with 2.  (The reference
// deals its batches out strided over the order, so there the thirds do not
// follow the workload; here they do.)  The lanes of a point split each
// stencil cell's candidates round-robin over the full 3^k stencil; the
// count pass sums their hits, the fill writes them in candidate order, one
// contiguous run per round of the group's lanes (ballot prefix).
#include "sj_common.hpp"
#include "sj_traverse.hpp"

namespace sj {

template <int K, int PASS>
__global__ __launch_bounds__(JB) void k_join_tpp(JoinArgs a, const uint32_t *order, uint32_t first, uint32_t count,
                                                 uint32_t T) {
    const uint32_t lane = __lane_id();
    const uint64_t gid = ((uint64_t)blockIdx.x * JB + threadIdx.x) / T;
    const uint32_t j = lane % T;  // T divides 64
    const uint32_t gbase = lane - j;
    const uint64_t gmask = (T == 64u ? ~0ull : ((1ull << T) - 1ull)) << gbase;
    if (PASS == PASS_FILL && !fill_gate(a)) return;
    LaneStats st;
    if (gid < count) {
        const int d = a.d;
        const uint32_t s = order[first + gid];
        const uint32_t cellA = a.cellOf[s];
        uint32_t c[K];
#pragma unroll
        for (int q = 0; q < K; q++) c[q] = a.cellCoord[(uint64_t)cellA * K + q];
        const uint64_t linA = a.cellLinear[cellA];
        const double *p = a.pts + (uint64_t)s * d;
        const bool ownW = s >= a.wBegin && s < a.wEnd;
        uint64_t ownPos = 0, ownLim = 0;
        if (PASS == PASS_FILL && ownW) {
            ownPos = a.offS[s] - a.outBase;
            ownLim = ownPos + a.ownCnt[s];
        }
        // count: this lane's hits; fill: the group's hits so far (uniform)
        uint32_t own = 0;
        uint64_t cand = 0, err = 0;
        if (PASS == PASS_COUNT || ownW) {
            int o[K];
#pragma unroll
            for (int q = 0; q < K; q++) o[q] = -1;
            int noff = 1;
#pragma unroll
            for (int q = 0; q < K; q++) noff *= 3;
            for (int it = 0; it < noff; ++it) {
                bool inb = true;
                int64_t dl = 0;
#pragma unroll
                for (int q = 0; q < K; q++) {
                    const int64_t v = (int64_t)c[q] + o[q];
                    inb = inb && v >= 0 && v < (int64_t)a.nCells[q];
                    dl += (int64_t)o[q] * a.mult[q];
                }
                const int64_t b = inb ? find_cell(a, (uint64_t)((int64_t)linA + dl)) : -1;
                if (b >= 0) {
                    const uint32_t t0 = a.cellStart[b], t1 = a.cellStart[b + 1];
                    // the group's lanes share t0, t1: one trip count per group
                    for (uint32_t base = t0; base < t1; base += T) {
                        const uint32_t t = base + j;
                        const bool hit = t < t1 && d2_fp64<0>(p, a.pts + (uint64_t)t * d, d) <= a.T;
                        if (PASS == PASS_COUNT) {
                            cand += t < t1 ? 1u : 0u;
                            own += hit ? 1u : 0u;
                        } else {
                            const uint64_t m = (__ballot(hit) & gmask) >> gbase;
                            if (hit) {
                                const uint64_t w = ownPos + own + (uint32_t)__popcll(m & ((1ull << j) - 1ull));
                                if (w < ownLim) a.ids[w] = (int32_t)a.origId[t];
                                else err++;
                            }
                            own += (uint32_t)__popcll(m);
                        }
                    }
                }
                bool carry = true;
#pragma unroll
                for (int q = 0; q < K; q++) {
                    if (carry) {
                        if (o[q] == 1) {
                            o[q] = -1;
                        } else {
                            o[q]++;
                            carry = false;
                        }
                    }
                }
            }
        }
        if (PASS == PASS_COUNT) {
            for (uint32_t off = T >> 1; off > 0; off >>= 1) own += __shfl_xor(own, (int)off, 64);
            if (j == 0) a.ownCnt[s] = own;
        } else if (ownW && j == 0 && own != ownLim - ownPos) {
            err++;
        }
        st.cand += cand;
        st.err += err;
    }
    flush_stats(a, st);
}

// Sizes of the cells in workload order and, from their exclusive scan, the
// sorted positions of all points in that order.
__global__ void k_tpp_sizes(const uint32_t *corder, const uint32_t *cellStart, uint64_t nne, uint32_t *size) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (i >= nne) return;
    const uint32_t c = corder[i];
    size[i] = cellStart[c + 1] - cellStart[c];
}

__global__ void k_tpp_points(const uint32_t *corder, const uint32_t *cellStart, const uint32_t *rank, uint64_t nne,
                             uint32_t *order) {
    const uint64_t i = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (i >= nne) return;
    const uint32_t c = corder[i];
    const uint32_t p0 = cellStart[c], p1 = cellStart[c + 1];
    for (uint32_t p = p0; p < p1; p++) order[rank[i] + (p - p0)] = p;
}

using TppKernel = void (*)(JoinArgs, const uint32_t *, uint32_t, uint32_t, uint32_t);

template <int K>
static TppKernel pick_tpp_k(int pass) {
    return pass == PASS_COUNT ? k_join_tpp<K, PASS_COUNT> : k_join_tpp<K, PASS_FILL>;
}

static TppKernel pick_tpp(int k, int pass) {
    switch (k) {
    case 1: return pick_tpp_k<1>(pass);
    case 2: return pick_tpp_k<2>(pass);
    case 3: return pick_tpp_k<3>(pass);
    case 4: return pick_tpp_k<4>(pass);
    case 5: return pick_tpp_k<5>(pass);
    case 6: return pick_tpp_k<6>(pass);
    case 7: return pick_tpp_k<7>(pass);
    case 8: return pick_tpp_k<8>(pass);
    }
    return nullptr;
}

// Point order of the schedule (n sorted positions) from the cell order.
sj_status tpp_point_order(sj_grid_s *g, const uint32_t *corder, DBuf<uint32_t> &order) {
    Pool *pool = &g->ctx->pool;
    hipStream_t s = g->ctx->stream;
    const uint64_t nne = g->nNonEmpty;
    const unsigned nb = (unsigned)((nne + JB - 1) / JB ? (nne + JB - 1) / JB : 1);
    DBuf<uint32_t> size, rank;
    SJ_TRY(size.alloc(pool, nne));
    SJ_TRY(rank.alloc(pool, nne + 1));
    SJ_TRY(order.alloc(pool, g->n));
    hipLaunchKernelGGL(k_tpp_sizes, dim3(nb), dim3(JB), 0, s, corder, g->cellStart.p, nne, size.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u32(pool, s, size.p, nne, rank.p));
    hipLaunchKernelGGL(k_tpp_points, dim3(nb), dim3(JB), 0, s, corder, g->cellStart.p, rank.p, nne, order.p);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// The three launches of a pass: thirds of the order with 16, 8 and 2 lanes
// per point.
sj_status launch_join_tpp(const JoinArgs &a, int k, int pass, uint64_t n, hipStream_t s) {
    TppKernel kfn = pick_tpp(k, pass);
    if (!kfn) {
        set_error("threads per point: k=%d above the device stencil limit", k);
        return SJ_ERR_UNSUPPORTED;
    }
    const uint32_t cut[4] = {0u, (uint32_t)(n / 3), (uint32_t)(2 * n / 3), (uint32_t)n};
    const uint32_t lanes[3] = {16u, 8u, 2u};
    for (int third = 0; third < 3; third++) {
        const uint32_t cnt = cut[third + 1] - cut[third];
        if (cnt == 0) continue;
        const uint64_t threads = (uint64_t)cnt * lanes[third];
        const unsigned blocks = (unsigned)((threads + JB - 1) / JB);
        hipLaunchKernelGGL(kfn, dim3(blocks), dim3(JB), 0, s, a, a.tppOrder, cut[third], cnt, lanes[third]);
        SJ_LAUNCH_CHECK();
    }
    return SJ_OK;
}

}  // namespace sj
