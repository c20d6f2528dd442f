// Plain FP64 wide traversal kernel and its launch; the tile walk itself
// is in sj_wide.hpp.
#include "sj_wide.hpp"

namespace sj {

template <int PASS>
__global__ __launch_bounds__(JB) void k_join_wide(JoinArgs a) {
    __shared__ WideLds lds[WPB];
    if (PASS == PASS_FILL && !fill_gate(a)) return;
    const unsigned wave = threadIdx.x >> 6;
    const unsigned lane = __lane_id();
    WideLds &L = lds[wave];
    const uint32_t nTiles = *a.nTilesDev;
    LaneStats st;
    if (a.tileOrder) {
        const uint32_t ns = min(gridDim.x, (uint32_t)QSLICES);
        const uint32_t sl = blockIdx.x % ns;
        const uint32_t nbl = (gridDim.x - sl + ns - 1) / ns;
        uint64_t i = (uint64_t)(blockIdx.x / ns) * WPB + wave;
        for (;;) {
            const uint64_t q = queue_entry(i, sl, ns);
            if (q >= nTiles) break;
            uint32_t nx = 0;
            if (lane == 0) nx = atomicAdd(a.qctr + sl * QSTRIDE, 1u);
            process_tile_wide<PASS>(a, a.tileOrder[q], L, st);
            i = (uint64_t)nbl * WPB + readlane_u32(nx, 0);
        }
    } else {
        for (uint32_t tile = blockIdx.x * WPB + wave; tile < nTiles; tile += gridDim.x * WPB)
            process_tile_wide<PASS>(a, tile, L, st);
    }
    flush_stats(a, st);
}

// Every grid the K == D tile walk does not cover, up to WIDE_MAXD dims.
bool wide_supported(int d, int k) { return d <= WIDE_MAXD && (d > DEV_MAXK || d != k); }

sj_status launch_join_wide(const JoinArgs &a, int pass, int cu_count, hipStream_t s) {
    const uint64_t len = (uint64_t)a.qEnd - a.qBegin;
    uint64_t blocks = std::min<uint64_t>((uint64_t)cu_count * 8, (len + 63) / 64 / WPB + 1);
    if (blocks == 0) blocks = 1;
    if (pass == PASS_COUNT)
        hipLaunchKernelGGL(k_join_wide<PASS_COUNT>, dim3((unsigned)blocks), dim3(JB), 0, s, a);
    else
        hipLaunchKernelGGL(k_join_wide<PASS_FILL>, dim3((unsigned)blocks), dim3(JB), 0, s, a);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

}  // namespace sj
