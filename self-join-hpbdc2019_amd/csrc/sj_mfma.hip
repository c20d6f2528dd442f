// Matrix-core traversal kernels and their host side (decision constants,
// dispatch); the tile walk itself is in sj_mfma.hpp.
#include <cmath>
#include <cstring>

#include "sj_mfma.hpp"

namespace sj {

template <int DP, int PASS, bool RP>
__device__ __forceinline__ void join_body_mfma(const JoinArgs &a, MfmaLds<DP> &L) {
    if (PASS == PASS_FILL && !fill_gate(a)) return;
    const unsigned lane = __lane_id();
    mfma_lds_clear(L);
    const uint32_t nTiles = *a.nTilesDev;
    LaneStats st;
    TileRec R;
    if (a.tileOrder) {
        const uint32_t ns = min(gridDim.x, (uint32_t)QSLICES);
        const uint32_t sl = blockIdx.x % ns;
        const uint32_t nbl = (gridDim.x - sl + ns - 1) / ns;
        uint64_t i = (uint64_t)(blockIdx.x / ns) * WPB + (threadIdx.x >> 6);
        for (;;) {
            const uint64_t q = queue_entry(i, sl, ns);
            if (q >= nTiles) break;
            uint32_t nx = 0;
            if (lane == 0) nx = atomicAdd(a.qctr + sl * QSTRIDE, 1u);
            process_tile_mfma<DP, PASS, RP>(a, a.tileOrder[q], L, st, R);
            i = (uint64_t)nbl * WPB + readlane_u32(nx, 0);
        }
    } else {
        for (uint32_t tile = blockIdx.x * WPB + (threadIdx.x >> 6); tile < nTiles; tile += gridDim.x * WPB)
            process_tile_mfma<DP, PASS, RP>(a, tile, L, st, R);
    }
    flush_stats(a, st);
    // mask words this wave's chunks needed, plus its last partial block
    if (PASS == PASS_COUNT && a.recDir && lane == 0) {
        const uint32_t sl = (blockIdx.x * WPB + (threadIdx.x >> 6)) % REC_NSLICE;
        atomicAdd(&a.recCtr[REC_SLICE0 + sl * REC_STRIDE + 1], (unsigned long long)(R.want + REC_BLK));
    }
}

template <int DP, int PASS>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(2))) void k_join_mfma(JoinArgs a) {
    __shared__ MfmaLds<DP> lds[WPB];
    join_body_mfma<DP, PASS, false>(a, lds[threadIdx.x >> 6]);
}

// Fill replaying the count pass's hit record
template <int DP>
__global__ __launch_bounds__(JB) __attribute__((amdgpu_waves_per_eu(2))) void k_join_mfma_rp(JoinArgs a) {
    __shared__ MfmaLds<DP> lds[WPB];
    join_body_mfma<DP, PASS_FILL, true>(a, lds[threadIdx.x >> 6]);
}

// ---------------------------------------------------------------- host
int mfma_padded_dims(int d) {
    if (d < 1 || d > MFMA_MAX_D) return 0;
    return d <= 4 ? 4 : (d <= 8 ? 8 : (d <= 16 ? 16 : (d <= 32 ? 32 : 64)));
}

// Error bound of the matrix-core squared distance against the canonical
// FP64 one, as alpha*s + beta with s = |q|^2 + |c|^2 of the centred FP32
// vectors (u = 2^-24, d <= |q| + |c| <= sqrt(2s) for the true distance d):
//  * centring in FP64 and one rounding to FP32: each vector is off by at
//    most 1.0001 u |v|, so the distance by E <= 1.0002 u (|q| + |c|), and
//    d^2 by at most 2 d E + E^2 <= 4.0008 u s + 4 u^2 s;
//  * centring in FP32 from the FP32 copy, coordinates within R of the
//    origin: a vector is off by at most 1.0001 u R + u |v|, so
//    E <= u (2.0002 R + 1.0001 sqrt(2s)) and, with
//    sqrt(2s) <= (2s / sqrt(T) + sqrt(T)) / 2, d^2 moves by at most
//    u (4.0004 R / sqrt(T) + 4.0004) s + 2.0002 u R sqrt(T)
//    + u^2 (8.002 R^2 + 8.001 s);
//  * norms, dot product (DP terms, any order) and s - 2 q.c in FP32:
//    (2*gamma + 3u)*s with gamma = 1.01*DP*u;
//  * the canonical FP64 sum is within rho = 1.01*(d+2)*2^-53 of the exact
//    value relatively: it enters through tlo = T/(1+rho) and
//    thi = T/(1-rho), each widened by 4u for the FP32 compare;
//  * 5% on alpha and beta covers their own FP32 evaluation and the
//    rounding of s; DP * 2^-120 covers denormal flushing in the matrix core.
MfmaBand mfma_band(double T, double R, int d) {
    MfmaBand b;
    const double u = std::ldexp(1.0, -24);
    const int DP = mfma_padded_dims(d);
    const double gam = 1.01 * DP * u;
    double alpha = (2.0 * gam + 3.0 * u), beta = 0.0;
    if (R >= 0.0) {
        const double rt = std::sqrt(T);
        alpha += u * (4.0004 * R / rt + 4.0004) + 8.001 * u * u;
        beta += 2.0002 * u * R * rt + 8.002 * u * u * R * R;
    } else {
        alpha += 4.0008 * u + 8.0 * u * u;
    }
    alpha *= 1.05;
    beta = beta * 1.05 + DP * std::ldexp(1.0, -120);
    const double rho = 1.01 * (d + 2) * std::ldexp(1.0, -53);
    const double tlo = T / (1.0 + rho) * (1.0 - 4.0 * u);
    const double thi = T / (1.0 - rho) * (1.0 + 4.0 * u);
    float f = (float)alpha;
    if ((double)f < alpha) f = std::nextafter(f, INFINITY);
    b.alpha = f;
    f = (float)beta;
    if ((double)f < beta) f = std::nextafter(f, INFINITY);
    b.beta = f;
    f = (float)tlo;
    if ((double)f > tlo) f = std::nextafter(f, -INFINITY);
    b.tlo = f;
    f = (float)thi;
    if ((double)f < thi) f = std::nextafter(f, INFINITY);
    b.thi = f;
    b.t32 = (float)T;
    b.DP = DP;
    return b;
}

void mfma_consts(JoinArgs &a, double T, double R) {
    const MfmaBand b = mfma_band(T, R, a.d);
    a.mAlpha = b.alpha;
    a.mBeta = b.beta;
    a.mTlo = b.tlo;
    a.mThi = b.thi;
    a.mT32 = b.t32;
    a.mDP = b.DP;
    a.mSrc32 = R >= 0.0 ? 1 : 0;
}

using MfmaKernel = void (*)(JoinArgs);
static MfmaKernel pick_mfma(int DP, int pass, bool replay) {
    if (replay) {
        switch (DP) {
        case 4: return k_join_mfma_rp<4>;
        case 8: return k_join_mfma_rp<8>;
        case 16: return k_join_mfma_rp<16>;
        case 32: return k_join_mfma_rp<32>;
        case 64: return k_join_mfma_rp<64>;
        }
        return nullptr;
    }
    switch (DP) {
    case 4: return pass == PASS_COUNT ? k_join_mfma<4, PASS_COUNT> : k_join_mfma<4, PASS_FILL>;
    case 8: return pass == PASS_COUNT ? k_join_mfma<8, PASS_COUNT> : k_join_mfma<8, PASS_FILL>;
    case 16: return pass == PASS_COUNT ? k_join_mfma<16, PASS_COUNT> : k_join_mfma<16, PASS_FILL>;
    case 32: return pass == PASS_COUNT ? k_join_mfma<32, PASS_COUNT> : k_join_mfma<32, PASS_FILL>;
    case 64: return pass == PASS_COUNT ? k_join_mfma<64, PASS_COUNT> : k_join_mfma<64, PASS_FILL>;
    }
    return nullptr;
}

sj_status launch_join_mfma(const JoinArgs &a, int pass, int cu_count, hipStream_t s, bool replay) {
    MfmaKernel kfn = pick_mfma(a.mDP, pass, replay);
    if (!kfn) {
        set_error("no matrix-core kernel for padded dims %d", a.mDP);
        return SJ_ERR_UNSUPPORTED;
    }
    const uint64_t len = (uint64_t)a.qEnd - a.qBegin;
    uint64_t blocks = std::min<uint64_t>((uint64_t)cu_count * 8, (len + 63) / 64 / WPB + 1);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(kfn, dim3((unsigned)blocks), dim3(JB), 0, s, a);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

}  // namespace sj
