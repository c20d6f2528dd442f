// Exclusive scans with the total appended (out[n] = sum of in[0..n)) for the
// stream-ordered paths of grid construction and joins, in two launches and
// no initialisation pass:
//   1. per-tile sums (tile = 256 threads x 16 items);
//   2. every tile adds the sums of the tiles before it, read directly (at
//      most SCAN_DIRECT_TILES of them, i.e. up to 16M items), and scans
//      itself; the last tile appends the total.
// Longer inputs scan the tile sums in one block in between (three
// launches).  The input is either one array or the sum of two (row
// offsets from own + mirrored counts).
#include "sj_common.hpp"

namespace sj {

namespace {

constexpr int SB = 256;                              // threads per tile
constexpr int SI = 16;                               // items per thread
constexpr uint64_t ST = (uint64_t)SB * SI;           // items per tile
constexpr uint64_t SCAN_DIRECT_TILES = 4096;

struct In1 {
    const uint32_t *a;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return a[i]; }
};
struct In2 {
    const uint32_t *a, *b;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return (uint64_t)a[i] + b[i]; }
};

// head flag of a sorted key array: 1 where a run of equal keys starts
template <class K>
struct InHead {
    const K *k;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return i == 0 || k[i] != k[i - 1] ? 1u : 0u; }
};

template <class T>
__device__ __forceinline__ T wave_incl(T v) {
    const unsigned lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T u = __shfl_up(v, off, 64);
        if (lane >= (unsigned)off) v += u;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_total(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block sum of one value per thread (every thread gets it)
template <class T>
__device__ __forceinline__ T block_total(T v, T *ws) {
    v = wave_total(v);
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    T t = 0;
#pragma unroll
    for (int w = 0; w < SB / 64; w++) t += ws[w];
    __syncthreads();
    return t;
}

template <class In, class Tacc>
__global__ __launch_bounds__(SB) void k_tile_sums(In in, uint64_t n, Tacc *sums) {
    __shared__ Tacc ws[SB / 64];
    const uint64_t base = (uint64_t)blockIdx.x * ST;
    Tacc acc = 0;
#pragma unroll
    for (int j = 0; j < SI; j++) {
        const uint64_t i = base + (uint64_t)j * SB + threadIdx.x;
        if (i < n) acc += (Tacc)in(i);
    }
    acc = block_total(acc, ws);
    if (threadIdx.x == 0) sums[blockIdx.x] = acc;
}

// one block: sums[0..nb) -> exclusive prefixes in place
template <class Tacc>
__global__ __launch_bounds__(SB) void k_tile_sums_scan(Tacc *sums, uint64_t nb) {
    __shared__ Tacc ws[SB / 64];
    Tacc carry = 0;
    for (uint64_t b0 = 0; b0 < nb; b0 += SB) {
        const uint64_t i = b0 + threadIdx.x;
        const Tacc v = i < nb ? sums[i] : 0;
        const Tacc inc = wave_incl(v);
        if (__lane_id() == 63) ws[threadIdx.x >> 6] = inc;
        __syncthreads();
        Tacc before = 0, all = 0;
#pragma unroll
        for (int w = 0; w < SB / 64; w++) {
            const Tacc x = ws[w];
            if (w < (int)(threadIdx.x >> 6)) before += x;
            all += x;
        }
        __syncthreads();
        if (i < nb) sums[i] = carry + before + inc - v;
        carry += all;
    }
}

// direct: the carry is the sum of sums[0..blockIdx); else sums holds it
template <class In, class Tacc, bool DIRECT>
__global__ __launch_bounds__(SB) void k_tile_scan(In in, uint64_t n, const Tacc *sums, Tacc *out) {
    __shared__ Tacc ws[SB / 64];
    __shared__ Tacc tile[ST];
    const uint64_t base = (uint64_t)blockIdx.x * ST;
    Tacc carry;
    if (DIRECT) {
        Tacc c = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += SB) c += sums[b];
        carry = block_total(c, ws);
    } else {
        carry = sums[blockIdx.x];
    }
    // coalesced loads into LDS, then each thread scans SI consecutive items
#pragma unroll
    for (int j = 0; j < SI; j++) {
        const uint64_t i = base + (uint64_t)j * SB + threadIdx.x;
        tile[j * SB + threadIdx.x] = i < n ? (Tacc)in(i) : (Tacc)0;
    }
    __syncthreads();
    Tacc v[SI];
    Tacc local = 0;
#pragma unroll
    for (int j = 0; j < SI; j++) {
        v[j] = tile[threadIdx.x * SI + j];
        local += v[j];
    }
    const Tacc inc = wave_incl(local);
    if (__lane_id() == 63) ws[threadIdx.x >> 6] = inc;
    __syncthreads();
    Tacc before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SB / 64; w++) {
        const Tacc x = ws[w];
        if (w < (int)(threadIdx.x >> 6)) before += x;
        all += x;
    }
    Tacc ex = carry + before + inc - local;
#pragma unroll
    for (int j = 0; j < SI; j++) {
        tile[threadIdx.x * SI + j] = ex;
        ex += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SI; j++) {
        const uint64_t i = base + (uint64_t)j * SB + threadIdx.x;
        if (i < n) out[i] = tile[j * SB + threadIdx.x];
    }
    if (base + ST >= n && threadIdx.x == 0) out[n] = carry + all;
}

sj_status zero_one(hipStream_t s, uint64_t *p) { return fill_u64(s, p, 1, 0ull); }
sj_status zero_one(hipStream_t s, uint32_t *p) { return fill_u32(s, p, 1, 0u); }

template <class In, class Tacc>
sj_status scan_two_pass(Pool *pool, hipStream_t s, In in, uint64_t n, Tacc *out) {
    if (n == 0) return zero_one(s, out);  // out[0] = the empty total
    const uint64_t nb = (n + ST - 1) / ST;
    if (nb > 0xFFFFFFFFull) {
        set_error("scan: %llu items exceed the tile index range", (unsigned long long)n);
        return SJ_ERR_OVERFLOW;
    }
    DBuf<Tacc> sums;
    SJ_TRY(sums.alloc(pool, nb));
    hipLaunchKernelGGL((k_tile_sums<In, Tacc>), dim3((unsigned)nb), dim3(SB), 0, s, in, n, sums.p);
    SJ_LAUNCH_CHECK();
    if (nb <= SCAN_DIRECT_TILES) {
        hipLaunchKernelGGL((k_tile_scan<In, Tacc, true>), dim3((unsigned)nb), dim3(SB), 0, s, in, n, sums.p, out);
    } else {
        hipLaunchKernelGGL((k_tile_sums_scan<Tacc>), dim3(1), dim3(SB), 0, s, sums.p, nb);
        SJ_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_tile_scan<In, Tacc, false>), dim3((unsigned)nb), dim3(SB), 0, s, in, n, sums.p, out);
    }
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

}  // namespace

sj_status scan_exclusive_sum2_u32_to_u64(Pool *pool, hipStream_t s, const uint32_t *a, const uint32_t *b,
                                         uint64_t n, uint64_t *out) {
    return scan_two_pass<In2, uint64_t>(pool, s, In2{a, b}, n, out);
}

sj_status scan_exclusive_u32_to_u64(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n, uint64_t *out) {
    return scan_two_pass<In1, uint64_t>(pool, s, In1{in}, n, out);
}

sj_status scan_exclusive_u32_to_u32(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n, uint32_t *out) {
    return scan_two_pass<In1, uint32_t>(pool, s, In1{in}, n, out);
}

sj_status scan_exclusive_heads(Pool *pool, hipStream_t s, const uint32_t *keys, uint64_t n, uint32_t *out) {
    return scan_two_pass<InHead<uint32_t>, uint32_t>(pool, s, InHead<uint32_t>{keys}, n, out);
}

sj_status scan_exclusive_heads(Pool *pool, hipStream_t s, const uint64_t *keys, uint64_t n, uint32_t *out) {
    return scan_two_pass<InHead<uint64_t>, uint32_t>(pool, s, InHead<uint64_t>{keys}, n, out);
}

}  // namespace sj
