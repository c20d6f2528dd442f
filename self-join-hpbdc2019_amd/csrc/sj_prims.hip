// Device primitives of libsjamd: error state, caching pool, exclusive scan
// and a stable LSD radix sort of (key, value) pairs, written for wave64.
#include <algorithm>
#include <cstring>

#include "sj_common.hpp"

namespace sj {

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

const char *last_error() { return g_err.c_str(); }

// -------------------------------------------------------------------- pool
static size_t size_class(size_t b) {
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
}

static void pool_free(bool pinned, void *p) {
    if (pinned) (void)hipHostFree(p);
    else (void)hipFree(p);
}

Pool::~Pool() {
    for (auto &kv : free_)
        for (void *p : kv.second) pool_free(pinned_, p);
    for (auto &kv : live_) pool_free(pinned_, kv.first);
}

void Pool::arena_begin(void *base, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    abase_ = static_cast<char *>(base);
    acap_ = bytes;
    atop_ = apeak_ = 0;
    an_ = 0;
}

void Pool::arena_end() {
    std::lock_guard<std::mutex> lk(mu_);
    abase_ = nullptr;
    acap_ = atop_ = 0;
    an_ = 0;
}

sj_status Pool::alloc(size_t bytes, void **out) {
    std::lock_guard<std::mutex> lk(mu_);
    if (abase_) {
        const size_t off = (atop_ + 255) & ~(size_t)255;
        const size_t need = bytes ? bytes : 1;
        if (an_ == ARENA_BLOCKS || off > acap_ || need > acap_ - off) {
            set_error("static scratch exhausted: %zu bytes requested at offset %zu of %zu (%d blocks)", need, off,
                      acap_, an_);
            return SJ_ERR_NOMEM;
        }
        ablk_[an_++] = {off, true};
        atop_ = off + need;
        apeak_ = std::max(apeak_, atop_);
        *out = abase_ + off;
        return SJ_OK;
    }
    size_t c = size_class(bytes);
    auto it = free_.find(c);
    if (it != free_.end() && !it->second.empty()) {
        *out = it->second.back();
        it->second.pop_back();
        live_[*out] = c;
        return SJ_OK;
    }
    void *p = nullptr;
    auto get = [&]() { return pinned_ ? hipHostMalloc(&p, c, hipHostMallocDefault) : hipMalloc(&p, c); };
    hipError_t e = get();
    if (e != hipSuccess) {
        // release cached blocks and retry once
        (void)hipGetLastError();
        for (auto &kv : free_)
            for (void *q : kv.second) pool_free(pinned_, q);
        free_.clear();
        e = get();
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("%s allocation of %zu bytes failed", pinned_ ? "pinned host" : "device", c);
            return SJ_ERR_NOMEM;
        }
    }
    live_[p] = c;
    *out = p;
    return SJ_OK;
}

void Pool::release(void *p) {
    std::lock_guard<std::mutex> lk(mu_);
    char *q = static_cast<char *>(p);
    if (abase_ && q >= abase_ && q < abase_ + acap_) {
        const size_t off = (size_t)(q - abase_);
        for (int i = an_ - 1; i >= 0; i--)
            if (ablk_[i].live && ablk_[i].off == off) {
                ablk_[i].live = false;
                break;
            }
        while (an_ > 0 && !ablk_[an_ - 1].live) atop_ = ablk_[--an_].off;
        return;
    }
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_[it->second].push_back(p);
    live_.erase(it);
}

void Pool::trim() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : free_)
        for (void *p : kv.second) pool_free(pinned_, p);
    free_.clear();
}

int device_cu_count(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 256;
    return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
}

// -------------------------------------------------------------------- fill
template <class T>
__global__ void k_fill(T *p, uint64_t n, T v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 8192) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

sj_status fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v) {
    if (!n) return SJ_OK;
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

sj_status fill_u64(hipStream_t s, uint64_t *p, uint64_t n, uint64_t v) {
    if (!n) return SJ_OK;
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// -------------------------------------------------------------------- scan
// Exclusive scan in three launches: per-tile totals, scan of the totals in
// one block, per-tile rescan with the carried offset.  Tile = 256 x 16.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint64_t SCAN_TILE = (uint64_t)SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const unsigned lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T u = __shfl_up(v, off, 64);
        if (lane >= (unsigned)off) v += u;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *total) {
    __shared__ T wsum[SCAN_BLOCK / 64];
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x >> 6;
    T inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (unsigned w = 0; w < SCAN_BLOCK / 64; w++) {
        T x = wsum[w];
        if (w < wave) before += x;
        all += x;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

template <class Tin, class Tacc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const Tin *in, uint64_t n, Tacc *sums) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    Tacc acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        if (i < n) acc += (Tacc)in[i];
    }
    acc = wave_sum(acc);
    __shared__ Tacc ws[SCAN_BLOCK / 64];
    if (lane_id() == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Tacc t = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; w++) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

// single block: exclusive scan of sums[0..nb) in place; sums[nb] = total
template <class Tacc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(Tacc *sums, uint64_t nb) {
    Tacc carry = 0;
    for (uint64_t base = 0; base < nb; base += SCAN_BLOCK) {
        uint64_t i = base + threadIdx.x;
        Tacc v = i < nb ? sums[i] : 0;
        Tacc tot;
        Tacc ex = block_exclusive_scan(v, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

template <class Tin, class Tacc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const Tin *in, uint64_t n, const Tacc *sums,
                                                           Tacc *out, uint64_t nb) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    Tacc carry = sums[blockIdx.x];
    // blocked layout within the tile: thread t owns items [t*ITEMS, +ITEMS)
    Tacc v[SCAN_ITEMS];
    __shared__ Tin tile[SCAN_TILE];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        tile[j * SCAN_BLOCK + threadIdx.x] = i < n ? in[i] : (Tin)0;
    }
    __syncthreads();
    Tacc local = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = (Tacc)tile[threadIdx.x * SCAN_ITEMS + j];
        local += v[j];
    }
    Tacc tot;
    Tacc ex = block_exclusive_scan(local, &tot) + carry;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint64_t i = base + (uint64_t)threadIdx.x * SCAN_ITEMS + j;
        if (i < n) out[i] = ex;
        ex += v[j];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = sums[nb];
}

uint64_t scan_temp_count(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// Same scan with a caller-owned block-sum buffer of scan_temp_count(n)
// entries: allocates nothing, so it can be recorded into a graph.
sj_status scan_exclusive_u32_to_u64_tmp(hipStream_t s, const uint32_t *in, uint64_t n, uint64_t *out,
                                        uint64_t *sums) {
    if (n == 0) {
        hipLaunchKernelGGL(k_fill<uint64_t>, dim3(1), dim3(64), 0, s, out, (uint64_t)1, (uint64_t)0);
        SJ_LAUNCH_CHECK();
        return SJ_OK;
    }
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL((k_scan_reduce<uint32_t, uint64_t>), dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n,
                       sums);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scan_sums<uint64_t>), dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nb);
    SJ_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scan_apply<uint32_t, uint64_t>), dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n,
                       sums, out, nb);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

// -------------------------------------------------------------- radix sort
// LSD, 8-bit digits, three launches per pass.  Tile = 256 threads x 16
// rounds.  Upsweep counts the digits of every tile (hist[digit][tile]); one
// block per digit scans its row over the tiles (histScan) and keeps the
// digit's total; the downsweep takes each (digit, tile)'s output base as the
// digit's base (a 256-entry scan of the totals, done by every block) plus
// that prefix, ranks each round of 256 keys stably with wave64 ballots
// (peer mask of equal digits, popcount below the lane) plus a prefix over
// the four waves, places the tile's keys in LDS in digit order and writes
// them out run by run.
constexpr int RS_BLOCK = 256;
constexpr int RS_ROUNDS = 16;
constexpr uint64_t RS_TILE = (uint64_t)RS_BLOCK * RS_ROUNDS;
static_assert(RS_TILE == RADIX_TILE, "sort tile");

template <class K>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_hist(const K *keys, uint64_t n, int shift,
                                                         uint32_t *hist, uint32_t ntiles) {
    __shared__ uint32_t bins[256];
    bins[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int j = 0; j < RS_ROUNDS; j++) {
        uint64_t i = base + (uint64_t)j * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&bins[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = bins[threadIdx.x];
}

// block d: exclusive scan of hist[d][0..ntiles) into histScan[d][..], the
// row's total into totals[d]
constexpr int RD_ITEMS = 8;
__global__ __launch_bounds__(RS_BLOCK) void k_radix_digit_scan(const uint32_t *hist, uint32_t ntiles,
                                                               uint32_t *histScan, uint32_t *totals) {
    const uint64_t row = (uint64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < ntiles; c0 += RS_BLOCK * RD_ITEMS) {
        // thread t owns entries [c0 + t * ITEMS, + ITEMS)
        uint32_t v[RD_ITEMS], local = 0;
#pragma unroll
        for (int j = 0; j < RD_ITEMS; j++) {
            const uint32_t i = c0 + threadIdx.x * RD_ITEMS + j;
            v[j] = i < ntiles ? hist[row + i] : 0u;
            local += v[j];
        }
        uint32_t tot;
        uint32_t ex = carry + block_exclusive_scan<uint32_t>(local, &tot);
#pragma unroll
        for (int j = 0; j < RD_ITEMS; j++) {
            const uint32_t i = c0 + threadIdx.x * RD_ITEMS + j;
            if (i < ntiles) histScan[row + i] = ex;
            ex += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__device__ __forceinline__ uint32_t popc_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <class K>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_scatter(const K *kin, const uint32_t *vin, K *kout,
                                                            uint32_t *vout, uint64_t n, int shift,
                                                            const uint32_t *histScan, const uint32_t *hist,
                                                            const uint32_t *totals, uint32_t ntiles) {
    __shared__ uint32_t base[256];   // next output position of each digit
    __shared__ uint32_t gbase[256];  // the tile's first output position of each digit
    __shared__ uint32_t lbase[256];  // the tile's first slot of each digit in digit order
    __shared__ uint32_t wcnt[RS_BLOCK / 64][256];
    __shared__ K sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    const unsigned wave = threadIdx.x >> 6;
    uint32_t allKeys;
    const uint32_t digitBase = block_exclusive_scan<uint32_t>(totals[threadIdx.x], &allKeys);
    const uint32_t gb = digitBase + histScan[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    base[threadIdx.x] = gb;
    gbase[threadIdx.x] = gb;
    uint32_t tileCount;
    lbase[threadIdx.x] =
        block_exclusive_scan<uint32_t>(hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x], &tileCount);
    const uint64_t tbase = (uint64_t)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ROUNDS; j++) {
        const uint64_t i = tbase + (uint64_t)j * RS_BLOCK + threadIdx.x;
        const bool valid = i < n;
        K key = valid ? kin[i] : (K)0;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t dg = (uint32_t)(key >> shift) & 255u;
#pragma unroll
        for (int w = 0; w < RS_BLOCK / 64; w++) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = popc_below(peers);
        if (valid && rank == 0) wcnt[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t s = base[threadIdx.x];
#pragma unroll
            for (int w = 0; w < RS_BLOCK / 64; w++) {
                uint32_t c = wcnt[w][threadIdx.x];
                wcnt[w][threadIdx.x] = s;
                s += c;
            }
            base[threadIdx.x] = s;
        }
        __syncthreads();
        if (valid) {
            // the key's output position, and its slot in the tile's digit order
            const uint32_t pos = wcnt[wave][dg] + rank;
            const uint32_t lp = lbase[dg] + (pos - gbase[dg]);
            sk[lp] = key;
            sv[lp] = val;
        }
        __syncthreads();
    }
    // the tile in digit order: a run of equal digits goes to consecutive
    // output positions, so the writes coalesce
    const uint32_t cnt = (uint32_t)min<uint64_t>(RS_TILE, n - tbase);
    for (uint32_t i = threadIdx.x; i < cnt; i += RS_BLOCK) {
        const K key = sk[i];
        const uint32_t dg = (uint32_t)(key >> shift) & 255u;
        const uint32_t pos = gbase[dg] + (i - lbase[dg]);
        kout[pos] = key;
        vout[pos] = sv[i];
    }
}

uint32_t radix_tiles(uint64_t n) { return (uint32_t)((n + RS_TILE - 1) / RS_TILE); }

template <class K>
static sj_status radix_sort_pairs(Pool *pool, hipStream_t s, K **keys, K **keys_alt, uint32_t **vals,
                                  uint32_t **vals_alt, uint64_t n, int bits, const uint32_t *hist0) {
    if (n <= 1 || bits <= 0) return SJ_OK;
    if (n > 0xFFFFFFFFull) {
        set_error("radix sort supports at most 2^32-1 keys");
        return SJ_ERR_OVERFLOW;
    }
    const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    DBuf<uint32_t> hist, histScan;
    SJ_TRY(hist.alloc(pool, (uint64_t)256 * ntiles));
    SJ_TRY(histScan.alloc(pool, (uint64_t)256 * ntiles + 256));  // per-digit prefixes, then the totals
    uint32_t *totals = histScan.p + (uint64_t)256 * ntiles;
    for (int shift = 0; shift < bits; shift += 8) {
        // the lowest digit's histogram may come with the keys
        const uint32_t *h = shift == 0 && hist0 ? hist0 : hist.p;
        if (h == hist.p) {
            hipLaunchKernelGGL((k_radix_hist<K>), dim3(ntiles), dim3(RS_BLOCK), 0, s, *keys, n, shift, hist.p,
                               ntiles);
            SJ_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_radix_digit_scan, dim3(256), dim3(RS_BLOCK), 0, s, h, ntiles, histScan.p, totals);
        SJ_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_radix_scatter<K>), dim3(ntiles), dim3(RS_BLOCK), 0, s, *keys, *vals,
                           *keys_alt, *vals_alt, n, shift, histScan.p, h, totals, ntiles);
        SJ_LAUNCH_CHECK();
        std::swap(*keys, *keys_alt);
        std::swap(*vals, *vals_alt);
    }
    return SJ_OK;
}

sj_status radix_sort_pairs_u64(Pool *pool, hipStream_t s, uint64_t **keys, uint64_t **keys_alt,
                               uint32_t **vals, uint32_t **vals_alt, uint64_t n, int bits, const uint32_t *hist0) {
    return radix_sort_pairs<uint64_t>(pool, s, keys, keys_alt, vals, vals_alt, n, bits, hist0);
}

sj_status radix_sort_pairs_u32(Pool *pool, hipStream_t s, uint32_t **keys, uint32_t **keys_alt,
                               uint32_t **vals, uint32_t **vals_alt, uint64_t n, int bits, const uint32_t *hist0) {
    return radix_sort_pairs<uint32_t>(pool, s, keys, keys_alt, vals, vals_alt, n, bits, hist0);
}

}  // namespace sj
