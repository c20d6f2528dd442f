// Internal declarations shared by the HIP translation units of libsjamd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sjamd.h"

namespace sj {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);

#define SJ_TRY_HIP(expr)                                                          \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::sj::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                  \
            return SJ_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define SJ_TRY(expr)                        \
    do {                                    \
        sj_status s_ = (expr);              \
        if (s_ != SJ_OK) return s_;         \
    } while (0)

#define SJ_LAUNCH_CHECK()                                                         \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            ::sj::set_error("kernel launch failed: %s (%s:%d)",                   \
                            hipGetErrorString(e_), __FILE__, __LINE__);           \
            return SJ_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

// ------------------------------------------------------------ device pool
// Caching allocator: power-of-two size classes, freed blocks are kept for
// reuse so steady-state joins make no hipMalloc calls.  A pinned pool
// hands out page-locked host memory (hipHostMalloc) instead.
//
// Arena mode (arena_begin .. arena_end) serves every allocation from one
// caller-given device range instead, in order and 256-byte aligned, with a
// fixed bookkeeping table: no hipMalloc and no host allocation.  A release
// frees the topmost block and any released blocks under it.
class Pool {
  public:
    explicit Pool(bool pinned = false) : pinned_(pinned) {}
    ~Pool();
    sj_status alloc(size_t bytes, void **out);
    void release(void *p);
    void trim();
    void arena_begin(void *base, size_t bytes);
    void arena_end();
    bool arena_active() const { return abase_ != nullptr; }
    size_t arena_peak() const { return apeak_; }
    // bytes left above the topmost block (0 outside arena mode)
    size_t arena_free() const {
        const size_t off = (atop_ + 255) & ~(size_t)255;
        return abase_ && off < acap_ ? acap_ - off : 0;
    }

  private:
    bool pinned_;
    std::mutex mu_;
    std::map<size_t, std::vector<void *>> free_;
    std::map<void *, size_t> live_;
    static constexpr int ARENA_BLOCKS = 256;
    struct ArenaBlock {
        size_t off;
        bool live;
    };
    char *abase_ = nullptr;
    size_t acap_ = 0, atop_ = 0, apeak_ = 0;
    ArenaBlock ablk_[ARENA_BLOCKS];
    int an_ = 0;
};

// RAII device buffer drawn from a Pool.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    Pool *pool = nullptr;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), pool(o.pool) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        reset();
        p = o.p; n = o.n; pool = o.pool;
        o.p = nullptr; o.n = 0;
        return *this;
    }
    ~DBuf() { reset(); }
    sj_status alloc(Pool *pl, size_t count) {
        reset();
        pool = pl;
        void *q = nullptr;
        sj_status s = pool->alloc((count ? count : 1) * sizeof(T), &q);
        if (s != SJ_OK) return s;
        p = static_cast<T *>(q);
        n = count;
        return SJ_OK;
    }
    void reset() {
        if (p && pool) pool->release(p);
        p = nullptr;
        n = 0;
    }
};

// ---------------------------------------------------------- grid params
// Written by the device grid builder, read back once by the host.
struct GridParams {
    double minArr[SJ_MAXK];
    double maxArr[SJ_MAXK];
    uint32_t nCells[SJ_MAXK];
    uint64_t mult[SJ_MAXK];  // prod_{j<i} nCells[j]
    uint64_t totalCells;
    uint32_t flags;          // bit0 non-finite coordinate, bit1 overflow
    uint32_t pad;
    // written by the later build kernels: their error flags (GP_*) and the
    // non-empty cell count, read back with the parameters in one copy
    uint32_t err;
    uint32_t nne;
};
enum { GP_NONFINITE = 1u, GP_OVERFLOW = 2u };

// Small pinned host staging area for scalar readbacks.
struct HostStage {
    void *p = nullptr;
    size_t bytes = 0;
};

// ------------------------------------------------------------- launches
uint64_t scan_temp_count(uint64_t n);
sj_status scan_exclusive_u32_to_u64_tmp(hipStream_t s, const uint32_t *in, uint64_t n, uint64_t *out,
                                        uint64_t *sums);
sj_status scan_exclusive_u32_to_u64(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n,
                                    uint64_t *out /* n+1 */);
sj_status scan_exclusive_u32_to_u32(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n,
                                    uint32_t *out /* n+1 */);
// exclusive scan of the run-head flags of sorted keys (1 where keys[i] !=
// keys[i-1]), total (the number of runs) at out[n]
sj_status scan_exclusive_heads(Pool *pool, hipStream_t s, const uint32_t *keys, uint64_t n, uint32_t *out);
sj_status scan_exclusive_heads(Pool *pool, hipStream_t s, const uint64_t *keys, uint64_t n, uint32_t *out);
// exclusive scan of a[i] + b[i], total at out[n]
sj_status scan_exclusive_sum2_u32_to_u64(Pool *pool, hipStream_t s, const uint32_t *a, const uint32_t *b,
                                         uint64_t n, uint64_t *out /* n+1 */);
// Stable LSD radix sort of (key,val) by the low `bits` bits of key.  The
// sorted data ends in (*keys,*vals); the other buffers are scratch.  hist0
// (when set): the caller's counts of the lowest 8-bit digit per sort tile of
// RADIX_TILE keys, digit-major (hist0[digit * radix_tiles(n) + tile]).
constexpr uint32_t RADIX_TILE = 4096;
uint32_t radix_tiles(uint64_t n);
sj_status radix_sort_pairs_u64(Pool *pool, hipStream_t s, uint64_t **keys, uint64_t **keys_alt,
                               uint32_t **vals, uint32_t **vals_alt, uint64_t n, int bits,
                               const uint32_t *hist0 = nullptr);
sj_status radix_sort_pairs_u32(Pool *pool, hipStream_t s, uint32_t **keys, uint32_t **keys_alt,
                               uint32_t **vals, uint32_t **vals_alt, uint64_t n, int bits,
                               const uint32_t *hist0 = nullptr);
sj_status fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v);
sj_status fill_u64(hipStream_t s, uint64_t *p, uint64_t n, uint64_t v);

int device_cu_count(int device);

}  // namespace sj

// --------------------------------------------------------------- handles
struct sj_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    sj::Pool pool;
    sj::Pool hpool{true};        // pinned host memory: result batches, staging
    hipStream_t copyStream = nullptr;  // result-batch downloads (created on first use)
    uint64_t *pinned = nullptr;  // 64 scalars of pinned readback space
    hipEvent_t ev = nullptr;     // marks a readback into `pinned` the host waits for
    hipEvent_t tev[10] = {};     // phase timers of the joins (pairs)
    int cu_count = 256;
    // hit-record space of the joins on this context (directory entries,
    // mask words): grows to the demand a join reports when it did not fit
    uint64_t recDirCap = 0, recMaskCap = 0;
    // neighbour IDs of the largest result seen (plus a margin): a join whose
    // result fits fills without first reading |S| back to the host
    uint64_t idsHint = 0;
    // key width of the last grid built on this context, with its shape and
    // eps: a build of the same shape sorts with it before the grid
    // parameters reach the host, and sorts again if they need more bits
    struct {
        bool valid = false, wide = false;
        int d = 0, k = 0, bits = 0;
        uint64_t n = 0;
        double eps = 0.0;
    } gridHint;
};

struct sj_grid_s {
    sj_context ctx = nullptr;
    uint64_t n = 0;
    int d = 0, k = 0;
    double eps = 0.0;
    sj::GridParams gp{};
    uint64_t nNonEmpty = 0;
    bool dense = false;
    sj::DBuf<sj::GridParams> d_gp;
    sj::DBuf<double> sortedPts;      // n x d, cell order
    sj::DBuf<uint32_t> sortedIds;    // indexLookupArr
    sj::DBuf<uint32_t> cellOf;       // cell of each sorted position
    sj::DBuf<uint32_t> cellStart;    // nNonEmpty + 1
    sj::DBuf<uint64_t> cellLinear;   // nNonEmpty
    sj::DBuf<uint32_t> cellCoord;    // nNonEmpty x k
    sj::DBuf<uint32_t> posLB;        // totalCells+1 (dense grids): first sorted position with linear ID >= L
    sj::DBuf<float> sortedPts32;     // lazily built for the FP32 form
    double ms_build = 0.0;
    bool transient = false;          // made for one join, which may take its buffers
    // arena mode: device bytes the join still needs after the grid; the
    // dense lookup table is built only if it fits beside them
    uint64_t reserveAfter = 0;
};

struct sj_batch_host {
    sj::DBuf<int32_t> ids;            // neighbour IDs of the batch rows (pinned host)
    sj::DBuf<uint32_t> rowIds;        // original IDs of the rows (pinned host; .n rows)
    sj::DBuf<uint64_t> rowOffsets;    // rows+1, into ids (pinned host)
};

struct sj_result_s {
    sj_context ctx = nullptr;
    // caller-owned outputs (the allocation-free entry point): offsets[n+1],
    // ids[extIdsCap], rowIds[n]; null: the result allocates its own
    uint64_t *extOffsets = nullptr;
    int32_t *extIds = nullptr;
    uint32_t *extRowIds = nullptr;
    uint64_t extIdsCap = 0;
    uint64_t n = 0;
    uint64_t total = 0;
    bool on_device = false;
    sj::DBuf<uint64_t> d_offsets;      // n+1, rows in cell order
    sj::DBuf<int32_t> d_ids;           // total
    sj::DBuf<uint32_t> d_rowIds;       // original point ID of each row
    sj::DBuf<uint64_t> d_offsetsOrig;  // n+1, original ID order (built on demand)
    sj::DBuf<int32_t> d_idsOrig;
    std::vector<sj_batch_host> batches;
    sj_stats stats{};
};
