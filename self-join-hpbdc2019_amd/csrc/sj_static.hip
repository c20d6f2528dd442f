// The allocation-free self-join (sj_self_join_static): grid build and join
// whose device scratch comes from one file-scope device array instead of
// hipMalloc, for callers that cannot allocate (code under graph capture,
// tight serving loops, harnesses that forbid it).  The outputs are the
// caller's arrays.  No counterpart in the reference.
#include <cstring>
#include <mutex>

#include "sj_common.hpp"

namespace sj {
sj_status grid_build(sj_context ctx, const double *d_points, uint64_t n, int d, int k, double eps,
                     sj_grid_s *g);
sj_status join(sj_grid_s *g, const sj_join_params *P, sj_result_s *r);
int mfma_padded_dims(int d);
}  // namespace sj

using namespace sj;

// The scratch of every allocation-free call on a device.  Each device the
// library runs on holds its own copy until the process ends.
static constexpr size_t SCRATCH_BYTES = (size_t)SJ_STATIC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

// Device address of `scratch` per device, from hipGetSymbolAddress with
// that device current (the array's name is only a host-side stand-in).
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};
hipError_t g_rc[MAX_DEVICES] = {};
bool g_done[MAX_DEVICES] = {};

// Upper bounds of the device bytes the grid build (grid_need) and the join
// (join_need) draw from the pool for n points, every buffer counted as if
// none were released and padded to 256 bytes: the input check of the entry
// point.  The dense lookup table is not counted: it is built only when it
// fits beside the join's share.
constexpr uint64_t PAD = 256 * 64;  // alignment of up to 64 blocks

uint64_t scan_state(uint64_t n) { return n / 8 + (1u << 20); }  // look-back scan state
uint64_t radix_hist(uint64_t n) { return 3 * (n / 4) + 8192; }  // histograms of one sort (+ the first, given)

uint64_t grid_need(uint64_t n, int d, int k) {
    uint64_t b = 1024 * (2 * 12 + 1) * 8 + 1024;       // min/max partials, flags, grid params
    b += 4 * n * 8 + radix_hist(n);                     // keys, alternate keys (64-bit), values
    b += 2 * n * 4 + 4 + scan_state(n);                 // head flags, their scan
    b += n * 4 + (n + 1) * 4 + n * 8;                   // cellOf, cellStart, cellLinear
    b += n * (uint64_t)d * 8 + n * (uint64_t)k * 4;     // sorted points, cell coordinates
    return b + PAD;
}

uint64_t join_need(uint64_t n, int d, const sj_join_params *P) {
    uint64_t b = 0;
    if (P->precision == SJ_PREC_FP32_RECHECK) {         // FP32 copy (matrix cores: padded rows)
        const int dp = mfma_padded_dims(d);
        b += n * (uint64_t)(dp > d ? dp : d) * 4 + 256;
    }
    b += 3 * (n + 1) * 4 + 4 + scan_state(n);           // tiles: heads, scan, starts
    const uint64_t bound = n / 64 + 1 + n;              // work-queue entries
    if (P->schedule == SJ_SCHED_WORKQUEUE) b += 4 * bound * 4 + radix_hist(bound) + 8 * 64 * 4;
    b += 2 * n * 4 + 64;                                // own and mirrored counts, statistics
    b += scan_state(n) + 19 * 8 + n * 8;                // offsets scan, readback, mirrored cursors
    return b + PAD;
}

}  // namespace

namespace sj {
// Called by sj_context_create with `device` current: looks up this
// device's copy of the scratch once.
void static_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void *p = nullptr;
    g_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[device] = g_rc[device] == hipSuccess ? p : nullptr;
    g_done[device] = true;
}
}  // namespace sj

extern "C" {

SJ_API uint64_t sj_static_scratch_bytes(void) { return SCRATCH_BYTES; }

SJ_API uint64_t sj_static_need(uint64_t n, int32_t d, int32_t k, const sj_join_params *params) {
    if (!params) return 0;
    return grid_need(n, d, k) + join_need(n, d, params);
}

SJ_API sj_status sj_self_join_static(sj_context ctx, const double *d_points, uint64_t n, int32_t d, int32_t k,
                                     double eps, const sj_join_params *params, uint64_t *d_offsets,
                                     int32_t *d_ids, uint64_t ids_capacity, uint32_t *d_rowIds,
                                     uint64_t *total) {
    if (total) *total = 0;
    if (!ctx || !params || !d_offsets || !d_ids || !d_rowIds || !total) {
        set_error("sj_self_join_static: null argument");
        return SJ_ERR_INVALID;
    }
    if (params->capacity != 0) {
        set_error("sj_self_join_static: result batches (capacity > 0) need host buffers; use sj_self_join");
        return SJ_ERR_UNSUPPORTED;
    }
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES) {
        set_error("sj_self_join_static: device %d", ctx->device);
        return SJ_ERR_INVALID;
    }
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_done[ctx->device] || !g_ptr[ctx->device]) {
            set_error("sj_self_join_static: no scratch on device %d (hipGetSymbolAddress: %s)", ctx->device,
                      g_done[ctx->device] ? hipGetErrorString(g_rc[ctx->device]) : "not looked up");
            return SJ_ERR_HIP;
        }
        base = g_ptr[ctx->device];
    }
    const uint64_t need = grid_need(n, d, k) + join_need(n, d, params);
    if (need > SCRATCH_BYTES) {
        set_error("sj_self_join_static: %llu points of %d dims need up to %llu bytes of scratch; the static "
                  "scratch holds %llu (use sj_self_join)",
                  (unsigned long long)n, d, (unsigned long long)need, (unsigned long long)SCRATCH_BYTES);
        return SJ_ERR_OVERFLOW;
    }
    // no hit record: its space would not fit beside the grid
    sj_join_params P = *params;
    P.hit_record = -1;
    sj_status s;
    ctx->pool.arena_begin(base, SCRATCH_BYTES);
    {
        sj_grid_s g;
        sj_result_s r;
        g.reserveAfter = join_need(n, d, params);
        s = grid_build(ctx, d_points, n, d, k, eps, &g);
        if (s == SJ_OK) {
            r.extOffsets = d_offsets;
            r.extIds = d_ids;
            r.extIdsCap = ids_capacity;
            r.extRowIds = d_rowIds;
            s = join(&g, &P, &r);
            *total = r.total;
        }
        // every kernel of the call is done before the scratch is handed on
        (void)hipStreamSynchronize(ctx->stream);
    }
    ctx->pool.arena_end();
    return s;
}

}  // extern "C"
