// extern "C" entry points of libsjamd (declared in include/sjamd.h).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>

#include <thread>

#include "sj_common.hpp"
#include "sj_traverse.hpp"

namespace sj {
const char *last_error();
sj_status grid_build(sj_context ctx, const double *d_points, uint64_t n, int d, int k, double eps,
                     sj_grid_s *g);
sj_status grid_nd_mask_host(sj_grid_s *g, std::vector<uint32_t> &mask, std::vector<uint32_t> &offsets);
sj_status join(sj_grid_s *g, const sj_join_params *P, sj_result_s *r);
sj_status grid_workload(sj_grid_s *g, int lid, int64_t *nb, uint32_t *order);
sj_status batch_sample_count(sj_grid_s *g, uint64_t stride, uint64_t *count);
sj_status batch_sample_count_wq(sj_grid_s *g, int lid, uint64_t stride, uint64_t *count);
double eps_threshold(double eps);
sj_status import_csv(const char *path, int d, double **pts, uint64_t *n);
sj_status sort_in_nd_bins(double *pts, uint64_t n, int d, int stable);
sj_status result_to_orig(sj_result_s *r);
void static_scratch_init(int device);
}  // namespace sj

using namespace sj;

namespace {

struct ResultHostCache {
    std::vector<uint64_t> offsets;
    std::vector<int32_t> ids;
};

// device-resident table downloaded as one host batch (rows in cell order)
// the device table as one host batch: three downloads into pinned memory
// from the context's pool, queued together on the context's stream
sj_status device_as_batch(sj_result r, sj_batch_host &B) {
    const uint64_t n = r->n;
    hipStream_t s = r->ctx->stream;
    SJ_TRY(B.rowOffsets.alloc(&r->ctx->hpool, n + 1));
    SJ_TRY(B.rowIds.alloc(&r->ctx->hpool, n));
    SJ_TRY(B.ids.alloc(&r->ctx->hpool, r->total));
    SJ_TRY_HIP(hipMemcpyAsync(B.rowOffsets.p, r->d_offsets.p, (n + 1) * 8, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipMemcpyAsync(B.rowIds.p, r->d_rowIds.p, n * 4, hipMemcpyDeviceToHost, s));
    if (r->total) SJ_TRY_HIP(hipMemcpyAsync(B.ids.p, r->d_ids.p, r->total * 4, hipMemcpyDeviceToHost, s));
    SJ_TRY_HIP(hipStreamSynchronize(s));
    return SJ_OK;
}

// host copy in original point-ID order
sj_status ensure_host_csr(sj_result r, std::vector<uint64_t> &off, std::vector<int32_t> &ids) {
    off.resize(r->n + 1);
    ids.resize(r->total);
    if (r->on_device) {
        // materialise the host copy once, kept as batch 0
        if (r->batches.empty()) {
            r->batches.resize(1);
            SJ_TRY(device_as_batch(r, r->batches[0]));
        }
    }
    std::vector<uint64_t> cnt(r->n, 0);
    for (auto &B : r->batches)
        for (size_t q = 0; q < B.rowIds.n; q++) cnt[B.rowIds.p[q]] = B.rowOffsets.p[q + 1] - B.rowOffsets.p[q];
    uint64_t s = 0;
    for (uint64_t i = 0; i < r->n; i++) {
        off[i] = s;
        s += cnt[i];
    }
    off[r->n] = s;
    for (auto &B : r->batches)
        for (size_t q = 0; q < B.rowIds.n; q++)
            std::copy(B.ids.p + B.rowOffsets.p[q], B.ids.p + B.rowOffsets.p[q + 1], ids.begin() + off[B.rowIds.p[q]]);
    return SJ_OK;
}

}  // namespace

extern "C" {

SJ_API const char *sj_last_error(void) { return sj::last_error(); }

SJ_API double sj_eps_threshold(double eps) { return sj::eps_threshold(eps); }

SJ_API int32_t sj_fp32_band(double T, double R, int32_t D, float *lo, float *hi, float *t32) {
    const Fp32Margin m = fp32_margin(T, R, D);
    if (lo) *lo = m.lo;
    if (hi) *hi = m.hi;
    if (t32) *t32 = m.t32;
    return m.usable ? 1 : 0;
}

SJ_API int32_t sj_mfma_band(double T, double R, int32_t d, float *alpha, float *beta, float *tlo, float *thi) {
    if (mfma_padded_dims(d) == 0) return 0;
    const MfmaBand b = mfma_band(T, R, d);
    if (alpha) *alpha = b.alpha;
    if (beta) *beta = b.beta;
    if (tlo) *tlo = b.tlo;
    if (thi) *thi = b.thi;
    return b.DP;
}

SJ_API sj_status sj_context_create(int32_t device, void *stream, sj_context *out) {
    if (!out) {
        set_error("sj_context_create: null output");
        return SJ_ERR_INVALID;
    }
    *out = nullptr;
    int ndev = 0;
    SJ_TRY_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("sj_context_create: device %d of %d", device, ndev);
        return SJ_ERR_INVALID;
    }
    SJ_TRY_HIP(hipSetDevice(device));
    // the device address of this device's allocation-free scratch, looked up
    // once per device with the device current
    static_scratch_init(device);
    sj_context c = new (std::nothrow) sj_context_s();
    if (!c) return SJ_ERR_NOMEM;
    c->device = device;
    if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
        c->own_stream = false;
    } else {
        // a blocking stream: its work is ordered after earlier work of the
        // legacy default stream, where a caller may have produced d_points
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);
        if (e != hipSuccess) {
            delete c;
            set_error("hipStreamCreate: %s", hipGetErrorString(e));
            return SJ_ERR_HIP;
        }
        c->own_stream = true;
    }
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&c->pinned), 64 * sizeof(uint64_t),
                                 hipHostMallocDefault);
    if (e != hipSuccess) {
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        set_error("hipHostMalloc: %s", hipGetErrorString(e));
        return SJ_ERR_NOMEM;
    }
    e = hipEventCreateWithFlags(&c->ev, hipEventDisableTiming);
    for (int i = 0; i < 10 && e == hipSuccess; i++) e = hipEventCreate(&c->tev[i]);
    if (e != hipSuccess) {
        for (int i = 0; i < 10; i++)
            if (c->tev[i]) (void)hipEventDestroy(c->tev[i]);
        if (c->ev) (void)hipEventDestroy(c->ev);
        (void)hipHostFree(c->pinned);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        set_error("hipEventCreate: %s", hipGetErrorString(e));
        return SJ_ERR_HIP;
    }
    c->cu_count = device_cu_count(device);
    *out = c;
    return SJ_OK;
}

SJ_API sj_status sj_context_set_stream(sj_context ctx, void *stream) {
    if (!ctx || !stream) {
        set_error("sj_context_set_stream: null argument");
        return SJ_ERR_INVALID;
    }
    SJ_TRY_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    ctx->stream = static_cast<hipStream_t>(stream);
    ctx->own_stream = false;
    return SJ_OK;
}

SJ_API void sj_context_destroy(sj_context ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copyStream) {
        (void)hipStreamSynchronize(ctx->copyStream);
        (void)hipStreamDestroy(ctx->copyStream);
    }
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->ev) (void)hipEventDestroy(ctx->ev);
    for (int i = 0; i < 10; i++)
        if (ctx->tev[i]) (void)hipEventDestroy(ctx->tev[i]);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

SJ_API sj_status sj_grid_build(sj_context ctx, const double *d_points, uint64_t n, int32_t d, int32_t k,
                               double eps, sj_grid *out) {
    if (!out) {
        set_error("sj_grid_build: null output");
        return SJ_ERR_INVALID;
    }
    *out = nullptr;
    sj_grid g = new (std::nothrow) sj_grid_s();
    if (!g) return SJ_ERR_NOMEM;
    const auto t0 = std::chrono::steady_clock::now();
    sj_status s = grid_build(ctx, d_points, n, d, k, eps, g);
    if (s != SJ_OK) {
        delete g;
        return s;
    }
    g->ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = g;
    return SJ_OK;
}

SJ_API sj_status sj_grid_get_info(sj_grid g, sj_grid_info *info) {
    if (!g || !info) {
        set_error("sj_grid_get_info: null argument");
        return SJ_ERR_INVALID;
    }
    memset(info, 0, sizeof *info);
    info->n = g->n;
    info->d = g->d;
    info->k = g->k;
    info->eps = g->eps;
    for (int j = 0; j < g->k; j++) {
        info->minArr[j] = g->gp.minArr[j];
        info->maxArr[j] = g->gp.maxArr[j];
        info->nCells[j] = g->gp.nCells[j];
    }
    info->totalCells = g->gp.totalCells;
    info->nNonEmpty = g->nNonEmpty;
    info->dense_lookup = g->dense ? 1 : 0;
    return SJ_OK;
}

SJ_API sj_status sj_grid_copy_arrays(sj_grid g, uint64_t *cellLinear, int32_t *indexMin, int32_t *indexMax,
                                     uint32_t *indexLookup, uint32_t *cellCoord) {
    if (!g) {
        set_error("sj_grid_copy_arrays: null grid");
        return SJ_ERR_INVALID;
    }
    const uint64_t nne = g->nNonEmpty;
    SJ_TRY_HIP(hipStreamSynchronize(g->ctx->stream));
    if (cellLinear) SJ_TRY_HIP(hipMemcpy(cellLinear, g->cellLinear.p, nne * 8, hipMemcpyDeviceToHost));
    if (indexMin || indexMax) {
        std::vector<uint32_t> st(nne + 1);
        SJ_TRY_HIP(hipMemcpy(st.data(), g->cellStart.p, (nne + 1) * 4, hipMemcpyDeviceToHost));
        for (uint64_t c = 0; c < nne; c++) {
            if (indexMin) indexMin[c] = (int32_t)st[c];
            if (indexMax) indexMax[c] = (int32_t)st[c + 1] - 1;
        }
    }
    if (indexLookup) SJ_TRY_HIP(hipMemcpy(indexLookup, g->sortedIds.p, g->n * 4, hipMemcpyDeviceToHost));
    if (cellCoord)
        SJ_TRY_HIP(hipMemcpy(cellCoord, g->cellCoord.p, nne * (uint64_t)g->k * 4, hipMemcpyDeviceToHost));
    return SJ_OK;
}

SJ_API sj_status sj_grid_nd_mask(sj_grid g, uint32_t *mask, uint32_t *offsets, uint64_t *nElems) {
    if (!g) {
        set_error("sj_grid_nd_mask: null grid");
        return SJ_ERR_INVALID;
    }
    std::vector<uint32_t> m, o;
    SJ_TRY_HIP(hipStreamSynchronize(g->ctx->stream));
    SJ_TRY(grid_nd_mask_host(g, m, o));
    if (nElems) *nElems = m.size();
    if (mask) std::copy(m.begin(), m.end(), mask);
    if (offsets) std::copy(o.begin(), o.end(), offsets);
    return SJ_OK;
}

SJ_API void sj_grid_free(sj_grid g) {
    if (!g) return;
    (void)hipStreamSynchronize(g->ctx->stream);
    delete g;
}

SJ_API sj_status sj_join(sj_grid g, const sj_join_params *params, sj_result *out) {
    if (!g || !params || !out) {
        set_error("sj_join: null argument");
        return SJ_ERR_INVALID;
    }
    *out = nullptr;
    sj_result r = new (std::nothrow) sj_result_s();
    if (!r) return SJ_ERR_NOMEM;
    sj_status s = join(g, params, r);
    if (s != SJ_OK) {
        // no kernel or batch download may still touch the result's buffers
        (void)hipStreamSynchronize(g->ctx->stream);
        if (g->ctx->copyStream) (void)hipStreamSynchronize(g->ctx->copyStream);
        delete r;
        return s;
    }
    r->stats.ms_grid = g->ms_build;
    *out = r;
    return SJ_OK;
}

SJ_API sj_status sj_self_join(sj_context ctx, const double *d_points, uint64_t n, int32_t d, int32_t k,
                              double eps, const sj_join_params *params, sj_result *out) {
    sj_grid g = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    SJ_TRY(sj_grid_build(ctx, d_points, n, d, k, eps, &g));
    g->transient = true;
    sj_status s = sj_join(g, params, out);
    sj_grid_free(g);
    if (s == SJ_OK)
        (*out)->stats.ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

SJ_API sj_status sj_result_get_stats(sj_result r, sj_stats *st) {
    if (!r || !st) {
        set_error("sj_result_get_stats: null argument");
        return SJ_ERR_INVALID;
    }
    *st = r->stats;
    return SJ_OK;
}

SJ_API uint64_t sj_result_total_pairs(sj_result r) { return r ? r->total : 0; }
SJ_API uint64_t sj_result_num_points(sj_result r) { return r ? r->n : 0; }

SJ_API sj_status sj_result_device_csr(sj_result r, const uint64_t **d_offsets, const int32_t **d_ids) {
    if (!r) {
        set_error("sj_result_device_csr: null result");
        return SJ_ERR_INVALID;
    }
    if (!r->on_device) {
        set_error("sj_result_device_csr: result was produced in host batches");
        return SJ_ERR_STATE;
    }
    SJ_TRY(result_to_orig(r));
    if (d_offsets) *d_offsets = r->d_offsetsOrig.p;
    if (d_ids) *d_ids = r->d_idsOrig.p;
    return SJ_OK;
}

SJ_API sj_status sj_result_device_table(sj_result r, const uint64_t **d_offsets, const int32_t **d_ids,
                                        const uint32_t **d_rowIds) {
    if (!r) {
        set_error("sj_result_device_table: null result");
        return SJ_ERR_INVALID;
    }
    if (!r->on_device) {
        set_error("sj_result_device_table: result was produced in host batches");
        return SJ_ERR_STATE;
    }
    if (d_offsets) *d_offsets = r->d_offsets.p;
    if (d_ids) *d_ids = r->d_ids.p;
    if (d_rowIds) *d_rowIds = r->d_rowIds.p;
    return SJ_OK;
}

SJ_API sj_status sj_result_copy_csr(sj_result r, uint64_t *offsets, int32_t *ids) {
    if (!r) {
        set_error("sj_result_copy_csr: null result");
        return SJ_ERR_INVALID;
    }
    std::vector<uint64_t> off;
    std::vector<int32_t> v;
    SJ_TRY(ensure_host_csr(r, off, v));
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    if (ids) std::copy(v.begin(), v.end(), ids);
    return SJ_OK;
}

SJ_API sj_status sj_result_copy_pairs(sj_result r, int32_t *keys, int32_t *vals) {
    if (!r) {
        set_error("sj_result_copy_pairs: null result");
        return SJ_ERR_INVALID;
    }
    std::vector<uint64_t> off;
    std::vector<int32_t> v;
    SJ_TRY(ensure_host_csr(r, off, v));
    for (uint64_t p = 0; p < r->n; p++)
        for (uint64_t i = off[p]; i < off[p + 1]; i++) {
            if (keys) keys[i] = (int32_t)p;
            if (vals) vals[i] = v[i];
        }
    return SJ_OK;
}

SJ_API sj_status sj_result_neighbor_table(sj_result r, sj_neighbor_lookup *table) {
    if (!r || !table) {
        set_error("sj_result_neighbor_table: null argument");
        return SJ_ERR_INVALID;
    }
    if (r->on_device && r->batches.empty()) {
        r->batches.resize(1);
        SJ_TRY(device_as_batch(r, r->batches[0]));
    }
    uint64_t rows = 0;
    for (size_t b = 0; b < r->batches.size(); b++) {
        if (r->batches[b].ids.n > 0x7fffffffull) {
            set_error("neighbor table: batch above 2^31-1 pairs");
            return SJ_ERR_OVERFLOW;
        }
        rows += r->batches[b].rowIds.n;
    }
    // rows of no batch (none when the batches cover every point): empty
    if (rows != r->n)
        for (uint64_t p = 0; p < r->n; p++) table[p] = {(int32_t)p, 0, -1, -1};
    // each batch's rows, split over a few host threads for large batches
    for (size_t b = 0; b < r->batches.size(); b++) {
        const auto &B = r->batches[b];
        auto part = [&](uint64_t q0, uint64_t q1) {
            for (uint64_t q = q0; q < q1; q++)
                table[B.rowIds.p[q]] = {(int32_t)B.rowIds.p[q], (int32_t)B.rowOffsets.p[q],
                                        (int32_t)B.rowOffsets.p[q + 1] - 1, (int32_t)b};
        };
        const uint64_t m = B.rowIds.n;
        const unsigned nt = m >= (1u << 20) ? 8u : 1u;
        if (nt == 1) {
            part(0, m);
            continue;
        }
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) th.emplace_back(part, m * t / nt, m * (t + 1) / nt);
        for (auto &t : th) t.join();
    }
    return SJ_OK;
}

SJ_API sj_status sj_result_batch_data(sj_result r, uint64_t batch, const int32_t **ids, uint64_t *count) {
    if (!r || !ids || !count) {
        set_error("sj_result_batch_data: null argument");
        return SJ_ERR_INVALID;
    }
    if (r->on_device && r->batches.empty()) {
        r->batches.resize(1);
        SJ_TRY(device_as_batch(r, r->batches[0]));
    }
    if (batch >= r->batches.size()) {
        set_error("sj_result_batch_data: batch %llu of %zu", (unsigned long long)batch, r->batches.size());
        return SJ_ERR_INVALID;
    }
    *ids = r->batches[batch].ids.p;
    *count = r->batches[batch].ids.n;
    return SJ_OK;
}

SJ_API void sj_result_free(sj_result r) {
    if (!r) return;
    if (r->ctx) (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
}

SJ_API sj_status sj_workload(sj_grid g, int32_t lid, int64_t *nbPoints, uint32_t *order) {
    if (!g) {
        set_error("sj_workload: null grid");
        return SJ_ERR_INVALID;
    }
    return grid_workload(g, lid, nbPoints, order);
}

// Buffer/batch policy of GPU.cu:1029-1055 (callGPUBatchEst) and
// GPU.cu:1245-1271 (callGPUBatchEstWorkQueue): 3 streams, a 40M-pair
// buffer, alpha 0.05 (2 alpha and a third of the estimate per buffer when
// the estimate is below three buffers), buffer returned 1.5x.  An estimate
// of 0 gives one batch of one pair (the reference divides by zero there).
static void batch_policy(uint64_t cnt, uint64_t stride, uint64_t *sampleCount, uint64_t *estimated,
                         uint64_t *estimatedWithAlpha, uint64_t *numBatches, uint64_t *bufferSize) {
    const uint64_t est = cnt * stride;
    uint32_t buf = 40000000u;
    const double alpha = 0.05;
    const uint64_t withAlpha = (uint64_t)((double)est * (1.0 + alpha * 1.0));
    if (est < (uint64_t)buf * 3u) buf = (uint32_t)((double)est * (1.0 + alpha * 2.0) / 3.0);
    if (buf == 0) buf = 1;
    uint64_t nb = (uint64_t)std::ceil(((1.0 + alpha) * (double)est) / (double)buf);
    if (nb == 0) nb = 1;
    if (sampleCount) *sampleCount = cnt;
    if (estimated) *estimated = est;
    if (estimatedWithAlpha) *estimatedWithAlpha = withAlpha;
    if (numBatches) *numBatches = nb;
    if (bufferSize) *bufferSize = (uint64_t)(uint32_t)(1.5 * (double)buf);
}

SJ_API sj_status sj_batch_estimate(sj_grid g, uint64_t stride, uint64_t *sampleCount, uint64_t *estimated,
                                   uint64_t *estimatedWithAlpha, uint64_t *numBatches, uint64_t *bufferSize) {
    if (!g) {
        set_error("sj_batch_estimate: null grid");
        return SJ_ERR_INVALID;
    }
    if (stride == 0) stride = 100;
    uint64_t cnt = 0;
    SJ_TRY(batch_sample_count(g, stride, &cnt));
    batch_policy(cnt, stride, sampleCount, estimated, estimatedWithAlpha, numBatches, bufferSize);
    return SJ_OK;
}

SJ_API sj_status sj_batch_estimate_workqueue(sj_grid g, int32_t lid, uint64_t stride, uint64_t *sampleCount,
                                             uint64_t *estimated, uint64_t *estimatedWithAlpha,
                                             uint64_t *numBatches, uint64_t *bufferSize) {
    if (!g) {
        set_error("sj_batch_estimate_workqueue: null grid");
        return SJ_ERR_INVALID;
    }
    if (stride == 0) stride = 100;
    uint64_t cnt = 0;
    SJ_TRY(batch_sample_count_wq(g, lid ? 1 : 0, stride, &cnt));
    batch_policy(cnt, stride, sampleCount, estimated, estimatedWithAlpha, numBatches, bufferSize);
    return SJ_OK;
}

SJ_API sj_status sj_import_csv(const char *path, int32_t d, double **points, uint64_t *n) {
    return import_csv(path, d, points, n);
}

SJ_API sj_status sj_sort_in_nd_bins(double *points, uint64_t n, int32_t d, int32_t stable) {
    return sort_in_nd_bins(points, n, d, stable);
}

SJ_API void sj_free_host(void *p) { free(p); }

}  // extern "C"
