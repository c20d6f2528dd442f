// Many independent self-joins (different n, d, k and eps) in one
// persistent launch per pass, recorded once into a hipGraph.  The count
// kernel keeps a hit record that the fill kernel replays.
//
// The reference runs one join per process (main.cu:36-226); a batch of
// small joins there is a sequence of under-filled launches.  Here a plan
// builds every instance's grid, FP32 copy and row-aligned tiles, lays the
// instances' rows end to end (instance i owns rows [base_i, base_i + n_i)
// of the shared count, offset and cursor arrays) and sizes the shared
// neighbour-ID array with one count run.  The graph then holds the whole
// join of every instance:
//   zero counters -> count kernel over all (instance, tile) work items ->
//   row totals -> one scan over all rows -> mirrored-hit cursors ->
//   fill kernel over all work items -> per-instance relative offsets.
// Each replay recomputes the joins of the planned grids.  The autotuner
// times replays over stencil form, schedule, persistent blocks per CU and
// queue chunk, and keeps the fastest.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "sj_traverse.hpp"

namespace sj {
const char *last_error();
sj_status grid_build(sj_context ctx, const double *d_points, uint64_t n, int d, int k, double eps,
                     sj_grid_s *g);

__global__ void k_many_totals(const uint32_t *own, const uint32_t *mir, uint64_t N, uint32_t *tot) {
    const uint64_t j = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (j < N) tot[j] = own[j] + mir[j];
}

__global__ void k_many_cursors(const uint64_t *G, const uint32_t *own, uint64_t N, unsigned long long *cur) {
    const uint64_t j = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (j < N) cur[j] = G[j] + own[j];
}

// rel[relBase_i + j] = G[base_i + j] - G[base_i], j in [0, n_i]; relBase_i =
// base_i + i; instance of an entry by binary search over relBase
__global__ void k_many_rebase(const uint64_t *G, const uint64_t *base, const uint64_t *relBase, int count,
                              uint64_t E, uint64_t *rel, uint64_t *idsBase) {
    const uint64_t e = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (e >= E) return;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (relBase[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const uint64_t j = e - relBase[lo];
    const uint64_t g0 = G[base[lo]];
    rel[e] = G[base[lo] + j] - g0;
    if (j == 0) idsBase[lo] = g0;
}

__global__ void k_iota_u32(uint32_t *v, uint64_t n) {
    const uint64_t j = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (j < n) v[j] = (uint32_t)j;
}

__global__ void k_gather_work(const uint32_t *order, const uint32_t *inst, const uint32_t *tile, uint64_t n,
                              uint32_t *instQ, uint32_t *tileQ) {
    const uint64_t j = (uint64_t)blockIdx.x * JB + threadIdx.x;
    if (j >= n) return;
    instQ[j] = inst[order[j]];
    tileQ[j] = tile[order[j]];
}

static unsigned nb(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + JB - 1) / JB); }

struct ManyInst {
    sj_grid_s *g = nullptr;
    uint64_t n = 0, base = 0;
    uint32_t nTiles = 0;
    int prec = 0;
    DBuf<uint32_t> tileStart, nTilesDev;
};

}  // namespace sj

using namespace sj;

struct sj_many_s {
    sj_context ctx = nullptr;
    sj_join_params P{};
    std::vector<ManyInst> inst;
    uint64_t N = 0, total = 0;
    uint32_t W = 0;
    std::vector<JoinArgs> args;
    DBuf<JoinArgs> dargs;
    DBuf<uint32_t> workInst, workTile, workInstQ, workTileQ;
    DBuf<uint32_t> own, mir, tot, qctr;
    DBuf<uint64_t> G, scanTmp, rel, base, relBase, idsBase;
    DBuf<unsigned long long> mirCur, stat;
    DBuf<int32_t> ids;
    sj_many_tuning tune{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> hIdsBase;
    std::vector<unsigned long long> hStat;
    double ms_plan = 0.0, ms_grid = 0.0, ms_run = 0.0;
    int mfma = 0;  // FP32 instances: 1 with, 2 without matrix-core walks
    // hit record shared by the instances: the count kernel records, the fill
    // kernel replays (the wide walk's instances evaluate again)
    bool rec = false;
    DBuf<uint32_t> tBound;
    DBuf<uint64_t> tileDir, recDir, recMask;
    DBuf<unsigned long long> recCtr;
    uint64_t dirCap = 0, maskCap = 0;

    ~sj_many_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &x : inst) delete x.g;
    }
};

namespace {

unsigned persistent_blocks(const sj_many_s *m) {
    const int bpc = m->tune.blocks_per_cu > 0 ? m->tune.blocks_per_cu : 8;
    const uint64_t want = (uint64_t)m->ctx->cu_count * bpc;
    const uint64_t need = std::max<uint64_t>(1, ((uint64_t)m->W + WPB - 1) / WPB);
    return (unsigned)std::min(want, need);
}

// The whole join of every instance, enqueued on the context stream.
// `fill` = false stops after the offsets (used to size the ID array).
sj_status enqueue_pipeline(sj_many_s *m, bool fill) {
    hipStream_t s = m->ctx->stream;
    const uint64_t N = m->N;
    const int cnt = (int)m->inst.size();
    const bool queue = m->tune.schedule == SJ_SCHED_WORKQUEUE;
    const uint32_t *wi = queue ? m->workInstQ.p : m->workInst.p;
    const uint32_t *wt = queue ? m->workTileQ.p : m->workTile.p;
    const uint32_t chunk = m->tune.queue_chunk > 0 ? (uint32_t)m->tune.queue_chunk : 1u;
    const unsigned blocks = persistent_blocks(m);
    SJ_TRY(fill_u32(s, m->own.p, N, 0u));
    SJ_TRY(fill_u32(s, m->mir.p, N, 0u));
    SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(m->stat.p), (uint64_t)cnt * 8, 0ull));
    SJ_TRY(fill_u32(s, m->qctr.p, 1, 0u));
    if (m->rec) SJ_TRY(fill_u64(s, reinterpret_cast<uint64_t *>(m->recCtr.p), REC_NCTR, 0ull));
    SJ_TRY(launch_join_many(m->dargs.p, wi, wt, m->W, m->qctr.p, chunk, queue ? 1 : 0, PASS_COUNT, m->mfma, blocks,
                           s));
    hipLaunchKernelGGL(k_many_totals, dim3(nb(N)), dim3(JB), 0, s, m->own.p, m->mir.p, N, m->tot.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(scan_exclusive_u32_to_u64_tmp(s, m->tot.p, N, m->G.p, m->scanTmp.p));
    if (!fill) return SJ_OK;
    hipLaunchKernelGGL(k_many_cursors, dim3(nb(N)), dim3(JB), 0, s, m->G.p, m->own.p, N, m->mirCur.p);
    SJ_LAUNCH_CHECK();
    SJ_TRY(fill_u32(s, m->qctr.p, 1, 0u));
    SJ_TRY(launch_join_many(m->dargs.p, wi, wt, m->W, m->qctr.p, chunk, queue ? 1 : 0, PASS_FILL, m->mfma, blocks,
                           s, m->rec));
    const uint64_t E = N + (uint64_t)cnt;
    hipLaunchKernelGGL(k_many_rebase, dim3(nb(E)), dim3(JB), 0, s, m->G.p, m->base.p, m->relBase.p, cnt, E,
                       m->rel.p, m->idsBase.p);
    SJ_LAUNCH_CHECK();
    return SJ_OK;
}

sj_status upload_args(sj_many_s *m) {
    for (size_t i = 0; i < m->inst.size(); i++) m->args[i].form = m->tune.form;
    SJ_TRY_HIP(hipMemcpyAsync(m->dargs.p, m->args.data(), m->args.size() * sizeof(JoinArgs),
                              hipMemcpyHostToDevice, m->ctx->stream));
    SJ_TRY_HIP(hipStreamSynchronize(m->ctx->stream));
    return SJ_OK;
}

sj_status capture(sj_many_s *m) {
    hipStream_t s = m->ctx->stream;
    if (m->exec) {
        (void)hipGraphExecDestroy(m->exec);
        m->exec = nullptr;
    }
    if (m->graph) {
        (void)hipGraphDestroy(m->graph);
        m->graph = nullptr;
    }
    SJ_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const sj_status st = enqueue_pipeline(m, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st != SJ_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error("many-join graph capture: %s", hipGetErrorString(e));
        return SJ_ERR_HIP;
    }
    m->graph = g;
    SJ_TRY_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    return SJ_OK;
}

// Mask words the record needed in the last count (the slices' demand).
sj_status rec_want(sj_many_s *m, uint64_t *want, uint64_t *bad) {
    std::vector<unsigned long long> hc(REC_NCTR);
    SJ_TRY_HIP(hipMemcpy(hc.data(), m->recCtr.p, REC_NCTR * 8, hipMemcpyDeviceToHost));
    uint64_t w = 0;
    for (uint32_t sl = 0; sl < REC_NSLICE; sl++) w += hc[REC_SLICE0 + sl * REC_STRIDE + 1];
    *want = w;
    *bad = hc[REC_BAD];
    return SJ_OK;
}

// Record space for `words` mask words (the slices may fill unevenly: the
// callers ask for several times the demand seen).
sj_status rec_space(sj_many_s *m, uint64_t words) {
    if (words <= m->maskCap) return SJ_OK;
    SJ_TRY(m->recMask.alloc(&m->ctx->pool, words));
    m->maskCap = words;
    for (auto &a : m->args) {
        a.recMask = m->recMask.p;
        a.maskCap = words;
    }
    return SJ_OK;
}

// One replay plus the checks: |S| unchanged, no inconsistent writes.  A
// record that outgrew its space (every fill item then reports an error)
// grows the space, the graph is recorded again and the replay repeated.
sj_status run_once(sj_many_s *m, double *ms) {
    hipStream_t s = m->ctx->stream;
    hipEvent_t e0, e1;
    SJ_TRY_HIP(hipEventCreate(&e0));
    SJ_TRY_HIP(hipEventCreate(&e1));
    (void)hipEventRecord(e0, s);
    hipError_t le = hipGraphLaunch(m->exec, s);
    (void)hipEventRecord(e1, s);
    if (le == hipSuccess) le = hipStreamSynchronize(s);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (le != hipSuccess) {
        set_error("many-join graph launch: %s", hipGetErrorString(le));
        return SJ_ERR_HIP;
    }
    if (ms) *ms = t;
    const size_t cnt = m->inst.size();
    uint64_t tot = 0;
    SJ_TRY_HIP(hipMemcpy(&tot, m->G.p + m->N, 8, hipMemcpyDeviceToHost));
    m->hStat.resize(cnt * 8);
    SJ_TRY_HIP(hipMemcpy(m->hStat.data(), m->stat.p, cnt * 8 * 8, hipMemcpyDeviceToHost));
    m->hIdsBase.resize(cnt);
    SJ_TRY_HIP(hipMemcpy(m->hIdsBase.data(), m->idsBase.p, cnt * 8, hipMemcpyDeviceToHost));
    if (tot != m->total) {
        set_error("many-join: |S| changed between plan (%llu) and run (%llu)", (unsigned long long)m->total,
                  (unsigned long long)tot);
        return SJ_ERR_STATE;
    }
    if (m->rec) {
        uint64_t want = 0, bad = 0;
        SJ_TRY(rec_want(m, &want, &bad));
        if (bad) {
            if (m->maskCap >= (1ull << 36)) {
                set_error("many-join: the hit record outgrew %llu words", (unsigned long long)m->maskCap);
                return SJ_ERR_NOMEM;
            }
            SJ_TRY(rec_space(m, std::max<uint64_t>(2 * m->maskCap, 3 * want)));
            SJ_TRY(upload_args(m));
            SJ_TRY(capture(m));
            return run_once(m, ms);
        }
    }
    for (size_t i = 0; i < cnt; i++)
        if (m->hStat[i * 8 + STAT_ERR]) {
            set_error("many-join: instance %zu: fill pass disagreed with the count pass", i);
            return SJ_ERR_STATE;
        }
    return SJ_OK;
}

sj_status plan(sj_many_s *m, const sj_instance *in, int count) {
    sj_context ctx = m->ctx;
    Pool *pool = &ctx->pool;
    hipStream_t s = ctx->stream;
    const sj_join_params &P = m->P;
    m->inst.resize(count);
    m->args.resize(count);
    uint64_t N = 0;
    uint64_t W = 0;
    for (int i = 0; i < count; i++) {
        const sj_instance &x = in[i];
        if (x.k < 1 || x.k > DEV_MAXK) {
            set_error("many-join: instance %d: k=%d outside 1..%d", i, x.k, DEV_MAXK);
            return SJ_ERR_UNSUPPORTED;
        }
        ManyInst &I = m->inst[i];
        I.g = new (std::nothrow) sj_grid_s();
        if (!I.g) return SJ_ERR_NOMEM;
        const auto tg = std::chrono::steady_clock::now();
        sj_status st = grid_build(ctx, x.points, x.n, x.d, x.k, x.eps, I.g);
        I.g->ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
        m->ms_grid += I.g->ms_build;
        if (st != SJ_OK) {
            set_error("many-join: instance %d: %s", i, last_error());
            return st;
        }
        I.n = x.n;
        I.base = N;
        N += x.n;
        JoinArgs a = base_args(I.g, P.form, eps_threshold(x.eps));
        int prec = P.precision;
        a.useTiles = 1;
        if (prec == SJ_PREC_FP32_RECHECK) SJ_TRY(prepare_fp32(I.g, a, &prec));
        if (P.precision == SJ_PREC_FP32_RECHECK && prec == SJ_PREC_FP64) {
            set_error("many-join: instance %d: the FP32 margin is not usable (d=%d, k=%d)", i, x.d, x.k);
            return SJ_ERR_UNSUPPORTED;
        }
        I.prec = prec;
        if (prec == SJ_PREC_FP64) {
            a.pts32 = nullptr;
            a.mDP = 0;
        }
        SJ_TRY(build_tiles(I.g, a, 0u, (uint32_t)x.n, I.tileStart, I.nTilesDev));
        a.tileStart = I.tileStart.p;
        a.nTilesDev = I.nTilesDev.p;
        m->args[i] = a;
    }
    SJ_TRY_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> hInst, hTile;
    for (int i = 0; i < count; i++) {
        uint32_t nt = 0;
        SJ_TRY_HIP(hipMemcpy(&nt, m->inst[i].nTilesDev.p, 4, hipMemcpyDeviceToHost));
        m->inst[i].nTiles = nt;
        W += nt;
    }
    if (W > 0xFFFFFFFFull || N >= 0xFFFFFFFFull) {
        set_error("many-join: %llu rows / %llu tiles exceed 32-bit indices", (unsigned long long)N,
                  (unsigned long long)W);
        return SJ_ERR_OVERFLOW;
    }
    m->N = N;
    m->W = (uint32_t)W;
    m->mfma = P.precision == SJ_PREC_FP32_RECHECK ? 1 : 0;
    if (m->mfma) {
        // no instance on the matrix cores: the FP32 LDS tile walks only
        bool anyMfma = false;
        for (int i = 0; i < count; i++) anyMfma = anyMfma || m->args[i].mDP > 0;
        if (!anyMfma) m->mfma = 2;
    }
    hInst.reserve(W);
    hTile.reserve(W);
    for (int i = 0; i < count; i++)
        for (uint32_t t = 0; t < m->inst[i].nTiles; t++) {
            hInst.push_back((uint32_t)i);
            hTile.push_back(t);
        }

    // shared row arrays
    SJ_TRY(m->own.alloc(pool, N));
    SJ_TRY(m->mir.alloc(pool, N));
    SJ_TRY(m->tot.alloc(pool, N));
    SJ_TRY(m->G.alloc(pool, N + 1));
    SJ_TRY(m->scanTmp.alloc(pool, scan_temp_count(N)));
    SJ_TRY(m->mirCur.alloc(pool, N));
    SJ_TRY(m->stat.alloc(pool, (uint64_t)count * 8));
    SJ_TRY(m->rel.alloc(pool, N + count));
    SJ_TRY(m->qctr.alloc(pool, 1));
    SJ_TRY(m->idsBase.alloc(pool, count));
    std::vector<uint64_t> hb(count), hrb(count);
    for (int i = 0; i < count; i++) {
        hb[i] = m->inst[i].base;
        hrb[i] = m->inst[i].base + (uint64_t)i;
    }
    SJ_TRY(m->base.alloc(pool, count));
    SJ_TRY(m->relBase.alloc(pool, count));
    SJ_TRY_HIP(hipMemcpy(m->base.p, hb.data(), count * 8, hipMemcpyHostToDevice));
    SJ_TRY_HIP(hipMemcpy(m->relBase.p, hrb.data(), count * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < count; i++) {
        JoinArgs &a = m->args[i];
        const uint64_t b = m->inst[i].base;
        a.ownCnt = m->own.p + b;
        a.mirCnt = m->mir.p + b;
        a.stat = m->stat.p + (uint64_t)i * 8;
        a.offS = m->G.p + b;
        a.outBase = 0;
        a.mirCur = m->mirCur.p + b;
    }

    // work items in instance order, and by descending tile cost for the queue
    SJ_TRY(m->workInst.alloc(pool, W));
    SJ_TRY(m->workTile.alloc(pool, W));
    SJ_TRY(m->workInstQ.alloc(pool, W));
    SJ_TRY(m->workTileQ.alloc(pool, W));
    if (W) {
        SJ_TRY_HIP(hipMemcpy(m->workInst.p, hInst.data(), W * 4, hipMemcpyHostToDevice));
        SJ_TRY_HIP(hipMemcpy(m->workTile.p, hTile.data(), W * 4, hipMemcpyHostToDevice));
        DBuf<uint32_t> key, key2, val, val2;
        SJ_TRY(key.alloc(pool, W));
        SJ_TRY(key2.alloc(pool, W));
        SJ_TRY(val.alloc(pool, W));
        SJ_TRY(val2.alloc(pool, W));
        uint64_t w0 = 0;
        for (int i = 0; i < count; i++) {
            SJ_TRY(launch_tile_cost(m->args[i], m->args[i].k, m->inst[i].nTiles, key.p + w0, val.p + w0, s));
            w0 += m->inst[i].nTiles;
        }
        hipLaunchKernelGGL(k_iota_u32, dim3(nb(W)), dim3(JB), 0, s, val.p, (uint64_t)W);
        SJ_LAUNCH_CHECK();
        uint32_t *kp = key.p, *ka = key2.p, *vp = val.p, *va = val2.p;
        SJ_TRY(radix_sort_pairs_u32(pool, s, &kp, &ka, &vp, &va, W, 8));
        hipLaunchKernelGGL(k_gather_work, dim3(nb(W)), dim3(JB), 0, s, vp, m->workInst.p, m->workTile.p,
                           (uint64_t)W, m->workInstQ.p, m->workTileQ.p);
        SJ_LAUNCH_CHECK();
        SJ_TRY_HIP(hipStreamSynchronize(s));
    }

    // hit record: per-tile bounds of every instance, end to end, scanned into
    // directory bases; one directory and word space for all instances
    if (P.hit_record >= 0 && W) {
        SJ_TRY(m->tBound.alloc(pool, W));
        SJ_TRY(m->tileDir.alloc(pool, (uint64_t)W + 1));
        uint64_t w0 = 0;
        for (int i = 0; i < count; i++) {
            const uint32_t cw = m->args[i].mDP > 0 ? 32u : 64u;
            SJ_TRY(launch_tile_bound(m->args[i], m->args[i].k, m->inst[i].nTiles, cw, m->tBound.p + w0, s));
            w0 += m->inst[i].nTiles;
        }
        SJ_TRY(scan_exclusive_u32_to_u64(pool, s, m->tBound.p, W, m->tileDir.p));
        SJ_TRY_HIP(hipMemcpyAsync(&m->dirCap, m->tileDir.p + W, 8, hipMemcpyDeviceToHost, s));
        SJ_TRY_HIP(hipStreamSynchronize(s));
        m->dirCap = std::max<uint64_t>(m->dirCap, 1);
        SJ_TRY(m->recDir.alloc(pool, (uint64_t)REC_DW * m->dirCap));
        SJ_TRY(m->recCtr.alloc(pool, REC_NCTR));
        SJ_TRY(rec_space(m, std::max<uint64_t>(1ull << 23, 16 * N)));
        w0 = 0;
        for (int i = 0; i < count; i++) {
            JoinArgs &a = m->args[i];
            a.recDir = m->recDir.p;
            a.recMask = m->recMask.p;
            a.tileDir = m->tileDir.p + w0;
            a.recCtr = m->recCtr.p;
            a.dirCap = m->dirCap;
            a.maskCap = m->maskCap;
            w0 += m->inst[i].nTiles;
        }
        m->rec = true;
    }

    // size the shared ID array with one eager count run (which also sizes
    // the hit record: three times the words it needed)
    SJ_TRY(m->dargs.alloc(pool, count));
    SJ_TRY(upload_args(m));
    SJ_TRY(enqueue_pipeline(m, false));
    SJ_TRY_HIP(hipStreamSynchronize(s));
    SJ_TRY_HIP(hipMemcpy(&m->total, m->G.p + N, 8, hipMemcpyDeviceToHost));
    if (m->rec) {
        uint64_t want = 0, bad = 0;
        SJ_TRY(rec_want(m, &want, &bad));
        SJ_TRY(rec_space(m, 3 * want));
    }
    SJ_TRY(m->ids.alloc(pool, m->total));
    for (int i = 0; i < count; i++) m->args[i].ids = m->ids.p;
    SJ_TRY(upload_args(m));
    SJ_TRY(capture(m));
    return run_once(m, &m->ms_run);
}

bool valid_tuning(const sj_many_tuning &t) {
    return t.form >= 0 && t.form <= 3 && (t.schedule == SJ_SCHED_STRIDED || t.schedule == SJ_SCHED_WORKQUEUE) &&
           t.blocks_per_cu >= 0 && t.blocks_per_cu <= 64 && t.queue_chunk >= 0 && t.queue_chunk <= 64;
}

}  // namespace

SJ_API sj_status sj_many_create(sj_context ctx, const sj_instance *instances, int32_t count,
                                const sj_join_params *params, sj_many *out) {
    if (!out || !ctx || !instances || count <= 0 || !params) {
        set_error("sj_many_create: invalid arguments");
        return SJ_ERR_INVALID;
    }
    *out = nullptr;
    if (params->form < 0 || params->form > 3 || params->precision < 0 || params->precision > 1 ||
        params->schedule < 0 || params->schedule > 1 || (params->block_size != 0 && params->block_size != JB) ||
        params->capacity != 0) {
        set_error("sj_many_create: unsupported params (block_size 0/%d, capacity 0)", JB);
        return SJ_ERR_INVALID;
    }
    const auto t0 = std::chrono::steady_clock::now();
    sj_many m = new (std::nothrow) sj_many_s();
    if (!m) return SJ_ERR_NOMEM;
    m->ctx = ctx;
    m->P = *params;
    m->tune.form = params->form;
    m->tune.schedule = params->schedule;
    m->tune.blocks_per_cu = 8;
    m->tune.queue_chunk = params->queue_chunk > 0 && params->queue_chunk <= 64 ? params->queue_chunk : 1;
    const sj_status st = plan(m, instances, count);
    if (st != SJ_OK) {
        delete m;
        return st;
    }
    m->ms_plan = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = m;
    return SJ_OK;
}

SJ_API sj_status sj_many_run(sj_many m) {
    if (!m) {
        set_error("sj_many_run: null plan");
        return SJ_ERR_INVALID;
    }
    return run_once(m, &m->ms_run);
}

SJ_API sj_status sj_many_set_tuning(sj_many m, const sj_many_tuning *t) {
    if (!m || !t || !valid_tuning(*t)) {
        set_error("sj_many_set_tuning: invalid tuning");
        return SJ_ERR_INVALID;
    }
    m->tune = *t;
    SJ_TRY(upload_args(m));
    SJ_TRY(capture(m));
    return run_once(m, &m->ms_run);
}

SJ_API sj_status sj_many_autotune(sj_many m, int32_t tune_form, int32_t reps, sj_many_tuning *chosen) {
    if (!m) {
        set_error("sj_many_autotune: null plan");
        return SJ_ERR_INVALID;
    }
    if (reps <= 0) reps = 3;
    auto measure = [&](const sj_many_tuning &t, double *best) -> sj_status {
        m->tune = t;
        SJ_TRY(upload_args(m));
        SJ_TRY(capture(m));
        double b = 1e30;
        for (int r = 0; r < reps; r++) {
            double ms = 0.0;
            SJ_TRY(run_once(m, &ms));
            b = std::min(b, ms);
        }
        *best = b;
        return SJ_OK;
    };
    const sj_many_tuning start = m->tune;
    sj_many_tuning best = start;
    double bestMs = 1e30;
    // stage 1: stencil form x schedule
    std::vector<int> forms;
    if (tune_form) forms = {SJ_FORM_BUNICOMP, SJ_FORM_LID, SJ_FORM_FULL};
    else forms = {start.form};
    for (int f : forms)
        for (int sc : {SJ_SCHED_STRIDED, SJ_SCHED_WORKQUEUE}) {
            sj_many_tuning t = start;
            t.form = f;
            t.schedule = sc;
            t.queue_chunk = 1;
            t.blocks_per_cu = 8;
            double ms;
            SJ_TRY(measure(t, &ms));
            if (ms < bestMs) {
                bestMs = ms;
                best = t;
            }
        }
    // stage 2: persistent blocks per CU x queue chunk
    for (int bpc : {4, 8, 16, 32})
        for (int ch : {1, 2, 4}) {
            if (best.schedule == SJ_SCHED_STRIDED && ch != 1) continue;
            sj_many_tuning t = best;
            t.blocks_per_cu = bpc;
            t.queue_chunk = ch;
            double ms;
            SJ_TRY(measure(t, &ms));
            if (ms < bestMs) {
                bestMs = ms;
                best = t;
            }
        }
    m->tune = best;
    SJ_TRY(upload_args(m));
    SJ_TRY(capture(m));
    SJ_TRY(run_once(m, &m->ms_run));
    if (chosen) *chosen = best;
    return SJ_OK;
}

SJ_API sj_status sj_many_get_tuning(sj_many m, sj_many_tuning *t) {
    if (!m || !t) {
        set_error("sj_many_get_tuning: null argument");
        return SJ_ERR_INVALID;
    }
    *t = m->tune;
    return SJ_OK;
}

SJ_API sj_status sj_many_instance(sj_many m, int32_t i, const uint64_t **d_offsets, const int32_t **d_ids,
                                  const uint32_t **d_rowIds, uint64_t *total) {
    if (!m || i < 0 || (size_t)i >= m->inst.size()) {
        set_error("sj_many_instance: no instance %d", i);
        return SJ_ERR_INVALID;
    }
    const ManyInst &I = m->inst[i];
    const uint64_t g0 = m->hIdsBase[i];
    const uint64_t g1 = (size_t)i + 1 < m->inst.size() ? m->hIdsBase[i + 1] : m->total;
    if (d_offsets) *d_offsets = m->rel.p + I.base + (uint64_t)i;
    if (d_ids) *d_ids = m->ids.p + g0;
    if (d_rowIds) *d_rowIds = I.g->sortedIds.p;
    if (total) *total = g1 - g0;
    return SJ_OK;
}

SJ_API sj_status sj_many_stats(sj_many m, int32_t i, sj_stats *st) {
    if (!m || !st || i < -1 || (i >= 0 && (size_t)i >= m->inst.size())) {
        set_error("sj_many_stats: invalid arguments");
        return SJ_ERR_INVALID;
    }
    memset(st, 0, sizeof *st);
    const size_t lo = i < 0 ? 0 : (size_t)i, hi = i < 0 ? m->inst.size() : (size_t)i + 1;
    for (size_t j = lo; j < hi; j++) {
        st->candidates += m->hStat[j * 8 + STAT_CAND];
        st->fp32_band_pairs += m->hStat[j * 8 + STAT_BAND];
        st->fp32_misclassified += m->hStat[j * 8 + STAT_MIS];
        const uint64_t g0 = m->hIdsBase[j];
        const uint64_t g1 = j + 1 < m->inst.size() ? m->hIdsBase[j + 1] : m->total;
        st->total_pairs += g1 - g0;
        st->ms_grid += m->inst[j].g->ms_build;
    }
    st->num_batches = 1;
    return SJ_OK;
}

SJ_API sj_status sj_many_times(sj_many m, double *ms_plan, double *ms_grid, double *ms_run) {
    if (!m) {
        set_error("sj_many_times: null plan");
        return SJ_ERR_INVALID;
    }
    if (ms_plan) *ms_plan = m->ms_plan;
    if (ms_grid) *ms_grid = m->ms_grid;
    if (ms_run) *ms_run = m->ms_run;
    return SJ_OK;
}

SJ_API int32_t sj_many_count(sj_many m) { return m ? (int32_t)m->inst.size() : 0; }

SJ_API void sj_many_free(sj_many m) { delete m; }
