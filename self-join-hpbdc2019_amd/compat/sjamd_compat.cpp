// Reference-named C++ host interface over the C ABI (see sjamd_compat.hpp).
#include "sjamd_compat.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ostream>
#include <vector>

#include "sjamd.h"

namespace {

int g_d = 2, g_k = 2;  // params.h:9-10 defaults
int g_status = 0;
sj_context g_ctx = nullptr;

bool fail(const char *where, sj_status s) {
    g_status = (int)s;
    if (s != SJ_OK) fprintf(stderr, "\nError: %s: %s (status %d)\n", where, sj_last_error(), (int)s);
    return s != SJ_OK;
}

sj_context context() {
    if (!g_ctx && fail("sj_context_create", sj_context_create(0, nullptr, &g_ctx))) return nullptr;
    return g_ctx;
}

// Points of the caller's vector<vector<>> in one device buffer.
struct DevicePoints {
    double *p = nullptr;
    uint64_t n = 0;
    bool ok = false;
    explicit DevicePoints(const std::vector<std::vector<DTYPE>> &v) {
        n = v.size();
        std::vector<double> flat(n * (size_t)g_d);
        for (uint64_t i = 0; i < n; i++) {
            if (v[i].size() < (size_t)g_d) {
                fprintf(stderr, "\nError: point %llu has %zu coordinates, GPUNUMDIM=%d\n",
                        (unsigned long long)i, v[i].size(), g_d);
                g_status = SJ_ERR_INVALID;
                return;
            }
            std::copy(v[i].begin(), v[i].begin() + g_d, flat.begin() + i * g_d);
        }
        if (hipMalloc(&p, std::max<size_t>(flat.size(), 1) * sizeof(double)) != hipSuccess) {
            g_status = SJ_ERR_NOMEM;
            p = nullptr;
            return;
        }
        if (!flat.empty() && hipMemcpy(p, flat.data(), flat.size() * sizeof(double), hipMemcpyHostToDevice) !=
                                 hipSuccess) {
            g_status = SJ_ERR_HIP;
            return;
        }
        ok = true;
    }
    ~DevicePoints() {
        if (p) (void)hipFree(p);
    }
};

sj_grid build_grid(std::vector<std::vector<DTYPE>> *pts, DTYPE eps, DevicePoints &dp) {
    sj_context ctx = context();
    if (!ctx || !dp.ok) return nullptr;
    sj_grid g = nullptr;
    if (fail("sj_grid_build", sj_grid_build(ctx, dp.p, dp.n, g_d, g_k, eps, &g))) return nullptr;
    return g;
}

}  // namespace

void sjamdSetDims(int gpunumdim, int numindexeddim) {
    g_d = gpunumdim;
    g_k = numindexeddim;
}
int sjamdGpuNumDim() { return g_d; }
int sjamdNumIndexedDim() { return g_k; }
int sjamdLastStatus() { return g_status; }

void warmUpGPU() { (void)context(); }

void importNDDataset(std::vector<std::vector<DTYPE>> *dataPoints, char *fname) {
    double *pts = nullptr;
    uint64_t n = 0;
    if (fail("importNDDataset", sj_import_csv(fname, g_d, &pts, &n))) return;
    printf("\nData import: Total data points: %llu", (unsigned long long)n);
    for (uint64_t i = 0; i < n; i++) dataPoints->emplace_back(pts + i * g_d, pts + (i + 1) * g_d);
    sj_free_host(pts);
}

void sortInNDBins(std::vector<std::vector<DTYPE>> *dataPoints) {
    const uint64_t n = dataPoints->size();
    if (!n) return;
    const int d = (int)(*dataPoints)[0].size();
    std::vector<double> flat(n * d);
    for (uint64_t i = 0; i < n; i++) std::copy((*dataPoints)[i].begin(), (*dataPoints)[i].end(), flat.begin() + i * d);
    if (fail("sortInNDBins", sj_sort_in_nd_bins(flat.data(), n, d, 0))) return;
    for (uint64_t i = 0; i < n; i++) std::copy(flat.begin() + i * d, flat.begin() + (i + 1) * d, (*dataPoints)[i].begin());
}

void generateNDGridDimensions(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon, DTYPE *minArr,
                              DTYPE *maxArr, unsigned int *nCells, uint64_t *totalCells) {
    DevicePoints dp(*NDdataPoints);
    sj_grid g = build_grid(NDdataPoints, epsilon, dp);
    if (!g) return;
    sj_grid_info info;
    if (!fail("sj_grid_get_info", sj_grid_get_info(g, &info))) {
        for (int j = 0; j < g_k; j++) {
            minArr[j] = info.minArr[j];
            maxArr[j] = info.maxArr[j];
            nCells[j] = info.nCells[j];
        }
        *totalCells = info.totalCells;
    }
    sj_grid_free(g);
}

void populateNDGridIndexAndLookupArray(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon,
                                       struct gridCellLookup **gridCellLookupArr, struct grid **index,
                                       unsigned int *indexLookupArr, DTYPE *minArr, unsigned int *nCells,
                                       uint64_t totalCells, unsigned int *nNonEmptyCells,
                                       unsigned int **gridCellNDMask, unsigned int *gridCellNDMaskOffsets,
                                       unsigned int *nNDMaskElems) {
    (void)minArr;
    (void)nCells;
    (void)totalCells;
    DevicePoints dp(*NDdataPoints);
    sj_grid g = build_grid(NDdataPoints, epsilon, dp);
    if (!g) return;
    sj_grid_info info;
    if (fail("sj_grid_get_info", sj_grid_get_info(g, &info))) {
        sj_grid_free(g);
        return;
    }
    const uint64_t nne = info.nNonEmpty;
    std::vector<uint64_t> lin(nne);
    std::vector<int32_t> imin(nne), imax(nne);
    if (fail("sj_grid_copy_arrays",
             sj_grid_copy_arrays(g, lin.data(), imin.data(), imax.data(), indexLookupArr, nullptr))) {
        sj_grid_free(g);
        return;
    }
    *index = new grid[nne];
    *gridCellLookupArr = new gridCellLookup[nne];
    for (uint64_t c = 0; c < nne; c++) {
        (*index)[c] = {imin[c], imax[c]};
        (*gridCellLookupArr)[c].idx = (unsigned int)c;
        (*gridCellLookupArr)[c].gridLinearID = lin[c];
    }
    uint64_t ne = 0;
    if (!fail("sj_grid_nd_mask", sj_grid_nd_mask(g, nullptr, nullptr, &ne))) {
        *gridCellNDMask = new unsigned int[ne];
        fail("sj_grid_nd_mask", sj_grid_nd_mask(g, *gridCellNDMask, gridCellNDMaskOffsets, &ne));
        *nNDMaskElems = (unsigned int)ne;
    }
    *nNonEmptyCells = (unsigned int)nne;
    sj_grid_free(g);
}

void distanceTableNDGridBatches(int searchMode, std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE *epsilon,
                                struct grid *index, struct gridCellLookup *gridCellLookupArr,
                                unsigned int *nNonEmptyCells, DTYPE *minArr, unsigned int *nCells,
                                unsigned int *indexLookupArr, struct neighborTableLookup *neighborTable,
                                std::vector<struct neighborDataPtrs> *pointersToNeighbors,
                                uint64_t *totalNeighbors, unsigned int *gridCellNDMask,
                                unsigned int *gridCellNDMaskOffsets, unsigned int *nNDMaskElems) {
    (void)index; (void)gridCellLookupArr; (void)nNonEmptyCells; (void)minArr; (void)nCells;
    (void)indexLookupArr; (void)gridCellNDMask; (void)gridCellNDMaskOffsets; (void)nNDMaskElems;
    *totalNeighbors = 0;
    sj_join_params P;
    memset(&P, 0, sizeof P);
    switch (searchMode) {  // main.cu:88-150
    case 3: P.form = SJ_FORM_FULL; P.schedule = SJ_SCHED_STRIDED; break;
    case 4: P.form = SJ_FORM_UNICOMP; P.schedule = SJ_SCHED_STRIDED; break;
    case 5: P.form = SJ_FORM_BUNICOMP; P.schedule = SJ_SCHED_STRIDED; break;
    case 6: P.form = SJ_FORM_LID; P.schedule = SJ_SCHED_STRIDED; break;
    // 7: full stencil over the workload-sorted points (here through the work
    // queue); 8: several lanes per point; 9: the work queue
    case 7: case 9: P.form = SJ_FORM_FULL; P.schedule = SJ_SCHED_WORKQUEUE; break;
    case 8: P.form = SJ_FORM_FULL; P.schedule = SJ_SCHED_THREADS_PER_POINT; break;
    case 10: P.form = SJ_FORM_LID; P.schedule = SJ_SCHED_WORKQUEUE; break;
    default:
        fprintf(stderr, "\nError: search mode %d is not in 3..10\n", searchMode);
        g_status = SJ_ERR_INVALID;
        return;
    }
    P.precision = SJ_PREC_FP64;
    DevicePoints dp(*NDdataPoints);
    sj_grid g = build_grid(NDdataPoints, *epsilon, dp);
    if (!g) return;
    // result buffer size from the reference estimators: the work-queue one
    // for modes 9 and 10, the strided sample otherwise (GPU.cu:1748-1758)
    uint64_t cnt = 0, est = 0, esta = 0, nb = 0, buf = 0;
    const sj_status est_st =
        (searchMode == 9 || searchMode == 10)
            ? sj_batch_estimate_workqueue(g, searchMode == 10, 100, &cnt, &est, &esta, &nb, &buf)
            : sj_batch_estimate(g, 100, &cnt, &est, &esta, &nb, &buf);
    if (fail("sj_batch_estimate", est_st)) {
        sj_grid_free(g);
        return;
    }
    P.capacity = buf;
    sj_result r = nullptr;
    sj_status st = sj_join(g, &P, &r);
    if (st == SJ_ERR_OVERFLOW) {
        // one row larger than the estimated buffer: no batching
        P.capacity = 0;
        st = sj_join(g, &P, &r);
    }
    sj_grid_free(g);
    if (fail("sj_join", st)) return;
    const uint64_t n = NDdataPoints->size();
    std::vector<sj_neighbor_lookup> tab(n);
    if (fail("sj_result_neighbor_table", sj_result_neighbor_table(r, tab.data()))) {
        sj_result_free(r);
        return;
    }
    sj_stats stats;
    sj_result_get_stats(r, &stats);
    std::vector<int *> data(stats.num_batches, nullptr);
    for (uint64_t b = 0; b < stats.num_batches; b++) {
        const int32_t *ids = nullptr;
        uint64_t count = 0;
        if (fail("sj_result_batch_data", sj_result_batch_data(r, b, &ids, &count))) {
            sj_result_free(r);
            return;
        }
        data[b] = new int[count ? count : 1];
        std::copy(ids, ids + count, data[b]);
        pointersToNeighbors->push_back({data[b], (int)count});
    }
    for (uint64_t p = 0; p < n; p++) {
        const sj_neighbor_lookup &t = tab[p];
        neighborTable[p].pointID = t.pointID;
        neighborTable[p].indexmin = t.indexmin;
        neighborTable[p].indexmax = t.indexmax;
        neighborTable[p].dataPtr = t.batch >= 0 ? data[t.batch] : nullptr;
    }
    *totalNeighbors = sj_result_total_pairs(r);
    printf("\n~~~TOTAL RESULT SET SIZE ON HOST:  %llu", (unsigned long long)*totalNeighbors);
    sj_result_free(r);
    g_status = 0;
}

namespace {
unsigned long long batch_est(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon, int workqueue, int lid,
                             unsigned int *retNumBatches, unsigned int *retGPUBufferSize) {
    DevicePoints dp(*NDdataPoints);
    sj_grid g = build_grid(NDdataPoints, epsilon, dp);
    if (!g) return 0;
    uint64_t cnt = 0, est = 0, esta = 0, nb = 0, buf = 0;
    const sj_status st = workqueue ? sj_batch_estimate_workqueue(g, lid, 100, &cnt, &est, &esta, &nb, &buf)
                                   : sj_batch_estimate(g, 100, &cnt, &est, &esta, &nb, &buf);
    sj_grid_free(g);
    if (fail("sj_batch_estimate", st)) return 0;
    printf("\nGPU: result set size for estimating the number of batches (sampled): %llu",
           (unsigned long long)cnt);
    printf("\nNumber of batches: %llu, buffer size: %llu", (unsigned long long)nb, (unsigned long long)buf);
    if (retNumBatches) *retNumBatches = (unsigned int)nb;
    if (retGPUBufferSize) *retGPUBufferSize = (unsigned int)buf;
    g_status = 0;
    return esta;
}
}  // namespace

unsigned long long callGPUBatchEst(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon,
                                   unsigned int *retNumBatches, unsigned int *retGPUBufferSize) {
    return batch_est(NDdataPoints, epsilon, 0, 0, retNumBatches, retGPUBufferSize);
}

unsigned long long callGPUBatchEstWorkQueue(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon,
                                            int searchMode, unsigned int *retNumBatches,
                                            unsigned int *retGPUBufferSize) {
    return batch_est(NDdataPoints, epsilon, 1, searchMode == 6 || searchMode == 10 ? 1 : 0, retNumBatches,
                     retGPUBufferSize);
}

// GPU.cu:2595-2637, without reading pointIDKey[cnt].
void constructNeighborTableKeyValueWithPtrs(int *pointIDKey, int *pointInDistValue,
                                            struct neighborTableLookup *neighborTable, int *pointersToNeighbors,
                                            unsigned int *cnt) {
    std::copy(pointInDistValue, pointInDistValue + *cnt, pointersToNeighbors);
    unsigned int i = 0;
    while (i < *cnt) {
        unsigned int j = i;
        while (j + 1 < *cnt && pointIDKey[j + 1] == pointIDKey[i]) j++;
        const int key = pointIDKey[i];
        neighborTable[key].pointID = key;
        neighborTable[key].indexmin = (int)i;
        neighborTable[key].indexmax = (int)j;
        neighborTable[key].dataPtr = pointersToNeighbors;
        i = j + 1;
    }
}

namespace {
struct ModeText {
    int mode;
    const char *file, *name;
};
// main.cu:88-150
const ModeText kModes[] = {
    {3, "gpu_global_stats.txt", "global memory kernel"},
    {4, "gpu_unicomp_stats.txt", "Unicomp"},
    {5, "gpu_b-unicomp_stats.txt", "B-Unicomp"},
    {6, "gpu_linID-unicomp_stats.txt", "Linear ID Unicomp"},
    {7, "gpu_sorting-cells_stats.txt", "sorting cells by workload"},
    {8, "gpu_dynamic-threads-fixed_stats.txt", "dynamic threads per points (fixed)"},
    {9, "gpu_working-queue_stats.txt", "working queue"},
    {10, "gpu_working-queue_lid-unicomp_stats.txt", "working queue + LID-Unicomp"},
};
}  // namespace

const char *searchModeStatsFile(int searchMode) {
    for (const auto &m : kModes)
        if (m.mode == searchMode) return m.file;
    return nullptr;
}

const char *searchModeName(int searchMode) {
    for (const auto &m : kModes)
        if (m.mode == searchMode) return m.name;
    return nullptr;
}

void writeStatsLine(std::ostream &out, double totalTime, const char *inputFname, DTYPE epsilon,
                    uint64_t totalNeighbors) {
    out << totalTime << ", " << inputFname << ", " << epsilon << ", " << totalNeighbors
        << ", GPUNUMDIM/NUMINDEXEDDIM/DTYPE/BLOCKSIZE/THperPOINT/SORT: " << g_d << ", " << g_k << ", "
        << "double" << ", " << 256 << ", " << 1 << ", " << 1 << std::endl;
}

void printNeighborTable(FILE *out, struct neighborTableLookup *neighborTable, size_t n) {
    for (size_t i = 0; i < n; i++) {
        neighborTableLookup &t = neighborTable[i];
        if (t.dataPtr && t.indexmax >= t.indexmin) std::sort(t.dataPtr + t.indexmin, t.dataPtr + t.indexmax + 1);
        fprintf(out, "\npoint id: %d, neighbors: ", (int)i);
        fprintf(out, "nb neighbor %d\n", t.indexmax - t.indexmin + 1);
        for (int j = t.indexmin; j < t.indexmax; j++) fprintf(out, "%d,", t.dataPtr[j]);
    }
    fprintf(out, "\n\n\n");
}
