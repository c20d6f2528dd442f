/*
 * sjamd.h -- C ABI of the MI355X (gfx950) grid-indexed epsilon self-join.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes on every
 * call, no exits.  Device pointers are HIP device addresses; `stream` is a
 * hipStream_t passed as void* (NULL: the context creates its own).
 *
 * Ordering: every entry point enqueues its work on the context's stream
 * and reads the caller's device buffers there.  Input produced on another
 * stream must be complete, or that stream joined to the context's stream
 * (an event wait), before the call.  A context created with stream = NULL
 * uses a blocking stream, so work the caller enqueued earlier on the legacy
 * default stream (stream 0) is complete before the context's work starts.
 *
 * Each entry point names the reference (benoitgallet/self-join-hpbdc2019)
 * interface it replaces.
 */
#ifndef SJAMD_H
#define SJAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SJ_API __attribute__((visibility("default")))
#define SJ_MAXK 12 /* indexed dimensions supported by the 3^k stencil */
#define SJ_MAXD 64 /* point dimensions */

typedef enum sj_status {
    SJ_OK = 0,
    SJ_ERR_INVALID = 1,     /* bad argument (eps <= 0, NaN coordinate, k > d ...) */
    SJ_ERR_NOMEM = 2,       /* device or host allocation failed */
    SJ_ERR_HIP = 3,         /* HIP runtime error */
    SJ_ERR_OVERFLOW = 4,    /* cell count / coordinate / pair count overflow */
    SJ_ERR_UNSUPPORTED = 5, /* configuration not supported */
    SJ_ERR_IO = 6,          /* file could not be read */
    SJ_ERR_STATE = 7        /* result not available in the requested view */
} sj_status;

/* Traversal forms: reference search modes 3 (FULL), 4 (UNICOMP),
 * 5 (BUNICOMP), 6 and 10 (LID); main.cu:88-150, GPU.cu:2162-2283. */
typedef enum sj_form {
    SJ_FORM_FULL = 0,
    SJ_FORM_UNICOMP = 1,
    SJ_FORM_BUNICOMP = 2,
    SJ_FORM_LID = 3
} sj_form;

typedef enum sj_precision {
    SJ_PREC_FP64 = 0,        /* reference predicate, kernel.cu:341-349 */
    SJ_PREC_FP32_RECHECK = 1 /* FP32 bulk + FP64 re-decision of the margin band */
} sj_precision;

/* Query-point scheduling: strided batches (modes 3-6, kernel.cu:978), the
 * workload-sorted atomic work queue (modes 7, 9, 10, kernel.cu:2202-2236),
 * or several lanes per query point over the workload order (mode 8,
 * kernel.cu:2072-2188: 16, 8 and 2 lanes per point for the heaviest,
 * middle and lightest third; full stencil and FP64 only). */
typedef enum sj_schedule {
    SJ_SCHED_STRIDED = 0,
    SJ_SCHED_WORKQUEUE = 1,
    SJ_SCHED_THREADS_PER_POINT = 2
} sj_schedule;

typedef struct sj_context_s *sj_context;
typedef struct sj_grid_s *sj_grid;
typedef struct sj_result_s *sj_result;

typedef struct sj_join_params {
    int32_t form;         /* sj_form */
    int32_t precision;    /* sj_precision */
    int32_t schedule;     /* sj_schedule */
    int32_t sort_rows;    /* 1: each neighbour row ascending by point ID */
    uint64_t capacity;    /* result pairs per batch; 0 = whole result on device */
    int32_t block_size;   /* threads per block of the traversal kernels (0: 256) */
    int32_t queue_chunk;  /* query points per work-queue grab (0: 64) */
    int32_t cells_per_tile; /* reserved for tiled traversal (0: auto) */
    int32_t hit_record;   /* 0: the fill replays the hit masks the count pass
                           * recorded, when they fit the context's record
                           * space (which grows to the demand seen); -1: the
                           * fill evaluates every distance again */
} sj_join_params;

typedef struct sj_grid_info {
    uint64_t n;
    int32_t d, k;
    double eps;
    double minArr[SJ_MAXK];   /* min - eps per indexed dim */
    double maxArr[SJ_MAXK];   /* max + eps per indexed dim */
    uint32_t nCells[SJ_MAXK];
    uint64_t totalCells;
    uint64_t nNonEmpty;
    int32_t dense_lookup;     /* 1: O(1) cell table, 0: binary search */
    int32_t reserved;
} sj_grid_info;

typedef struct sj_stats {
    uint64_t total_pairs;        /* |S|, reference totalNeighbors */
    uint64_t num_batches;
    uint64_t candidates;         /* distance evaluations of the count pass */
    uint64_t fp32_band_pairs;    /* candidates re-decided in FP64 */
    uint64_t fp32_misclassified; /* pairs whose plain FP32 decision was wrong */
    uint64_t overflow_retries;
    double ms_grid;              /* device time, grid construction */
    double ms_workload;          /* device time, workload estimate + order */
    double ms_count;             /* device time, count pass */
    double ms_scan;              /* device time, offsets */
    double ms_fill;              /* device time, fill pass (all batches) */
    double ms_sort_rows;         /* device time, canonical row sort */
    double ms_total;             /* wall time of the call */
    uint64_t fill_from_record;   /* 1: the fill replayed the count pass's hit record */
} sj_stats;

/* Row of the reference neighbour table (structs.h:41-47): neighbours of
 * pointID are ids[indexmin..indexmax] of batch `batch`. */
typedef struct sj_neighbor_lookup {
    int32_t pointID;
    int32_t indexmin;
    int32_t indexmax;
    int32_t batch;
} sj_neighbor_lookup;

/* Last error message of the calling thread. */
SJ_API const char *sj_last_error(void);

/* Replaces warmUpGPU (GPU.h:54, GPU.cu:2572): selects the device, creates
 * the stream and the caching device pool. */
SJ_API sj_status sj_context_create(int32_t device, void *stream, sj_context *out);
SJ_API void sj_context_destroy(sj_context ctx);
SJ_API sj_status sj_context_set_stream(sj_context ctx, void *stream);

/* Replaces generateNDGridDimensions + populateNDGridIndexAndLookupArray
 * (main.cu:30-31,240-503).  d_points: n x d row-major float64 in device
 * memory; the first k coordinates are indexed.  Built on the device. */
SJ_API sj_status sj_grid_build(sj_context ctx, const double *d_points, uint64_t n,
                               int32_t d, int32_t k, double eps, sj_grid *out);
SJ_API sj_status sj_grid_get_info(sj_grid g, sj_grid_info *info);
/* Host copies of the grid arrays (main.cu:240-413 outputs): cellLinear =
 * gridCellLookupArr[].gridLinearID, indexMin/indexMax = index[], indexLookup =
 * indexLookupArr, cellCoord = per-cell coordinates [nNonEmpty x k].  Any
 * pointer may be NULL. */
SJ_API sj_status sj_grid_copy_arrays(sj_grid g, uint64_t *cellLinear, int32_t *indexMin,
                                     int32_t *indexMax, uint32_t *indexLookup,
                                     uint32_t *cellCoord);
/* gridCellNDMask / gridCellNDMaskOffsets / nNDMaskElems (main.cu:381-401).
 * mask may be NULL to query nElems. */
SJ_API sj_status sj_grid_nd_mask(sj_grid g, uint32_t *mask, uint32_t *offsets, uint64_t *nElems);
SJ_API void sj_grid_free(sj_grid g);

/* Replaces distanceTableNDGridBatches (GPU.h:37-40, GPU.cu:1297-2569).
 * Produces the neighbour table; see the sj_result_* views. */
SJ_API sj_status sj_join(sj_grid g, const sj_join_params *params, sj_result *out);

/* Grid build + join in one call (the reference's timed region,
 * main.cu:182-195, with the grid built on the device). */
SJ_API sj_status sj_self_join(sj_context ctx, const double *d_points, uint64_t n, int32_t d,
                              int32_t k, double eps, const sj_join_params *params,
                              sj_result *out);

SJ_API sj_status sj_result_get_stats(sj_result r, sj_stats *stats);
SJ_API uint64_t sj_result_total_pairs(sj_result r);
SJ_API uint64_t sj_result_num_points(sj_result r);
/* Device neighbour table as produced (no copy): rows in cell order; row r
 * lists the neighbours of point rowIds[r] in ids[offsets[r]..offsets[r+1]).
 * This is the device form of the reference's neighbour table
 * (neighborTableLookup + data arrays, structs.h:41-57).  Available when the
 * result is a single device-resident batch. */
SJ_API sj_status sj_result_device_table(sj_result r, const uint64_t **d_offsets,
                                        const int32_t **d_ids, const uint32_t **d_rowIds);
/* Device CSR in original point-ID order: offsets[n+1] (uint64), ids[|S|],
 * built from the table on the first call (one extra pass over the pairs). */
SJ_API sj_status sj_result_device_csr(sj_result r, const uint64_t **d_offsets,
                                      const int32_t **d_ids);
/* Host copy of the CSR (any batching). */
SJ_API sj_status sj_result_copy_csr(sj_result r, uint64_t *offsets, int32_t *ids);
/* Key/value view, key-sorted: replaces the thrust::sort_by_key output
 * (GPU.cu:2336-2349).  keys, vals: |S| entries each. */
SJ_API sj_status sj_result_copy_pairs(sj_result r, int32_t *keys, int32_t *vals);
/* Reference neighbour table (constructNeighborTableKeyValueWithPtrs,
 * GPU.h:52, GPU.cu:2595-2637): table[n]; data of batch b via
 * sj_result_batch_data. */
SJ_API sj_status sj_result_neighbor_table(sj_result r, sj_neighbor_lookup *table);
SJ_API sj_status sj_result_batch_data(sj_result r, uint64_t batch, const int32_t **ids,
                                      uint64_t *count);
SJ_API void sj_result_free(sj_result r);

/* Workload per non-empty cell (sortByWorkLoad / sortByWorkLoadLidUnicomp,
 * kernel.cu:33-250) and the descending order (GPU.cu:1695-1698). */
SJ_API sj_status sj_workload(sj_grid g, int32_t lid, int64_t *nbPoints, uint32_t *order);

/* Batch estimator (callGPUBatchEst / callGPUBatchEstWorkQueue, GPU.h:42-50,
 * GPU.cu:854-1284): full-stencil hit count of every stride-th point
 * (stride 0: 100, the reference's sample rate), estimated = count * stride,
 * estimatedWithAlpha = estimated * 1.05 (the value callGPUBatchEst returns),
 * and the buffer/batch policy of GPU.cu:1029-1055 (numBatches, bufferSize =
 * 1.5x the per-batch buffer).  Any output pointer may be NULL. */
SJ_API sj_status sj_batch_estimate(sj_grid g, uint64_t stride, uint64_t *sampleCount,
                                   uint64_t *estimated, uint64_t *estimatedWithAlpha,
                                   uint64_t *numBatches, uint64_t *bufferSize);
/* Work-queue estimator (callGPUBatchEstWorkQueue, GPU.h:47-50, GPU.cu:1072-1284,
 * kernel.cu:785-942): the sample is the first n/stride points of the
 * workload-ordered point list (cells by descending workload, ties by
 * ascending cell index; points of a cell in ascending ID), the order modes 9
 * (lid = 0, full-stencil workload) and 10 (lid = 1, LID workload) run in
 * (GPU.cu:1685-1692).  Same outputs and policy as sj_batch_estimate. */
SJ_API sj_status sj_batch_estimate_workqueue(sj_grid g, int32_t lid, uint64_t stride,
                                             uint64_t *sampleCount, uint64_t *estimated,
                                             uint64_t *estimatedWithAlpha, uint64_t *numBatches,
                                             uint64_t *bufferSize);

/* Replaces importNDDataset (import_dataset.cpp:34-73).  *points is
 * allocated with malloc; release with sj_free_host. */
SJ_API sj_status sj_import_csv(const char *path, int32_t d, double **points, uint64_t *n);
/* Replaces sortInNDBins (import_dataset.cpp:16-30,76-80).  stable=0 uses
 * std::sort with the reference comparator (libstdc++ order); stable=1 a
 * stable sort. */
SJ_API sj_status sj_sort_in_nd_bins(double *points, uint64_t n, int32_t d, int32_t stable);
SJ_API void sj_free_host(void *p);

/* ---- Many joins in one graph ------------------------------------------
 * No counterpart in the reference, which runs one join per process
 * (main.cu:36-226): a batch of independent self-joins of different size,
 * dimension and eps, joined by one persistent kernel per pass over all
 * instances' tiles and recorded once into a hipGraph. */
typedef struct sj_instance {
    const double *points;  /* device, n x d float64, row-major */
    uint64_t n;
    int32_t d, k;          /* k indexed dims, 1..8 */
    double eps;
} sj_instance;

typedef struct sj_many_tuning {
    int32_t form;          /* sj_form */
    int32_t schedule;      /* sj_schedule */
    int32_t blocks_per_cu; /* persistent blocks of 256 threads per CU (0: 8) */
    int32_t queue_chunk;   /* work-queue tiles per grab (0: 1) */
} sj_many_tuning;

typedef struct sj_many_s *sj_many;

/* Plan: builds every grid and tile list, sizes the shared neighbour-ID
 * array with one count run, records the graph and runs it once.  params:
 * form, precision, schedule, queue_chunk; capacity must be 0. */
SJ_API sj_status sj_many_create(sj_context ctx, const sj_instance *instances, int32_t count,
                                const sj_join_params *params, sj_many *out);
/* One graph launch: count, offsets and fill of every instance. */
SJ_API sj_status sj_many_run(sj_many m);
/* Times graph replays over stencil form (when tune_form != 0), schedule,
 * blocks per CU and queue chunk; keeps and returns the fastest. */
SJ_API sj_status sj_many_autotune(sj_many m, int32_t tune_form, int32_t reps, sj_many_tuning *chosen);
SJ_API sj_status sj_many_set_tuning(sj_many m, const sj_many_tuning *tuning);
SJ_API sj_status sj_many_get_tuning(sj_many m, sj_many_tuning *tuning);
/* Device neighbour table of instance i, as sj_result_device_table:
 * offsets[n_i + 1] relative to ids, rows in cell order, rowIds the
 * original point ID of each row. */
SJ_API sj_status sj_many_instance(sj_many m, int32_t i, const uint64_t **d_offsets, const int32_t **d_ids,
                                  const uint32_t **d_rowIds, uint64_t *total);
/* Counters of instance i, or of the whole batch for i = -1. */
SJ_API sj_status sj_many_stats(sj_many m, int32_t i, sj_stats *stats);
/* Wall time of the plan, of its grid builds, and device time of the last
 * graph replay, in ms. */
SJ_API sj_status sj_many_times(sj_many m, double *ms_plan, double *ms_grid, double *ms_run);
SJ_API int32_t sj_many_count(sj_many m);
SJ_API void sj_many_free(sj_many m);

/* ---- Allocation-free self-join -----------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): sj_self_join with every device buffer of the grid
 * build and the join taken from one static device array of
 * SJ_STATIC_SCRATCH_MIB MiB per device (held from the library's load on
 * that device until the process ends), and the outputs written to the
 * caller's arrays: offsets[n+1] (uint64) and rowIds[n] (rows in cell order,
 * as sj_result_device_table), ids[ids_capacity] (int32).  *total receives
 * |S|; above ids_capacity the call returns SJ_ERR_OVERFLOW with *total set,
 * and nothing is written to ids.  Calls no hipMalloc, hipFree or other
 * allocating routine; the context must exist (sj_context_create looks up
 * the array's device address).  Inputs whose scratch bound
 * (sj_static_need) exceeds the array, and capacity > 0 in params, are
 * rejected with SJ_ERR_OVERFLOW / SJ_ERR_UNSUPPORTED before any device
 * work.  The fill evaluates distances again (no hit record).  Every call
 * on a device shares that device's array: calls that overlap on one device,
 * from any context or thread, race. */
#define SJ_STATIC_SCRATCH_MIB 256
SJ_API sj_status sj_self_join_static(sj_context ctx, const double *d_points, uint64_t n, int32_t d,
                                     int32_t k, double eps, const sj_join_params *params,
                                     uint64_t *d_offsets, int32_t *d_ids, uint64_t ids_capacity,
                                     uint32_t *d_rowIds, uint64_t *total);
/* Bytes of static scratch sj_self_join_static may need for n points (an
 * upper bound), and the size of the array. */
SJ_API uint64_t sj_static_need(uint64_t n, int32_t d, int32_t k, const sj_join_params *params);
SJ_API uint64_t sj_static_scratch_bytes(void);

/* Eps threshold T = max{x : sqrt_rn(x) <= eps}; sqrt(d2) <= eps iff d2 <= T. */
SJ_API double sj_eps_threshold(double eps);

/* Band of the FP32 precision form on the VALU tile walk (no reference
 * counterpart; exposed so the bound can be checked on the host).  Each
 * coordinate is stored as y = fl32(x - c) with |x - c| <= R for the tile's
 * FP64 centre c; s32 is the FMA sum of the D squared FP32 differences in
 * dimension order.  Then d2_64 <= T implies s32 <= hi and d2_64 > T implies
 * s32 >= lo, so s32 < lo is a hit, s32 > hi a miss, and [lo, hi] is
 * re-decided in FP64.  t32 = fl32(T) is the plain FP32 decision whose
 * errors sj_stats.fp32_misclassified counts.  Returns 1 when the band is
 * usable (narrow enough), else 0 (the join then runs the FP64 predicate). */
SJ_API int32_t sj_fp32_band(double T, double R, int32_t D, float *lo, float *hi, float *t32);

/* Decision constants of the FP32 form on the matrix cores (no reference
 * counterpart).  With FP32 vectors centred on the tile's first point, s =
 * |q|^2 + |c|^2, d2 = s - 2 q.c (any summation order) and M = alpha s +
 * beta: d2 + M <= tlo implies d2_64 <= T, d2 - M > thi implies d2_64 > T.
 * R >= 0: staging from the FP32 copy relative to the grid origin, whose
 * vectors are within R of it (Euclidean); R < 0: centring in FP64.
 * Returns the padded operand width (4..64), 0 when d is above 64. */
SJ_API int32_t sj_mfma_band(double T, double R, int32_t d, float *alpha, float *beta, float *tlo, float *thi);

#ifdef __cplusplus
}
#endif
#endif /* SJAMD_H */
