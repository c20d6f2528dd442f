/*
 * CPU oracle for the grid-indexed epsilon self-join.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (benoitgallet/self-join-hpbdc2019) used as the checker for the
 * HIP product and as the timed host-core baseline.  Only tests/, the smoke
 * entry point and the benchmark's cpu_baseline leg may load it; the product
 * library never links or calls it.
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef SJ_ORACLE_H
#define SJ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SJO_MAXK 12

enum { SJO_FORM_FULL = 0, SJO_FORM_UNICOMP = 1, SJO_FORM_BUNICOMP = 2, SJO_FORM_LID = 3 };

typedef struct {
    int d, k;
    double eps;
    uint64_t n;
    double minArr[SJO_MAXK];
    double maxArr[SJO_MAXK];
    uint32_t nCells[SJO_MAXK];
    uint64_t totalCells;
    uint64_t nNonEmpty;
    uint64_t *cellLinear;    /* [nNonEmpty] ascending linear IDs            */
    int32_t *indexMin;       /* [nNonEmpty] first slot in indexLookup       */
    int32_t *indexMax;       /* [nNonEmpty] last slot (inclusive)           */
    uint32_t *indexLookup;   /* [n] point IDs grouped by cell, ascending    */
    uint32_t *cellCoord;     /* [nNonEmpty*k] per-cell integer coordinates  */
    uint32_t *ndMask;        /* occupied coordinates per dim, concatenated  */
    uint32_t ndMaskOffsets[2 * SJO_MAXK];
    uint64_t nNDMaskElems;
} sjo_grid;

/* Predicate P */
double sjo_eps_threshold(double eps);
double sjo_dist2(const double *a, const double *b, int d);
void sjo_dist2_batch(const double *a, const double *b, uint64_t m, int d, double *out);

/* G2 */
uint64_t sjo_linear_id(const uint32_t *idx, const uint32_t *nCells, int k);

/* G1 */
int sjo_grid_dims(const double *pts, uint64_t n, int d, int k, double eps,
                  double *minArr, double *maxArr, uint32_t *nCells,
                  uint64_t *totalCells);
/* G3 */
int sjo_grid_build(const double *pts, uint64_t n, int d, int k, double eps,
                   sjo_grid *g);
void sjo_grid_free(sjo_grid *g);

/* Stencil claim rules (T3-T6). */
int sjo_claims(int form, int k, const uint32_t *cellCoord, const int *off,
               const uint32_t *nCells);

/* Grid join -> CSR in original point-ID order.
 * offsets: caller array of n+1; ids allocated here (free with sjo_free).
 * sort_rows != 0 sorts each row ascending (canonical form). */
int sjo_join(const sjo_grid *g, const double *pts, int form, int nthreads,
             int sort_rows, uint64_t *offsets, int32_t **ids);

/* Brute force O(n^2) with predicate P; same CSR product, rows sorted. */
int sjo_brute(const double *pts, uint64_t n, int d, double eps, int nthreads,
              uint64_t *offsets, int32_t **ids);

/* W1: per non-empty cell (in cellLinear order) the number of candidate
 * points of its stencil; lid != 0 counts only cells with linear ID >= own. */
int sjo_workload(const sjo_grid *g, int lid, int64_t *nbPoints);

/* W1 + W2: cells ordered by descending workload, ties by ascending cell
 * index (a stable order). */
int sjo_workload_order(const sjo_grid *g, int lid, uint32_t *order);

/* B1: sampled estimate (every 100th point, full stencil). */
uint64_t sjo_sample_count(const sjo_grid *g, const double *pts, uint64_t stride);
/* B1, work-queue form: the first n/stride points of the workload-ordered
 * point list (sjo_workload_order with lid, points of a cell in ID order). */
uint64_t sjo_sample_count_workqueue(const sjo_grid *g, const double *pts, uint64_t stride, int lid);
void sjo_batch_policy(uint64_t sampleCount, uint64_t offsetRate,
                      uint64_t *estimated, uint64_t *estimatedWithAlpha,
                      uint64_t *numBatches, uint64_t *bufferSize);

/* B3: run-length neighbor table from key-sorted (key,value) pairs. */
typedef struct { int32_t pointID, indexmin, indexmax; int32_t batch; } sjo_ntl;
int sjo_neighbor_table(const int32_t *keys, const int32_t *vals, uint64_t cnt,
                       int32_t batch, sjo_ntl *table);

/* L1: dataset import and the optional lexicographic pre-sort. */
int sjo_import_csv(const char *path, int d, double **pts, uint64_t *n);
int sjo_sort_in_nd_bins_stable(double *pts, uint64_t n, int d);

void sjo_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
