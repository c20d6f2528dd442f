/*
 * CPU oracle for the grid-indexed epsilon self-join (TEST INFRASTRUCTURE).
 *
 * Restates the reference algorithm of benoitgallet/self-join-hpbdc2019 in
 * plain C + OpenMP.  It is the checker for the HIP product and the timed
 * host-core baseline; it is never linked into the product library.
 *
 * Deliberate, documented deviations from the reference (all are defects of
 * the reference that a checker must not inherit):
 *   - result pairs always carry original point IDs (reference mixes the
 *     workload-sorted and original namespaces, kernel.cu:978-988,2504);
 *   - Unicomp emits the mirrored pair (q,p), not (p,q) twice
 *     (kernel.cu:356-360 via stamploops.h:9);
 *   - neighbours outside [0,nCells) are skipped (reference relies on an
 *     unsigned wrap, kernel.cu:1458);
 *   - the workload kernel uses the point stride d, not k (kernel.cu:68);
 *   - ties of the workload sort and of the bin pre-sort are broken by index
 *     (reference uses unstable std::sort, GPU.cu:1697, import_dataset.cpp:78).
 *
 * Build: gcc -O3 -fopenmp -ffp-contract=off (explicit fma() only).
 */
#include "sj_oracle.h"

#include <errno.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SJO_OK 0
#define SJO_EINVAL 1
#define SJO_ENOMEM 2
#define SJO_EOVERFLOW 4
#define SJO_EUNSUP 5
#define SJO_EIO 6

void sjo_free(void *p) { free(p); }

/* ------------------------------------------------------------------ */
/* Predicate P: kernel.cu:341-349 (evalPoint), :421-429, :515-523.     */
/* acc += (x_q,l - x_p,l)^2 with nvcc's default contraction into fma,  */
/* then sqrt(acc) <= eps, which is acc <= T for the largest T with     */
/* sqrt_rn(T) <= eps (sqrt_rn is monotone).                            */
/* ------------------------------------------------------------------ */
double sjo_eps_threshold(double eps)
{
    double t = eps * eps;
    if (!(eps >= 0.0) || isinf(eps)) return NAN;
    if (sqrt(t) <= eps) {
        for (;;) {
            double nx = nextafter(t, INFINITY);
            if (sqrt(nx) <= eps) t = nx; else break;
        }
    } else {
        while (!(sqrt(t) <= eps)) t = nextafter(t, 0.0);
    }
    return t;
}

double sjo_dist2(const double *a, const double *b, int d)
{
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        double df = b[l] - a[l];
        acc = fma(df, df, acc);
    }
    return acc;
}

/* Predicate P over m pairs: out[i] = sjo_dist2(a + i*d, b + i*d, d). */
void sjo_dist2_batch(const double *a, const double *b, uint64_t m, int d, double *out)
{
    for (uint64_t i = 0; i < m; i++) out[i] = sjo_dist2(a + i * d, b + i * d, d);
}

/* ------------------------------------------------------------------ */
/* G2: main.cu:423-439 (dim 0 fastest).                                */
/* ------------------------------------------------------------------ */
uint64_t sjo_linear_id(const uint32_t *idx, const uint32_t *nCells, int k)
{
    uint64_t id = 0, mul = 1;
    for (int i = 0; i < k; i++) {
        id += (uint64_t)idx[i] * mul;
        mul *= nCells[i];
    }
    return id;
}

/* ------------------------------------------------------------------ */
/* G1: main.cu:446-503.                                                */
/* ------------------------------------------------------------------ */
int sjo_grid_dims(const double *pts, uint64_t n, int d, int k, double eps,
                  double *minArr, double *maxArr, uint32_t *nCells,
                  uint64_t *totalCells)
{
    if (n == 0 || k < 1 || k > SJO_MAXK || k > d) return SJO_EINVAL;
    if (!(eps > 0.0) || isinf(eps)) return SJO_EINVAL;
    for (int j = 0; j < k; j++) { minArr[j] = pts[j]; maxArr[j] = pts[j]; }
    for (uint64_t i = 0; i < n; i++)
        for (int j = 0; j < d; j++)
            if (isnan(pts[i * d + j]) || isinf(pts[i * d + j])) return SJO_EINVAL;
    for (uint64_t i = 1; i < n; i++)
        for (int j = 0; j < k; j++) {
            double x = pts[i * d + j];
            if (x < minArr[j]) minArr[j] = x;
            if (x > maxArr[j]) maxArr[j] = x;
        }
    for (int j = 0; j < k; j++) { minArr[j] -= eps; maxArr[j] += eps; }
    uint64_t tot = 1;
    for (int j = 0; j < k; j++) {
        double c = ceil((maxArr[j] - minArr[j]) / eps);
        if (!(c < 4294967296.0)) return SJO_EOVERFLOW;
        nCells[j] = (uint32_t)c;
        if (__builtin_mul_overflow(tot, (uint64_t)nCells[j], &tot)) return SJO_EOVERFLOW;
    }
    *totalCells = tot;
    return SJO_OK;
}

/* per-point cell coordinate, main.cu:263,294 */
static inline int cell_coord(double x, double mn, double eps, uint32_t *c)
{
    double v = (x - mn) / eps;
    if (!(v >= 0.0) || !(v < 4294967296.0)) return SJO_EOVERFLOW;
    *c = (uint32_t)v;
    return SJO_OK;
}

/* stable LSD radix sort of (key,id) pairs by key, 16-bit digits */
static int radix_sort_pairs(uint64_t *key, uint32_t *id, uint64_t n, uint64_t maxkey)
{
    uint64_t *k2 = malloc(n * sizeof(uint64_t));
    uint32_t *i2 = malloc(n * sizeof(uint32_t));
    uint64_t *cnt = malloc(65536 * sizeof(uint64_t));
    if (!k2 || !i2 || !cnt) { free(k2); free(i2); free(cnt); return SJO_ENOMEM; }
    int bits = 0;
    while (bits < 64 && (maxkey >> bits)) bits++;
    for (int sh = 0; sh < bits; sh += 16) {
        memset(cnt, 0, 65536 * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) cnt[(key[i] >> sh) & 0xFFFF]++;
        uint64_t s = 0;
        for (int b = 0; b < 65536; b++) { uint64_t c = cnt[b]; cnt[b] = s; s += c; }
        for (uint64_t i = 0; i < n; i++) {
            uint64_t p = cnt[(key[i] >> sh) & 0xFFFF]++;
            k2[p] = key[i]; i2[p] = id[i];
        }
        memcpy(key, k2, n * sizeof(uint64_t));
        memcpy(id, i2, n * sizeof(uint32_t));
    }
    free(k2); free(i2); free(cnt);
    return SJO_OK;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------ */
/* G3: main.cu:240-413.  Unique linear IDs ascending (std::set,        */
/* :257-273), points bucketed in ascending point ID (:291-313),        */
/* index[c] = {indexmin,indexmax} inclusive (:330-344,368-373),        */
/* gridCellLookupArr[c] = {c, linearID} (:371-372), per-dim mask of    */
/* occupied coordinates with [min,max] offsets (:286,297,381-401).     */
/* ------------------------------------------------------------------ */
int sjo_grid_build(const double *pts, uint64_t n, int d, int k, double eps,
                   sjo_grid *g)
{
    memset(g, 0, sizeof(*g));
    if (n > 0x7fffffffULL) return SJO_EOVERFLOW;
    int rc = sjo_grid_dims(pts, n, d, k, eps, g->minArr, g->maxArr, g->nCells, &g->totalCells);
    if (rc) return rc;
    g->d = d; g->k = k; g->eps = eps; g->n = n;

    uint64_t *key = malloc(n * sizeof(uint64_t));
    uint32_t *id = malloc(n * sizeof(uint32_t));
    if (!key || !id) { free(key); free(id); return SJO_ENOMEM; }
    int bad = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t c[SJO_MAXK];
        for (int j = 0; j < k; j++)
            if (cell_coord(pts[i * d + j], g->minArr[j], eps, &c[j])) bad = 1;
        key[i] = bad ? 0 : sjo_linear_id(c, g->nCells, k);
        id[i] = (uint32_t)i;
    }
    if (bad) { free(key); free(id); return SJO_EOVERFLOW; }
    rc = radix_sort_pairs(key, id, n, g->totalCells ? g->totalCells - 1 : 0);
    if (rc) { free(key); free(id); return rc; }

    uint64_t nne = 0;
    for (uint64_t i = 0; i < n; i++)
        if (i == 0 || key[i] != key[i - 1]) nne++;
    g->nNonEmpty = nne;
    g->cellLinear = malloc(nne * sizeof(uint64_t));
    g->indexMin = malloc(nne * sizeof(int32_t));
    g->indexMax = malloc(nne * sizeof(int32_t));
    g->cellCoord = malloc(nne * k * sizeof(uint32_t));
    g->indexLookup = id;
    if (!g->cellLinear || !g->indexMin || !g->indexMax || !g->cellCoord) {
        free(key); sjo_grid_free(g); return SJO_ENOMEM;
    }
    uint64_t c = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (i == 0 || key[i] != key[i - 1]) {
            if (i) g->indexMax[c - 1] = (int32_t)(i - 1);
            g->cellLinear[c] = key[i];
            g->indexMin[c] = (int32_t)i;
            uint64_t p = id[i];
            for (int j = 0; j < k; j++)
                cell_coord(pts[p * d + j], g->minArr[j], eps, &g->cellCoord[c * k + j]);
            c++;
        }
    }
    g->indexMax[c - 1] = (int32_t)(n - 1);
    free(key);

    /* per-dimension occupied coordinate mask */
    uint64_t total = 0;
    uint32_t *tmp = malloc(nne * sizeof(uint32_t));
    uint32_t *mask = malloc((uint64_t)k * nne * sizeof(uint32_t));
    if (!tmp || !mask) { free(tmp); free(mask); sjo_grid_free(g); return SJO_ENOMEM; }
    for (int j = 0; j < k; j++) {
        for (uint64_t q = 0; q < nne; q++) tmp[q] = g->cellCoord[q * k + j];
        qsort(tmp, nne, sizeof(uint32_t), cmp_u32);
        g->ndMaskOffsets[2 * j] = (uint32_t)total;
        for (uint64_t q = 0; q < nne; q++)
            if (q == 0 || tmp[q] != tmp[q - 1]) mask[total++] = tmp[q];
        g->ndMaskOffsets[2 * j + 1] = (uint32_t)(total - 1);
    }
    free(tmp);
    g->ndMask = mask;
    g->nNDMaskElems = total;
    return SJO_OK;
}

void sjo_grid_free(sjo_grid *g)
{
    free(g->cellLinear); free(g->indexMin); free(g->indexMax);
    free(g->indexLookup); free(g->cellCoord); free(g->ndMask);
    memset(g, 0, sizeof(*g));
}

/* ------------------------------------------------------------------ */
/* Stencil claim rules.  off[] in {-1,0,1}^k, not all zero.            */
/*  FULL     kernel.cu:1076-1093 (every neighbour cell)                */
/*  UNICOMP  stamploops.h:1-202: for the highest dim m with off[m]!=0, */
/*           the cell coordinate c[m] must be odd                      */
/*  BUNICOMP kernel.cu:1346-1392 for k=2,3; other k use the balanced   */
/*           rule  [off >lex 0] XOR ([sum off even] AND parity(sum c)) */
/*  LID      kernel.cu:1513-1536: linearID(neighbour) > linearID(own)  */
/* ------------------------------------------------------------------ */
static int lex_positive(int k, const int *off)
{
    for (int i = k - 1; i >= 0; i--)
        if (off[i]) return off[i] > 0;
    return 0;
}

int sjo_claims(int form, int k, const uint32_t *c, const int *o, const uint32_t *nCells)
{
    switch (form) {
    case SJO_FORM_FULL:
        return 1;
    case SJO_FORM_UNICOMP: {
        for (int i = k - 1; i >= 0; i--)
            if (o[i]) return (int)(c[i] & 1u);
        return 0;
    }
    case SJO_FORM_BUNICOMP: {
        uint64_t s = 0;
        for (int i = 0; i < k; i++) s += c[i];
        int odd = (int)(s & 1u);
        if (k == 2) {
            if (odd) return o[1] >= 0 && o[0] != 0;
            return o[1] > 0 || (o[1] < 0 && o[0] == 0);
        }
        if (k == 3) {
            if (odd)
                return (o[0] != 0 && o[1] >= 0 && o[2] >= 0) ||
                       (o[0] == 0 && o[1] >= 0 && o[2] > 0) ||
                       (o[1] > 0 && o[2] < 0);
            return (o[1] > 0 && o[2] >= 0) ||
                   (o[0] == 0 && o[1] < 0 && o[2] >= 0) ||
                   (o[0] != 0 && o[1] > 0 && o[2] < 0) ||
                   (o[1] == 0 && o[2] > 0);
        }
        int se = 0;
        for (int i = 0; i < k; i++) se += o[i];
        int even = ((se % 2) + 2) % 2 == 0;
        return lex_positive(k, o) ^ (even && odd);
    }
    case SJO_FORM_LID: {
        uint32_t nb[SJO_MAXK];
        for (int i = 0; i < k; i++) nb[i] = c[i] + (uint32_t)o[i];
        return sjo_linear_id(nb, nCells, k) > sjo_linear_id(c, nCells, k);
    }
    }
    return 0;
}

/* cell lookup: dense table for small grids, binary search otherwise
 * (the reference binary-searches gridCellLookupArr, kernel.cu:387-396) */
typedef struct {
    const sjo_grid *g;
    int32_t *dense;
} lookup_t;

static int lookup_init(lookup_t *L, const sjo_grid *g)
{
    L->g = g;
    L->dense = NULL;
    if (g->totalCells <= (1ULL << 26) || g->totalCells <= 16 * g->n) {
        L->dense = malloc(g->totalCells * sizeof(int32_t));
        if (!L->dense) return SJO_ENOMEM;
        memset(L->dense, 0xff, g->totalCells * sizeof(int32_t));
        for (uint64_t c = 0; c < g->nNonEmpty; c++) L->dense[g->cellLinear[c]] = (int32_t)c;
    }
    return SJO_OK;
}

static inline int64_t lookup_cell(const lookup_t *L, uint64_t lin)
{
    if (L->dense) return L->dense[lin];
    uint64_t lo = 0, hi = L->g->nNonEmpty;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (L->g->cellLinear[mid] < lin) lo = mid + 1; else hi = mid;
    }
    if (lo < L->g->nNonEmpty && L->g->cellLinear[lo] == lin) return (int64_t)lo;
    return -1;
}

typedef struct {
    int k;
    uint64_t nOff;
    int *off; /* nOff * k */
} stencil_t;

static int stencil_init(stencil_t *S, int k)
{
    uint64_t total = 1;
    for (int i = 0; i < k; i++) total *= 3;
    S->k = k;
    S->nOff = total - 1;
    S->off = malloc((total ? total : 1) * k * sizeof(int));
    if (!S->off) return SJO_ENOMEM;
    uint64_t w = 0;
    for (uint64_t t = 0; t < total; t++) {
        uint64_t r = t;
        int nz = 0;
        int o[SJO_MAXK];
        for (int i = 0; i < k; i++) { o[i] = (int)(r % 3) - 1; r /= 3; nz |= o[i]; }
        if (!nz) continue;
        memcpy(S->off + w * k, o, k * sizeof(int));
        w++;
    }
    return SJO_OK;
}

/* neighbour cell of c by offset o, -1 when outside the grid or empty */
static inline int64_t neighbor(const lookup_t *L, const uint32_t *c, const int *o)
{
    const sjo_grid *g = L->g;
    uint32_t nb[SJO_MAXK];
    for (int i = 0; i < g->k; i++) {
        int64_t v = (int64_t)c[i] + o[i];
        if (v < 0 || v >= (int64_t)g->nCells[i]) return -1;
        nb[i] = (uint32_t)v;
    }
    return lookup_cell(L, sjo_linear_id(nb, g->nCells, g->k));
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------ */
/* Join: kernel.cu:949-1095 (FULL), 1102-1242 (UNICOMP),               */
/* 1249-1396 (B-UNICOMP), 1403-1538 (LID).  Origin cell: every ordered */
/* pair (p,q) incl. p==q (evaluateCellUnicompOrigin, kernel.cu:445).   */
/* Claimed neighbour cells emit (p,q) and (q,p)                        */
/* (evalPointUnicompAdjacent, kernel.cu:525-529).  The product is the  */
/* per-point neighbour table (GPU.cu:2595-2637) as CSR in original IDs */
/* built by count + scan + fill.                                       */
/* ------------------------------------------------------------------ */
int sjo_join(const sjo_grid *g, const double *pts, int form, int nthreads,
             int sort_rows, uint64_t *offsets, int32_t **ids_out)
{
    *ids_out = NULL;
    if (form < 0 || form > 3) return SJO_EINVAL;
    const uint64_t n = g->n;
    const int d = g->d, k = g->k;
    const double T = sjo_eps_threshold(g->eps);
    if (nthreads > 0) omp_set_num_threads(nthreads);

    lookup_t L;
    stencil_t S;
    if (lookup_init(&L, g)) return SJO_ENOMEM;
    if (stencil_init(&S, k)) { free(L.dense); return SJO_ENOMEM; }
    uint64_t *cnt = calloc(n + 1, sizeof(uint64_t));
    uint64_t *cur = calloc(n, sizeof(uint64_t));
    if (!cnt || !cur) { free(cnt); free(cur); free(L.dense); free(S.off); return SJO_ENOMEM; }

    const int64_t nne = (int64_t)g->nNonEmpty;
    const int mirror = form != SJO_FORM_FULL;
    for (int pass = 0; pass < 2; pass++) {
        int32_t *ids = pass ? *ids_out : NULL;
#pragma omp parallel for schedule(dynamic, 8)
        for (int64_t a = 0; a < nne; a++) {
            const uint32_t *ca = g->cellCoord + a * k;
            /* origin cell */
            for (int32_t i = g->indexMin[a]; i <= g->indexMax[a]; i++) {
                uint32_t p = g->indexLookup[i];
                for (int32_t j = g->indexMin[a]; j <= g->indexMax[a]; j++) {
                    uint32_t q = g->indexLookup[j];
                    if (sjo_dist2(pts + (uint64_t)p * d, pts + (uint64_t)q * d, d) <= T) {
                        if (!pass) __atomic_fetch_add(&cnt[p], 1, __ATOMIC_RELAXED);
                        else ids[offsets[p] + __atomic_fetch_add(&cur[p], 1, __ATOMIC_RELAXED)] = (int32_t)q;
                    }
                }
            }
            for (uint64_t s = 0; s < S.nOff; s++) {
                const int *o = S.off + s * k;
                if (!sjo_claims(form, k, ca, o, g->nCells)) continue;
                int64_t b = neighbor(&L, ca, o);
                if (b < 0) continue;
                for (int32_t i = g->indexMin[a]; i <= g->indexMax[a]; i++) {
                    uint32_t p = g->indexLookup[i];
                    for (int32_t j = g->indexMin[b]; j <= g->indexMax[b]; j++) {
                        uint32_t q = g->indexLookup[j];
                        if (sjo_dist2(pts + (uint64_t)p * d, pts + (uint64_t)q * d, d) <= T) {
                            if (!pass) {
                                __atomic_fetch_add(&cnt[p], 1, __ATOMIC_RELAXED);
                                if (mirror) __atomic_fetch_add(&cnt[q], 1, __ATOMIC_RELAXED);
                            } else {
                                ids[offsets[p] + __atomic_fetch_add(&cur[p], 1, __ATOMIC_RELAXED)] = (int32_t)q;
                                if (mirror)
                                    ids[offsets[q] + __atomic_fetch_add(&cur[q], 1, __ATOMIC_RELAXED)] = (int32_t)p;
                            }
                        }
                    }
                }
            }
        }
        if (!pass) {
            uint64_t s = 0;
            for (uint64_t i = 0; i < n; i++) { offsets[i] = s; s += cnt[i]; }
            offsets[n] = s;
            *ids_out = malloc((s ? s : 1) * sizeof(int32_t));
            if (!*ids_out) { free(cnt); free(cur); free(L.dense); free(S.off); return SJO_ENOMEM; }
        }
    }
    if (sort_rows) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t p = 0; p < (int64_t)n; p++)
            qsort(*ids_out + offsets[p], offsets[p + 1] - offsets[p], sizeof(int32_t), cmp_i32);
    }
    free(cnt); free(cur); free(L.dense); free(S.off);
    return SJO_OK;
}

/* Brute force over all pairs with the same predicate: an independent check
   of the grid joins on small inputs.  Rows sorted. */
int sjo_brute(const double *pts, uint64_t n, int d, double eps, int nthreads,
              uint64_t *offsets, int32_t **ids_out)
{
    *ids_out = NULL;
    if (!(eps >= 0.0)) return SJO_EINVAL;
    const double T = sjo_eps_threshold(eps);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    uint64_t *cnt = calloc(n + 1, sizeof(uint64_t));
    if (!cnt) return SJO_ENOMEM;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < (int64_t)n; p++) {
        uint64_t c = 0;
        for (uint64_t q = 0; q < n; q++)
            c += sjo_dist2(pts + (uint64_t)p * d, pts + q * d, d) <= T;
        cnt[p] = c;
    }
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; i++) { offsets[i] = s; s += cnt[i]; }
    offsets[n] = s;
    free(cnt);
    int32_t *ids = malloc((s ? s : 1) * sizeof(int32_t));
    if (!ids) return SJO_ENOMEM;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < (int64_t)n; p++) {
        uint64_t w = offsets[p];
        for (uint64_t q = 0; q < n; q++)
            if (sjo_dist2(pts + (uint64_t)p * d, pts + q * d, d) <= T) ids[w++] = (int32_t)q;
    }
    *ids_out = ids;
    return SJO_OK;
}

/* ------------------------------------------------------------------ */
/* W1: kernel.cu:33-133 (sortByWorkLoad) and :141-250 (LID variant):   */
/* sum of the sizes of the non-empty stencil cells, own cell included. */
/* ------------------------------------------------------------------ */
int sjo_workload(const sjo_grid *g, int lid, int64_t *nb)
{
    lookup_t L;
    stencil_t S;
    if (lookup_init(&L, g)) return SJO_ENOMEM;
    if (stencil_init(&S, g->k)) { free(L.dense); return SJO_ENOMEM; }
    const int k = g->k;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < (int64_t)g->nNonEmpty; a++) {
        const uint32_t *ca = g->cellCoord + a * k;
        int64_t s = g->indexMax[a] - g->indexMin[a] + 1;
        for (uint64_t t = 0; t < S.nOff; t++) {
            const int *o = S.off + t * k;
            int64_t b = neighbor(&L, ca, o);
            if (b < 0) continue;
            if (lid && g->cellLinear[b] < g->cellLinear[a]) continue;
            s += g->indexMax[b] - g->indexMin[b] + 1;
        }
        nb[a] = s;
    }
    free(L.dense); free(S.off);
    return SJO_OK;
}

typedef struct { int64_t w; uint32_t c; } wl_t;
static int cmp_wl(const void *x, const void *y)
{
    const wl_t *a = x, *b = y;
    if (a->w != b->w) return a->w > b->w ? -1 : 1;
    return (a->c > b->c) - (a->c < b->c);
}

/* W2: GPU.cu:1695-1698 descending by nbPoints (ties: ascending cell) */
int sjo_workload_order(const sjo_grid *g, int lid, uint32_t *order)
{
    int64_t *w = malloc(g->nNonEmpty * sizeof(int64_t));
    wl_t *t = malloc(g->nNonEmpty * sizeof(wl_t));
    if (!w || !t) { free(w); free(t); return SJO_ENOMEM; }
    int rc = sjo_workload(g, lid, w);
    if (rc) { free(w); free(t); return rc; }
    for (uint64_t c = 0; c < g->nNonEmpty; c++) { t[c].w = w[c]; t[c].c = (uint32_t)c; }
    qsort(t, g->nNonEmpty, sizeof(wl_t), cmp_wl);
    for (uint64_t c = 0; c < g->nNonEmpty; c++) order[c] = t[c].c;
    free(w); free(t);
    return SJO_OK;
}

/* ------------------------------------------------------------------ */
/* B1: kernel.cu:602-781 counts full-stencil hits of every stride-th   */
/* point (stride = 1/sampleRate = 100, GPU.cu:873-875,934-936).        */
/* ------------------------------------------------------------------ */
uint64_t sjo_sample_count(const sjo_grid *g, const double *pts, uint64_t stride)
{
    lookup_t L;
    stencil_t S;
    if (lookup_init(&L, g)) return 0;
    if (stencil_init(&S, g->k)) { free(L.dense); return 0; }
    const int d = g->d, k = g->k;
    const double T = sjo_eps_threshold(g->eps);
    /* number of sampled threads: N * sampleRate truncated (GPU.cu:890) */
    uint64_t nsamp = (uint64_t)((double)g->n * (1.0 / (double)stride));
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t t = 0; t < (int64_t)nsamp; t++) {
        uint64_t p = (uint64_t)t * stride;
        const double *x = pts + p * d;
        uint32_t c[SJO_MAXK];
        for (int j = 0; j < k; j++) cell_coord(x[j], g->minArr[j], g->eps, &c[j]);
        int zero[SJO_MAXK] = {0};
        for (uint64_t s = 0; s <= S.nOff; s++) {
            const int *o = s < S.nOff ? S.off + s * k : zero;
            int64_t b = neighbor(&L, c, o);
            if (b < 0) continue;
            for (int32_t j = g->indexMin[b]; j <= g->indexMax[b]; j++)
                total += sjo_dist2(x, pts + (uint64_t)g->indexLookup[j] * d, d) <= T;
        }
    }
    free(L.dense); free(S.off);
    return total;
}

/* B1, work-queue form: kernel.cu:785-942 takes point IDs 0..nsamp-1 of   */
/* the workload-ordered point list sortedCells (GPU.cu:1697-1716: cells  */
/* by descending workload, points of a cell in index order) through the  */
/* counter counterEstimator; nsamp = N * 0.01 (GPU.cu:1104-1106).  The    */
/* ordering workload is the LID one for modes 6 and 10, else the full    */
/* one (GPU.cu:1685-1692).  Cell ties are broken by ascending cell index. */
uint64_t sjo_sample_count_workqueue(const sjo_grid *g, const double *pts, uint64_t stride, int lid)
{
    if (stride == 0) stride = 1;
    const uint64_t nsamp = (uint64_t)((double)g->n * (1.0 / (double)stride));
    if (nsamp == 0) return 0;
    uint32_t *order = malloc((g->nNonEmpty ? g->nNonEmpty : 1) * sizeof(uint32_t));
    uint32_t *list = malloc(nsamp * sizeof(uint32_t));
    if (!order || !list) { free(order); free(list); return 0; }
    if (sjo_workload_order(g, lid, order)) { free(order); free(list); return 0; }
    uint64_t r = 0;
    for (uint64_t i = 0; i < g->nNonEmpty && r < nsamp; i++)
        for (int32_t j = g->indexMin[order[i]]; j <= g->indexMax[order[i]] && r < nsamp; j++)
            list[r++] = g->indexLookup[j];
    lookup_t L;
    stencil_t S;
    if (lookup_init(&L, g)) { free(order); free(list); return 0; }
    if (stencil_init(&S, g->k)) { free(L.dense); free(order); free(list); return 0; }
    const int d = g->d, k = g->k;
    const double T = sjo_eps_threshold(g->eps);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t t = 0; t < (int64_t)nsamp; t++) {
        const double *x = pts + (uint64_t)list[t] * d;
        uint32_t c[SJO_MAXK];
        for (int j = 0; j < k; j++) cell_coord(x[j], g->minArr[j], g->eps, &c[j]);
        int zero[SJO_MAXK] = {0};
        for (uint64_t s = 0; s <= S.nOff; s++) {
            const int *o = s < S.nOff ? S.off + s * k : zero;
            int64_t b = neighbor(&L, c, o);
            if (b < 0) continue;
            for (int32_t j = g->indexMin[b]; j <= g->indexMax[b]; j++)
                total += sjo_dist2(x, pts + (uint64_t)g->indexLookup[j] * d, d) <= T;
        }
    }
    free(L.dense); free(S.off); free(order); free(list);
    return total;
}

/* GPU.cu:1029-1055: est = cnt*offsetRate; buffer 40M ("size in HPBDC
 * paper"), alpha 0.05; small results use est*(1+2alpha)/3; batches =
 * ceil((1+alpha)*est/buffer); the returned buffer is 1.5x.  Guarded
 * against est == 0 (the reference divides by zero there). */
void sjo_batch_policy(uint64_t sampleCount, uint64_t offsetRate,
                      uint64_t *estimated, uint64_t *estimatedWithAlpha,
                      uint64_t *numBatches, uint64_t *bufferSize)
{
    const uint64_t streams = 3;
    uint64_t est = sampleCount * offsetRate;
    uint32_t buf = 40000000u;
    const double alpha = 0.05;
    *estimated = est;
    *estimatedWithAlpha = (uint64_t)((double)est * (1.0 + alpha * 1.0));
    if (est < (uint64_t)buf * streams)
        buf = (uint32_t)((double)est * (1.0 + (alpha * 2.0)) / (double)streams);
    if (buf == 0) buf = 1;
    uint64_t nb = (uint64_t)ceil(((1.0 + alpha) * (double)est * 1.0) / ((double)buf * 1.0));
    if (nb == 0) nb = 1;
    *numBatches = nb;
    *bufferSize = (uint64_t)(uint32_t)(1.5 * (double)buf);
}

/* ------------------------------------------------------------------ */
/* B3: GPU.cu:2595-2637 run-length over key-sorted pairs.  Each run    */
/* sets table[key] = {key, first, last} (inclusive).  Unlike the       */
/* reference it does not read keys[cnt].                               */
/* ------------------------------------------------------------------ */
int sjo_neighbor_table(const int32_t *keys, const int32_t *vals, uint64_t cnt,
                       int32_t batch, sjo_ntl *table)
{
    (void)vals;
    uint64_t i = 0;
    while (i < cnt) {
        uint64_t j = i;
        while (j + 1 < cnt && keys[j + 1] == keys[i]) j++;
        int32_t key = keys[i];
        table[key].pointID = key;
        table[key].indexmin = (int32_t)i;
        table[key].indexmax = (int32_t)j;
        table[key].batch = batch;
        i = j + 1;
    }
    return SJO_OK;
}

/* ------------------------------------------------------------------ */
/* L1: import_dataset.cpp:34-73.  Fields split on ','; inside a field, */
/* whitespace-separated numbers are read until one fails to parse      */
/* (std::stringstream >> double); then rows of d values are formed and */
/* a trailing partial row is dropped.                                  */
/* ------------------------------------------------------------------ */
static const char *scan_number(const char *s, const char *end)
{
    /* grammar accepted by num_get for double: [+-]? digits [. digits]
     * [(e|E) [+-]? digits]; at least one mantissa digit */
    const char *p = s;
    if (p < end && (*p == '+' || *p == '-')) p++;
    int digits = 0;
    while (p < end && *p >= '0' && *p <= '9') { p++; digits++; }
    if (p < end && *p == '.') {
        p++;
        while (p < end && *p >= '0' && *p <= '9') { p++; digits++; }
    }
    if (!digits) return NULL;
    if (p < end && (*p == 'e' || *p == 'E')) {
        const char *q = p + 1;
        if (q < end && (*q == '+' || *q == '-')) q++;
        int ed = 0;
        while (q < end && *q >= '0' && *q <= '9') { q++; ed++; }
        if (!ed) return NULL;
        p = q;
    }
    return p;
}

int sjo_import_csv(const char *path, int d, double **pts_out, uint64_t *n_out)
{
    *pts_out = NULL; *n_out = 0;
    if (d < 1) return SJO_EINVAL;
    FILE *f = fopen(path, "rb");
    if (!f) return SJO_EIO;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = malloc(sz + 1);
    if (!buf) { fclose(f); return SJO_ENOMEM; }
    if (fread(buf, 1, sz, f) != (size_t)sz) { free(buf); fclose(f); return SJO_EIO; }
    fclose(f);
    buf[sz] = 0;
    uint64_t cap = 1024, cnt = 0;
    double *vals = malloc(cap * sizeof(double));
    if (!vals) { free(buf); return SJO_ENOMEM; }
    const char *p = buf, *end = buf + sz;
    while (p < end) {
        const char *fe = memchr(p, ',', end - p);
        if (!fe) fe = end;
        const char *q = p;
        for (;;) {
            while (q < fe && (*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r' || *q == '\v' || *q == '\f')) q++;
            if (q >= fe) break;
            const char *ne = scan_number(q, fe);
            if (!ne) break;
            char tmp[512];
            size_t len = (size_t)(ne - q);
            if (len >= sizeof tmp) break;
            memcpy(tmp, q, len);
            tmp[len] = 0;
            if (cnt == cap) {
                cap *= 2;
                double *nv = realloc(vals, cap * sizeof(double));
                if (!nv) { free(vals); free(buf); return SJO_ENOMEM; }
                vals = nv;
            }
            vals[cnt++] = strtod(tmp, NULL);
            q = ne;
        }
        p = fe + 1;
    }
    free(buf);
    *n_out = cnt / (uint64_t)d;
    *pts_out = vals;
    return SJO_OK;
}

/* import_dataset.cpp:16-30,76-80: lexicographic order of int(x_i) over
 * all d dims.  std::sort is unstable; the oracle uses a stable merge sort
 * (ties keep input order). */
static const double *g_sortpts;
static int g_sortd;
static int cmp_bins(const uint64_t a, const uint64_t b)
{
    for (int i = 0; i < g_sortd; i++) {
        int x = (int)g_sortpts[a * g_sortd + i], y = (int)g_sortpts[b * g_sortd + i];
        if (x < y) return -1;
        if (x > y) return 1;
    }
    return 0;
}

static void merge_sort_idx(uint64_t *a, uint64_t *tmp, uint64_t n)
{
    if (n < 2) return;
    uint64_t h = n / 2;
    merge_sort_idx(a, tmp, h);
    merge_sort_idx(a + h, tmp, n - h);
    uint64_t i = 0, j = h, w = 0;
    while (i < h && j < n) tmp[w++] = cmp_bins(a[j], a[i]) < 0 ? a[j++] : a[i++];
    while (i < h) tmp[w++] = a[i++];
    while (j < n) tmp[w++] = a[j++];
    memcpy(a, tmp, n * sizeof(uint64_t));
}

int sjo_sort_in_nd_bins_stable(double *pts, uint64_t n, int d)
{
    for (uint64_t i = 0; i < n * (uint64_t)d; i++)
        if (!(fabs(pts[i]) < 2147483648.0)) return SJO_EOVERFLOW;
    uint64_t *idx = malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t *tmp = malloc((n ? n : 1) * sizeof(uint64_t));
    double *cp = malloc((n ? n : 1) * d * sizeof(double));
    if (!idx || !tmp || !cp) { free(idx); free(tmp); free(cp); return SJO_ENOMEM; }
    for (uint64_t i = 0; i < n; i++) idx[i] = i;
    g_sortpts = pts; g_sortd = d;
    merge_sort_idx(idx, tmp, n);
    memcpy(cp, pts, n * d * sizeof(double));
    for (uint64_t i = 0; i < n; i++) memcpy(pts + i * d, cp + idx[i] * d, d * sizeof(double));
    free(idx); free(tmp); free(cp);
    return SJO_OK;
}
