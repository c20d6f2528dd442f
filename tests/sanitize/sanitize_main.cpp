// Drives the CPU oracle and the library's host C++ (CSV import, bin
// pre-sort) under AddressSanitizer and UndefinedBehaviorSanitizer: every
// entry point on small random, degenerate and malformed inputs, with the
// results cross-checked so the calls do real work.  Exit status 0: clean.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sj_oracle.h"
#include "sjamd.h"

namespace sj {
// the library's error sink (sj_prims.hip) for the host-only build
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    va_end(ap);
    (void)fmt;
}
sj_status import_csv(const char *path, int d, double **pts, uint64_t *n);
sj_status sort_in_nd_bins(double *pts, uint64_t n, int d, int stable);
}  // namespace sj

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            fails++;                                                      \
        }                                                                 \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand() {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * 0x1.0p-53;
}

static void join_checks(int n, int d, int k, double eps, double scale) {
    std::vector<double> pts((size_t)n * d);
    for (auto &v : pts) v = urand() * scale;
    sjo_grid g;
    memset(&g, 0, sizeof g);
    CHECK(sjo_grid_build(pts.data(), n, d, k, eps, &g) == 0);
    std::vector<uint64_t> off(n + 1), boff(n + 1);
    int32_t *bids = nullptr;
    CHECK(sjo_brute(pts.data(), n, d, eps, 2, boff.data(), &bids) == 0);
    for (int form = 0; form < 4; form++) {
        int32_t *ids = nullptr;
        CHECK(sjo_join(&g, pts.data(), form, 2, 1, off.data(), &ids) == 0);
        CHECK(off[n] == boff[n]);
        sjo_free(ids);
    }
    std::vector<int64_t> w(g.nNonEmpty + 1);
    std::vector<uint32_t> order(g.nNonEmpty + 1);
    for (int lid = 0; lid < 2; lid++) {
        CHECK(sjo_workload(&g, lid, w.data()) == 0);
        CHECK(sjo_workload_order(&g, lid, order.data()) == 0);
        (void)sjo_sample_count_workqueue(&g, pts.data(), 7, lid);
    }
    (void)sjo_sample_count(&g, pts.data(), 3);
    // the key/value view of the brute-force table and its run-length table
    std::vector<int32_t> keys(boff[n]);
    for (int p = 0; p < n; p++)
        for (uint64_t i = boff[p]; i < boff[p + 1]; i++) keys[i] = p;
    std::vector<sjo_ntl> tab(n);
    CHECK(sjo_neighbor_table(keys.data(), bids, boff[n], 0, tab.data()) == 0);
    for (int p = 0; p < n; p++) CHECK(tab[p].indexmax - tab[p].indexmin + 1 == (int)(boff[p + 1] - boff[p]));
    std::vector<double> d2(n);
    sjo_dist2_batch(pts.data(), pts.data() + d, (uint64_t)(n - 1), d, d2.data());
    sjo_free(bids);
    sjo_grid_free(&g);
}

int main() {
    join_checks(600, 2, 2, 0.05, 1.0);
    join_checks(400, 3, 3, 0.12, 1.0);
    join_checks(300, 6, 4, 0.6, 1.0);
    join_checks(200, 8, 2, 0.9, 1.0);
    join_checks(50, 2, 2, 10.0, 1.0);   // every point in one cell
    join_checks(1, 3, 3, 0.1, 1.0);     // a single point
    // degenerate grid arguments are rejected, not read past
    {
        double p[4] = {0, 0, 1, 1};
        sjo_grid g;
        memset(&g, 0, sizeof g);
        CHECK(sjo_grid_build(p, 2, 2, 2, 0.0, &g) != 0);
        CHECK(sjo_grid_build(p, 2, 2, 3, 0.5, &g) != 0);
        p[1] = NAN;
        CHECK(sjo_grid_build(p, 2, 2, 2, 0.5, &g) != 0);
        CHECK(std::isnan(sjo_eps_threshold(-1.0)));
    }
    uint64_t est, esta, nb, buf;
    sjo_batch_policy(0, 100, &est, &esta, &nb, &buf);
    CHECK(nb == 1);
    sjo_batch_policy(1280000, 100, &est, &esta, &nb, &buf);
    CHECK(nb == 4 && buf == 60000000ull);

    // CSV import: whitespace, exponents, a bad token and a partial row
    const char *path = "sanitize_points.csv";
    FILE *f = fopen(path, "w");
    fprintf(f, "0.5, 1e-3\n  2.25,-4\n7,8,9\n\n1.5,x\n3");
    fclose(f);
    for (int pass = 0; pass < 2; pass++) {
        double *a = nullptr;
        uint64_t na = 0;
        const int rc = pass ? sjo_import_csv(path, 2, &a, &na) : (int)sj::import_csv(path, 2, &a, &na);
        CHECK(rc == 0);
        if (a && na) {
            std::vector<double> s(a, a + na * 2);
            CHECK(sj::sort_in_nd_bins(s.data(), na, 2, 0) == SJ_OK);
            CHECK(sj::sort_in_nd_bins(s.data(), na, 2, 1) == SJ_OK);
            CHECK(sjo_sort_in_nd_bins_stable(s.data(), na, 2) == 0);
        }
        if (pass) sjo_free(a);
        else free(a);
    }
    double *a = nullptr;
    uint64_t na = 0;
    CHECK(sj::import_csv("no/such/file.csv", 2, &a, &na) == SJ_ERR_IO);
    CHECK(sj::import_csv(path, 0, &a, &na) == SJ_ERR_INVALID);
    remove(path);
    if (fails) {
        fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    printf("sanitize_host: clean\n");
    return 0;
}
