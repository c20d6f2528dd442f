// Dataset import and the lexicographic bin pre-sort (host C++).
//   importNDDataset   import_dataset.cpp:34-73
//   sortInNDBins      import_dataset.cpp:16-30,76-80
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "sjamd.h"

namespace sj {
void set_error(const char *fmt, ...);

// Fields are split on ','; each field is read with operator>>(double)
// until extraction fails; values form rows of d and a trailing partial row
// is dropped (import_dataset.cpp:40-70).
sj_status import_csv(const char *path, int d, double **pts, uint64_t *n) {
    if (!path || !pts || !n || d < 1) {
        set_error("sj_import_csv: invalid arguments");
        return SJ_ERR_INVALID;
    }
    *pts = nullptr;
    *n = 0;
    std::ifstream in(path);
    if (!in) {
        set_error("sj_import_csv: cannot open %s", path);
        return SJ_ERR_IO;
    }
    std::vector<double> all;
    for (std::string f; std::getline(in, f, ',');) {
        std::stringstream ss(f);
        double v;
        while (ss >> v) {
            all.push_back(v);
            if (ss.peek() == ',') ss.ignore();
        }
    }
    const uint64_t rows = all.size() / (uint64_t)d;
    double *out = static_cast<double *>(malloc((rows ? rows : 1) * d * sizeof(double)));
    if (!out) {
        set_error("sj_import_csv: out of host memory");
        return SJ_ERR_NOMEM;
    }
    std::copy(all.begin(), all.begin() + rows * d, out);
    *pts = out;
    *n = rows;
    return SJ_OK;
}

// Lexicographic order of int(x_i) over all d coordinates.  stable=0 runs
// std::sort with the reference comparator over an index array (same
// comparison sequence as sorting the rows themselves, hence the same
// libstdc++ permutation); stable=1 keeps ties in input order.
sj_status sort_in_nd_bins(double *pts, uint64_t n, int d, int stable) {
    if ((!pts && n) || d < 1) {
        set_error("sj_sort_in_nd_bins: invalid arguments");
        return SJ_ERR_INVALID;
    }
    for (uint64_t i = 0; i < n * (uint64_t)d; i++)
        if (!(std::fabs(pts[i]) < 2147483648.0)) {
            set_error("sj_sort_in_nd_bins: coordinate outside int range");
            return SJ_ERR_OVERFLOW;
        }
    std::vector<std::vector<double>> rows(n, std::vector<double>(d));
    for (uint64_t i = 0; i < n; i++) std::copy(pts + i * d, pts + (i + 1) * d, rows[i].begin());
    auto cmp = [d](const std::vector<double> &a, const std::vector<double> &b) {
        for (int i = 0; i < d; i++) {
            if (int(a[i]) < int(b[i])) return true;
            if (int(a[i]) > int(b[i])) return false;
        }
        return false;
    };
    if (stable)
        std::stable_sort(rows.begin(), rows.end(), cmp);
    else
        std::sort(rows.begin(), rows.end(), cmp);
    for (uint64_t i = 0; i < n; i++) std::copy(rows[i].begin(), rows[i].end(), pts + i * d);
    return SJ_OK;
}

}  // namespace sj
