// Driver with the reference program's flow (main.cu:36-226): import,
// bin pre-sort, grid dimensions, grid index, batched join, neighbour table.
// Usage: sjamd_selfjoin dataset eps ndim searchmode [nindexed] [dump.txt]
// The optional dump lists every point's neighbours, sorted, one line each.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sjamd_compat.hpp"

int main(int argc, char *argv[]) {
    if (argc < 5) {
        printf("\n\nIncorrect number of input parameters.  \nShould be dataset file, epsilon, number of "
               "dimensions, searchmode [, indexed dimensions, dump file]\n");
        return 1;
    }
    const double epsilon = atof(argv[2]);
    const int ndim = atoi(argv[3]);
    const int mode = atoi(argv[4]);
    const int nidx = argc > 5 ? atoi(argv[5]) : ndim;
    const char *dump = argc > 6 ? argv[6] : nullptr;
    sjamdSetDims(ndim, nidx);

    std::vector<std::vector<DTYPE>> pts;
    importNDDataset(&pts, argv[1]);
    if (sjamdLastStatus() || pts.empty()) return 2;
    sortInNDBins(&pts);
    warmUpGPU();

    std::vector<DTYPE> minArr(nidx), maxArr(nidx);
    std::vector<unsigned int> nCells(nidx);
    uint64_t totalCells = 0;
    generateNDGridDimensions(&pts, epsilon, minArr.data(), maxArr.data(), nCells.data(), &totalCells);
    printf("\nGrid: total cells (including empty) %llu", (unsigned long long)totalCells);

    grid *index = nullptr;
    gridCellLookup *lookup = nullptr;
    unsigned int *mask = nullptr, nMask = 0, nNonEmpty = 0;
    std::vector<unsigned int> maskOffsets(2 * nidx);
    std::vector<unsigned int> indexLookupArr(pts.size());
    populateNDGridIndexAndLookupArray(&pts, epsilon, &lookup, &index, indexLookupArr.data(), minArr.data(),
                                      nCells.data(), totalCells, &nNonEmpty, &mask, maskOffsets.data(), &nMask);

    std::vector<neighborTableLookup> table(pts.size());
    std::vector<neighborDataPtrs> batches;
    uint64_t totalNeighbors = 0;
    double eps = epsilon;
    const auto t0 = std::chrono::steady_clock::now();
    distanceTableNDGridBatches(mode, &pts, &eps, index, lookup, &nNonEmpty, minArr.data(), nCells.data(),
                               indexLookupArr.data(), table.data(), &batches, &totalNeighbors, mask,
                               maskOffsets.data(), &nMask);
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("\nTime: %f\ntotalNeighbors: %llu\nbatches: %zu\n", t, (unsigned long long)totalNeighbors,
           batches.size());
    if (sjamdLastStatus()) return 3;

    if (dump) {
        FILE *f = fopen(dump, "w");
        if (!f) return 4;
        for (size_t i = 0; i < pts.size(); i++) {
            std::vector<int> v;
            if (table[i].dataPtr)
                v.assign(table[i].dataPtr + table[i].indexmin, table[i].dataPtr + table[i].indexmax + 1);
            std::sort(v.begin(), v.end());
            fprintf(f, "%zu:", i);
            for (int x : v) fprintf(f, " %d", x);
            fprintf(f, "\n");
        }
        fclose(f);
    }
    for (auto &b : batches) delete[] b.dataPtr;
    delete[] index;
    delete[] lookup;
    delete[] mask;
    return 0;
}
