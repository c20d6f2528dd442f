// Driver with the reference program's flow (main.cu:36-226): import,
// bin pre-sort, grid dimensions, grid index, batched join, neighbour table,
// the per-mode stats line and the neighbour-table listing.
// Usage: sjamd_selfjoin dataset eps ndim searchmode [nindexed] [table.txt]
// The stats line is appended to gpu_<mode>_stats.txt in the working
// directory (main.cu:88-150,203-208).  The listing (main.cu:212-222) goes to
// table.txt when given (the reference prints it to stdout, which at 2M
// points is gigabytes of text).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "sjamd_compat.hpp"

int main(int argc, char *argv[]) {
    if (argc < 5) {
        printf("\n\nIncorrect number of input parameters.  \nShould be dataset file, epsilon, number of "
               "dimensions, searchmode [, indexed dimensions, table file]\n");
        return 1;
    }
    const double epsilon = atof(argv[2]);
    const int ndim = atoi(argv[3]);
    const int mode = atoi(argv[4]);
    const int nidx = argc > 5 ? atoi(argv[5]) : ndim;
    const char *dump = argc > 6 ? argv[6] : nullptr;
    sjamdSetDims(ndim, nidx);
    printf("\nDataset file: %s", argv[1]);
    printf("\nEpsilon: %f", epsilon);
    printf("\nNumber of dimensions (NDIM): %d\n", ndim);

    std::vector<std::vector<DTYPE>> pts;
    importNDDataset(&pts, argv[1]);
    if (sjamdLastStatus() || pts.empty()) return 2;
    sortInNDBins(&pts);

    std::ofstream gpu_stats;
    const char *stats_name = searchModeStatsFile(mode);
    if (stats_name) {
        printf("\n\n# Search mode is : %s #\n\n", searchModeName(mode));
        gpu_stats.open(stats_name, std::ios::app);
    }
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
    printf("\n~~~Time: %f", t);
    printf("\ntotalNeighbors: %llu\nbatches: %zu\n", (unsigned long long)totalNeighbors, batches.size());
    if (sjamdLastStatus()) return 3;
    if (stats_name) {
        writeStatsLine(gpu_stats, t, argv[1], epsilon, totalNeighbors);
        gpu_stats.close();
    }

    if (dump) {
        FILE *f = fopen(dump, "w");
        if (!f) return 4;
        printNeighborTable(f, table.data(), pts.size());
        fclose(f);
    }
    for (auto &b : batches) delete[] b.dataPtr;
    delete[] index;
    delete[] lookup;
    delete[] mask;
    return 0;
}
