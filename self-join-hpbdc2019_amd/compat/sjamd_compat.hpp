// C++ host interface with the names, argument meaning and output layout of
// the reference program (benoitgallet/self-join-hpbdc2019), implemented on
// top of the C ABI in sjamd.h.  A program written against the reference's
// GPU.h / main.cu prototypes links against libsjamd_compat.so instead.
//
// Differences a caller can see:
//   * GPUNUMDIM / NUMINDEXEDDIM are runtime values (sjamdSetDims) instead of
//     compile-time macros (params.h:9-10);
//   * result pairs always carry original point IDs (the reference mixes
//     namespaces when SORT_BY_WORKLOAD is set, kernel.cu:978-988);
//   * errors are reported through sjamdLastStatus() besides the message on
//     stderr; the reference prints and carries on (GPU.cu:1338-1341).
#pragma once

#include <cstdint>
#include <cstdio>
#include <ostream>
#include <vector>

#ifndef DTYPE
#define DTYPE double
#endif

// structs.h:41-47
struct neighborTableLookup {
    int pointID;
    int indexmin;
    int indexmax;
    int *dataPtr;
};

// structs.h:54-57
struct neighborDataPtrs {
    int *dataPtr;
    int sizeOfDataArr;
};

// structs.h:81-85
struct grid {
    int indexmin;
    int indexmax;
};

// structs.h:88-97
struct gridCellLookup {
    unsigned int idx;
    uint64_t gridLinearID;
    bool operator<(const gridCellLookup &other) const { return gridLinearID < other.gridLinearID; }
};

// structs.h:99-102
struct schedulingCell {
    int nbPoints;
    int cellId;
};

// Runtime replacement of GPUNUMDIM / NUMINDEXEDDIM (params.h:9-10).
void sjamdSetDims(int gpunumdim, int numindexeddim);
int sjamdGpuNumDim();
int sjamdNumIndexedDim();
// Status of the last call (0 = success, else an sj_status code).
int sjamdLastStatus();

// import_dataset.cpp:34
void importNDDataset(std::vector<std::vector<DTYPE>> *dataPoints, char *fname);
// import_dataset.cpp:76 (libstdc++ std::sort with the reference comparator)
void sortInNDBins(std::vector<std::vector<DTYPE>> *dataPoints);

// main.cu:446-503.  minArr, maxArr, nCells: NUMINDEXEDDIM entries.
void generateNDGridDimensions(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon, DTYPE *minArr,
                              DTYPE *maxArr, unsigned int *nCells, uint64_t *totalCells);

// main.cu:240-413.  Allocates *gridCellLookupArr, *index, *gridCellNDMask
// with new[]; indexLookupArr and gridCellNDMaskOffsets are caller arrays
// (N and 2*NUMINDEXEDDIM entries).
void populateNDGridIndexAndLookupArray(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon,
                                       struct gridCellLookup **gridCellLookupArr, struct grid **index,
                                       unsigned int *indexLookupArr, DTYPE *minArr, unsigned int *nCells,
                                       uint64_t totalCells, unsigned int *nNonEmptyCells,
                                       unsigned int **gridCellNDMask, unsigned int *gridCellNDMaskOffsets,
                                       unsigned int *nNDMaskElems);

// GPU.h:37-40.  searchMode 3..10 (main.cu:88-150).  Fills neighborTable[N]
// and appends one data array per batch to pointersToNeighbors (the caller
// owns those arrays, as in the reference).  The grid arguments describe the
// caller's host grid; the device rebuilds the same grid from the points.
void distanceTableNDGridBatches(int searchMode, std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE *epsilon,
                                struct grid *index, struct gridCellLookup *gridCellLookupArr,
                                unsigned int *nNonEmptyCells, DTYPE *minArr, unsigned int *nCells,
                                unsigned int *indexLookupArr, struct neighborTableLookup *neighborTable,
                                std::vector<struct neighborDataPtrs> *pointersToNeighbors,
                                uint64_t *totalNeighbors, unsigned int *gridCellNDMask,
                                unsigned int *gridCellNDMaskOffsets, unsigned int *nNDMaskElems);

// GPU.h:42-45 (callGPUBatchEst): sampled result-size estimate of every
// 100th point; returns the estimate with alpha and sets the batch count and
// the 1.5x buffer size (GPU.cu:1029-1063).  The reference's device-array
// arguments are replaced by the points: the device builds its own grid.
unsigned long long callGPUBatchEst(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon,
                                   unsigned int *retNumBatches, unsigned int *retGPUBufferSize);
// GPU.h:47-50 (callGPUBatchEstWorkQueue): the sample is the first 1% of the
// workload-ordered points (LID workload for searchMode 6 and 10, full
// otherwise, GPU.cu:1685-1692).
unsigned long long callGPUBatchEstWorkQueue(std::vector<std::vector<DTYPE>> *NDdataPoints, DTYPE epsilon,
                                            int searchMode, unsigned int *retNumBatches,
                                            unsigned int *retGPUBufferSize);

// GPU.h:52: run-length neighbour table over key-sorted (key,value) pairs.
void constructNeighborTableKeyValueWithPtrs(int *pointIDKey, int *pointInDistValue,
                                            struct neighborTableLookup *neighborTable, int *pointersToNeighbors,
                                            unsigned int *cnt);

// GPU.h:54
void warmUpGPU();

// ---- the reference program's outputs (main.cu:88-150, 203-222) --------
// Stats file of a search mode ("gpu_b-unicomp_stats.txt" for mode 5, ...),
// NULL outside 3..10.
const char *searchModeStatsFile(int searchMode);
// Banner text of a search mode ("B-Unicomp" for mode 5, ...), NULL outside 3..10.
const char *searchModeName(int searchMode);
// One stats line (main.cu:203-206): time, dataset, epsilon, |S|, then
// GPUNUMDIM, NUMINDEXEDDIM, DTYPE, BLOCKSIZE, THREADPERPOINT and
// SORT_BY_WORKLOAD as the reference build reports them (params.h defaults:
// 256, 1, 1).
void writeStatsLine(std::ostream &out, double totalTime, const char *inputFname, DTYPE epsilon,
                    uint64_t totalNeighbors);
// Neighbour-table listing (main.cu:212-222): every point's neighbours
// sorted in place, then "\npoint id: i, neighbors: nb neighbor c\n" and the
// IDs at positions indexmin .. indexmax-1, each followed by a comma (the
// reference's loop stops one short of indexmax, so the last, largest
// neighbour is not listed).
void printNeighborTable(FILE *out, struct neighborTableLookup *neighborTable, size_t n);
