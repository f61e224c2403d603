/*
 * gpu_join.h -- the reference's host entry point, kept as the drop-in surface.
 * Implemented in tensor_bf_main_amd/csrc/host/gpu_join.cpp on top of the C ABI
 * (bfjoin.h).  C++ linkage, as in the reference.
 */
#ifndef BFJ_GPU_JOIN_H
#define BFJ_GPU_JOIN_H

#include <stdint.h>
#include <vector>

#include "bfjoin.h"
#include "params.h"

/*
 * Replaces GPUJoinMainBruteForce (gpu_join.h:36-42, gpu_join.cu:210-322).
 * dataset: (*nbQueryPoints + ADDITIONAL_POINTS) x COMPUTE_DIM doubles, row-major,
 * first INPUT_DATA_DIM columns used (main.cu:39-57 builds it).  On return
 * *totalNeighbors holds the number of ordered pairs (i, j), self included, with
 * sqrt(d2(i, j)) <= *epsilon.  An unknown searchMode prints an error and exits,
 * as gpu_join.cu:311-315 does; HIP errors are printed and the total is 0.
 */
void GPUJoinMainBruteForce(unsigned int searchMode, unsigned int device, INPUT_DATA_TYPE *dataset,
                           unsigned int *nbQueryPoints, ACCUM_TYPE *epsilon, uint64_t *totalNeighbors);

/*
 * gpu_join.h:27-34.  The reference's body is commented out and its CLI rejects
 * searchMode 5 (main.cu:64-83); this keeps the symbol and reports that.
 */
void GPUJoinMainBruteForceNvidia(unsigned int searchMode, unsigned int device, INPUT_DATA_TYPE *dataset,
                                 INPUT_DATA_TYPE *datasetTranspose, unsigned int *nbQueryPoints,
                                 ACCUM_TYPE *epsilon, uint64_t *totalNeighbors);

/*
 * Run-time variant (no counterpart in the reference): dimensions, a double eps,
 * optional per-point counts (self excluded, length n) and statistics.  Returns a
 * bfj_status instead of exiting.
 */
int GPUJoinMainBruteForceEx(unsigned int searchMode, unsigned int device, const double *dataset, unsigned int n,
                            unsigned int dim, unsigned int computeDim, double epsilon, uint64_t *totalNeighbors,
                            int64_t *counts, bfj_stats *stats);

/*
 * Result-set variant (no counterpart in the reference, which only counts): every
 * unordered pair i < j within epsilon, appended to *pairs as (i, j), in no
 * particular order.  *totalNeighbors gets the reference's ordered, self-inclusive
 * total (n + 2 * pairs).  Returns a bfj_status.
 */
int GPUJoinMainBruteForcePairs(unsigned int searchMode, unsigned int device, const double *dataset, unsigned int n,
                               unsigned int dim, unsigned int computeDim, double epsilon,
                               std::vector<uint32_t> *pairs, uint64_t *totalNeighbors, bfj_stats *stats);

/*
 * Bug-for-bug variant (no counterpart as a function in the reference): the total of
 * searchMode 1-4 computed with the reference's own arithmetic at computePrec /
 * accumPrec (params.h COMPUTE_PREC / ACCUM_PREC), bfj_reference_mode_join.  Returns a
 * bfj_status.
 */
int GPUJoinMainBruteForceReference(unsigned int searchMode, unsigned int device, const double *dataset, unsigned int n,
                                   unsigned int dim, unsigned int computeDim, double epsilon, int computePrec,
                                   int accumPrec, uint64_t *totalNeighbors);

/*
 * Accuracy of the last join on `device` through this file's entry points (no
 * counterpart in the reference): bfj_pass_errors on the device's context.  Call it
 * right after the join.  Returns a bfj_status.
 */
int GPUJoinPassErrors(unsigned int device, uint64_t maxCandidates, int64_t sampleTiles, uint64_t seed,
                      bfj_errors *errors);

/* searchMode -> bfj_dtype, or -1 for modes that are errors (5, 6, unknown). */
int searchModeDtype(unsigned int searchMode);

#endif /* BFJ_GPU_JOIN_H */
