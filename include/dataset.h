/*
 * dataset.h -- dataset readers.  importDataset keeps the reference signature
 * (dataset.h:8-10, dataset.cu:10-39); the others add a run-time dimension and a
 * binary format.  Implemented in tensor_bf_main_amd/csrc/host/dataset.cpp.
 */
#ifndef BFJ_DATASET_H
#define BFJ_DATASET_H

#include <stdint.h>
#include <vector>

#include "params.h"

/* dataset.cu:10-39: values separated by ", " or blanks, grouped into rows of
 * INPUT_DATA_DIM values regardless of line breaks; a trailing partial row is
 * dropped; a file that cannot be opened prints an error and exits(1). */
void importDataset(std::vector<std::vector<INPUT_DATA_TYPE>> *inputVector, char *filename);

/* Same reader with the row length given at run time.  Returns 0, or -1 when the
 * file cannot be opened (no exit). */
int importDatasetDim(std::vector<std::vector<double>> *inputVector, const char *filename, int dim);

/* Binary format: 8-byte magic "BFJDATA1", int64 n, int64 d, then n*d little-endian
 * doubles, row-major.  Returns 0 on success, -1 on I/O or format errors. */
int importDatasetBinary(std::vector<double> *rows, int64_t *n, int64_t *d, const char *filename);
int exportDatasetBinary(const double *rows, int64_t n, int64_t d, const char *filename);

#endif /* BFJ_DATASET_H */
