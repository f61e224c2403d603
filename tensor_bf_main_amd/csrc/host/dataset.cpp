// dataset.cpp -- dataset readers.
//
// importDataset follows the reading rules of the reference (dataset.cu:10-39):
// each step reads one number with "%lf, " semantics (blanks skipped, one number,
// then a comma and the blanks after it when a comma follows), rows are formed
// every `dim` values whatever the line structure, a trailing partial row is
// dropped, and reading stops at the first token that is not a number.
#include "dataset.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

// One "%lf, " step on a buffered stream.  Returns false at EOF or on a non-number.
bool read_value(FILE *f, double *v)
{
    if (std::fscanf(f, "%lf", v) != 1) return false;
    int c = std::fgetc(f);
    if (c == ',') {
        do {
            c = std::fgetc(f);
        } while (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f');
    }
    if (c != EOF) std::ungetc(c, f);
    return true;
}

} // namespace

int importDatasetDim(std::vector<std::vector<double>> *inputVector, const char *filename, int dim)
{
    if (!inputVector || !filename || dim < 1) return -1;
    FILE *f = std::fopen(filename, "r");
    if (!f) return -1;
    std::vector<double> row;
    row.reserve(dim);
    double v = 0.0;
    while (read_value(f, &v)) {
        row.push_back(v);
        if ((int)row.size() == dim) {
            inputVector->push_back(row);
            row.clear();
        }
    }
    std::fclose(f);
    return 0;
}

void importDataset(std::vector<std::vector<INPUT_DATA_TYPE>> *inputVector, char *filename)
{
    if (importDatasetDim(inputVector, filename, INPUT_DATA_DIM) != 0) {
        std::fprintf(stderr, "[Dataset] ~ Could not open the input file %s\n", filename);
        std::exit(1);
    }
}

static const char kMagic[8] = {'B', 'F', 'J', 'D', 'A', 'T', 'A', '1'};

int importDatasetBinary(std::vector<double> *rows, int64_t *n, int64_t *d, const char *filename)
{
    if (!rows || !n || !d || !filename) return -1;
    FILE *f = std::fopen(filename, "rb");
    if (!f) return -1;
    char magic[8];
    int64_t hdr[2];
    int rc = -1;
    if (std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 && std::fread(hdr, 8, 2, f) == 2 &&
        hdr[0] >= 0 && hdr[1] >= 1 && hdr[1] <= (int64_t)1 << 20) {
        rows->resize((size_t)(hdr[0] * hdr[1]));
        if (std::fread(rows->data(), 8, rows->size(), f) == rows->size()) {
            *n = hdr[0];
            *d = hdr[1];
            rc = 0;
        }
    }
    std::fclose(f);
    return rc;
}

int exportDatasetBinary(const double *rows, int64_t n, int64_t d, const char *filename)
{
    if ((!rows && n > 0) || n < 0 || d < 1 || !filename) return -1;
    FILE *f = std::fopen(filename, "wb");
    if (!f) return -1;
    const int64_t hdr[2] = {n, d};
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, 8, 2, f) == 2 &&
              (n == 0 || std::fwrite(rows, 8, (size_t)(n * d), f) == (size_t)(n * d));
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}
