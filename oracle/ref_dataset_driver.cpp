// Driver for the reference dataset reader compiled from its own sources into
// oracle/_ref/.  Prints the parsed rows so tests can compare the product's
// reader with the reference's on the same file.  Test infrastructure only.
#include <cstdio>
#include <vector>
#include "dataset.h"

int main(int argc, char **argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: ref_import <file>\n"); return 2; }
    std::vector<std::vector<INPUT_DATA_TYPE>> rows;
    importDataset(&rows, argv[1]);
    std::printf("%zu %d\n", rows.size(), (int)INPUT_DATA_DIM);
    for (const auto &r : rows) {
        for (size_t k = 0; k < r.size(); ++k) std::printf(k ? " %.17g" : "%.17g", (double)r[k]);
        std::printf("\n");
    }
    return 0;
}
