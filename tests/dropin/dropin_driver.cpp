// dropin_driver.cpp -- a caller of the reference's host API, written against the
// reference's own types and compile-time configuration (test program, not product).
//
// Built with INPUT_DATA_DIM 18, COMPUTE_DIM 32 and ACCUM_PREC 32 (the configuration the
// reference's sample files test100.txt / test128.txt need), it does what the
// reference's main.cu does with them:
//   * eps parsed into ACCUM_TYPE (main.cu:20);
//   * rows read by importDataset(std::vector<std::vector<INPUT_DATA_TYPE>> *, char *)
//     (dataset.h:8-10);
//   * the (n + ADDITIONAL_POINTS) x COMPUTE_DIM zero-padded host array built
//     (main.cu:39-57);
//   * GPUJoinMainBruteForce(searchMode, device, dataset, &nbQueryPoints, &epsilon,
//     &totalNeighbors) called (gpu_join.h:36-42, main.cu:94) once per search mode on
//     the command line.
// It prints "mode <m> total <t>" per call and checks that the call left its inputs
// (the padded array, nbQueryPoints, epsilon) unchanged, as the reference's callee does.
//
// usage: dropin_driver <file> <eps> <searchMode>...
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dataset.h"
#include "gpu_join.h"
#include "params.h"

static uint64_t fingerprint(const INPUT_DATA_TYPE *p, size_t count)
{
    uint64_t h = 1469598103934665603ull; // FNV-1a over the bytes
    const unsigned char *b = reinterpret_cast<const unsigned char *>(p);
    for (size_t k = 0; k < count * sizeof(INPUT_DATA_TYPE); ++k) h = (h ^ b[k]) * 1099511628211ull;
    return h;
}

int main(int argc, char *argv[])
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <file> <eps> <searchMode>...\n", argv[0]);
        return 2;
    }
    char filename[256];
    std::snprintf(filename, sizeof filename, "%s", argv[FILENAME_ARG]);
    ACCUM_TYPE epsilon = (ACCUM_TYPE)std::atof(argv[EPSILON_ARG]);

    std::vector<std::vector<INPUT_DATA_TYPE>> inputVector;
    importDataset(&inputVector, filename);
    unsigned int nbQueryPoints = (unsigned int)inputVector.size();

    // Rows at a stride of COMPUTE_DIM, zeros past INPUT_DATA_DIM, then ADDITIONAL_POINTS zero rows.
    const size_t cells = (size_t)(nbQueryPoints + ADDITIONAL_POINTS) * COMPUTE_DIM;
    std::vector<INPUT_DATA_TYPE> database(cells, (INPUT_DATA_TYPE)0.0);
    for (unsigned int i = 0; i < nbQueryPoints; ++i)
        std::memcpy(&database[(size_t)i * COMPUTE_DIM], inputVector[i].data(), INPUT_DATA_DIM * sizeof(INPUT_DATA_TYPE));

    const uint64_t before = fingerprint(database.data(), cells);
    const unsigned int n_before = nbQueryPoints;
    const ACCUM_TYPE eps_before = epsilon;
    std::printf("rows %u dim %d compute_dim %d eps %.17g\n", nbQueryPoints, INPUT_DATA_DIM, COMPUTE_DIM,
                (double)epsilon);
    for (int a = SEARCHMODE_ARG; a < argc; ++a) {
        const unsigned int searchMode = (unsigned int)std::atoi(argv[a]);
        uint64_t totalResult = 12345; // the callee must overwrite it
        GPUJoinMainBruteForce(searchMode, 0, database.data(), &nbQueryPoints, &epsilon, &totalResult);
        std::printf("mode %u total %llu\n", searchMode, (unsigned long long)totalResult);
        if (fingerprint(database.data(), cells) != before || nbQueryPoints != n_before || epsilon != eps_before) {
            std::fprintf(stderr, "mode %u: the call changed its inputs\n", searchMode);
            return 1;
        }
    }
    return 0;
}
