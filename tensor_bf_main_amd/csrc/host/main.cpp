// main.cpp -- the reference's CLI (main.cu:16-128) on top of gpu_join.h.
//
//   bfjoin_main <file> <eps> <searchMode> [device] [--dim D] [--compute-dim C]
//               [--binary] [--counts out.txt] [--pairs out.txt] [--log file] [--parse-only]
//               [--eps-double] [--reference-arithmetic [--compute-prec P] [--accum-prec P]]
//
// Same positional arguments, console lines and CSV result log as the reference;
// the options give the row length at run time (the reference fixes it at compile
// time, params.h:4-5), read the binary format, and write per-point counts.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "dataset.h"
#include "gpu_join.h"
#include "params.h"

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static const char *dtype_name(int dt)
{
    switch (dt) {
    case BFJ_DTYPE_FP64: return "fp64";
    case BFJ_DTYPE_FP16: return "fp16";
    case BFJ_DTYPE_BF16: return "bf16";
    case BFJ_DTYPE_FP8: return "fp8";
    }
    return "none";
}

int main(int argc, char *argv[])
{
    if (argc < 4) {
        std::cerr << "usage: " << argv[0]
                  << " <file> <eps> <searchMode> [device] [--dim D] [--compute-dim C] [--binary]"
                     " [--counts out.txt] [--pairs out.txt] [--log file] [--eps-double]"
                     " [--reference-arithmetic [--compute-prec P] [--accum-prec P]]\n";
        return 1;
    }
    const char *filename = argv[FILENAME_ARG];
    // main.cu:20 parses eps into ACCUM_TYPE (float unless ACCUM_PREC is 64): the same
    // argument string gives the same eps as the reference, and the same eps as the
    // drop-in GPUJoinMainBruteForce, which takes an ACCUM_TYPE*.  --eps-double keeps
    // the full double instead.
    const double eps_arg = std::atof(argv[EPSILON_ARG]);
    double epsilon = (double)(ACCUM_TYPE)eps_arg;
    const unsigned int searchMode = (unsigned)std::atoi(argv[SEARCHMODE_ARG]);
    unsigned int device = 0;
    int dim = INPUT_DATA_DIM, computeDim = -1;
    bool binary = false, parse_only = false;
    std::string counts_path, pairs_path, log_path = "tensor_brute-force.txt";
    bool ref_arith = BFJ_REFERENCE_ARITHMETIC != 0;
    int computePrec = COMPUTE_PREC, accumPrec = ACCUM_PREC;
    int pos = 0;
    for (int a = 4; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--dim") && a + 1 < argc) dim = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--compute-dim") && a + 1 < argc) computeDim = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--binary")) binary = true;
        else if (!std::strcmp(argv[a], "--parse-only")) parse_only = true;
        else if (!std::strcmp(argv[a], "--counts") && a + 1 < argc) counts_path = argv[++a];
        else if (!std::strcmp(argv[a], "--pairs") && a + 1 < argc) pairs_path = argv[++a];
        else if (!std::strcmp(argv[a], "--log") && a + 1 < argc) log_path = argv[++a];
        else if (!std::strcmp(argv[a], "--reference-arithmetic")) ref_arith = true;
        else if (!std::strcmp(argv[a], "--eps-double")) epsilon = eps_arg;
        else if (!std::strcmp(argv[a], "--compute-prec") && a + 1 < argc) computePrec = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--accum-prec") && a + 1 < argc) accumPrec = std::atoi(argv[++a]);
        else if (pos++ == 0) device = (unsigned)std::atoi(argv[a]); // DEVICE_ARG
        else {
            std::cerr << "[Main] ~ Error: unknown argument " << argv[a] << "\n";
            return 1;
        }
    }

    // main.cu:62-108: modes 5 and 6 are errors before any work is done.
    if (searchMode == SM_NVIDIA) {
        std::cerr << "[Main] ~ Error: removed\n";
        return 1;
    }
    if (searchMode == SM_CPU) {
        std::cerr << "[Main] ~ Error: Work in Progress\n";
        return 1;
    }
    if (searchModeDtype(searchMode) < 0) {
        std::cerr << "[Main] ~ Error: Unknown search mode\n";
        return 1;
    }
    if (Q * C != 256 || Q % 8 != 0 || C % 8 != 0) { // main.cu:89-93
        std::cout << "Set Q & C values as (16 & 16) or (32 & 8) or (8 & 32)" << std::endl;
        return 1;
    }

    /***** Import dataset *****/
    const double tRead0 = now_s();
    std::vector<double> flat;
    int64_t n64 = 0, d64 = 0;
    if (binary) {
        if (importDatasetBinary(&flat, &n64, &d64, filename) != 0) {
            std::fprintf(stderr, "[Dataset] ~ Could not read the binary file %s\n", filename);
            return 1;
        }
        dim = (int)d64;
    } else {
        std::vector<std::vector<double>> inputVector;
        if (importDatasetDim(&inputVector, filename, dim) != 0) {
            std::fprintf(stderr, "[Dataset] ~ Could not open the input file %s\n", filename);
            return 1;
        }
        n64 = (int64_t)inputVector.size();
        flat.resize((size_t)n64 * dim);
        for (int64_t i = 0; i < n64; ++i)
            for (int k = 0; k < dim; ++k) flat[(size_t)i * dim + k] = inputVector[(size_t)i][(size_t)k];
    }
    const double timeRead = now_s() - tRead0;
    std::cout << "[Main | Time] ~ Time to read the dataset: " << timeRead << '\n';
    const unsigned int nbQueryPoints = (unsigned int)n64;
    if (parse_only) {
        // Rows as read, full precision, then stop before any device work.
        std::printf("[Main] ~ rows %lld dim %d\n", (long long)n64, dim);
        for (int64_t i = 0; i < n64; ++i)
            for (int k = 0; k < dim; ++k) std::printf(k + 1 < dim ? "%.17g " : "%.17g\n", flat[(size_t)i * dim + k]);
        return 0;
    }

    // main.cu:39-57: rows zero-padded to the compute stride, ADDITIONAL_POINTS zero rows appended.
    if (computeDim < dim) computeDim = (dim == INPUT_DATA_DIM && COMPUTE_DIM >= dim) ? COMPUTE_DIM : dim;
    std::vector<double> database((size_t)(nbQueryPoints + ADDITIONAL_POINTS) * computeDim, 0.0);
    for (size_t i = 0; i < nbQueryPoints; ++i)
        for (int k = 0; k < dim; ++k) database[i * computeDim + k] = flat[i * dim + k];

    /***** Compute the distance similarity join *****/
    std::vector<int64_t> counts(counts_path.empty() ? 0 : nbQueryPoints);
    bfj_stats stats;
    std::memset(&stats, 0, sizeof stats);
    uint64_t totalResult = 0;
    const double tJoin0 = now_s();
    const bool ref_mode = ref_arith && searchMode >= SM_GPU && searchMode <= SM_TENSOR_MQ;
    if (ref_mode && (!counts_path.empty() || !pairs_path.empty())) {
        std::cerr << "[Main] ~ Error: the reference's arithmetic gives a total only (no --counts / --pairs)\n";
        return 1;
    }
    // The reference passes eps as ACCUM_TYPE (gpu_join.cu:270); the library rounds it.
    const int rc = ref_mode ? GPUJoinMainBruteForceReference(searchMode, device, database.data(), nbQueryPoints,
                                                             (unsigned)dim, (unsigned)computeDim, epsilon, computePrec,
                                                             accumPrec, &totalResult)
                            : GPUJoinMainBruteForceEx(searchMode, device, database.data(), nbQueryPoints,
                                                      (unsigned)dim, (unsigned)computeDim, epsilon, &totalResult,
                                                      counts.empty() ? nullptr : counts.data(), &stats);
    const double timeJoin = now_s() - tJoin0;
    if (rc != BFJ_OK) return 1;
    // Accuracy of the pass's approximate distances (after the timed region): the max
    // |d2_approx - d2_fp64| over its candidate pairs, for the log.
    bfj_errors errs;
    std::memset(&errs, 0, sizeof errs);
    if (!ref_mode && GPUJoinPassErrors(device, 1u << 22, 256, 12345, &errs) != BFJ_OK) return 1;
    if (errs.violations) {
        std::cerr << "[Main] ~ Error: " << errs.violations << " pairs outside the error bound\n";
        return 1;
    }
    std::cout << "[Main | Result] ~ Time to join: " << timeJoin << '\n';
    std::cout << "[Main | Result] ~ Total result set size: " << totalResult << '\n';

    if (!counts_path.empty()) {
        std::ofstream cf(counts_path);
        for (int64_t c : counts) cf << c << '\n';
    }
    if (!pairs_path.empty()) {
        // The result set itself: "i j" per line, i < j, in no particular order.
        std::vector<uint32_t> pairs;
        uint64_t total = 0;
        if (GPUJoinMainBruteForcePairs(searchMode, device, database.data(), nbQueryPoints, (unsigned)dim,
                                       (unsigned)computeDim, epsilon, &pairs, &total, nullptr) != BFJ_OK)
            return 1;
        std::ofstream pf(pairs_path);
        for (size_t k = 0; k + 1 < pairs.size(); k += 2) pf << pairs[k] << ' ' << pairs[k + 1] << '\n';
        std::cout << "[Main | Result] ~ Pairs written: " << pairs.size() / 2 << '\n';
    }

    // main.cu:115-127: append one CSV row, header first if the file is empty.  The
    // reference's eleven columns come first; the rest are new.
    const double pairs = (double)nbQueryPoints * (double)nbQueryPoints;
    std::ifstream in(log_path);
    const bool empty = !in.good() || in.peek() == std::ifstream::traits_type::eof();
    in.close();
    std::ofstream out(log_path, std::ios::out | std::ios::app);
    if (empty)
        out << "Dataset, epsilon, searchMode, executionTime, totalNeighbors, inputDim, computeDim, blockSize, "
               "warpPerBlock, computePrec, accumPrec, dtype, pairsPerSec, bandPairs, maxAbsD2Err, tileRows, "
               "kernelMs\n";
    out << filename << ", " << epsilon << ", " << searchMode << ", " << timeJoin << ", " << totalResult << ", " << dim
        << ", " << computeDim << ", " << BLOCKSIZE << ", " << WARP_PER_BLOCK << ", " << computePrec << ", "
        << accumPrec << ", " << (ref_mode ? "reference" : dtype_name(stats.dtype)) << ", " << (timeJoin > 0 ? pairs / timeJoin : 0.0) << ", "
        << stats.band_pairs << ", " << errs.cand_max_abs_err << ", " << stats.tile_rows << ", " << stats.ms_total
        << std::endl;
    return 0;
}
