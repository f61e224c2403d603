// gpu_join.cpp -- the reference's host entry point (gpu_join.cu:210-322) on top of
// the C ABI.  No HIP header is included: everything device-side is behind bfjoin.h.
#include "gpu_join.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace {

// One library context per device for the process (the reference re-selects the
// device and re-allocates on every call, gpu_join.cu:218-267).
std::mutex g_mu;
std::vector<bfj_ctx *> g_ctx;

bfj_ctx *context_for(unsigned int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx.size() <= device) g_ctx.resize(device + 1, nullptr);
    if (!g_ctx[device]) {
        bfj_ctx *c = nullptr;
        if (bfj_init((int)device, &c) != BFJ_OK) {
            std::fprintf(stderr, "[GPU] ~ Error: %s\n", bfj_last_error());
            return nullptr;
        }
        g_ctx[device] = c;
    }
    return g_ctx[device];
}

} // namespace

int searchModeDtype(unsigned int searchMode)
{
    switch (searchMode) {
    case SM_GPU:
    case SM_CUDA_ALT:
    case SM_MFMA_FP64:
        return BFJ_DTYPE_FP64;
    case SM_TENSOR:
    case SM_MFMA_FP16:
        return BFJ_DTYPE_FP16;
    case SM_TENSOR_MQ:
        return COMPUTE_PREC == 64 ? BFJ_DTYPE_FP64 : BFJ_DTYPE_FP16;
    case SM_MFMA_BF16:
        return BFJ_DTYPE_BF16;
    case SM_MFMA_FP8:
        return BFJ_DTYPE_FP8;
    default:
        return -1;
    }
}

int GPUJoinMainBruteForceEx(unsigned int searchMode, unsigned int device, const double *dataset, unsigned int n,
                            unsigned int dim, unsigned int computeDim, double epsilon, uint64_t *totalNeighbors,
                            int64_t *counts, bfj_stats *stats)
{
    const int dtype = searchModeDtype(searchMode);
    if (dtype < 0) {
        std::fprintf(stderr, "[GPU] ~ Error: Unknown search mode\n");
        return BFJ_ERR_INVALID;
    }
    bfj_ctx *ctx = context_for(device);
    if (!ctx) return BFJ_ERR_DEVICE;
    uint64_t total = 0;
    const int rc = bfj_self_join(ctx, dataset, (int64_t)n, (int32_t)dim, (int64_t)computeDim, &epsilon, 1, dtype, 0,
                                 (int64_t)n, counts, &total, stats);
    if (rc != BFJ_OK) {
        std::fprintf(stderr, "[GPU] ~ Error: %s\n", bfj_last_error());
        return rc;
    }
    if (totalNeighbors) *totalNeighbors = total;
    return BFJ_OK;
}

int GPUJoinMainBruteForcePairs(unsigned int searchMode, unsigned int device, const double *dataset, unsigned int n,
                               unsigned int dim, unsigned int computeDim, double epsilon,
                               std::vector<uint32_t> *pairs, uint64_t *totalNeighbors, bfj_stats *stats)
{
    const int dtype = searchModeDtype(searchMode);
    if (dtype < 0) {
        std::fprintf(stderr, "[GPU] ~ Error: Unknown search mode\n");
        return BFJ_ERR_INVALID;
    }
    bfj_ctx *ctx = context_for(device);
    if (!ctx) return BFJ_ERR_DEVICE;
    // Start at 16 pairs per point; the library reports the size it needs when that
    // is not enough, and the call is repeated once with it.
    uint64_t cap = std::max<uint64_t>(16ull * n, 1024), need = 0, total = 0;
    int rc = BFJ_OK;
    for (int attempt = 0; attempt < 2; ++attempt) {
        pairs->assign(2 * cap, 0u);
        rc = bfj_self_join_pairs(ctx, dataset, (int64_t)n, (int32_t)dim, (int64_t)computeDim, epsilon, dtype,
                                 pairs->data(), cap, &need, &total, stats);
        if (rc != BFJ_ERR_CAPACITY) break;
        cap = need;
    }
    if (rc != BFJ_OK) {
        std::fprintf(stderr, "[GPU] ~ Error: %s\n", bfj_last_error());
        pairs->clear();
        return rc;
    }
    pairs->resize(2 * need);
    if (totalNeighbors) *totalNeighbors = total;
    return BFJ_OK;
}

int GPUJoinMainBruteForceReference(unsigned int searchMode, unsigned int device, const double *dataset, unsigned int n,
                                   unsigned int dim, unsigned int computeDim, double epsilon, int computePrec,
                                   int accumPrec, uint64_t *totalNeighbors)
{
    bfj_ctx *ctx = context_for(device);
    if (!ctx) return BFJ_ERR_DEVICE;
    uint64_t total = 0;
    const int rc = bfj_reference_mode_join(ctx, dataset, (int64_t)n, (int32_t)dim, (int64_t)computeDim, epsilon,
                                           (int32_t)searchMode, computePrec, accumPrec, &total);
    if (rc != BFJ_OK) {
        std::fprintf(stderr, "[GPU] ~ Error: %s\n", bfj_last_error());
        return rc;
    }
    if (totalNeighbors) *totalNeighbors = total;
    return BFJ_OK;
}

int GPUJoinPassErrors(unsigned int device, uint64_t maxCandidates, int64_t sampleTiles, uint64_t seed,
                      bfj_errors *errors)
{
    bfj_ctx *ctx = context_for(device);
    if (!ctx) return BFJ_ERR_DEVICE;
    const int rc = bfj_pass_errors(ctx, maxCandidates, sampleTiles, seed, errors);
    if (rc != BFJ_OK) std::fprintf(stderr, "[GPU] ~ Error: %s\n", bfj_last_error());
    return rc;
}

void GPUJoinMainBruteForce(unsigned int searchMode, unsigned int device, INPUT_DATA_TYPE *dataset,
                           unsigned int *nbQueryPoints, ACCUM_TYPE *epsilon, uint64_t *totalNeighbors)
{
    if (searchModeDtype(searchMode) < 0) {
        std::fprintf(stderr, "[GPU] ~ Error: Unknown search mode\n"); // gpu_join.cu:311-315
        std::exit(1);
    }
    *totalNeighbors = 0;
    if (BFJ_REFERENCE_ARITHMETIC && searchMode >= SM_GPU && searchMode <= SM_TENSOR_MQ) {
        (void)GPUJoinMainBruteForceReference(searchMode, device, dataset, *nbQueryPoints, INPUT_DATA_DIM, COMPUTE_DIM,
                                             (double)*epsilon, COMPUTE_PREC, ACCUM_PREC, totalNeighbors);
        return;
    }
    (void)GPUJoinMainBruteForceEx(searchMode, device, dataset, *nbQueryPoints, INPUT_DATA_DIM, COMPUTE_DIM,
                                  (double)*epsilon, totalNeighbors, nullptr, nullptr);
}

void GPUJoinMainBruteForceNvidia(unsigned int, unsigned int, INPUT_DATA_TYPE *, INPUT_DATA_TYPE *, unsigned int *,
                                 ACCUM_TYPE *, uint64_t *totalNeighbors)
{
    std::fprintf(stderr, "[GPU] ~ Error: removed\n");
    if (totalNeighbors) *totalNeighbors = 0;
}
