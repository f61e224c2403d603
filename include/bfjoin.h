/*
 * bfjoin.h -- C ABI of the MI355X brute-force Euclidean epsilon self-join.
 *
 * This is the drop-in boundary under the reference's host driver.  The
 * reference (jeevana-swaroop/tensor_bf_main) has one host entry point,
 *     void GPUJoinMainBruteForce(unsigned searchMode, unsigned device,
 *                                INPUT_DATA_TYPE *dataset, unsigned *nbQueryPoints,
 *                                ACCUM_TYPE *epsilon, uint64_t *totalNeighbors);
 * (gpu_join.h:36-42, body gpu_join.cu:210-322), which allocates device memory,
 * copies the padded dataset in, runs the prep kernels (convertDataset,
 * kernel_join.cu:44-54; preComputedSquaredCoordinates, kernel_join.cu:57-88),
 * launches one join kernel (kernel_join.cu:168-677) and copies back one uint64.
 * The functions below replace that body; include/gpu_join.h keeps the
 * reference signature on top of them.
 *
 * Conventions: every function returns a bfj_status; on failure the reason is
 * available from bfj_last_error() (thread-local).  No C++ exception crosses
 * this boundary.  Plain pointers and sizes only.
 *
 * Semantics (kernel_join.cu:184-198, gpu_join.cu:318-321): for every ordered
 * pair (i, j) of rows, i in [row_begin, row_end), j in [0, n), the pair is a hit
 * when sqrt(d2(i, j)) <= eps, where d2 is the FP64 direct form
 * sum_k fma-accumulated (x_i[k] - x_j[k])^2 in ascending k.  counts[] exclude
 * the self pair; totals[] include it (the reference's totalNeighbors).  Every
 * dtype returns exactly these counts: low-precision dot products only decide
 * the pairs whose error band excludes the threshold; the rest are re-evaluated
 * in FP64 on the GPU.
 */
#ifndef BFJOIN_H
#define BFJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: bfj_stats gained fp64_pairs.  3: bfj_stats lost max_abs_d2_err (see bfj_pass_errors)
   and gained candidates; bfj_set_capacity, bfj_fail_allocation, bfj_pass_errors,
   bfj_plan_tiles and bfj_self_join_device_noalloc.  4: bfj_stats gained slices;
   bfj_set_stream_limit.  5: candidate chunks of BFJ_CHUNK_RECORDS = 4096 records
   (bfj_set_capacity counted chunks of 1024 before).  (Another struct layout or unit:
   callers built against an earlier version must be rebuilt.) */
#define BFJ_ABI_VERSION 5

typedef struct bfj_ctx bfj_ctx;
typedef struct bfj_plan bfj_plan;

typedef enum bfj_status {
    BFJ_OK = 0,
    BFJ_ERR_INVALID = 1,     /* bad argument (null pointer, n < 0, d < 1, ld < d, NaN eps, ...) */
    BFJ_ERR_DEVICE = 2,      /* HIP runtime error or bad device index */
    BFJ_ERR_NOMEM = 3,       /* device allocation failed */
    BFJ_ERR_UNSUPPORTED = 4, /* dimension or dtype outside what the kernels are built for */
    BFJ_ERR_INTERNAL = 5,
    BFJ_ERR_CAPACITY = 6     /* output buffer too small; the call reports the size it needs */
} bfj_status;

/* Arithmetic of the dot products.  The band refine is FP64 in every case. */
typedef enum bfj_dtype {
    BFJ_DTYPE_FP64 = 0, /* FP64 MFMA (v_mfma_f64_16x16x4), FP64 norms */
    BFJ_DTYPE_FP16 = 1, /* FP16 MFMA (v_mfma_f32_16x16x32_f16), FP32 accumulation and norms */
    BFJ_DTYPE_BF16 = 2, /* BF16 MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulation and norms */
    BFJ_DTYPE_FP8 = 3   /* FP8 e4m3, per-row power-of-two scales (v_mfma_scale_f32_16x16x128_f8f6f4) */
} bfj_dtype;

#define BFJ_MAX_EPS 8   /* epsilon values per dataset in one pass */
#define BFJ_MAX_DIM 256 /* largest d the MFMA kernels are built for */
#define BFJ_CHUNK_RECORDS 4096 /* candidate records (16 bytes each) per chunk of the candidate stream */

typedef struct bfj_stats {
    uint64_t pairs;          /* ordered pairs evaluated: (row_end - row_begin) * n, summed over datasets */
    uint64_t band_pairs;     /* pairs in the join's error band for some eps (union over eps) */
    uint64_t band_capacity;  /* refine list capacity used for the pass */
    int32_t band_retries;    /* passes repeated because the refine list overflowed */
    int32_t tile_rows;       /* query rows held per wavefront in the chosen tile */
    int32_t k_padded;        /* dot-product length after zero padding */
    int32_t dtype;
    double max_band_halfwidth; /* widest (H - L)/2 over rows and eps, input units */
    double ms_prep;          /* device time: centre, cast, norms, thresholds */
    double ms_join;          /* device time: MFMA join with fused epsilon epilogue */
    double ms_refine;        /* device time: band tests, middle stage, FP64 refine, count finalisation */
    double ms_total;         /* device time of the whole pass */
    uint64_t fp64_pairs;     /* band pairs re-evaluated in FP64: those the middle stage (FP16/BF16/FP8
                                passes) left undecided; band_pairs for FP64 passes */
    uint64_t candidates;     /* pair records the join streamed to the band tests (the pairs it
                                could not reject) */
    int32_t slices;          /* pieces the join's work ran in (1: one pass; more when the
                                candidate stream exceeded the stream limit) */
    int32_t reserved;
} bfj_stats;

/* One dataset of a batch.  data/counts are host or device pointers depending on
 * the call; totals is always host memory. */
typedef struct bfj_dataset {
    const double *data; /* n x ld row-major, first d columns used */
    int64_t n;
    int32_t d;
    int64_t ld;
    int32_t n_eps;
    const double *eps;  /* [n_eps] host */
    int64_t *counts;    /* [n_eps][n], self excluded; may be NULL */
    uint64_t *totals;   /* [n_eps], self included; may be NULL */
} bfj_dataset;

/* Accuracy of a pass's approximate squared distances (bfj_pass_errors). */
typedef struct bfj_errors {
    uint64_t cand_total;      /* candidate records the pass streamed */
    uint64_t cand_pairs;      /* of those, checked here */
    uint64_t tile_pairs;      /* pairs checked in the sampled tiles */
    uint64_t violations;      /* checked pairs whose FP64 distance lies outside the interval the
                                 band tests assumed for them: 0 when the bound holds */
    double cand_max_abs_err;  /* max |S_hat - d2| over the candidates checked, input units */
    double tile_max_abs_err;  /* max |S_hat - d2| over the sampled tiles, input units */
    double max_acc_ratio;     /* sampled tiles, FP16/BF16/FP8 passes: max |S_hat - S_cast| / E_i,
                                 S_cast the exact squared distance of the cast rows and E_i the
                                 accumulation bound (<= 1 when it holds); -1 for FP64 passes */
    double max_quant_ratio;   /* sampled tiles, FP16/BF16/FP8 passes: max |dist_cast - dist| /
                                 (q_i + q_j), the quantisation radius (<= 1); -1 for FP64 passes */
} bfj_errors;

const char *bfj_last_error(void);
int bfj_abi_version(void);
int bfj_device_count(int *count);

/* Context: one per host thread; owns a HIP stream and all device buffers
 * (the reference leaks them, gpu_join.cu:221-267). */
int bfj_init(int device, bfj_ctx **ctx);
int bfj_destroy(bfj_ctx *ctx);

/* Replaces gpu_join.cu:210-322 for one dataset held in host memory.
 * data: n x ld doubles (host).  counts: [n_eps][row_end-row_begin] (host, may be NULL).
 * totals: [n_eps] (host, may be NULL).  row_begin/row_end select the query rows
 * (the shard of one GPU); candidates are always all n rows. */
int bfj_self_join(bfj_ctx *ctx, const double *data, int64_t n, int32_t d, int64_t ld,
                  const double *eps, int32_t n_eps, int32_t dtype,
                  int64_t row_begin, int64_t row_end,
                  int64_t *counts, uint64_t *totals, bfj_stats *stats);

/* Same with device-resident data and counts (counts: device int64 [n_eps][rows]).
 * Runs on the context stream and returns after it has synchronised. */
int bfj_self_join_device(bfj_ctx *ctx, const double *d_data, int64_t n, int32_t d, int64_t ld,
                         const double *eps, int32_t n_eps, int32_t dtype,
                         int64_t row_begin, int64_t row_end,
                         int64_t *d_counts, uint64_t *totals, bfj_stats *stats);

/* bfj_self_join_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the call
 * (code that runs under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and status codes as bfj_self_join_device; the
 * pass's working buffers (packed operands, norms, candidate stream, band lists,
 * counters) come from one file-scope device array of BFJ_NOALLOC_SCRATCH_BYTES that
 * each GPU holds until the process ends, not from hipMalloc.  Limits: the dataset
 * must fit that array (about 1.5 KiB per row at d = 128 FP16 besides the candidate
 * stream and band lists, which take the rest: n up to roughly 16k rows; a larger
 * dataset returns BFJ_ERR_CAPACITY).  A pass whose candidate stream or band list
 * does not fit runs in slices within the array (bfj_set_stream_limit) instead of
 * growing them.  Calls on one GPU share its array, so the library runs them one
 * at a time (a per-device lock); bfj_pass_errors does not apply to them. */
#define BFJ_NOALLOC_SCRATCH_BYTES (64ull << 20)
int bfj_self_join_device_noalloc(bfj_ctx *ctx, const double *d_data, int64_t n, int32_t d, int64_t ld,
                                 const double *eps, int32_t n_eps, int32_t dtype,
                                 int64_t row_begin, int64_t row_end,
                                 int64_t *d_counts, uint64_t *totals, bfj_stats *stats);

/* The result set instead of counts: every unordered pair {i, j}, i < j, of the n
 * rows with sqrt(d2(i, j)) <= eps (the reference's predicate, kernel_join.cu:194),
 * written to pairs[2k], pairs[2k+1] (host, i < j) in no particular order.  The
 * reference only counts pairs (kernel_join.cu:184, gpu_join.cu:321); its ordered
 * self-inclusive total equals n + 2 * *n_pairs (eps >= 0).  *n_pairs is always set;
 * when it exceeds capacity the call returns BFJ_ERR_CAPACITY and the caller
 * retries with a buffer of that many pairs.  totals (may be NULL) as in
 * bfj_self_join with n_eps = 1. */
int bfj_self_join_pairs(bfj_ctx *ctx, const double *data, int64_t n, int32_t d, int64_t ld,
                        double eps, int32_t dtype, uint32_t *pairs, uint64_t capacity,
                        uint64_t *n_pairs, uint64_t *total, bfj_stats *stats);

/* The reference's own arithmetic instead of the exact predicate, for bug-for-bug
 * totals: search_mode 1 (distanceCalculationBruteForceCuda, kernel_join.cu:168-200),
 * 2 (...CudaAlt, kernel_join.cu:203-239), 3 (...TensorBasic, kernel_join.cu:245-325)
 * or 4 (...TensorHalf, kernel_join.cu:328-524), computed with the reference's
 * COMPUTE_PREC (compute_prec: 16, 32 or 64) and ACCUM_PREC (accum_prec: 32 or 64),
 * params.h:10-12.  Modes 3 and 4 exist for 16 / 32 only.  *total gets the ordered,
 * self-inclusive count of the reference (gpu_join.cu:318-321).  Modes 3 and 4 sum each
 * 16-column chunk's products in column order in FP32, an assumption about the tensor
 * cores the reference ran on.  d <= 256. */
int bfj_reference_mode_join(bfj_ctx *ctx, const double *data, int64_t n, int32_t d, int64_t ld,
                            double eps, int32_t search_mode, int32_t compute_prec, int32_t accum_prec,
                            uint64_t *total);

/* Many independent datasets, each with its own eps list, in one persistent launch.
 * device_resident != 0: sets[i].data / counts are device pointers. */
int bfj_self_join_batch(bfj_ctx *ctx, const bfj_dataset *sets, int32_t n_sets, int32_t dtype,
                        int32_t device_resident, bfj_stats *stats);

/* Plans: all buffers allocated, join shapes tuned (bfj_plan_tiles) and the whole pass
 * captured in one hipGraph at creation; bfj_plan_run replays it.  Datasets are
 * device-resident. */
int bfj_plan_create(bfj_ctx *ctx, const bfj_dataset *sets, int32_t n_sets, int32_t dtype, bfj_plan **plan);
/* The join shape a plan runs with, per K-chunk group (datasets whose d pads to the same
 * number of K-chunks): kc, rows per wavefront, resident blocks per CU and candidate
 * panel bytes.  bfj_plan_create picks them for every group that has no shape set with
 * bfj_set_tile_rows: it times whole passes over a seeded sample of the group's own
 * datasets at their own eps, for every built row count x blocks per CU, then two
 * other panel sizes.  Writes up to max_groups entries; *n_groups gets the count. */
int bfj_plan_tiles(const bfj_plan *plan, int32_t *kc, int32_t *rows, int32_t *blocks_per_cu, int32_t *panel_bytes,
                   int32_t max_groups, int32_t *n_groups);
int bfj_plan_run(bfj_plan *plan, bfj_stats *stats);
int bfj_plan_destroy(bfj_plan *plan);

/* Device buffers on the context's device (for callers without a GPU allocator
 * of their own; any device pointer from the same HIP runtime works too).
 * Copies run on the context stream and return when complete. */
int bfj_buffer_alloc(bfj_ctx *ctx, uint64_t bytes, void **d_ptr);
int bfj_buffer_free(bfj_ctx *ctx, void *d_ptr);
int bfj_buffer_upload(bfj_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes);
int bfj_buffer_download(bfj_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes);

/* Tile choice.  bfj_autotune times every built tile shape for (dtype, d) on a
 * synthetic problem of n points and keeps the fastest in the context; later
 * joins use it.  bfj_set_tile_rows forces a shape (0 = heuristic). */
int bfj_autotune(bfj_ctx *ctx, int32_t dtype, int32_t d, int64_t n, int32_t *tile_rows_out);
int bfj_set_tile_rows(bfj_ctx *ctx, int32_t dtype, int32_t tile_rows);
/* Locality order (no counterpart in the reference; default on): a pass whose rows
 * are all queries packs them grouped by nearest of 128 pivot rows, so the pairs near
 * eps fall into few tiles.  Results do not depend on it; enable = 0 keeps input order. */
int bfj_set_reorder(bfj_ctx *ctx, int32_t enable);

/* Capacities of the candidate stream (chunks of BFJ_CHUNK_RECORDS records) and of the band lists
 * (records) for the next pass; 0 leaves one as it is.  No counterpart in the
 * reference, which has neither.  A pass that needs more reports it, and the library
 * grows the buffer and repeats the pass (stats band_retries): results never depend
 * on these values, only the number of passes does.  Used to pre-size the buffers
 * for a known workload, or to force the grow-and-repeat path in tests. */
int bfj_set_capacity(bfj_ctx *ctx, uint64_t candidate_chunks, uint64_t band_records);

/* Device memory the candidate stream and the band lists of a pass may take, in bytes
 * (0, the default: half of the device's memory).  No counterpart in the reference,
 * which streams no candidates.  A pass whose stream would need more runs the join's
 * work items in slices, each followed by its band tests, middle stage and FP64
 * refine; the counts are the same (stats slices > 1).  A no-allocation call slices
 * within its fixed buffers the same way.  Sliced passes synchronise with the host
 * after every slice and are not captured: a plan that needed slices runs them
 * eagerly from then on.  A limit too small for one work item's candidates fails with
 * BFJ_ERR_CAPACITY.  A growth of the stream that the device cannot allocate (other
 * allocations hold its memory) also runs the pass in slices, under half the bytes that
 * failed, until this function is called again. */
int bfj_set_stream_limit(bfj_ctx *ctx, uint64_t bytes);

/* Fault injection for tests of the error paths (no counterpart in the reference):
 * the countdown-th device allocation the library makes from now on (over every context
 * of the process) fails as if the device were out of memory, and the call that made it
 * returns BFJ_ERR_NOMEM; 0 turns it off.  A context stays usable after such a failure. */
int bfj_fail_allocation(int64_t countdown);

/* Accuracy of the last pass run on this context (no counterpart in the reference,
 * which reports no error): sigma^2 S_hat, the squared distance the pass's dot
 * products gave, against the FP64 direct form d2 of kernel_join.cu:184-189, over
 *   (a) the pass's candidate records (the pairs it could not reject: near or inside
 *       eps), all of them, or every k-th chunk so that about max_candidates are
 *       checked (0: all), and
 *   (b) sample_tiles seeded 16 x 16 tiles of all pairs, spread over the datasets by
 *       their pair counts and recomputed with the pass's own MFMA sequence.
 * Every checked pair is also tested against the interval the band tests assumed.
 * After a pass that ran in slices, (a) covers the records of its last slice.
 * Runs after the pass, outside its timing; it needs the pass's buffers and rows,
 * so call it before any other call on the context and before freeing the rows. */
int bfj_pass_errors(bfj_ctx *ctx, uint64_t max_candidates, int64_t sample_tiles, uint64_t seed, bfj_errors *out);
int bfj_tile_rows_options(int32_t dtype, int32_t d, int32_t *options, int32_t max_options, int32_t *n_options);

#ifdef __cplusplus
}
#endif

#endif /* BFJOIN_H */
