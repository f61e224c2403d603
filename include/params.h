/*
 * params.h -- compile-time defaults of the reference's host API (params.h of
 * jeevana-swaroop/tensor_bf_main), kept under the same names so its driver code
 * compiles unchanged against this library.  Every value can be overridden with
 * -D; the library itself takes n, d, the dtype and the eps list at run time.
 *
 * Left out on purpose: the single-letter GEMM-sample macros M, N, K
 * (params.h:41-50) and the WMMA/DMMA tiling constants of the two unbuilt NVIDIA
 * GEMM samples (params.h:52-142).  The join kernels here never include this file.
 */
#ifndef BFJ_PARAMS_H
#define BFJ_PARAMS_H

#ifndef INPUT_DATA_DIM
#define INPUT_DATA_DIM 90 /* params.h:4 */
#endif
#ifndef COMPUTE_DIM
#define COMPUTE_DIM 96 /* params.h:5: rows are zero-padded to this stride */
#endif

#ifndef BLOCKSIZE
#define BLOCKSIZE 256 /* params.h:7 (reported in the CSV log only) */
#endif
#ifndef WARP_PER_BLOCK
#define WARP_PER_BLOCK 4 /* params.h:8 (reported in the CSV log only) */
#endif

#ifndef INPUT_DATA_PREC
#define INPUT_DATA_PREC 64 /* params.h:10 */
#endif
#ifndef COMPUTE_PREC
#define COMPUTE_PREC 16 /* params.h:11: 16 selects the FP16-MFMA path for searchMode 4, 64 the FP64 one */
#endif
#ifndef ACCUM_PREC
#define ACCUM_PREC 32 /* params.h:12: type of the epsilon argument */
#endif

/* params.h:14-15: queries x candidates per warp in the reference's MQ kernel;
 * main.cu:89-93 checks Q*C == 256.  Kept for that check only. */
#ifndef Q
#define Q 16
#endif
#ifndef C
#define C 16
#endif

/* 0: searchModes 1-4 return the exact FP64 answer.  1: they count with the
 * reference's own COMPUTE_PREC / ACCUM_PREC arithmetic instead (bug-for-bug totals,
 * bfj_reference_mode_join).  Not in the reference, which only has the latter. */
#ifndef BFJ_REFERENCE_ARITHMETIC
#define BFJ_REFERENCE_ARITHMETIC 0
#endif

#ifndef ADDITIONAL_POINTS
#define ADDITIONAL_POINTS 15 /* params.h:20: zero rows main.cu appends after the data */
#endif

/* params.h:160-164: positional CLI arguments */
#define NB_ARGS_MAX 5
#define FILENAME_ARG 1
#define EPSILON_ARG 2
#define SEARCHMODE_ARG 3
#define DEVICE_ARG 4

/* params.h:166-171: search modes.  Every mode returns the exact FP64 answer unless
 * BFJ_REFERENCE_ARITHMETIC is set. */
#define SM_GPU 1       /* reference: CUDA-core direct form     -> here FP64 MFMA + FP64 band refine */
#define SM_CUDA_ALT 2  /* reference: CUDA-core expanded form   -> here FP64 MFMA + FP64 band refine */
#define SM_TENSOR 3    /* reference: WMMA identity trick, half -> here FP16 MFMA + FP64 band refine */
#define SM_TENSOR_MQ 4 /* reference: WMMA MQ (FP16 or DMMA FP64 by COMPUTE_PREC) -> FP16 / FP64 MFMA */
#define SM_NVIDIA 5    /* reference: removed (main.cu:64-83); still an error */
#define SM_CPU 6       /* reference: "Work in Progress" (main.cu:97-102); still an error */
/* New modes: the dot-product dtype chosen explicitly. */
#define SM_MFMA_FP64 7
#define SM_MFMA_FP16 8
#define SM_MFMA_BF16 9
#define SM_MFMA_FP8 10

#define WARP_SIZE 64 /* a gfx950 wavefront (the reference's value is 32) */

/* params.h:174-203: element types.  The library's rows are always FP64. */
#if INPUT_DATA_PREC == 64
#define INPUT_DATA_TYPE double
#else
#error "the join reads FP64 rows: INPUT_DATA_PREC must be 64"
#endif

#if ACCUM_PREC == 64
#define ACCUM_TYPE double
#else
#define ACCUM_TYPE float
#endif

#endif /* BFJ_PARAMS_H */
