// refine.hip -- FP64 band refine and count finalisation.
//
// Every pair the join kernel could not decide is re-evaluated here in the
// reference's own arithmetic: the direct form of distanceCalculationBruteForceCuda
// (kernel_join.cu:184-189), t = q[k] - c[k]; acc = fma(t, t, acc) in ascending k,
// compared with T, the largest double whose square root is <= eps, which is
// the predicate sqrt(acc) <= eps of kernel_join.cu:194.  These are the only
// FP64 steps of the low-precision paths besides the threshold computation.
#include "bfj_internal.h"
#include <math.h>

__global__ __launch_bounds__(BFJ_BLOCK) void k_refine(PassArgs a)
{
    const uint32_t cnt = *a.band_count;
    const uint32_t m = cnt < a.band_capacity ? cnt : a.band_capacity;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < m; t += stride) {
        const BandEntry be = a.band[t];
        const int si = (int)(be.meta >> 16);
        const uint32_t mask = be.meta & 0xffu;
        const SetDesc &s = a.sets[si];
        const double *q = s.x + (int64_t)be.i * s.ld;
        const double *c = s.x + (int64_t)be.j * s.ld;
        double acc = 0.0;
        for (int k = 0; k < s.d; ++k) {
            const double d = q[k] - c[k];
            acc = fma(d, d, acc);
        }
        for (int e = 0; e < s.n_eps; ++e)
            if (((mask >> e) & 1u) && acc <= s.T[e]) atomicAdd(&a.cnt[s.cnt_off + (int64_t)e * s.npad + be.i], 1u);
        const double sigma = a.scal[si].sigma;
        const double approx = sigma * sigma * (a.hmax[s.row_off + be.i] - 2.0 * (double)be.z);
        const double err = fabs(approx - acc);
        if (err >= 0.0) atomicMax(&a.scal[si].err_bits, (unsigned long long)__double_as_longlong(err));
    }
}

void bfj_launch_refine(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_refine, dim3(2048), dim3(BFJ_BLOCK), 0, st, a);
}

// Per-row counts (self excluded) into the caller's int64 array and the
// reference's self-inclusive totals (gpu_join.cu:320-321).
__global__ void k_finalize(PassArgs a)
{
    __shared__ unsigned long long red[BFJ_MAXEPS][BFJ_PREP_ROWS];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const int64_t i = a.blk_row0[b] + threadIdx.x;
    const bool query = i < s.n && i >= s.row_begin && i < s.row_end;
    const int64_t rows = s.row_end - s.row_begin;
    for (int e = 0; e < s.n_eps; ++e) {
        unsigned long long c = 0;
        if (query) {
            c = a.cnt[s.cnt_off + (int64_t)e * s.npad + i];
            const long long self = (s.T[e] >= 0.0) ? 1 : 0; // d2(i, i) == 0 exactly
            if (s.out) s.out[(int64_t)e * rows + (i - s.row_begin)] = (long long)c - self;
        }
        red[e][threadIdx.x] = c;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)s.n_eps) {
        unsigned long long t = 0;
        for (int r = 0; r < BFJ_PREP_ROWS; ++r) t += red[threadIdx.x][r];
        if (t) atomicAdd(&a.scal[si].total[threadIdx.x], t);
    }
}

void bfj_launch_finalize(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(a.n_row_blocks), dim3(BFJ_PREP_ROWS), 0, st, a);
}
