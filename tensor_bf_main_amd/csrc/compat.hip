// compat.hip -- the reference's own arithmetic for its four search modes.
//
// The exact paths (bfj_self_join*) answer the FP64 predicate sqrt(d2) <= eps for every
// dtype.  The reference instead counts with whatever COMPUTE_PREC / ACCUM_PREC it was
// built for (params.h:10-12, default FP16 compute and FP32 accumulation), so its
// totals depend on the search mode.  These kernels restate that arithmetic for callers
// that need the reference's totals bug for bug:
//
//   mode 1  distanceCalculationBruteForceCuda     kernel_join.cu:168-200
//           d = C(q_k - c_k); acc = A(acc + A(C(d * d)))                 k = 0 .. d-1
//   mode 2  distanceCalculationBruteForceCudaAlt  kernel_join.cu:203-239
//           acc = A(acc + A(A(A(C(C(-2 q_k) * c_k)) + A(q_k) * A(q_k)) + A(C(c_k * c_k))))
//   mode 3  distanceCalculationBruteForceTensorBasic  kernel_join.cu:245-325 (C = FP16, A = FP32)
//           per 16-column chunk: D_k = FP16(q_k - c_k) (the identity-matrix MMA with an FP16
//           accumulator), then acc = acc + sum_k D_k^2 (Gram diagonal, FP32 accumulator)
//   mode 4  distanceCalculationBruteForceTensorHalf   kernel_join.cu:328-524 (C = FP16, A = FP32)
//           per chunk: t = cn_j + sum_k C(-2 q_k) * c_k (MMA on the candidate chunk norm),
//           acc = (acc + t) + cn_i; chunk norms are FP32 sums of squares (kernel_join.cu:78-86)
//   hit     sqrt(acc) <= A(eps) (modes 2 and 4: sqrt(|acc|)), ordered pairs, self included
//
// When C and A are one type, nvcc's default contraction (-fmad=true) fuses a product
// into the sum that consumes it: mode 1 is acc = fma(t, t, acc) (for FP64, exactly the
// FP64 direct form of the exact paths) and mode 2 acc = acc + fma(c, c, fma(C(-2 q), c, q2)).
//
// C and A are the compute and accumulation types; every operation rounds to nearest
// even in its type, as IEEE binary16/32/64 arithmetic does on both vendors' ALUs
// (-ffp-contract=off keeps products and sums apart).  Modes 1 and 2 are therefore fully
// determined.  Modes 3 and 4 run on NVIDIA tensor cores, whose internal summation order
// and precision are not documented: here the 16 exact products of a chunk are summed in
// column order in FP32 and then added to the accumulator.  The sqrt test is done as
// acc <= T with T the largest A value whose square root is <= A(eps), which is the same
// predicate for a correctly rounded square root.
#include "bfj_internal.h"
#include <cmath>

#define BFJ_COMPAT_TJ 16 // candidate rows per LDS tile

// Rows in the compute type, padded with zeros to dpad columns (convertDataset,
// kernel_join.cu:44-54: a cast of each FP64 value, round to nearest even).
template <class CT>
__global__ void k_compat_convert(const double *x, int64_t n, int d, int64_t ld, int dpad, CT *out)
{
    const int64_t total = n * dpad;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / dpad;
        const int k = (int)(t - i * dpad);
        out[t] = k < d ? (CT)x[i * ld + k] : (CT)0;
    }
}

// FP32 squared norm of every 16-column chunk of every row: a sequential FP32 sum of
// the exact squares (preComputedSquaredCoordinates, kernel_join.cu:78-86).
__global__ void k_compat_chunknorm(const _Float16 *X, int64_t n, int dpad, float *cn)
{
    const int nch = dpad / 16;
    const int64_t total = n * nch;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const _Float16 *p = X + t * 16; // row t / nch, chunk t % nch: contiguous 16 values
        float s = 0.f;
        for (int k = 0; k < 16; ++k) s += (float)p[k] * (float)p[k];
        cn[t] = s;
    }
}

__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One thread per query row, candidates through LDS tiles of TJ rows.  Each candidate
// keeps its own accumulator, updated in ascending column order as in the reference.
template <int MODE, class CT, class AT>
__global__ __launch_bounds__(BFJ_BLOCK) void k_compat_join(const CT *X, const float *cn, int64_t n, int d, int dpad,
                                                           AT thr, unsigned long long *cnt)
{
    constexpr int TJ = BFJ_COMPAT_TJ;
    constexpr bool SAME = sizeof(CT) == sizeof(AT); // compute and accumulation in one type
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    CT *tile = reinterpret_cast<CT *>(lds_raw);                                  // [TJ][dpad]
    float *tcn = reinterpret_cast<float *>(lds_raw + (size_t)TJ * dpad * sizeof(CT)); // [TJ][dpad/16]
    const int nch = dpad / 16;
    const int64_t i = blockIdx.x * (int64_t)BFJ_BLOCK + threadIdx.x;
    const bool live = i < n;
    const CT *q = X + (live ? i : 0) * (int64_t)dpad;
    unsigned long long hits = 0;
    for (int64_t j0 = 0; j0 < n; j0 += TJ) {
        const int tj = (int)min((int64_t)TJ, n - j0);
        __syncthreads(); // the previous tile is no longer read
        for (int t = threadIdx.x; t < tj * dpad; t += BFJ_BLOCK) tile[t] = X[j0 * dpad + t];
        if (MODE == 4)
            for (int t = threadIdx.x; t < tj * nch; t += BFJ_BLOCK) tcn[t] = cn[j0 * nch + t];
        __syncthreads();
        AT acc[TJ];
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[j] = (AT)0;
        if (MODE == 1 || MODE == 2) {
            for (int k = 0; k < d; ++k) {
                const CT qk = q[k];
                const CT qm2 = (CT)(-2.0) * qk;
                const AT q2 = (AT)qk * (AT)qk;
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    const CT c = tile[j * dpad + k];
                    if constexpr (SAME) { // nvcc's default contraction fuses each product into its sum
                        if (MODE == 1) {
                            const CT t = qk - c;
                            acc[j] = fma_t(t, t, acc[j]);
                        } else {
                            acc[j] = acc[j] + fma_t((AT)c, (AT)c, fma_t((AT)qm2, (AT)c, q2));
                        }
                    } else {
                        if (MODE == 1) {
                            const CT t = qk - c;
                            const CT p = t * t;
                            acc[j] = acc[j] + (AT)p;
                        } else {
                            const CT pc = qm2 * c;
                            const CT cc = c * c;
                            acc[j] = acc[j] + (((AT)pc + q2) + (AT)cc);
                        }
                    }
                }
            }
        } else {
            for (int ch = 0; ch < nch; ++ch) {
                AT s[TJ];
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int k = ch * 16 + r;
                    const CT qk = q[k];
                    const CT qm2 = (CT)(-2.0) * qk;
#pragma unroll
                    for (int j = 0; j < TJ; ++j) {
                        const CT c = tile[j * dpad + k];
                        AT p;
                        if (MODE == 3) {
                            const CT dd = qk - c;
                            p = (AT)dd * (AT)dd;
                        } else {
                            p = (AT)qm2 * (AT)c;
                        }
                        s[j] = r == 0 ? p : s[j] + p;
                    }
                }
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    if (MODE == 3) acc[j] = acc[j] + s[j];
                    else acc[j] = (acc[j] + (tcn[j * nch + ch] + s[j])) + cn[(live ? i : 0) * nch + ch];
                }
            }
        }
        if (live) {
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                const AT v = (MODE == 2 || MODE == 4) ? (acc[j] < (AT)0 ? -acc[j] : acc[j]) : acc[j];
                if (j < tj && v <= thr) ++hits;
            }
        }
    }
    // Block sum, one atomic per block.
    __shared__ unsigned long long red[BFJ_BLOCK / BFJ_WAVE];
    for (int o = 32; o > 0; o >>= 1) hits += __shfl_xor(hits, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = hits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < BFJ_BLOCK / BFJ_WAVE; ++w) t += red[w];
        if (t) atomicAdd(cnt, t);
    }
}

// Largest value t of type T with sqrt(t) <= e (T's correctly rounded square root):
// sqrt(acc) <= e holds exactly when acc <= t.
template <class T>
static T sqrt_threshold(T e)
{
    if (!(e >= (T)0)) return (T)-1; // no acc >= 0 passes; NaN eps compares false
    if (e == (T)INFINITY) return (T)INFINITY;
    T t = e * e;
    while (std::sqrt(t) > e) t = std::nextafter(t, (T)-INFINITY);
    while (std::sqrt(std::nextafter(t, (T)INFINITY)) <= e && std::nextafter(t, (T)INFINITY) < (T)INFINITY)
        t = std::nextafter(t, (T)INFINITY);
    return t;
}

template <int MODE, class CT, class AT>
static hipError_t launch_mode(const CT *X, const float *cn, int64_t n, int d, int dpad, double eps,
                              unsigned long long *cnt, hipStream_t st)
{
    const AT thr = sqrt_threshold<AT>((AT)eps); // the reference passes eps as ACCUM_TYPE
    const size_t lds = (size_t)BFJ_COMPAT_TJ * dpad * sizeof(CT) + (size_t)BFJ_COMPAT_TJ * (dpad / 16) * sizeof(float);
    const unsigned blocks = (unsigned)((n + BFJ_BLOCK - 1) / BFJ_BLOCK);
    hipLaunchKernelGGL((k_compat_join<MODE, CT, AT>), dim3(blocks), dim3(BFJ_BLOCK), lds, st, X, cn, n, d, dpad, thr, cnt);
    return hipGetLastError();
}

template <class CT>
static hipError_t convert(const double *x, int64_t n, int d, int64_t ld, int dpad, void *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_compat_convert<CT>, dim3(2048), dim3(BFJ_BLOCK), 0, st, x, n, d, ld, dpad, (CT *)out);
    return hipGetLastError();
}

// Converted rows (X, n x dpad of the compute type), chunk norms (cn, n x dpad/16
// floats, mode 4 only) and the hit counter are device buffers of the caller.
// Returns 1 for an unsupported (mode, compute, accum) combination.
int bfj_compat_run(const double *x, int64_t n, int d, int64_t ld, double eps, int mode, int cprec, int aprec, void *X,
                   float *cn, unsigned long long *cnt, hipStream_t st)
{
    const int dpad = (d + 15) / 16 * 16;
    hipError_t e = hipSuccess;
    if (n == 0) return 0;
    if (cprec == 16) e = convert<_Float16>(x, n, d, ld, dpad, X, st);
    else if (cprec == 32) e = convert<float>(x, n, d, ld, dpad, X, st);
    else e = convert<double>(x, n, d, ld, dpad, X, st);
    if (e != hipSuccess) return 2;
    if (mode == 4) {
        hipLaunchKernelGGL(k_compat_chunknorm, dim3(1024), dim3(BFJ_BLOCK), 0, st, (const _Float16 *)X, n, dpad, cn);
        if (hipGetLastError() != hipSuccess) return 2;
    }
#define BFJ_M(M, CP, AP, CT, AT) \
    if (mode == M && cprec == CP && aprec == AP) e = launch_mode<M, CT, AT>((const CT *)X, cn, n, d, dpad, eps, cnt, st); else
    BFJ_M(1, 16, 32, _Float16, float)
    BFJ_M(1, 16, 64, _Float16, double)
    BFJ_M(1, 32, 32, float, float)
    BFJ_M(1, 32, 64, float, double)
    BFJ_M(1, 64, 32, double, float)
    BFJ_M(1, 64, 64, double, double)
    BFJ_M(2, 16, 32, _Float16, float)
    BFJ_M(2, 16, 64, _Float16, double)
    BFJ_M(2, 32, 32, float, float)
    BFJ_M(2, 32, 64, float, double)
    BFJ_M(2, 64, 32, double, float)
    BFJ_M(2, 64, 64, double, double)
    BFJ_M(3, 16, 32, _Float16, float)
    BFJ_M(4, 16, 32, _Float16, float)
    return 1;
#undef BFJ_M
    return e == hipSuccess ? 0 : 2;
}
