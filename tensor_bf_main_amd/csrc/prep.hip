// prep.hip -- per-pass preparation kernels.
//
// Replaces convertDataset (kernel_join.cu:44-54: cast FP64 rows to the compute
// type) and preComputedSquaredCoordinates (kernel_join.cu:57-88: squared norms
// per row chunk).  Here the rows are first centred on the column means (the
// distances do not change, the magnitudes that set every rounding error do),
// scaled by a power of two, cast, and written straight into MFMA fragment
// order.  For every row the kernel also measures, in FP64, how far the cast row
// lies from the exact centred row; those radii and the norms set the
// per-row error band that makes the low-precision decisions rigorous.
//
// Band derivation (kernel space; the join kernel computes
//   Z_ij = MFMA(C = g_i; xhat_i . xhat_j) - n_j  ~=  0.5 * (2 g_i + N_i - S_ij),
//   S_ij = ||xhat_i - xhat_j||^2):
//   * quantisation: | ||xhat_i - xhat_j|| - ||y_i - y_j|| | <= q_i + q_j,
//     q_i >= ||y_i - xhat_i|| measured per row, y the exact centred row.
//   * oracle rounding: the FP64 direct form differs from the exact squared
//     distance by at most rho = gamma_{d+2}(2^-53) relatively.
//   * accumulation: the MFMA forms acc = g_i + xhat_i . xhat_j (g_i is its C
//     operand) and Z = fl(acc - n_j), n_j = 0.5 fl32(N_j).  Then
//     |Z - (g_i - 0.5 N_j + xhat_i . xhat_j)| <= E_i / 2 with
//     E_i = 2 (1.01 K' + 1) u (|g_i| + sqrt(N_i N_max)) + 4 u N_max (+ FP64 norm
//     terms): the standard bound for K' = 2K + 8 round-to-nearest operations in any
//     order at unit roundoff u (2^-24 for FP32 accumulation, 2^-53 for FP64) on
//     partial sums within |g_i| + sum |p_k| <= |g_i| + sqrt(N_i N_max), plus the
//     final subtraction and the FP32 rounding of n_j.
//   So S_hat = (2 g_i + N_i) - 2 Z satisfies |S_hat - S_ij| <= E_i, and for eps e:
//     certainly within eps   if S_hat + E_i <= L_e = (sqrt(T_e/(1+rho)) - r)^2
//     certainly outside      if S_hat - E_i >  H_e = (sqrt(T_e/(1-rho)) + r)^2
//   with r = q_i + q_j (one bound for both rows of the pair).  g_i = 0.5 (Hmax_i +
//   E_i - N_i) rounded up, Hmax_i the largest H of r = q_i + q_max (>= H(r_ij) for
//   every j), makes Z < 0 certainly outside for every eps and both rows: one Z per
//   unordered pair decides for i and for j.  FP8 passes split Hmax into a row part
//   (in g_i) and a column part (in the join's candidate threshold for n_j);
//   k_thresholds has the details.
#include "bfj_internal.h"
#include <math.h>
#include <float.h>

__global__ void k_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

void bfj_launch_zero(void *p, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    size_t n16 = bytes / 16, ntail = bytes % 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)p, n16,
                       (unsigned char *)p + n16 * 16, ntail);
}

// Blocks of one dataset all raise the same word, and same-address atomics are performed
// one after another in memory; a block whose value the word already holds skips its
// atomic.  The word only grows, so a stale read can only lead to an unneeded atomic.
__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v)
{
    if (!(v >= 0.0)) return;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (__hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= b) return;
    atomicMax(addr, b);
}

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// ---- column means -----------------------------------------------------------
// Partial sums over 64-row blocks, then an in-order sum per dataset: the mean is
// the same bits on every run.
// The same pass keeps each block's column maxima and minima (colsum holds three
// [row block][BFJ_MAXDIM] arrays: sums, maxima, minima) for the scale.
__global__ void k_colsum(PassArgs a)
{
    // Threads as [row lane][column]: cw columns (d rounded up to a power of two) by
    // BFJ_BLOCK / cw row lanes, so narrow rows keep every thread busy (d = 16: 16 rows
    // at a time instead of 16 threads of 256).  Row lane rr sums rows rr, rr + rl, ...
    // in order, and the row lanes' partials are then added in lane order: a fixed
    // order, so the sums are the same bits on every run.
    static_assert(BFJ_MAXDIM <= BFJ_BLOCK, "one thread per column at the widest rows");
    __shared__ double red[3][BFJ_BLOCK];
    const int b = blockIdx.x;
    const SetDesc &s = a.sets[a.blk_set[b]];
    const int64_t r0 = a.blk_row0[b];
    const int64_t plane = (int64_t)a.n_row_blocks * BFJ_MAXDIM;
    int cw = 1;
    while (cw < s.d) cw <<= 1; // d <= BFJ_MAXDIM = BFJ_BLOCK
    const int rl = BFJ_BLOCK / cw;
    const int k = threadIdx.x % cw, rr = threadIdx.x / cw;
    double acc = 0.0, mx = -INFINITY, mn = INFINITY;
    if (k < s.d) {
        for (int64_t i = r0 + rr; i < r0 + BFJ_PREP_ROWS && i < s.n; i += rl) {
            const double v = s.x[i * s.ld + k];
            acc += v;
            mx = fmax(mx, v);
            mn = fmin(mn, v);
        }
    }
    red[0][threadIdx.x] = acc;
    red[1][threadIdx.x] = mx;
    red[2][threadIdx.x] = mn;
    __syncthreads();
    if (rr == 0 && k < s.d) {
        for (int q = 1; q < rl; ++q) {
            acc += red[0][q * cw + k];
            mx = fmax(mx, red[1][q * cw + k]);
            mn = fmin(mn, red[2][q * cw + k]);
        }
        a.colsum[(int64_t)b * BFJ_MAXDIM + k] = acc;
        a.colsum[plane + (int64_t)b * BFJ_MAXDIM + k] = mx;
        a.colsum[2 * plane + (int64_t)b * BFJ_MAXDIM + k] = mn;
    }
}

// One block per (dataset, column): the block sums the column's partial sums in a
// fixed order (strided per thread, then a tree), so the mean is reproducible.
// It also reduces the column's extremes to max |x - mu| over the column, which goes
// into the dataset's max |x - mu|: fl(x - mu) is monotonic in x, so this is the
// largest fl(|x - mu|) over the column's elements exactly.  NaN or inf in the
// column makes the sum (and the mean) or an extreme non-finite; that marks the
// dataset.
__global__ void k_colmean(PassArgs a)
{
    __shared__ double red[BFJ_BLOCK], rmx[BFJ_BLOCK], rmn[BFJ_BLOCK];
    const int si = blockIdx.x, k = blockIdx.y;
    const SetDesc &s = a.sets[si];
    const int64_t plane = (int64_t)a.n_row_blocks * BFJ_MAXDIM;
    double acc = 0.0, mx = -INFINITY, mn = INFINITY;
    if (k < s.d) {
        const int64_t b0 = a.set_blk0[si];
        for (int b = threadIdx.x; b < a.set_nblk[si]; b += BFJ_BLOCK) {
            const int64_t o = (b0 + b) * BFJ_MAXDIM + k;
            acc += a.colsum[o];
            mx = fmax(mx, a.colsum[plane + o]);
            mn = fmin(mn, a.colsum[2 * plane + o]);
        }
    }
    red[threadIdx.x] = acc;
    rmx[threadIdx.x] = mx;
    rmn[threadIdx.x] = mn;
    __syncthreads();
    for (int h = BFJ_BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[threadIdx.x] += red[threadIdx.x + h];
            rmx[threadIdx.x] = fmax(rmx[threadIdx.x], rmx[threadIdx.x + h]);
            rmn[threadIdx.x] = fmin(rmn[threadIdx.x], rmn[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool col = s.n > 0 && k < s.d;
        const double mu = col ? red[0] / (double)s.n : 0.0;
        a.mu[(int64_t)si * BFJ_MAXDIM + k] = mu;
        if (col) {
            const double m = fmax(rmx[0] - mu, mu - rmn[0]);
            if (!isfinite(mu) || !isfinite(rmx[0]) || !isfinite(rmn[0]) || !isfinite(m))
                atomicOr(&a.scal[si].nonfinite, 1ull);
            else
                atomic_max_nonneg(&a.scal[si].maxabs_bits, m);
        }
    }
}

// ---- scale ------------------------------------------------------------------
// A copy of every kernel-space row at higher precision than the join's operands,
// row-major by position (BFJ_MAXDIM values per row), and its radius ||y_i - copy_i||
// bounded as in k_quantize16.  The middle stage (refine.hip) decides most of the join's
// band from it, reading whole chunks, so the columns from d up to a multiple of 32 are
// written as zeros.  This kernel writes the FP16 copy of FP8 passes (L lanes per row, 8
// columns per lane); k_quantize16 writes the FP32 copy of FP16/BF16 passes from the
// values it already holds.
__global__ __launch_bounds__(BFJ_BLOCK) void k_mid_pack(PassArgs a)
{
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const double *mu = a.mu + (int64_t)si * BFJ_MAXDIM;
    const double inv_sigma = 1.0 / a.scal[si].sigma; // exact: sigma is a power of two
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int units = (s.d + 31) / 32 * 4; // 8-column units, whole 32-column chunks
    int L = 1;
    while (L < units) L <<= 1;
    const int rpp = BFJ_WAVE / L;
    const int m = lane & (L - 1), lr = lane / L;
    const int k0 = 8 * m;
    const int64_t wrow0 = a.blk_row0[b] + wave * (BFJ_PREP_ROWS / 4);
    for (int rr = 0; rr < BFJ_PREP_ROWS / 4; rr += rpp) {
        const int64_t i = wrow0 + rr + lr;
        const bool in = rr + lr < BFJ_PREP_ROWS / 4 && i < s.npad && m < units;
        const bool real = in && i < s.n;
        const int64_t src = real ? (int64_t)bfj_row_at(a, s, i) : 0;
        double sq_e = 0.0, sq_y = 0.0;
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + j;
            const double v = (real && k < s.d) ? (s.x[src * s.ld + k] - mu[k]) * inv_sigma : 0.0;
            const _Float16 q = (_Float16)v;
            const double de = v - (double)q;
            const uint32_t b16 = __builtin_bit_cast(uint16_t, q);
            w[j / 2] = (j & 1) ? (w[j / 2] | (b16 << 16)) : b16;
            sq_e = fma(de, de, sq_e);
            sq_y = fma(v, v, sq_y);
        }
        if (in) *reinterpret_cast<uint4 *>(a.mid_rows + ((s.row_off + i) * BFJ_MAXDIM + k0) * 2) = make_uint4(w[0], w[1], w[2], w[3]);
        for (int o = L >> 1; o > 0; o >>= 1) {
            sq_e += __shfl_xor(sq_e, o);
            sq_y += __shfl_xor(sq_y, o);
        }
        if (m == 0 && rr + lr < BFJ_PREP_ROWS / 4 && i < s.npad)
            a.mid_q[s.row_off + i] = real ? (sqrt(sq_e) + 2.3e-16 * sqrt(sq_y)) * (1.0 + 1e-9) + 1e-300 : 0.0;
    }
}

// sigma = 2^ceil(log2(max|x - mu|)) so the centred, scaled rows lie in [-1, 1]
// for every dtype: the FP16/FP8 ranges are respected and no norm can overflow.
__global__ void k_sigma(PassArgs a)
{
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= a.n_sets) return;
    const double m = __longlong_as_double((long long)a.scal[si].maxabs_bits);
    double sigma = 1.0;
    if (m > 0.0 && isfinite(m)) {
        int e;
        frexp(m, &e); // m in [2^(e-1), 2^e)
        sigma = ldexp(1.0, e);
    }
    a.scal[si].sigma = sigma;
}

// ---- cast + pack + norms + quantisation radius ---------------------------------
__device__ inline uint8_t e4m3_encode(double v)
{
    // OCP FP8 e4m3 (bias 7, max 448, subnormals m * 2^-9), round to nearest even.
    const uint8_t sgn = v < 0.0 ? 0x80 : 0x00;
    double a = fabs(v);
    if (a > 448.0) a = 448.0;
    if (a == 0.0) return sgn;
    int e;
    frexp(a, &e);
    int E = e - 1; // a in [2^E, 2^(E+1))
    if (E < -6) {
        const double m = rint(ldexp(a, 9));
        if (m >= 8.0) return sgn | (uint8_t)(1 << 3);
        return sgn | (uint8_t)m;
    }
    double m = rint(ldexp(a, 3 - E)); // in [8, 16]
    if (m >= 16.0) { E += 1; m = 8.0; }
    if (E > 8 || (E == 8 && m > 14.0)) { E = 8; m = 14.0; }
    return sgn | (uint8_t)(((E + 7) << 3) | ((int)m - 8));
}

// One wavefront per row; 4 wavefronts x 16 rows per block.
__global__ __launch_bounds__(BFJ_BLOCK) void k_quantize(PassArgs a)
{
    __shared__ double redq[BFJ_BLOCK / BFJ_WAVE], redn[BFJ_BLOCK / BFJ_WAVE];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const double *mu = a.mu + (int64_t)si * BFJ_MAXDIM;
    const double sigma = a.scal[si].sigma;
    const double inv_sigma = 1.0 / sigma; // exact: sigma is a power of two
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int KCH = bfj_kchunk(a.dtype);
    const int FB = bfj_frag_bytes(a.dtype);
    const int ES = a.dtype == BFJ_FP64 ? 8 : (a.dtype == BFJ_FP8 ? 1 : 2);
    const int EPL = FB / ES; // elements per lane per chunk
    const int Kpad = s.kc * KCH;
    double qmax = 0.0, nmax = 0.0;
    for (int rr = 0; rr < BFJ_PREP_ROWS / 4; ++rr) {
        const int64_t i = a.blk_row0[b] + wave * (BFJ_PREP_ROWS / 4) + rr;
        if (i >= s.npad) break;
        const bool real = i < s.n;
        const int64_t src = real ? (int64_t)bfj_row_at(a, s, i) : 0; // the row packed at position i
        // Centred, scaled row values for this lane's k = lane + 64 t.
        double y[BFJ_MAXDIM / BFJ_WAVE];
        double rmax = 0.0;
#pragma unroll
        for (int t = 0; t < BFJ_MAXDIM / BFJ_WAVE; ++t) {
            const int k = lane + BFJ_WAVE * t;
            double v = 0.0;
            if (real && k < s.d) v = (s.x[src * s.ld + k] - mu[k]) * inv_sigma;
            y[t] = v;
            rmax = fmax(rmax, fabs(v));
        }
        int scale_byte = 127;
        double rs = 1.0;
        if (a.dtype == BFJ_FP8) {
            rmax = wave_max(rmax);
            if (rmax > 0.0) {
                int e;
                frexp(rmax / 448.0, &e); // rmax/448 in [2^(e-1), 2^e)
                int ex = e;
                if (ldexp(1.0, e - 1) == rmax / 448.0) ex = e - 1;
                if (ex < -127) ex = -127;
                if (ex > 127) ex = 127;
                scale_byte = 127 + ex;
                rs = ldexp(1.0, ex);
            }
        }
        double sq_e = 0.0, sq_n = 0.0, sq_y = 0.0;
        unsigned char *base = a.pack + s.pack_off + (i / 16) * (int64_t)s.kc * BFJ_WAVE * FB;
#pragma unroll
        for (int t = 0; t < BFJ_MAXDIM / BFJ_WAVE; ++t) {
            const int k = lane + BFJ_WAVE * t;
            if (k >= Kpad) break;
            const double v = y[t];
            double xh;
            const int c = k / KCH, kl = k % KCH, g = kl / EPL, vv = kl % EPL;
            unsigned char *dst = base + ((int64_t)c * BFJ_WAVE + (i % 16) + 16 * g) * FB + vv * ES;
            if (a.dtype == BFJ_FP64) {
                xh = v;
                *(double *)dst = xh;
            } else if (a.dtype == BFJ_FP16) {
                const _Float16 h = (_Float16)v;
                xh = (double)h;
                *(_Float16 *)dst = h;
            } else if (a.dtype == BFJ_BF16) {
                const __bf16 h = (__bf16)(float)v;
                xh = (double)(float)h;
                *(__bf16 *)dst = h;
            } else {
                const uint8_t q = e4m3_encode(v / rs);
                xh = e4m3_decode(q) * rs;
                *(uint8_t *)dst = q;
            }
            const double de = v - xh;
            sq_e = fma(de, de, sq_e);
            sq_n = fma(xh, xh, sq_n);
            sq_y = fma(v, v, sq_y);
        }
        sq_e = wave_sum(sq_e);
        sq_n = wave_sum(sq_n);
        sq_y = wave_sum(sq_y);
        if (lane == 0) {
            const int64_t row = s.row_off + i;
            // Radius: measured cast error, plus the FP64 rounding of (x - mu), inflated
            // to cover the FP64 evaluation of the sums.
            // (+ 1e-300 covers underflow of the scaled values, which are at most 1.)
            const double q = real ? (sqrt(sq_e) + 2.3e-16 * sqrt(sq_y)) * (1.0 + 1e-9) + 1e-300 : 0.0;
            a.qerr[row] = q;
            a.nrm[row] = sq_n;
            if (a.dtype == BFJ_FP8) a.scales[row] = (uint8_t)scale_byte;
            if (a.dtype == BFJ_FP64) ((double *)a.nq)[row] = real ? 0.5 * sq_n : 1e300;
            else ((float *)a.nq)[row] = real ? 0.5f * (float)sq_n : 1e30f;
            if (real) {
                qmax = fmax(qmax, q);
                nmax = fmax(nmax, sq_n);
            }
        }
    }
    if (lane == 0) {
        redq[wave] = qmax;
        redn[wave] = nmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BFJ_BLOCK / BFJ_WAVE; ++w) {
            qmax = fmax(qmax, redq[w]);
            nmax = fmax(nmax, redn[w]);
        }
        atomic_max_nonneg(&a.scal[si].qmax_bits, qmax);
        atomic_max_nonneg(&a.scal[si].nmax_bits, nmax);
    }
}

// FP16 / BF16: one group of L lanes per row, each lane 8 consecutive columns, which
// are exactly one 16-byte unit of the packed fragments (16x16x32 layout: K-chunk
// k/32, lane-quarter (k%32)/8), so a row is read with 16-byte loads and written
// with one 16-byte store per lane; 64/L rows per wavefront at a time.  The row's
// sums are then reduced over its L lanes only.  Same outputs as k_quantize.
template <bool BF>
__global__ __launch_bounds__(BFJ_BLOCK) void k_quantize16(PassArgs a)
{
    __shared__ double redq[BFJ_BLOCK / BFJ_WAVE], redn[BFJ_BLOCK / BFJ_WAVE];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const double *mu = a.mu + (int64_t)si * BFJ_MAXDIM;
    const double inv_sigma = 1.0 / a.scal[si].sigma; // exact: sigma is a power of two
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int units = s.kc * 4;                  // 8-column units of a padded row
    int L = 4;
    while (L < units) L <<= 1;                   // lanes per row: a power of two <= 64
    const int rpp = BFJ_WAVE / L;                // rows per pass
    const int m = lane & (L - 1), lr = lane / L; // this lane's unit and row of the pass
    const int k0 = 8 * m;
    double qmax = 0.0, nmax = 0.0;
    const int64_t wrow0 = a.blk_row0[b] + wave * (BFJ_PREP_ROWS / 4);
    for (int rr = 0; rr < BFJ_PREP_ROWS / 4; rr += rpp) {
        const int64_t i = wrow0 + rr + lr;
        const bool in = rr + lr < BFJ_PREP_ROWS / 4 && i < s.npad && m < units;
        const bool real = in && i < s.n;
        const int64_t src = real ? (int64_t)bfj_row_at(a, s, i) : 0;
        double y[8];
        if (real && s.al16 && k0 + 8 <= s.d) {
            const double2 *p2 = reinterpret_cast<const double2 *>(s.x + src * s.ld + k0);
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const double2 w2 = p2[h];
                y[2 * h] = (w2.x - mu[k0 + 2 * h]) * inv_sigma;
                y[2 * h + 1] = (w2.y - mu[k0 + 2 * h + 1]) * inv_sigma;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + j;
                y[j] = (real && k < s.d) ? (s.x[src * s.ld + k] - mu[k]) * inv_sigma : 0.0;
            }
        }
        double sq_e = 0.0, sq_n = 0.0, sq_y = 0.0, sq_c = 0.0;
        if (a.mid) {
            // The middle stage's FP32 copy (units = 4 kc covers d up to a multiple of
            // 32, zeros past d) and its cast error.
            float f[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                f[j] = (float)y[j];
                const double dc = y[j] - (double)f[j];
                sq_c = fma(dc, dc, sq_c);
            }
            if (in && k0 + 8 <= BFJ_MAXDIM) {
                uint4 *dst = reinterpret_cast<uint4 *>(a.mid_rows + ((s.row_off + i) * BFJ_MAXDIM + k0) * 4);
                dst[0] = __builtin_bit_cast(uint4, bfj_f4{f[0], f[1], f[2], f[3]});
                dst[1] = __builtin_bit_cast(uint4, bfj_f4{f[4], f[5], f[6], f[7]});
            }
        }
        uint32_t w[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            uint16_t bits[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const double v = y[2 * h + u];
                double xh;
                if (BF) {
                    const __bf16 q = (__bf16)(float)v;
                    xh = (double)(float)q;
                    bits[u] = __builtin_bit_cast(uint16_t, q);
                } else {
                    const _Float16 q = (_Float16)v;
                    xh = (double)q;
                    bits[u] = __builtin_bit_cast(uint16_t, q);
                }
                const double de = v - xh;
                sq_e = fma(de, de, sq_e);
                sq_n = fma(xh, xh, sq_n);
                sq_y = fma(v, v, sq_y);
            }
            w[h] = (uint32_t)bits[0] | ((uint32_t)bits[1] << 16);
        }
        if (in) {
            const int c = m >> 2, g = m & 3;
            uint4 *dst = reinterpret_cast<uint4 *>(a.pack + s.pack_off) +
                         ((i / 16) * s.kc + c) * BFJ_WAVE + (i % 16) + 16 * g;
            *dst = make_uint4(w[0], w[1], w[2], w[3]);
        }
        for (int o = L >> 1; o > 0; o >>= 1) {
            sq_e += __shfl_xor(sq_e, o);
            sq_n += __shfl_xor(sq_n, o);
            sq_y += __shfl_xor(sq_y, o);
            if (a.mid) sq_c += __shfl_xor(sq_c, o);
        }
        if (m == 0 && rr + lr < BFJ_PREP_ROWS / 4 && i < s.npad) {
            const int64_t row = s.row_off + i;
            // Radius: measured cast error, plus the FP64 rounding of (x - mu), inflated
            // to cover the FP64 evaluation of the sums (as in k_quantize).
            const double q = real ? (sqrt(sq_e) + 2.3e-16 * sqrt(sq_y)) * (1.0 + 1e-9) + 1e-300 : 0.0;
            if (a.mid) a.mid_q[row] = real ? (sqrt(sq_c) + 2.3e-16 * sqrt(sq_y)) * (1.0 + 1e-9) + 1e-300 : 0.0;
            a.qerr[row] = q;
            a.nrm[row] = sq_n;
            ((float *)a.nq)[row] = real ? 0.5f * (float)sq_n : 1e30f;
            if (real) {
                qmax = fmax(qmax, q);
                nmax = fmax(nmax, sq_n);
            }
        }
    }
    qmax = wave_max(qmax);
    nmax = wave_max(nmax);
    if (lane == 0) {
        redq[wave] = qmax;
        redn[wave] = nmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w2 = 1; w2 < BFJ_BLOCK / BFJ_WAVE; ++w2) {
            qmax = fmax(qmax, redq[w2]);
            nmax = fmax(nmax, redn[w2]);
        }
        atomic_max_nonneg(&a.scal[si].qmax_bits, qmax);
        atomic_max_nonneg(&a.scal[si].nmax_bits, nmax);
    }
}

// ---- thresholds -----------------------------------------------------------------
__device__ inline float f32_up(double v)
{
    if (isnan(v)) return __int_as_float(0x7fc00000);
    if (v >= 3.0e38) return INFINITY;
    if (v <= -3.0e38) return -3.0e38f;
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

__device__ inline float f32_down(double v)
{
    if (isnan(v)) return __int_as_float(0x7fc00000);
    if (v <= -3.0e38) return -INFINITY;
    if (v >= 3.0e38) return 3.0e38f;
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}

// One thread per row.
__global__ void k_thresholds(PassArgs a)
{
    __shared__ double redh[BFJ_PREP_ROWS];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const SetScalars &sc = a.scal[si];
    const int64_t i = a.blk_row0[b] + threadIdx.x;
    double halfband = 0.0;
    if (i < s.npad) {
        const int64_t row = s.row_off + i;
        const bool fp64 = a.dtype == BFJ_FP64;
        const bool query = i < s.n && i >= s.row_begin && i < s.row_end;
        const double sigma2 = sc.sigma * sc.sigma;
        const double qmax = __longlong_as_double((long long)sc.qmax_bits);
        const double nmax = __longlong_as_double((long long)sc.nmax_bits);
        const double Ni = a.nrm[row];
        const double r = a.qerr[row] + qmax;
        const int K = s.kc * bfj_kchunk(a.dtype);
        const double u = fp64 ? 0x1p-53 : 0x1p-24;
        const double rho = 1.01 * (s.d + 2) * 0x1p-53;
        const double u64 = 0x1p-53;
        // Every eps is one of: none (T < 0: no pair), all (T at least the largest
        // possible FP64 distance of the dataset: every pair, counted in k_finalize),
        // or banded.  Only banded eps shape g, so Z never saturates.
        // Largest oracle d2 in kernel space: ||y_i - y_j|| <= 2 (sqrt(N_max) + q_max).
        const double dmax = 2.0 * (sqrt(nmax) + qmax) * (1.0 + 0x1p-40);
        const double Dmax2 = dmax * dmax * (1.0 + rho) * (1.0 + 0x1p-40);
        double Hc[BFJ_MAXEPS], Lc[BFJ_MAXEPS];
        uint32_t all_mask = 0;
        // Hmax: the largest banded H of this row (r = q_i + q_max).  Every pair (i, j)
        // is decided with its own r_ij = q_i + q_j <= q_i + q_max, so H(r_ij) <= Hmax
        // and Z < 0 is a certain reject of the pair, for both of its rows.
        double Hmax = -INFINITY, shi_max = -INFINITY;
        for (int e = 0; e < s.n_eps; ++e) {
            const double T = s.T[e] / sigma2; // kernel space (exact power-of-two scaling)
            Hc[e] = Lc[e] = -INFINITY;
            if (T < 0.0) {
                // no pair: both bounds stay -inf
            } else if (T >= Dmax2) {
                all_mask |= 1u << e;
            } else {
                const double slo = sqrt(T / (1.0 + rho)) * (1.0 - 0x1p-50);
                const double shi = sqrt(T / (1.0 - rho)) * (1.0 + 0x1p-50);
                Lc[e] = slo > r ? (slo - r) * (slo - r) * (1.0 - 0x1p-49) : -INFINITY;
                Hc[e] = (shi + r) * (shi + r) * (1.0 + 0x1p-49);
                if (query) halfband = fmax(halfband, 0.5 * (Hc[e] - (Lc[e] > 0.0 ? Lc[e] : 0.0)) * sigma2);
                Hmax = fmax(Hmax, Hc[e]);
                shi_max = fmax(shi_max, shi);
            }
        }
        if (i == 0 && all_mask) atomicOr(&a.scal[si].all_mask, (unsigned long long)all_mask);
        // FP8 passes split the reject threshold into a row part and a column part: for
        // every banded eps and every partner j,
        //   (shi_e + q_i + q_j)^2 <= (shi_max + q_i)^2 + [2 (shi_max + q_max) q_j + q_j^2]
        // (q_i <= q_max; both parts rounded up, with room for the (1 + 2^-49) and the
        // (1 + 2^-50) of the band tests).  g_i carries the row part Arow; the join's
        // candidate threshold n'_j = n_j - Bcol_j / 2 (rounded down) carries the column
        // part, and dnq_j = n_j - n'_j lets the band tests rebuild Z = acc - n_j.  A pair
        // whose q_j is below q_max is then rejected by the join instead of streamed
        // (FP8 workload 4: 40 % fewer candidates).  The other dtypes keep Hmax and
        // n'_j = n_j: their radii are small against the thresholds, so the split
        // removes a few per cent of the candidates, less than the extra gather of dnq
        // per record costs k_classify.
        const bool split = a.dtype == BFJ_FP8 && Hmax != -INFINITY;
        const double up = 1.0 + 0x1p-46;
        const double qi_u = a.qerr[row] * up, qm_u = qmax * up;
        const double Arow = split ? (shi_max + qi_u) * (shi_max + qi_u) * up : Hmax;
        const double Bcol = split ? (2.0 * (shi_max + qm_u) * qi_u + qi_u * qi_u) * up : 0.0;
        const double Bmax = split ? (2.0 * (shi_max + qm_u) * qm_u + qm_u * qm_u) * up : 0.0;
        double dn = 0.0;
        if (split && i < s.n) {
            const double nqv = (double)((float *)a.nq)[row];
            double t = nqv - 0.5 * Bcol;
            t -= fabs(t) * 0x1p-50 + 0x1p-1000; // below the exact difference (4x its rounding)
            const float nqn = f32_down(t);
            ((float *)a.nq)[row] = nqn;
            dn = nqv - (double)nqn;
        }
        a.dnq[row] = dn;
        double g; // FP64 value of the per-row initial term actually stored
        double E = 0.0;
        if (!query || Hmax == -INFINITY) {
            g = fp64 ? -1e300 : -1e30;
        } else {
            // Accumulation bound; |g| is bounded from the previous estimate.  The MFMA
            // sums g_i and the K products (partial sums within |g| + sqrt(N_i N_max));
            // n_j enters once, in the final subtraction and its own FP32 rounding.
            const double Kp = 2.0 * K + 8.0;
            const double coef = 2.0 * (1.01 * Kp + 1.0) * u;
            const double cross = sqrt(Ni * nmax);
            double gabs = 0.5 * fabs(Arow - Ni);
            for (int it = 0; it < 2; ++it) {
                // 4 u Bmax: the final subtraction acc - n'_j, whose operand is up to
                // Bmax / 2 further from acc than n_j.
                E = coef * (gabs + cross) + 4.0 * u * (nmax + Bmax) + 4.0 * (K + 4) * u64 * (Ni + nmax + gabs) +
                    K * 0x1p-120;
                gabs = 0.5 * fabs(Arow - Ni) + E;
            }
            E *= 1.01;
            // Rounded up whatever its sign: 2 g + N_i - E >= Arow, so Z' = acc - n'_j < 0
            // gives S_hat = 2 g + N_i - 2 (Z' - dnq_j) > Arow + E + Bcol_j, a certain reject
            // (S >= S_hat - E > Arow + Bcol_j >= H(r_ij)) for every banded eps.
            const double x = 0.5 * (Arow + E - Ni);
            const double gv = x + fabs(x) * 0x1p-50 + 0x1p-1000;
            g = fp64 ? gv : (double)f32_up(gv);
        }
        if (fp64) ((double *)a.gq)[row] = g;
        else ((float *)a.gq)[row] = (float)g;
        a.he[row] = make_double2(2.0 * g + Ni, E);
        a.row_set[row] = (uint16_t)si;
        if (i == 0) {
            // Per-dataset radii of the banded eps; k_classify forms the bounds of each
            // pair from them and the pair's own r = q_i + q_j.
            uint32_t band_mask = 0;
            for (int e = 0; e < s.n_eps; ++e) {
                const double T = s.T[e] / sigma2;
                const bool banded = T >= 0.0 && !((all_mask >> e) & 1u);
                if (banded) band_mask |= 1u << e;
                a.scal[si].slo[e] = banded ? sqrt(T / (1.0 + rho)) * (1.0 - 0x1p-50) : 0.0;
                a.scal[si].shi[e] = banded ? sqrt(T / (1.0 - rho)) * (1.0 + 0x1p-50) : 0.0;
            }
            a.scal[si].band_mask = band_mask;
        }
    }
    redh[threadIdx.x] = halfband;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int t = 0; t < BFJ_PREP_ROWS; ++t) m = fmax(m, redh[t]);
        atomic_max_nonneg(&a.scal[si].halfband_bits, m);
    }
}

void bfj_launch_prep(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_colsum, dim3(a.n_row_blocks), dim3(BFJ_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_colmean, dim3(a.n_sets, BFJ_MAXDIM), dim3(BFJ_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_sigma, dim3((a.n_sets + 255) / 256), dim3(256), 0, st, a);
    bfj_launch_order(a, st);
    if (a.dtype == BFJ_FP16) hipLaunchKernelGGL(k_quantize16<false>, dim3(a.n_row_blocks), dim3(BFJ_BLOCK), 0, st, a);
    else if (a.dtype == BFJ_BF16) hipLaunchKernelGGL(k_quantize16<true>, dim3(a.n_row_blocks), dim3(BFJ_BLOCK), 0, st, a);
    else hipLaunchKernelGGL(k_quantize, dim3(a.n_row_blocks), dim3(BFJ_BLOCK), 0, st, a);
    if (a.mid && !a.mid_fp32) hipLaunchKernelGGL(k_mid_pack, dim3(a.n_row_blocks), dim3(BFJ_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_thresholds, dim3(a.n_row_blocks), dim3(BFJ_PREP_ROWS), 0, st, a);
}
