// errstat.hip -- accuracy statistics of a finished pass (bfj_pass_errors).
//
// The pass decides every pair from S_hat = 2 g_i + N_i - 2 Z (kernel space), the
// squared distance its dot products give, with |S_hat - S| <= E_i for S the squared
// distance of the cast rows, and |dist(cast rows) - dist(exact rows)| <= q_i + q_j.
// These kernels measure, after the pass and outside any timed region, how far
// sigma^2 S_hat lies from the reference's FP64 direct form d2 (kernel_join.cu:184-189,
// the same FMA chain as the oracle and k_refine), and check on real data that the
// interval the band tests assume does hold d2:
//   * k_errcand: the pass's own candidate records (the pairs the join could not
//     reject: those near or inside eps), all or every k-th chunk;
//   * k_errtile: seeded 16 x 16 tiles of all pairs, recomputed with the join's MFMA
//     sequence (same instruction, operands, C operand g_i and K-chunk order, so the
//     same Z), where the cast rows are also decoded so that the accumulation bound
//     E_i and the quantisation radius are each checked on their own.
#include "mfma_ops.h"
#include <math.h>

__device__ inline void errmax(unsigned long long *bits, double v)
{
    if (v >= 0.0) atomicMax(bits, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_fmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// FP64 direct form of rows p and q over d columns, k ascending (kernel_join.cu:184-189).
__device__ inline double direct_d2(const double *p, const double *q, int d)
{
    double acc = 0.0;
    for (int k = 0; k < d; ++k) {
        const double t = p[k] - q[k];
        acc = fma(t, t, acc);
    }
    return acc;
}

// Does d2 (input units) lie in the interval the band tests take for the pair: the
// exact distance D = dist(y_i, y_j) (kernel space) within [sqrt(lo) - r, sqrt(hi) + r]
// for S in [lo, hi], and d2 = sigma^2 D^2 (1 +- rho)?  (The same bounds as k_classify.)
__device__ inline bool in_certified(double d2, double sig2, double shat, double E, double z, double hx, double dj,
                                    double r, int d)
{
    const double mg = (fabs(hx) + 2.0 * fabs(z) + 2.0 * dj + E) * 0x1p-50;
    const double lo = shat - E - mg, hi = shat + E + mg;
    const double rho = 1.01 * (d + 2) * 0x1p-53;
    const double Dlo = sqrt(d2 / sig2 / (1.0 + rho)), Dhi = sqrt(d2 / sig2 / (1.0 - rho));
    const double blo = sqrt(lo > 0.0 ? lo : 0.0) - r, bhi = sqrt(hi > 0.0 ? hi : 0.0) + r;
    const double tol = 0x1p-44 * (bhi + Dhi) + 1e-300; // FP64 rounding of these few steps
    return Dhi >= blo - tol && Dlo <= bhi + tol;
}

// One wavefront per chunk of the candidate stream (chunk ch = t * stride), one lane
// per record.
__global__ __launch_bounds__(BFJ_BLOCK) void k_errcand(PassArgs a, ErrOut *out, uint32_t n_chunks, uint32_t stride)
{
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (BFJ_BLOCK / BFJ_WAVE) + (threadIdx.x >> 6);
    const uint32_t n_waves = gridDim.x * (BFJ_BLOCK / BFJ_WAVE);
    const uint32_t n_pick = (n_chunks + stride - 1) / stride;
    double emax = -1.0;
    unsigned long long checked = 0, bad = 0;
    for (uint32_t t = wave; t < n_pick; t += n_waves) {
        const uint32_t ch = t * stride;
        const uint32_t fill = min(a.chunk_fill[ch], (uint32_t)BFJ_CHUNK);
        const Candidate *recs = a.cand + (size_t)ch * BFJ_CHUNK;
        for (uint32_t k = lane; k < fill; k += BFJ_WAVE) {
            const Candidate c = recs[k];
            const uint32_t col = c.col & ~BFJ_COL_SYM;
            const int si = (int)a.row_set[c.row];
            const SetDesc &s = a.sets[si];
            const double sig2 = a.scal[si].sigma * a.scal[si].sigma;
            const double *p = s.x + (int64_t)bfj_row_at(a, s, c.row - s.row_off) * s.ld;
            const double *q = s.x + (int64_t)bfj_row_at(a, s, col - s.row_off) * s.ld;
            const double d2 = direct_d2(p, q, s.d);
            const double2 h = a.he[c.row];
            const double dj = a.dnq[col]; // the record holds Z' = acc - n'_j = Z + dnq_j
            const double shat = h.x - 2.0 * c.z + 2.0 * dj;
            emax = fmax(emax, fabs(sig2 * shat - d2));
            const double r = (a.qerr[c.row] + a.qerr[col]) * (1.0 + 0x1p-50);
            bad += in_certified(d2, sig2, shat, h.y, c.z, h.x, dj, r, s.d) ? 0 : 1;
            ++checked;
        }
    }
    emax = wave_fmax(emax);
    checked = wave_sum_u64(checked);
    bad = wave_sum_u64(bad);
    if (lane == 0) {
        errmax(&out->cand_max_bits, emax);
        if (checked) atomicAdd(&out->cand_pairs, checked);
        if (bad) atomicAdd(&out->violations, bad);
    }
}

void bfj_launch_errcand(const PassArgs &a, ErrOut *out, uint32_t n_chunks, uint32_t stride, hipStream_t st)
{
    if (!n_chunks) return;
    hipLaunchKernelGGL(k_errcand, dim3(2048), dim3(BFJ_BLOCK), 0, st, a, out, n_chunks, stride < 1 ? 1 : stride);
}

// Cast value number v of a lane's fragment of K-chunk c (the packed layout of
// k_quantize / k_quantize16: lane L holds row L % 16, values k = c KCH + (L / 16) EPL + v).
template <class Tr>
__device__ inline float frag_value(const typename Tr::frag_t &f, int v, int scale_byte)
{
    if constexpr (Tr::dtype == BFJ_FP16 || Tr::dtype == BFJ_BF16) {
        return (float)f[v];
    } else if constexpr (Tr::dtype == BFJ_FP8) {
        const uint8_t b = (uint8_t)((uint32_t)f[v / 4] >> (8 * (v % 4)));
        return (float)(e4m3_decode(b) * ldexp(1.0, scale_byte - 127));
    } else {
        return (float)f[v];
    }
}

// One wavefront per sampled tile: the join's accumulators for the 16 x 16 pairs of row
// block rb against candidate block cb, then per pair the FP64 direct form, and (low
// precision) the exact squared distance of the cast rows from their decoded values.
template <class Tr>
__global__ __launch_bounds__(BFJ_WAVE) void k_errtile(PassArgs a, const ErrTile *tiles, uint32_t n_tiles, ErrOut *out)
{
    typedef typename Tr::frag_t frag_t;
    typedef typename Tr::acc_t acc_t;
    typedef typename Tr::val_t val_t;
    constexpr bool LOW = Tr::dtype != BFJ_FP64;
    constexpr int KCH = Tr::dtype == BFJ_FP64 ? 8 : (Tr::dtype == BFJ_FP8 ? 128 : 32);
    constexpr int EPL = Tr::dtype == BFJ_FP64 ? 2 : (Tr::dtype == BFJ_FP8 ? 32 : 8);
    __shared__ float vals[2][16][BFJ_MAXDIM]; // decoded cast rows of the two blocks
    const int lane = threadIdx.x;
    double emax = -1.0, amax = -1.0, qmax = -1.0;
    unsigned long long checked = 0, bad = 0;
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const ErrTile w = tiles[t];
        const SetDesc &s = a.sets[w.set];
        const frag_t *P = reinterpret_cast<const frag_t *>(a.pack + s.pack_off);
        const val_t *nq = reinterpret_cast<const val_t *>(a.nq) + s.row_off;
        const val_t *gq = reinterpret_cast<const val_t *>(a.gq) + s.row_off;
        const uint8_t *scb = a.scales + s.row_off;
        const int sa = Tr::dtype == BFJ_FP8 ? (int)scb[w.rb * 16 + (lane & 15)] : 127;
        const int sb = Tr::dtype == BFJ_FP8 ? (int)scb[w.cb * 16 + (lane & 15)] : 127;
        acc_t G;
#pragma unroll
        for (int r = 0; r < 4; ++r) G[r] = gq[w.rb * 16 + Tr::out_row(lane, r)];
        acc_t acc = G;
        for (int c = 0; c < s.kc; ++c) {
            const frag_t fa = P[((int64_t)w.rb * s.kc + c) * BFJ_WAVE + lane];
            const frag_t fb = P[((int64_t)w.cb * s.kc + c) * BFJ_WAVE + lane];
            acc = Tr::mma(fa, fb, acc, sa, sb); // the join's chain: C = g_i first, then K-chunks in order
            if (LOW) {
                const int row = lane & 15, k0 = c * KCH + (lane >> 4) * EPL;
#pragma unroll
                for (int v = 0; v < EPL; ++v) {
                    vals[0][row][k0 + v] = frag_value<Tr>(fa, v, sa);
                    vals[1][row][k0 + v] = frag_value<Tr>(fb, v, sb);
                }
            }
        }
        __syncthreads();
        const val_t nj = nq[w.cb * 16 + (lane & 15)];
        const double sig2 = a.scal[w.set].sigma * a.scal[w.set].sigma;
        const int jl = lane & 15, pj = w.cb * 16 + jl;
        const int kp = s.kc * KCH;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int il = Tr::out_row(lane, r), pi = w.rb * 16 + il;
            if (pi >= s.n || pj >= s.n) continue;
            const uint32_t ri = bfj_row_at(a, s, pi), rj = bfj_row_at(a, s, pj);
            if ((int64_t)ri < s.row_begin || (int64_t)ri >= s.row_end) continue; // not a query row: no g_i
            const double z = (double)(acc[r] - nj); // Z' as the join forms it (in val_t)
            const double2 h = a.he[s.row_off + pi];
            const double dj = a.dnq[s.row_off + pj];
            const double shat = h.x - 2.0 * z + 2.0 * dj;
            const double d2 = direct_d2(s.x + (int64_t)ri * s.ld, s.x + (int64_t)rj * s.ld, s.d);
            emax = fmax(emax, fabs(sig2 * shat - d2));
            const double qi = a.qerr[s.row_off + pi], qj = a.qerr[s.row_off + pj];
            bad += in_certified(d2, sig2, shat, h.y, z, h.x, dj, (qi + qj) * (1.0 + 0x1p-50), s.d) ? 0 : 1;
            ++checked;
            if (LOW) {
                double sq = 0.0; // exact up to FP64 rounding: differences of FP32-exact values
                for (int k = 0; k < kp; ++k) {
                    const double t2 = (double)vals[0][il][k] - (double)vals[1][jl][k];
                    sq = fma(t2, t2, sq);
                }
                if (h.y > 0.0) amax = fmax(amax, fabs(shat - sq) / h.y);
                if (qi + qj > 0.0) qmax = fmax(qmax, fabs(sqrt(sq) - sqrt(d2 / sig2)) / (qi + qj));
            }
        }
        __syncthreads();
    }
    emax = wave_fmax(emax);
    amax = wave_fmax(amax);
    qmax = wave_fmax(qmax);
    checked = wave_sum_u64(checked);
    bad = wave_sum_u64(bad);
    if (lane == 0) {
        errmax(&out->tile_max_bits, emax);
        errmax(&out->acc_ratio_bits, amax);
        errmax(&out->quant_ratio_bits, qmax);
        if (checked) atomicAdd(&out->tile_pairs, checked);
        if (bad) atomicAdd(&out->violations, bad);
    }
}

void bfj_launch_errtile(const PassArgs &a, const ErrTile *tiles, uint32_t n_tiles, ErrOut *out, hipStream_t st)
{
    if (!n_tiles) return;
    const unsigned grid = n_tiles < 8192u ? n_tiles : 8192u;
    switch (a.dtype) {
    case BFJ_FP16: hipLaunchKernelGGL(k_errtile<TrF16>, dim3(grid), dim3(BFJ_WAVE), 0, st, a, tiles, n_tiles, out); break;
    case BFJ_BF16: hipLaunchKernelGGL(k_errtile<TrBF16>, dim3(grid), dim3(BFJ_WAVE), 0, st, a, tiles, n_tiles, out); break;
    case BFJ_FP8: hipLaunchKernelGGL(k_errtile<TrF8>, dim3(grid), dim3(BFJ_WAVE), 0, st, a, tiles, n_tiles, out); break;
    default: hipLaunchKernelGGL(k_errtile<TrF64>, dim3(grid), dim3(BFJ_WAVE), 0, st, a, tiles, n_tiles, out); break;
    }
}
