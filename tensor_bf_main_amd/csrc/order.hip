// order.hip -- a locality order for the rows of a dataset before they are packed.
//
// The join tests every pair; what an order changes is where the few pairs near the
// threshold land.  In input order they are scattered, so about one 64x16 tile in ten
// holds one and takes the slow candidate path.  Rows grouped by their nearest of
// BFJ_PIVOTS pivot rows put those pairs into few tiles (measured on the 64-centre
// GMM workload: 1.6 % of the tiles hot instead of 11 %), and the candidate records,
// the refine's row gathers and the count atomics of a chunk then fall on few rows.
//
// The order is a permutation perm of the dataset-local rows: position p of the
// packed fragments, norms, radii and counts holds the dataset's row perm[p].  Every
// result is mapped back (k_finalize writes counts at perm[p], the refine reads FP64
// row perm[p], pairs are reported as perm[i], perm[j]), so results do not depend on
// it.  A dataset that is not reordered (row-range passes, small n, or when the
// context turns it off) gets the identity.
//
// Steps, all deterministic:
//   k_order_pivots   pivot p = row p*n/P, centred and scaled like the packed rows,
//                    as FP16 (zero-padded to a multiple of 32 columns), and its
//                    squared norm
//   k_order_label    nearest pivot of every row: an FP16 MFMA of 16 rows x 128
//                    pivots per wavefront (argmin of |p|^2 - 2 x.p, ties to the
//                    lower pivot), and the 128-bin histogram of each 64-row block
//   k_order_scan*    exclusive offsets: bin-major over the dataset, block-major
//                    within a bin (three passes over [blocks][bins])
//   k_order_scatter  a row's position = its (block, bin) offset + its rank among the
//                    block's rows of that bin
#include "bfj_internal.h"

typedef _Float16 ord_h8 __attribute__((ext_vector_type(8)));
typedef float ord_f4 __attribute__((ext_vector_type(4)));

#define ORD_SEG BFJ_ORD_SEG

__device__ inline bool ord_active(const SetDesc &s) { return s.reorder != 0; }

// One block per dataset: BFJ_PIVOTS rows, centred, scaled by sigma (|y| <= 1), FP16,
// zero-padded to a multiple of 32 columns, and their squared norms in FP32.
__global__ void k_order_pivots(PassArgs a)
{
    const int si = blockIdx.x;
    const SetDesc &s = a.sets[si];
    if (!ord_active(s)) return;
    const double *mu = a.mu + (int64_t)si * BFJ_MAXDIM;
    const double inv_sigma = 1.0 / a.scal[si].sigma; // exact: a power of two
    const int K = (s.d + 31) / 32 * 32; // d <= BFJ_MAXDIM
    _Float16 *pv = a.pivots + (int64_t)si * BFJ_PIVOTS * BFJ_MAXDIM;
    for (int t = threadIdx.x; t < BFJ_PIVOTS * K; t += blockDim.x) {
        const int p = t / K, k = t % K;
        const int64_t row = (int64_t)p * s.n / BFJ_PIVOTS;
        const double v = k < s.d ? (s.x[row * s.ld + k] - mu[k]) * inv_sigma : 0.0;
        pv[p * BFJ_MAXDIM + k] = (_Float16)v;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < BFJ_PIVOTS; p += blockDim.x) {
        float acc = 0.f;
        for (int k = 0; k < K; ++k) {
            const float v = (float)pv[p * BFJ_MAXDIM + k];
            acc += v * v;
        }
        a.pivot_norm[si * BFJ_PIVOTS + p] = acc;
    }
}

// One 256-thread block per 64-row prep block: each wavefront labels 16 rows.
__global__ __launch_bounds__(BFJ_BLOCK) void k_order_label(PassArgs a)
{
    __shared__ __attribute__((aligned(16))) _Float16 pch[BFJ_PIVOTS * 32]; // pivots x one K-chunk
    __shared__ uint32_t hist[BFJ_PIVOTS];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    if (!ord_active(s)) return; // uniform over the block
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = threadIdx.x; t < BFJ_PIVOTS; t += BFJ_BLOCK) hist[t] = 0;
    const double *mu = a.mu + (int64_t)si * BFJ_MAXDIM;
    const double inv_sigma = 1.0 / a.scal[si].sigma;
    const int64_t r = a.blk_row0[b] + w * 16 + (lane & 15); // this lane's A row
    const bool real = r < s.n;
    const _Float16 *pv = a.pivots + (int64_t)si * BFJ_PIVOTS * BFJ_MAXDIM;
    const int nch = (s.d + 31) / 32;
    ord_f4 acc[BFJ_PIVOTS / 16];
#pragma unroll
    for (int t = 0; t < BFJ_PIVOTS / 16; ++t) acc[t] = ord_f4{0.f, 0.f, 0.f, 0.f};
    // A fragments of every K-chunk first (row r, columns 32c + 8 (lane / 16) + j), so that
    // all of the row's loads are in flight together rather than one chunk per pair of
    // barriers; 16-byte loads when the row's 8 values are whole and aligned.
    ord_h8 avs[BFJ_MAXDIM / 32];
#pragma unroll
    for (int c = 0; c < BFJ_MAXDIM / 32; ++c) {
        if (c >= nch) break; // uniform
        const int k0 = c * 32 + 8 * (lane >> 4);
        ord_h8 av;
        if (real && s.al16 && k0 + 8 <= s.d) {
            const double2 *src = reinterpret_cast<const double2 *>(s.x + r * s.ld + k0);
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const double2 w2 = src[h];
                av[2 * h] = (_Float16)((w2.x - mu[k0 + 2 * h]) * inv_sigma);
                av[2 * h + 1] = (_Float16)((w2.y - mu[k0 + 2 * h + 1]) * inv_sigma);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + j;
                const double v = (real && k < s.d) ? (s.x[r * s.ld + k] - mu[k]) * inv_sigma : 0.0;
                av[j] = (_Float16)v;
            }
        }
        avs[c] = av;
    }
#pragma unroll
    for (int c = 0; c < BFJ_MAXDIM / 32; ++c) {
        if (c >= nch) break; // uniform
        __syncthreads(); // previous chunk consumed
        // Pivot columns [32c, 32c + 32) of every pivot: 128 x 32 halves.
        for (int t = threadIdx.x; t < BFJ_PIVOTS * 4; t += BFJ_BLOCK) {
            const int p = t >> 2, q = t & 3;
            *(ord_h8 *)(pch + p * 32 + q * 8) = *(const ord_h8 *)(pv + p * BFJ_MAXDIM + c * 32 + q * 8);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < BFJ_PIVOTS / 16; ++t) {
            const ord_h8 bv = *(const ord_h8 *)(pch + (t * 16 + (lane & 15)) * 32 + 8 * (lane >> 4));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(avs[c], bv, acc[t], 0, 0, 0);
        }
    }
    // acc[t][i] of lane l: row 4 (l / 16) + i of the wavefront's 16, pivot 16 t + l % 16.
    float best[4];
    int arg[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        best[i] = INFINITY;
        arg[i] = 0;
    }
#pragma unroll
    for (int t = 0; t < BFJ_PIVOTS / 16; ++t) {
        const int p = t * 16 + (lane & 15);
        const float pn = a.pivot_norm[si * BFJ_PIVOTS + p];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float sc = pn - 2.f * acc[t][i];
            if (sc < best[i]) { // ascending p per lane: ties keep the lower pivot
                best[i] = sc;
                arg[i] = p;
            }
        }
    }
    // Argmin over the 16 lanes that hold the same rows (lowest pivot on ties).
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float ob = __shfl_xor(best[i], o);
            const int oa = __shfl_xor(arg[i], o);
            if (ob < best[i] || (ob == best[i] && oa < arg[i])) {
                best[i] = ob;
                arg[i] = oa;
            }
        }
    if ((lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = a.blk_row0[b] + w * 16 + 4 * (lane >> 4) + i;
            if (row < s.n) {
                a.label[s.row_off + row] = (uint8_t)arg[i];
                atomicAdd(&hist[arg[i]], 1u);
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < BFJ_PIVOTS; t += BFJ_BLOCK) a.ohist[(int64_t)b * BFJ_PIVOTS + t] = hist[t];
}

// Scan pass 1: the per-bin sums of each segment of ORD_SEG consecutive blocks of a
// dataset.  Grid (segments, datasets), one thread per bin.
__global__ void k_order_scan1(PassArgs a)
{
    const int si = blockIdx.y, g = blockIdx.x, bin = threadIdx.x;
    const SetDesc &s = a.sets[si];
    const int b0 = a.set_blk0[si] + g * ORD_SEG, b1 = a.set_blk0[si] + a.set_nblk[si];
    if (!ord_active(s) || b0 >= b1) return;
    uint32_t sum = 0;
    for (int b = b0; b < b0 + ORD_SEG && b < b1; ++b) sum += a.ohist[(int64_t)b * BFJ_PIVOTS + bin];
    a.oseg[((int64_t)si * a.oseg_stride + g) * BFJ_PIVOTS + bin] = sum;
}

// Scan pass 2: one block per dataset, one thread per bin.  The segment sums become
// each segment's first position within the dataset: the start of its bin (bins in
// ascending order) plus the rows of that bin in earlier segments.
__global__ void k_order_scan2(PassArgs a)
{
    __shared__ uint32_t tot[BFJ_PIVOTS];
    const int si = blockIdx.x, bin = threadIdx.x;
    const SetDesc &s = a.sets[si];
    if (!ord_active(s)) return;
    const int nseg = (a.set_nblk[si] + ORD_SEG - 1) / ORD_SEG;
    uint32_t *seg = a.oseg + (int64_t)si * a.oseg_stride * BFJ_PIVOTS;
    uint32_t t = 0;
    for (int g = 0; g < nseg; ++g) t += seg[g * BFJ_PIVOTS + bin];
    tot[bin] = t;
    __syncthreads();
    uint32_t run = 0;
    for (int q = 0; q < bin; ++q) run += tot[q];
    for (int g = 0; g < nseg; ++g) {
        const uint32_t c = seg[g * BFJ_PIVOTS + bin];
        seg[g * BFJ_PIVOTS + bin] = run;
        run += c;
    }
}

// Scan pass 3: within each segment, every (block, bin) count becomes that block's
// first position for the bin.  Grid (segments, datasets), one thread per bin.
__global__ void k_order_scan3(PassArgs a)
{
    const int si = blockIdx.y, g = blockIdx.x, bin = threadIdx.x;
    const SetDesc &s = a.sets[si];
    const int b0 = a.set_blk0[si] + g * ORD_SEG, b1 = a.set_blk0[si] + a.set_nblk[si];
    if (!ord_active(s) || b0 >= b1) return;
    uint32_t run = a.oseg[((int64_t)si * a.oseg_stride + g) * BFJ_PIVOTS + bin];
    for (int b = b0; b < b0 + ORD_SEG && b < b1; ++b) {
        uint32_t *h = a.ohist + (int64_t)b * BFJ_PIVOTS + bin;
        const uint32_t c = *h;
        *h = run;
        run += c;
    }
}

// Position of every row: its (block, bin) start plus its rank among the block's rows
// with that bin (rows in ascending order).  Padding positions [n, npad) and datasets
// that are not reordered map to themselves.  One thread per row of a 64-row block.
__global__ void k_order_scatter(PassArgs a)
{
    __shared__ uint8_t lab[BFJ_PREP_ROWS];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const int t = threadIdx.x;
    const int64_t r = a.blk_row0[b] + t;
    if (!ord_active(s)) {
        if (r < s.npad) a.perm[s.row_off + r] = (uint32_t)r;
        return;
    }
    const bool real = r < s.n;
    lab[t] = real ? a.label[s.row_off + r] : 0;
    __syncthreads();
    if (r >= s.npad) return;
    if (!real) {
        a.perm[s.row_off + r] = (uint32_t)r;
        return;
    }
    const int L = lab[t];
    uint32_t rank = 0;
    for (int u = 0; u < t; ++u) rank += lab[u] == L ? 1u : 0u; // rows u < t of this block are all real
    const uint32_t pos = a.ohist[(int64_t)b * BFJ_PIVOTS + L] + rank;
    a.perm[s.row_off + pos] = (uint32_t)r;
}

void bfj_launch_order(const PassArgs &a, hipStream_t st)
{
    const int max_seg = a.oseg_stride;
    hipLaunchKernelGGL(k_order_pivots, dim3(a.n_sets), dim3(BFJ_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_order_label, dim3(a.n_row_blocks), dim3(BFJ_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_order_scan1, dim3(max_seg, a.n_sets), dim3(BFJ_PIVOTS), 0, st, a);
    hipLaunchKernelGGL(k_order_scan2, dim3(a.n_sets), dim3(BFJ_PIVOTS), 0, st, a);
    hipLaunchKernelGGL(k_order_scan3, dim3(max_seg, a.n_sets), dim3(BFJ_PIVOTS), 0, st, a);
    hipLaunchKernelGGL(k_order_scatter, dim3(a.n_row_blocks), dim3(BFJ_PREP_ROWS), 0, st, a);
}
