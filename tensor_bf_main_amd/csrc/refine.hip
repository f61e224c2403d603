// refine.hip -- classification of the join's candidates, FP64 band refine and
// count finalisation.
//
// k_classify reads the candidate stream the join wrote (pairs with Z >= 0) and
// applies the per-epsilon tests of the error bound: certain-in pairs are counted,
// pairs inside the band go to the refine list.
//
// Every pair the join kernel could not decide is re-evaluated here in the
// reference's own arithmetic: the direct form of distanceCalculationBruteForceCuda
// (kernel_join.cu:184-189), t = q[k] - c[k]; acc = fma(t, t, acc) in ascending k,
// compared with T, the largest double whose square root is <= eps, which is
// the predicate sqrt(acc) <= eps of kernel_join.cu:194.  These are the only
// FP64 steps of the low-precision paths besides the threshold computation.
//
// Layout of the work: one wavefront takes 64 band records, one per lane.  The
// two rows of every record are streamed through LDS in chunks of RCH doubles:
// the wavefront loads the 128 row segments of a chunk with coalesced 8-byte
// loads (16 lanes per 128-byte segment) and stores them transposed,
// lds[k][pair], so that each lane then runs its own sequential FMA chain with
// conflict-free LDS reads.  One HBM/L2 line per row segment instead of one per
// lane per word.
#include "bfj_internal.h"
#include <math.h>

// ---- classification of the join's candidate stream --------------------------------
// One wavefront per chunk.  Per record, row and eps: certainly within eps is
// counted, inside the band goes to the refine list, otherwise out.  A record of a
// symmetric sweep decides for its row and for its candidate row.  Counts are summed
// per distinct row of the wavefront before one atomic per row and eps.
__device__ inline int set_of_row(const PassArgs &a, uint32_t row)
{
    int lo = 0, hi = a.n_sets - 1;
    while (lo < hi) { // last set with row_off <= row
        const int mid = (lo + hi + 1) >> 1;
        if ((uint64_t)a.sets[mid].row_off <= row) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// cnt[row][e] += 1 for every lane and every eps bit of `mask`; lanes with the same
// global row are summed first (a vote loop over the distinct rows of the wavefront).
// `li` is the dataset-local row of `key` in dataset `s`.  All lanes take part.
__device__ inline void add_counts(const PassArgs &a, uint32_t key, const SetDesc &s, uint32_t li, uint32_t mask,
                                  int lane)
{
    unsigned long long pending = __ballot(mask != 0u);
    if (!pending) return;
    uint32_t any = mask;
    for (int o = 32; o > 0; o >>= 1) any |= (uint32_t)__shfl_xor((int)any, o);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t k = (uint32_t)__shfl((int)key, leader);
        const bool same = key == k;
        for (int e = 0; e < BFJ_MAXEPS; ++e) {
            if (!((any >> e) & 1u)) continue;
            const unsigned long long m = __ballot(same && ((mask >> e) & 1u));
            if (m && lane == leader) atomicAdd(&a.cnt[s.cnt_off + (int64_t)e * s.npad + li], (uint32_t)__popcll(m));
        }
        pending &= ~__ballot(same);
    }
}

// Decide the eps of one row from the bounds [lo, hi] of the pair's squared distance.
__device__ inline void decide(const double2 *b, int n_eps, double lo, double hi, uint32_t &in, uint32_t &band)
{
    for (int e = 0; e < n_eps; ++e) {
        const double2 p = b[e];
        if (hi <= p.x) in |= 1u << e;
        else if (lo <= p.y) band |= 1u << e;
    }
}

__global__ __launch_bounds__(BFJ_BLOCK) void k_classify(PassArgs a)
{
    const int lane = threadIdx.x & 63;
    const uint32_t n_chunks = min(*a.chunk_count, a.chunk_capacity);
    const uint32_t wave_id = blockIdx.x * (BFJ_BLOCK / BFJ_WAVE) + (threadIdx.x >> 6);
    const uint32_t n_waves = gridDim.x * (BFJ_BLOCK / BFJ_WAVE);
    for (uint32_t ch = wave_id; ch < n_chunks; ch += n_waves) {
        const uint32_t fill = min(a.chunk_fill[ch], (uint32_t)BFJ_CHUNK);
        for (uint32_t t0 = 0; t0 < fill; t0 += BFJ_WAVE) {
            const uint32_t t = t0 + lane;
            const bool live = t < fill;
            Candidate c;
            c.row = 0;
            c.col = 0;
            c.z = -1.0;
            if (live) c = a.cand[(size_t)ch * BFJ_CHUNK + t];
            const int si = live ? (a.n_sets == 1 ? 0 : set_of_row(a, c.row)) : 0;
            const SetDesc &s = a.sets[si];
            const uint32_t i = c.row - (uint32_t)s.row_off;
            const uint32_t j = c.col & ~BFJ_COL_SYM;
            uint32_t in_i = 0, band_i = 0, in_j = 0, band_j = 0;
            if (live) {
                // S_hat = 2 g_i + N_i - 2 Z, |S_hat - S| <= E_i; mg covers the FP64
                // rounding of these few operations.
                const double2 h = a.he[c.row];
                const double shat = h.x - 2.0 * c.z;
                const double mg = (fabs(h.x) + 2.0 * fabs(c.z) + h.y) * 0x1p-50;
                const double lo = shat - h.y - mg, hi = shat + h.y + mg;
                decide(a.lh + s.lh_off + (int64_t)i * s.n_eps, s.n_eps, lo, hi, in_i, band_i);
                if (c.col & BFJ_COL_SYM) decide(a.lh + s.lh_off + (int64_t)j * s.n_eps, s.n_eps, lo, hi, in_j, band_j);
            }
            add_counts(a, c.row, s, i, in_i, lane);
            add_counts(a, (uint32_t)s.row_off + j, s, j, in_j, lane);
            // Band pairs to the refine list: one atomic per wavefront.
            const bool band = (band_i | band_j) != 0u;
            const unsigned long long bb = __ballot(band);
            if (bb) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(a.band_count, (uint32_t)__popcll(bb));
                base = __shfl(base, 0);
                if (band) {
                    const uint32_t idx = base + (uint32_t)__popcll(bb & ((1ull << lane) - 1ull));
                    if (idx < a.band_capacity) {
                        BandEntry be;
                        be.i = i;
                        be.j = j;
                        be.z = (float)c.z;
                        be.meta = band_i | (band_j << 8) | ((uint32_t)si << 16);
                        a.band[idx] = be;
                    }
                }
            }
        }
    }
}

void bfj_launch_classify(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_classify, dim3(2048), dim3(BFJ_BLOCK), 0, st, a);
}

#define RCH 16                 // doubles per row chunk (one 128-byte line)
#define RPAD (BFJ_WAVE + 1)    // LDS row length: 65 doubles keeps both phases conflict-free
#define RWAVES 2              // wavefronts per block: 2 x 17 KiB of staged rows

__global__ __launch_bounds__(RWAVES * BFJ_WAVE) void k_refine(PassArgs a)
{
    __shared__ double sq[RWAVES][RCH][RPAD];
    __shared__ double sc[RWAVES][RCH][RPAD];
    __shared__ const double *sptr[RWAVES][2 * BFJ_WAVE];
    __shared__ int sd[RWAVES][BFJ_WAVE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t cnt = *a.band_count;
    const uint32_t m = cnt < a.band_capacity ? cnt : a.band_capacity;
    const uint32_t groups = (m + BFJ_WAVE - 1) / BFJ_WAVE;
    const uint32_t wave_id = blockIdx.x * RWAVES + w, n_waves = gridDim.x * RWAVES;
    for (uint32_t g = wave_id; g < groups; g += n_waves) {
        const uint32_t t = g * BFJ_WAVE + lane;
        const bool live = t < m;
        BandEntry be;
        be.i = be.j = 0;
        be.z = 0.f;
        be.meta = 0;
        if (live) be = a.band[t];
        const int si = (int)(be.meta >> 16);
        const SetDesc &s = a.sets[si];
        const int d = live ? s.d : 0;
        sptr[w][2 * lane] = s.x + (int64_t)be.i * s.ld;
        sptr[w][2 * lane + 1] = s.x + (int64_t)be.j * s.ld;
        sd[w][lane] = d;
        int dmax = d;
        for (int o = 32; o > 0; o >>= 1) dmax = max(dmax, __shfl_xor(dmax, o));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double acc = 0.0;
        for (int k0 = 0; k0 < dmax; k0 += RCH) {
            // Load: segment sg = (pair p, row r) covers k0..k0+RCH-1 of that row;
            // 4 segments per instruction, 16 lanes each.
#pragma unroll 4
            for (int it = 0; it < 2 * BFJ_WAVE / 4; ++it) {
                const int sg = it * 4 + (lane >> 4);
                const int p = sg >> 1, r = sg & 1;
                const int k = k0 + (lane & 15);
                const double v = (k < sd[w][p]) ? sptr[w][sg][k] : 0.0;
                if (r == 0) sq[w][lane & 15][p] = v;
                else sc[w][lane & 15][p] = v;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int kend = min(RCH, d - k0);
            for (int kk = 0; kk < kend; ++kk) {
                const double df = sq[w][kk][lane] - sc[w][kk][lane];
                acc = fma(df, df, acc);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        double err = -1.0;
        // Hits per (row, eps) for both rows of the record; lanes of the same row are
        // summed first (records of one group share few rows).
        uint32_t hits_i = 0, hits_j = 0;
        if (live) {
            for (int e = 0; e < s.n_eps; ++e)
                if (acc <= s.T[e]) {
                    hits_i |= be.meta & (1u << e);
                    hits_j |= (be.meta >> 8) & (1u << e);
                }
        }
        add_counts(a, (uint32_t)s.row_off + be.i, s, be.i, hits_i, lane);
        add_counts(a, (uint32_t)s.row_off + be.j, s, be.j, hits_j, lane);
        if (live) {
            const double sigma = a.scal[si].sigma;
            const double approx = sigma * sigma * (a.he[s.row_off + be.i].x - 2.0 * (double)be.z);
            err = fabs(approx - acc);
        }
        // Largest approximation error of the group: one atomic per dataset present.
        const int si0 = __shfl(si, 0);
        if (__ballot(live && si != si0) == 0ull) {
            double mx = err;
            for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
            if (lane == 0 && mx >= 0.0) atomicMax(&a.scal[si0].err_bits, (unsigned long long)__double_as_longlong(mx));
        } else if (err >= 0.0) {
            atomicMax(&a.scal[si].err_bits, (unsigned long long)__double_as_longlong(err));
        }
    }
}

void bfj_launch_refine(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_refine, dim3(4096), dim3(RWAVES * BFJ_WAVE), 0, st, a);
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
            // "all" eps: every candidate, self included, is within eps (bounded in
            // k_thresholds); the join did not count them.
            const bool all = (a.scal[si].all_mask >> e) & 1ull;
            c = all ? (unsigned long long)s.n : a.cnt[s.cnt_off + (int64_t)e * s.npad + i];
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
