// join_kernel.h -- the MFMA epsilon-join kernel, templated on dot-product dtype
// (Tr), K-chunks (KC) and query rows held per wavefront (R).
//
// Replaces the reference's join kernels kernel_join.cu:328-524 (WMMA half,
// "multi-query") and kernel_join.cu:532-675 (DMMA double), and the tiled GEMM
// samples they were meant to grow into (wmmaTensorCoresGemm.cu:97-296,
// dmmaTensorCoresGemm.cu:113-287), with one gfx950 design:
//
//   * A wavefront owns a strip of R query rows.  Their operand fragments stay
//     in VGPRs for the whole pass (R/16 row blocks x KC chunks).
//   * It streams every candidate block of 16 rows: KC fragment loads, each one
//     coalesced 1-2 KiB read of the pre-packed operand array.
//   * The accumulator of each 16x16 tile starts at C0 = g_i - n_j, so after
//     the MFMAs Z = 0.5*(H_i - d2_approx(i, j)) with no further arithmetic: the
//     squared norms and the threshold are folded into the initial value
//     (the reference pre-loads candidate chunk norms the same way,
//     kernel_join.cu:433-437).
//   * Z < 0 means "certainly out".  One max-reduction over the tile and one
//     wave vote skip almost every tile.  Tiles with some Z >= 0 go to the slow
//     path: certain-in pairs bump per-row counters in LDS, pairs inside the
//     error band are appended to the FP64 refine list (one atomic per wave).
//   * Per-row counts are flushed with one global atomic per row per strip,
//     instead of one atomic per hit (kernel_join.cu:484, :662).
#pragma once
#include "bfj_internal.h"

struct TrF16 {
    static constexpr bool prefetch = true; // double-buffer the candidate fragments
    static constexpr int dtype = BFJ_FP16;
    typedef bfj_h8 frag_t;
    typedef bfj_f4 acc_t;
    typedef float val_t;
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int, int)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    // Output register r of lane l holds tile row 4*(l/16)+r, column l%16.
    __device__ static inline int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct TrBF16 {
    static constexpr bool prefetch = true; // double-buffer the candidate fragments
    static constexpr int dtype = BFJ_BF16;
    typedef bfj_b8 frag_t;
    typedef bfj_f4 acc_t;
    typedef float val_t;
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int, int)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct TrF8 {
    static constexpr bool prefetch = true; // double-buffer the candidate fragments
    static constexpr int dtype = BFJ_FP8;
    typedef bfj_i8 frag_t;
    typedef bfj_f4 acc_t;
    typedef float val_t;
    // Operand format 0 = FP8 e4m3 for A and B; sa/sb are E8M0 row scales.
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int sa, int sb)
    {
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
    }
    __device__ static inline int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct TrF64 {
    static constexpr bool prefetch = false; // 64-cycle MFMAs hide the loads of one block
    static constexpr int dtype = BFJ_FP64;
    typedef bfj_d2 frag_t; // two K=4 steps of v_mfma_f64_16x16x4
    typedef bfj_d4 acc_t;
    typedef double val_t;
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int, int)
    {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[1], c, 0, 0, 0);
        return c;
    }
    // The FP64 shape stores tile row 4*r + l/16 in register r of lane l.
    __device__ static inline int out_row(int lane, int r) { return 4 * r + (lane >> 4); }
};

template <class V> struct Pair2;
template <> struct Pair2<float> { typedef float2 T; };
template <> struct Pair2<double> { typedef double2 T; };

__device__ inline int bfj_lane() { return __lane_id(); }

template <class Tr, int KC, int R>
__device__ __forceinline__ void join_strip(const JoinArgs &a, const SetDesc &s, int rb0, uint32_t *lcnt,
                                           typename Tr::val_t *zt)
{
    typedef typename Tr::frag_t frag_t;
    typedef typename Tr::acc_t acc_t;
    typedef typename Tr::val_t val_t;
    typedef typename Pair2<val_t>::T pair_t;
    constexpr int RB = R / 16;
    const int lane = bfj_lane();
    const int64_t nblk = s.npad / 16;
    const frag_t *P = reinterpret_cast<const frag_t *>(a.pack + s.pack_off);
    const val_t *nq = reinterpret_cast<const val_t *>(a.nq) + s.row_off;
    const val_t *gq = reinterpret_cast<const val_t *>(a.gq) + s.row_off;
    const pair_t *ov = reinterpret_cast<const pair_t *>(a.ov) + s.ov_off;
    const uint8_t *sc = (Tr::dtype == BFJ_FP8) ? a.scales + s.row_off : nullptr;

    // Strip fragments and per-row initial values, resident for the whole pass.
    frag_t A[RB][KC];
    val_t G[RB][4];
    int SA[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int64_t blk = rb0 + rb;
        const bool ok = blk < nblk;
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            if (ok) A[rb][c] = P[(blk * KC + c) * BFJ_WAVE + lane];
            else A[rb][c] = frag_t{};
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = blk * 16 + Tr::out_row(lane, r);
            G[rb][r] = ok ? gq[row] : (val_t)-1e30;
        }
        SA[rb] = (Tr::dtype == BFJ_FP8 && ok) ? (int)sc[blk * 16 + (lane & 15)] : 127;
    }

    const int n_eps = s.n_eps;
    frag_t B[KC];
    val_t nj = nq[lane & 15];
    int SB = (Tr::dtype == BFJ_FP8) ? (int)sc[lane & 15] : 127;
    if constexpr (Tr::prefetch) {
#pragma unroll
        for (int c = 0; c < KC; ++c) B[c] = P[c * BFJ_WAVE + lane];
    }

    for (int64_t cb = 0; cb < nblk; ++cb) {
        // Prefetch the next candidate block (fragments, norms, scales) while this
        // one is in the MFMA pipe.
        frag_t Bn[Tr::prefetch ? KC : 1];
        const int64_t nb = (cb + 1 < nblk) ? cb + 1 : cb;
        const val_t njn = nq[nb * 16 + (lane & 15)];
        const int SBn = (Tr::dtype == BFJ_FP8) ? (int)sc[nb * 16 + (lane & 15)] : 127;
        if constexpr (Tr::prefetch) {
#pragma unroll
            for (int c = 0; c < KC; ++c) Bn[c] = P[(nb * KC + c) * BFJ_WAVE + lane];
        } else {
#pragma unroll
            for (int c = 0; c < KC; ++c) B[c] = P[(cb * KC + c) * BFJ_WAVE + lane];
        }

        acc_t acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[rb][r] = G[rb][r] - nj;
#pragma unroll
        for (int c = 0; c < KC; ++c)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = Tr::mma(A[rb][c], B[c], acc[rb], SA[rb], SB);

        // Inputs are NaN-free (the pass rejects non-finite rows), so with
        // -fno-honor-nans this is a tree of v_max3.  -0.0 >= 0 holds, so a zero Z
        // of either sign reaches the exact slow-path tests.
        val_t m = acc[0][0];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) m = fmax(m, acc[rb][r]);

        if (__builtin_expect(__ballot(m >= (val_t)0) != 0ull, 0)) {
            // Slow path, taken by the few tiles where some pair is not certainly
            // out: park only the hot 16x16 tiles in this wavefront's LDS (one
            // 16-byte store per lane each) and walk them with a rolled loop, so
            // the rare path costs neither registers nor code in the fast one.
            uint32_t hot = 0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const val_t mr = fmax(fmax(acc[rb][0], acc[rb][1]), fmax(acc[rb][2], acc[rb][3]));
                if (__ballot(mr >= (val_t)0) != 0ull) {
                    hot |= 1u << rb;
                    *reinterpret_cast<acc_t *>(zt + (rb * BFJ_WAVE + lane) * 4) = acc[rb];
                }
            }
            const int64_t col = cb * 16 + (lane & 15);
            const bool col_ok = col < s.n;
            hot = (uint32_t)__builtin_amdgcn_readfirstlane((int)hot);
#pragma unroll 1
            for (int q = 0; q < RB * 4; ++q) {
                if (!((hot >> (q >> 2)) & 1u)) {
                    q |= 3; // skip the rest of this row block
                    continue;
                }
                const val_t z = zt[((q >> 2) * BFJ_WAVE + lane) * 4 + (q & 3)];
                const bool cand = z >= (val_t)0 && col_ok;
                if (__ballot(cand) == 0ull) continue;
                const int li = (q >> 2) * 16 + Tr::out_row(lane, q & 3);
                const int64_t i = (int64_t)rb0 * 16 + li;
                uint32_t band = 0;
                if (cand) {
                    for (int e = 0; e < n_eps; ++e) {
                        const pair_t t = ov[i * n_eps + e];
                        if (z >= t.y) atomicAdd(&lcnt[e * R + li], 1u);
                        else if (z >= t.x) band |= 1u << e;
                    }
                }
                const unsigned long long bal = __ballot(band != 0u);
                if (bal) {
                    const int leader = __ffsll((long long)bal) - 1;
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(a.band_count, (uint32_t)__popcll(bal));
                    base = __shfl(base, leader);
                    if (band) {
                        const uint32_t idx = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                        if (idx < a.band_capacity) {
                            BandEntry be;
                            be.i = (uint32_t)i;
                            be.j = (uint32_t)col;
                            be.z = (float)z;
                            be.meta = band | ((uint32_t)s.set_id << 16);
                            a.band[idx] = be;
                        }
                    }
                }
            }
        }
        nj = njn;
        SB = SBn;
        if constexpr (Tr::prefetch) {
#pragma unroll
            for (int c = 0; c < KC; ++c) B[c] = Bn[c];
        }
    }

    // Flush per-row counts of the strip: one global atomic per non-zero row.
    __threadfence_block();
    for (int t = lane; t < n_eps * R; t += BFJ_WAVE) {
        const uint32_t v = lcnt[t];
        if (v) {
            const int e = t / R, li = t % R;
            const int64_t i = (int64_t)rb0 * 16 + li;
            atomicAdd(&a.cnt[s.cnt_off + (int64_t)e * s.npad + i], v);
            lcnt[t] = 0;
        }
    }
    __threadfence_block();
}

// Dynamic LDS per wavefront: BFJ_MAXEPS x R row counters, then the R x 16 tile
// of Z values parked by the slow path.
template <class Tr, int R>
__host__ __device__ constexpr int join_lds_per_wave()
{
    return BFJ_MAXEPS * R * 4 + R * 16 * (int)sizeof(typename Tr::val_t);
}

// Persistent kernel: every wavefront takes strips from a global queue until it
// is empty; the grid may also be sized to one wavefront per strip.
template <class Tr, int KC, int R>
__global__ __launch_bounds__(BFJ_BLOCK) void k_join(JoinArgs a)
{
    extern __shared__ unsigned char lds_raw[];
    const int wave = threadIdx.x / BFJ_WAVE;
    const int lane = bfj_lane();
    unsigned char *mine = lds_raw + wave * join_lds_per_wave<Tr, R>();
    uint32_t *lcnt = reinterpret_cast<uint32_t *>(mine);
    typename Tr::val_t *zt = reinterpret_cast<typename Tr::val_t *>(mine + BFJ_MAXEPS * R * 4);
    for (int t = lane; t < BFJ_MAXEPS * R; t += BFJ_WAVE) lcnt[t] = 0;
    __threadfence_block();
    for (;;) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(a.item_counter, 1u);
        // Wave-uniform from here on: descriptor loads become scalar loads.
        item = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(item, 0));
        if (item >= (uint32_t)a.n_items) break;
        const WorkItem w = a.items[item];
        const SetDesc s = a.sets[w.set];
        join_strip<Tr, KC, R>(a, s, w.rb0, lcnt, zt);
    }
}
