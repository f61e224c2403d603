// join_kernel.h -- the MFMA epsilon-join kernel, templated on dot-product dtype
// (Tr), K-chunks (KC) and query rows held per wavefront (R).
//
// Replaces the reference's join kernels kernel_join.cu:328-524 (WMMA half,
// "multi-query") and kernel_join.cu:532-675 (DMMA double), and the tiled GEMM
// samples they were meant to grow into (wmmaTensorCoresGemm.cu:97-296,
// dmmaTensorCoresGemm.cu:113-287), with one gfx950 design:
//
//   * A block of 4 wavefronts owns a strip of 4R query rows; each wavefront
//     keeps the operand fragments of its R rows in VGPRs for the whole pass.
//   * The block streams every candidate block of 16 rows through an NB-deep
//     LDS ring filled by direct-to-LDS loads (global_load_lds): one coalesced
//     load per block instead of one per wavefront (the reference reads
//     candidates from global memory in every warp, kernel_join.cu:432), no
//     VGPRs spent on staging, one barrier per candidate block.
//   * The first MFMA of each row block takes g_i as its C operand, so after the
//     MFMAs acc = g_i + xhat_i . xhat_j and Z = acc - n_j = 0.5*(2 g_i + N_i -
//     d2_approx(i, j)): the squared norms and the threshold are folded in with
//     no VALU work before the MFMAs, and the test Z >= 0 is acc >= n_j (the
//     reference pre-loads candidate chunk norms, kernel_join.cu:433-437).
//     For FP8, n_j is the candidate threshold k_thresholds writes (prep.hip): half
//     the squared norm lowered by the row's share of the quantisation margin, so
//     that a row with a small cast error rejects more.
//   * Z < 0 means "certainly out" for every epsilon.  One max tree and one wave
//     vote skip almost every tile.  The pairs of a hot tile with Z >= 0 are
//     appended to a candidate stream (row, col, Z) with wave-compacted plain
//     stores into per-wavefront chunks of 4096 records (one atomic per chunk),
//     so the hot path never waits on memory or on another wavefront; the
//     per-epsilon tests, counting and band selection run afterwards in
//     k_classify (refine.hip).  The reference instead does one global atomic
//     per hit inside the join (kernel_join.cu:484, :662).
#pragma once
#include "bfj_internal.h"
#include "mfma_ops.h"

__device__ inline int bfj_lane() { return __lane_id(); }

#ifndef BFJ_SLOT_BYTES
#define BFJ_SLOT_BYTES 16384 // at most this many candidate fragment bytes per ring slot
#endif
#ifndef BFJ_SLOT_MFMA
#define BFJ_SLOT_MFMA 128 // MFMAs per wavefront per ring slot to aim for
#endif
#ifndef BFJ_DIRECT_B
#define BFJ_DIRECT_B 0 // 1: candidate fragments read straight into VGPRs, no LDS ring
#endif
#ifndef BFJ_RING_BYTES
#define BFJ_RING_BYTES (24 * 1024) // target LDS bytes of the candidate ring
#endif

// LDS layout of one join block (bytes), shared by the kernel and the host.
template <class Tr, int KC, int R>
struct JoinLds {
    typedef typename Tr::val_t val_t;
    static constexpr int FB = (int)sizeof(typename Tr::frag_t);
    static constexpr int TILE = KC * BFJ_WAVE * FB; // candidate fragments of one 16-row block
    // A ring slot holds CPT consecutive candidate blocks: one wait, one barrier and
    // one set of loads per slot instead of per block.  Enough blocks that each
    // wavefront issues about BFJ_SLOT_MFMA MFMAs (R/16 x KC per block) between
    // barriers, within BFJ_SLOT_BYTES.  Measured at FP16 R = 128, d = 128: 4 blocks
    // (128 MFMAs) per slot take 1.4 % off the join against 2.
    static constexpr int CPT_MFMA = BFJ_SLOT_MFMA / ((R / 16) * KC) < 1 ? 1 : BFJ_SLOT_MFMA / ((R / 16) * KC);
    static constexpr int CPT_BYTES = BFJ_SLOT_BYTES / TILE;
    static constexpr int CPT_RAW = CPT_MFMA < CPT_BYTES ? CPT_MFMA : CPT_BYTES;
    // (FP64 slots hold at most 2 blocks, so a slot's norms fit one 64-lane load.)
    static constexpr int CPT_MAX = (int)sizeof(val_t) == 8 ? 2 : 4;
    // 1, 2 or 4: panels and strips are multiples of 4 blocks, so items are whole slots.
    static constexpr int CPT_CLAMP = CPT_RAW < 1 ? 1 : (CPT_RAW > CPT_MAX ? CPT_MAX : CPT_RAW);
    static constexpr int CPT = CPT_CLAMP >= 4 ? 4 : (CPT_CLAMP >= 2 ? 2 : 1);
    static constexpr int UNITS = CPT * TILE / 16; // 16-byte units of a slot
    // Direct-to-LDS loads per wavefront per slot: the fragment area is padded to a
    // whole number of block-wide 16-byte loads so every wavefront issues the same
    // count, which keeps the vmcnt bookkeeping a compile-time constant.
    static constexpr int UPW = (UNITS + BFJ_JBLOCK - 1) / BFJ_JBLOCK;
    static constexpr int FRAG = UPW * BFJ_JBLOCK * 16;
    static constexpr int NQB = CPT * 16 * (int)sizeof(val_t); // candidate norms of the slot
    static constexpr int BUF = FRAG + NQB + CPT * 16;          // + FP8 scales
    // Ring depth: about 24 KiB per block, at least double buffering.  Measured: a
    // deeper ring buys less than the resident blocks its LDS would cost.
    static constexpr int NB_RAW = BFJ_RING_BYTES / BUF;
    static constexpr int NB = NB_RAW < 2 ? 2 : (NB_RAW > 16 ? 16 : NB_RAW);
    // DMA instructions per slot: every wavefront issues UPW fragment loads; wavefront 0
    // also loads the norms and the last wavefront the FP8 scales.
    static constexpr int LPT_MAX = UPW + 1 + (Tr::dtype == BFJ_FP8 ? 1 : 0);
    static constexpr int RING = NB * BUF;
    __host__ __device__ static constexpr int total(int) { return BFJ_DIRECT_B ? 16 : RING; }
};

// Per-wavefront cursor into the candidate stream: chunks of BFJ_CHUNK records
// are handed out by one atomic each; the records themselves are plain stores
// that nothing waits on.
struct CandCursor {
    Candidate *dst; // next slot of the current chunk (uniform)
    uint32_t chunk; // current chunk (0xffffffff: none yet)
    uint32_t fill;  // records written into it
    bool live;      // the chunk lies within the capacity (else its records are dropped)
};

__device__ inline void cand_close(const JoinArgs &a, const CandCursor &cc, int lane)
{
    if (cc.chunk != 0xffffffffu && cc.chunk < a.chunk_capacity && lane == 0) a.chunk_fill[cc.chunk] = cc.fill;
}

// Append the pairs of one 16x16 output block with acc >= n_j (Z >= 0) as candidate
// records, wave-compacted: one lane mask per output register (a single compare each),
// at most one chunk grab, then one 16-byte store per pair at the chunk's uniform base
// plus a 32-bit lane offset.  The caller has set the accumulators of the pairs not to
// stream to -inf (in a symmetric sweep only pairs above the diagonal, col > row, are
// kept, since the tile that holds (j, i) is dropped and self pairs are counted in
// k_finalize).  row0 is the global row of the block's first row, colrec the record's
// column word.
template <class Tr>
__device__ __forceinline__ void cand_push_block(const JoinArgs &a, CandCursor &cc, const typename Tr::acc_t &v,
                                                typename Tr::val_t nj, uint32_t row0, uint32_t colrec, int lane)
{
    unsigned long long bal[4];
    uint32_t off[4], nrec = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        bal[r] = __builtin_amdgcn_ballot_w64(v[r] >= nj);
        off[r] = nrec;
        nrec += (uint32_t)__popcll(bal[r]);
    }
    if (nrec == 0) return;
    if (cc.fill + nrec > BFJ_CHUNK) { // nrec <= 4 x 64 fits a fresh chunk
        cand_close(a, cc, lane);
        // 64-bit counter, saturated to a 32-bit index: once past the capacity every
        // later chunk is dropped, never written over an earlier one.
        uint32_t c = 0;
        if (lane == 0) {
            const unsigned long long c64 = atomicAdd(a.chunk_count, 1ull);
            c = c64 < 0xfffffffeull ? (uint32_t)c64 : 0xfffffffeu;
        }
        cc.chunk = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(c, 0));
        cc.fill = 0;
        cc.live = cc.chunk < a.chunk_capacity;
        cc.dst = a.cand + (size_t)(cc.live ? cc.chunk : 0u) * BFJ_CHUNK;
    }
    if (cc.live) {
        Candidate *dst = cc.dst;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (v[r] >= nj) { // this lane's bit of bal[r]
                const uint32_t k = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(bal[r] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal[r], off[r]));
                Candidate rec;
                rec.row = row0 + (uint32_t)Tr::out_row(lane, r);
                rec.col = colrec;
                rec.z = (double)(v[r] - nj);
                dst[k] = rec;
            }
        }
    }
    cc.dst += nrec;
    cc.fill += nrec;
}

template <class Tr, int KC, int R>
__device__ __forceinline__ void join_strip(const JoinArgs &a, const SetDesc &s, int rb0, int cb0, int cb1,
                                           unsigned char *lds,
                                           CandCursor &cc)
{
    typedef typename Tr::frag_t frag_t;
    typedef typename Tr::acc_t acc_t;
    typedef typename Tr::val_t val_t;
    typedef JoinLds<Tr, KC, R> L;
    constexpr int RB = R / 16;
    const int lane = bfj_lane();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / BFJ_WAVE));
    const int64_t nblk = s.npad / 16;
    const int wrb0 = rb0 + wave * RB; // first row block of this wavefront
    const unsigned char *P = a.pack + s.pack_off; // fragment f starts at byte f * 64 * FB
    constexpr int FB = (int)sizeof(frag_t);
    const val_t *nq = reinterpret_cast<const val_t *>(a.nq) + s.row_off;
    const val_t *gq = reinterpret_cast<const val_t *>(a.gq) + s.row_off;
    const uint8_t *sc = (Tr::dtype == BFJ_FP8) ? a.scales + s.row_off : nullptr;

    // Fragments and initial values of this wavefront's rows, resident in VGPRs.
    frag_t A[RB][KC];
    acc_t G[RB]; // g_i per output register: the C operand of each row block's first MFMA
    int SA[RB];
    bool any_rows = false;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int64_t blk = wrb0 + rb;
        const bool ok = blk < nblk;
        any_rows |= ok;
#pragma unroll
        for (int c = 0; c < KC; ++c) A[rb][c] = ok ? Tr::load(P + (blk * KC + c) * BFJ_WAVE * FB, lane) : frag_t{};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = blk * 16 + Tr::out_row(lane, r);
            G[rb][r] = ok ? gq[row] : (val_t)-1e30;
        }
        SA[rb] = (Tr::dtype == BFJ_FP8 && ok) ? (int)sc[blk * 16 + (lane & 15)] : 127;
    }

    // Candidate tiles stream through an NB-deep LDS ring filled by direct-to-LDS
    // loads (global_load_lds, no VGPRs), NB-1 slots ahead of the MFMAs.  Wavefront
    // w loads units w*64 + JBLOCK*u + lane of each slot (units past the slot re-read
    // unit 0 into the padding); wavefront 0 also loads the slot's norms and (FP8)
    // the last wavefront its scales.
    typedef __attribute__((address_space(3))) void lds_void;
    const unsigned char *srcb = P;
    const unsigned char *nqb = reinterpret_cast<const unsigned char *>(nq);
    static_assert(L::NQB <= 4 * BFJ_WAVE, "one 64-lane load carries a slot's norms");
    const bool nq_wave = wave == 0;
    const bool sc_wave = Tr::dtype == BFJ_FP8 && wave == BFJ_JWAVES - 1;
    // Ring unit U of a slot (16 bytes, in LDS order) is half h of lane operand l of
    // fragment f (mfma_ops.h: the ring keeps fragments in 16-byte halves), which the
    // packed layout holds at unit f * 64H + l H + h, H = FB / 16 halves per operand.
    uint32_t voff[L::UPW];                   // this lane's byte offset in a slot, per load
#pragma unroll
    for (int u = 0; u < L::UPW; ++u) {
        constexpr int H = FB / 16;
        const int unit = u * BFJ_JBLOCK + wave * BFJ_WAVE + lane;
        const int f = unit / (BFJ_WAVE * H), w = unit % (BFJ_WAVE * H);
        const int src = f * BFJ_WAVE * H + (w % BFJ_WAVE) * H + w / BFJ_WAVE;
        voff[u] = (uint32_t)(unit < L::UNITS ? src : 0) * 16u;
    }
    // Slot t of the item holds blocks cb0 + t*CPT ... + CPT-1; t and slot are uniform,
    // so the addresses are a scalar base plus a lane offset.  Blocks past the
    // dataset's last one read padding that the buffers carry for this (never used).
    auto issue = [&](int t, int slot) {
        unsigned char *b = lds + slot * L::BUF;
        const int cbase = cb0 + t * L::CPT;
        const unsigned char *gt = srcb + (size_t)cbase * L::TILE;
#pragma unroll
        for (int u = 0; u < L::UPW; ++u)
            __builtin_amdgcn_global_load_lds(gt + voff[u], (lds_void *)(b + (u * BFJ_JBLOCK + wave * BFJ_WAVE) * 16), 16,
                                             0, 0);
        if (nq_wave && lane < L::NQB / 4)
            __builtin_amdgcn_global_load_lds(nqb + (size_t)cbase * 16 * sizeof(val_t) + lane * 4,
                                             (lds_void *)(b + L::FRAG), 4, 0, 0);
        if (sc_wave && lane < 4 * L::CPT)
            __builtin_amdgcn_global_load_lds(sc + (size_t)cbase * 16 + lane * 4, (lds_void *)(b + L::FRAG + L::NQB), 4,
                                             0, 0);
    };
    // vmcnt <= (NB-2)*LPT: of this wavefront's loads only the NB-2 newest tiles
    // may still be in flight, so the oldest one (the tile about to be read) landed.
    // The count depends on how many loads this wavefront issues per slot (uniform).
    static_assert((L::NB - 2) * L::LPT_MAX < 64, "vmcnt field");
    const int lpt_extra = (nq_wave ? 1 : 0) + (sc_wave ? 1 : 0);
    auto wait_ring = [&]() {
        constexpr int N0 = (L::NB - 2) * L::UPW, N1 = (L::NB - 2) * (L::UPW + 1), N2 = (L::NB - 2) * (L::UPW + 2);
        if (N0 == N2 || lpt_extra == 0) __builtin_amdgcn_s_waitcnt((N0 & 15) | ((N0 >> 4) << 14) | (7 << 4) | (15 << 8));
        else if (lpt_extra == 1) __builtin_amdgcn_s_waitcnt((N1 & 15) | ((N1 >> 4) << 14) | (7 << 4) | (15 << 8));
        else __builtin_amdgcn_s_waitcnt((N2 & 15) | ((N2 >> 4) << 14) | (7 << 4) | (15 << 8));
    };

#ifndef BFJ_PIPE_MFMA
#define BFJ_PIPE_MFMA 16
#endif
    // Read-ahead of the next block's fragments pays where a block has enough MFMA
    // work to cover an LDS read.  Measured: -1 % on the FP16 join at d = 128 (32
    // MFMAs per block), -3 % at BF16 d = 256, -2 % for FP8 at 4 MFMAs of K = 128 per
    // block, but +4 % on a batch of small-d FP16 tiles with 4-8 MFMAs per block,
    // where it costs more in registers than it hides.
    constexpr bool PIPE = RB * KC * (Tr::dtype == BFJ_FP8 ? 4 : 1) >= BFJ_PIPE_MFMA;
    // One candidate block cb against this wavefront's rows.  With `more`, next(c)
    // gives the next block's K-chunk c, which is read into B[c] as soon as this
    // block's MFMAs on B[c] have issued: the next block's LDS reads are in flight
    // under this block's remaining MFMAs instead of stalling its first one, with no
    // extra registers.
    auto block = [&](int cb, frag_t (&B)[KC], val_t nj, int SB, bool more, auto &&next) {
        // acc = g_i + xhat_i . xhat_j; the first MFMA reads g_i as its C operand, so no
        // VALU work precedes the MFMAs.  Z = acc - n_j is only formed for hot tiles:
        // acc < n_j is Z < 0 exactly, and max commutes with subtracting the per-lane n_j.
        acc_t acc[RB];
        // Raised issue priority while this wavefront has MFMAs and the max tree to
        // issue, so the SIMD's matrix pipe goes to the wavefronts that can feed it
        // rather than to one that is loading or pushing candidates.  Measured: -0.8 %
        // at FP16 R = 64 (3 resident blocks per CU), but +4 % at BF16 R = 32 and
        // +10 % for FP64, so only the wide low-precision strips use it.  At FP16
        // R = 128 the two s_setprio also bound the scheduler's regions: without them
        // the kernel needs 276 registers and spills into AGPRs (1 wave per SIMD).
        constexpr bool PRIO = Tr::dtype != BFJ_FP64 && R >= 64;
        if (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[rb] = Tr::mma(A[rb][0], B[0], G[rb], SA[rb], SB);
        // The scheduling barriers keep each read right after the MFMAs it follows
        // (left alone, the scheduler sinks them all to the block's end).
        if (PIPE && more) {
            B[0] = next(0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int c = 1; c < KC; ++c) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = Tr::mma(A[rb][c], B[c], acc[rb], SA[rb], SB);
            if (PIPE && more) {
                B[c] = next(c);
                if (c + 1 < KC) __builtin_amdgcn_sched_barrier(0);
            }
        }

        // Inputs are NaN-free (the pass rejects non-finite rows), so with -fno-honor-nans
        // these are v_max3 trees: one per row block, then one over the blocks.
        val_t mrb[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) mrb[rb] = fmax(fmax(acc[rb][0], acc[rb][1]), fmax(acc[rb][2], acc[rb][3]));
        val_t m = mrb[0];
#pragma unroll
        for (int rb = 1; rb < RB; ++rb) m = fmax(m, mrb[rb]);

        if (PRIO) __builtin_amdgcn_s_setprio(0);
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(m >= nj) != 0ull, 0)) {
            // Rare: stream every pair with acc >= n_j (Z >= 0) as a candidate record;
            // k_classify applies the per-eps tests to Z = acc - n_j.  Only the row blocks
            // that hold such a pair are visited; in a symmetric sweep the tiles with
            // cb < rb belong to another tile.  The row block's accumulators are picked
            // with register selects (rb is uniform), so one push body serves them all.
            const bool sym = s.sym != 0;
            uint32_t hot = 0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) hot |= (uint32_t)(__builtin_amdgcn_ballot_w64(mrb[rb] >= nj) != 0ull) << rb;
            // In a symmetric sweep keep the row blocks rb <= cb - wrb0: all of them once
            // the candidate block has passed this wavefront's last row block.
            if (sym && cb < wrb0 + RB - 1) {
                const int lim = cb - wrb0;
                hot &= lim < 0 ? 0u : (2u << lim) - 1u;
            }
            // Padding columns need no mask: their n_j is 1e30 (prep), beyond any
            // acc = g_i + xhat_i . 0.
            const uint32_t colrec = (uint32_t)(s.row_off + cb * 16 + (lane & 15)) | (sym ? BFJ_COL_SYM : 0u);
            // The diagonal block keeps only col > row: one lane mask per register (fixed
            // per lane), applied under a uniform branch marked unlikely, so that the
            // compiler keeps the branch instead of four selects on every push (FP8 join
            // -3.7 %).
            auto push = [&](acc_t v, int blk) {
                if (__builtin_expect(sym && blk == cb, 0)) {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        v[r] = ((lane & 15) <= Tr::out_row(lane, r)) ? -(val_t)__builtin_huge_val() : v[r];
                }
                cand_push_block<Tr>(a, cc, v, nj, (uint32_t)(s.row_off + blk * 16), colrec, lane);
            };
            if (RB == 1) {
                if (hot) push(acc[0], wrb0);
            } else {
                while (hot) {
                    const int rb = __builtin_ctz(hot);
                    hot &= hot - 1u;
                    acc_t v = acc[0];
#pragma unroll
                    for (int k = 1; k < RB; ++k)
                        if (rb == k) v = acc[k];
                    push(v, wrb0 + rb);
                }
            }
        }
    };

    if (BFJ_DIRECT_B) {
        // Candidate fragments straight from global memory (L2-resident panel) into
        // VGPRs, one block ahead; no ring, no barrier.
        frag_t Bn[KC];
        val_t njn = 0;
        int SBn = 127;
        auto fetch = [&](int cb) {
#pragma unroll
            for (int c = 0; c < KC; ++c) Bn[c] = Tr::load(P + ((int64_t)cb * KC + c) * BFJ_WAVE * FB, lane);
            njn = nq[cb * 16 + (lane & 15)];
            if (Tr::dtype == BFJ_FP8) SBn = (int)sc[cb * 16 + (lane & 15)];
        };
        fetch(cb0);
        auto none = [&](int) { return frag_t{}; };
        for (int cb = cb0; cb < cb1; ++cb) {
            frag_t B[KC];
#pragma unroll
            for (int c = 0; c < KC; ++c) B[c] = Bn[c];
            const val_t nj = njn;
            const int SB = SBn;
            if (cb + 1 < cb1) fetch(cb + 1);
            if (any_rows) block(cb, B, nj, SB, false, none);
        }
        return;
    }

    // Drain the strip's own loads here, so that inside the loop the only
    // outstanding vector-memory loads are the ring's.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // The item's panel [cb0, cb1) of candidate blocks, a whole number of ring slots.
    // (A dataset whose rows are all queries only has panels from about the strip's
    // first block on: every unordered pair lies in exactly one tile (rb, cb) with
    // cb >= rb, which decides it for both rows.)  Item slot t lives in ring buffer
    // t mod NB.  The last NB-1 prefetches run past the item into memory that holds
    // the next blocks or the buffers' tail padding; they are never used.
    const int nt = (cb1 - cb0) / L::CPT;
#pragma unroll
    for (int t = 0; t < L::NB - 1; ++t) issue(t, t);
    auto step = [&](int t, int rcur) {
        wait_ring();
        // All quarters of slot t landed (each wavefront waited for its own loads
        // above); every wavefront is done with slot t-1, whose fragments fed its
        // MFMAs before it got here.  No release fence: the ring is written only by
        // the direct-to-LDS loads, and a release would drain the ones in flight.
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        issue(t + L::NB - 1, rcur == 0 ? L::NB - 1 : rcur - 1); // into the buffer of slot t-1
        if (!any_rows) return;
        const unsigned char *b = lds + rcur * L::BUF;
        auto frag = [&](int h, int c) { return Tr::load_ring(b + (h * KC + c) * BFJ_WAVE * FB, lane); };
        auto norm = [&](int h) { return reinterpret_cast<const val_t *>(b + L::FRAG)[h * 16 + (lane & 15)]; };
        auto scale = [&](int h) {
            return (Tr::dtype == BFJ_FP8) ? (int)b[L::FRAG + L::NQB + h * 16 + (lane & 15)] : 127;
        };
        frag_t B[KC];
        if (PIPE) {
            // The slot's first fragments and all its norms (and FP8 scales) up front;
            // each block then reads the next one's fragments under its own MFMAs.
            val_t njs[L::CPT];
            int SBs[L::CPT];
#pragma unroll
            for (int c = 0; c < KC; ++c) B[c] = frag(0, c);
#pragma unroll
            for (int h = 0; h < L::CPT; ++h) {
                njs[h] = norm(h);
                SBs[h] = scale(h);
            }
#pragma unroll
            for (int h = 0; h < L::CPT; ++h)
                block(cb0 + t * L::CPT + h, B, njs[h], SBs[h], h + 1 < L::CPT, [&](int c) { return frag(h + 1, c); });
        } else {
#pragma unroll
            for (int h = 0; h < L::CPT; ++h) {
#pragma unroll
                for (int c = 0; c < KC; ++c) B[c] = frag(h, c);
                block(cb0 + t * L::CPT + h, B, norm(h), scale(h), false, [&](int c) { return frag(h, c); });
            }
        }
    };
    // Unrolled by NB so that every ring address is a constant.
    for (int t0 = 0; t0 < nt; t0 += L::NB) {
#pragma unroll
        for (int j = 0; j < L::NB; ++j)
            if (t0 + j < nt) step(t0 + j, j);
    }
    // The prefetched tail loads must land before the next strip reuses the ring.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Persistent kernel: every block takes 4R-row strips from a global queue until
// it is empty.  All waves of a block stay together (the candidate tiles are
// shared), so the queue is read by one thread and broadcast through LDS.
template <class Tr, int KC, int R>
#ifdef BFJ_WAVES_PER_EU
#define BFJ_JOIN_ATTR __attribute__((amdgpu_waves_per_eu(BFJ_WAVES_PER_EU, BFJ_WAVES_PER_EU)))
#else
#define BFJ_JOIN_ATTR
#endif
__global__ __launch_bounds__(BFJ_JBLOCK) BFJ_JOIN_ATTR void k_join(JoinArgs a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    __shared__ uint32_t item_s;
    CandCursor cc;
    cc.dst = a.cand;
    cc.chunk = 0xffffffffu;
    cc.fill = BFJ_CHUNK;
    cc.live = false;
    const int lane = bfj_lane();
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(a.item_counter, 1u);
        __syncthreads();
        const uint32_t item = (uint32_t)__builtin_amdgcn_readfirstlane((int)item_s);
        __syncthreads();
        if (item >= (uint32_t)a.n_items) break;
        const WorkItem w = a.items[item];
        const SetDesc s = a.sets[w.set];
        join_strip<Tr, KC, R>(a, s, w.rb0, w.cb0, w.cb1, lds_raw, cc);
    }
    cand_close(a, cc, lane);
}
