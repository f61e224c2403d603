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
// Layout of the work: one lane per band record.  Each lane streams the two rows
// of its pair through registers with 16-byte loads, RK doubles at a time, the
// next chunk in flight while the current one feeds the FMA chain; the many
// wavefronts per SIMD hide the rest of the latency.  Rows of the pairs of one
// wavefront are shared in L2 (a strip's records cluster on few rows).
#include "bfj_internal.h"
#include <math.h>

// ---- classification of the join's candidate stream --------------------------------
// One wavefront per chunk.  Per record, row and eps: certainly within eps is
// counted, inside the band goes to the refine list, otherwise out.  A record of a
// symmetric sweep decides for its row and for its candidate row.  Counts go through
// a per-wavefront LDS table (below); add_bucket is the path for passes whose count
// array is not indexable in 32 bits.
//
// add_bucket: cnt[bucket][row] += 1 for every lane with a bucket (>= 0); lanes with the
// same global row and bucket are summed first (a vote loop over the distinct keys of
// the wavefront), then one atomic per key.  `li` is the dataset-local row of `key` in
// dataset `s`.  All lanes take part.
template <class Set>
__device__ inline void add_bucket(const PassArgs &a, uint32_t key, const Set &s, uint32_t li, int bucket, int lane)
{
    const unsigned long long k = bucket >= 0 ? (((unsigned long long)key << 3) | (unsigned)bucket) : ~0ull;
    unsigned long long pending = __ballot(bucket >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const unsigned long long kk = __shfl(k, leader);
        const unsigned long long m = __ballot(k == kk);
        if (lane == leader) atomicAdd(&a.cnt[s.cnt_off + (int64_t)bucket * s.npad + li], (uint32_t)__popcll(m));
        pending &= ~m;
    }
}

// Per-wavefront count table in LDS: cnt index -> pending count, open addressing with
// linear probing.  Counts of keys that repeat over many records (the rows of a strip,
// the columns of a dense tile) add up here and reach cnt with one global atomic per
// key per flush instead of one per key per 64 records.  Keys are cnt indices (used
// when the whole cnt array is indexable in 32 bits; the caller checks).  A key that
// finds no slot within BFJ_CTAB_PROBES probes goes straight to cnt.
#ifndef BFJ_CTAB
#define BFJ_CTAB 512 // slots per wavefront (a power of two)
#endif
#define BFJ_CTAB_PROBES 8
#define BFJ_CTAB_EMPTY 0xffffffffu
struct CountTable {
    uint32_t key[BFJ_CTAB];
    uint32_t cnt[BFJ_CTAB];
    uint16_t used[BFJ_CTAB]; // slots claimed since the last flush, in claim order
    uint32_t n_used;
};

__device__ inline void ctab_clear(CountTable &t, int lane)
{
    for (int h = lane; h < BFJ_CTAB; h += BFJ_WAVE) {
        t.key[h] = BFJ_CTAB_EMPTY;
        t.cnt[h] = 0;
    }
    if (lane == 0) t.n_used = 0;
    __builtin_amdgcn_wave_barrier();
}

// cnt[idx] += 1 for the lanes with `has`, through the table.
__device__ inline void ctab_add(const PassArgs &a, CountTable &t, uint32_t idx, bool has)
{
    if (!has) return;
    uint32_t h = (idx * 2654435761u) >> (32 - __builtin_ctz(BFJ_CTAB));
    for (int p = 0; p < BFJ_CTAB_PROBES; ++p) {
        uint32_t old = BFJ_CTAB_EMPTY;
        __hip_atomic_compare_exchange_strong(&t.key[h], &old, idx, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == BFJ_CTAB_EMPTY)
            t.used[__hip_atomic_fetch_add(&t.n_used, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)] =
                (uint16_t)h;
        if (old == BFJ_CTAB_EMPTY || old == idx) {
            // Workgroup scope: an LDS atomic the compiler cannot merge with the
            // global one below into a flat atomic.
            __hip_atomic_fetch_add(&t.cnt[h], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return;
        }
        h = (h + 1) & (BFJ_CTAB - 1);
    }
    atomicAdd(&a.cnt[idx], 1u);
}

// Moves the pending counts to cnt (one atomic per claimed slot) and empties the
// table.  All lanes take part.
__device__ inline void ctab_flush(const PassArgs &a, CountTable &t, int lane)
{
    __builtin_amdgcn_wave_barrier();
    const uint32_t n = __hip_atomic_load(&t.n_used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    for (uint32_t q = lane; q < n; q += BFJ_WAVE) {
        const uint32_t h = t.used[q];
        atomicAdd(&a.cnt[t.key[h]], t.cnt[h]);
        t.key[h] = BFJ_CTAB_EMPTY;
        t.cnt[h] = 0;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) t.n_used = 0;
    __builtin_amdgcn_wave_barrier();
}

// cnt[bi][i] += 1 and cnt[bj][j] += 1 for rows i, j of dataset s (bucket -1: no
// count), through the table when use_tab holds (uniform), else by the vote loop.
// All lanes take part.
template <class Set>
__device__ inline void count_rows(const PassArgs &a, CountTable &tab, bool use_tab, const Set &s, uint32_t i,
                                  int bi, uint32_t j, int bj, int lane)
{
    if (use_tab) {
        // Room for this call's keys (at most 2 per lane) at under 3/4 load.
        if (__hip_atomic_load(&tab.n_used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) >
            (uint32_t)(BFJ_CTAB * 3 / 4 - 2 * BFJ_WAVE))
            ctab_flush(a, tab, lane);
        ctab_add(a, tab, (uint32_t)(s.cnt_off + (int64_t)bi * s.npad + i), bi >= 0);
        ctab_add(a, tab, (uint32_t)(s.cnt_off + (int64_t)bj * s.npad + j), bj >= 0);
    } else {
        add_bucket(a, (uint32_t)s.row_off + i, s, i, bi, lane);
        add_bucket(a, (uint32_t)s.row_off + j, s, j, bj, lane);
    }
}

// Pair-list mode: append the pair at packed positions (i, j) for the lanes with
// `take` (one atomic per wavefront), as dataset rows, smaller first.  Pairs past the
// capacity are counted, not written.  All lanes take part.
__device__ inline void append_pairs(const PassArgs &a, const SetDesc &s, bool take, uint32_t pi, uint32_t pj,
                                    int lane)
{
    uint32_t i = 0, j = 0;
    if (take) {
        i = bfj_row_at(a, s, pi);
        j = bfj_row_at(a, s, pj);
        if (i > j) {
            const uint32_t t = i;
            i = j;
            j = t;
        }
    }
    const unsigned long long bal = __ballot(take);
    if (!bal) return;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(a.pair_count, (unsigned long long)__popcll(bal));
    base = __shfl(base, 0);
    if (take) {
        const unsigned long long idx = base + (unsigned long long)__popcll(bal & ((1ull << lane) - 1ull));
        if (idx < a.pair_capacity) a.pairs[idx] = make_uint2(i, j);
    }
}

// Decide every banded eps of a pair from the bounds [lo, hi] of its squared
// distance S (kernel space) and its quantisation radius r = q_i + q_j:
//   certainly within eps if hi <= L = (slo - r)^2, certainly outside if lo > H = (shi + r)^2
// (rounded outwards), in the band otherwise.
template <class Set>
__device__ inline void decide(const Set &sc, uint32_t band_mask, int n_eps, double r, double lo, double hi,
                              uint32_t &in, uint32_t &band)
{
#pragma unroll
    for (int e = 0; e < BFJ_MAXEPS; ++e) {
        if (e >= n_eps) break;
        if (!((band_mask >> e) & 1u)) continue;
        const double slo = sc.slo[e], shi = sc.shi[e];
        const double L = slo > r ? (slo - r) * (slo - r) * (1.0 - 0x1p-49) : -INFINITY;
        const double H = (shi + r) * (shi + r) * (1.0 + 0x1p-49);
        // Flags by arithmetic, not through a selected reference: that form puts in
        // and band in scratch memory.
        const bool cin = hi <= L;
        in |= (uint32_t)cin << e;
        band |= (uint32_t)(!cin && lo <= H) << e;
    }
}

// The decisions for one candidate record (lanes with live == false decide nothing).
struct Decided {
    Candidate c;
    int si;
    uint32_t i, j, in_i, band_i, in_j, band_j;
};

// What classify needs of a dataset: its descriptor's offsets and its per-eps bounds.
struct SetView {
    int64_t row_off, cnt_off, npad;
    int n_eps;
    uint32_t band_mask;
    double slo[BFJ_MAXEPS], shi[BFJ_MAXEPS];
};

// Through the constant address space: scalar loads when si is uniform.  Both arrays
// are read-only while classify runs.
typedef const __attribute__((address_space(4))) SetDesc *bfj_cset;
typedef const __attribute__((address_space(4))) SetScalars *bfj_cscal;

__device__ inline SetView set_view_uniform(const PassArgs &a, int si)
{
    SetView v;
    const bfj_cset s = (bfj_cset)a.sets + si;
    const bfj_cscal sc = (bfj_cscal)a.scal + si;
    v.row_off = s->row_off;
    v.cnt_off = s->cnt_off;
    v.npad = s->npad;
    v.n_eps = s->n_eps;
    v.band_mask = (uint32_t)sc->band_mask;
#pragma unroll
    for (int e = 0; e < BFJ_MAXEPS; ++e) {
        v.slo[e] = sc->slo[e];
        v.shi[e] = sc->shi[e];
    }
    return v;
}

// The loads of one candidate record: the record and the per-row data of both its rows
// (the record holds both as global rows, so nothing here waits on the dataset's
// descriptor).  Branch-free, so that the loads of several groups go out together.
// Lanes past the chunk's fill re-read its last record and decide nothing.
struct Fetched {
    Candidate c;
    double2 h;     // he[row]
    double qi, qj; // quantisation radii of both rows
    double dj;     // dnq[col]: n_j minus the join's candidate threshold
    int si;        // dataset
    bool live;
};

// SPLIT: the pass lowered the join's candidate thresholds (FP8, k_thresholds), so the
// record's Z' needs dnq of its column; a separate instance keeps the other passes'
// registers (and occupancy) as they were.
template <bool SPLIT>
__device__ inline void fetch_record(const PassArgs &a, const Candidate *recs, uint32_t t, uint32_t fill, Fetched &f)
{
    f.live = t < fill;
    f.c = recs[f.live ? t : fill - 1];
    f.si = (int)a.row_set[f.c.row];
    f.h = a.he[f.c.row];
    f.qi = a.qerr[f.c.row];
    f.qj = a.qerr[f.c.col & ~BFJ_COL_SYM];
    f.dj = SPLIT ? a.dnq[f.c.col & ~BFJ_COL_SYM] : 0.0;
}

// s: the record's dataset (row_off); b: its bounds (slo, shi); n_eps, band_mask of it.
template <class Set, class Bounds>
__device__ inline Decided decide_fetched(const Fetched &f, const Set &s, const Bounds &b, int n_eps,
                                         uint32_t band_mask)
{
    Decided o;
    o.c = f.c;
    o.si = f.si;
    o.i = f.c.row - (uint32_t)s.row_off;
    o.j = (f.c.col & ~BFJ_COL_SYM) - (uint32_t)s.row_off;
    o.in_i = o.band_i = o.in_j = o.band_j = 0;
    if (f.live) {
        // S_hat = 2 g_i + N_i - 2 Z with Z = acc - n_j = Z' - dnq_j (the record holds
        // Z' = acc - n'_j), |S_hat - S| <= E_i; mg covers the FP64 rounding of these
        // few operations.
        const double shat = f.h.x - 2.0 * f.c.z + 2.0 * f.dj;
        const double mg = (fabs(f.h.x) + 2.0 * fabs(f.c.z) + 2.0 * f.dj + f.h.y) * 0x1p-50;
        const double lo = shat - f.h.y - mg, hi = shat + f.h.y + mg;
        // |dist(xhat_i, xhat_j) - dist(y_i, y_j)| <= q_i + q_j: the same bounds hold
        // for both rows of the pair.
        const double r = (f.qi + f.qj) * (1.0 + 0x1p-50);
        decide(b, band_mask, n_eps, r, lo, hi, o.in_i, o.band_i);
        if (f.c.col & BFJ_COL_SYM) {
            o.in_j = o.in_i;
            o.band_j = o.band_i;
        }
    }
    return o;
}

#ifndef BFJ_CLS_GROUPS
#define BFJ_CLS_GROUPS 4 // 64-record groups of a chunk whose loads are in flight together
#endif

// The band-list block a wavefront of k_classify is filling.  A wavefront reserves
// a.band_block slots at a time (one atomic per block on the shared counter instead
// of one per chunk, and no pass that waits for it per chunk) and writes its band
// records into them in order; at the end it marks the rest of its last block as
// holes.  base and fill are wave-uniform.
struct BandCursor {
    unsigned long long base; // first slot of the block (~0: none yet)
    uint32_t fill;           // slots of it written
    unsigned long long records;
};

__device__ inline unsigned long long band_reserve(const PassArgs &a, int lane)
{
    unsigned long long b = 0;
    if (lane == 0) b = atomicAdd(a.band_count, (unsigned long long)a.band_block);
    return __shfl(b, 0);
}

// Appends the entry of every lane with `take` (ranks in lane order), opening a new
// block when this one is full; band_block >= 64, so one group needs at most one new
// block.  Slots past the capacity are reserved but not written (the host sees the
// counter and repeats the pass).  All lanes take part.
__device__ inline void band_append(const PassArgs &a, BandCursor &bc, bool take, const BandEntry &be, int lane)
{
    const unsigned long long bb = __ballot(take);
    if (!bb) return;
    const uint32_t nb = (uint32_t)__popcll(bb);
    if (bc.base == ~0ull) {
        bc.base = band_reserve(a, lane);
        bc.fill = 0;
    }
    const uint32_t room = a.band_block - bc.fill;
    unsigned long long next = ~0ull;
    if (nb > room) next = band_reserve(a, lane);
    if (take) {
        const uint32_t k = (uint32_t)__popcll(bb & ((1ull << lane) - 1ull));
        const unsigned long long idx = k < room ? bc.base + bc.fill + k : next + (k - room);
        if (idx < a.band_capacity) a.band[idx] = be;
    }
    if (nb > room) {
        bc.base = next;
        bc.fill = nb - room;
    } else {
        bc.fill += nb;
    }
    bc.records += nb;
}

// The end of a wavefront's part: holes in the rest of its block, and its record count.
__device__ inline void band_close(const PassArgs &a, const BandCursor &bc, int lane)
{
    if (bc.base != ~0ull) {
        BandEntry hole;
        hole.i = hole.j = 0;
        hole.z = 0.f;
        hole.meta = 0u;
        for (uint32_t t = bc.fill + lane; t < a.band_block; t += BFJ_WAVE) {
            const unsigned long long idx = bc.base + t;
            if (idx < a.band_capacity) a.band[idx] = hole;
        }
    }
    if (lane == 0 && bc.records) atomicAdd(a.band_records, bc.records);
}

// body(decided, dataset, group) for the groups g0 + k, k < BFJ_CLS_GROUPS, with bit k
// of gmask (uniform).  When every record of these groups is from one dataset (the
// common case: a chunk is filled by one join wavefront), its descriptor and scalars
// are read once through scalar loads; otherwise per lane.  All lanes take part.
template <bool SPLIT, class Body>
__device__ inline void classify_groups(const PassArgs &a, const Candidate *recs, uint32_t fill, uint32_t g0,
                                       uint32_t gmask, int lane, Body &&body)
{
    constexpr int G = BFJ_CLS_GROUPS;
    Fetched f[G];
#pragma unroll
    for (int k = 0; k < G; ++k) fetch_record<SPLIT>(a, recs, (g0 + k) * BFJ_WAVE + lane, fill, f[k]);
    const int si0 = __builtin_amdgcn_readfirstlane(f[0].si);
    bool same = true;
#pragma unroll
    for (int k = 0; k < G; ++k) same &= f[k].si == si0;
    if (__ballot(!same) == 0ull) {
        const SetView v = set_view_uniform(a, si0);
#pragma unroll
        for (int k = 0; k < G; ++k)
            if ((gmask >> k) & 1u) body(decide_fetched(f[k], v, v, v.n_eps, v.band_mask), v, g0 + k);
    } else {
        // Per lane, the bounds read as the eps loop needs them.
#pragma unroll
        for (int k = 0; k < G; ++k)
            if ((gmask >> k) & 1u) {
                const SetDesc &s = a.sets[f[k].si];
                const SetScalars &sc = a.scal[f[k].si];
                body(decide_fetched(f[k], s, sc, s.n_eps, (uint32_t)sc.band_mask), s, g0 + k);
            }
    }
}

template <bool SPLIT>
__global__ __launch_bounds__(BFJ_BLOCK) void k_classify(PassArgs a)
{
    constexpr int G = BFJ_CLS_GROUPS;
    const int lane = threadIdx.x & 63;
    __shared__ CountTable tabs[BFJ_BLOCK / BFJ_WAVE];
    CountTable &tab = tabs[threadIdx.x >> 6];
    const bool use_tab = a.cnt_len < (int64_t)BFJ_CTAB_EMPTY;
    if (use_tab) ctab_clear(tab, lane);
    const unsigned long long chunks_used = *a.chunk_count;
    const uint32_t n_chunks = chunks_used < a.chunk_capacity ? (uint32_t)chunks_used : a.chunk_capacity;
    unsigned long long records_read = 0;
    BandCursor bc;
    bc.base = ~0ull;
    bc.fill = 0;
    bc.records = 0;
    const uint32_t wave_id = blockIdx.x * (BFJ_BLOCK / BFJ_WAVE) + (threadIdx.x >> 6);
    const uint32_t n_waves = gridDim.x * (BFJ_BLOCK / BFJ_WAVE);
    for (uint32_t ch = wave_id; ch < n_chunks; ch += n_waves) {
        const uint32_t fill = min(a.chunk_fill[ch], (uint32_t)BFJ_CHUNK);
        records_read += fill;
        const Candidate *recs = a.cand + (size_t)ch * BFJ_CHUNK;
        const uint32_t n_groups = (fill + BFJ_WAVE - 1) / BFJ_WAVE; // at most 16
        // Count the certain-in pairs; write the records with an eps in the band.
        auto body = [&](const Decided &d, const auto &s, uint32_t) {
            // A record with an eps in the band is decided whole by the later stages;
            // otherwise it counts once, under the first eps that certainly holds it.
            const bool band = (d.band_i | d.band_j) != 0u;
            const int bi = (!band && d.in_i) ? __builtin_ctz(d.in_i) : -1;
            const int bj = (!band && d.in_j) ? __builtin_ctz(d.in_j) : -1;
            count_rows(a, tab, use_tab, s, d.i, bi, d.j, bj, lane);
            if (a.emit_pairs) append_pairs(a, a.sets[d.si], !band && (d.in_i & 1u) && d.i != d.j, d.i, d.j, lane);
            BandEntry be;
            be.i = d.i;
            be.j = d.j;
            be.z = (float)d.c.z;
            be.meta = (d.band_i | d.band_j) | ((d.c.col & BFJ_COL_SYM) ? 0x8000u : 0u) | ((uint32_t)d.si << 16);
            band_append(a, bc, band, be, lane);
        };
        for (uint32_t g0 = 0; g0 < n_groups; g0 += G) {
            const uint32_t left = n_groups - g0;
            classify_groups<SPLIT>(a, recs, fill, g0, left >= (uint32_t)G ? (1u << G) - 1u : (1u << left) - 1u, lane,
                                   body);
        }
    }
    band_close(a, bc, lane);
    if (use_tab) ctab_flush(a, tab, lane);
    if (lane == 0 && records_read) atomicAdd(a.cand_records, records_read);
}

void bfj_launch_classify(const PassArgs &a, hipStream_t st)
{
    const dim3 grid(BFJ_CLASSIFY_WAVES / (BFJ_BLOCK / BFJ_WAVE));
    if (a.dtype == BFJ_FP8) hipLaunchKernelGGL(k_classify<true>, grid, dim3(BFJ_BLOCK), 0, st, a);
    else hipLaunchKernelGGL(k_classify<false>, grid, dim3(BFJ_BLOCK), 0, st, a);
}

#define RK 8 // doubles per row per register chunk

// Chunk of RK doubles of one row into registers: 16-byte loads when the dataset's
// rows are 16-byte aligned (al16), 8-byte loads otherwise.
typedef const __attribute__((address_space(1))) double *bfj_gptr; // global memory: global_load, not flat_load
typedef const __attribute__((address_space(1))) bfj_d2 *bfj_gptr2;

__device__ inline void load_chunk(bfj_gptr p, bool al16, double (&u)[RK])
{
    if (al16) {
#pragma unroll
        for (int h = 0; h < RK / 2; ++h) {
            const bfj_d2 w = *(bfj_gptr2)(p + 2 * h);
            u[2 * h] = w.x;
            u[2 * h + 1] = w.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < RK; ++k) u[k] = p[k];
    }
}

// The reference's direct form over one chunk, k ascending.
__device__ inline double fma_chunk(const double (&u)[RK], const double (&v)[RK], double acc)
{
#pragma unroll
    for (int k = 0; k < RK; ++k) {
        const double t = u[k] - v[k];
        acc = fma(t, t, acc);
    }
    return acc;
}

// Refine one band record per lane (lanes with live == false take part in the votes
// and decide nothing): the sequential FMA chain of the pair's two FP64 rows and the
// bucket of the first eps (ascending) whose threshold holds it, counted for its row
// and (symmetric record) for the other row.  All lanes take part.
__device__ inline void refine_one(const PassArgs &a, const BandEntry &be, bool live, int lane, CountTable &tab,
                                  bool use_tab)
{
    const int si = (int)(be.meta >> 16);
    const SetDesc &s = a.sets[si];
    const int d = live ? s.d : 0;
    // be.i / be.j are packed positions; the FP64 rows are the dataset's rows held there.
    const bfj_gptr p = (bfj_gptr)(s.x + (int64_t)(live ? bfj_row_at(a, s, be.i) : 0) * s.ld);
    const bfj_gptr q = (bfj_gptr)(s.x + (int64_t)(live ? bfj_row_at(a, s, be.j) : 0) * s.ld);
    // Each lane runs the sequential FMA chain of its own pair.  Whole chunks go
    // through two register buffers, the next chunk's loads in flight while the
    // current one is consumed; the last d % RK values are loaded one by one.
    double acc = 0.0;
    const int nfull = d / RK;
    double u0[RK], v0[RK], u1[RK], v1[RK];
    const bool al16 = s.al16 != 0;
    if (nfull > 0) {
        load_chunk(p, al16, u0);
        load_chunk(q, al16, v0);
    }
    for (int c = 0; c < nfull; c += 2) {
        if (c + 1 < nfull) {
            load_chunk(p + (c + 1) * RK, al16, u1);
            load_chunk(q + (c + 1) * RK, al16, v1);
        }
        acc = fma_chunk(u0, v0, acc);
        if (c + 1 >= nfull) break;
        if (c + 2 < nfull) {
            load_chunk(p + (c + 2) * RK, al16, u0);
            load_chunk(q + (c + 2) * RK, al16, v0);
        }
        acc = fma_chunk(u1, v1, acc);
    }
    for (int k = nfull * RK; k < d; ++k) {
        const double tk = p[k] - q[k];
        acc = fma(tk, tk, acc);
    }
    // The pair counts once, under the first eps (ascending) whose threshold holds
    // its FP64 distance, for its row and (symmetric record) for the other row.
    int k = -1;
    if (live)
        for (int e = 0; e < s.n_eps; ++e)
            if (acc <= s.T[e]) {
                k = e;
                break;
            }
    count_rows(a, tab, use_tab, s, be.i, k, be.j, (be.meta & 0x8000u) ? k : -1, lane);
    if (a.emit_pairs) append_pairs(a, s, k == 0 && be.i != be.j, be.i, be.j, lane);
}

// Each wavefront takes runs of BFJ_REFINE_RUN consecutive records (classify writes a
// chunk's records together, so the rows of a run repeat and their counts meet in
// the wavefront's count table), the runs dealt round-robin over the wavefronts.
#ifndef BFJ_REFINE_RUN
#define BFJ_REFINE_RUN 512
#endif
__global__ __launch_bounds__(BFJ_BLOCK) void k_refine(PassArgs a)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long cnt = *a.band_count;
    const uint64_t m = cnt < a.band_capacity ? cnt : a.band_capacity;
    const uint64_t wave = blockIdx.x * (BFJ_BLOCK / BFJ_WAVE) + (threadIdx.x >> 6);
    const uint64_t step = (uint64_t)gridDim.x * (BFJ_BLOCK / BFJ_WAVE) * BFJ_REFINE_RUN;
    __shared__ CountTable tabs[BFJ_BLOCK / BFJ_WAVE];
    CountTable &tab = tabs[threadIdx.x >> 6];
    const bool use_tab = a.cnt_len < (int64_t)BFJ_CTAB_EMPTY;
    if (use_tab) ctab_clear(tab, lane);
    // base is uniform over the wavefront, so every lane takes part in the votes.
    for (uint64_t run = wave * BFJ_REFINE_RUN; run < m; run += step) {
        const uint64_t end = run + BFJ_REFINE_RUN < m ? run + BFJ_REFINE_RUN : m;
        for (uint64_t base = run; base < end; base += BFJ_WAVE) {
            const uint64_t t = base + lane;
            const bool live = t < end;
            BandEntry be;
            be.i = be.j = 0;
            be.z = 0.f;
            be.meta = 0;
            if (live) be = a.band[t];
            refine_one(a, be, live && bfj_band_live(be), lane, tab, use_tab);
        }
    }
    if (use_tab) ctab_flush(a, tab, lane);
}

// Pair-list mode when eps covers every pair: all (i, j), i < j, row i's pairs at
// offset i (n - 1) - i (i - 1) / 2, in order.
__global__ void k_all_pairs(uint2 *pairs, int64_t n)
{
    for (int64_t i = blockIdx.x; i < n - 1; i += gridDim.x) {
        const int64_t off = i * (n - 1) - i * (i - 1) / 2;
        for (int64_t j = i + 1 + threadIdx.x; j < n; j += blockDim.x)
            pairs[off + (j - i - 1)] = make_uint2((uint32_t)i, (uint32_t)j);
    }
}

void bfj_launch_all_pairs(uint2 *pairs, int64_t n, hipStream_t st)
{
    if (n > 1) hipLaunchKernelGGL(k_all_pairs, dim3(4096), dim3(BFJ_BLOCK), 0, st, pairs, n);
}

void bfj_launch_refine(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_refine, dim3(4096), dim3(BFJ_BLOCK), 0, st, a);
}

// Per-row counts (self excluded) into the caller's int64 array and the
// reference's self-inclusive totals (gpu_join.cu:320-321).
__global__ void k_finalize(PassArgs a)
{
    __shared__ unsigned long long red[BFJ_MAXEPS][BFJ_PREP_ROWS];
    const int b = blockIdx.x;
    const int si = a.blk_set[b];
    const SetDesc &s = a.sets[si];
    const int64_t i = a.blk_row0[b] + threadIdx.x; // packed position
    const int64_t oi = i < s.n ? (int64_t)bfj_row_at(a, s, i) : i; // the dataset row held there
    const bool query = i < s.n && oi >= s.row_begin && oi < s.row_end;
    const int64_t rows = s.row_end - s.row_begin;
    unsigned long long run = 0; // buckets 0..e: pairs within the e-th smallest eps
    for (int e = 0; e < s.n_eps; ++e) {
        unsigned long long c = 0;
        if (query) {
            run += a.cnt[s.cnt_off + (int64_t)e * s.npad + i];
            // "all" eps: every candidate, self included, is within eps (bounded in
            // k_thresholds); the join did not count them.
            const bool all = (a.scal[si].all_mask >> e) & 1ull;
            c = all ? (unsigned long long)s.n : run;
            // A symmetric sweep streams no self pairs: d2(i, i) == 0 is within every eps >= 0.
            if (!all && s.sym && s.T[e] >= 0.0) c += 1;
            const long long self = (s.T[e] >= 0.0) ? 1 : 0; // d2(i, i) == 0 exactly
            if (s.out) s.out[(int64_t)s.eperm[e] * rows + (oi - s.row_begin)] = (long long)c - self;
        }
        red[e][threadIdx.x] = c;
    }
    __syncthreads();
    // The block's totals go to its own slot (k_totals adds them up per dataset): one
    // same-address atomic per block would be performed one after another in memory.
    if (threadIdx.x < BFJ_MAXEPS) {
        unsigned long long t = 0;
        if (threadIdx.x < (unsigned)s.n_eps)
            for (int r = 0; r < BFJ_PREP_ROWS; ++r) t += red[threadIdx.x][r];
        a.blk_tot[(int64_t)b * BFJ_MAXEPS + (threadIdx.x < (unsigned)s.n_eps ? s.eperm[threadIdx.x] : threadIdx.x)] = t;
    }
}

// One wavefront per dataset: its totals, the sums of its row blocks' slots.
__global__ __launch_bounds__(BFJ_WAVE) void k_totals(PassArgs a)
{
    const int si = blockIdx.x, lane = threadIdx.x;
    const int n_eps = a.sets[si].n_eps;
    const int64_t b0 = a.set_blk0[si], nb = a.set_nblk[si];
    for (int e = 0; e < n_eps; ++e) {
        unsigned long long t = 0;
        for (int64_t k = lane; k < nb; k += BFJ_WAVE) t += a.blk_tot[(b0 + k) * BFJ_MAXEPS + e];
        for (int o = BFJ_WAVE / 2; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (lane == 0) a.scal[si].total[e] = t;
    }
}

void bfj_launch_finalize(const PassArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(a.n_row_blocks), dim3(BFJ_PREP_ROWS), 0, st, a);
    if (a.n_sets > 0) hipLaunchKernelGGL(k_totals, dim3(a.n_sets), dim3(BFJ_WAVE), 0, st, a);
}

// ---- middle stage ---------------------------------------------------------------------
// The join's band is set by its operands' precision: e4m3 keeps 3 mantissa bits (a
// radius q_i of ~3.6 % of a row's norm), BF16 8 and FP16 11.  Before the FP64 refine,
// each band record is decided again from copies of its two rows at higher precision
// (k_mid_pack): FP16 for FP8 passes (512 bytes per row at d = 256 instead of 2 KiB of
// FP64), FP32 for FP16/BF16 passes (half the FP64 bytes).  S = ||c_i - c_j||^2 is
// summed in FP64: each difference and square is within 2^-52 relative of exact (exact
// for FP16), so the sum is within (d + 3) 2^-53 S, far inside the 2^-40 margin taken;
// and |dist(c_i, c_j) - dist(y_i, y_j)| <= qc_i + qc_j bounds the exact distance as in
// k_classify, with a radius 2^-8 (FP16 over e4m3) to 2^-13 (FP32 over FP16) of the
// join's.  Certain pairs are counted here; the rest go to the FP64 refine unchanged.
// The same runs of records per wavefront and count table as k_refine.

// Sum of (p_k - q_k)^2 over one chunk of both rows (4 16-byte units: 32 FP16 or 16
// FP32 values), in FP64.
template <bool F32>
__device__ inline double mid_chunk(const uint4 (&p)[4], const uint4 (&q)[4], double S)
{
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (F32) {
            const bfj_f4 fp = __builtin_bit_cast(bfj_f4, p[u]);
            const bfj_f4 fq = __builtin_bit_cast(bfj_f4, q[u]);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double dk = (double)fp[k] - (double)fq[k];
                S = fma(dk, dk, S);
            }
        } else {
            const bfj_h8 hp = __builtin_bit_cast(bfj_h8, p[u]);
            const bfj_h8 hq = __builtin_bit_cast(bfj_h8, q[u]);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double dk = (double)hp[k] - (double)hq[k];
                S = fma(dk, dk, S);
            }
        }
    }
    return S;
}

// The middle stage of 64 band records, one per lane (`live`: the lane holds one):
// S = ||c_i - c_j||^2 from the stage's copies, then decide, count or forward to the
// FP64 refine.  The records of lanes src < lim are read (holes among them too: a
// hole reads row 0 and is discarded).  All lanes take part.
template <bool F32>
__device__ inline void mid_records(const PassArgs &a, CountTable &tab, bool use_tab, const BandEntry &be, bool live,
                                   int lim, int lane)
{
    const unsigned long long below = (1ull << lane) - 1ull;
    const SetDesc &s = a.sets[be.meta >> 16];
    // S = ||c_i - c_j||^2 of the 64 records in 8 rounds of 8.  Each record's two
    // rows are read by a group of 8 lanes, lane h taking the row's 16-byte units h,
    // h + 8, ..., so one load covers 128 contiguous bytes of each of 8 rows (one
    // lane per record would touch 64 cache lines per load).  The group's partial
    // sums meet by butterfly and go back to the record's own lane.  k_mid_pack and
    // k_quantize16 wrote the columns past d up to a whole unit count as zeros.
    constexpr int ES = F32 ? 4 : 2; // bytes per value of the copy
    const int grp = lane >> 3, h = lane & 7;
    double S = 0.0;
    for (int b = 0; b < 8; ++b) {
        const int src = 8 * b + grp;
        const uint32_t ri = __shfl(be.i, src), rj = __shfl(be.j, src), rm = __shfl(be.meta, src);
        double Sg = 0.0;
        if (src < lim) {
            const SetDesc &rs = a.sets[rm >> 16];
            const int units = F32 ? (rs.d + 15) / 16 * 4 : (rs.d + 31) / 32 * 4;
            const uint4 *p = reinterpret_cast<const uint4 *>(a.mid_rows + (rs.row_off + ri) * BFJ_MAXDIM * ES);
            const uint4 *q = reinterpret_cast<const uint4 *>(a.mid_rows + (rs.row_off + rj) * BFJ_MAXDIM * ES);
            for (int u0 = h; u0 < units; u0 += 32) {
                // Unconditional loads (the index clamped to the row's last unit)
                // so that all eight are in flight together; units past the row
                // are then zeroed.
                uint4 pp[4], qq[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int u = u0 + 8 * k;
                    const int uc = u < units ? u : units - 1;
                    pp[k] = p[uc];
                    qq[k] = q[uc];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (u0 + 8 * k >= units) pp[k] = qq[k] = make_uint4(0u, 0u, 0u, 0u);
                Sg = mid_chunk<F32>(pp, qq, Sg);
            }
        }
        Sg += __shfl_xor(Sg, 1);
        Sg += __shfl_xor(Sg, 2);
        Sg += __shfl_xor(Sg, 4);
        const double Sb = __shfl(Sg, 8 * h);
        if (grp == b) S = Sb;
    }
    uint32_t in = 0, band = 0;
    if (live) {
        const SetScalars &sc = a.scal[be.meta >> 16];
        const double r = (a.mid_q[s.row_off + be.i] + a.mid_q[s.row_off + be.j]) * (1.0 + 0x1p-50);
        decide(sc, (uint32_t)sc.band_mask, s.n_eps, r, S * (1.0 - 0x1p-40), S * (1.0 + 0x1p-40), in, band);
    }
    // Still in a band: decided whole by the FP64 refine.  Otherwise the record
    // counts once, under the first eps that certainly holds it (if any).
    const bool fwd = live && band != 0u;
    const int k = (live && !fwd && in) ? __builtin_ctz(in) : -1;
    count_rows(a, tab, use_tab, s, be.i, k, be.j, (be.meta & 0x8000u) ? k : -1, lane);
    if (a.emit_pairs) append_pairs(a, s, k == 0 && be.i != be.j, be.i, be.j, lane);
    const unsigned long long fb = __ballot(fwd);
    if (fb) {
        unsigned long long b0 = 0;
        if (lane == 0) b0 = atomicAdd(a.mid_count, (unsigned long long)__popcll(fb));
        b0 = __shfl(b0, 0);
        if (fwd) {
            const unsigned long long idx = b0 + (unsigned long long)__popcll(fb & below);
            if (idx < a.band_capacity) a.mid_band[idx] = be;
        }
    }
}

// Records a wavefront stages and regroups by query row at a time, in the FP16-copy
// (FP8 pass) middle stage (a multiple of 64).
#ifndef BFJ_MID_SORT
#define BFJ_MID_SORT 256
#endif
template <bool F32>
__global__ __launch_bounds__(BFJ_BLOCK) void k_mid(PassArgs a)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long cnt = *a.band_count;
    const uint64_t m = cnt < a.band_capacity ? cnt : a.band_capacity;
    const uint64_t wave = blockIdx.x * (BFJ_BLOCK / BFJ_WAVE) + (threadIdx.x >> 6);
    const uint64_t step = (uint64_t)gridDim.x * (BFJ_BLOCK / BFJ_WAVE) * BFJ_REFINE_RUN;
    __shared__ CountTable tabs[BFJ_BLOCK / BFJ_WAVE];
    CountTable &tab = tabs[threadIdx.x >> 6];
    const bool use_tab = a.cnt_len < (int64_t)BFJ_CTAB_EMPTY;
    if (use_tab) ctab_clear(tab, lane);
    if constexpr (F32) {
        // FP32 copies (FP16 and BF16 passes): the records in stream order, 64 at a time.
        for (uint64_t run = wave * BFJ_REFINE_RUN; run < m; run += step) {
            const uint64_t end = run + BFJ_REFINE_RUN < m ? run + BFJ_REFINE_RUN : m;
            for (uint64_t base = run; base < end; base += BFJ_WAVE) {
                const uint64_t t = base + lane;
                BandEntry be;
                be.i = be.j = 0;
                be.z = 0.f;
                be.meta = 0;
                if (t < end) be = a.band[t];
                const bool live = t < end && bfj_band_live(be);
                mid_records<F32>(a, tab, use_tab, be, live, (int)(end - base < BFJ_WAVE ? end - base : BFJ_WAVE),
                                 lane);
            }
        }
    } else {
        // FP16 copies (FP8 passes): the live records of each BFJ_MID_SORT-record window are
        // first regrouped in LDS by the low 6 bits of their query row i (a counting sort:
        // one LDS atomic per record gives its rank in its bucket, a wave scan the bucket
        // offsets); holes are dropped.  The band records of an FP8 pass are sparse over
        // the strips' rows, so in stream order the 8 records of a round read mostly
        // different query rows; the records of one join chunk come from a strip of R <= 64
        // consecutive rows, so after the regrouping a bucket holds one row and the 8
        // records of a round mostly share row i, whose loads are then the same addresses in
        // one instruction, fetched once.
        constexpr int PER = BFJ_MID_SORT / BFJ_WAVE; // records per lane while regrouping
        __shared__ BandEntry srt[BFJ_BLOCK / BFJ_WAVE][BFJ_MID_SORT];
        __shared__ uint32_t bkt[BFJ_BLOCK / BFJ_WAVE][BFJ_WAVE];
        BandEntry *mys = srt[threadIdx.x >> 6];
        uint32_t *myb = bkt[threadIdx.x >> 6];
        for (uint64_t run = wave * BFJ_REFINE_RUN; run < m; run += step) {
            const uint64_t end = run + BFJ_REFINE_RUN < m ? run + BFJ_REFINE_RUN : m;
            for (uint64_t win = run; win < end; win += BFJ_MID_SORT) {
                myb[lane] = 0;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                BandEntry rr[PER];
                uint32_t rank[PER];
                int key[PER];
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const uint64_t t = win + (uint64_t)(k * BFJ_WAVE + lane);
                    rr[k].i = rr[k].j = 0;
                    rr[k].z = 0.f;
                    rr[k].meta = 0;
                    if (t < end) rr[k] = a.band[t];
                    const bool lv = t < end && bfj_band_live(rr[k]);
                    key[k] = lv ? (int)(rr[k].i & (BFJ_WAVE - 1)) : -1;
                    rank[k] = lv ? __hip_atomic_fetch_add(&myb[key[k]], 1u, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_WORKGROUP)
                                 : 0u;
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                // Exclusive scan of the 64 bucket counts; nl = live records of the window.
                const uint32_t cnt_b = myb[lane];
                uint32_t incl = cnt_b;
#pragma unroll
                for (int o = 1; o < BFJ_WAVE; o <<= 1) {
                    const uint32_t v = __shfl_up(incl, o);
                    if (lane >= o) incl += v;
                }
                const uint32_t nl = (uint32_t)__shfl((int)incl, BFJ_WAVE - 1);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                myb[lane] = incl - cnt_b;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                // Bucket offset + rank < nl <= BFJ_MID_SORT.
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (key[k] >= 0) mys[myb[key[k]] + rank[k]] = rr[k];
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                for (uint32_t base = 0; base < nl; base += BFJ_WAVE) {
                    const bool live = base + lane < nl;
                    BandEntry be;
                    be.i = be.j = 0;
                    be.z = 0.f;
                    be.meta = 0;
                    if (live) be = mys[base + lane];
                    mid_records<F32>(a, tab, use_tab, be, live, (int)(nl - base < BFJ_WAVE ? nl - base : BFJ_WAVE),
                                     lane);
                }
                // The staging and the counters are rewritten only after these reads.
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (use_tab) ctab_flush(a, tab, lane);
}

void bfj_launch_mid(const PassArgs &a, hipStream_t st)
{
    if (a.mid_fp32) hipLaunchKernelGGL(k_mid<true>, dim3(4096), dim3(BFJ_BLOCK), 0, st, a);
    else hipLaunchKernelGGL(k_mid<false>, dim3(4096), dim3(BFJ_BLOCK), 0, st, a);
}
