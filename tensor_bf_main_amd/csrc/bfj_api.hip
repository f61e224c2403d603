// bfj_api.hip -- C ABI and host orchestration of one join pass.
//
// Replaces the body of GPUJoinMainBruteForce (gpu_join.cu:210-322):
//   reference                               here
//   cudaMalloc + H2D of the padded rows     device buffers owned by a context, grown
//     (:221-222), never freed                 on demand, freed by bfj_destroy
//   convertDataset / preComputedSquared-    prep kernels (prep.hip): centre, scale,
//     Coordinates (:224-255)                  cast into MFMA fragment order, norms,
//                                             per-row error radii and thresholds
//   one join kernel per searchMode          MFMA join with fused epsilon epilogue
//     (:271-316)                              (join_kernel.h), then the FP64 band
//                                             refine and count finalisation (refine.hip)
//   cudaDeviceSynchronize + D2H of one      per-row int64 counts + totals per eps;
//     uint64 (:318-321)                       status codes instead of exit(1)
#include "bfj_internal.h"
#include "../../include/bfjoin.h"
static_assert(BFJ_CHUNK == BFJ_CHUNK_RECORDS, "candidate chunk size of the kernels and of the public header");

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HC(expr)                                                                               \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? BFJ_ERR_NOMEM : BFJ_ERR_DEVICE, "%s: %s", #expr, \
                        hipGetErrorString(e_));                                                \
    } while (0)

#define RC(expr)                   \
    do {                           \
        int r_ = (expr);           \
        if (r_ != BFJ_OK) return r_; \
    } while (0)

// ---------------------------------------------------------------------------
// dispatch to the per-dtype translation units
int bfj_compat_run(const double *x, int64_t n, int d, int64_t ld, double eps, int mode, int cprec, int aprec, void *X,
                   float *cn, unsigned long long *cnt, hipStream_t st);
int bfj_join_launch_f16(const JoinLaunch &, const JoinArgs &, hipStream_t);
int bfj_join_launch_bf16(const JoinLaunch &, const JoinArgs &, hipStream_t);
int bfj_join_launch_f8(const JoinLaunch &, const JoinArgs &, hipStream_t);
int bfj_join_launch_f64(const JoinLaunch &, const JoinArgs &, hipStream_t);
bool bfj_join_has_f16(int, int);
bool bfj_join_has_bf16(int, int);
bool bfj_join_has_f8(int, int);
bool bfj_join_has_f64(int, int);
int bfj_join_occ_f16(int, int, int);
int bfj_join_occ_bf16(int, int, int);
int bfj_join_occ_f8(int, int, int);
int bfj_join_occ_f64(int, int, int);

int bfj_join_launch(const JoinLaunch &L, const JoinArgs &a, hipStream_t s)
{
    switch (L.dtype) {
    case BFJ_FP16: return bfj_join_launch_f16(L, a, s);
    case BFJ_BF16: return bfj_join_launch_bf16(L, a, s);
    case BFJ_FP8: return bfj_join_launch_f8(L, a, s);
    case BFJ_FP64: return bfj_join_launch_f64(L, a, s);
    }
    return 1;
}

bool bfj_join_has(int dtype, int kc, int rows)
{
    switch (dtype) {
    case BFJ_FP16: return bfj_join_has_f16(kc, rows);
    case BFJ_BF16: return bfj_join_has_bf16(kc, rows);
    case BFJ_FP8: return bfj_join_has_f8(kc, rows);
    case BFJ_FP64: return bfj_join_has_f64(kc, rows);
    }
    return false;
}

int bfj_join_max_blocks_per_cu(int dtype, int kc, int rows, int lds_bytes)
{
    switch (dtype) {
    case BFJ_FP16: return bfj_join_occ_f16(kc, rows, lds_bytes);
    case BFJ_BF16: return bfj_join_occ_bf16(kc, rows, lds_bytes);
    case BFJ_FP8: return bfj_join_occ_f8(kc, rows, lds_bytes);
    case BFJ_FP64: return bfj_join_occ_f64(kc, rows, lds_bytes);
    }
    return 0;
}

static const int kRowOptions[] = {16, 32, 64, 96, 128};

// Smallest built K-chunk count >= need for this dtype (0 if none).
static int pick_kc(int dtype, int d)
{
    const int need = (d + bfj_kchunk(dtype) - 1) / bfj_kchunk(dtype);
    for (int kc = need; kc <= 32; ++kc)
        for (int r : kRowOptions)
            if (bfj_join_has(dtype, kc, r)) return kc;
    return 0;
}

int bfj_join_lds_f16(int, int, int);
int bfj_join_lds_bf16(int, int, int);
int bfj_join_lds_f8(int, int, int);
int bfj_join_lds_f64(int, int, int);

// Dynamic LDS of one join block: candidate tile ring + per-wavefront state.
static int lds_for(int dtype, int kc, int rows, int n_eps)
{
    switch (dtype) {
    case BFJ_FP16: return bfj_join_lds_f16(kc, rows, n_eps);
    case BFJ_BF16: return bfj_join_lds_bf16(kc, rows, n_eps);
    case BFJ_FP8: return bfj_join_lds_f8(kc, rows, n_eps);
    case BFJ_FP64: return bfj_join_lds_f64(kc, rows, n_eps);
    }
    return 0;
}

// Work items of one K-chunk group for a panel of `panel` candidate blocks (the loop
// that builds them, counted only).
static int64_t count_items(const std::vector<SetDesc> &sets, int kc, int rows, int panel)
{
    const int strip = BFJ_JWAVES * rows / 16;
    int64_t n = 0;
    for (const SetDesc &s : sets) {
        if (s.kc != kc || s.row_end <= s.row_begin) continue;
        const int nblk = (int)(s.npad / 16);
        const int rb_begin = (int)(s.row_begin / 16), rb_end = (int)((s.row_end + 15) / 16);
        for (int c0 = 0; c0 < nblk; c0 += panel) {
            const int c1 = std::min(nblk, c0 + panel);
            for (int rb = rb_begin; rb < rb_end; rb += strip)
                if ((s.sym ? std::max(c0, rb & ~3) : c0) < c1) ++n;
        }
    }
    return n;
}

// Default tile rows R: the widest strip among those with the most resident blocks per
// CU, counted up to a cap (latency hiding across blocks matters more than strip width
// up to there).  Measured on MI355X:
//  * FP8 (twice the MFMA rate, so more issue and LDS work per flop) wants 4 blocks per
//    CU, FP16/BF16/FP64 3 in input order (R = 64 at d = 128);
//  * in the locality order hot tiles are rare and 2 blocks of wider strips win
//    (FP16 d = 128: R = 128, 76 ms against 78 ms at R = 64);
//  * FP16/BF16 with 1-2 K-chunks (d <= 64): a candidate block is too little MFMA work
//    to cover its epilogue, so resident wavefronts matter more than strip width
//    (n = 1M: R = 64 beats R = 128 by 8 % at d = 16 and d = 64).
static int heuristic_rows(int dtype, int kc, int n_eps, bool reordered)
{
    const bool low = dtype == BFJ_FP16 || dtype == BFJ_BF16;
    const int cap = dtype == BFJ_FP8 ? 4 : (low && kc <= 2 ? 4 : (low && reordered ? 2 : 3));
    int best = 0, best_occ = -1;
    for (int r : kRowOptions)
        if (bfj_join_has(dtype, kc, r)) {
            const int occ = std::min(cap, bfj_join_max_blocks_per_cu(dtype, kc, r, lds_for(dtype, kc, r, n_eps)));
            if (occ >= best_occ) {
                best = r;
                best_occ = occ;
            }
        }
    return best;
}

// Largest double T with sqrt(T) <= eps: the predicate of kernel_join.cu:194
// (sqrt(d2) <= eps) becomes d2 <= T.  eps < 0 matches nothing (-1), eps = inf
// matches everything.
static double eps_threshold(double eps)
{
    if (eps < 0.0) return -1.0;
    if (std::isinf(eps)) return INFINITY;
    double t = eps * eps;
    if (std::isinf(t)) t = 1.7976931348623157e308;
    while (t > 0.0 && std::sqrt(t) > eps) t = std::nextafter(t, -INFINITY);
    for (;;) {
        const double u = std::nextafter(t, INFINITY);
        if (std::isinf(u) || std::sqrt(u) > eps) break;
        t = u;
    }
    return t;
}

// ---------------------------------------------------------------------------
// context
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    uint64_t id = 0; // unique per allocation: a freed and re-made buffer never keeps its id
};

struct Group {
    int kc, rows, lds;
    int cap;         // resident join blocks per CU of the persistent grid (0: as many as fit)
    int panel_bytes; // candidate fragment bytes per work item
    int32_t item0, n_items;
};

struct bfj_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    DevBuf pack, nq, gq, he, row_set, qerr, dnq, nrm, scales, cnt, band, counters, colsum, mu, sets, scal, items, tabs, xcopy,
        outbuf, cand, chunk_fill, pairs, ref_x, ref_aux, perm, label, ohist, oseg, pivots, pivot_norm, mid_rows,
        mid_q, mid_band, blk_tot;
    int reorder = 1;   // pack whole-dataset passes in a locality order (order.hip)
    int mid_stage = 1; // low-precision passes: middle stage before the FP64 refine (refine.hip)
    uint64_t band_capacity = 0;  // records of the band list and of the middle stage's list
    uint64_t pair_capacity = 0; // pair-list mode (bfj_self_join_pairs) only
    int emit_pairs = 0;
    uint32_t chunk_capacity = 0; // candidate chunks of BFJ_CHUNK records
    uint64_t stream_limit = 0;   // bytes the candidate stream and band lists may take (0: half of mem_total)
    uint64_t mem_total = 0;      // the device's memory
    uint64_t mem_ceiling = 0;    // stream bytes known not to allocate (a failed growth; 0: none)
    int tile_rows[4][33] = {};  // [dtype][kc] forced/tuned R (0 = heuristic)
    int tile_cap[4][33] = {};   // [dtype][kc] tuned resident blocks per CU (0 = as many as fit)
    int tile_panel[4][33] = {}; // [dtype][kc] tuned panel bytes (0 = BFJ_PANEL_BYTES)
    // Work items of the last plan shape (repeated calls on the same shapes reuse them,
    // and skip the upload while the device copy is current).  Item lists are known by
    // a unique id, never by address: a list freed with its plan can be replaced by a
    // new one at the same address.
    std::vector<int64_t> item_key;
    std::shared_ptr<const std::vector<WorkItem>> item_cache;
    uint64_t item_cache_id = 0;
    std::vector<Group> group_cache;
    uint64_t items_staged_id = 0;     // item list whose copy the items buffer holds (0: none)
    uint64_t items_staged_buf_id = 0; // allocation of the items buffer it was copied into
    // The last pass that ran to completion on this context (bfj_pass_errors reads its
    // buffers); cleared when another pass starts or its rows may be gone.
    std::unique_ptr<struct PassPlan> last;
    DevBuf errbuf;
    void *scratch = nullptr; // device address of this GPU's no-allocation scratch array
    int noalloc = 0;         // a no-allocation call is running: buffers come from scratch, never grow
};

// Bumped whenever a device buffer moves: a captured graph holds raw pointers.
static std::atomic<uint64_t> g_buf_gen{1};
// Source of the DevBuf and work-item list ids.
static std::atomic<uint64_t> g_obj_id{1};

// The no-allocation entry point's device memory: one file-scope array, which every GPU
// the library runs on holds (its own copy) until the process ends.  In host code the
// name is only a stand-in: the array's device address on each GPU comes from
// hipGetSymbolAddress with that GPU current (scratch_address) and is the only
// address host code ever uses for it.
#define BFJ_SCRATCH_BYTES ((size_t)BFJ_NOALLOC_SCRATCH_BYTES)
#define BFJ_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[BFJ_SCRATCH_BYTES];
static std::mutex g_scratch_mu;                          // guards g_scratch_ptr
static void *g_scratch_ptr[BFJ_MAX_DEVICES];             // per device: its array's address
static std::mutex g_scratch_call_mu[BFJ_MAX_DEVICES];    // per device: one no-allocation call at a time

// While a no-allocation call runs on this thread, ensure() carves buffers from its
// device's scratch array (a bump allocator) instead of calling hipMalloc.
struct Arena {
    unsigned char *base;
    size_t size, used;
};
static thread_local Arena *t_arena = nullptr;

// Allocation-failure injection (bfj_fail_allocation): the countdown-th device allocation
// from the time it was set fails as if the device were out of memory.
static std::atomic<int64_t> g_fail_countdown{0};

static int ensure(DevBuf &b, size_t bytes)
{
    if (bytes <= b.cap) return BFJ_OK;
    if (t_arena) {
        const size_t off = (t_arena->used + 255) & ~(size_t)255;
        if (off > t_arena->size || bytes > t_arena->size - off)
            return fail(BFJ_ERR_CAPACITY, "the pass needs more than the %zu bytes of the no-allocation scratch",
                        t_arena->size);
        b.p = t_arena->base + off;
        b.cap = bytes;
        b.id = g_obj_id.fetch_add(1);
        t_arena->used = off + bytes;
        return BFJ_OK;
    }
    g_buf_gen.fetch_add(1);
    if (b.p) HC(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    b.id = 0;
    size_t want = bytes + bytes / 8 + 256;
    if (g_fail_countdown.load() > 0 && g_fail_countdown.fetch_sub(1) == 1)
        return fail(BFJ_ERR_NOMEM, "hipMalloc(%zu bytes) failed: allocation failure injected", want);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return fail(BFJ_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    b.id = g_obj_id.fetch_add(1);
    return BFJ_OK;
}

static void release(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    b.id = 0;
}

// ---------------------------------------------------------------------------
// one pass over a set of datasets

struct HostSet {
    const double *x; // device
    int64_t n, ld, row_begin, row_end;
    int32_t d, n_eps;
    double eps[BFJ_MAXEPS];
    int64_t *out; // device or NULL
};

struct PassPlan {
    int dtype = 0;
    std::vector<SetDesc> sets;
    std::shared_ptr<const std::vector<WorkItem>> items; // shared with the context's cache
    uint64_t items_id = 0;                               // unique id of that list
    std::vector<Group> groups;
    std::vector<int32_t> blk_set, set_blk0, set_nblk;
    std::vector<int64_t> blk_row0;
    int64_t rows_total = 0, pack_bytes = 0, cnt_total = 0;
    uint64_t pairs = 0;
    int max_rows = 0, max_kc = 0;
    int sliced = 0;                     // the candidate stream does not fit: the join runs in slices
    unsigned long long stream_chunks = 0; // candidate chunks of the last sliced run (sizes the next one's slices)
};

static int validate(const HostSet &h, int dtype)
{
    if (!h.x && h.n > 0) return fail(BFJ_ERR_INVALID, "data is NULL");
    if (h.n < 0) return fail(BFJ_ERR_INVALID, "n < 0");
    if (h.n > (int64_t)0x7fffffff) return fail(BFJ_ERR_UNSUPPORTED, "n > 2^31-1");
    if (h.d < 1) return fail(BFJ_ERR_INVALID, "d < 1");
    if (h.d > BFJ_MAX_DIM) return fail(BFJ_ERR_UNSUPPORTED, "d = %d > %d", h.d, BFJ_MAX_DIM);
    if (h.ld < h.d) return fail(BFJ_ERR_INVALID, "ld < d");
    if (h.n_eps < 1 || h.n_eps > BFJ_MAXEPS) return fail(BFJ_ERR_INVALID, "n_eps must be in [1, %d]", BFJ_MAXEPS);
    for (int e = 0; e < h.n_eps; ++e)
        if (std::isnan(h.eps[e])) return fail(BFJ_ERR_INVALID, "eps[%d] is NaN", e);
    if (h.row_begin < 0 || h.row_end > h.n || h.row_begin > h.row_end)
        return fail(BFJ_ERR_INVALID, "row range [%lld, %lld) outside [0, %lld]", (long long)h.row_begin,
                    (long long)h.row_end, (long long)h.n);
    if (dtype < BFJ_FP64 || dtype > BFJ_FP8) return fail(BFJ_ERR_INVALID, "unknown dtype %d", dtype);
    if (!pick_kc(dtype, h.d)) return fail(BFJ_ERR_UNSUPPORTED, "no kernel built for dtype %d, d = %d", dtype, h.d);
    return BFJ_OK;
}

// Panel bytes for a group with no panel size set: BFJ_PANEL_BYTES, halved (down to
// 32 candidate blocks per item) while the group has fewer than two items per block of
// its persistent grid.  A small dataset otherwise runs a few long items on part of the
// chip (n = 10,000, FP64: 175 items of up to 256 blocks for 512 resident blocks).
static int small_pass_panel(const bfj_ctx *ctx, const PassPlan &p, const Group &g, int FB)
{
    const int64_t tile_bytes = (int64_t)g.kc * BFJ_WAVE * FB;
    int occ = bfj_join_max_blocks_per_cu(p.dtype, g.kc, g.rows, g.lds);
    if (g.cap > 0) occ = std::min(occ, g.cap);
    const int64_t want = 2ll * std::max(1, occ) * ctx->n_cu;
    int64_t bytes = BFJ_PANEL_BYTES;
    while (bytes / tile_bytes / 16 * 16 > 32 &&
           count_items(p.sets, g.kc, g.rows, (int)std::max<int64_t>(32, bytes / tile_bytes / 16 * 16)) < want)
        bytes /= 2;
    return (int)bytes;
}

static int build_plan(bfj_ctx *ctx, const std::vector<HostSet> &hs, int dtype, PassPlan &p, int force_rows)
{
    p = PassPlan();
    p.dtype = dtype;
    const int FB = bfj_frag_bytes(dtype);
    for (size_t si = 0; si < hs.size(); ++si) {
        const HostSet &h = hs[si];
        RC(validate(h, dtype));
        SetDesc s;
        memset(&s, 0, sizeof s);
        s.x = h.x;
        s.n = h.n;
        s.npad = (h.n + 63) / 64 * 64; // whole 4-block ring slots: no partial slot at the end
        s.ld = h.ld;
        s.row_begin = h.row_begin;
        s.row_end = h.row_end;
        s.d = h.d;
        s.kc = pick_kc(dtype, h.d);
        s.n_eps = h.n_eps;
        s.set_id = (int32_t)si;
        s.sym = (h.row_begin == 0 && h.row_end == h.n) ? 1 : 0;
        s.al16 = ((uintptr_t)h.x % 16 == 0 && h.ld % 2 == 0) ? 1 : 0;
        // A locality order only pays once a dataset spans many tiles.
        s.reorder = (ctx->reorder && s.sym && h.n >= 512) ? 1 : 0;
        s.out = h.out;
        // Thresholds in ascending order: a pair within eps k is within every later one,
        // so the kernels count each pair once, under the first eps that holds it.
        std::vector<std::pair<double, int>> te;
        for (int e = 0; e < h.n_eps; ++e) te.push_back({eps_threshold(h.eps[e]), e});
        std::stable_sort(te.begin(), te.end(),
                         [](const std::pair<double, int> &x, const std::pair<double, int> &y) { return x.first < y.first; });
        for (int e = 0; e < h.n_eps; ++e) {
            s.T[e] = te[e].first;
            s.eperm[e] = (int8_t)te[e].second;
        }
        s.row_off = p.rows_total;
        s.pack_off = p.pack_bytes;
        s.cnt_off = p.cnt_total;
        p.rows_total += s.npad;
        p.pack_bytes += (s.npad / 16) * (int64_t)s.kc * BFJ_WAVE * FB;
        p.cnt_total += s.npad * s.n_eps;
        p.pairs += (uint64_t)(h.row_end - h.row_begin) * (uint64_t)h.n;
        p.max_kc = std::max(p.max_kc, s.kc);
        // Row blocks for the row-parallel prep / finalize kernels.
        p.set_blk0.push_back((int32_t)p.blk_set.size());
        int32_t nb = 0;
        for (int64_t r0 = 0; r0 < s.npad; r0 += BFJ_PREP_ROWS, ++nb) {
            p.blk_set.push_back((int32_t)si);
            p.blk_row0.push_back(r0);
        }
        p.set_nblk.push_back(nb);
        p.sets.push_back(s);
    }
    if (p.sets.size() > 65535) return fail(BFJ_ERR_UNSUPPORTED, "more than 65535 datasets in one pass");
    // Candidate records and kernels index rows of a pass with 32-bit integers.
    if (p.rows_total > (int64_t)0x7fffffff)
        return fail(BFJ_ERR_UNSUPPORTED, "%lld padded rows in one pass (at most 2^31 - 1)", (long long)p.rows_total);
    // Work items grouped by K-chunk count, longest strips first (greedy LPT through the queue).
    std::vector<int> kcs;
    for (auto &s : p.sets)
        if (std::find(kcs.begin(), kcs.end(), s.kc) == kcs.end()) kcs.push_back(s.kc);
    std::sort(kcs.begin(), kcs.end());
    // Tile shape per K-chunk group, then the items (reused when the shapes repeat).
    std::vector<Group> groups;
    std::vector<int64_t> key = {dtype, (int64_t)p.sets.size()};
    for (int kc : kcs) {
        Group g;
        memset(&g, 0, sizeof g);
        g.kc = kc;
        int n_eps_max = 1;
        bool reordered = true;
        for (auto &s : p.sets)
            if (s.kc == kc) {
                n_eps_max = std::max(n_eps_max, (int)s.n_eps);
                reordered = reordered && s.reorder;
            }
        int rows = force_rows ? force_rows : ctx->tile_rows[dtype][kc];
        if (!rows || !bfj_join_has(dtype, kc, rows)) rows = heuristic_rows(dtype, kc, n_eps_max, reordered);
        if (!bfj_join_has(dtype, kc, rows)) return fail(BFJ_ERR_UNSUPPORTED, "no tile for dtype %d kc %d", dtype, kc);
        g.rows = rows;
        g.lds = lds_for(dtype, kc, rows, n_eps_max);
        g.cap = ctx->tile_cap[dtype][kc];
        g.panel_bytes = ctx->tile_panel[dtype][kc] ? ctx->tile_panel[dtype][kc] : BFJ_PANEL_BYTES;
        if (!ctx->tile_panel[dtype][kc]) g.panel_bytes = small_pass_panel(ctx, p, g, FB);
        groups.push_back(g);
        key.insert(key.end(), {kc, rows, g.lds, g.cap, g.panel_bytes});
    }
    for (auto &s : p.sets) key.insert(key.end(), {s.kc, s.npad, s.sym, s.row_begin, s.row_end});
    if (ctx->item_cache && key == ctx->item_key) {
        p.items = ctx->item_cache;
        p.items_id = ctx->item_cache_id;
        p.groups = ctx->group_cache;
        for (auto &g : p.groups) p.max_rows = std::max(p.max_rows, g.rows);
        return BFJ_OK;
    }
    auto items = std::make_shared<std::vector<WorkItem>>();
    for (Group &g : groups) {
        const int kc = g.kc, rows = g.rows;
        g.item0 = (int32_t)items->size();
        // Items: one block strip (JWAVES wavefronts x R rows) against one panel of
        // about 1 MiB of candidate fragments.  Items run panel by panel, so the blocks
        // in flight on an XCD sweep the same panel and its tiles are fetched into that
        // XCD's L2 once instead of once per strip.  Datasets largest first.
        const int64_t tile_bytes = (int64_t)kc * BFJ_WAVE * FB;
        const int panel = (int)std::max<int64_t>(32, (int64_t)g.panel_bytes / tile_bytes / 16 * 16);
        const int strip = BFJ_JWAVES * rows / 16;
        std::vector<size_t> order;
        for (size_t si = 0; si < p.sets.size(); ++si)
            if (p.sets[si].kc == kc && p.sets[si].row_end > p.sets[si].row_begin) order.push_back(si);
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return p.sets[x].npad > p.sets[y].npad; });
        for (size_t si : order) {
            const SetDesc &s = p.sets[si];
            const int nblk = (int)(s.npad / 16);
            const int rb_begin = (int)(s.row_begin / 16), rb_end = (int)((s.row_end + 15) / 16);
            for (int c0 = 0; c0 < nblk; c0 += panel) {
                const int c1 = std::min(nblk, c0 + panel);
                for (int rb = rb_begin; rb < rb_end; rb += strip) {
                    WorkItem w;
                    w.set = (int32_t)si;
                    w.rb0 = rb;
                    // Panels, strips and nblk are multiples of 4 blocks, so an item is whole
                    // ring slots (a slot holds 1, 2 or 4 blocks).  Blocks left of the
                    // diagonal that this adds are dropped by the kernel.
                    w.cb0 = s.sym ? std::max(c0, rb & ~3) : c0;
                    w.cb1 = c1;
                    if (w.cb0 < w.cb1) items->push_back(w);
                }
            }
        }
        if (items->size() > (size_t)0x7fffffff) return fail(BFJ_ERR_UNSUPPORTED, "too many join work items");
        g.n_items = (int32_t)(items->size() - g.item0);
        p.max_rows = std::max(p.max_rows, rows);
        if (g.n_items) p.groups.push_back(g);
    }
    p.items = items;
    p.items_id = g_obj_id.fetch_add(1);
    ctx->item_key = key;
    ctx->item_cache = items;
    ctx->item_cache_id = p.items_id;
    ctx->group_cache = p.groups;
    return BFJ_OK;
}

// Bytes the candidate stream and the band lists of a pass may take.
static uint64_t stream_limit(const bfj_ctx *ctx)
{
    const uint64_t lim = ctx->stream_limit ? ctx->stream_limit : ctx->mem_total / 2;
    return ctx->mem_ceiling ? std::min(lim, ctx->mem_ceiling) : lim;
}

// Band lists of a pass: the band list, and the middle stage's list when it runs.
static uint64_t band_lists(const bfj_ctx *ctx, const PassPlan &p)
{
    return (p.dtype != BFJ_FP64 && ctx->mid_stage) ? 2 : 1;
}

// Capacities of the candidate stream and the band lists cut in proportion to fit the
// stream limit (a pass that needs more then overflows and runs in slices).
static void clamp_to_limit(bfj_ctx *ctx, const PassPlan &p)
{
    if (ctx->noalloc) return;
    const double bytes = (double)ctx->chunk_capacity * BFJ_CHUNK * sizeof(Candidate) +
                         (double)ctx->band_capacity * sizeof(BandEntry) * (double)band_lists(ctx, p);
    const double limit = (double)stream_limit(ctx);
    if (bytes <= limit) return;
    const double f = limit / bytes;
    ctx->chunk_capacity = (uint32_t)std::max(1.0, (double)ctx->chunk_capacity * f);
    ctx->band_capacity = (uint64_t)std::max(1024.0, (double)ctx->band_capacity * f);
}

// Buffers sized for the plan; descriptors uploaded.
static int stage_plan(bfj_ctx *ctx, PassPlan &p)
{
    const int dt = p.dtype;
    const size_t vb = bfj_acc_bytes(dt);
    const size_t nrows = (size_t)std::max<int64_t>(p.rows_total, 1);
    // The join prefetches up to NB-1 ring slots past the end of an item: after the
    // last dataset that lands in this padding (never used).
    RC(ensure(ctx->pack, (size_t)p.pack_bytes + (256 << 10)));
    RC(ensure(ctx->nq, (nrows + 1024) * vb));
    RC(ensure(ctx->gq, nrows * vb));
    RC(ensure(ctx->he, nrows * 16));
    RC(ensure(ctx->row_set, nrows * 2));
    RC(ensure(ctx->qerr, nrows * 8));
    RC(ensure(ctx->dnq, nrows * 8));
    RC(ensure(ctx->nrm, nrows * 8));
    RC(ensure(ctx->scales, nrows + 1024));
    RC(ensure(ctx->cnt, std::max<int64_t>(p.cnt_total, 1) * 4));
    if (ctx->band_capacity == 0)
        ctx->band_capacity = std::min<uint64_t>(std::max<uint64_t>((uint64_t)p.rows_total * 64, 1u << 20), 1u << 28);
    if (ctx->chunk_capacity == 0) {
        // Start at 96 candidates per row (1.5 GiB at a million rows; about 100 neighbours
        // per point fit) and at least 16 Mi records (256 MiB: two chunks for each of
        // 2048 join wavefronts); a pass that needs more is repeated with the size it
        // reported.
        const uint64_t want = std::max<uint64_t>((uint64_t)p.rows_total * 96, 16ull << 20) / BFJ_CHUNK;
        ctx->chunk_capacity = (uint32_t)std::min<uint64_t>(want, (1ull << 34) / BFJ_CHUNK);
    }
    clamp_to_limit(ctx, p);
    RC(ensure(ctx->band, (size_t)ctx->band_capacity * sizeof(BandEntry)));
    if (p.dtype != BFJ_FP64 && ctx->mid_stage) {
        RC(ensure(ctx->mid_rows, nrows * BFJ_MAXDIM * (p.dtype == BFJ_FP8 ? 2 : 4)));
        RC(ensure(ctx->mid_q, nrows * 8));
        RC(ensure(ctx->mid_band, (size_t)ctx->band_capacity * sizeof(BandEntry)));
    }
    RC(ensure(ctx->cand, (size_t)ctx->chunk_capacity * BFJ_CHUNK * sizeof(Candidate)));
    RC(ensure(ctx->chunk_fill, (size_t)ctx->chunk_capacity * 4));
    RC(ensure(ctx->counters, BFJ_CTR_SLOTS * 8));
    RC(ensure(ctx->colsum, std::max<size_t>(p.blk_set.size(), 1) * BFJ_MAXDIM * 8 * 3)); // sums, maxima, minima
    RC(ensure(ctx->mu, p.sets.size() * BFJ_MAXDIM * 8));
    RC(ensure(ctx->sets, p.sets.size() * sizeof(SetDesc)));
    RC(ensure(ctx->scal, p.sets.size() * sizeof(SetScalars)));
    const size_t n_items = p.items ? p.items->size() : 0;
    RC(ensure(ctx->items, std::max<size_t>(n_items, 1) * sizeof(WorkItem)));
    const size_t nb = p.blk_set.size(), ns = p.sets.size();
    const size_t tab_bytes = nb * 4 + nb * 8 + ns * 4 * 2 + 64;
    RC(ensure(ctx->tabs, tab_bytes));
    RC(ensure(ctx->blk_tot, std::max<size_t>(nb, 1) * BFJ_MAXEPS * 8));
    int max_nblk = 1;
    for (int32_t v : p.set_nblk) max_nblk = std::max(max_nblk, (int)v);
    const size_t max_seg = (size_t)(max_nblk + BFJ_ORD_SEG - 1) / BFJ_ORD_SEG;
    RC(ensure(ctx->perm, nrows * 4));
    RC(ensure(ctx->label, nrows));
    RC(ensure(ctx->ohist, std::max<size_t>(nb, 1) * BFJ_PIVOTS * 4));
    RC(ensure(ctx->oseg, std::max<size_t>(ns, 1) * max_seg * BFJ_PIVOTS * 4));
    RC(ensure(ctx->pivots, std::max<size_t>(ns, 1) * BFJ_PIVOTS * BFJ_MAXDIM * 2));
    RC(ensure(ctx->pivot_norm, std::max<size_t>(ns, 1) * BFJ_PIVOTS * 4));
    HC(hipMemcpyAsync(ctx->sets.p, p.sets.data(), ns * sizeof(SetDesc), hipMemcpyHostToDevice, ctx->stream));
    if (n_items && (ctx->items_staged_id != p.items_id || ctx->items_staged_buf_id != ctx->items.id)) {
        ctx->items_staged_id = 0;
        HC(hipMemcpyAsync(ctx->items.p, p.items->data(), n_items * sizeof(WorkItem), hipMemcpyHostToDevice,
                          ctx->stream));
        ctx->items_staged_id = p.items_id;
        ctx->items_staged_buf_id = ctx->items.id;
    }
    unsigned char *t = (unsigned char *)ctx->tabs.p;
    HC(hipMemcpyAsync(t, p.blk_row0.data(), nb * 8, hipMemcpyHostToDevice, ctx->stream));
    HC(hipMemcpyAsync(t + nb * 8, p.blk_set.data(), nb * 4, hipMemcpyHostToDevice, ctx->stream));
    HC(hipMemcpyAsync(t + nb * 12, p.set_blk0.data(), ns * 4, hipMemcpyHostToDevice, ctx->stream));
    HC(hipMemcpyAsync(t + nb * 12 + ns * 4, p.set_nblk.data(), ns * 4, hipMemcpyHostToDevice, ctx->stream));
    return BFJ_OK;
}

// k_classify's wavefronts reserve band slots in blocks; a block as large as a quarter
// of the capacity spread over all of them (64..1024 slots) bounds the holes its partly
// filled last blocks leave.
static uint32_t band_block_for(uint64_t band_capacity)
{
    const uint64_t share = band_capacity / (4ull * BFJ_CLASSIFY_WAVES);
    uint32_t blk = 64;
    while (blk < 1024 && (uint64_t)blk * 2 <= share) blk *= 2;
    return blk;
}

static PassArgs pass_args(bfj_ctx *ctx, const PassPlan &p)
{
    PassArgs a;
    memset(&a, 0, sizeof a);
    const size_t nb = p.blk_set.size(), ns = p.sets.size();
    unsigned char *t = (unsigned char *)ctx->tabs.p;
    a.sets = (const SetDesc *)ctx->sets.p;
    a.scal = (SetScalars *)ctx->scal.p;
    a.blk_tot = (unsigned long long *)ctx->blk_tot.p;
    a.n_sets = (int32_t)ns;
    a.dtype = p.dtype;
    a.blk_row0 = (const int64_t *)t;
    a.blk_set = (const int32_t *)(t + nb * 8);
    a.set_blk0 = (const int32_t *)(t + nb * 12);
    a.set_nblk = (const int32_t *)(t + nb * 12 + ns * 4);
    a.n_row_blocks = (int32_t)nb;
    a.colsum = (double *)ctx->colsum.p;
    a.mu = (double *)ctx->mu.p;
    a.pack = (unsigned char *)ctx->pack.p;
    a.nq = ctx->nq.p;
    a.gq = ctx->gq.p;
    a.he = (double2 *)ctx->he.p;
    a.row_set = (uint16_t *)ctx->row_set.p;
    a.qerr = (double *)ctx->qerr.p;
    a.dnq = (double *)ctx->dnq.p;
    a.nrm = (double *)ctx->nrm.p;
    a.scales = (uint8_t *)ctx->scales.p;
    a.cnt = (uint32_t *)ctx->cnt.p;
    a.cnt_len = p.cnt_total;
    unsigned long long *ctr = (unsigned long long *)ctx->counters.p;
    a.band = (BandEntry *)ctx->band.p;
    a.band_count = ctr + BFJ_CTR_BAND;
    a.band_capacity = ctx->band_capacity;
    a.band_records = ctr + BFJ_CTR_BANDREC;
    a.band_block = band_block_for(ctx->band_capacity);
    a.mid = (p.dtype != BFJ_FP64 && ctx->mid_stage) ? 1 : 0;
    a.mid_fp32 = p.dtype != BFJ_FP8 ? 1 : 0;
    a.mid_rows = (unsigned char *)ctx->mid_rows.p;
    a.mid_q = (double *)ctx->mid_q.p;
    a.mid_band = (BandEntry *)ctx->mid_band.p;
    a.mid_count = ctr + BFJ_CTR_MID;
    a.cand = (const Candidate *)ctx->cand.p;
    a.chunk_fill = (const uint32_t *)ctx->chunk_fill.p;
    a.chunk_count = ctr + BFJ_CTR_CHUNKS;
    a.chunk_capacity = ctx->chunk_capacity;
    a.cand_records = ctr + BFJ_CTR_CANDS;
    a.emit_pairs = ctx->emit_pairs;
    a.pairs = (uint2 *)ctx->pairs.p;
    a.pair_count = ctr + BFJ_CTR_PAIRS;
    a.pair_capacity = ctx->pair_capacity;
    int max_nblk = 1;
    for (int32_t v : p.set_nblk) max_nblk = std::max(max_nblk, (int)v);
    a.perm = (uint32_t *)ctx->perm.p;
    a.label = (uint8_t *)ctx->label.p;
    a.ohist = (uint32_t *)ctx->ohist.p;
    a.oseg = (uint32_t *)ctx->oseg.p;
    a.oseg_stride = (max_nblk + BFJ_ORD_SEG - 1) / BFJ_ORD_SEG;
    a.pivots = (_Float16 *)ctx->pivots.p;
    a.pivot_norm = (float *)ctx->pivot_norm.p;
    return a;
}

// The pass's counts, scalars and counters zeroed, then the prep kernels.
static int enqueue_prep(bfj_ctx *ctx, const PassPlan &p, const PassArgs &a)
{
    hipStream_t st = ctx->stream;
    bfj_launch_zero(ctx->cnt.p, (size_t)std::max<int64_t>(p.cnt_total, 1) * 4, st);
    bfj_launch_zero(ctx->scal.p, p.sets.size() * sizeof(SetScalars), st);
    bfj_launch_zero(ctx->counters.p, BFJ_CTR_SLOTS * 8, st);
    if (p.blk_set.size()) bfj_launch_prep(a, st);
    HC(hipGetLastError());
    return BFJ_OK;
}

// The join over the work items [lo, hi) of the pass's list: one launch per K-chunk
// group the range meets, each with its own work-queue head (zeroed beforehand).
static int enqueue_join(bfj_ctx *ctx, const PassPlan &p, const PassArgs &a, size_t lo, size_t hi)
{
    hipStream_t st = ctx->stream;
    uint32_t *item_ctr = (uint32_t *)((unsigned long long *)ctx->counters.p + BFJ_CTR_ITEMS);
    for (size_t gi = 0; gi < p.groups.size(); ++gi) {
        const Group &g = p.groups[gi];
        const size_t b = std::max(lo, (size_t)g.item0), e = std::min(hi, (size_t)g.item0 + (size_t)g.n_items);
        if (b >= e) continue;
        JoinArgs j;
        memset(&j, 0, sizeof j);
        j.sets = a.sets;
        j.items = (const WorkItem *)ctx->items.p + b;
        j.n_items = (int32_t)(e - b);
        j.item_counter = item_ctr + gi % 32;
        j.pack = a.pack;
        j.nq = a.nq;
        j.gq = a.gq;
        j.scales = a.scales;
        j.cand = (Candidate *)ctx->cand.p;
        j.chunk_fill = (uint32_t *)ctx->chunk_fill.p;
        j.chunk_count = (unsigned long long *)ctx->counters.p + BFJ_CTR_CHUNKS;
        j.chunk_capacity = ctx->chunk_capacity;
        JoinLaunch L;
        L.dtype = p.dtype;
        L.kc = g.kc;
        L.rows = g.rows;
        L.lds_bytes = g.lds;
        int occ = bfj_join_max_blocks_per_cu(p.dtype, g.kc, g.rows, g.lds);
        if (g.cap > 0) occ = std::min(occ, g.cap);
        const int persistent = std::max(1, occ) * ctx->n_cu;
        L.grid_blocks = std::min(persistent, j.n_items);
        if (L.grid_blocks < 1) L.grid_blocks = 1;
        const int rc = bfj_join_launch(L, j, st);
        if (rc == 1) return fail(BFJ_ERR_INTERNAL, "join kernel dtype %d kc %d rows %d not built", p.dtype, g.kc, g.rows);
        if (rc) return fail(BFJ_ERR_DEVICE, "join launch failed: %s", hipGetErrorString(hipGetLastError()));
    }
    return BFJ_OK;
}

// Band tests of the candidate stream, the middle stage and the FP64 refine.
static void enqueue_downstream(const PassArgs &a, hipStream_t st)
{
    bfj_launch_classify(a, st);
    if (a.mid) {
        bfj_launch_mid(a, st);
        PassArgs r = a;
        r.band = a.mid_band;
        r.band_count = a.mid_count;
        bfj_launch_refine(r, st);
    } else {
        bfj_launch_refine(a, st);
    }
}

// Enqueue the whole pass on the context stream (no host synchronisation, no
// memset nodes: safe to capture in a graph).
static int enqueue_pass(bfj_ctx *ctx, const PassPlan &p, bool timed)
{
    ctx->last.reset();
    const PassArgs a = pass_args(ctx, p);
    hipStream_t st = ctx->stream;
    if (timed) HC(hipEventRecord(ctx->ev[0], st));
    RC(enqueue_prep(ctx, p, a));
    if (timed) HC(hipEventRecord(ctx->ev[1], st));
    RC(enqueue_join(ctx, p, a, 0, p.items ? p.items->size() : 0));
    if (timed) HC(hipEventRecord(ctx->ev[2], st));
    enqueue_downstream(a, st);
    bfj_launch_finalize(a, st);
    HC(hipGetLastError());
    if (timed) HC(hipEventRecord(ctx->ev[3], st));
    return BFJ_OK;
}

static double bits_to_double(unsigned long long b)
{
    double d;
    memcpy(&d, &b, 8);
    return d;
}

// The kernels assume finite rows; a pass over NaN or inf data is reported, not returned.
static int check_finite(const std::vector<SetScalars> &sc)
{
    for (size_t i = 0; i < sc.size(); ++i)
        if (sc[i].nonfinite) return fail(BFJ_ERR_INVALID, "dataset %zu contains NaN or infinite values", i);
    return BFJ_OK;
}

// After a pass: if the candidate stream or the refine list overflowed (records
// past the capacity were dropped), grow both to what the pass asked for and set
// *again so that the caller repeats the pass.  When the grown buffers would exceed
// the stream limit (or a no-allocation call cannot grow them), set *slice instead:
// the caller then runs the pass in slices (run_sliced).
#define BFJ_MAX_CHUNKS ((1ull << 36) / BFJ_CHUNK) // 64 Gi candidate records (1 TiB): more is refused
#define BFJ_MAX_BAND (1ull << 34)     // band-list slots (256 GiB per list): more is refused
static int grow_on_overflow(bfj_ctx *ctx, const PassPlan &p, const unsigned long long *ctr, bool *again,
                            bool *slice)
{
    *again = false;
    *slice = false;
    const unsigned long long band_count = ctr[BFJ_CTR_BAND], chunks = ctr[BFJ_CTR_CHUNKS];
    const bool over_c = chunks > ctx->chunk_capacity, over_b = band_count > ctx->band_capacity;
    if (!over_c && !over_b) return BFJ_OK;
    const unsigned long long want_c =
        over_c ? std::min<unsigned long long>(chunks + chunks / 4 + 64, BFJ_MAX_CHUNKS) : ctx->chunk_capacity;
    const unsigned long long want_b =
        over_b ? std::min<unsigned long long>(band_count + band_count / 2 + 1024, BFJ_MAX_BAND) : ctx->band_capacity;
    const double bytes = (double)want_c * BFJ_CHUNK * sizeof(Candidate) +
                         (double)want_b * sizeof(BandEntry) * (double)band_lists(ctx, p);
    if (ctx->noalloc || bytes > (double)stream_limit(ctx) || chunks > BFJ_MAX_CHUNKS || band_count > BFJ_MAX_BAND) {
        *slice = true;
        return BFJ_OK;
    }
    // A growth that the device cannot allocate (other allocations hold its memory) runs
    // the pass in slices instead, under a ceiling of half the bytes that failed.
    auto grow = [&](DevBuf &b, size_t n) {
        const int rc = ensure(b, n);
        if (rc == BFJ_ERR_NOMEM) {
            ctx->mem_ceiling = (uint64_t)(bytes / 2);
            *slice = true;
        }
        return rc == BFJ_ERR_NOMEM ? BFJ_OK : rc;
    };
    if (over_c) {
        ctx->chunk_capacity = (uint32_t)want_c;
        RC(grow(ctx->cand, (size_t)ctx->chunk_capacity * BFJ_CHUNK * sizeof(Candidate)));
        if (!*slice) RC(grow(ctx->chunk_fill, (size_t)ctx->chunk_capacity * 4));
        *again = true;
    }
    if (over_b && !*slice) {
        // band_count counts reserved slots: the records plus the holes of classify's
        // last blocks.  The larger capacity may give larger blocks (pass_args); while
        // a block is above its 64-slot minimum the holes stay within a quarter of the
        // capacity, so 1.5x the slots holds the same pass again (records + capacity / 4
        // <= slots + capacity / 4 <= capacity).  At the minimum the holes are at most
        // 8192 x 64 slots, which the growth covers within a repeat or two more.
        ctx->band_capacity = want_b;
        RC(grow(ctx->band, (size_t)ctx->band_capacity * sizeof(BandEntry)));
        if (ctx->mid_band.p && !*slice) RC(grow(ctx->mid_band, (size_t)ctx->band_capacity * sizeof(BandEntry)));
        *again = true;
    }
    if (*slice) *again = false;
    return BFJ_OK;
}

// Totals and statistics of a finished pass (no overflow): ctr holds its counters,
// retries the passes repeated before it.
static int collect_pass(bfj_ctx *ctx, const PassPlan &p, const unsigned long long *ctr, int retries, bool timed,
                        std::vector<SetScalars> &sc, bfj_stats *stats)
{
    sc.assign(p.sets.size(), SetScalars());
    HC(hipMemcpy(sc.data(), ctx->scal.p, sc.size() * sizeof(SetScalars), hipMemcpyDeviceToHost));
    RC(check_finite(sc));
    if (!stats) return BFJ_OK;
    memset(stats, 0, sizeof *stats);
    stats->pairs = p.pairs;
    stats->band_pairs = ctr[BFJ_CTR_BANDREC];
    stats->fp64_pairs = (p.dtype != BFJ_FP64 && ctx->mid_stage) ? ctr[BFJ_CTR_MID] : ctr[BFJ_CTR_BANDREC];
    stats->candidates = ctr[BFJ_CTR_CANDS];
    stats->band_capacity = ctx->band_capacity;
    stats->band_retries = retries;
    stats->tile_rows = p.max_rows;
    stats->k_padded = p.max_kc * bfj_kchunk(p.dtype);
    stats->dtype = p.dtype;
    for (auto &s : sc) stats->max_band_halfwidth = std::max(stats->max_band_halfwidth, bits_to_double(s.halfband_bits));
    if (timed) {
        float t01 = 0, t12 = 0, t23 = 0, t03 = 0;
        HC(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        HC(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
        HC(hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]));
        HC(hipEventElapsedTime(&t03, ctx->ev[0], ctx->ev[3]));
        stats->ms_prep = t01;
        stats->ms_join = t12;
        stats->ms_refine = t23;
        stats->ms_total = t03;
    }
    return BFJ_OK;
}

// The counters of the pass just enqueued, after the stream has drained.
static int read_counters(bfj_ctx *ctx, unsigned long long *ctr)
{
    HC(hipMemcpyAsync(ctr, ctx->counters.p, BFJ_CTR_ITEMS * 8, hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    return BFJ_OK;
}

// The totals and statistics of a pass that ran to completion; it becomes the
// context's last pass.
static int finish_pass(bfj_ctx *ctx, const PassPlan &p, const unsigned long long *ctr, int retries, bool timed,
                       std::vector<std::vector<uint64_t>> *totals_out, bfj_stats *stats)
{
    std::vector<SetScalars> sc;
    RC(collect_pass(ctx, p, ctr, retries, timed, sc, stats));
    if (stats) stats->slices = 1;
    ctx->last.reset(new PassPlan(p));
    if (totals_out) {
        totals_out->assign(p.sets.size(), {});
        for (size_t si = 0; si < p.sets.size(); ++si)
            (*totals_out)[si].assign(sc[si].total, sc[si].total + p.sets[si].n_eps);
    }
    return BFJ_OK;
}

// A pass whose candidate stream does not fit the stream limit (or the fixed buffers
// of a no-allocation call): prep once, then the join's work items in slices, each
// slice followed by its band tests, middle stage and FP64 refine, which add to the
// same per-row counts; finalize once at the end.  Every pair is still decided by
// exactly one record of one slice, so the counts are those of a single pass.
//
// The band lists are sized so that a slice whose join fits the candidate stream fits
// them too, and no slice is repeated once its counts were added: a slice of c chunks
// holds at most cC records (C = BFJ_CHUNK), and k_classify leaves holes only in the last
// block of each wavefront that had a chunk (min(c, 8192) blocks of band_block_for(B)
// slots).  With B = 4/3 x CK + 64 min(K, 8192) + 1024 slots for K chunks, the records and
// holes of any c <= K chunks fit: at 64-slot blocks directly, and at larger blocks
// the holes are within B/4 (band_block_for) and the records within 3B/4.  A slice
// whose join needs more chunks than that is repeated with half its items before
// anything else runs on it.  Host synchronisation after each slice's join and band
// tests: eager only, never captured.  chunks_hint, the chunks the whole pass needs
// (0: unknown), sizes the first slice.
static int run_sliced(bfj_ctx *ctx, PassPlan &p, unsigned long long chunks_hint,
                      std::vector<std::vector<uint64_t>> *totals_out, bfj_stats *stats, int retries)
{
    const uint64_t W = BFJ_CLASSIFY_WAVES;
    // Buffers the device cannot allocate halve the ceiling and are tried again (at most
    // 8 times, then the out-of-memory error stands).
    for (int tries = 0; !ctx->noalloc; ++tries) {
        // Candidate chunks K and band slots 4/3 x CK + 64 min(K, 8192) + 1024 per
        // list, together within the stream limit.
        const double lists = (double)band_lists(ctx, p);
        const double per_chunk =
            (double)BFJ_CHUNK * sizeof(Candidate) + lists * sizeof(BandEntry) * (BFJ_CHUNK * 4.0 / 3.0 + 64.0);
        const double fixed = lists * sizeof(BandEntry) * 1025.0;
        const double k = std::min(((double)stream_limit(ctx) - fixed) / per_chunk, (double)BFJ_MAX_CHUNKS);
        if (k < 1.0)
            return fail(BFJ_ERR_CAPACITY, "the stream limit (%llu bytes) holds no candidate chunk and its band lists",
                        (unsigned long long)stream_limit(ctx));
        ctx->chunk_capacity = (uint32_t)k;
        ctx->band_capacity = ((uint64_t)ctx->chunk_capacity * BFJ_CHUNK * 4 + 2) / 3 +
                             64 * std::min<uint64_t>(ctx->chunk_capacity, W) + 1024;
        const int rc = stage_plan(ctx, p);
        if (rc == BFJ_OK) break;
        if (rc != BFJ_ERR_NOMEM || tries == 7) return rc;
        ctx->mem_ceiling = stream_limit(ctx) / 2;
    }
    // Chunks a slice may fill: its records and holes within the band capacity B.
    const uint64_t B = ctx->band_capacity, blk = band_block_for(B);
    uint64_t kr = B / (BFJ_CHUNK + blk);
    if (kr >= W) kr = (B - W * blk) / BFJ_CHUNK;
    kr = std::min<uint64_t>(kr, ctx->chunk_capacity);
    if (kr == 0)
        return fail(BFJ_ERR_CAPACITY, "a band list of %llu slots holds no slice of the candidate stream",
                    (unsigned long long)B);
    ctx->last.reset();
    const PassArgs a = pass_args(ctx, p);
    hipStream_t st = ctx->stream;
    const size_t n_items = p.items ? p.items->size() : 0;
    HC(hipEventRecord(ctx->ev[0], st));
    RC(enqueue_prep(ctx, p, a));
    HC(hipEventRecord(ctx->ev[1], st));
    static_assert(BFJ_CTR_BAND == 0 && BFJ_CTR_CHUNKS == 1 && BFJ_CTR_PAIRS == 2 && BFJ_CTR_MID == 3,
                  "counter slots zeroed per slice");
    unsigned long long *cdev = (unsigned long long *)ctx->counters.p;
    unsigned long long tot[BFJ_CTR_ITEMS] = {};
    double ms_join = 0.0, ms_down = 0.0;
    int slices = 0;
    // Items per slice: aimed at 3/4 of kr chunks.
    size_t step = chunks_hint ? (size_t)std::max(1.0, 0.75 * (double)n_items * (double)kr / (double)chunks_hint)
                              : n_items;
    for (size_t lo = 0; lo < n_items;) {
        size_t hi = std::min(n_items, lo + std::max<size_t>(step, 1));
        unsigned long long ctr[BFJ_CTR_ITEMS] = {};
        for (;;) {
            // Every counter but the pair list's, which the slices share, starts at zero.
            bfj_launch_zero(cdev, 16, st);
            bfj_launch_zero(cdev + 3, 8, st);
            bfj_launch_zero(cdev + 4, (BFJ_CTR_SLOTS - 4) * 8, st);
            HC(hipEventRecord(ctx->ev[2], st));
            RC(enqueue_join(ctx, p, a, lo, hi));
            HC(hipEventRecord(ctx->ev[3], st));
            RC(read_counters(ctx, ctr));
            if (ctr[BFJ_CTR_CHUNKS] <= kr) break;
            if (hi - lo == 1)
                return fail(BFJ_ERR_CAPACITY,
                            "one join work item streams %llu candidate chunks; a slice holds %llu (stream limit %llu "
                            "bytes)",
                            ctr[BFJ_CTR_CHUNKS], (unsigned long long)kr, (unsigned long long)stream_limit(ctx));
            hi = lo + (hi - lo) / 2;
            ++retries;
        }
        float t = 0.f;
        HC(hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]));
        ms_join += t;
        enqueue_downstream(a, st);
        HC(hipGetLastError());
        HC(hipEventRecord(ctx->ev[4], st));
        RC(read_counters(ctx, ctr));
        HC(hipEventElapsedTime(&t, ctx->ev[3], ctx->ev[4]));
        ms_down += t;
        if (ctr[BFJ_CTR_BAND] > B || ctr[BFJ_CTR_MID] > B)
            return fail(BFJ_ERR_INTERNAL, "a slice of %llu chunks overflowed its band list (%llu slots of %llu)",
                        ctr[BFJ_CTR_CHUNKS], std::max(ctr[BFJ_CTR_BAND], ctr[BFJ_CTR_MID]), (unsigned long long)B);
        for (int k : {BFJ_CTR_BAND, BFJ_CTR_CHUNKS, BFJ_CTR_MID, BFJ_CTR_CANDS, BFJ_CTR_BANDREC}) tot[k] += ctr[k];
        ++slices;
        const double grow = ctr[BFJ_CTR_CHUNKS] ? 0.75 * (double)kr / (double)ctr[BFJ_CTR_CHUNKS] : 4.0;
        step = (size_t)std::max(1.0, (double)(hi - lo) * std::min(grow, 4.0));
        lo = hi;
    }
    bfj_launch_finalize(a, st);
    HC(hipGetLastError());
    HC(hipEventRecord(ctx->ev[4], st));
    HC(hipStreamSynchronize(st));
    p.stream_chunks = tot[BFJ_CTR_CHUNKS];
    RC(finish_pass(ctx, p, tot, retries, false, totals_out, stats));
    if (stats) {
        float t01 = 0.f, t04 = 0.f;
        HC(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        HC(hipEventElapsedTime(&t04, ctx->ev[0], ctx->ev[4]));
        stats->ms_prep = t01;
        stats->ms_join = ms_join;
        stats->ms_refine = ms_down;
        stats->ms_total = t04;
        stats->slices = slices;
    }
    return BFJ_OK;
}

// Run a staged pass to completion, growing the candidate stream / refine list and
// repeating the pass when one overflowed, or running it in slices when they would
// outgrow the stream limit.  Totals per set land in totals_out[set][eps].
static int run_pass(bfj_ctx *ctx, PassPlan &p, std::vector<std::vector<uint64_t>> *totals_out, bfj_stats *stats,
                    bool timed = true)
{
    int retries = 0;
    unsigned long long hint = p.stream_chunks;
    while (!p.sliced) {
        RC(enqueue_pass(ctx, p, timed));
        unsigned long long ctr[BFJ_CTR_ITEMS] = {};
        RC(read_counters(ctx, ctr));
        bool again = false, slice = false;
        RC(grow_on_overflow(ctx, p, ctr, &again, &slice));
        if (slice) {
            p.sliced = 1;
            hint = ctr[BFJ_CTR_CHUNKS];
            ++retries;
            break;
        }
        if (!again) return finish_pass(ctx, p, ctr, retries, timed, totals_out, stats);
        if (++retries > 4)
            return fail(BFJ_ERR_INTERNAL, "candidate stream / band list keep overflowing (%llu chunks, %llu slots)",
                        ctr[BFJ_CTR_CHUNKS], ctr[BFJ_CTR_BAND]);
        RC(stage_plan(ctx, p)); // the grown buffers may have moved
    }
    return run_sliced(ctx, p, hint, totals_out, stats, retries);
}

// Uniform [0, 1) rows for the tile autotuner (splitmix64).
__global__ void k_synth(double *x, int64_t total, uint64_t seed)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)total; i += stride) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        x[i] = (double)(z >> 11) * 0x1p-53;
    }
}

// Every device buffer of a context.
static std::vector<DevBuf *> all_bufs(bfj_ctx *c)
{
    return {&c->pack, &c->nq, &c->gq, &c->he, &c->row_set, &c->qerr, &c->dnq, &c->nrm, &c->scales, &c->cnt,
            &c->band, &c->counters, &c->colsum, &c->mu, &c->sets, &c->scal, &c->items, &c->tabs,
            &c->xcopy, &c->outbuf, &c->cand, &c->chunk_fill, &c->pairs, &c->ref_x, &c->ref_aux,
            &c->perm, &c->label, &c->ohist, &c->oseg, &c->pivots, &c->pivot_norm, &c->mid_rows,
            &c->mid_q, &c->mid_band, &c->errbuf, &c->blk_tot};
}

// This device's scratch address, looked up once per device with the device current.
static int scratch_address(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= BFJ_MAX_DEVICES)
        return fail(BFJ_ERR_UNSUPPORTED, "device %d: the no-allocation scratch is kept for %d devices", device,
                    BFJ_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(BFJ_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        }
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return BFJ_OK;
}

namespace {

// For the length of a no-allocation call: the context's buffers, capacities and cached
// work items are set aside and every buffer starts empty (ensure() then carves from the
// arena); the destructor puts them back.
struct NoallocScope {
    bfj_ctx *c;
    std::vector<DevBuf> saved;
    uint64_t band_capacity, items_staged_id, items_staged_buf_id, item_cache_id;
    uint32_t chunk_capacity;
    std::vector<int64_t> item_key;
    std::shared_ptr<const std::vector<WorkItem>> item_cache;
    std::vector<Group> group_cache;
    explicit NoallocScope(bfj_ctx *ctx) : c(ctx)
    {
        for (DevBuf *b : all_bufs(c)) {
            saved.push_back(*b);
            *b = DevBuf();
        }
        band_capacity = c->band_capacity;
        chunk_capacity = c->chunk_capacity;
        items_staged_id = c->items_staged_id;
        items_staged_buf_id = c->items_staged_buf_id;
        item_key = c->item_key;
        item_cache = c->item_cache;
        item_cache_id = c->item_cache_id;
        group_cache = c->group_cache;
        c->last.reset();
        c->noalloc = 1;
    }
    void clear_bufs()
    {
        for (DevBuf *b : all_bufs(c)) *b = DevBuf();
    }
    ~NoallocScope()
    {
        size_t k = 0;
        for (DevBuf *b : all_bufs(c)) *b = saved[k++];
        c->band_capacity = band_capacity;
        c->chunk_capacity = chunk_capacity;
        c->items_staged_id = items_staged_id;
        c->items_staged_buf_id = items_staged_buf_id;
        c->item_key = item_key;
        c->item_cache = item_cache;
        c->item_cache_id = item_cache_id;
        c->group_cache = group_cache;
        c->last.reset(); // the pass's buffers were scratch: nothing to measure afterwards
        c->noalloc = 0;
        t_arena = nullptr;
    }
};

} // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

const char *bfj_last_error(void) { return g_err.c_str(); }

int bfj_abi_version(void) { return BFJ_ABI_VERSION; }

int bfj_device_count(int *count)
{
    if (!count) return fail(BFJ_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return BFJ_OK;
}

int bfj_init(int device, bfj_ctx **out)
{
    if (!out) return fail(BFJ_ERR_INVALID, "ctx out-pointer is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(BFJ_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= n) return fail(BFJ_ERR_DEVICE, "device %d out of range [0, %d)", device, n);
    HC(hipSetDevice(device));
    bfj_ctx *c = new bfj_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        for (auto &ev : c->ev)
            if (e == hipSuccess) e = hipEventCreate(&ev);
    }
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    size_t mem = 0;
    if (e == hipSuccess) e = hipDeviceTotalMem(&mem, device);
    c->mem_total = mem;
    if (e != hipSuccess) {
        bfj_destroy(c);
        return fail(BFJ_ERR_DEVICE, "context setup failed: %s", hipGetErrorString(e));
    }
    const int rc = scratch_address(device, &c->scratch);
    if (rc != BFJ_OK) {
        const std::string msg = g_err;
        bfj_destroy(c);
        return fail(rc, "%s", msg.c_str());
    }
    *out = c;
    return BFJ_OK;
}

int bfj_destroy(bfj_ctx *c)
{
    if (!c) return BFJ_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf *b : all_bufs(c)) release(*b);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BFJ_OK;
}

static int join_one(bfj_ctx *ctx, const double *d_x, int64_t n, int32_t d, int64_t ld, const double *eps,
                    int32_t n_eps, int32_t dtype, int64_t row_begin, int64_t row_end, int64_t *d_counts,
                    uint64_t *totals, bfj_stats *stats, int force_rows)
{
    HostSet h;
    memset(&h, 0, sizeof h);
    h.x = d_x;
    h.n = n;
    h.d = d;
    h.ld = ld;
    h.row_begin = row_begin;
    h.row_end = row_end;
    h.n_eps = n_eps;
    h.out = d_counts;
    for (int e = 0; e < n_eps && e < BFJ_MAXEPS; ++e) h.eps[e] = eps[e];
    if (n == 0) { // an empty dataset: nothing to launch, every total is 0
        RC(validate(h, dtype));
        if (totals)
            for (int e = 0; e < n_eps; ++e) totals[e] = 0;
        if (stats) memset(stats, 0, sizeof *stats);
        return BFJ_OK;
    }
    PassPlan p;
    RC(build_plan(ctx, {h}, dtype, p, force_rows));
    RC(stage_plan(ctx, p));
    std::vector<std::vector<uint64_t>> tot;
    RC(run_pass(ctx, p, &tot, stats));
    if (totals)
        for (int e = 0; e < n_eps; ++e) totals[e] = tot[0][e];
    return BFJ_OK;
}

// Bytes of n rows of d doubles at a stride of ld: the last row ends after its d-th
// value, which may be the end of the caller's buffer (a column slice of a wider array).
static size_t rows_bytes(int64_t n, int64_t d, int64_t ld)
{
    return n > 0 ? ((size_t)(n - 1) * (size_t)ld + (size_t)d) * 8 : 0;
}

static int common_checks(bfj_ctx *ctx, const double *data, int64_t n, const double *eps, int32_t n_eps)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    if (!eps) return fail(BFJ_ERR_INVALID, "eps is NULL");
    if (n_eps < 1 || n_eps > BFJ_MAXEPS) return fail(BFJ_ERR_INVALID, "n_eps must be in [1, %d]", BFJ_MAXEPS);
    if (n < 0) return fail(BFJ_ERR_INVALID, "n < 0");
    if (n > 0 && !data) return fail(BFJ_ERR_INVALID, "data is NULL");
    HC(hipSetDevice(ctx->device));
    return BFJ_OK;
}

int bfj_self_join_device(bfj_ctx *ctx, const double *d_data, int64_t n, int32_t d, int64_t ld, const double *eps,
                         int32_t n_eps, int32_t dtype, int64_t row_begin, int64_t row_end, int64_t *d_counts,
                         uint64_t *totals, bfj_stats *stats)
{
    RC(common_checks(ctx, d_data, n, eps, n_eps));
    return join_one(ctx, d_data, n, d, ld, eps, n_eps, dtype, row_begin, row_end, d_counts, totals, stats, 0);
}

int bfj_self_join_device_noalloc(bfj_ctx *ctx, const double *d_data, int64_t n, int32_t d, int64_t ld,
                                 const double *eps, int32_t n_eps, int32_t dtype, int64_t row_begin, int64_t row_end,
                                 int64_t *d_counts, uint64_t *totals, bfj_stats *stats)
{
    RC(common_checks(ctx, d_data, n, eps, n_eps));
    if (!ctx->scratch) return fail(BFJ_ERR_DEVICE, "no scratch address for device %d", ctx->device);
    std::lock_guard<std::mutex> lk(g_scratch_call_mu[ctx->device]);
    NoallocScope scope(ctx);
    Arena arena{(unsigned char *)ctx->scratch, BFJ_SCRATCH_BYTES, 0};
    t_arena = &arena;
    if (n > 0) {
        // Sizing pass (no device work): the buffers every pass of this shape needs, with
        // the smallest candidate stream and band lists; the rest of the scratch then goes
        // to those two: 60 % to the candidate stream, or a third for FP8, whose
        // candidates are nearly all band records (each kept in two lists of 16 bytes).
        HostSet h;
        memset(&h, 0, sizeof h);
        h.x = d_data;
        h.n = n;
        h.d = d;
        h.ld = ld;
        h.row_begin = row_begin;
        h.row_end = row_end;
        h.n_eps = n_eps;
        h.out = d_counts;
        for (int e = 0; e < n_eps; ++e) h.eps[e] = eps[e];
        PassPlan p;
        RC(build_plan(ctx, {h}, dtype, p, 0));
        // The smallest stream a no-allocation pass runs with: 64 Ki records.
        const size_t min_chunks = ((size_t)1 << 16) / BFJ_CHUNK;
        ctx->chunk_capacity = (uint32_t)min_chunks;
        ctx->band_capacity = 1024;
        const int rc = stage_plan(ctx, p);
        if (rc == BFJ_ERR_CAPACITY)
            return fail(BFJ_ERR_CAPACITY, "n = %lld, d = %d needs more than the %zu bytes of the no-allocation scratch",
                        (long long)n, d, BFJ_SCRATCH_BYTES);
        RC(rc);
        const size_t fixed = arena.used + 64 * 1024; // + alignment of every buffer
        const size_t rest = BFJ_SCRATCH_BYTES > fixed ? BFJ_SCRATCH_BYTES - fixed : 0;
        const size_t lists = (p.dtype != BFJ_FP64 && ctx->mid_stage) ? 2 : 1;
        const size_t cand_share = p.dtype == BFJ_FP8 ? rest / 3 : rest / 10 * 6;
        const size_t chunks = cand_share / ((size_t)BFJ_CHUNK * sizeof(Candidate) + 4);
        const size_t band = (rest - cand_share) / (lists * sizeof(BandEntry));
        if (chunks < min_chunks || band < 1024)
            return fail(BFJ_ERR_CAPACITY, "n = %lld, d = %d leaves too little of the %zu-byte no-allocation scratch "
                        "for the candidate stream", (long long)n, d, BFJ_SCRATCH_BYTES);
        ctx->chunk_capacity = (uint32_t)std::min<size_t>(chunks, BFJ_MAX_CHUNKS);
        ctx->band_capacity = band;
        scope.clear_bufs();
        arena.used = 0;
    }
    return join_one(ctx, d_data, n, d, ld, eps, n_eps, dtype, row_begin, row_end, d_counts, totals, stats, 0);
}

int bfj_self_join(bfj_ctx *ctx, const double *data, int64_t n, int32_t d, int64_t ld, const double *eps,
                  int32_t n_eps, int32_t dtype, int64_t row_begin, int64_t row_end, int64_t *counts,
                  uint64_t *totals, bfj_stats *stats)
{
    RC(common_checks(ctx, data, n, eps, n_eps));
    if (d < 1 || ld < d) return fail(BFJ_ERR_INVALID, "need 1 <= d <= ld");
    if (row_begin < 0 || row_end > n || row_begin > row_end) return fail(BFJ_ERR_INVALID, "bad row range");
    ctx->last.reset();
    const size_t xbytes = rows_bytes(n, d, ld);
    RC(ensure(ctx->xcopy, std::max<size_t>(xbytes, 8)));
    if (n > 0) HC(hipMemcpyAsync(ctx->xcopy.p, data, xbytes, hipMemcpyHostToDevice, ctx->stream));
    const size_t obytes = (size_t)std::max<int64_t>(row_end - row_begin, 1) * n_eps * 8;
    RC(ensure(ctx->outbuf, obytes));
    RC(join_one(ctx, (const double *)ctx->xcopy.p, n, d, ld, eps, n_eps, dtype, row_begin, row_end,
                counts ? (int64_t *)ctx->outbuf.p : nullptr, totals, stats, 0));
    if (counts && row_end > row_begin)
        HC(hipMemcpy(counts, ctx->outbuf.p, (size_t)(row_end - row_begin) * n_eps * 8, hipMemcpyDeviceToHost));
    return BFJ_OK;
}

int bfj_self_join_pairs(bfj_ctx *ctx, const double *data, int64_t n, int32_t d, int64_t ld, double eps, int32_t dtype,
                        uint32_t *pairs, uint64_t capacity, uint64_t *n_pairs, uint64_t *total, bfj_stats *stats)
{
    RC(common_checks(ctx, data, n, &eps, 1));
    if (!n_pairs) return fail(BFJ_ERR_INVALID, "n_pairs is NULL");
    *n_pairs = 0;
    if (capacity > 0 && !pairs) return fail(BFJ_ERR_INVALID, "pairs is NULL");
    if (d < 1 || ld < d) return fail(BFJ_ERR_INVALID, "need 1 <= d <= ld");
    ctx->last.reset();
    if (n > 0) {
        RC(ensure(ctx->xcopy, rows_bytes(n, d, ld)));
        HC(hipMemcpyAsync(ctx->xcopy.p, data, rows_bytes(n, d, ld), hipMemcpyHostToDevice, ctx->stream));
    }
    RC(ensure(ctx->pairs, (size_t)std::max<uint64_t>(capacity, 1) * 8));
    ctx->pair_capacity = capacity;
    ctx->emit_pairs = 1;
    uint64_t tot = 0;
    const int rc = join_one(ctx, (const double *)ctx->xcopy.p, n, d, ld, &eps, 1, dtype, 0, n, nullptr, &tot, stats, 0);
    ctx->emit_pairs = 0;
    ctx->pair_capacity = 0;
    RC(rc);
    if (n == 0) {
        if (total) *total = 0;
        return BFJ_OK;
    }
    unsigned long long found = 0;
    SetScalars sc;
    HC(hipMemcpy(&found, (unsigned long long *)ctx->counters.p + BFJ_CTR_PAIRS, 8, hipMemcpyDeviceToHost));
    HC(hipMemcpy(&sc, ctx->scal.p, sizeof sc, hipMemcpyDeviceToHost));
    const bool all = (sc.all_mask & 1ull) != 0; // eps beyond every distance: the join left them to us
    if (all) found = n > 1 ? (unsigned long long)n * (unsigned long long)(n - 1) / 2 : 0;
    *n_pairs = found;
    if (total) *total = tot;
    if (found > capacity)
        return fail(BFJ_ERR_CAPACITY, "pairs buffer holds %llu pairs, %llu needed", (unsigned long long)capacity, found);
    if (all) bfj_launch_all_pairs((uint2 *)ctx->pairs.p, n, ctx->stream);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(ctx->stream));
    if (found) HC(hipMemcpy(pairs, ctx->pairs.p, found * 8, hipMemcpyDeviceToHost));
    return BFJ_OK;
}

static int batch_hostsets(bfj_ctx *ctx, const bfj_dataset *sets, int32_t n_sets, int device_resident,
                          std::vector<HostSet> &hs, std::vector<size_t> &xoff, std::vector<size_t> &ooff)
{
    if (!sets || n_sets < 1) return fail(BFJ_ERR_INVALID, "no datasets");
    hs.clear();
    xoff.clear();
    ooff.clear();
    size_t xb = 0, ob = 0;
    for (int i = 0; i < n_sets; ++i) {
        const bfj_dataset &s = sets[i];
        if (!s.eps || s.n_eps < 1 || s.n_eps > BFJ_MAXEPS)
            return fail(BFJ_ERR_INVALID, "dataset %d: eps missing or n_eps outside [1, %d]", i, BFJ_MAXEPS);
        HostSet h;
        memset(&h, 0, sizeof h);
        h.x = s.data;
        h.n = s.n;
        h.d = s.d;
        h.ld = s.ld;
        h.row_begin = 0;
        h.row_end = s.n;
        h.n_eps = s.n_eps;
        h.out = device_resident ? s.counts : nullptr;
        for (int e = 0; e < s.n_eps; ++e) h.eps[e] = s.eps[e];
        hs.push_back(h);
        xoff.push_back(xb);
        ooff.push_back(ob);
        if (s.n > 0 && s.ld >= s.d && s.d >= 1) xb += (rows_bytes(s.n, s.d, s.ld) + 15) / 16 * 16;
        if (s.n > 0) ob += (size_t)s.n * s.n_eps * 8;
    }
    (void)ctx;
    return BFJ_OK;
}

int bfj_reference_mode_join(bfj_ctx *ctx, const double *data, int64_t n, int32_t d, int64_t ld, double eps,
                            int32_t search_mode, int32_t compute_prec, int32_t accum_prec, uint64_t *total)
{
    RC(common_checks(ctx, data, n, &eps, 1));
    if (!total) return fail(BFJ_ERR_INVALID, "total is NULL");
    if (d < 1 || ld < d) return fail(BFJ_ERR_INVALID, "need 1 <= d <= ld");
    if (d > BFJ_MAXDIM) return fail(BFJ_ERR_UNSUPPORTED, "d = %d > %d", d, BFJ_MAXDIM);
    if (n > (int64_t)0x7fffffff) return fail(BFJ_ERR_UNSUPPORTED, "n > 2^31-1");
    const bool prec_ok = (compute_prec == 16 || compute_prec == 32 || compute_prec == 64) &&
                         (accum_prec == 32 || accum_prec == 64);
    if (search_mode < 1 || search_mode > 4) return fail(BFJ_ERR_INVALID, "search_mode %d is not 1..4", search_mode);
    if (!prec_ok || (search_mode >= 3 && (compute_prec != 16 || accum_prec != 32)))
        return fail(BFJ_ERR_UNSUPPORTED, "search_mode %d has no reference arithmetic for compute %d / accum %d",
                    search_mode, compute_prec, accum_prec);
    *total = 0;
    if (n == 0) return BFJ_OK;
    ctx->last.reset(); // xcopy is about to change under the last pass's rows
    const int dpad = (d + 15) / 16 * 16;
    RC(ensure(ctx->xcopy, rows_bytes(n, d, ld)));
    HC(hipMemcpyAsync(ctx->xcopy.p, data, rows_bytes(n, d, ld), hipMemcpyHostToDevice, ctx->stream));
    RC(ensure(ctx->ref_x, (size_t)n * dpad * (compute_prec / 8)));
    // Hit counter first, then the FP32 chunk norms of mode 4.
    RC(ensure(ctx->ref_aux, 256 + (size_t)n * (dpad / 16) * sizeof(float)));
    bfj_launch_zero(ctx->ref_aux.p, 8, ctx->stream);
    unsigned long long *cnt = (unsigned long long *)ctx->ref_aux.p;
    float *cn = (float *)((char *)ctx->ref_aux.p + 256);
    const int rc = bfj_compat_run((const double *)ctx->xcopy.p, n, d, ld, eps, search_mode, compute_prec, accum_prec,
                                  ctx->ref_x.p, cn, cnt, ctx->stream);
    if (rc == 1) return fail(BFJ_ERR_UNSUPPORTED, "no reference-arithmetic kernel for this combination");
    if (rc != 0) return fail(BFJ_ERR_DEVICE, "reference-arithmetic launch failed");
    unsigned long long h = 0;
    HC(hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    *total = (uint64_t)h;
    return BFJ_OK;
}

int bfj_self_join_batch(bfj_ctx *ctx, const bfj_dataset *sets, int32_t n_sets, int32_t dtype, int32_t device_resident,
                        bfj_stats *stats)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    HC(hipSetDevice(ctx->device));
    std::vector<HostSet> hs;
    std::vector<size_t> xoff, ooff;
    RC(batch_hostsets(ctx, sets, n_sets, device_resident, hs, xoff, ooff));
    ctx->last.reset();
    if (!device_resident) {
        // Copy every dataset into one device buffer; counts come back through another.
        size_t xb = 0, ob = 0;
        for (int i = 0; i < n_sets; ++i) {
            if (sets[i].n > 0 && sets[i].ld >= sets[i].d && sets[i].d >= 1)
                xb = xoff[i] + rows_bytes(sets[i].n, sets[i].d, sets[i].ld);
            if (sets[i].n > 0) ob = ooff[i] + (size_t)sets[i].n * sets[i].n_eps * 8;
        }
        RC(ensure(ctx->xcopy, std::max<size_t>(xb, 8)));
        RC(ensure(ctx->outbuf, std::max<size_t>(ob, 8)));
        for (int i = 0; i < n_sets; ++i) {
            if (sets[i].n > 0 && !sets[i].data) return fail(BFJ_ERR_INVALID, "dataset %d: data is NULL", i);
            if (sets[i].n > 0 && sets[i].ld >= sets[i].d && sets[i].d >= 1)
                HC(hipMemcpyAsync((char *)ctx->xcopy.p + xoff[i], sets[i].data,
                                  rows_bytes(sets[i].n, sets[i].d, sets[i].ld), hipMemcpyHostToDevice, ctx->stream));
            hs[i].x = (const double *)((char *)ctx->xcopy.p + xoff[i]);
            hs[i].out = sets[i].counts ? (int64_t *)((char *)ctx->outbuf.p + ooff[i]) : nullptr;
        }
    }
    PassPlan p;
    RC(build_plan(ctx, hs, dtype, p, 0));
    RC(stage_plan(ctx, p));
    std::vector<std::vector<uint64_t>> tot;
    RC(run_pass(ctx, p, &tot, stats));
    for (int i = 0; i < n_sets; ++i) {
        if (sets[i].totals)
            for (int e = 0; e < sets[i].n_eps; ++e) sets[i].totals[e] = tot[i][e];
        if (!device_resident && sets[i].counts && sets[i].n > 0)
            HC(hipMemcpy(sets[i].counts, (char *)ctx->outbuf.p + ooff[i], (size_t)sets[i].n * sets[i].n_eps * 8,
                         hipMemcpyDeviceToHost));
    }
    return BFJ_OK;
}

// ---- tile search -------------------------------------------------------------------
// The shape of one K-chunk group of a batch, searched on a seeded sample of the group's
// own datasets (about 3e10 pairs: small passes are ruled by fixed costs, not the join's
// shape) at their own eps.  Each shape is timed as whole passes (device time of prep,
// join, band tests and counting; the fastest of two).  The search starts from the
// heuristic's shape and moves only for a gain above 2 %: R x resident join blocks per
// CU, then the panel size.
static int tune_group(bfj_ctx *ctx, const std::vector<HostSet> &group, int dtype, int kc)
{
    std::vector<size_t> order(group.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    uint64_t seed = 0x5eed0000ull + (uint64_t)kc;
    for (size_t k = order.size(); k > 1; --k) { // Fisher-Yates with splitmix64
        uint64_t z = (seed += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        std::swap(order[k - 1], order[(size_t)(z % k)]);
    }
    std::vector<HostSet> sample;
    double pairs = 0.0;
    int n_eps_max = 1;
    for (size_t k : order) {
        if (!sample.empty() && pairs >= 3e10) break;
        HostSet h = group[k];
        h.out = nullptr; // tuning passes write no counts
        sample.push_back(h);
        pairs += (double)h.n * (double)h.n;
        n_eps_max = std::max(n_eps_max, (int)h.n_eps);
    }
    auto time_shape = [&](int rows, int cap, int panel, double *ms) -> int {
        ctx->tile_rows[dtype][kc] = rows;
        ctx->tile_cap[dtype][kc] = cap;
        ctx->tile_panel[dtype][kc] = panel;
        PassPlan p;
        RC(build_plan(ctx, sample, dtype, p, 0));
        RC(stage_plan(ctx, p));
        *ms = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            bfj_stats st;
            RC(run_pass(ctx, p, nullptr, &st));
            *ms = std::min(*ms, st.ms_total);
        }
        return BFJ_OK;
    };
    // The heuristic's shape first (rows 0: heuristic_rows; all blocks that fit; default panel).
    bool reordered = ctx->reorder != 0;
    for (const HostSet &h : sample) reordered = reordered && h.n >= 512;
    int best_r = heuristic_rows(dtype, kc, n_eps_max, reordered), best_cap = 0, best_panel = BFJ_PANEL_BYTES;
    double best_ms;
    RC(time_shape(best_r, 0, BFJ_PANEL_BYTES, &best_ms));
    for (int r : kRowOptions) {
        if (!bfj_join_has(dtype, kc, r)) continue;
        const int occ = std::max(1, bfj_join_max_blocks_per_cu(dtype, kc, r, lds_for(dtype, kc, r, n_eps_max)));
        std::vector<int> caps = {0}; // as many blocks as fit, then fewer (down to 2)
        for (int c : {4, 3, 2})
            if (c < occ) caps.push_back(c);
        for (int cap : caps) {
            if (r == best_r && cap == 0) continue; // timed above
            double ms;
            RC(time_shape(r, cap, BFJ_PANEL_BYTES, &ms));
            if (ms < 0.98 * best_ms) {
                best_ms = ms;
                best_r = r;
                best_cap = cap;
            }
        }
    }
    for (int panel : {BFJ_PANEL_BYTES / 2, BFJ_PANEL_BYTES * 2}) {
        double ms;
        RC(time_shape(best_r, best_cap, panel, &ms));
        if (ms < 0.98 * best_ms) {
            best_ms = ms;
            best_panel = panel;
        }
    }
    ctx->tile_rows[dtype][kc] = best_r;
    ctx->tile_cap[dtype][kc] = best_cap;
    ctx->tile_panel[dtype][kc] = best_panel;
    ctx->last.reset();
    return BFJ_OK;
}

// ---- plans: one hipGraph per batch ---------------------------------------------
struct bfj_plan {
    bfj_ctx *ctx = nullptr;
    PassPlan pass;
    std::vector<bfj_dataset> sets;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t captured_gen = 0; // buffer generation the graph was captured with
    uint32_t captured_chunks = 0; // capacities baked into the captured kernel arguments
    uint64_t captured_band = 0;
};

static int capture(bfj_plan *pl)
{
    bfj_ctx *ctx = pl->ctx;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    pl->exec = nullptr;
    pl->graph = nullptr;
    HC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_pass(ctx, pl->pass, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != BFJ_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(BFJ_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    pl->graph = g;
    HC(hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0));
    pl->captured_gen = g_buf_gen.load();
    pl->captured_chunks = ctx->chunk_capacity;
    pl->captured_band = ctx->band_capacity;
    return BFJ_OK;
}

int bfj_plan_create(bfj_ctx *ctx, const bfj_dataset *sets, int32_t n_sets, int32_t dtype, bfj_plan **out)
{
    if (!ctx || !out) return fail(BFJ_ERR_INVALID, "ctx or plan out-pointer is NULL");
    *out = nullptr;
    HC(hipSetDevice(ctx->device));
    std::vector<HostSet> hs;
    std::vector<size_t> xoff, ooff;
    RC(batch_hostsets(ctx, sets, n_sets, 1, hs, xoff, ooff));
    // Tile shapes: every K-chunk group without a tuned or forced shape is tuned on a
    // sample of its own datasets, at their own eps (tune_group).
    if (dtype >= BFJ_FP64 && dtype <= BFJ_FP8) {
        std::vector<int> done;
        for (const HostSet &h : hs) {
            if (validate(h, dtype) != BFJ_OK || h.n == 0) continue; // build_plan reports it
            const int kc = pick_kc(dtype, h.d);
            if (ctx->tile_rows[dtype][kc] || std::find(done.begin(), done.end(), kc) != done.end()) continue;
            done.push_back(kc);
            std::vector<HostSet> group;
            for (const HostSet &o : hs)
                if (validate(o, dtype) == BFJ_OK && o.n > 0 && pick_kc(dtype, o.d) == kc) group.push_back(o);
            RC(tune_group(ctx, group, dtype, kc));
        }
    }
    bfj_plan *pl = new bfj_plan();
    pl->ctx = ctx;
    pl->sets.assign(sets, sets + n_sets);
    int rc = build_plan(ctx, hs, dtype, pl->pass, 0);
    if (rc == BFJ_OK) rc = stage_plan(ctx, pl->pass);
    if (rc == BFJ_OK) {
        // One eager pass sizes the refine list, then the pass is captured (unless it
        // runs in slices, which are eager).
        rc = run_pass(ctx, pl->pass, nullptr, nullptr);
    }
    if (rc == BFJ_OK && !pl->pass.sliced) rc = capture(pl);
    if (rc != BFJ_OK) {
        bfj_plan_destroy(pl);
        return rc;
    }
    *out = pl;
    return BFJ_OK;
}

int bfj_plan_run(bfj_plan *pl, bfj_stats *stats)
{
    if (!pl) return fail(BFJ_ERR_INVALID, "plan is NULL");
    bfj_ctx *ctx = pl->ctx;
    HC(hipSetDevice(ctx->device));
    // The context's buffers are shared with eager calls: restage if they moved.
    RC(stage_plan(ctx, pl->pass));
    std::vector<std::vector<uint64_t>> tot;
    auto give_totals = [&]() {
        for (size_t i = 0; i < pl->sets.size(); ++i)
            if (pl->sets[i].totals)
                for (int e = 0; e < pl->sets[i].n_eps; ++e) pl->sets[i].totals[e] = tot[i][(size_t)e];
    };
    // A plan whose candidate stream outgrew the stream limit runs in slices, eagerly.
    if (pl->pass.sliced) {
        RC(run_pass(ctx, pl->pass, &tot, stats));
        give_totals();
        return BFJ_OK;
    }
    if (pl->captured_gen != g_buf_gen.load() || pl->captured_chunks != ctx->chunk_capacity ||
        pl->captured_band != ctx->band_capacity)
        RC(capture(pl));
    for (int attempt = 0; attempt < 5; ++attempt) {
        ctx->last.reset();
        HC(hipGraphLaunch(pl->exec, ctx->stream));
        unsigned long long ctr[BFJ_CTR_ITEMS] = {};
        RC(read_counters(ctx, ctr));
        bool again = false, slice = false;
        RC(grow_on_overflow(ctx, pl->pass, ctr, &again, &slice));
        if (slice) {
            pl->pass.sliced = 1;
            RC(run_sliced(ctx, pl->pass, ctr[BFJ_CTR_CHUNKS], &tot, stats, attempt + 1));
            give_totals();
            return BFJ_OK;
        }
        if (!again) {
            RC(finish_pass(ctx, pl->pass, ctr, attempt, true, &tot, stats));
            give_totals();
            return BFJ_OK;
        }
        RC(stage_plan(ctx, pl->pass));
        RC(capture(pl));
    }
    return fail(BFJ_ERR_INTERNAL, "candidate stream / band list keep overflowing");
}

int bfj_plan_tiles(const bfj_plan *pl, int32_t *kc, int32_t *rows, int32_t *blocks_per_cu, int32_t *panel_bytes,
                   int32_t max_groups, int32_t *n_groups)
{
    if (!pl || !n_groups) return fail(BFJ_ERR_INVALID, "plan or n_groups is NULL");
    const std::vector<Group> &gs = pl->pass.groups;
    *n_groups = (int32_t)gs.size();
    for (int32_t k = 0; k < (int32_t)gs.size() && k < max_groups; ++k) {
        const Group &g = gs[(size_t)k];
        const int occ = bfj_join_max_blocks_per_cu(pl->pass.dtype, g.kc, g.rows, g.lds);
        if (kc) kc[k] = g.kc;
        if (rows) rows[k] = g.rows;
        if (blocks_per_cu) blocks_per_cu[k] = g.cap > 0 ? std::min(g.cap, occ) : occ;
        if (panel_bytes) panel_bytes[k] = g.panel_bytes;
    }
    return BFJ_OK;
}

int bfj_plan_destroy(bfj_plan *pl)
{
    if (!pl) return BFJ_OK;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    delete pl;
    return BFJ_OK;
}

// ---- device buffers -------------------------------------------------------------------
int bfj_buffer_alloc(bfj_ctx *ctx, uint64_t bytes, void **d_ptr)
{
    if (!ctx || !d_ptr) return fail(BFJ_ERR_INVALID, "ctx or out-pointer is NULL");
    *d_ptr = nullptr;
    HC(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *d_ptr = nullptr;
        (void)hipGetLastError();
        return fail(BFJ_ERR_NOMEM, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
    }
    return BFJ_OK;
}

int bfj_buffer_free(bfj_ctx *ctx, void *d_ptr)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    if (!d_ptr) return BFJ_OK;
    HC(hipSetDevice(ctx->device));
    HC(hipStreamSynchronize(ctx->stream));
    HC(hipFree(d_ptr));
    return BFJ_OK;
}

int bfj_buffer_upload(bfj_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes)
{
    if (!ctx || (bytes && (!d_dst || !h_src))) return fail(BFJ_ERR_INVALID, "NULL argument");
    if (!bytes) return BFJ_OK;
    HC(hipSetDevice(ctx->device));
    HC(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    return BFJ_OK;
}

int bfj_buffer_download(bfj_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes)
{
    if (!ctx || (bytes && (!d_src || !h_dst))) return fail(BFJ_ERR_INVALID, "NULL argument");
    if (!bytes) return BFJ_OK;
    HC(hipSetDevice(ctx->device));
    HC(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    return BFJ_OK;
}

// ---- tiles ------------------------------------------------------------------------
int bfj_tile_rows_options(int32_t dtype, int32_t d, int32_t *options, int32_t max_options, int32_t *n_options)
{
    if (!n_options) return fail(BFJ_ERR_INVALID, "n_options is NULL");
    *n_options = 0;
    if (dtype < BFJ_FP64 || dtype > BFJ_FP8) return fail(BFJ_ERR_INVALID, "unknown dtype %d", dtype);
    if (d < 1 || d > BFJ_MAX_DIM) return fail(BFJ_ERR_UNSUPPORTED, "d = %d outside [1, %d]", d, BFJ_MAX_DIM);
    const int kc = pick_kc(dtype, d);
    int k = 0;
    for (int r : kRowOptions)
        if (kc && bfj_join_has(dtype, kc, r)) {
            if (options && k < max_options) options[k] = r;
            ++k;
        }
    *n_options = k;
    return BFJ_OK;
}

int bfj_set_reorder(bfj_ctx *ctx, int32_t enable)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    ctx->reorder = enable ? 1 : 0;
    return BFJ_OK;
}

int bfj_set_tile_rows(bfj_ctx *ctx, int32_t dtype, int32_t tile_rows)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    if (dtype < BFJ_FP64 || dtype > BFJ_FP8) return fail(BFJ_ERR_INVALID, "unknown dtype %d", dtype);
    for (int kc = 0; kc <= 32; ++kc) {
        ctx->tile_rows[dtype][kc] = tile_rows;
        ctx->tile_cap[dtype][kc] = 0;
        ctx->tile_panel[dtype][kc] = 0;
    }
    return BFJ_OK;
}

int bfj_autotune(bfj_ctx *ctx, int32_t dtype, int32_t d, int64_t n, int32_t *tile_rows_out)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    if (dtype < BFJ_FP64 || dtype > BFJ_FP8) return fail(BFJ_ERR_INVALID, "unknown dtype %d", dtype);
    if (d < 1 || d > BFJ_MAX_DIM) return fail(BFJ_ERR_UNSUPPORTED, "d = %d outside [1, %d]", d, BFJ_MAX_DIM);
    if (n < 256) n = 256;
    HC(hipSetDevice(ctx->device));
    const int kc = pick_kc(dtype, d);
    if (!kc) return fail(BFJ_ERR_UNSUPPORTED, "no kernel for dtype %d d %d", dtype, d);
    // Synthetic uniform rows; a small eps keeps the epilogue on its fast path, as in
    // the joins the tiles are tuned for.
    double *x = nullptr;
    HC(hipMalloc(&x, (size_t)n * d * 8));
    hipLaunchKernelGGL(k_synth, dim3(1024), dim3(256), 0, ctx->stream, x, (int64_t)n * d, 12345ull);
    const double eps = 1e-3 * std::sqrt((double)d);
    int best = 0;
    double best_ms = 1e300;
    for (int r : kRowOptions) {
        if (!bfj_join_has(dtype, kc, r)) continue;
        double ms_min = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            bfj_stats st;
            int rc = join_one(ctx, x, n, d, d, &eps, 1, dtype, 0, n, nullptr, nullptr, &st, r);
            if (rc != BFJ_OK) {
                (void)hipFree(x);
                return rc;
            }
            ms_min = std::min(ms_min, st.ms_join);
        }
        if (ms_min < best_ms) {
            best_ms = ms_min;
            best = r;
        }
    }
    (void)hipFree(x);
    ctx->last.reset(); // its rows are gone
    ctx->tile_rows[dtype][kc] = best;
    if (tile_rows_out) *tile_rows_out = best;
    return BFJ_OK;
}

int bfj_fail_allocation(int64_t countdown)
{
    if (countdown < 0) return fail(BFJ_ERR_INVALID, "countdown < 0");
    g_fail_countdown.store(countdown);
    return BFJ_OK;
}

int bfj_set_stream_limit(bfj_ctx *ctx, uint64_t bytes)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    ctx->stream_limit = bytes;
    ctx->mem_ceiling = 0;
    // Stream buffers grown past the new limit are given back; the next pass sizes them
    // within it.
    if ((double)ctx->cand.cap + (double)ctx->band.cap + (double)ctx->mid_band.cap > (double)stream_limit(ctx)) {
        HC(hipSetDevice(ctx->device));
        HC(hipStreamSynchronize(ctx->stream));
        ctx->last.reset();
        for (DevBuf *b : {&ctx->cand, &ctx->chunk_fill, &ctx->band, &ctx->mid_band}) release(*b);
        g_buf_gen.fetch_add(1);
    }
    return BFJ_OK;
}

int bfj_set_capacity(bfj_ctx *ctx, uint64_t candidate_chunks, uint64_t band_records)
{
    if (!ctx) return fail(BFJ_ERR_INVALID, "ctx is NULL");
    if (candidate_chunks > BFJ_MAX_CHUNKS) return fail(BFJ_ERR_INVALID, "candidate_chunks > %llu", BFJ_MAX_CHUNKS);
    if (band_records > BFJ_MAX_BAND) return fail(BFJ_ERR_INVALID, "band_records > %llu", BFJ_MAX_BAND);
    // The buffers keep their size; the kernels only use the first `capacity` entries.
    if (candidate_chunks) ctx->chunk_capacity = (uint32_t)candidate_chunks;
    if (band_records) ctx->band_capacity = band_records;
    return BFJ_OK;
}

// Seeded 16 x 16 tiles over the datasets of a pass, about `want` of them, spread by
// each dataset's share of the pairs: symmetric datasets over blocks rb <= cb,
// row-range passes over the query blocks against all blocks.
static std::vector<ErrTile> sample_tiles(const PassPlan &p, int64_t want, uint64_t seed)
{
    std::vector<ErrTile> out;
    if (want <= 0) return out;
    auto next = [&seed]() {
        uint64_t z = (seed += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    double total = 0.0;
    for (const SetDesc &s : p.sets) total += (double)(s.row_end - s.row_begin) * (double)s.n;
    if (total <= 0.0) return out;
    for (size_t si = 0; si < p.sets.size(); ++si) {
        const SetDesc &s = p.sets[si];
        const double w = (double)(s.row_end - s.row_begin) * (double)s.n;
        if (w <= 0.0) continue;
        const int64_t k = (int64_t)std::ceil((double)want * w / total);
        const int64_t nb = (s.n + 15) / 16;
        const int64_t qb0 = s.row_begin / 16, qb1 = (s.row_end + 15) / 16;
        for (int64_t t = 0; t < k; ++t) {
            ErrTile e;
            e.set = (int32_t)si;
            e.pad = 0;
            if (s.sym) {
                int64_t a = (int64_t)(next() % (uint64_t)nb), b = (int64_t)(next() % (uint64_t)nb);
                e.rb = (int32_t)std::min(a, b);
                e.cb = (int32_t)std::max(a, b);
            } else {
                e.rb = (int32_t)(qb0 + (int64_t)(next() % (uint64_t)(qb1 - qb0)));
                e.cb = (int32_t)(next() % (uint64_t)nb);
            }
            out.push_back(e);
        }
    }
    return out;
}

int bfj_pass_errors(bfj_ctx *ctx, uint64_t max_candidates, int64_t sample_tiles_n, uint64_t seed, bfj_errors *out)
{
    if (!ctx || !out) return fail(BFJ_ERR_INVALID, "ctx or out is NULL");
    memset(out, 0, sizeof *out);
    if (sample_tiles_n < 0) return fail(BFJ_ERR_INVALID, "sample_tiles < 0");
    if (!ctx->last) return fail(BFJ_ERR_INVALID, "no finished pass on this context to measure");
    HC(hipSetDevice(ctx->device));
    const PassPlan &p = *ctx->last;
    const PassArgs a = pass_args(ctx, p);
    unsigned long long ctr[BFJ_CTR_ITEMS] = {};
    RC(read_counters(ctx, ctr));
    const uint32_t n_chunks =
        (uint32_t)std::min<unsigned long long>(ctr[BFJ_CTR_CHUNKS], (unsigned long long)ctx->chunk_capacity);
    const unsigned long long total = ctr[BFJ_CTR_CANDS];
    uint32_t stride = 1;
    if (max_candidates && total > max_candidates)
        stride = (uint32_t)std::min<unsigned long long>((total + max_candidates - 1) / max_candidates, 0xffffffffull);
    const std::vector<ErrTile> tiles = sample_tiles(p, sample_tiles_n, seed);
    if (tiles.size() > (size_t)0x7fffffff) return fail(BFJ_ERR_INVALID, "too many sample tiles");
    const size_t tile_off = 256; // ErrOut first
    RC(ensure(ctx->errbuf, tile_off + std::max<size_t>(tiles.size(), 1) * sizeof(ErrTile)));
    ErrOut *d_out = (ErrOut *)ctx->errbuf.p;
    ErrTile *d_tiles = (ErrTile *)((char *)ctx->errbuf.p + tile_off);
    bfj_launch_zero(d_out, sizeof(ErrOut), ctx->stream);
    if (!tiles.empty())
        HC(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(ErrTile), hipMemcpyHostToDevice, ctx->stream));
    bfj_launch_errcand(a, d_out, n_chunks, stride, ctx->stream);
    bfj_launch_errtile(a, d_tiles, (uint32_t)tiles.size(), d_out, ctx->stream);
    HC(hipGetLastError());
    ErrOut h;
    HC(hipMemcpyAsync(&h, d_out, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    out->cand_total = total;
    out->cand_pairs = h.cand_pairs;
    out->tile_pairs = h.tile_pairs;
    out->violations = h.violations;
    out->cand_max_abs_err = h.cand_pairs ? bits_to_double(h.cand_max_bits) : 0.0;
    out->tile_max_abs_err = h.tile_pairs ? bits_to_double(h.tile_max_bits) : 0.0;
    const bool low = p.dtype != BFJ_FP64;
    out->max_acc_ratio = low && h.tile_pairs ? bits_to_double(h.acc_ratio_bits) : -1.0;
    out->max_quant_ratio = low && h.tile_pairs ? bits_to_double(h.quant_ratio_bits) : -1.0;
    return BFJ_OK;
}

} // extern "C"
