// bfj_internal.h -- device-side data structures shared by the prep, join and
// refine kernels and the host orchestration (bfj_api.hip).
//
// Data layout in HBM for one pass (all datasets of a batch concatenated):
//   x          caller's FP64 rows, n x ld, row-major (never copied on the device path)
//   pack       dot-product operands in MFMA fragment order: for every block of 16
//              rows and every K-chunk, 64 lanes x FRAG_BYTES contiguous, so one
//              wavefront load of a fragment is a single coalesced 1-2 KiB read
//   nq[row]    0.5 * ||xhat_j||^2 (FP32 or FP64), +BIG for padding rows
//   gq[row]    g_i = 0.5 * (H* + E_i - N_i) rounded up (H*: largest banded H of the
//              dataset); -BIG for rows that are not queries
//   he[row]    (2 g_i + N_i, E_i): S_hat = he.x - 2 Z estimates S with |S_hat - S| <= E_i
//   cnt[k][row] per-row pair counts by bucket: pairs whose smallest eps (in ascending
//              order) that holds them is k; eps e counts buckets 0..e (uint32)
//   band list  (i, j, z, mask) records of pairs whose band straddles a threshold
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BFJ_MAXEPS 8
#define BFJ_WAVE 64
#define BFJ_BLOCK 256
#ifndef BFJ_JWAVES
#define BFJ_JWAVES 4 // wavefronts per join block (they share the candidate ring)
#endif
#define BFJ_JBLOCK (BFJ_JWAVES * BFJ_WAVE)

typedef _Float16 bfj_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bfj_b8 __attribute__((ext_vector_type(8)));
typedef float bfj_f4 __attribute__((ext_vector_type(4)));
typedef double bfj_d2 __attribute__((ext_vector_type(2)));
typedef double bfj_d4 __attribute__((ext_vector_type(4)));
typedef int bfj_i8 __attribute__((ext_vector_type(8)));
typedef int bfj_i4 __attribute__((ext_vector_type(4)));

enum { BFJ_FP64 = 0, BFJ_FP16 = 1, BFJ_BF16 = 2, BFJ_FP8 = 3 };

// One dataset of a pass.  Row-indexed arrays are addressed with row_off + i,
// where i is the dataset-local row; rows are padded to a multiple of 16.
struct SetDesc {
    const double *x;   // FP64 input rows (device)
    int64_t n;         // points
    int64_t npad;      // n rounded up to 64 (whole 4-block ring slots)
    int64_t ld;        // row stride of x in doubles
    int64_t row_begin; // query rows [row_begin, row_end)
    int64_t row_end;
    int64_t row_off;   // offset in the row-indexed arrays
    int64_t pack_off;  // byte offset of this dataset's fragments in pack
    int64_t cnt_off;   // offset into cnt; cnt[cnt_off + e * npad + i]
    int64_t *out;      // device int64 counts [n_eps][row_end-row_begin], self excluded (may be NULL)
    int32_t d;         // true dimension
    int32_t kc;        // K-chunks after padding
    int32_t n_eps;
    int32_t set_id;
    int32_t sym;       // every row is a query: each unordered pair is evaluated once
    int32_t al16;      // x and ld keep every row 16-byte aligned
    int32_t reorder;   // rows packed in a locality order (perm); 0: position = row
    double T[BFJ_MAXEPS]; // input-space thresholds, ascending: hit iff d2 <= T
    int8_t eperm[BFJ_MAXEPS]; // T[k] belongs to the caller's eps[eperm[k]]
};

// Per-dataset device scalars produced by the prep kernels.
struct SetScalars {
    double sigma;          // power-of-two scale: kernel space = (x - mu) / sigma
    unsigned long long maxabs_bits; // max |x - mu| (non-negative double as bits)
    unsigned long long qmax_bits;   // max quantisation error radius (kernel space)
    unsigned long long nmax_bits;   // max ||xhat||^2 (kernel space)
    unsigned long long halfband_bits; // max (H - L) / 2 (input space)
    unsigned long long nonfinite;     // non-zero if some row holds NaN or inf
    unsigned long long all_mask;      // eps whose T exceeds every possible distance: all pairs hit
    unsigned long long band_mask;     // eps decided by the join + refine (neither none nor all)
    double slo[BFJ_MAXEPS];           // sqrt(T/(1+rho)) rounded down (kernel space)
    double shi[BFJ_MAXEPS];           // sqrt(T/(1-rho)) rounded up
    unsigned long long total[BFJ_MAXEPS]; // totals, self included
};

// A pair whose certain-in / certain-out tests were both inconclusive.
struct BandEntry {
    uint32_t i;     // dataset-local row of the tile (the one Z was computed for)
    uint32_t j;     // dataset-local candidate row
    float z;        // Z value of the pair from the join kernel (kernel space)
    uint32_t meta;  // bits 0..7: eps in the band; bit 15: counts for row j too; 16..31: set index
};
// A band-list entry with no eps in the band (meta bits 0..7 clear) is a hole: the
// unused tail of a block that k_classify reserved.  Consumers skip holes.
__host__ __device__ inline bool bfj_band_live(const BandEntry &e) { return (e.meta & 0xffu) != 0u; }

// A join work item: one strip of 4R query rows of one dataset against one panel of
// candidate blocks [cb0, cb1).
struct WorkItem {
    int32_t set;
    int32_t rb0; // first 16-row block of the strip
    int32_t cb0; // candidate blocks [cb0, cb1)
    int32_t cb1;
};

// A pair the join could not reject (Z >= 0 for the widest banded eps).  The join
// streams these into per-wavefront chunks; k_classify applies the per-eps tests.
struct Candidate {
    uint32_t row; // global row (row_off + dataset-local row)
    uint32_t col; // bits 0..30: global candidate row (row_off + dataset-local row); bit 31: the pair counts for it too
    double z;     // Z of the pair (kernel space)
};
#define BFJ_COL_SYM 0x80000000u

// Candidate records per chunk: one atomic on the chunk counter per 4096 records keeps
// the join's grabs of fresh chunks off the critical path (the public header's
// BFJ_CHUNK_RECORDS, asserted equal where both are included).
#define BFJ_CHUNK 4096
#ifndef BFJ_PANEL_BYTES
#define BFJ_PANEL_BYTES (1 << 20) // candidate fragments per join work item
#endif

struct JoinArgs {
    const SetDesc *sets;
    const WorkItem *items;
    int32_t n_items;
    uint32_t *item_counter; // persistent-kernel queue head (zeroed before launch)
    const unsigned char *pack;
    const void *nq;          // float or double per row
    const void *gq;          // float or double per row
    const uint8_t *scales;   // FP8 per-row E8M0 scale bytes (NULL otherwise)
    Candidate *cand;         // candidate stream
    uint32_t *chunk_fill;    // records written into each chunk
    unsigned long long *chunk_count; // chunks handed out (may exceed chunk_capacity: overflow)
    uint32_t chunk_capacity;
};

// Counters of a pass: one buffer zeroed at the start of every pass, 64-bit slots.
// Every count that can grow with the number of pairs is 64-bit, so none can wrap
// and hide an overflow from the host (which compares them with the capacities).
#define BFJ_CTR_BAND 0   // band-list slots reserved by k_classify (records and holes)
#define BFJ_CTR_CHUNKS 1 // candidate chunks handed out by the join
#define BFJ_CTR_PAIRS 2  // pairs found in pair-list mode
#define BFJ_CTR_MID 3    // records the middle stage leaves for the FP64 refine
#define BFJ_CTR_CANDS 4  // candidate records the join wrote (statistics)
#define BFJ_CTR_BANDREC 5 // band records k_classify wrote (statistics)
#define BFJ_CLASSIFY_WAVES (2048 * 4) // k_classify's grid: 2048 blocks of 4 wavefronts
#define BFJ_CTR_ITEMS 8  // from this slot on: 32-bit work-queue heads, one per join launch
#define BFJ_CTR_SLOTS 32 // 256 bytes

// Host-side launchers (one translation unit per dtype; see join_*.hip).
struct JoinLaunch {
    int dtype;
    int kc;
    int rows;           // R: query rows per wavefront
    int grid_blocks;    // 0: one wavefront per item; >0: persistent grid
    int lds_bytes;      // dynamic LDS for per-row counts
};

int bfj_join_launch(const JoinLaunch &L, const JoinArgs &a, hipStream_t s);
bool bfj_join_has(int dtype, int kc, int rows);
int bfj_join_max_blocks_per_cu(int dtype, int kc, int rows, int lds_bytes);

// Fragment geometry per dtype: K values per chunk and bytes per lane per chunk.
__host__ __device__ inline int bfj_kchunk(int dtype) { return dtype == BFJ_FP64 ? 8 : (dtype == BFJ_FP8 ? 128 : 32); }
__host__ __device__ inline int bfj_frag_bytes(int dtype) { return dtype == BFJ_FP8 ? 32 : 16; }
__host__ __device__ inline int bfj_acc_bytes(int dtype) { return dtype == BFJ_FP64 ? 8 : 4; }

// Prep / refine / finalize launchers (prep.hip, refine.hip).
struct PassArgs {
    const SetDesc *sets;       // device
    SetScalars *scal;          // device, one per set
    unsigned long long *blk_tot; // [row block][BFJ_MAXEPS] totals of each block (k_finalize)
    int32_t n_sets;
    int32_t dtype;
    const int32_t *blk_set;    // block -> set table for row-parallel kernels
    const int64_t *blk_row0;   // block -> first dataset-local row
    int32_t n_row_blocks;      // blocks in those tables (64 rows each)
    const int32_t *set_blk0;   // set -> its first block in the tables
    const int32_t *set_nblk;   // set -> number of its blocks
    double *colsum;            // [3][row_blocks][BFJ_MAXDIM] per-block column sums, maxima, minima
    double *mu;                // [n_sets][BFJ_MAXDIM]
    unsigned char *pack;
    void *nq, *gq;
    double2 *he;               // per row
    uint16_t *row_set;         // per row: index of its dataset
    double *qerr, *nrm;        // per row
    double *dnq;               // per row: n_j minus the join's candidate threshold for it (FP8; else 0)
    uint8_t *scales;
    uint32_t *cnt;
    int64_t cnt_len;
    BandEntry *band;
    unsigned long long *band_count;    // slots reserved: k_classify takes blocks of band_block
    uint64_t band_capacity;
    uint32_t band_block;               // slots per k_classify reservation (a power of two, 64..1024)
    unsigned long long *band_records;  // band records written (statistics)
    // Low-precision passes: a middle stage between classify and the FP64 refine
    // (refine.hip) on higher-precision copies of the rows: FP16 for FP8 passes, FP32
    // for FP16/BF16 passes.
    int32_t mid;                       // 1: run it
    int32_t mid_fp32;                  // the copy is FP32 (else FP16)
    unsigned char *mid_rows;           // [row][BFJ_MAXDIM] copy of the kernel-space rows, by position
    double *mid_q;                     // [row] radius of that copy
    BandEntry *mid_band;               // records it leaves for the FP64 refine (capacity band_capacity)
    unsigned long long *mid_count;
    const Candidate *cand;
    const uint32_t *chunk_fill;
    const unsigned long long *chunk_count;
    uint32_t chunk_capacity;
    unsigned long long *cand_records;  // k_classify adds the records it read (statistics)
    int32_t emit_pairs;                // pair-list mode: one dataset, eps 0, every row a query
    uint2 *pairs;                      // (i, j), i < j
    unsigned long long *pair_count;    // pairs found (may exceed pair_capacity)
    unsigned long long pair_capacity;
    // Row order (order.hip): position -> dataset-local row, per row_off + position.
    uint32_t *perm;
    uint8_t *label;                    // nearest pivot per row (input order)
    uint32_t *ohist;                   // [prep block][BFJ_PIVOTS] counts, then offsets
    uint32_t *oseg;                    // [set][oseg_stride][BFJ_PIVOTS] segment offsets
    int32_t oseg_stride;               // scan segments of the largest dataset
    _Float16 *pivots;                  // [set][BFJ_PIVOTS][BFJ_MAXDIM]
    float *pivot_norm;                 // [set][BFJ_PIVOTS]
};

#define BFJ_PIVOTS 128   // pivot rows of the locality order
#define BFJ_ORD_SEG 64   // prep blocks per scan segment of the locality order

// Dataset-local row held at position i (the identity unless the set is reordered).
__device__ inline uint32_t bfj_row_at(const PassArgs &a, const SetDesc &s, int64_t i)
{
    return a.perm[s.row_off + i];
}

#define BFJ_MAXDIM 256
#define BFJ_PREP_ROWS 64   // rows per block in row-parallel prep kernels

// OCP FP8 e4m3 byte -> its value (bias 7, subnormals m * 2^-9; the encoder in prep.hip
// never produces the NaN code 0x7f / 0xff).
__device__ inline double e4m3_decode(uint8_t b)
{
    const double sg = (b & 0x80) ? -1.0 : 1.0;
    const int ef = (b >> 3) & 15, m = b & 7;
    if (ef == 0) return sg * ldexp((double)m, -9);
    return sg * ldexp((double)(8 + m), ef - 10);
}

void bfj_launch_zero(void *p, size_t bytes, hipStream_t s);
void bfj_launch_prep(const PassArgs &a, hipStream_t s);
void bfj_launch_classify(const PassArgs &a, hipStream_t s);
void bfj_launch_refine(const PassArgs &a, hipStream_t s);
void bfj_launch_mid(const PassArgs &a, hipStream_t s);
void bfj_launch_finalize(const PassArgs &a, hipStream_t s);
void bfj_launch_all_pairs(uint2 *pairs, int64_t n, hipStream_t s);
void bfj_launch_order(const PassArgs &a, hipStream_t s);

// Accuracy statistics of a finished pass (errstat.hip).
struct ErrTile {
    int32_t set, rb, cb, pad; // dataset, row block and candidate block of one sampled 16 x 16 tile
};
struct ErrOut { // zeroed before the statistics kernels; maxima as bits of non-negative doubles
    unsigned long long cand_pairs, tile_pairs, violations;
    unsigned long long cand_max_bits, tile_max_bits, acc_ratio_bits, quant_ratio_bits;
};
void bfj_launch_errcand(const PassArgs &a, ErrOut *out, uint32_t n_chunks, uint32_t stride, hipStream_t s);
void bfj_launch_errtile(const PassArgs &a, const ErrTile *tiles, uint32_t n_tiles, ErrOut *out, hipStream_t s);
