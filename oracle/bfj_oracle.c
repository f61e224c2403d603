/*
 * bfj_oracle.c -- CPU FP64 oracle for the brute-force Euclidean epsilon self-join.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it.  The product path
 * (tensor_bf_main_amd/) never links, loads or calls it.
 *
 * What it restates from the reference (jeevana-swaroop/tensor_bf_main):
 *   - distance arithmetic: kernel_join.cu:184-189 (distanceCalculationBruteForceCuda,
 *     searchMode 1).  For a query q and candidate c the squared distance is
 *     accumulated over k = 0..d-1 in ascending order as
 *         t = q[k] - c[k];  acc = fma(t, t, acc);
 *     which is what nvcc's default FMA contraction makes of
 *     `acc += (q[k]-c[k])*(q[k]-c[k])` with double COMPUTE/ACCUM types.
 *   - predicate: kernel_join.cu:194, `sqrt(acc) <= epsilon`.  It is evaluated
 *     here as `acc <= T` with T = bfo_threshold(eps), the largest double whose
 *     correctly rounded square root is <= eps; sqrt is monotone, so the two
 *     predicates agree for every acc >= 0.
 *   - count semantics: kernel_join.cu:184 loops candidates over all
 *     nbQueryPoints including the query itself, and gpu_join.cu:318-321 returns
 *     the total number of ordered hits (self included).  Here the per-point
 *     counts exclude self and the totals include it (total = sum(counts) + n
 *     whenever eps >= 0).
 *
 * Build: see oracle/Makefile (-O3 -ffp-contract=off -mfma -fopenmp).  The
 * explicit fma() calls are correctly rounded whether or not they vectorise, so
 * every lane of a vector loop gives the bits the scalar loop gives.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BFO_CB 8 /* candidates per transposed block */

/* Largest double T with sqrt(T) <= eps (reference predicate, kernel_join.cu:194).
 * eps < 0 gives -1 (nothing matches, not even d2 == 0); eps = +inf gives +inf. */
double bfo_threshold(double eps)
{
    if (isnan(eps)) return NAN;
    if (eps < 0.0) return -1.0;
    if (isinf(eps)) return INFINITY;
    double t = eps * eps;
    if (isinf(t)) t = DBL_MAX;
    while (t > 0.0 && sqrt(t) > eps) t = nextafter(t, -INFINITY);
    for (;;) {
        double u = nextafter(t, INFINITY);
        if (isinf(u) || sqrt(u) > eps) break;
        t = u;
    }
    return t;
}

/* Direct-form squared distance, kernel_join.cu:185-189 order. */
double bfo_pair_d2(const double *q, const double *c, int32_t d)
{
    double acc = 0.0;
    for (int32_t k = 0; k < d; ++k) {
        double t = q[k] - c[k];
        acc = fma(t, t, acc);
    }
    return acc;
}

int bfo_pair_d2_batch(const double *data, int64_t n, int32_t d, int64_t ld,
                      const int64_t *pi, const int64_t *pj, int64_t m, double *out)
{
    if (!data || !pi || !pj || !out || n < 0 || d < 1 || ld < d || m < 0) return 1;
    for (int64_t t = 0; t < m; ++t) {
        if (pi[t] < 0 || pi[t] >= n || pj[t] < 0 || pj[t] >= n) return 2;
        out[t] = bfo_pair_d2(data + pi[t] * ld, data + pj[t] * ld, d);
    }
    return 0;
}

/* Candidates transposed into blocks of BFO_CB: blk[b][k][jj] = data[(b*CB+jj)*ld + k].
 * Missing tail candidates are zero rows; their results are discarded. */
static double *transpose_blocks(const double *data, int64_t n, int32_t d, int64_t ld, int64_t *nblk_out)
{
    int64_t nblk = (n + BFO_CB - 1) / BFO_CB;
    double *blk = (double *)calloc((size_t)(nblk > 0 ? nblk : 1) * (size_t)d * BFO_CB, sizeof(double));
    if (!blk) return NULL;
    for (int64_t b = 0; b < nblk; ++b)
        for (int jj = 0; jj < BFO_CB; ++jj) {
            int64_t j = b * BFO_CB + jj;
            if (j >= n) break;
            for (int32_t k = 0; k < d; ++k) blk[(b * d + k) * BFO_CB + jj] = data[j * ld + k];
        }
    *nblk_out = nblk;
    return blk;
}

/* One query row against every candidate: counts_row[e] += #{j != i : d2(i,j) <= T[e]}. */
static void query_row(const double *q, int64_t i, const double *blk, int64_t nblk, int64_t n, int32_t d,
                      const double *T, int32_t n_eps, int64_t *counts_row /* [n_eps] */)
{
    for (int64_t b = 0; b < nblk; ++b) {
        const double *cb = blk + b * (int64_t)d * BFO_CB;
        double acc[BFO_CB];
        for (int jj = 0; jj < BFO_CB; ++jj) acc[jj] = 0.0;
        for (int32_t k = 0; k < d; ++k) {
            const double qk = q[k];
            const double *ck = cb + (int64_t)k * BFO_CB;
            for (int jj = 0; jj < BFO_CB; ++jj) {
                double t = qk - ck[jj];
                acc[jj] = fma(t, t, acc[jj]);
            }
        }
        for (int jj = 0; jj < BFO_CB; ++jj) {
            int64_t j = b * BFO_CB + jj;
            if (j >= n || j == i) continue;
            for (int32_t e = 0; e < n_eps; ++e) counts_row[e] += (acc[jj] <= T[e]);
        }
    }
}

static int check_args(const double *data, int64_t n, int32_t d, int64_t ld, const double *eps, int32_t n_eps)
{
    if (n < 0 || d < 1 || ld < d || n_eps < 1 || !eps) return 1;
    if (n > 0 && !data) return 1;
    for (int32_t e = 0; e < n_eps; ++e)
        if (isnan(eps[e])) return 1;
    return 0;
}

/*
 * Rows `rows[0..n_rows)` against all n candidates.
 * counts[e * n_rows + r] = neighbours of rows[r] within eps[e], self excluded.
 */
int bfo_join_rows(const double *data, int64_t n, int32_t d, int64_t ld,
                  const int64_t *rows, int64_t n_rows, const double *eps, int32_t n_eps,
                  int64_t *counts, int32_t nthreads)
{
    if (check_args(data, n, d, ld, eps, n_eps) || n_rows < 0 || (n_rows > 0 && (!rows || !counts))) return 1;
    for (int64_t r = 0; r < n_rows; ++r)
        if (rows[r] < 0 || rows[r] >= n) return 2;
    double T[64];
    if (n_eps > 64) return 1;
    for (int32_t e = 0; e < n_eps; ++e) T[e] = bfo_threshold(eps[e]);
    int64_t nblk = 0;
    double *blk = transpose_blocks(data, n, d, ld, &nblk);
    if (!blk) return 3;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_rows; ++r) {
        int64_t c[64];
        for (int32_t e = 0; e < n_eps; ++e) c[e] = 0;
        query_row(data + rows[r] * ld, rows[r], blk, nblk, n, d, T, n_eps, c);
        for (int32_t e = 0; e < n_eps; ++e) counts[(int64_t)e * n_rows + r] = c[e];
    }
    free(blk);
    return 0;
}

/*
 * Full self-join over query rows [row_begin, row_end) against all n candidates.
 * counts[e * (row_end - row_begin) + (i - row_begin)]: self excluded.
 * totals[e]: reference semantics (gpu_join.cu:320-321) over the row range,
 *            ordered pairs with self included.
 */
int bfo_join(const double *data, int64_t n, int32_t d, int64_t ld, const double *eps, int32_t n_eps,
             int64_t row_begin, int64_t row_end, int64_t *counts, uint64_t *totals, int32_t nthreads)
{
    if (check_args(data, n, d, ld, eps, n_eps) || !totals) return 1;
    if (row_begin < 0 || row_end > n || row_begin > row_end) return 2;
    int64_t m = row_end - row_begin;
    if (m > 0 && !counts) return 1;
    int64_t *rows = (int64_t *)malloc((size_t)(m > 0 ? m : 1) * sizeof(int64_t));
    if (!rows) return 3;
    for (int64_t r = 0; r < m; ++r) rows[r] = row_begin + r;
    int rc = bfo_join_rows(data, n, d, ld, rows, m, eps, n_eps, counts, nthreads);
    free(rows);
    if (rc) return rc;
    for (int32_t e = 0; e < n_eps; ++e) {
        uint64_t t = 0;
        for (int64_t r = 0; r < m; ++r) t += (uint64_t)counts[(int64_t)e * m + r];
        if (bfo_threshold(eps[e]) >= 0.0) t += (uint64_t)m; /* self pairs: d2 == 0 */
        totals[e] = t;
    }
    return 0;
}

int bfo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------
 * The reference's own arithmetic for its search modes 1-4, as a checker for the
 * library's bfj_reference_mode_join.  Values of every type are held in doubles
 * (binary16 and binary32 values are exact there) and every operation is rounded
 * once to the type it runs in:
 *   mode 1  kernel_join.cu:168-200   t = C(q_k - c_k); acc = A(acc + A(C(t * t)))
 *   mode 2  kernel_join.cu:203-239   acc = A(acc + A(A(A(C(C(-2 q_k) c_k)) + A(q_k)A(q_k)) + A(C(c_k c_k))))
 *   mode 3  kernel_join.cu:245-325   per 16-column chunk: D_k = FP16(q_k - c_k),
 *                                    acc = acc + (sum of D_k^2 in column order, FP32)
 *   mode 4  kernel_join.cu:328-524   per chunk: t = cn_c + (sum of FP16(-2 q_k) c_k, FP32),
 *                                    acc = (acc + t) + cn_q; cn = FP32 chunk sums of squares
 *                                    (kernel_join.cu:78-86)
 * When the compute and accumulation types are the same, nvcc's default contraction
 * (-fmad=true) fuses a product into the sum that consumes it: mode 1 becomes
 * acc = fma(t, t, acc) (the FP64 oracle above for 64/64) and mode 2
 * acc = acc + fma(c, c, fma(C(-2 q), c, q2)).  Modes 3 and 4 ran on tensor cores whose
 * summation order is undocumented; the column-order FP32 sum is an assumption, so those
 * two restate the library's definition rather than the reference's hardware.
 * The hit test sqrt(acc) <= A(eps) (modes 2 and 4 on |acc|) is acc <= T_A. */

/* Round to binary16, nearest even, from a double in one step (x86 has no direct
 * conversion; going through float first could round twice). */
static double bfo_h16(double x)
{
    if (isnan(x)) return x;
    const double ax = fabs(x);
    if (ax >= 65520.0) return copysign(INFINITY, x); /* 65504 + half an ulp rounds up to inf */
    double ulp;
    if (ax < 0x1p-14) ulp = 0x1p-24; /* subnormal binade */
    else {
        int e;
        (void)frexp(ax, &e); /* ax in [2^(e-1), 2^e) */
        ulp = ldexp(1.0, e - 1 - 10);
    }
    return nearbyint(x / ulp) * ulp; /* exact scaling; nearbyint rounds half to even */
}

/* Exported for the tests that pin the rounding against numpy's float64 -> float16. */
double bfo_round_half(double x) { return bfo_h16(x); }

static double bfo_to_c(double x, int cp) { return cp == 16 ? bfo_h16(x) : cp == 32 ? (double)(float)x : x; }
static double bfo_to_a(double x, int ap) { return ap == 32 ? (double)(float)x : x; }
/* a - b and a * b in the compute type (operands already of that type). */
static double bfo_csub(double a, double b, int cp)
{
    if (cp == 16) return bfo_h16(a - b); /* exact in double, then one rounding */
    if (cp == 32) return (double)((float)a - (float)b);
    return a - b;
}
static double bfo_cmul(double a, double b, int cp)
{
    if (cp == 16) return bfo_h16(a * b);
    if (cp == 32) return (double)((float)a * (float)b);
    return a * b;
}
static double bfo_aadd(double a, double b, int ap) { return ap == 32 ? (double)((float)a + (float)b) : a + b; }
static double bfo_amul(double a, double b, int ap) { return ap == 32 ? (double)((float)a * (float)b) : a * b; }
static double bfo_afma(double a, double b, double c, int ap)
{
    return ap == 32 ? (double)fmaf((float)a, (float)b, (float)c) : fma(a, b, c);
}

static double bfo_threshold_f(double eps)
{
    const float e = (float)eps;
    if (isnan(e)) return NAN;
    if (e < 0.0f) return -1.0;
    if (isinf(e)) return INFINITY;
    float t = e * e;
    if (isinf(t)) t = FLT_MAX;
    while (t > 0.0f && sqrtf(t) > e) t = nextafterf(t, -INFINITY);
    for (;;) {
        float u = nextafterf(t, INFINITY);
        if (isinf(u) || sqrtf(u) > e) break;
        t = u;
    }
    return (double)t;
}

int bfo_reference_total(const double *data, int64_t n, int32_t d, int64_t ld, double eps, int32_t mode, int32_t cp,
                        int32_t ap, uint64_t *total)
{
    if (!total || n < 0 || d < 1 || ld < d || (n > 0 && !data)) return 1;
    if (mode < 1 || mode > 4) return 2;
    if ((cp != 16 && cp != 32 && cp != 64) || (ap != 32 && ap != 64)) return 2;
    if (mode >= 3 && (cp != 16 || ap != 32)) return 2;
    const int dpad = (d + 15) / 16 * 16, nch = dpad / 16;
    const double thr = ap == 32 ? bfo_threshold_f(eps) : bfo_threshold(eps);
    double *X = (double *)calloc((size_t)(n > 0 ? n : 1) * dpad, sizeof(double));
    double *cn = (double *)calloc((size_t)(n > 0 ? n : 1) * nch, sizeof(double));
    if (!X || !cn) {
        free(X);
        free(cn);
        return 3;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < d; ++k) X[i * dpad + k] = bfo_to_c(data[i * ld + k], cp); /* convertDataset */
    for (int64_t i = 0; i < n; ++i)
        for (int ch = 0; ch < nch; ++ch) {
            float s = 0.0f;
            for (int r = 0; r < 16; ++r) {
                const float v = (float)X[i * dpad + ch * 16 + r];
                s = fmaf(v, v, s); /* v*v is exact, so fused or not gives the same sum */
            }
            cn[i * nch + ch] = (double)s;
        }
    uint64_t hits = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : hits)
    for (int64_t i = 0; i < n; ++i) {
        const double *q = X + i * dpad;
        for (int64_t j = 0; j < n; ++j) {
            const double *c = X + j * dpad;
            double acc = 0.0;
            if (mode == 1) {
                for (int k = 0; k < d; ++k) {
                    const double t = bfo_csub(q[k], c[k], cp);
                    acc = cp == ap ? bfo_afma(t, t, acc, ap) : bfo_aadd(acc, bfo_to_a(bfo_cmul(t, t, cp), ap), ap);
                }
            } else if (mode == 2) {
                for (int k = 0; k < d; ++k) {
                    const double qm2 = bfo_cmul(-2.0, q[k], cp);
                    const double q2 = bfo_amul(bfo_to_a(q[k], ap), bfo_to_a(q[k], ap), ap);
                    double t;
                    if (cp == ap) t = bfo_afma(c[k], c[k], bfo_afma(qm2, c[k], q2, ap), ap);
                    else
                        t = bfo_aadd(bfo_aadd(bfo_to_a(bfo_cmul(qm2, c[k], cp), ap), q2, ap),
                                     bfo_to_a(bfo_cmul(c[k], c[k], cp), ap), ap);
                    acc = bfo_aadd(acc, t, ap);
                }
            } else {
                for (int ch = 0; ch < nch; ++ch) {
                    float s = 0.0f;
                    for (int r = 0; r < 16; ++r) {
                        const int k = ch * 16 + r;
                        float p;
                        if (mode == 3) {
                            const float dd = (float)bfo_h16(q[k] - c[k]);
                            p = dd * dd;
                        } else {
                            p = (float)bfo_h16(-2.0 * q[k]) * (float)c[k];
                        }
                        s = r == 0 ? p : s + p;
                    }
                    if (mode == 3) acc = (double)((float)acc + s);
                    else acc = (double)(((float)acc + ((float)cn[j * nch + ch] + s)) + (float)cn[i * nch + ch]);
                }
            }
            const double v = (mode == 2 || mode == 4) ? fabs(acc) : acc;
            if (v <= thr) ++hits;
        }
    }
    free(X);
    free(cn);
    *total = hits;
    return 0;
}
