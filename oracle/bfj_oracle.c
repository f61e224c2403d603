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
