/*
 * All-host-cores variant of the oracle's GMRES(m) + block-Jacobi ILU(0):
 * GMRES.java:127-176 over BlockDiagonalPreconditioner.java:46-75 wrapping
 * ILU.java:80-125 / UnitLowerCompRowMatrix.java:49-68 /
 * UpperCompRowMatrix.java:49-68.
 *
 * TEST INFRASTRUCTURE ONLY: the second leg of the timed CPU baseline (MTJ
 * itself is single-threaded; this is what the same algorithm gets from every
 * host core).  Nothing in the shipped library links or calls it.
 *
 * Parallel over rows for the SpMV and the vector updates, over blocks for
 * the factorisation and the sweeps.  Each block's rows are factored and
 * swept in the reference's order, so the factors and every preconditioner
 * application equal the sequential oracle's bit for bit.  Dot products and
 * norms sum fixed contiguous chunks per thread and then the chunk sums in
 * thread order: deterministic for a thread count, but not the reference's
 * left-to-right order, so iterates agree with or_gmres to round-off only.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define OMP_MAX_THREADS 512

static double pdot(int n, const double *x, const double *y)
{
    double part[OMP_MAX_THREADS];
    int nt = 1;
#pragma omp parallel
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const long lo = (long)n * t / T, hi = (long)n * (t + 1) / T;
        double s = 0;
        for (long i = lo; i < hi; ++i)
            s += x[i] * y[i];
        part[t] = s;
        if (t == 0)
            nt = T;
    }
    double s = 0;
    for (int t = 0; t < nt; ++t)
        s += part[t];
    return s;
}

static double pnorm2(int n, const double *x) { return sqrt(pdot(n, x, x)); }

/* Math.max: NaN if either argument is NaN */
static double jmax(double a, double b)
{
    if (a != a)
        return a;
    if (b != b)
        return b;
    return a > b ? a : b;
}

/* y = A x with CompRowMatrix.mult's per-row order and zero rule
 * (CompRowMatrix.java:301-333) */
static void pmult(int n, const int *rowptr, const int *colind, const double *data,
                  const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double dot = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            dot += data[j] * x[colind[j]];
        y[i] = dot != 0 ? dot : 0.0;
    }
}

/* u = b - A x (multAdd(-1, x, u) after u = b, CompRowMatrix.java:336-353) */
static void presidual(int n, const int *rowptr, const int *colind, const double *data,
                      const double *b, const double *x, double *u)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double dot = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            dot += data[j] * x[colind[j]];
        u[i] = b[i] + -1 * dot;
    }
}

static void pset_scaled(int n, double *x, double alpha, const double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        x[i] = alpha == 0 ? 0.0 : alpha * y[i];
}

static void padd_scaled(int n, double *x, double alpha, const double *y)
{
    if (alpha == 0)
        return;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        x[i] += alpha * y[i];
}

/* ILU(0) of rows r0..r1-1 of a block-diagonal matrix: ILU.java:86-125 with
 * k running over the block (no entry of the block couples outside it) */
static int ilu_factor_rows(int r0, int r1, int nnz, const int *rowptr, const int *colind,
                           double *lu, const int *diagind)
{
    for (int k = r0 + 1; k < r1; ++k)
        for (int i = rowptr[k]; i < diagind[k]; ++i) {
            const int index = colind[i];
            const double LUii = lu[diagind[index]];
            if (LUii == 0)
                return OR_ZERO_PIVOT;
            const double LUki = (lu[i] /= LUii);
            for (int j = diagind[index] + 1, l = rowptr[k] + 1; j < rowptr[index + 1]; ++j) {
                while (l < rowptr[k + 1] && colind[l] < colind[j])
                    l++;
                if (l < nnz && colind[l] == colind[j])
                    lu[l] -= LUki * lu[j];
            }
        }
    return OR_OK;
}

typedef struct {
    int n, nblocks;
    int *bs, *rowptr, *colind, *diagind;
    double *lu, *y;
} bjilu_t;

static void bjilu_apply(const bjilu_t *M, const double *b, double *x)
{
    const int *rowptr = M->rowptr, *colind = M->colind, *diagind = M->diagind;
    const double *lu = M->lu;
    double *y = M->y;
#pragma omp parallel for schedule(dynamic, 1)
    for (int blk = 0; blk < M->nblocks; ++blk) {
        const int r0 = M->bs[blk], r1 = M->bs[blk + 1];
        for (int i = r0; i < r1; ++i) {
            double sum = 0;
            for (int j = rowptr[i]; j < diagind[i]; ++j)
                sum += lu[j] * y[colind[j]];
            y[i] = b[i] - sum;
        }
        for (int i = r1 - 1; i >= r0; --i) {
            double sum = 0;
            for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j)
                sum += lu[j] * x[colind[j]];
            x[i] = (y[i] - sum) / lu[diagind[i]];
        }
    }
}

static void bjilu_free(bjilu_t *M)
{
    free(M->bs);
    free(M->rowptr);
    free(M->colind);
    free(M->diagind);
    free(M->lu);
    free(M->y);
}

static int bjilu_setup(bjilu_t *M, int n, const int *rowptr, const int *colind,
                       const double *data, int nblocks)
{
    memset(M, 0, sizeof(*M));
    M->n = n;
    M->nblocks = nblocks;
    M->bs = (int *)malloc(sizeof(int) * (nblocks + 1));
    if (or_block_starts(n, nblocks, M->bs))
        return OR_BAD_ARGUMENT;
    const int nnz = or_block_filter_count(n, rowptr, colind, nblocks, M->bs);
    M->rowptr = (int *)malloc(sizeof(int) * (n + 1));
    M->colind = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    M->lu = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    M->diagind = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    M->y = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    or_block_filter(n, rowptr, colind, data, nblocks, M->bs, M->rowptr, M->colind, M->lu,
                    NULL);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int k = 0; k < n; ++k) {
        const int d = or_binary_search(M->colind, k, M->rowptr[k], M->rowptr[k + 1]);
        if (d < 0)
            bad |= 1;
        M->diagind[k] = d;
    }
    if (bad)
        return OR_MISSING_DIAGONAL;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int blk = 0; blk < nblocks; ++blk)
        bad |= ilu_factor_rows(M->bs[blk], M->bs[blk + 1], nnz, M->rowptr, M->colind, M->lu,
                               M->diagind) != OR_OK;
    return bad ? OR_ZERO_PIVOT : OR_OK;
}

typedef struct {
    double c, s;
} rot_t;

static rot_t rot_make(double x, double y)
{
    /* GivensRotation(x, y) (GivensRotation.java:42-60) */
    rot_t g;
    const double roe = fabs(x) > fabs(y) ? x : y;
    const double scale = fabs(x) + fabs(y);
    if (scale != 0) {
        const double xs = x / scale, ys = y / scale;
        double r = scale * sqrt(xs * xs + ys * ys);
        if (roe < 0)
            r *= -1;
        g.c = x / r;
        g.s = y / r;
    } else {
        g.c = 1;
        g.s = 0;
    }
    return g;
}

static void rot_apply(rot_t g, double *a, int i1, int i2)
{
    const double t = g.c * a[i1] + g.s * a[i2];
    a[i2] = -g.s * a[i1] + g.c * a[i2];
    a[i1] = t;
}

/* GMRES(m) (GMRES.java:127-176) + block-Jacobi ILU(0) over equal row
 * ranges, DefaultIterationMonitor semantics (rtol/atol/dtol/maxIter, NaN). */
int or_gmres_bjilu_omp(int n, const int *rowptr, const int *colind, const double *data,
                       const double *b, double *x, const or_opts *opts, or_stats *stats,
                       int nthreads)
{
    if (nthreads > 0)
        omp_set_num_threads(nthreads < OMP_MAX_THREADS ? nthreads : OMP_MAX_THREADS);
    const int m = opts->restart > 0 ? opts->restart : 30;
    bjilu_t M;
    int st = bjilu_setup(&M, n, rowptr, colind, data, opts->nblocks);
    memset(stats, 0, sizeof(*stats));
    if (st) {
        bjilu_free(&M);
        stats->status = st;
        return st;
    }
    double *w = (double *)calloc((size_t)n * (3 + m + 1), sizeof(double));
    double *ww = w, *u = w + n, *r = w + 2 * (size_t)n, *V = w + 3 * (size_t)n;
    double *s = (double *)calloc(m + 1, sizeof(double));
    double *Hm = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    rot_t *rot = (rot_t *)calloc(m + 1, sizeof(rot_t));
    const int ld = m + 1;
    int iter = 0;
    double initR = 0, res = 0;
    /* DefaultIterationMonitor.convergedI + AbstractIterationMonitor.converged */
#define MON(rv, conv)                                                                   \
    do {                                                                                \
        res = (rv);                                                                     \
        conv = 0;                                                                       \
        if (iter == 0)                                                                  \
            initR = res;                                                                \
        if (res < jmax(opts->rtol * initR, opts->atol))                                 \
            conv = 1;                                                                   \
        else if (res > opts->dtol * initR)                                              \
            st = OR_DIVERGENCE;                                                         \
        else if (iter >= opts->max_iter)                                                \
            st = OR_ITERATIONS;                                                         \
        else if (res != res)                                                            \
            st = OR_DIVERGENCE;                                                         \
    } while (0)
    presidual(n, rowptr, colind, data, b, x, u);
    bjilu_apply(&M, u, r);
    double normr = pnorm2(n, r);
    int conv = 0;
    for (iter = 0;; iter++) {
        MON(pnorm2(n, r), conv);
        if (st || conv)
            break;
        pset_scaled(n, V, 1 / normr, r);
        for (int k = 0; k <= m; ++k)
            s[k] = 0;
        s[0] = normr;
        int i = 0;
        for (;; i++, iter++) {
            if (!(i < m))
                break;
            MON(fabs(s[i]), conv);
            if (st || conv)
                break;
            pmult(n, rowptr, colind, data, V + (size_t)i * n, u);
            bjilu_apply(&M, u, ww);
            for (int k = 0; k <= i; k++) {
                Hm[k + i * ld] = pdot(n, ww, V + (size_t)k * n);
                padd_scaled(n, ww, -Hm[k + i * ld], V + (size_t)k * n);
            }
            Hm[i + 1 + i * ld] = pnorm2(n, ww);
            pset_scaled(n, V + (size_t)(i + 1) * n, 1. / Hm[i + 1 + i * ld], ww);
            for (int k = 0; k < i; ++k)
                rot_apply(rot[k], Hm + i * ld, k, k + 1);
            rot[i] = rot_make(Hm[i + i * ld], Hm[i + 1 + i * ld]);
            rot_apply(rot[i], Hm + i * ld, i, i + 1);
            rot_apply(rot[i], s, i, i + 1);
        }
        if (st)
            break;
        conv = 0;
        /* upper-triangular solve of the i x i system (dtrtrs) */
        for (int k = 0; k < i; ++k)
            if (Hm[k + k * ld] == 0)
                st = OR_SINGULAR;
        if (st)
            break;
        for (int k = i - 1; k >= 0; --k)
            if (s[k] != 0) {
                s[k] = s[k] / Hm[k + k * ld];
                for (int q = 0; q < k; ++q)
                    s[q] = s[q] - s[k] * Hm[q + k * ld];
            }
        for (int j = 0; j < i; j++)
            padd_scaled(n, x, s[j], V + (size_t)j * n);
        presidual(n, rowptr, colind, data, b, x, u);
        bjilu_apply(&M, u, r);
        normr = pnorm2(n, r);
    }
#undef MON
    stats->iterations = iter;
    stats->residual = res;
    stats->status = st;
    free(w);
    free(s);
    free(Hm);
    free(rot);
    bjilu_free(&M);
    return st;
}
