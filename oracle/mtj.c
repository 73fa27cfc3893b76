/*
 * Oracle (test infrastructure only): scalar restatement of the MTJ sparse
 * path.  Every loop keeps the reference's statement order; compile with
 * -O2 -ffp-contract=off.  See oracle.h for the file list.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* java.lang.Math.max for doubles: NaN wins. */
static double java_max(double a, double b)
{
    if (a != a)
        return a;
    if (b != b)
        return b;
    return a > b ? a : b;
}

/* mtj sparse/Arrays.java:101-113 (binarySearch over [begin,end)). */
int or_binary_search(const int *index, int key, int begin, int end)
{
    end--;
    while (begin <= end) {
        int mid = (end + begin) >> 1;
        if (index[mid] < key)
            begin = mid + 1;
        else if (index[mid] > key)
            end = mid - 1;
        else
            return mid;
    }
    return -1;
}

/* CompRowMatrix.construct(int[][]) (CompRowMatrix.java:165-190): copy the
 * per-row column lists, range-check, sort each row ascending. */
int or_csr_construct(int n, int ncols, const int *rowlen, const int *cols,
                     int *rowptr, int *colind)
{
    rowptr[0] = 0;
    int src = 0;
    for (int i = 1; i <= n; ++i) {
        rowptr[i] = rowptr[i - 1] + rowlen[i - 1];
        for (int j = rowptr[i - 1]; j < rowptr[i]; ++j, ++src) {
            int c = cols[src];
            if (c < 0 || c >= ncols)
                return OR_BAD_ARGUMENT;
            colind[j] = c;
        }
        /* insertion sort of the row (integers: any sort gives the same) */
        for (int j = rowptr[i - 1] + 1; j < rowptr[i]; ++j) {
            int v = colind[j], k = j - 1;
            while (k >= rowptr[i - 1] && colind[k] > v) {
                colind[k + 1] = colind[k];
                --k;
            }
            colind[k + 1] = v;
        }
    }
    return OR_OK;
}

/* CompRowMatrix.getIndex/add/set (CompRowMatrix.java:396-437). */
static int csr_index(const int *rowptr, const int *colind, int row, int col)
{
    int i = or_binary_search(colind, col, rowptr[row], rowptr[row + 1]);
    if (i != -1 && colind[i] == col)
        return i;
    return -1;
}

int or_csr_add(const int *rowptr, const int *colind, double *data, int row,
               int col, double value)
{
    int i = csr_index(rowptr, colind, row, col);
    if (i < 0)
        return OR_NOT_IN_PATTERN;
    data[i] += value;
    return OR_OK;
}

int or_csr_set(const int *rowptr, const int *colind, double *data, int row,
               int col, double value)
{
    int i = csr_index(rowptr, colind, row, col);
    if (i < 0)
        return OR_NOT_IN_PATTERN;
    data[i] = value;
    return OR_OK;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Matrices.zeroRows (Matrices.java:631-652): every stored entry of a listed
 * row becomes 0 (diagonal -> `diagonal`); then the diagonal is set again. */
int or_zero_rows(int n, const int *rowptr, const int *colind, double *data,
                 double diagonal, const int *rows, int nrows)
{
    int *rs = (int *)malloc(sizeof(int) * (nrows > 0 ? nrows : 1));
    memcpy(rs, rows, sizeof(int) * nrows);
    qsort(rs, nrows, sizeof(int), cmp_int);
    for (int r = 0; r < n; ++r) {
        int lo = 0, hi = nrows - 1, found = 0;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            if (rs[mid] < r)
                lo = mid + 1;
            else if (rs[mid] > r)
                hi = mid - 1;
            else {
                found = 1;
                break;
            }
        }
        if (!found)
            continue;
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j)
            data[j] = (colind[j] == r) ? diagonal : 0.0;
    }
    free(rs);
    if (diagonal != 0)
        for (int k = 0; k < nrows; ++k)
            if (or_csr_set(rowptr, colind, data, rows[k], rows[k], diagonal))
                return OR_NOT_IN_PATTERN;
    return OR_OK;
}

/* CompRowMatrix.mult(Vector,Vector), DenseVector branch
 * (CompRowMatrix.java:301-320): y zeroed, y[i] stored only if dot != 0. */
void or_csr_mult(int n, const int *rowptr, const int *colind,
                 const double *data, const double *x, double *y)
{
    for (int i = 0; i < n; ++i)
        y[i] = 0.0;
    for (int i = 0; i < n; ++i) {
        double dot = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            dot += data[j] * x[colind[j]];
        if (dot != 0)
            y[i] = dot;
    }
}

void or_csr_mult_f32(int n, const int *rowptr, const int *colind,
                     const float *data, const float *x, float *y)
{
    for (int i = 0; i < n; ++i)
        y[i] = 0.0f;
    for (int i = 0; i < n; ++i) {
        float dot = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            dot += data[j] * x[colind[j]];
        if (dot != 0)
            y[i] = dot;
    }
}

/* CompRowMatrix.multAdd (CompRowMatrix.java:336-353). */
void or_csr_multadd(int n, const int *rowptr, const int *colind,
                    const double *data, double alpha, const double *x,
                    double *y)
{
    for (int i = 0; i < n; ++i) {
        double dot = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            dot += data[j] * x[colind[j]];
        y[i] += alpha * dot;
    }
}

/* CompRowMatrix.transMult (CompRowMatrix.java:356-372): row-order scatter. */
void or_csr_transmult(int n, int ncols, const int *rowptr, const int *colind,
                      const double *data, const double *x, double *y)
{
    for (int i = 0; i < ncols; ++i)
        y[i] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            y[colind[j]] += data[j] * x[i];
}

/* CompRowMatrix.transMultAdd (CompRowMatrix.java:375-394):
 * y = alpha*(y/alpha + A'x) with DenseVector.scale as a plain product. */
void or_csr_transmultadd(int n, int ncols, const int *rowptr,
                         const int *colind, const double *data, double alpha,
                         const double *x, double *y)
{
    double inv = 1. / alpha;
    for (int i = 0; i < ncols; ++i)
        y[i] *= inv;
    for (int i = 0; i < n; ++i)
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            y[colind[j]] += data[j] * x[i];
    for (int i = 0; i < ncols; ++i)
        y[i] *= alpha;
}

/* ---- DenseVector (DenseVector.java:216-295) ----------------------------- */
double or_dot(int n, const double *x, const double *y)
{
    double dot = 0.;
    for (int i = 0; i < n; ++i)
        dot += x[i] * y[i];
    return dot;
}

double or_norm2(int n, const double *x)
{
    double norm = 0;
    for (int i = 0; i < n; ++i)
        norm += x[i] * x[i];
    return sqrt(norm);
}

/* x = alpha*y  (set(alpha,y); alpha == 0 zeroes) */
static void v_set_scaled(int n, double *x, double alpha, const double *y)
{
    if (alpha == 0) {
        for (int i = 0; i < n; ++i)
            x[i] = 0;
        return;
    }
    for (int i = 0; i < n; ++i)
        x[i] = alpha * y[i];
}

/* x += alpha*y  (add(alpha,y); alpha == 0 is a no-op) */
static void v_add_scaled(int n, double *x, double alpha, const double *y)
{
    if (alpha == 0)
        return;
    for (int i = 0; i < n; i++)
        x[i] += alpha * y[i];
}

static void v_add(int n, double *x, const double *y)
{
    for (int i = 0; i < n; i++)
        x[i] += y[i];
}

static void v_copy(int n, double *x, const double *y)
{
    memcpy(x, y, sizeof(double) * n);
}

/* ---- ILU(0) (ILU.java:86-140) ------------------------------------------- */
#define ILU_FACTOR_BODY(T)                                                    \
    for (int k = 0; k < n; ++k) {                                            \
        int d = or_binary_search(colind, k, rowptr[k], rowptr[k + 1]);      \
        if (d < 0) {                                                         \
            if (bad_row)                                                     \
                *bad_row = k;                                                \
            return OR_MISSING_DIAGONAL;                                      \
        }                                                                    \
        diagind[k] = d;                                                      \
    }                                                                        \
    const int nnz = rowptr[n];                                               \
    for (int k = 1; k < n; ++k)                                              \
        for (int i = rowptr[k]; i < diagind[k]; ++i) {                       \
            int index = colind[i];                                           \
            T LUii = lu[diagind[index]];                                     \
            if (LUii == 0) {                                                 \
                if (bad_row)                                                 \
                    *bad_row = i;                                            \
                return OR_ZERO_PIVOT;                                        \
            }                                                                \
            T LUki = (lu[i] /= LUii);                                        \
            for (int j = diagind[index] + 1, l = rowptr[k] + 1;              \
                 j < rowptr[index + 1]; ++j) {                               \
                while (l < rowptr[k + 1] && colind[l] < colind[j])           \
                    l++;                                                     \
                /* the reference reads colind[l] also at l == rowptr[k+1] */ \
                if (l < nnz && colind[l] == colind[j])                       \
                    lu[l] -= LUki * lu[j];                                   \
            }                                                                \
        }                                                                    \
    return OR_OK;

int or_ilu_factor(int n, const int *rowptr, const int *colind, double *lu,
                  int *diagind, int *bad_row)
{
    ILU_FACTOR_BODY(double)
}

int or_ilu_factor_f32(int n, const int *rowptr, const int *colind, float *lu,
                      int *diagind, int *bad_row)
{
    ILU_FACTOR_BODY(float)
}

/* ---- ICC (ICC.java:71-125) ---------------------------------------------- */
/* R.set(A) then factor(): a dense work row Rk per row k, eliminated against
 * every earlier row i (skipped when Rk[i] / R(i,i) == 0) with row i's stored
 * upper entries; row k's entries from the diagonal on become Rk / sqrt(Rk[k]).
 * Entries left of the diagonal keep A's values.  bad_row: the 1-based row of
 * the error (as the reference's messages). */
int or_icc_factor(int n, const int *rowptr, const int *colind, double *data, int *diagind,
                  int *bad_row)
{
    for (int k = 0; k < n; ++k) {
        int d = or_binary_search(colind, k, rowptr[k], rowptr[k + 1]);
        if (d < 0) {
            if (bad_row)
                *bad_row = k + 1;
            return OR_MISSING_DIAGONAL;
        }
        diagind[k] = d;
    }
    double *Rk = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int k = 0; k < n; ++k) {
        for (int i = 0; i < n; ++i)
            Rk[i] = 0;
        for (int i = rowptr[k]; i < rowptr[k + 1]; ++i)
            Rk[colind[i]] = data[i];
        for (int i = 0; i < k; ++i) {
            const double Rii = data[diagind[i]];
            if (Rii == 0) {
                if (bad_row)
                    *bad_row = i + 1;
                free(Rk);
                return OR_ZERO_PIVOT;
            }
            const double Rki = Rk[i] / Rii;
            if (Rki == 0)
                continue;
            for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j)
                Rk[colind[j]] -= Rki * data[j];
        }
        if (Rk[k] == 0) {
            if (bad_row)
                *bad_row = k + 1;
            free(Rk);
            return OR_ZERO_PIVOT;
        }
        const double sqRkk = sqrt(Rk[k]);
        for (int i = diagind[k]; i < rowptr[k + 1]; ++i)
            data[i] = Rk[colind[i]] / sqRkk;
    }
    free(Rk);
    return OR_OK;
}

/* ICC.apply = transApply: UpperCompRowMatrix.transSolve (:70-90) then
 * UpperCompRowMatrix.solve (:49-67) with R's upper part. */
void or_icc_apply(int n, const int *rowptr, const int *colind, const double *R,
                  const int *diagind, const double *b, double *x, double *y)
{
    memcpy(y, b, sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        y[i] /= R[diagind[i]];
        for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j)
            y[colind[j]] -= R[j] * y[i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double sum = 0;
        for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j)
            sum += R[j] * x[colind[j]];
        x[i] = (y[i] - sum) / R[diagind[i]];
    }
}

/* ILU.apply: UnitLowerCompRowMatrix.solve (:49-68) then
 * UpperCompRowMatrix.solve (:49-68). */
#define ILU_APPLY_BODY(T)                                                     \
    for (int i = 0; i < n; ++i) {                                            \
        T sum = 0;                                                           \
        for (int j = rowptr[i]; j < diagind[i]; ++j)                         \
            sum += lu[j] * y[colind[j]];                                     \
        y[i] = b[i] - sum;                                                   \
    }                                                                        \
    for (int i = n - 1; i >= 0; --i) {                                       \
        T sum = 0;                                                           \
        for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j)                 \
            sum += lu[j] * x[colind[j]];                                     \
        x[i] = (y[i] - sum) / lu[diagind[i]];                                \
    }

void or_ilu_apply(int n, const int *rowptr, const int *colind,
                  const double *lu, const int *diagind, const double *b,
                  double *x, double *y)
{
    ILU_APPLY_BODY(double)
}

void or_ilu_apply_f32(int n, const int *rowptr, const int *colind,
                      const float *lu, const int *diagind, const float *b,
                      float *x, float *y)
{
    ILU_APPLY_BODY(float)
}

/* ILU.transApply: UpperCompRowMatrix.transSolve (:70-91) then
 * UnitLowerCompRowMatrix.transSolve (:70-88), column sweeps. */
void or_ilu_transapply(int n, const int *rowptr, const int *colind,
                       const double *lu, const int *diagind, const double *b,
                       double *x, double *y)
{
    v_copy(n, y, b);
    for (int i = 0; i < n; ++i) {
        y[i] /= lu[diagind[i]];
        for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j)
            y[colind[j]] -= lu[j] * y[i];
    }
    v_copy(n, x, y);
    for (int i = n - 1; i >= 0; --i)
        for (int j = rowptr[i]; j < diagind[i]; ++j)
            x[colind[j]] -= lu[j] * x[i];
}

/* Block-Jacobi: DistRowMatrix ownership of contiguous row ranges
 * (DistRowMatrix.java:56-76); the block preconditioner sees only the
 * diagonal block A_rr (BlockDiagonalPreconditioner.java:46-75).  ILU(0) of
 * the block-diagonal part equals per-block ILU(0) with column indices
 * shifted, so the blocks are kept in global numbering here. */
int or_block_starts(int n, int nblocks, int *bstart)
{
    if (nblocks < 1 || nblocks > n)
        return OR_BAD_ARGUMENT;
    for (int r = 0; r <= nblocks; ++r)
        bstart[r] = (int)(((long long)n * r) / nblocks);
    return OR_OK;
}

static int block_of(int row, int nblocks, const int *bstart)
{
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= row)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

int or_block_filter_count(int n, const int *rowptr, const int *colind,
                          int nblocks, const int *bstart)
{
    int cnt = 0;
    for (int i = 0; i < n; ++i) {
        int b = block_of(i, nblocks, bstart);
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            if (colind[j] >= bstart[b] && colind[j] < bstart[b + 1])
                cnt++;
    }
    return cnt;
}

int or_block_filter(int n, const int *rowptr, const int *colind,
                    const double *data, int nblocks, const int *bstart,
                    int *browptr, int *bcolind, double *bdata, int *src_index)
{
    int cnt = 0;
    browptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        int b = block_of(i, nblocks, bstart);
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            if (colind[j] >= bstart[b] && colind[j] < bstart[b + 1]) {
                bcolind[cnt] = colind[j];
                if (bdata)
                    bdata[cnt] = data[j];
                if (src_index)
                    src_index[cnt] = j;
                cnt++;
            }
        browptr[i + 1] = cnt;
    }
    return OR_OK;
}

/* ---- Preconditioner objects --------------------------------------------- */
typedef struct {
    int kind, n;
    int *rowptr, *colind, *diagind;
    double *lu, *y;
    or_amg *amg;
    /* SSOR: the matrix and parameters; diagonal: 1 / A(i,i) */
    const int *arp, *aci;
    const double *av;
    or_amg_params sp;
    double *invdiag;
} prec_t;

static void prec_free(prec_t *p)
{
    if (p->kind == 3) {
        or_amg_free(p->amg);
    } else if (p->kind == 4) {
        free(p->invdiag);
    } else if (p->kind == 5) {
        /* nothing owned */
    } else if (p->kind != 0) {
        free(p->rowptr);
        free(p->colind);
        free(p->diagind);
        free(p->lu);
    }
    free(p->y);
}

/* setMatrix: copy A (or its block-diagonal part) and factor. */
static int prec_setup(prec_t *p, int n, const int *rowptr, const int *colind,
                      const double *data, const or_opts *o, int *bad_row)
{
    memset(p, 0, sizeof(*p));
    p->kind = o->prec_kind;
    p->n = n;
    p->y = (double *)malloc(sizeof(double) * n);
    if (p->kind == 0)
        return OR_OK;
    if (p->kind == 3)
        return or_amg_setup(n, rowptr, colind, data, &o->amg, &p->amg, bad_row);
    if (p->kind == 4) {
        /* DiagonalPreconditioner.setMatrix (DiagonalPreconditioner.java:67-78) */
        p->invdiag = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) {
            int d = or_binary_search(colind, i, rowptr[i], rowptr[i + 1]);
            double v = d >= 0 ? data[d] : 0.0;
            if (v == 0) {
                if (bad_row)
                    *bad_row = i;
                return OR_ZERO_PIVOT;
            }
            p->invdiag[i] = 1 / v;
        }
        return OR_OK;
    }
    if (p->kind == 5) {
        for (int i = 0; i < n; ++i)
            if (or_binary_search(colind, i, rowptr[i], rowptr[i + 1]) < 0) {
                if (bad_row)
                    *bad_row = i + 1;
                return OR_MISSING_DIAGONAL;
            }
        p->arp = rowptr;
        p->aci = colind;
        p->av = data;
        p->sp = o->amg;
        return OR_OK;
    }
    int nnz;
    p->rowptr = (int *)malloc(sizeof(int) * (n + 1));
    if (p->kind == 6) {
        /* ICC: R = copy of A, factored in place */
        nnz = rowptr[n];
        p->colind = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        p->lu = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        memcpy(p->rowptr, rowptr, sizeof(int) * (n + 1));
        memcpy(p->colind, colind, sizeof(int) * nnz);
        memcpy(p->lu, data, sizeof(double) * nnz);
        p->diagind = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
        return or_icc_factor(n, p->rowptr, p->colind, p->lu, p->diagind, bad_row);
    }
    if (p->kind == 1) {
        nnz = rowptr[n];
        p->colind = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        p->lu = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        memcpy(p->rowptr, rowptr, sizeof(int) * (n + 1));
        memcpy(p->colind, colind, sizeof(int) * nnz);
        memcpy(p->lu, data, sizeof(double) * nnz);
    } else if (p->kind == 2) {
        int *bs = (int *)malloc(sizeof(int) * (o->nblocks + 1));
        if (or_block_starts(n, o->nblocks, bs)) {
            free(bs);
            return OR_BAD_ARGUMENT;
        }
        nnz = or_block_filter_count(n, rowptr, colind, o->nblocks, bs);
        p->colind = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        p->lu = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        or_block_filter(n, rowptr, colind, data, o->nblocks, bs, p->rowptr,
                        p->colind, p->lu, NULL);
        free(bs);
    } else {
        return OR_BAD_ARGUMENT;
    }
    p->diagind = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    return or_ilu_factor(n, p->rowptr, p->colind, p->lu, p->diagind, bad_row);
}

static void prec_apply(prec_t *p, const double *b, double *x)
{
    if (p->kind == 0) {
        v_copy(p->n, x, b);
        return;
    }
    if (p->kind == 3) {
        /* AMG.apply starts the cycle from x's old content (AMG.java:208-216) */
        or_amg_apply(p->amg, b, x, 0);
        return;
    }
    if (p->kind == 4) {
        for (int i = 0; i < p->n; ++i)
            x[i] = b[i] * p->invdiag[i];
        return;
    }
    if (p->kind == 5) {
        /* SSOR.apply starts from x's old content (SSOR.java:143-194) */
        or_ssor_apply(p->n, p->arp, p->aci, p->av, p->sp.block_rows, p->sp.ordering,
                      p->sp.omega_pre_f, p->sp.omega_pre_r, p->sp.reverse, b, x, NULL);
        return;
    }
    if (p->kind == 6) {
        or_icc_apply(p->n, p->rowptr, p->colind, p->lu, p->diagind, b, x, p->y);
        return;
    }
    or_ilu_apply(p->n, p->rowptr, p->colind, p->lu, p->diagind, b, x, p->y);
}

/* ---- DefaultIterationMonitor (DefaultIterationMonitor.java:133-156,
 *      AbstractIterationMonitor.java:56-89) ------------------------------- */
typedef struct {
    int iter;
    double initR, rtol, atol, dtol, residual;
    int maxIter;
    double *hist;
    int hist_cap, hist_len;
} monitor_t;

static void mon_init(monitor_t *m, const or_opts *o, double *hist, int cap)
{
    memset(m, 0, sizeof(*m));
    m->rtol = o->rtol;
    m->atol = o->atol;
    m->dtol = o->dtol;
    m->maxIter = o->max_iter;
    m->hist = hist;
    m->hist_cap = cap;
}

/* returns status (0 = keep going / converged per *conv) */
static int mon_converged(monitor_t *m, double r, int *conv)
{
    if (m->hist && m->hist_len < m->hist_cap)
        m->hist[m->hist_len] = r;
    m->hist_len++;
    m->residual = r;
    *conv = 0;
    if (m->iter == 0)
        m->initR = r;
    if (r < java_max(m->rtol * m->initR, m->atol)) {
        *conv = 1;
        return OR_OK;
    }
    if (r > m->dtol * m->initR)
        return OR_DIVERGENCE;
    if (m->iter >= m->maxIter)
        return OR_ITERATIONS;
    if (r != r)
        return OR_DIVERGENCE;
    return OR_OK;
}

static void fill_stats(or_stats *s, const monitor_t *m, int status)
{
    if (!s)
        return;
    s->iterations = m->iter;
    s->residual = m->residual;
    s->status = status;
    s->hist_len = m->hist_len;
}

/* ---- BiCGstab (BiCGstab.java:65-116) ------------------------------------ */
int or_bicgstab(int n, const int *rowptr, const int *colind,
                const double *data, const double *b, double *x,
                const or_opts *opts, or_stats *stats, double *hist,
                int hist_cap)
{
    prec_t M;
    int bad = -1;
    int st = prec_setup(&M, n, rowptr, colind, data, opts, &bad);
    monitor_t it;
    mon_init(&it, opts, hist, hist_cap);
    if (st) {
        prec_free(&M);
        fill_stats(stats, &it, st);
        if (stats)
            stats->bad_row = bad;
        return st;
    }
    double *w = (double *)malloc(sizeof(double) * 9 * (size_t)n);
    double *p = w, *s = w + n, *phat = w + 2 * n, *shat = w + 3 * n,
           *t = w + 4 * n, *v = w + 5 * n, *temp = w + 6 * n, *r = w + 7 * n,
           *rtilde = w + 8 * n;
    memset(w, 0, sizeof(double) * 9 * (size_t)n);

    double rho_1 = 1, rho_2 = 1, alpha = 1, beta = 1, omega = 1;
    int conv = 0;

    v_copy(n, r, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    v_copy(n, rtilde, r);

    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        rho_1 = or_dot(n, rtilde, r);
        if (rho_1 == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        if (omega == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        if (it.iter == 0)
            v_copy(n, p, r);
        else {
            beta = (rho_1 / rho_2) * (alpha / omega);
            v_set_scaled(n, temp, -omega, v);
            v_add(n, temp, p);
            v_copy(n, p, r);
            v_add_scaled(n, p, beta, temp);
        }
        prec_apply(&M, p, phat);
        or_csr_mult(n, rowptr, colind, data, phat, v);
        alpha = rho_1 / or_dot(n, rtilde, v);
        v_copy(n, s, r);
        v_add_scaled(n, s, -alpha, v);

        st = mon_converged(&it, or_norm2(n, s), &conv);
        if (st)
            break;
        if (conv) {
            v_add_scaled(n, x, alpha, phat);
            break;
        }
        prec_apply(&M, s, shat);
        or_csr_mult(n, rowptr, colind, data, shat, t);
        omega = or_dot(n, t, s) / or_dot(n, t, t);
        v_add_scaled(n, x, alpha, phat);
        v_add_scaled(n, x, omega, shat);
        v_copy(n, r, s);
        v_add_scaled(n, r, -omega, t);
        rho_2 = rho_1;
    }
    free(w);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}

/* ---- GivensRotation (GivensRotation.java:42-93) ------------------------- */
typedef struct {
    double c, s;
} givens_t;

static givens_t givens_make(double x, double y)
{
    givens_t g;
    double roe = fabs(x) > fabs(y) ? x : y;
    double scale = fabs(x) + fabs(y);
    if (scale != 0) {
        double xs = x / scale;
        double ys = y / scale;
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

/* apply to entries (i1, i2) of a column-major array with stride ld */
static void givens_apply(givens_t g, double *a, int i1, int i2)
{
    double temp = g.c * a[i1] + g.s * a[i2];
    a[i2] = -g.s * a[i1] + g.c * a[i2];
    a[i1] = temp;
}

/* dtrtrs('U','N','N') + reference dtrsm Left/Upper/NoTrans on one column
 * (AbstractTriangDenseMatrix.java:227-256). */
static int upper_solve(int m, const double *H, int ld, double *s)
{
    for (int k = 0; k < m; ++k)
        if (H[k + k * ld] == 0)
            return OR_SINGULAR;
    for (int k = m - 1; k >= 0; --k) {
        if (s[k] != 0) {
            s[k] = s[k] / H[k + k * ld];
            for (int i = 0; i < k; ++i)
                s[i] = s[i] - s[k] * H[i + k * ld];
        }
    }
    return OR_OK;
}

/* ---- GMRES(m) (GMRES.java:127-176) -------------------------------------- */
int or_gmres(int n, const int *rowptr, const int *colind, const double *data,
             const double *b, double *x, const or_opts *opts, or_stats *stats,
             double *hist, int hist_cap)
{
    const int m = opts->restart > 0 ? opts->restart : 30;
    prec_t M;
    int bad = -1;
    int st = prec_setup(&M, n, rowptr, colind, data, opts, &bad);
    monitor_t it;
    mon_init(&it, opts, hist, hist_cap);
    if (st) {
        prec_free(&M);
        fill_stats(stats, &it, st);
        if (stats)
            stats->bad_row = bad;
        return st;
    }
    double *w = (double *)calloc((size_t)n * (3 + m + 1), sizeof(double));
    double *ww = w, *u = w + n, *r = w + 2 * (size_t)n;
    double *V = w + 3 * (size_t)n; /* (m+1) basis vectors */
    double *s = (double *)calloc(m + 1, sizeof(double));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    givens_t *rot = (givens_t *)calloc(m + 1, sizeof(givens_t));
    const int ld = m + 1;
    int conv = 0;

    v_copy(n, u, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, u);
    prec_apply(&M, u, r);
    double normr = or_norm2(n, r);
    prec_apply(&M, b, u); /* computed and discarded, as in the reference */

    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        v_set_scaled(n, V, 1 / normr, r);
        for (int k = 0; k <= m; ++k)
            s[k] = 0;
        s[0] = normr;
        int i = 0;
        for (;; i++, it.iter++) {
            if (!(i < m))
                break;
            st = mon_converged(&it, fabs(s[i]), &conv);
            if (st || conv)
                break;
            double *vi = V + (size_t)i * n;
            or_csr_mult(n, rowptr, colind, data, vi, u);
            prec_apply(&M, u, ww);
            for (int k = 0; k <= i; k++) {
                H[k + i * ld] = or_dot(n, ww, V + (size_t)k * n);
                v_add_scaled(n, ww, -H[k + i * ld], V + (size_t)k * n);
            }
            H[i + 1 + i * ld] = or_norm2(n, ww);
            v_set_scaled(n, V + (size_t)(i + 1) * n, 1. / H[i + 1 + i * ld],
                         ww);
            for (int k = 0; k < i; ++k)
                givens_apply(rot[k], H + i * ld, k, k + 1);
            rot[i] = givens_make(H[i + i * ld], H[i + 1 + i * ld]);
            givens_apply(rot[i], H + i * ld, i, i + 1);
            givens_apply(rot[i], s, i, i + 1);
        }
        if (st)
            break;
        conv = 0;
        if (upper_solve(i, H, ld, s)) {
            st = OR_SINGULAR;
            break;
        }
        for (int j = 0; j < i; j++)
            v_add_scaled(n, x, s[j], V + (size_t)j * n);
        v_copy(n, u, b);
        or_csr_multadd(n, rowptr, colind, data, -1, x, u);
        prec_apply(&M, u, r);
        normr = or_norm2(n, r);
    }
    free(w);
    free(s);
    free(H);
    free(rot);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}

/* ---- CG (CG.java:58-92) --------------------------------------------------- */
int or_cg(int n, const int *rowptr, const int *colind, const double *data,
          const double *b, double *x, const or_opts *opts, or_stats *stats,
          double *hist, int hist_cap)
{
    prec_t M;
    int bad = -1;
    int st = prec_setup(&M, n, rowptr, colind, data, opts, &bad);
    monitor_t it;
    mon_init(&it, opts, hist, hist_cap);
    if (st) {
        prec_free(&M);
        fill_stats(stats, &it, st);
        if (stats)
            stats->bad_row = bad;
        return st;
    }
    double *w = (double *)calloc((size_t)n * 4 + 1, sizeof(double));
    double *r = w, *z = w + n, *p = w + 2 * (size_t)n, *q = w + 3 * (size_t)n;
    double alpha = 0, beta = 0, rho = 0, rho_1 = 0;
    int conv = 0;
    v_copy(n, r, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        prec_apply(&M, r, z);
        rho = or_dot(n, r, z);
        if (it.iter == 0) {
            v_copy(n, p, z);
        } else {
            beta = rho / rho_1;
            /* p.scale(beta).add(z) */
            for (int i = 0; i < n; ++i) {
                double pi = beta == 0 ? 0.0 : (beta == 1 ? p[i] : beta * p[i]);
                p[i] = pi + z[i];
            }
        }
        or_csr_mult(n, rowptr, colind, data, p, q);
        alpha = rho / or_dot(n, p, q);
        v_add_scaled(n, x, alpha, p);
        v_add_scaled(n, r, -alpha, q);
        rho_1 = rho;
    }
    free(w);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}

/* M.transApply: ILU.transApply (ILU.java:72-78), AMG.transApply; the
 * diagonal, SSOR and identity preconditioners' transApply is apply. */
static void prec_tapply(prec_t *p, const double *b, double *x)
{
    if (p->kind == 1 || p->kind == 2) {
        or_ilu_transapply(p->n, p->rowptr, p->colind, p->lu, p->diagind, b, x, p->y);
        return;
    }
    if (p->kind == 3) {
        or_amg_apply(p->amg, b, x, 1);
        return;
    }
    prec_apply(p, b, x);
}

/* x = beta*x + y with AbstractVector.scale (beta 0 zeroes, 1 keeps) then
 * DenseVector.add */
static void v_scale_add(int n, double *x, double beta, const double *y)
{
    for (int i = 0; i < n; ++i) {
        double xi = beta == 0 ? 0.0 : (beta == 1 ? x[i] : beta * x[i]);
        x[i] = xi + y[i];
    }
}

/* x.scale(beta) (AbstractVector.scale) */
static void v_scale(int n, double *x, double beta)
{
    if (beta == 1)
        return;
    for (int i = 0; i < n; ++i)
        x[i] = beta == 0 ? 0.0 : beta * x[i];
}

#define SOLVER_PROLOGUE                                                          \
    prec_t M;                                                                    \
    int bad = -1;                                                                \
    int st = prec_setup(&M, n, rowptr, colind, data, opts, &bad);                \
    monitor_t it;                                                                \
    mon_init(&it, opts, hist, hist_cap);                                         \
    if (st) {                                                                    \
        prec_free(&M);                                                           \
        fill_stats(stats, &it, st);                                              \
        if (stats)                                                               \
            stats->bad_row = bad;                                                \
        return st;                                                               \
    }

/* ---- BiCG (BiCG.java:74-121) ---------------------------------------------- */
int or_bicg(int n, const int *rowptr, const int *colind, const double *data,
            const double *b, double *x, const or_opts *opts, or_stats *stats,
            double *hist, int hist_cap)
{
    SOLVER_PROLOGUE
    double *w = (double *)calloc((size_t)n * 8 + 1, sizeof(double));
    double *r = w, *rt = w + n, *z = w + 2 * (size_t)n, *zt = w + 3 * (size_t)n;
    double *p = w + 4 * (size_t)n, *pt = w + 5 * (size_t)n, *q = w + 6 * (size_t)n,
           *qt = w + 7 * (size_t)n;
    double rho_1 = 1, rho_2 = 1, alpha = 1, beta = 1;
    int conv = 0;
    v_copy(n, r, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    v_copy(n, rt, r);
    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        prec_apply(&M, r, z);
        prec_tapply(&M, rt, zt);
        rho_1 = or_dot(n, z, rt);
        if (rho_1 == 0.) {
            st = OR_BREAKDOWN;
            break;
        }
        if (it.iter == 0) {
            v_copy(n, p, z);
            v_copy(n, pt, zt);
        } else {
            beta = rho_1 / rho_2;
            v_scale_add(n, p, beta, z);
            v_scale_add(n, pt, beta, zt);
        }
        or_csr_mult(n, rowptr, colind, data, p, q);
        or_csr_transmult(n, n, rowptr, colind, data, pt, qt);
        alpha = rho_1 / or_dot(n, pt, q);
        v_add_scaled(n, x, alpha, p);
        v_add_scaled(n, r, -alpha, q);
        v_add_scaled(n, rt, -alpha, qt);
        rho_2 = rho_1;
    }
    (void)beta;
    free(w);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}

/* ---- CGS (CGS.java:78-121) ------------------------------------------------ */
int or_cgs(int n, const int *rowptr, const int *colind, const double *data,
           const double *b, double *x, const or_opts *opts, or_stats *stats,
           double *hist, int hist_cap)
{
    SOLVER_PROLOGUE
    double *w = (double *)calloc((size_t)n * 10 + 1, sizeof(double));
    double *r = w, *rt = w + n, *u = w + 2 * (size_t)n, *p = w + 3 * (size_t)n;
    double *ph = w + 4 * (size_t)n, *vh = w + 5 * (size_t)n, *q = w + 6 * (size_t)n;
    double *sum = w + 7 * (size_t)n, *uh = w + 8 * (size_t)n, *qh = w + 9 * (size_t)n;
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    int conv = 0;
    v_copy(n, r, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    v_copy(n, rt, r);
    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        rho_1 = or_dot(n, rt, r);
        if (rho_1 == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        if (it.iter == 0) {
            v_copy(n, u, r);
            v_copy(n, p, u);
        } else {
            beta = rho_1 / rho_2;
            v_copy(n, u, r);
            v_add_scaled(n, u, beta, q);
            v_copy(n, sum, q);
            v_add_scaled(n, sum, beta, p);
            v_copy(n, p, u);
            v_add_scaled(n, p, beta, sum);
        }
        prec_apply(&M, p, ph);
        or_csr_mult(n, rowptr, colind, data, ph, vh);
        alpha = rho_1 / or_dot(n, rt, vh);
        v_set_scaled(n, q, -alpha, vh);
        v_add(n, q, u);
        v_copy(n, sum, u);
        v_add(n, sum, q);
        prec_apply(&M, sum, uh);
        v_add_scaled(n, x, alpha, uh);
        or_csr_mult(n, rowptr, colind, data, uh, qh);
        v_add_scaled(n, r, -alpha, qh);
        rho_2 = rho_1;
    }
    free(w);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}

/* ---- QMR (QMR.java:136-243), M1 = M2 = M ----------------------------------- */
int or_qmr(int n, const int *rowptr, const int *colind, const double *data,
           const double *b, double *x, const or_opts *opts, or_stats *stats,
           double *hist, int hist_cap)
{
    SOLVER_PROLOGUE
    double *wk = (double *)calloc((size_t)n * 14 + 1, sizeof(double));
    double *V[14];
    for (int k = 0; k < 14; ++k)
        V[k] = wk + (size_t)k * n;
    double *r = V[0], *y = V[1], *z = V[2], *v = V[3], *w = V[4], *p = V[5], *q = V[6];
    double *d = V[7], *s = V[8], *v_tld = V[9], *w_tld = V[10], *y_tld = V[11],
           *z_tld = V[12], *p_tld = V[13];
    double rho = 0, rho_1 = 0, xi = 0, gamma = 1., gamma_1 = 0, theta = 0, theta_1 = 0,
           eta = -1., delta = 0, ep = 0, beta = 0;
    int conv = 0;
    v_copy(n, r, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    v_copy(n, v_tld, r);
    prec_apply(&M, v_tld, y);
    rho = or_norm2(n, y);
    v_copy(n, w_tld, r);
    prec_tapply(&M, w_tld, z);
    xi = or_norm2(n, z);
    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        if (rho == 0 || xi == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        v_set_scaled(n, v, 1 / rho, v_tld);
        v_scale(n, y, 1 / rho);
        v_set_scaled(n, w, 1 / xi, w_tld);
        v_scale(n, z, 1 / xi);
        delta = or_dot(n, z, y);
        if (delta == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        prec_apply(&M, y, y_tld);
        prec_tapply(&M, z, z_tld);
        if (it.iter == 0) {
            v_copy(n, p, y_tld);
            v_copy(n, q, z_tld);
        } else {
            v_scale_add(n, p, -xi * delta / ep, y_tld);
            v_scale_add(n, q, -rho * delta / ep, z_tld);
        }
        or_csr_mult(n, rowptr, colind, data, p, p_tld);
        ep = or_dot(n, q, p_tld);
        if (ep == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        beta = ep / delta;
        if (beta == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        v_set_scaled(n, v_tld, -beta, v);
        v_add(n, v_tld, p_tld);
        prec_apply(&M, v_tld, y);
        rho_1 = rho;
        rho = or_norm2(n, y);
        v_set_scaled(n, w_tld, -beta, w);
        or_csr_transmultadd(n, n, rowptr, colind, data, 1, q, w_tld);
        prec_tapply(&M, w_tld, z);
        xi = or_norm2(n, z);
        gamma_1 = gamma;
        theta_1 = theta;
        theta = rho / (gamma_1 * beta);
        gamma = 1 / sqrt(1 + theta * theta);
        if (gamma == 0) {
            st = OR_BREAKDOWN;
            break;
        }
        eta = -eta * rho_1 * gamma * gamma / (beta * gamma_1 * gamma_1);
        if (it.iter == 0) {
            v_set_scaled(n, d, eta, p);
            v_set_scaled(n, s, eta, p_tld);
        } else {
            const double val = theta_1 * theta_1 * gamma * gamma;
            v_scale(n, d, val);
            v_add_scaled(n, d, eta, p);
            v_scale(n, s, val);
            v_add_scaled(n, s, eta, p_tld);
        }
        v_add(n, x, d);
        v_add_scaled(n, r, -1, s);
    }
    free(wk);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}

/* ---- IR (IR.java:57-70) ------------------------------------------------- */
int or_ir(int n, const int *rowptr, const int *colind, const double *data,
          const double *b, double *x, const or_opts *opts, or_stats *stats,
          double *hist, int hist_cap)
{
    prec_t M;
    int bad = -1;
    int st = prec_setup(&M, n, rowptr, colind, data, opts, &bad);
    monitor_t it;
    mon_init(&it, opts, hist, hist_cap);
    if (st) {
        prec_free(&M);
        fill_stats(stats, &it, st);
        if (stats)
            stats->bad_row = bad;
        return st;
    }
    double *z = (double *)calloc(n, sizeof(double));
    double *r = (double *)calloc(n, sizeof(double));
    int conv = 0;
    v_copy(n, r, b);
    or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    for (it.iter = 0;; it.iter++) {
        st = mon_converged(&it, or_norm2(n, r), &conv);
        if (st || conv)
            break;
        prec_apply(&M, r, z);
        v_add(n, x, z);
        v_copy(n, r, b);
        or_csr_multadd(n, rowptr, colind, data, -1, x, r);
    }
    free(z);
    free(r);
    prec_free(&M);
    fill_stats(stats, &it, st);
    return st;
}
