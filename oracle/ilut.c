/*
 * Oracle (test infrastructure only): ILUT, the dual-threshold incomplete LU
 * of mtj/src/no/uib/cipr/matrix/sparse/ILUT.java, restated statement by
 * statement with MTJ's dense work row.  Compile with -O2 -ffp-contract=off.
 *
 * Each row is kept as MTJ keeps a SparseVector of the FlexCompRowMatrix:
 * ascending column indices with values; an insertion (SparseVector.set on an
 * absent index, SparseVector.java:157-163, 188-222) keeps the order, and
 * zero() (:235-238) clears the values but keeps the structure.
 *
 * Starting structure: ILUT.setMatrix runs LU.set(A) and LU.compact()
 * (ILUT.java:136-141); SparseVector.compact (:342-362) keeps the nonzero
 * entries.  Here every factorisation starts from A's nonzero entries of
 * each row, in column order.  (MTJ's compact only runs on rows whose arrays
 * have spare capacity, and then pads the row with (0, 0.0) entries after
 * its sorted part; rows without spare capacity keep explicit zeros, and a
 * FlexCompRowMatrix reused across setMatrix calls carries the previous
 * factor's structure as zeros.  Those capacity-dependent forms are not
 * restated: the factorisation here is a function of A's nonzero entries.)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int *index;
    double *data;
    int used, cap;
} svec;

/* SparseVector.getIndex (SparseVector.java:188-222): the position of ind,
 * inserted with value 0 if absent (rows stay sorted, so the insertion point
 * is the first index above ind). */
static int svec_pos(svec *v, int ind)
{
    int lo = 0, hi = v->used;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (v->index[mid] < ind)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < v->used && v->index[lo] == ind)
        return lo;
    if (v->used == v->cap) {
        int nc = v->cap ? 2 * v->cap : 4;
        v->index = (int *)realloc(v->index, sizeof(int) * nc);
        v->data = (double *)realloc(v->data, sizeof(double) * nc);
        v->cap = nc;
    }
    memmove(v->index + lo + 1, v->index + lo, sizeof(int) * (v->used - lo));
    memmove(v->data + lo + 1, v->data + lo, sizeof(double) * (v->used - lo));
    v->index[lo] = ind;
    v->data[lo] = 0.;
    v->used++;
    return lo;
}

static void svec_set(svec *v, int ind, double val)
{
    const int pos = svec_pos(v, ind);   /* may reallocate v->data */
    v->data[pos] = val;
}

typedef struct {
    int index;
    double value;
} entry_t;

/* IntDoubleEntry.compareTo (ILUT.java:289-297): descending |value|;
 * Collections.sort is stable and the lists are built in ascending index
 * order, so equal magnitudes keep ascending index. */
static int entry_cmp(const void *a, const void *b)
{
    const entry_t *x = (const entry_t *)a, *y = (const entry_t *)b;
    const double ax = fabs(x->value), ay = fabs(y->value);
    if (ax < ay)
        return 1;
    if (ax > ay)
        return -1;
    return (x->index > y->index) - (x->index < y->index);
}

/* ILUT.gather (ILUT.java:231-272). */
static void ilut_gather(const double *z, int n, svec *v, double taui, int d, int p,
                        entry_t *lower, entry_t *upper)
{
    int nl = 0, nu = 0;
    for (int e = 0; e < v->used; ++e) {
        if (v->index[e] < d)
            nl++;
        else if (v->index[e] > d)
            nu++;
    }
    for (int e = 0; e < v->used; ++e)
        v->data[e] = 0;
    int cl = 0, cu = 0;
    for (int i = 0; i < d; ++i)
        if (fabs(z[i]) > taui) {
            lower[cl].index = i;
            lower[cl++].value = z[i];
        }
    for (int i = d + 1; i < n; ++i)
        if (fabs(z[i]) > taui) {
            upper[cu].index = i;
            upper[cu++].value = z[i];
        }
    qsort(lower, (size_t)cl, sizeof(entry_t), entry_cmp);
    qsort(upper, (size_t)cu, sizeof(entry_t), entry_cmp);
    svec_set(v, d, z[d]);
    const int kl = nl + p < cl ? nl + p : cl;
    for (int i = 0; i < kl; ++i)
        svec_set(v, lower[i].index, lower[i].value);
    const int ku = nu + p < cu ? nu + p : cu;
    for (int i = 0; i < ku; ++i)
        svec_set(v, upper[i].index, upper[i].value);
}

/* ILUT.setMatrix + factor (ILUT.java:136-208).  Output: the factor in CSR
 * (orowptr[n+1], ocolind / oval of at most or_ilut_bound entries, odiag[n]
 * the diagonal positions).  bad_row: 1-based row of the error, as the
 * reference's messages ("Missing diagonal entry on row", "Zero diagonal
 * entry on row ... during ILU process"). */
int or_ilut_factor(int n, const int *rowptr, const int *colind, const double *val, double tau,
                   int p, int *orowptr, int *ocolind, double *oval, int *odiag, int *bad_row)
{
    svec *rows = (svec *)calloc((size_t)(n > 0 ? n : 1), sizeof(svec));
    int *diagind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int st = OR_OK;
    for (int r = 0; r < n; ++r) {
        svec *v = &rows[r];
        v->cap = rowptr[r + 1] - rowptr[r] + 4;
        v->index = (int *)malloc(sizeof(int) * v->cap);
        v->data = (double *)malloc(sizeof(double) * v->cap);
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j)
            if (val[j] != 0.) {
                v->index[v->used] = colind[j];
                v->data[v->used++] = val[j];
            }
    }
    /* diagonal indices (ILUT.java:148-155, findDiagonalIndex :210-213) */
    for (int k = 0; k < n; ++k) {
        diagind[k] = or_binary_search(rows[k].index, k, 0, rows[k].used);
        if (diagind[k] < 0) {
            if (bad_row)
                *bad_row = k + 1;
            st = OR_MISSING_DIAGONAL;
            goto done;
        }
    }
    {
        double *LUi = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        entry_t *lower = (entry_t *)malloc(sizeof(entry_t) * (size_t)(n > 0 ? n : 1));
        entry_t *upper = (entry_t *)malloc(sizeof(entry_t) * (size_t)(n > 0 ? n : 1));
        for (int i = 1; i < n && st == OR_OK; ++i) {
            svec *rowi = &rows[i];
            /* rowi.norm(Two) * tau: SparseVector.norm2 (SparseVector.java:278-283) */
            double nrm = 0;
            for (int e = 0; e < rowi->used; ++e)
                nrm += rowi->data[e] * rowi->data[e];
            const double taui = sqrt(nrm) * tau;
            /* scatter (ILUT.java:218-225) */
            for (int c = 0; c < n; ++c)
                LUi[c] = 0;
            for (int e = 0; e < rowi->used; ++e)
                LUi[rowi->index[e]] = rowi->data[e];
            for (int k = 0; k < i; ++k) {
                const svec *rowk = &rows[k];
                if (rowk->data[diagind[k]] == 0) {
                    if (bad_row)
                        *bad_row = k + 1;
                    st = OR_ZERO_PIVOT;
                    break;
                }
                const double LUik = LUi[k] / rowk->data[diagind[k]];
                if (fabs(LUik) <= taui)
                    continue;
                for (int j = diagind[k] + 1; j < rowk->used; ++j)
                    LUi[rowk->index[j]] -= LUik * rowk->data[j];
                LUi[k] = LUik;
            }
            if (st != OR_OK)
                break;
            ilut_gather(LUi, n, rowi, taui, i, p, lower, upper);
            if (diagind[i] >= rowi->used || rowi->index[diagind[i]] != i) {
                diagind[i] = or_binary_search(rowi->index, i, 0, rowi->used);
                if (diagind[i] < 0) {
                    if (bad_row)
                        *bad_row = i + 1;
                    st = OR_MISSING_DIAGONAL;
                }
            }
        }
        free(LUi);
        free(lower);
        free(upper);
    }
    if (st == OR_OK) {
        int pos = 0;
        orowptr[0] = 0;
        for (int r = 0; r < n; ++r) {
            for (int e = 0; e < rows[r].used; ++e) {
                ocolind[pos] = rows[r].index[e];
                oval[pos++] = rows[r].data[e];
            }
            orowptr[r + 1] = pos;
            odiag[r] = orowptr[r] + diagind[r];
        }
    }
done:
    for (int r = 0; r < n; ++r) {
        free(rows[r].index);
        free(rows[r].data);
    }
    free(rows);
    free(diagind);
    return st;
}

/* Upper bound on the factor's entries: per row A's entries, plus at most
 * nl + p kept lower and nu + p kept upper entries (ILUT.java:261-271), each
 * no more than the columns on that side. */
long long or_ilut_bound(int n, const int *rowptr, const int *colind, int p)
{
    long long t = 0;
    for (int r = 0; r < n; ++r) {
        long long nl = 0, nu = 0;
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            if (colind[j] < r)
                nl++;
            else if (colind[j] > r)
                nu++;
        }
        long long kl = nl + p, ku = nu + p;
        if (kl < 0)
            kl = 0;
        if (ku < 0)
            ku = 0;
        if (kl > r)
            kl = r;
        if (ku > n - 1 - r)
            ku = n - 1 - r;
        t += (rowptr[r + 1] - rowptr[r]) + kl + ku + 1;
    }
    return t;
}
