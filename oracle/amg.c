/*
 * CPU oracle: smoothed-aggregation AMG (TEST INFRASTRUCTURE ONLY).
 *
 * Restates mtj/src/no/uib/cipr/matrix/sparse/AMG.java (Vanek-Mandel-Brezina
 * smoothed aggregation, V/W cycle), SSOR.java (forward + reverse sweep that
 * starts from the output vector's old content) and the dense coarse solve
 * (DenseLU.java:83-197 -> LAPACK dgetf2/dgetrs, restated from the published
 * reference algorithms).
 *
 * Two aggregation modes:
 *   AMG_SEQUENTIAL (0): AMG.java:376-642 as written.  Phase 1 and phase 3
 *     are greedy in row order.  Phase 2 iterates each HashSet N(i); its tie
 *     rule ("first aggregate to reach the largest overlap") depends on the
 *     JVM's HashSet<Integer> iteration order, restated here as Java 8+
 *     HashMap order without tree bins (bucket (h ^ h>>>16) & (cap-1)
 *     ascending, insertion order inside a bucket, cap = 16 doubled while
 *     size > 0.75 cap).  Parity with a JVM run is unpinned for that order.
 *   AMG_PARALLEL (1): the data-parallel variant the GPU runs.  Phase 1
 *     roots form a maximal set with pairwise disjoint neighbourhoods chosen
 *     in rounds by hashed priority (a root wins when it holds the largest
 *     priority on every node of its N(i)); phase 2 is phase 2 above
 *     (already independent per node); phase 3 takes the remaining free
 *     nodes in the same prioritised rounds.  Aggregates are numbered by
 *     phase, round and root index.
 *
 * Smoother: SSOR on each level as in SSOR.java:143-194.  With block_rows > 0
 * each level's rows are split into contiguous blocks of about block_rows
 * rows; couplings that leave a block read the vector as it was when the
 * SSOR application started, couplings inside a block behave exactly as in
 * SSOR.java (lower: this sweep's values, upper: the previous values in the
 * forward sweep, this sweep's values in the reverse sweep).  Each row's sum
 * runs over its entries in column order.  One block per level reproduces
 * SSOR.java exactly.  ordering = 1 visits each block's rows by greedy colour
 * (multicolour SSOR: rows of one colour are independent, so a sweep is
 * as many parallel steps as there are colours); "earlier" then means
 * earlier in that order.
 *
 * The interpolation (AMG.java:810-876) and Galerkin operator
 * (createGalerkinSlow, :883-919) are restated with sparse accumulators whose
 * per-entry operation order equals the reference's dense loops: products
 * with a zero operand are skipped, which leaves every sum unchanged.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- Java 8 HashSet<Integer> iteration order -------------------------- */
static int hs_capacity(int size)
{
    int cap = 16;
    while (size > (cap * 3) / 4)
        cap <<= 1;
    return cap;
}

static unsigned hs_bucket(int v, int cap)
{
    unsigned h = (unsigned)v;
    return (h ^ (h >> 16)) & (unsigned)(cap - 1);
}

/* cols: insertion order (ascending column); out: iteration order */
void or_hashset_order(const int *cols, int cnt, int *out)
{
    const int cap = hs_capacity(cnt);
    for (int i = 0; i < cnt; ++i)
        out[i] = cols[i];
    /* stable insertion sort by bucket */
    for (int i = 1; i < cnt; ++i) {
        int v = out[i];
        unsigned bv = hs_bucket(v, cap);
        int j = i - 1;
        while (j >= 0 && hs_bucket(out[j], cap) > bv) {
            out[j + 1] = out[j];
            --j;
        }
        out[j + 1] = v;
    }
}

/* MurmurHash3 32-bit finaliser (Appleby, public domain description) */
static uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

uint64_t or_amg_priority(int i)
{
    return ((uint64_t)fmix32((uint32_t)i) << 32) | (uint32_t)i;
}

/* ---- strength: N(i) = {j : |a_ij| >= eps sqrt(a_ii a_jj)} (AMG.java:482-507) */
static int find_diag(int n, const int *rp, const int *ci, int *diag)
{
    for (int i = 0; i < n; ++i) {
        int d = or_binary_search(ci, i, rp[i], rp[i + 1]);
        if (d < 0)
            return i + 1;
        diag[i] = d;
    }
    return 0;
}

static void strength(int n, const int *rp, const int *ci, const double *v,
                     const int *diag, double eps, int *Np, int *Ni)
{
    int q = 0;
    Np[0] = 0;
    for (int i = 0; i < n; ++i) {
        double aii = v[diag[i]];
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            double aij = v[j];
            double ajj = v[diag[ci[j]]];
            if (fabs(aij) >= eps * sqrt(aii * ajj))
                Ni[q++] = ci[j];
        }
        Np[i + 1] = q;
    }
}

/* createInitialR (AMG.java:512-532) */
static void initial_R(int n, const int *rp, const int *ci, const double *v,
                      char *R)
{
    for (int i = 0; i < n; ++i) {
        char h = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j)
            if (ci[j] != i && v[j] != 0) {
                h = 1;
                break;
            }
        R[i] = h;
    }
}

/* ---- aggregation ------------------------------------------------------- */
/* pt[i] = aggregate of node i or -1.  Returns the aggregate count. */
static int aggregate(int n, const int *Np, const int *Ni, char *R, int mode,
                     int *pt, int *nroots_out)
{
    int c = 0;
    for (int i = 0; i < n; ++i)
        pt[i] = -1;
    int *owner = (int *)malloc(sizeof(int) * (n > 0 ? n : 1)); /* initial aggregate of j */
    for (int i = 0; i < n; ++i)
        owner[i] = -1;

    /* phase 1: initial aggregates N(i), all free (AMG.java:537-562) */
    if (mode == 0) {
        for (int i = 0; i < n; ++i) {
            if (!R[i])
                continue;
            int fr = 1;
            for (int q = Np[i]; q < Np[i + 1]; ++q)
                fr &= R[Ni[q]];
            if (fr) {
                for (int q = Np[i]; q < Np[i + 1]; ++q) {
                    R[Ni[q]] = 0;
                    owner[Ni[q]] = c;
                }
                ++c;
            }
        }
    } else {
        char *cand = (char *)malloc(n > 0 ? n : 1);
        char *sel = (char *)malloc(n > 0 ? n : 1);
        uint64_t *m = (uint64_t *)malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
        int *root_of = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
        for (int i = 0; i < n; ++i)
            root_of[i] = -1;
        for (int round = 0; round < 10000; ++round) {
            int any = 0;
            for (int i = 0; i < n; ++i) {
                char ok = R[i];
                for (int q = Np[i]; q < Np[i + 1] && ok; ++q)
                    ok &= R[Ni[q]];
                cand[i] = ok;
                any |= ok;
                m[i] = 0;
            }
            if (!any)
                break;
            for (int i = 0; i < n; ++i)
                if (cand[i]) {
                    uint64_t p = or_amg_priority(i);
                    for (int q = Np[i]; q < Np[i + 1]; ++q)
                        if (m[Ni[q]] < p)
                            m[Ni[q]] = p;
                }
            for (int i = 0; i < n; ++i) {
                sel[i] = 0;
                if (!cand[i])
                    continue;
                uint64_t p = or_amg_priority(i);
                char s = 1;
                for (int q = Np[i]; q < Np[i + 1]; ++q)
                    s &= (m[Ni[q]] == p);
                sel[i] = s;
            }
            for (int i = 0; i < n; ++i)
                if (sel[i])
                    for (int q = Np[i]; q < Np[i + 1]; ++q) {
                        R[Ni[q]] = 0;
                        root_of[Ni[q]] = i;
                    }
        }
        /* number the initial aggregates by root index */
        int *id = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
        for (int i = 0; i < n; ++i)
            id[i] = -1;
        for (int i = 0; i < n; ++i)
            if (root_of[i] == i)
                id[i] = c++;
        for (int j = 0; j < n; ++j)
            if (root_of[j] >= 0)
                owner[j] = id[root_of[j]];
        free(id);
        free(cand);
        free(sel);
        free(m);
        free(root_of);
    }
    const int nroots = c;
    for (int j = 0; j < n; ++j)
        pt[j] = owner[j];

    /* phase 2: enlargement (AMG.java:567-614), independent per node */
    {
        int *cnt = (int *)calloc(nroots > 0 ? nroots : 1, sizeof(int));
        int *ord = (int *)malloc(sizeof(int) * 1024);
        int ordcap = 1024;
        char *freed = (char *)calloc(n > 0 ? n : 1, 1);
        int *target = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) {
            target[i] = -1;
            if (!R[i])
                continue;
            int len = Np[i + 1] - Np[i];
            if (len > ordcap) {
                ordcap = len;
                ord = (int *)realloc(ord, sizeof(int) * ordcap);
            }
            or_hashset_order(Ni + Np[i], len, ord);
            int largest = 0, maxv = 0;
            for (int q = 0; q < len; ++q) {
                int k = owner[ord[q]];
                if (k < 0)
                    continue;
                cnt[k]++;
                if (cnt[k] > maxv) {
                    largest = k;
                    maxv = cnt[k];
                }
            }
            for (int q = 0; q < len; ++q)
                if (owner[ord[q]] >= 0)
                    cnt[owner[ord[q]]] = 0;
            if (maxv > 0) {
                freed[i] = 1;
                target[i] = largest;
            }
        }
        for (int i = 0; i < n; ++i)
            if (freed[i]) {
                R[i] = 0;
                pt[i] = target[i];
            }
        free(cnt);
        free(ord);
        free(freed);
        free(target);
    }

    /* phase 3: remnants (AMG.java:619-641) */
    if (mode == 0) {
        for (int i = 0; i < n; ++i) {
            if (!R[i])
                continue;
            int got = 0;
            for (int q = Np[i]; q < Np[i + 1]; ++q)
                if (R[Ni[q]]) {
                    R[Ni[q]] = 0;
                    pt[Ni[q]] = c;
                    got = 1;
                }
            if (got)
                ++c;
        }
    } else {
        char *cand = (char *)malloc(n > 0 ? n : 1);
        char *sel = (char *)malloc(n > 0 ? n : 1);
        uint64_t *m = (uint64_t *)malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
        for (int round = 0; round < 10000; ++round) {
            int any = 0;
            for (int i = 0; i < n; ++i) {
                char ok = 0;
                if (R[i])
                    for (int q = Np[i]; q < Np[i + 1]; ++q)
                        ok |= R[Ni[q]];
                cand[i] = ok;
                any |= ok;
                m[i] = 0;
            }
            if (!any)
                break;
            for (int i = 0; i < n; ++i)
                if (cand[i]) {
                    uint64_t p = or_amg_priority(i);
                    for (int q = Np[i]; q < Np[i + 1]; ++q)
                        if (R[Ni[q]] && m[Ni[q]] < p)
                            m[Ni[q]] = p;
                }
            for (int i = 0; i < n; ++i) {
                sel[i] = 0;
                if (!cand[i])
                    continue;
                uint64_t p = or_amg_priority(i);
                char s = 1;
                for (int q = Np[i]; q < Np[i + 1]; ++q)
                    if (R[Ni[q]])
                        s &= (m[Ni[q]] == p);
                sel[i] = s;
            }
            for (int i = 0; i < n; ++i)
                if (sel[i]) {
                    for (int q = Np[i]; q < Np[i + 1]; ++q)
                        if (R[Ni[q]])
                            pt[Ni[q]] = c;
                    ++c;
                }
            for (int j = 0; j < n; ++j)
                if (R[j] && pt[j] >= 0)
                    R[j] = 0;
        }
        free(cand);
        free(sel);
        free(m);
    }
    free(owner);
    if (nroots_out)
        *nroots_out = nroots;
    return c;
}

int or_amg_aggregate(int n, const int *rp, const int *ci, const double *v,
                     double eps, int mode, int *pt, int *nagg, int *nroots)
{
    int *diag = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (find_diag(n, rp, ci, diag)) {
        free(diag);
        return OR_MISSING_DIAGONAL;
    }
    int nnz = rp[n];
    int *Np = (int *)malloc(sizeof(int) * (n + 1));
    int *Ni = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    char *R = (char *)malloc(n > 0 ? n : 1);
    strength(n, rp, ci, v, diag, eps, Np, Ni);
    initial_R(n, rp, ci, v, R);
    *nagg = aggregate(n, Np, Ni, R, mode, pt, nroots);
    free(diag);
    free(Np);
    free(Ni);
    free(R);
    return OR_OK;
}

/* ---- smoothed prolongator (AMG.java:810-876), row form ------------------ */
/* Row i of P = (I - omega D^-1 A_F) P_tent; stored CSR with ascending
 * aggregate columns.  P arrays are allocated here. */
static void prolongator(int n, const int *rp, const int *ci, const double *v,
                        const int *diag, const int *Np, const int *Ni,
                        const int *pt, int c, double omega, int **Prp_o,
                        int **Pci_o, double **Pv_o)
{
    int cap = n * 8 + 16, q = 0;
    int *Prp = (int *)malloc(sizeof(int) * (n + 1));
    int *Pci = (int *)malloc(sizeof(int) * cap);
    double *Pv = (double *)malloc(sizeof(double) * cap);
    double *acc = (double *)calloc(c > 0 ? c : 1, sizeof(double));
    int *touched = (int *)malloc(sizeof(int) * 4096);
    int tcap = 4096;
    char *mark = (char *)calloc(c > 0 ? c : 1, 1);
    Prp[0] = 0;
    for (int i = 0; i < n; ++i) {
        if (pt[i] == -1) {
            Prp[i + 1] = q;
            continue;
        }
        int nt = 0;
        double weak = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            int col = ci[j];
            if (pt[col] == -1)
                continue;
            double aij = v[j];
            /* Ni.contains(col): N rows are ascending */
            int strong = or_binary_search(Ni, col, Np[i], Np[i + 1]) >= 0;
            if (aij != 0 && !strong) {
                weak += aij;
                continue;
            }
            int k = pt[col];
            if (!mark[k]) {
                mark[k] = 1;
                acc[k] = 0;
                if (nt == tcap) {
                    tcap *= 2;
                    touched = (int *)realloc(touched, sizeof(int) * tcap);
                }
                touched[nt++] = k;
            }
            acc[k] += aij;
        }
        const int own = pt[i];
        if (!mark[own]) {
            mark[own] = 1;
            acc[own] = 0;
            touched[nt++] = own;
        }
        acc[own] -= weak;
        const double scale = -omega / v[diag[i]];
        for (int t = 0; t < nt; ++t)
            acc[touched[t]] *= scale;
        acc[own] += 1;
        /* ascending aggregate order */
        for (int a = 1; a < nt; ++a) {
            int key = touched[a], b = a - 1;
            while (b >= 0 && touched[b] > key) {
                touched[b + 1] = touched[b];
                --b;
            }
            touched[b + 1] = key;
        }
        if (q + nt > cap) {
            cap = 2 * cap + nt;
            Pci = (int *)realloc(Pci, sizeof(int) * cap);
            Pv = (double *)realloc(Pv, sizeof(double) * cap);
        }
        for (int t = 0; t < nt; ++t) {
            int k = touched[t];
            if (acc[k] != 0) {
                Pci[q] = k;
                Pv[q] = acc[k];
                ++q;
            }
            mark[k] = 0;
        }
        Prp[i + 1] = q;
    }
    free(acc);
    free(touched);
    free(mark);
    *Prp_o = Prp;
    *Pci_o = Pci;
    *Pv_o = Pv;
}

/* CSR (n x c) -> CSC (the CompColMatrix I), rows ascending per column */
static void to_csc(int n, int c, const int *rp, const int *ci, const double *v,
                   int **cp_o, int **ri_o, double **cv_o)
{
    int nnz = rp[n];
    int *cp = (int *)calloc(c + 1, sizeof(int));
    int *ri = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    double *cv = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    for (int j = 0; j < nnz; ++j)
        cp[ci[j] + 1]++;
    for (int k = 0; k < c; ++k)
        cp[k + 1] += cp[k];
    int *pos = (int *)malloc(sizeof(int) * (c > 0 ? c : 1));
    for (int k = 0; k < c; ++k)
        pos[k] = cp[k];
    for (int i = 0; i < n; ++i)
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            int p = pos[ci[j]]++;
            ri[p] = i;
            cv[p] = v[j];
        }
    free(pos);
    *cp_o = cp;
    *ri_o = ri;
    *cv_o = cv;
}

/* Galerkin Ac = P' A P with createGalerkinSlow's per-entry order:
 * AP[r,k] = sum_{j in row r asc} A[r,j] P[j,k];
 * Ac[i,k] = sum_{r in column i of P asc} P[r,i] AP[r,k]; zeros not stored. */
static void galerkin(int n, const int *rp, const int *ci, const double *v, int c,
                     const int *Prp, const int *Pci, const double *Pv,
                     const int *Icp, const int *Iri, const double *Icv,
                     int **Arp_o, int **Aci_o, double **Av_o)
{
    /* AP in CSR */
    int cap = n * 16 + 16, q = 0;
    int *APrp = (int *)malloc(sizeof(int) * (n + 1));
    int *APci = (int *)malloc(sizeof(int) * cap);
    double *APv = (double *)malloc(sizeof(double) * cap);
    double *acc = (double *)calloc(c > 0 ? c : 1, sizeof(double));
    char *mark = (char *)calloc(c > 0 ? c : 1, 1);
    int tcap = 4096;
    int *touched = (int *)malloc(sizeof(int) * tcap);
    APrp[0] = 0;
    for (int r = 0; r < n; ++r) {
        int nt = 0;
        for (int j = rp[r]; j < rp[r + 1]; ++j) {
            int col = ci[j];
            for (int t = Prp[col]; t < Prp[col + 1]; ++t) {
                int k = Pci[t];
                if (!mark[k]) {
                    mark[k] = 1;
                    acc[k] = 0;
                    if (nt == tcap) {
                        tcap *= 2;
                        touched = (int *)realloc(touched, sizeof(int) * tcap);
                    }
                    touched[nt++] = k;
                }
                acc[k] += v[j] * Pv[t];
            }
        }
        for (int a = 1; a < nt; ++a) {
            int key = touched[a], b = a - 1;
            while (b >= 0 && touched[b] > key) {
                touched[b + 1] = touched[b];
                --b;
            }
            touched[b + 1] = key;
        }
        if (q + nt > cap) {
            cap = 2 * cap + nt;
            APci = (int *)realloc(APci, sizeof(int) * cap);
            APv = (double *)realloc(APv, sizeof(double) * cap);
        }
        for (int t = 0; t < nt; ++t) {
            int k = touched[t];
            /* CompRowMatrix.mult stores a row's dot only when nonzero */
            if (acc[k] != 0) {
                APci[q] = k;
                APv[q] = acc[k];
                ++q;
            }
            mark[k] = 0;
        }
        APrp[r + 1] = q;
    }
    /* Ac rows */
    int ccap = c * 32 + 16, cq = 0;
    int *Arp = (int *)malloc(sizeof(int) * (c + 1));
    int *Aci = (int *)malloc(sizeof(int) * ccap);
    double *Av = (double *)malloc(sizeof(double) * ccap);
    Arp[0] = 0;
    for (int i = 0; i < c; ++i) {
        int nt = 0;
        for (int t = Icp[i]; t < Icp[i + 1]; ++t) {
            int r = Iri[t];
            double pri = Icv[t];
            for (int s = APrp[r]; s < APrp[r + 1]; ++s) {
                int k = APci[s];
                if (!mark[k]) {
                    mark[k] = 1;
                    acc[k] = 0;
                    if (nt == tcap) {
                        tcap *= 2;
                        touched = (int *)realloc(touched, sizeof(int) * tcap);
                    }
                    touched[nt++] = k;
                }
                acc[k] += pri * APv[s];
            }
        }
        for (int a = 1; a < nt; ++a) {
            int key = touched[a], b = a - 1;
            while (b >= 0 && touched[b] > key) {
                touched[b + 1] = touched[b];
                --b;
            }
            touched[b + 1] = key;
        }
        if (cq + nt > ccap) {
            ccap = 2 * ccap + nt;
            Aci = (int *)realloc(Aci, sizeof(int) * ccap);
            Av = (double *)realloc(Av, sizeof(double) * ccap);
        }
        for (int t = 0; t < nt; ++t) {
            int k = touched[t];
            if (acc[k] != 0) {
                Aci[cq] = k;
                Av[cq] = acc[k];
                ++cq;
            }
            mark[k] = 0;
        }
        Arp[i + 1] = cq;
    }
    free(APrp);
    free(APci);
    free(APv);
    free(acc);
    free(mark);
    free(touched);
    *Arp_o = Arp;
    *Aci_o = Aci;
    *Av_o = Av;
}

/* ---- hierarchy ---------------------------------------------------------- */
typedef struct {
    int n, c;
    int *rp, *ci, *diag;
    double *v;
    int *Prp, *Pci;      /* P (n x c) CSR */
    double *Pv;
    int *Icp, *Iri;      /* I = P as CompColMatrix */
    double *Icv;
    int *pt;
    int nb;
    int *bstart;
    int *blk;            /* block of each row */
    int *order;          /* rows in sweep order, block by block */
    int *rank;           /* position of each row in its block's order */
} amg_level;

struct or_amg {
    or_amg_params p;
    int m;               /* levels; m-1 smoothed levels + dense coarsest */
    amg_level *L;
    int nc;
    double *lu;          /* column-major nc x nc */
    int *ipiv;
    double **u, **f, **r;
    double *xx, *xold;
    int transpose;
};

void or_amg_params_default(or_amg_params *p)
{
    p->omega_pre_f = 1;
    p->omega_pre_r = 1.85;
    p->omega_post_f = 1.85;
    p->omega_post_r = 1;
    p->nu1 = 1;
    p->nu2 = 1;
    p->gamma = 1;
    p->min = 40;
    p->omega = 2. / 3;
    p->reverse = 1;
    p->mode = 0;
    p->block_rows = 0;
    p->max_coarse = 4096;
    p->zero_guess = 0;
    p->ordering = 0;
}

/* Greedy colouring in row order of the symmetrised in-block graph: a row
 * takes the smallest colour none of its earlier in-block neighbours has. */
static void greedy_colour(const amg_level *L, int *colour)
{
    const int n = L->n;
    const int *rp = L->rp, *ci = L->ci, *blk = L->blk;
    int nnz = rp[n];
    /* transpose pattern for the symmetrisation */
    int *tp = (int *)calloc(n + 1, sizeof(int));
    int *ti = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    for (int j = 0; j < nnz; ++j)
        tp[ci[j] + 1]++;
    for (int i = 0; i < n; ++i)
        tp[i + 1] += tp[i];
    int *pos = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i)
        pos[i] = tp[i];
    for (int i = 0; i < n; ++i)
        for (int j = rp[i]; j < rp[i + 1]; ++j)
            ti[pos[ci[j]]++] = i;
    int *mark = (int *)malloc(sizeof(int) * (n + 2));
    for (int i = 0; i < n + 2; ++i)
        mark[i] = -1;
    for (int i = 0; i < n; ++i) {
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            int c = ci[j];
            if (c < i && blk[c] == blk[i])
                mark[colour[c]] = i;
        }
        for (int j = tp[i]; j < tp[i + 1]; ++j) {
            int c = ti[j];
            if (c < i && blk[c] == blk[i])
                mark[colour[c]] = i;
        }
        int col = 0;
        while (mark[col] == i)
            ++col;
        colour[i] = col;
    }
    free(tp);
    free(ti);
    free(pos);
    free(mark);
}

static void level_blocks(amg_level *L, int block_rows, int ordering)
{
    int n = L->n;
    int nb = 1;
    if (block_rows > 0 && n > 0)
        nb = (n + block_rows - 1) / block_rows;
    if (nb < 1)
        nb = 1;
    L->nb = nb;
    L->bstart = (int *)malloc(sizeof(int) * (nb + 1));
    for (int b = 0; b <= nb; ++b)
        L->bstart[b] = (int)(((long long)n * b) / nb);
    L->blk = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int b = 0; b < nb; ++b)
        for (int r = L->bstart[b]; r < L->bstart[b + 1]; ++r)
            L->blk[r] = b;
    L->order = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    L->rank = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (ordering == 0) {
        for (int r = 0; r < n; ++r)
            L->order[r] = r;
    } else {
        /* rows of each block by (colour, row) */
        int *colour = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
        greedy_colour(L, colour);
        int q = 0;
        for (int b = 0; b < nb; ++b) {
            int maxc = -1;
            for (int r = L->bstart[b]; r < L->bstart[b + 1]; ++r)
                if (colour[r] > maxc)
                    maxc = colour[r];
            for (int c = 0; c <= maxc; ++c)
                for (int r = L->bstart[b]; r < L->bstart[b + 1]; ++r)
                    if (colour[r] == c)
                        L->order[q++] = r;
        }
        free(colour);
    }
    for (int b = 0; b < nb; ++b)
        for (int t = L->bstart[b]; t < L->bstart[b + 1]; ++t)
            L->rank[L->order[t]] = t - L->bstart[b];
}

/* LAPACK dgetf2 (partial pivoting, right-looking) on column-major a */
static int dgetf2(int n, double *a, int *ipiv)
{
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = fabs(a[j + j * n]);
        for (int i = j + 1; i < n; ++i)
            if (fabs(a[i + j * n]) > amax) {
                amax = fabs(a[i + j * n]);
                jp = i;
            }
        ipiv[j] = jp;
        if (a[jp + j * n] != 0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) {
                    double t = a[j + c * n];
                    a[j + c * n] = a[jp + c * n];
                    a[jp + c * n] = t;
                }
            if (j < n - 1) {
                double rr = 1.0 / a[j + j * n];
                for (int i = j + 1; i < n; ++i)
                    a[i + j * n] *= rr;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < n - 1)
            for (int c = j + 1; c < n; ++c) {
                double yj = a[j + c * n];
                if (yj != 0) {
                    double temp = -1.0 * yj;
                    for (int i = j + 1; i < n; ++i)
                        a[i + c * n] = a[i + c * n] + a[i + j * n] * temp;
                }
            }
    }
    return info;
}

/* dgetrs 'N' / 'T' with one right-hand side */
static void dgetrs(int n, const double *a, const int *ipiv, double *b, int trans)
{
    if (!trans) {
        for (int j = 0; j < n; ++j)
            if (ipiv[j] != j) {
                double t = b[j];
                b[j] = b[ipiv[j]];
                b[ipiv[j]] = t;
            }
        for (int k = 0; k < n; ++k)
            if (b[k] != 0)
                for (int i = k + 1; i < n; ++i)
                    b[i] = b[i] - b[k] * a[i + k * n];
        for (int k = n - 1; k >= 0; --k)
            if (b[k] != 0) {
                b[k] = b[k] / a[k + k * n];
                for (int i = 0; i < k; ++i)
                    b[i] = b[i] - b[k] * a[i + k * n];
            }
    } else {
        /* U' (upper, transposed, non-unit), then L' (unit), then swaps back */
        for (int i = 0; i < n; ++i) {
            double t = b[i];
            for (int k = 0; k < i; ++k)
                t -= a[k + i * n] * b[k];
            b[i] = t / a[i + i * n];
        }
        for (int i = n - 1; i >= 0; --i) {
            double t = b[i];
            for (int k = i + 1; k < n; ++k)
                t -= a[k + i * n] * b[k];
            b[i] = t;
        }
        for (int j = n - 1; j >= 0; --j)
            if (ipiv[j] != j) {
                double t = b[j];
                b[j] = b[ipiv[j]];
                b[ipiv[j]] = t;
            }
    }
}

void or_amg_free(or_amg *A)
{
    if (!A)
        return;
    for (int k = 0; k < A->m; ++k) {
        amg_level *L = &A->L[k];
        free(L->rp);
        free(L->ci);
        free(L->v);
        free(L->diag);
        free(L->Prp);
        free(L->Pci);
        free(L->Pv);
        free(L->Icp);
        free(L->Iri);
        free(L->Icv);
        free(L->pt);
        free(L->bstart);
        free(L->blk);
        free(L->order);
        free(L->rank);
        free(A->u[k]);
        free(A->f[k]);
        free(A->r[k]);
    }
    free(A->L);
    free(A->u);
    free(A->f);
    free(A->r);
    free(A->lu);
    free(A->ipiv);
    free(A->xx);
    free(A->xold);
    free(A);
}

/* AMG.setMatrix (AMG.java:228-294) */
int or_amg_setup(int n, const int *rp, const int *ci, const double *v,
                 const or_amg_params *prm, or_amg **out, int *bad_row)
{
    or_amg *A = (or_amg *)calloc(1, sizeof(or_amg));
    A->p = *prm;
    int capL = 32;
    A->L = (amg_level *)calloc(capL, sizeof(amg_level));
    /* level 0 = copy of the input */
    amg_level *L0 = &A->L[0];
    int nnz = rp[n];
    L0->n = n;
    L0->rp = (int *)malloc(sizeof(int) * (n + 1));
    L0->ci = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    L0->v = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    memcpy(L0->rp, rp, sizeof(int) * (n + 1));
    memcpy(L0->ci, ci, sizeof(int) * nnz);
    memcpy(L0->v, v, sizeof(double) * nnz);
    int m = 1;
    for (int k = 0; A->L[k].n > prm->min; ++k) {
        amg_level *L = &A->L[k];
        int nk = L->n, nnzk = L->rp[nk];
        L->diag = (int *)malloc(sizeof(int) * (nk > 0 ? nk : 1));
        int bad = find_diag(nk, L->rp, L->ci, L->diag);
        if (bad) {
            if (bad_row)
                *bad_row = bad;
            A->m = m;
            or_amg_free(A);
            return OR_MISSING_DIAGONAL;
        }
        double eps = 0.08 * pow(0.5, k);
        int *Np = (int *)malloc(sizeof(int) * (nk + 1));
        int *Ni = (int *)malloc(sizeof(int) * (nnzk > 0 ? nnzk : 1));
        char *R = (char *)malloc(nk > 0 ? nk : 1);
        strength(nk, L->rp, L->ci, L->v, L->diag, eps, Np, Ni);
        initial_R(nk, L->rp, L->ci, L->v, R);
        L->pt = (int *)malloc(sizeof(int) * (nk > 0 ? nk : 1));
        int c = aggregate(nk, Np, Ni, R, prm->mode, L->pt, NULL);
        free(R);
        if (c == 0) {
            free(Np);
            free(Ni);
            break;
        }
        L->c = c;
        prolongator(nk, L->rp, L->ci, L->v, L->diag, Np, Ni, L->pt, c, prm->omega,
                    &L->Prp, &L->Pci, &L->Pv);
        free(Np);
        free(Ni);
        to_csc(nk, c, L->Prp, L->Pci, L->Pv, &L->Icp, &L->Iri, &L->Icv);
        if (m == capL) {
            capL *= 2;
            A->L = (amg_level *)realloc(A->L, sizeof(amg_level) * capL);
            memset(A->L + m, 0, sizeof(amg_level) * (capL - m));
            L = &A->L[k];
        }
        amg_level *Lc = &A->L[k + 1];
        memset(Lc, 0, sizeof(*Lc));
        Lc->n = c;
        galerkin(nk, L->rp, L->ci, L->v, c, L->Prp, L->Pci, L->Pv, L->Icp, L->Iri,
                 L->Icv, &Lc->rp, &Lc->ci, &Lc->v);
        m++;
    }
    A->m = m;
    /* coarsest: dense LU (AMG.java:268-271) */
    amg_level *Lc = &A->L[m - 1];
    A->nc = Lc->n;
    if (A->nc > prm->max_coarse) {
        or_amg_free(A);
        return OR_BAD_ARGUMENT;
    }
    A->lu = (double *)calloc((size_t)A->nc * A->nc + 1, sizeof(double));
    A->ipiv = (int *)malloc(sizeof(int) * (A->nc + 1));
    for (int i = 0; i < A->nc; ++i)
        for (int j = Lc->rp[i]; j < Lc->rp[i + 1]; ++j)
            A->lu[i + (size_t)Lc->ci[j] * A->nc] = Lc->v[j];
    if (dgetf2(A->nc, A->lu, A->ipiv)) {
        or_amg_free(A);
        return OR_SINGULAR;
    }
    A->u = (double **)calloc(m, sizeof(double *));
    A->f = (double **)calloc(m, sizeof(double *));
    A->r = (double **)calloc(m, sizeof(double *));
    int nmax = 1;
    for (int k = 0; k < m; ++k) {
        int nk = A->L[k].n;
        A->u[k] = (double *)calloc(nk + 1, sizeof(double));
        A->f[k] = (double *)calloc(nk + 1, sizeof(double));
        A->r[k] = (double *)calloc(nk + 1, sizeof(double));
        if (nk > nmax)
            nmax = nk;
        if (k < m - 1)
            level_blocks(&A->L[k], prm->block_rows, prm->ordering);
    }
    A->xx = (double *)calloc(nmax, sizeof(double));
    A->xold = (double *)calloc(nmax, sizeof(double));
    *out = A;
    return OR_OK;
}

/* SSOR.apply (SSOR.java:143-194), block form described in the header.
 * Rows are visited in L->order (natural order: SSOR.java); an in-block
 * coupling reads this sweep's value when its row was visited earlier in
 * the sweep, the previous value otherwise. */
static void ssor(const amg_level *L, double omegaF, double omegaR, int reverse,
                 const double *bd, double *xd, double *xx, double *xold)
{
    const int n = L->n;
    const int *rp = L->rp, *ci = L->ci, *dg = L->diag, *blk = L->blk;
    const int *order = L->order, *rank = L->rank;
    const double *data = L->v;
    memcpy(xx, xd, sizeof(double) * n);
    memcpy(xold, xd, sizeof(double) * n);
    for (int t = 0; t < n; ++t) {
        const int i = order[t];
        const int bi = blk[i], ri = rank[i];
        double sigma = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            if (j == dg[i])
                continue;
            int c = ci[j];
            double xc = blk[c] != bi ? xold[c] : (rank[c] < ri ? xx[c] : xd[c]);
            sigma += data[j] * xc;
        }
        sigma = (bd[i] - sigma) / data[dg[i]];
        xx[i] = xd[i] + omegaF * (sigma - xd[i]);
    }
    if (!reverse) {
        memcpy(xd, xx, sizeof(double) * n);
        return;
    }
    for (int t = n - 1; t >= 0; --t) {
        const int i = order[t];
        const int bi = blk[i], ri = rank[i];
        double sigma = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            if (j == dg[i])
                continue;
            int c = ci[j];
            double xc = blk[c] != bi ? xold[c] : (rank[c] > ri ? xd[c] : xx[c]);
            sigma += data[j] * xc;
        }
        sigma = (bd[i] - sigma) / data[dg[i]];
        xd[i] = xx[i] + omegaR * (sigma - xx[i]);
    }
}

/* AMG.cycle (AMG.java:302-328) */
static void cycle(or_amg *A, int k)
{
    if (k == A->m - 1) {
        /* directSolve (AMG.java:333-342) */
        memcpy(A->u[k], A->f[k], sizeof(double) * A->nc);
        dgetrs(A->nc, A->lu, A->ipiv, A->u[k], A->transpose);
        return;
    }
    amg_level *L = &A->L[k];
    const or_amg_params *p = &A->p;
    for (int i = 0; i < p->nu1; ++i)
        ssor(L, p->omega_pre_f, p->omega_pre_r, p->reverse, A->f[k], A->u[k], A->xx,
             A->xold);
    int c = A->L[k + 1].n;
    memset(A->u[k + 1], 0, sizeof(double) * c);
    memcpy(A->r[k], A->f[k], sizeof(double) * L->n);
    or_csr_multadd(L->n, L->rp, L->ci, L->v, -1, A->u[k], A->r[k]);
    /* I.transMult(r, f_c) (CompColMatrix.java:302-318) */
    for (int i = 0; i < c; ++i) {
        double dot = 0;
        for (int t = L->Icp[i]; t < L->Icp[i + 1]; ++t)
            dot += L->Icv[t] * A->r[k][L->Iri[t]];
        A->f[k + 1][i] = dot;
    }
    for (int g = 0; g < p->gamma; ++g)
        cycle(A, k + 1);
    /* I.multAdd(u_c, u) (CompColMatrix.java:277-299): scatter by column,
     * y scaled by 1/1 and 1 (exact) */
    for (int i = 0; i < c; ++i)
        for (int t = L->Icp[i]; t < L->Icp[i + 1]; ++t)
            A->u[k][L->Iri[t]] += L->Icv[t] * A->u[k + 1][i];
    for (int i = 0; i < p->nu2; ++i)
        ssor(L, p->omega_post_f, p->omega_post_r, p->reverse, A->f[k], A->u[k], A->xx,
             A->xold);
}

/* AMG.apply / transApply (AMG.java:208-226): start from x's content */
void or_amg_apply(or_amg *A, const double *b, double *x, int transpose)
{
    int n = A->L[0].n;
    if (A->p.zero_guess)
        memset(A->u[0], 0, sizeof(double) * n);
    else
        memcpy(A->u[0], x, sizeof(double) * n);
    memcpy(A->f[0], b, sizeof(double) * n);
    A->transpose = transpose;
    cycle(A, 0);
    memcpy(x, A->u[0], sizeof(double) * n);
}

int or_amg_levels(const or_amg *A, int *sizes, int *nnzs, int cap)
{
    for (int k = 0; k < A->m && k < cap; ++k) {
        sizes[k] = A->L[k].n;
        nnzs[k] = A->L[k].rp[A->L[k].n];
    }
    return A->m;
}

/* Copies level k's matrix (CSR) and, for k < m-1, its aggregates and P
 * (CSR).  Any output pointer may be NULL. */
int or_amg_get_level(const or_amg *A, int k, int *rp, int *ci, double *v, int *pt,
                     int *Prp, int *Pci, double *Pv)
{
    if (k < 0 || k >= A->m)
        return OR_BAD_ARGUMENT;
    const amg_level *L = &A->L[k];
    int n = L->n, nnz = L->rp[n];
    if (rp)
        memcpy(rp, L->rp, sizeof(int) * (n + 1));
    if (ci)
        memcpy(ci, L->ci, sizeof(int) * nnz);
    if (v)
        memcpy(v, L->v, sizeof(double) * nnz);
    if (k < A->m - 1) {
        if (pt)
            memcpy(pt, L->pt, sizeof(int) * n);
        if (Prp)
            memcpy(Prp, L->Prp, sizeof(int) * (n + 1));
        if (Pci)
            memcpy(Pci, L->Pci, sizeof(int) * L->Prp[n]);
        if (Pv)
            memcpy(Pv, L->Pv, sizeof(double) * L->Prp[n]);
    }
    return OR_OK;
}

int or_amg_p_nnz(const or_amg *A, int k)
{
    if (k < 0 || k >= A->m - 1)
        return 0;
    return A->L[k].Prp[A->L[k].n];
}

/* One SSOR application on a CSR matrix (test hook). */
int or_ssor_apply(int n, const int *rp, const int *ci, const double *v,
                  int block_rows, int ordering, double omegaF, double omegaR,
                  int reverse, const double *b, double *x, int *order_out)
{
    amg_level L;
    memset(&L, 0, sizeof(L));
    L.n = n;
    L.rp = (int *)rp;
    L.ci = (int *)ci;
    L.v = (double *)v;
    L.diag = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (find_diag(n, rp, ci, L.diag)) {
        free(L.diag);
        return OR_MISSING_DIAGONAL;
    }
    level_blocks(&L, block_rows, ordering);
    if (order_out)
        memcpy(order_out, L.order, sizeof(int) * n);
    double *xx = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *xo = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    ssor(&L, omegaF, omegaR, reverse, b, x, xx, xo);
    free(xx);
    free(xo);
    free(L.diag);
    free(L.bstart);
    free(L.blk);
    free(L.order);
    free(L.rank);
    return OR_OK;
}
