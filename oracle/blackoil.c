/*
 * CPU oracle: fully-implicit black-oil residual and Newton Jacobian on a
 * Cartesian grid, stored as 3x3 blocks in the row-interleaved block layout.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference simulator is
 * IMPES and has no fully-implicit Jacobian; this restatement defines the
 * discretisation the GPU assembly is checked against, and is itself pinned
 * by central finite differences of its own residual, by the antisymmetry of
 * the face fluxes, and by the Corey form of
 * rs/src/no/uib/cipr/rs/rock/CoreyRockFluid.java:129-152
 * (kr = ke * Se^n, Se = (S - Sr) / (1 - sum of Sr)).
 *
 * Model.  Components water, oil, gas in surface volumes; phases water, oil,
 * gas; gas dissolves in oil (Rs).  Unknowns per cell u = (p, Sw, Sg),
 * So = 1 - Sw - Sg.  Residual of component c in cell i:
 *   R_c,i = V/dt (acc_c(u_i) - acc_c(u_old_i)) + sum_j F_c(i->j)
 *   acc_w = phi Sw bw, acc_o = phi So bo, acc_g = phi (Sg bg + Rs So bo)
 *   F_a(i->j) = T_ij lambda_a,up b_a,up dPhi_a, dPhi_a = (p_i - p_j)
 *               - rho_a,avg g (z_i - z_j); gas adds T lambda_o bo Rs (up) dPhi_o
 * with TPFA transmissibilities T_ij = A / (h/k_i + h/k_j), phase upwinding
 * (the lower-indexed cell is upstream when its potential drop is >= 0),
 * b = 1/B linear in p (water, oil), b_g and Rs proportional to p, phi
 * linear in p, rho_w = rho_w,s bw, rho_o = (rho_o,s + Rs rho_g,s) bo,
 * rho_g = rho_g,s bg, z the cell-centre depth (k + 1/2) dz.
 *
 * Block layout.  Block row I with blocks b0..b1-1 (columns ascending: the
 * 7-point neighbours and I itself) is stored as 3 consecutive scalar rows;
 * the value (row a, block b, column c) sits at
 * 9 b0 + a * 3 (b1 - b0) + 3 (b - b0) + c, i.e. it is the CSR of the
 * expanded scalar matrix with sorted columns.  Rows a = water, oil, gas
 * equations; columns c = p, Sw, Sg.  Every entry is summed in a fixed
 * order: the diagonal block starts with the accumulation term and adds the
 * faces in ascending neighbour order; the residual likewise.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    double bw, dbw, bo, dbo, bg, dbg, rs, drs, phi, dphi;
    double rho[3], drho[3];
    /* mobilities lambda_a = kr_a / mu_a and their saturation derivatives */
    double lam[3], dlam_sw[3], dlam_sg[3];
    double acc[3];
} BoCell;

static double ipow(double x, int n)
{
    double r = 1.0;
    for (int k = 0; k < n; ++k)
        r = r * x;
    return r;
}

/* Corey: kr = ke Se^n, Se = (S - Sr) / denom; dkr/dS = ke n Se^(n-1) / denom */
static void corey(double S, double Sr, double denom, double ke, int n, double *kr,
                  double *dkr)
{
    const double se = (S - Sr) / denom;
    *kr = ke * ipow(se, n);
    *dkr = n > 0 ? ke * (double)n * ipow(se, n - 1) / denom : 0.0;
}

static void bo_cell(const or_bo_params *P, double p, double sw, double sg, BoCell *c)
{
    const double dp = p - P->pref;
    const double so = (1.0 - sw) - sg;
    c->bw = P->bw0 * (1.0 + P->cw * dp);
    c->dbw = P->bw0 * P->cw;
    c->bo = P->bo0 * (1.0 + P->co * dp);
    c->dbo = P->bo0 * P->co;
    c->bg = P->bg0 * p / P->pref;
    c->dbg = P->bg0 / P->pref;
    c->rs = P->rs0 * p / P->pref;
    c->drs = P->rs0 / P->pref;
    c->phi = P->phi0 * (1.0 + P->cr * dp);
    c->dphi = P->phi0 * P->cr;
    c->rho[0] = P->rhow_s * c->bw;
    c->drho[0] = P->rhow_s * c->dbw;
    c->rho[1] = (P->rhoo_s + c->rs * P->rhog_s) * c->bo;
    c->drho[1] = c->drs * P->rhog_s * c->bo + (P->rhoo_s + c->rs * P->rhog_s) * c->dbo;
    c->rho[2] = P->rhog_s * c->bg;
    c->drho[2] = P->rhog_s * c->dbg;
    const double denom = ((1.0 - P->swr) - P->sor) - P->sgr;
    double kr, dkr;
    corey(sw, P->swr, denom, P->kw, P->nw, &kr, &dkr);
    c->lam[0] = kr / P->muw;
    c->dlam_sw[0] = dkr / P->muw;
    c->dlam_sg[0] = 0.0;
    corey(so, P->sor, denom, P->ko, P->no, &kr, &dkr);
    c->lam[1] = kr / P->muo;
    c->dlam_sw[1] = -dkr / P->muo;
    c->dlam_sg[1] = -dkr / P->muo;
    corey(sg, P->sgr, denom, P->kg, P->ng, &kr, &dkr);
    c->lam[2] = kr / P->mug;
    c->dlam_sw[2] = 0.0;
    c->dlam_sg[2] = dkr / P->mug;
    c->acc[0] = c->phi * sw * c->bw;
    c->acc[1] = c->phi * so * c->bo;
    c->acc[2] = c->phi * (sg * c->bg + c->rs * so * c->bo);
}

/* d acc / d (p, Sw, Sg), rows = components */
static void bo_acc_jac(double sw, double sg, const BoCell *c, double J[9])
{
    const double so = (1.0 - sw) - sg;
    J[0] = c->dphi * sw * c->bw + c->phi * sw * c->dbw;
    J[1] = c->phi * c->bw;
    J[2] = 0.0;
    J[3] = c->dphi * so * c->bo + c->phi * so * c->dbo;
    J[4] = -(c->phi * c->bo);
    J[5] = -(c->phi * c->bo);
    J[6] = c->dphi * (sg * c->bg + c->rs * so * c->bo) +
           c->phi * (sg * c->dbg + c->drs * so * c->bo + c->rs * so * c->dbo);
    J[7] = -(c->phi * c->rs * c->bo);
    J[8] = c->phi * (c->bg - c->rs * c->bo);
}

/* Face between row cell i and neighbour j: the component fluxes F(i->j)
 * and their derivatives with respect to u_i (Ji) and u_j (Jj). */
static void bo_face(const or_bo_params *P, int i, int j, double T, double dz_ij,
                    const double *u, const BoCell *ci, const BoCell *cj, double F[3],
                    double Ji[9], double Jj[9])
{
    memset(Ji, 0, 9 * sizeof(double));
    memset(Jj, 0, 9 * sizeof(double));
    F[0] = F[1] = F[2] = 0.0;
    const double pi = u[3 * i], pj = u[3 * j];
    double dphi_o = 0.0, ddi_o = 0.0, ddj_o = 0.0;
    int up_o = 0;
    for (int a = 0; a < 3; ++a) {
        const double ravg = 0.5 * (ci->rho[a] + cj->rho[a]);
        const double dphi = (pi - pj) - ravg * P->g * dz_ij;
        const double ddi = 1.0 - 0.5 * ci->drho[a] * P->g * dz_ij;
        const double ddj = -1.0 - 0.5 * cj->drho[a] * P->g * dz_ij;
        const int up_i = (i < j) ? (dphi >= 0.0) : (dphi > 0.0);
        const BoCell *cu = up_i ? ci : cj;
        const double b = a == 0 ? cu->bw : a == 1 ? cu->bo : cu->bg;
        const double db = a == 0 ? cu->dbw : a == 1 ? cu->dbo : cu->dbg;
        const double m = cu->lam[a] * b;
        const double dm[3] = {cu->lam[a] * db, cu->dlam_sw[a] * b, cu->dlam_sg[a] * b};
        F[a] = T * m * dphi;
        double *Ju = up_i ? Ji : Jj;
        Ji[3 * a] = T * (m * ddi);
        Jj[3 * a] = T * (m * ddj);
        for (int c = 0; c < 3; ++c)
            Ju[3 * a + c] = Ju[3 * a + c] + T * (dm[c] * dphi);
        if (a == 1) {
            dphi_o = dphi;
            ddi_o = ddi;
            ddj_o = ddj;
            up_o = up_i;
        }
    }
    /* dissolved gas travels with the oil phase */
    const BoCell *cu = up_o ? ci : cj;
    const double mr = cu->lam[1] * cu->bo * cu->rs;
    const double dmr[3] = {cu->lam[1] * (cu->dbo * cu->rs + cu->bo * cu->drs),
                           cu->dlam_sw[1] * cu->bo * cu->rs, cu->dlam_sg[1] * cu->bo * cu->rs};
    F[2] = F[2] + T * mr * dphi_o;
    double *Ju = up_o ? Ji : Jj;
    Ji[6] = Ji[6] + T * (mr * ddi_o);
    Jj[6] = Jj[6] + T * (mr * ddj_o);
    for (int c = 0; c < 3; ++c)
        Ju[6 + c] = Ju[6 + c] + T * (dmr[c] * dphi_o);
}

/* the 7-point neighbour list of cell i in ascending order (including i) */
static int bo_neighbours(int nx, int ny, int nz, int i, int nb[7])
{
    const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    int k = 0;
    if (z > 0)
        nb[k++] = i - nx * ny;
    if (y > 0)
        nb[k++] = i - nx;
    if (x > 0)
        nb[k++] = i - 1;
    nb[k++] = i;
    if (x < nx - 1)
        nb[k++] = i + 1;
    if (y < ny - 1)
        nb[k++] = i + nx;
    if (z < nz - 1)
        nb[k++] = i + nx * ny;
    return k;
}

/* TPFA: T = A / (h / k_i + h / k_j) with the face area and half cell length
 * of the face's direction; k_z = kz_ratio k_h */
static double bo_trans(const or_bo_params *P, int nx, int ny, int i, int j, const double *kh)
{
    /* direction from the cells' layer and line (also right for nx or ny = 1) */
    const int plane = nx * ny;
    double area, h, ki, kj;
    if (i / plane == j / plane && i / nx == j / nx) {
        area = P->dy * P->dz;
        h = 0.5 * P->dx;
        ki = kh[i];
        kj = kh[j];
    } else if (i / plane == j / plane) {
        area = P->dx * P->dz;
        h = 0.5 * P->dy;
        ki = kh[i];
        kj = kh[j];
    } else {
        area = P->dx * P->dy;
        h = 0.5 * P->dz;
        ki = P->kz_ratio * kh[i];
        kj = P->kz_ratio * kh[j];
    }
    return area / (h / ki + h / kj);
}

int or_bo_pattern(int nx, int ny, int nz, int *browptr, int *bcolind)
{
    const int n = nx * ny * nz;
    int nnzb = 0;
    if (browptr)
        browptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        int nb[7];
        const int k = bo_neighbours(nx, ny, nz, i, nb);
        for (int q = 0; q < k; ++q) {
            if (bcolind)
                bcolind[nnzb] = nb[q];
            nnzb++;
        }
        if (browptr)
            browptr[i + 1] = nnzb;
    }
    return nnzb;
}

void or_bo_assemble(const or_bo_params *P, int nx, int ny, int nz, const double *kh,
                    const double *u, const double *u_old, double *val, double *R)
{
    const int n = nx * ny * nz;
    const double vdt = (P->dx * P->dy * P->dz) / P->dt;
    int base = 0;
    for (int i = 0; i < n; ++i) {
        int nb[7];
        const int k = bo_neighbours(nx, ny, nz, i, nb);
        BoCell ci, co;
        bo_cell(P, u[3 * i], u[3 * i + 1], u[3 * i + 2], &ci);
        bo_cell(P, u_old[3 * i], u_old[3 * i + 1], u_old[3 * i + 2], &co);
        double Jd[9], r[3];
        bo_acc_jac(u[3 * i + 1], u[3 * i + 2], &ci, Jd);
        for (int q = 0; q < 9; ++q)
            Jd[q] = vdt * Jd[q];
        for (int a = 0; a < 3; ++a)
            r[a] = vdt * (ci.acc[a] - co.acc[a]);
        const int zi = i / (nx * ny);
        int qd = 0;
        for (int q = 0; q < k; ++q) {
            const int j = nb[q];
            if (j == i) {
                qd = q;
                continue;
            }
            BoCell cj;
            bo_cell(P, u[3 * j], u[3 * j + 1], u[3 * j + 2], &cj);
            const int zj = j / (nx * ny);
            const double dz_ij = ((double)zi + 0.5) * P->dz - ((double)zj + 0.5) * P->dz;
            const double T = bo_trans(P, nx, ny, i, j, kh);
            double F[3], Ji[9], Jj[9];
            bo_face(P, i, j, T, dz_ij, u, &ci, &cj, F, Ji, Jj);
            for (int a = 0; a < 3; ++a)
                r[a] = r[a] + F[a];
            for (int e = 0; e < 9; ++e)
                Jd[e] = Jd[e] + Ji[e];
            if (val)
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c)
                        val[base + a * 3 * k + 3 * q + c] = Jj[3 * a + c];
        }
        if (val)
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c)
                    val[base + a * 3 * k + 3 * qd + c] = Jd[3 * a + c];
        if (R)
            for (int a = 0; a < 3; ++a)
                R[3 * i + a] = r[a];
        base += 9 * k;
    }
}

/* the expanded scalar pattern of the block pattern (row-interleaved) */
int or_bo_expand(int nb, const int *browptr, const int *bcolind, int *rowptr, int *colind)
{
    int p = 0;
    rowptr[0] = 0;
    for (int I = 0; I < nb; ++I) {
        for (int a = 0; a < 3; ++a) {
            for (int b = browptr[I]; b < browptr[I + 1]; ++b)
                for (int c = 0; c < 3; ++c)
                    colind[p++] = 3 * bcolind[b] + c;
            rowptr[3 * I + a + 1] = p;
        }
    }
    return p;
}
