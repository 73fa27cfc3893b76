// Fully-implicit black-oil residual and Newton Jacobian on the device, as a
// 3x3 block matrix (rsgpu_csr_create_bsr3 layout).
//
// The reference simulator is IMPES (rs/src/no/uib/cipr/rs/numerics/
// Discretisation.java:208-225 solves only the pressure system implicitly);
// it has no fully-implicit Jacobian.  The relative permeabilities follow
// CoreyRockFluid (rs/src/no/uib/cipr/rs/rock/CoreyRockFluid.java:129-152:
// kr = ke Se^n, Se = (S - Sr) / (1 - sum of Sr)).
//
// Model.  Components water, oil, gas in surface volumes; phases water, oil,
// gas, gas dissolved in oil (Rs); unknowns per cell (p, Sw, Sg), So = 1 - Sw
// - Sg.  R_c,i = V/dt (acc_c(u_i) - acc_c(u_old,i)) + sum_j F_c(i->j) with
// acc_w = phi Sw bw, acc_o = phi So bo, acc_g = phi (Sg bg + Rs So bo);
// F_a(i->j) = T_ij (lambda_a b_a)_up dPhi_a, dPhi_a = (p_i - p_j) -
// rho_a,avg g (z_i - z_j), and the dissolved gas T (lambda_o bo Rs)_up
// dPhi_o; TPFA T_ij = A / (h/k_i + h/k_j); upstream = the lower-indexed cell
// when its potential drop is >= 0.  b_w, b_o, phi linear in p; b_g, Rs
// proportional to p; rho_w = rho_w,s bw, rho_o = (rho_o,s + Rs rho_g,s) bo,
// rho_g = rho_g,s bg; z the cell-centre depth (k + 1/2) dz.
//
// Design.  One lane per cell (block row) computes its accumulation, its
// faces in ascending neighbour order, the residual and its seven 3x3 blocks
// (analytic derivatives), and writes them into the row-interleaved value
// array.  Each face is evaluated from both sides; the potential drop seen
// from j is exactly the negation of the one seen from i, so both sides agree
// on the upstream cell.  Every sum is taken in a fixed order (accumulation
// first, then faces by ascending neighbour), so the result does not depend
// on scheduling.  Only +, -, *, / are used (integer Corey exponents by
// repeated products), so with -ffp-contract=off the values equal the CPU
// restatement's bit for bit.
#include "spmv_core.hpp"

struct rsgpu_blackoil {
    rsgpu_ctx *ctx = nullptr;
    rsgpu_blackoil_params P{};
    int nx = 0, ny = 0, nz = 0, n = 0;
    DevBuf<double> kh;
    uint64_t matrix_id = 0;   // pattern id of the matrix rsgpu_blackoil_matrix made
};

namespace {

struct BoCell {
    double bw, dbw, bo, dbo, bg, dbg, rs, drs, phi, dphi;
    double rho[3], drho[3];
    double lam[3], dlam_sw[3], dlam_sg[3];
    double acc[3];
};

__device__ double ipow(double x, int n)
{
    double r = 1.0;
    for (int k = 0; k < n; ++k)
        r = r * x;
    return r;
}

// kr = ke Se^n, Se = (S - Sr) / denom; dkr/dS = ke n Se^(n-1) / denom
__device__ void corey(double S, double Sr, double denom, double ke, int n, double &kr,
                      double &dkr)
{
    const double se = (S - Sr) / denom;
    kr = ke * ipow(se, n);
    dkr = n > 0 ? ke * (double)n * ipow(se, n - 1) / denom : 0.0;
}

__device__ void bo_cell(const rsgpu_blackoil_params &P, double p, double sw, double sg,
                        BoCell &c)
{
    const double dp = p - P.pref;
    const double so = (1.0 - sw) - sg;
    c.bw = P.bw0 * (1.0 + P.cw * dp);
    c.dbw = P.bw0 * P.cw;
    c.bo = P.bo0 * (1.0 + P.co * dp);
    c.dbo = P.bo0 * P.co;
    c.bg = P.bg0 * p / P.pref;
    c.dbg = P.bg0 / P.pref;
    c.rs = P.rs0 * p / P.pref;
    c.drs = P.rs0 / P.pref;
    c.phi = P.phi0 * (1.0 + P.cr * dp);
    c.dphi = P.phi0 * P.cr;
    c.rho[0] = P.rhow_s * c.bw;
    c.drho[0] = P.rhow_s * c.dbw;
    c.rho[1] = (P.rhoo_s + c.rs * P.rhog_s) * c.bo;
    c.drho[1] = c.drs * P.rhog_s * c.bo + (P.rhoo_s + c.rs * P.rhog_s) * c.dbo;
    c.rho[2] = P.rhog_s * c.bg;
    c.drho[2] = P.rhog_s * c.dbg;
    const double denom = ((1.0 - P.swr) - P.sor) - P.sgr;
    double kr, dkr;
    corey(sw, P.swr, denom, P.kw, P.nw, kr, dkr);
    c.lam[0] = kr / P.muw;
    c.dlam_sw[0] = dkr / P.muw;
    c.dlam_sg[0] = 0.0;
    corey(so, P.sor, denom, P.ko, P.no, kr, dkr);
    c.lam[1] = kr / P.muo;
    c.dlam_sw[1] = -dkr / P.muo;
    c.dlam_sg[1] = -dkr / P.muo;
    corey(sg, P.sgr, denom, P.kg, P.ng, kr, dkr);
    c.lam[2] = kr / P.mug;
    c.dlam_sw[2] = 0.0;
    c.dlam_sg[2] = dkr / P.mug;
    c.acc[0] = c.phi * sw * c.bw;
    c.acc[1] = c.phi * so * c.bo;
    c.acc[2] = c.phi * (sg * c.bg + c.rs * so * c.bo);
}

// d acc / d (p, Sw, Sg); rows are the water, oil, gas components
__device__ void bo_acc_jac(double sw, double sg, const BoCell &c, double J[9])
{
    const double so = (1.0 - sw) - sg;
    J[0] = c.dphi * sw * c.bw + c.phi * sw * c.dbw;
    J[1] = c.phi * c.bw;
    J[2] = 0.0;
    J[3] = c.dphi * so * c.bo + c.phi * so * c.dbo;
    J[4] = -(c.phi * c.bo);
    J[5] = -(c.phi * c.bo);
    J[6] = c.dphi * (sg * c.bg + c.rs * so * c.bo) +
           c.phi * (sg * c.dbg + c.drs * so * c.bo + c.rs * so * c.dbo);
    J[7] = -(c.phi * c.rs * c.bo);
    J[8] = c.phi * (c.bg - c.rs * c.bo);
}

// fluxes F(i->j) and their derivatives with respect to u_i (Ji), u_j (Jj)
__device__ void bo_face(const rsgpu_blackoil_params &P, int i, int j, double T, double dz_ij,
                        double pi, double pj, const BoCell &ci, const BoCell &cj, double F[3],
                        double Ji[9], double Jj[9])
{
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        Ji[e] = 0.0;
        Jj[e] = 0.0;
    }
    F[0] = F[1] = F[2] = 0.0;
    double dphi_o = 0.0, ddi_o = 0.0, ddj_o = 0.0;
    bool up_o = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double ravg = 0.5 * (ci.rho[a] + cj.rho[a]);
        const double dphi = (pi - pj) - ravg * P.g * dz_ij;
        const double ddi = 1.0 - 0.5 * ci.drho[a] * P.g * dz_ij;
        const double ddj = -1.0 - 0.5 * cj.drho[a] * P.g * dz_ij;
        const bool up_i = (i < j) ? (dphi >= 0.0) : (dphi > 0.0);
        const BoCell &cu = up_i ? ci : cj;
        const double b = a == 0 ? cu.bw : a == 1 ? cu.bo : cu.bg;
        const double db = a == 0 ? cu.dbw : a == 1 ? cu.dbo : cu.dbg;
        const double m = cu.lam[a] * b;
        const double dm[3] = {cu.lam[a] * db, cu.dlam_sw[a] * b, cu.dlam_sg[a] * b};
        F[a] = T * m * dphi;
        Ji[3 * a] = T * (m * ddi);
        Jj[3 * a] = T * (m * ddj);
        double *Ju = up_i ? Ji : Jj;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Ju[3 * a + c] = Ju[3 * a + c] + T * (dm[c] * dphi);
        if (a == 1) {
            dphi_o = dphi;
            ddi_o = ddi;
            ddj_o = ddj;
            up_o = up_i;
        }
    }
    // dissolved gas travels with the oil phase
    const BoCell &cu = up_o ? ci : cj;
    const double mr = cu.lam[1] * cu.bo * cu.rs;
    const double dmr[3] = {cu.lam[1] * (cu.dbo * cu.rs + cu.bo * cu.drs),
                           cu.dlam_sw[1] * cu.bo * cu.rs, cu.dlam_sg[1] * cu.bo * cu.rs};
    F[2] = F[2] + T * mr * dphi_o;
    Ji[6] = Ji[6] + T * (mr * ddi_o);
    Jj[6] = Jj[6] + T * (mr * ddj_o);
    double *Ju = up_o ? Ji : Jj;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        Ju[6 + c] = Ju[6 + c] + T * (dmr[c] * dphi_o);
}

struct Grid3 {
    int nx, ny, nz;
};

// the 7-point neighbours of cell i in ascending order (including i)
__device__ int bo_neighbours(const Grid3 &G, int i, int nb[7])
{
    const int x = i % G.nx, y = (i / G.nx) % G.ny, z = i / (G.nx * G.ny);
    int k = 0;
    if (z > 0)
        nb[k++] = i - G.nx * G.ny;
    if (y > 0)
        nb[k++] = i - G.nx;
    if (x > 0)
        nb[k++] = i - 1;
    nb[k++] = i;
    if (x < G.nx - 1)
        nb[k++] = i + 1;
    if (y < G.ny - 1)
        nb[k++] = i + G.nx;
    if (z < G.nz - 1)
        nb[k++] = i + G.nx * G.ny;
    return k;
}

// TPFA transmissibility of the face between i and j (k_z = kz_ratio k_h)
__device__ double bo_trans(const rsgpu_blackoil_params &P, const Grid3 &G, int i, int j,
                           const double *kh)
{
    // direction from the cells' layer and line (also right for nx or ny = 1)
    const int plane = G.nx * G.ny;
    double area, h, ki, kj;
    if (i / plane == j / plane && i / G.nx == j / G.nx) {
        area = P.dy * P.dz;
        h = 0.5 * P.dx;
        ki = kh[i];
        kj = kh[j];
    } else if (i / plane == j / plane) {
        area = P.dx * P.dz;
        h = 0.5 * P.dy;
        ki = kh[i];
        kj = kh[j];
    } else {
        area = P.dx * P.dy;
        h = 0.5 * P.dz;
        ki = P.kz_ratio * kh[i];
        kj = P.kz_ratio * kh[j];
    }
    return area / (h / ki + h / kj);
}

__global__ void __launch_bounds__(128)
    k_bo_assemble(rsgpu_blackoil_params P, Grid3 G, const double *__restrict__ kh,
                  const double *__restrict__ u, const double *__restrict__ u_old,
                  const int *__restrict__ brp, double *__restrict__ val, double *__restrict__ R)
{
    const int n = G.nx * G.ny * G.nz;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double vdt = (P.dx * P.dy * P.dz) / P.dt;
    int nb[7];
    const int k = bo_neighbours(G, i, nb);
    BoCell ci, co;
    bo_cell(P, u[3 * i], u[3 * i + 1], u[3 * i + 2], ci);
    bo_cell(P, u_old[3 * i], u_old[3 * i + 1], u_old[3 * i + 2], co);
    double Jd[9], r[3];
    bo_acc_jac(u[3 * i + 1], u[3 * i + 2], ci, Jd);
#pragma unroll
    for (int q = 0; q < 9; ++q)
        Jd[q] = vdt * Jd[q];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        r[a] = vdt * (ci.acc[a] - co.acc[a]);
    const int zi = i / (G.nx * G.ny);
    const int base = val ? 9 * brp[i] : 0;
    int qd = 0;
    for (int q = 0; q < k; ++q) {
        const int j = nb[q];
        if (j == i) {
            qd = q;
            continue;
        }
        BoCell cj;
        bo_cell(P, u[3 * j], u[3 * j + 1], u[3 * j + 2], cj);
        const int zj = j / (G.nx * G.ny);
        const double dz_ij = ((double)zi + 0.5) * P.dz - ((double)zj + 0.5) * P.dz;
        const double T = bo_trans(P, G, i, j, kh);
        double F[3], Ji[9], Jj[9];
        bo_face(P, i, j, T, dz_ij, u[3 * i], u[3 * j], ci, cj, F, Ji, Jj);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            r[a] = r[a] + F[a];
#pragma unroll
        for (int e = 0; e < 9; ++e)
            Jd[e] = Jd[e] + Ji[e];
        if (val) {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    val[base + a * 3 * k + 3 * q + c] = Jj[3 * a + c];
        }
    }
    if (val) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                val[base + a * 3 * k + 3 * qd + c] = Jd[3 * a + c];
    }
    if (R) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            R[3 * i + a] = r[a];
    }
}

int bo_fail(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

} // namespace

extern "C" int rsgpu_blackoil_create(rsgpu_ctx *ctx, const rsgpu_blackoil_params *p, int nx,
                                     int ny, int nz, const double *kh_host,
                                     rsgpu_blackoil **out)
{
    if (!ctx || !p || !kh_host || !out || nx < 1 || ny < 1 || nz < 1)
        return RSGPU_BAD_ARGUMENT;
    if ((long long)nx * ny * nz * 21 > 0x7fffffffLL) {
        ctx->err = "grid too large for 32-bit scalar indices";
        return RSGPU_BAD_ARGUMENT;
    }
    if (p->nw < 0 || p->no < 0 || p->ng < 0 || p->nw > 16 || p->no > 16 || p->ng > 16) {
        ctx->err = "Corey exponents must be integers in [0, 16]";
        return RSGPU_BAD_ARGUMENT;
    }
    rsgpu_blackoil *B = new rsgpu_blackoil();
    try {
        RS_HIP(hipSetDevice(ctx->device));
        B->ctx = ctx;
        B->P = *p;
        B->nx = nx;
        B->ny = ny;
        B->nz = nz;
        B->n = nx * ny * nz;
        B->kh.upload(kh_host, B->n, ctx->stream);
        RS_HIP(hipStreamSynchronize(ctx->stream));
    } catch (const RsError &e) {
        delete B;
        return bo_fail(ctx, e);
    }
    *out = B;
    return RSGPU_OK;
}

extern "C" int rsgpu_blackoil_destroy(rsgpu_blackoil *B)
{
    if (B) {
        if (B->ctx)
            (void)hipStreamSynchronize(B->ctx->stream);
        delete B;
    }
    return RSGPU_OK;
}

extern "C" int rsgpu_blackoil_matrix(rsgpu_blackoil *B, rsgpu_csr **J_out)
{
    if (!B || !J_out)
        return RSGPU_BAD_ARGUMENT;
    const int nx = B->nx, ny = B->ny, nz = B->nz, n = B->n;
    std::vector<int> brp(n + 1), bcol;
    bcol.reserve((size_t)7 * n);
    brp[0] = 0;
    for (int i = 0; i < n; ++i) {
        const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        if (z > 0)
            bcol.push_back(i - nx * ny);
        if (y > 0)
            bcol.push_back(i - nx);
        if (x > 0)
            bcol.push_back(i - 1);
        bcol.push_back(i);
        if (x < nx - 1)
            bcol.push_back(i + 1);
        if (y < ny - 1)
            bcol.push_back(i + nx);
        if (z < nz - 1)
            bcol.push_back(i + nx * ny);
        brp[i + 1] = (int)bcol.size();
    }
    const int st = rsgpu_csr_create_bsr3(B->ctx, n, (int)bcol.size(), brp.data(), bcol.data(),
                                         J_out);
    if (st == RSGPU_OK)
        B->matrix_id = (*J_out)->pattern_id;
    return st;
}

extern "C" int rsgpu_blackoil_assemble(rsgpu_blackoil *B, const double *u_dev,
                                       const double *u_old_dev, rsgpu_csr *J,
                                       double *residual_dev)
{
    if (!B || !u_dev || !u_old_dev)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_ctx *ctx = B->ctx;
    if (J && (J->pattern_id != B->matrix_id || J->bs != 3)) {
        ctx->err = "the Jacobian must be the matrix rsgpu_blackoil_matrix made for this grid";
        return RSGPU_BAD_ARGUMENT;
    }
    try {
        RS_HIP(hipSetDevice(ctx->device));
        const Grid3 G{B->nx, B->ny, B->nz};
        hipLaunchKernelGGL(k_bo_assemble, dim3((B->n + 127) / 128), dim3(128), 0, ctx->stream,
                           B->P, G, B->kh.p, u_dev, u_old_dev, J ? J->brp.p : nullptr,
                           J ? J->val.p : nullptr, residual_dev);
        RS_CHECK_LAUNCH();
        if (J)
            J->values_changed();
        return RSGPU_OK;
    } catch (const RsError &e) {
        return bo_fail(ctx, e);
    }
}
