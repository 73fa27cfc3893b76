// IMPES pressure Jacobian on gfx950: control-surface secondaries, flux part,
// diagonal part and source terms of rs's pressure Newton step.
//
// Reference: rs/src/no/uib/cipr/rs/field/CS.java:94-170 (calculateMassDensity,
// calculateDarcyFlux, bodyForce), rs/src/no/uib/cipr/rs/numerics/
// Discretisation.java:284-289 (flux part, then diagonal part),
// :365-453 (assemblePressureJacobianDiagonal, regular and outlet sources),
// :459-561 (assemblePressureJacobianFlux), LinearSolver.java:226-235
// (addToJacobian = CompRowMatrix.add, addToResidual = DenseVector.add).
//
// Design.  The reference walks connections and scatters into rows; every
// matrix entry and residual element is a sum, starting from zero, of its
// contributions in call order.  Here the work splits into
//   k_impes_cs    one lane per connection: the CS secondaries
//                 (density average, potential sum over MD, upstream, flux);
//   k_impes_conn  one lane per connection: per MD entry the here/there
//                 Jacobian contributions Ji, Jj and per connection the
//                 residual contributions ri, rj, each in the reference's
//                 loop order (MD entry, phase, component, heat term), so
//                 every Ji, Jj, ri, rj is the Java value bit for bit;
//   k_impes_rows  one lane per cell: its row of A and its b element as the
//                 sums of its contributions over its incident connections
//                 in ascending order (the order of the reference's
//                 connection loop), then the diagonal term and the cell's
//                 source terms in application order.
// Only +, -, *, / are used (no fused multiply-add, -ffp-contract=off), so
// A and b equal the reference's loops bit for bit.  Bytes per cell are a
// few hundred (the secondaries), so the kernels are HBM/latency bound and
// small next to one solve.
#include <algorithm>
#include <vector>

#include "rsgpu_impl.hpp"

struct rsgpu_impes {
    rsgpu_ctx *ctx = nullptr;
    rsgpu_flux *F = nullptr;
    const rsgpu_flux *Fmf = nullptr;
    int ncomp = 0;
    DevBuf<double> jh, jt, rh, rt;        // per MD entry / per connection
    int nterms = 0;
    DevBuf<int> src_ptr, src_term, src_kind;
    DevBuf<double> src_q, src_qe;
    DevBuf<int> err;
};

namespace {

struct ImpesK {
    int n, nconn, ncomp, thermal;
    double dt, g, omega, ax, ay;
    rsgpu_impes_state S;
};

__device__ __forceinline__ double body_force(const ImpesK &K, int e)
{
    const double x = K.S.cx[e], y = K.S.cy[e], z = K.S.cz[e];
    const double distance = (K.ax - x) * (K.ax - x) + (K.ay - y) * (K.ay - y);
    return K.g * z - (K.omega * K.omega / 2) * distance;
}

__device__ __forceinline__ double ph_at(const double *a, int n, int ph, int e)
{
    return a[(size_t)ph * n + e];
}

__global__ void k_impes_cs(ImpesK K, const int *__restrict__ md_ptr,
                           const int *__restrict__ md_elem, const double *__restrict__ md_t,
                           const int *__restrict__ here, const int *__restrict__ there,
                           double *__restrict__ rho, int *__restrict__ upstream,
                           double *__restrict__ darcy)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K.nconn)
        return;
    const int n = K.n, nc = K.nconn, h = here[c], t = there[c];
    double r[3];
    for (int ph = 0; ph < 3; ++ph) {
        const double Sh = ph_at(K.S.saturation, n, ph, h);
        const double St = ph_at(K.S.saturation, n, ph, t);
        const double Ssum = Sh + St;
        const double rh = ph_at(K.S.mass_density, n, ph, h);
        const double rt = ph_at(K.S.mass_density, n, ph, t);
        const double av = (Sh * rh + St * rt) / Ssum;
        r[ph] = isnan(av) ? 0.0 : av;
        rho[(size_t)ph * nc + c] = r[ph];
    }
    const int q0 = md_ptr[c], q1 = md_ptr[c + 1];
    for (int ph = 0; ph < 3; ++ph) {
        double D = 0;
        for (int q = q0; q < q1; ++q) {
            const int ek = md_elem[q];
            const double psi = ph_at(K.S.phase_pressure, n, ph, ek) + r[ph] * body_force(K, ek);
            D += md_t[q] * psi;
        }
        const int up = (D > 0) ? h : t;
        upstream[(size_t)ph * nc + c] = up;
        darcy[(size_t)ph * nc + c] = ph_at(K.S.phase_mobility, n, ph, up) * D;
    }
}

__global__ void k_impes_conn(ImpesK K, const int *__restrict__ md_ptr,
                             const int *__restrict__ md_elem, const double *__restrict__ md_t,
                             const int *__restrict__ mf_ptr, const int *__restrict__ mf_elem,
                             const double *__restrict__ mf_t, const int *__restrict__ here,
                             const int *__restrict__ there, const double *__restrict__ rho,
                             const int *__restrict__ upstream, double *__restrict__ jh,
                             double *__restrict__ jt, double *__restrict__ rh,
                             double *__restrict__ rt)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K.nconn)
        return;
    const int n = K.n, nc = K.nconn, ncomp = K.ncomp;
    const int i = here[c], j = there[c];
    const bool thermal = K.thermal != 0;
    const double deltaHere = thermal ? K.S.heat_capacity[i] : 0.0;
    const double deltaThere = thermal ? K.S.heat_capacity[j] : 0.0;
    const double dRdT_here = thermal ? K.S.dRdT[i] : 0.0;
    const double dRdT_there = thermal ? K.S.dRdT[j] : 0.0;
    int up[3];
    double r[3];
    bool pres[3];
    for (int ph = 0; ph < 3; ++ph) {
        up[ph] = upstream[(size_t)ph * nc + c];
        r[ph] = rho[(size_t)ph * nc + c];
        pres[ph] = K.S.present[(size_t)ph * n + up[ph]] != 0;
    }
    double ri = 0, rj = 0;
    for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
        const int ik = md_elem[q];
        const double tk = md_t[q];
        const double b = body_force(K, ik);
        double Ji = 0, Jj = 0;
        for (int ph = 0; ph < 3; ++ph) {
            if (!pres[ph])
                continue;
            const double psi = ph_at(K.S.phase_pressure, n, ph, ik) + r[ph] * b;
            for (int nu = 0; nu < ncomp; ++nu) {
                const double phaseCompMob =
                    K.S.comp_mobility[((size_t)ph * ncomp + nu) * n + up[ph]];
                const double dRdNhere = K.S.dRdN[(size_t)nu * n + i];
                const double dRdNthere = K.S.dRdN[(size_t)nu * n + j];
                const double dNdt = -phaseCompMob * tk * psi;
                const double d2Ndtp = -phaseCompMob * tk;
                ri += dRdNhere * dNdt;
                rj += -dRdNthere * dNdt;
                Ji += dRdNhere * d2Ndtp;
                Jj += -dRdNthere * d2Ndtp;
            }
            if (thermal) {
                const double phaseEnergyMob = ph_at(K.S.energy_mobility, n, ph, up[ph]);
                const double dTdt = -phaseEnergyMob * tk * psi;
                const double d2Tdtp = -phaseEnergyMob * tk;
                ri += dRdT_here * dTdt / deltaHere;
                rj += -dRdT_there * dTdt / deltaThere;
                Ji += dRdT_here * d2Tdtp / deltaHere;
                Jj += -dRdT_there * d2Tdtp / deltaThere;
            }
        }
        jh[q] = Ji;
        jt[q] = Jj;
    }
    if (thermal)
        for (int q = mf_ptr[c]; q < mf_ptr[c + 1]; ++q) {
            const double F = mf_t[q] * K.S.temperature[mf_elem[q]];
            ri += -dRdT_here * F / deltaHere;
            rj += dRdT_there * F / deltaThere;
        }
    rh[c] = ri;
    rt[c] = rj;
}

__device__ __forceinline__ int find_col(const int *__restrict__ colind, int lo, int hi, int k)
{
    // first position in [lo, hi) with colind >= k
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (colind[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_impes_rows(ImpesK K, const int *__restrict__ cell_ptr,
                             const int *__restrict__ cell_conn,
                             const signed char *__restrict__ cell_sign,
                             const int *__restrict__ md_ptr, const int *__restrict__ md_elem,
                             const double *__restrict__ jh, const double *__restrict__ jt,
                             const double *__restrict__ rh, const double *__restrict__ rt,
                             const int *__restrict__ rowptr, const int *__restrict__ colind,
                             double *__restrict__ val, const int *__restrict__ src_ptr,
                             const int *__restrict__ src_term, const int *__restrict__ src_kind,
                             const double *__restrict__ src_q,
                             const double *__restrict__ src_qe, double *__restrict__ bvec,
                             int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= K.n)
        return;
    const int n = K.n, ncomp = K.ncomp;
    const int r0 = rowptr[r], r1 = rowptr[r + 1];
    for (int p = r0; p < r1; ++p)
        val[p] = 0.0;
    double acc = 0;
    // flux part: incident connections in ascending order
    for (int a = cell_ptr[r]; a < cell_ptr[r + 1]; ++a) {
        const int c = cell_conn[a];
        const bool h = cell_sign[a] > 0;
        const double *J = h ? jh : jt;
        for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
            const int k = md_elem[q];
            const int p = find_col(colind, r0, r1, k);
            if (p >= r1 || colind[p] != k) {
                atomicOr(err, 1);
                continue;
            }
            val[p] += J[q];
        }
        acc += h ? rh[c] : rt[c];
    }
    // diagonal part
    const double dp = K.S.pressure[r] - K.S.p0[r];
    const double dRdp = K.S.dRdp[r];
    acc += (K.S.R0[r] + dRdp * dp) / K.dt;
    {
        const int p = find_col(colind, r0, r1, r);
        if (p >= r1 || colind[p] != r)
            atomicOr(err, 1);
        else
            val[p] += dRdp / K.dt;
    }
    // source terms of this cell, in application order
    for (int a = src_ptr[r]; a < src_ptr[r + 1]; ++a) {
        const int t = src_term[a];
        const double *q = src_q + (size_t)t * 3 * ncomp;
        for (int nu = 0; nu < ncomp; ++nu) {
            double qnu;
            if (src_kind[t] == 0) {
                qnu = q[nu];
            } else {
                qnu = 0;
                for (int ph = 0; ph < 3; ++ph)
                    qnu += q[ph * ncomp + nu];
            }
            acc += K.S.dRdN[(size_t)nu * n + r] * qnu;
        }
        if (K.thermal)
            acc += K.S.dRdT[r] * src_qe[t] / K.S.heat_capacity[r];
    }
    bvec[r] = acc;
}

__global__ void k_impes_clear(int *err) { *err = 0; }

int impes_fail(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

ImpesK make_k(const rsgpu_impes *E, const rsgpu_impes_params *P, const rsgpu_impes_state *S)
{
    ImpesK K;
    K.n = E->F->ncells;
    K.nconn = E->F->nconn;
    K.ncomp = E->ncomp;
    K.thermal = P->thermal ? 1 : 0;
    K.dt = P->dt;
    K.g = P->g;
    K.omega = P->omega;
    K.ax = P->ax;
    K.ay = P->ay;
    K.S = *S;
    return K;
}

void check_state(const rsgpu_impes *E, const rsgpu_impes_params *P,
                 const rsgpu_impes_state *S, bool assemble)
{
    if (!P || !S)
        throw RsError{RSGPU_BAD_ARGUMENT, "null parameters or state"};
    if (E->F->ncells == 0)
        return;
    if (!S->phase_pressure || !S->saturation || !S->mass_density || !S->phase_mobility ||
        !S->cx || !S->cy || !S->cz)
        throw RsError{RSGPU_BAD_ARGUMENT, "missing control-volume arrays"};
    if (assemble) {
        if (!S->pressure || !S->present || !S->comp_mobility || !S->dRdN || !S->dRdp ||
            !S->p0 || !S->R0)
            throw RsError{RSGPU_BAD_ARGUMENT, "missing control-volume arrays"};
        if (P->thermal && (!S->heat_capacity || !S->dRdT || !S->energy_mobility ||
                           !S->temperature || !E->Fmf))
            throw RsError{RSGPU_BAD_ARGUMENT,
                          "thermal assembly needs the heat arrays and the Fourier lists"};
    }
}

}  // namespace

extern "C" int rsgpu_impes_create(rsgpu_flux *F, const rsgpu_flux *Fmf, int ncomp,
                                  rsgpu_impes **out)
{
    if (!F || !out || ncomp < 1)
        return RSGPU_BAD_ARGUMENT;
    if (Fmf && (Fmf->nconn != F->nconn || Fmf->ncells != F->ncells))
        return RSGPU_BAD_ARGUMENT;
    try {
        auto *E = new rsgpu_impes;
        E->ctx = F->ctx;
        E->F = F;
        E->Fmf = Fmf;
        E->ncomp = ncomp;
        E->jh.alloc(std::max(F->nentries, 1));
        E->jt.alloc(std::max(F->nentries, 1));
        E->rh.alloc(std::max(F->nconn, 1));
        E->rt.alloc(std::max(F->nconn, 1));
        E->err.alloc(1);
        std::vector<int> z(F->ncells + 1, 0);
        E->src_ptr.upload(z, F->ctx->stream);
        E->src_term.alloc(1);
        E->src_kind.alloc(1);
        E->src_q.alloc(1);
        E->src_qe.alloc(1);
        RS_HIP(hipStreamSynchronize(F->ctx->stream));
        *out = E;
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(F->ctx, e);
    }
}

extern "C" int rsgpu_impes_destroy(rsgpu_impes *E)
{
    if (!E)
        return RSGPU_BAD_ARGUMENT;
    (void)hipStreamSynchronize(E->ctx->stream);
    delete E;
    return RSGPU_OK;
}

extern "C" int rsgpu_impes_set_sources(rsgpu_impes *E, int nterms, const int32_t *elem,
                                       const int32_t *kind, const double *q,
                                       const double *qe)
{
    if (!E || nterms < 0 || (nterms && (!elem || !kind || !q || !qe)))
        return RSGPU_BAD_ARGUMENT;
    try {
        const int n = E->F->ncells;
        hipStream_t s = E->ctx->stream;
        std::vector<int> ptr(n + 1, 0), term(std::max(nterms, 1), 0);
        for (int t = 0; t < nterms; ++t) {
            if (elem[t] < 0 || elem[t] >= n)
                throw RsError{RSGPU_BAD_ARGUMENT, "source element out of range"};
            if (kind[t] != 0 && kind[t] != 1)
                throw RsError{RSGPU_BAD_ARGUMENT, "source kind must be 0 (regular) or 1 (outlet)"};
            ptr[elem[t] + 1]++;
        }
        for (int i = 0; i < n; ++i)
            ptr[i + 1] += ptr[i];
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int t = 0; t < nterms; ++t)   // stable: application order per cell
            term[pos[elem[t]]++] = t;
        RS_HIP(hipStreamSynchronize(s));   // earlier assemblies read the old lists
        E->src_ptr.upload(ptr, s);
        E->src_term.upload(term, s);
        if (nterms) {
            E->src_kind.upload(kind, nterms, s);
            E->src_q.upload(q, (size_t)nterms * 3 * E->ncomp, s);
            E->src_qe.upload(qe, nterms, s);
        }
        RS_HIP(hipStreamSynchronize(s));
        E->nterms = nterms;
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(E->ctx, e);
    }
}

extern "C" int rsgpu_impes_control_surfaces(rsgpu_impes *E, const rsgpu_impes_params *P,
                                            const rsgpu_impes_state *S, double *rho,
                                            int32_t *upstream, double *darcy)
{
    if (!E || (E->F->nconn && (!rho || !upstream || !darcy)))
        return RSGPU_BAD_ARGUMENT;
    try {
        check_state(E, P, S, false);
        const rsgpu_flux *F = E->F;
        if (F->nconn) {
            hipLaunchKernelGGL(k_impes_cs, dim3((F->nconn + 255) / 256), dim3(256), 0,
                               E->ctx->stream, make_k(E, P, S), F->md_ptr.p, F->md_elem.p,
                               F->md_t.p, F->here.p, F->there.p, rho, upstream, darcy);
            RS_CHECK_LAUNCH();
        }
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(E->ctx, e);
    }
}

extern "C" int rsgpu_impes_assemble(rsgpu_impes *E, rsgpu_csr *A, const rsgpu_impes_params *P,
                                    const rsgpu_impes_state *S, const double *rho,
                                    const int32_t *upstream, double *b)
{
    if (!E || !A || (E->F->nconn && (!rho || !upstream)) || (E->F->ncells && !b))
        return RSGPU_BAD_ARGUMENT;
    try {
        const rsgpu_flux *F = E->F;
        if (A->n != F->ncells || A->ncols != F->ncells)
            throw RsError{RSGPU_BAD_ARGUMENT, "matrix size differs from the mesh"};
        check_state(E, P, S, true);
        if (P->dt == 0.0)
            throw RsError{RSGPU_BAD_ARGUMENT, "time step must be nonzero"};
        hipStream_t s = E->ctx->stream;
        const ImpesK K = make_k(E, P, S);
        const rsgpu_flux *M = E->Fmf ? E->Fmf : F;   // MF lists read only when thermal
        hipLaunchKernelGGL(k_impes_clear, dim3(1), dim3(1), 0, s, E->err.p);
        if (F->nconn) {
            hipLaunchKernelGGL(k_impes_conn, dim3((F->nconn + 255) / 256), dim3(256), 0, s, K,
                               F->md_ptr.p, F->md_elem.p, F->md_t.p, M->md_ptr.p, M->md_elem.p,
                               M->md_t.p, F->here.p, F->there.p, rho, upstream, E->jh.p,
                               E->jt.p, E->rh.p, E->rt.p);
            RS_CHECK_LAUNCH();
        }
        if (F->ncells) {
            hipLaunchKernelGGL(k_impes_rows, dim3((F->ncells + 255) / 256), dim3(256), 0, s, K,
                               F->cell_ptr.p, F->cell_conn.p, F->cell_sign.p, F->md_ptr.p,
                               F->md_elem.p, E->jh.p, E->jt.p, E->rh.p, E->rt.p, A->rowptr.p,
                               A->colind.p, A->val.p, E->src_ptr.p, E->src_term.p,
                               E->src_kind.p, E->src_q.p, E->src_qe.p, b, E->err.p);
            RS_CHECK_LAUNCH();
        }
        int h = 0;
        RS_HIP(hipMemcpyAsync(&h, E->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        if (h)
            throw RsError{RSGPU_NOT_IN_PATTERN, "Jacobian entry outside the matrix pattern"};
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(E->ctx, e);
    }
}
