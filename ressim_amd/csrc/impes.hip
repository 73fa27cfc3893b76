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
//   k_impes_bf    one lane per cell: its body force (gravity, centrifuge);
//   k_impes_cs    one lane per connection: the CS secondaries
//                 (density average, potential sum over MD, upstream, flux);
//   k_impes_conn  one lane per connection: per MD entry the here/there
//                 Jacobian contributions Ji, Jj and per connection the
//                 residual contributions ri, rj, each in the reference's
//                 loop order (MD entry, phase, component, heat term), so
//                 every Ji, Jj, ri, rj is the Java value bit for bit;
//   k_impes_vals  one lane per entry of A: the sum of its contributions in
//                 connection order, from a per-entry list built once per
//                 matrix pattern (k_map_*: positions by binary search,
//                 counting sort, each entry's short list sorted by MD
//                 index, which is connection order);
//   k_impes_rows  one lane per cell: its b element as the sum of its
//                 connections' ri / rj in ascending connection order, then
//                 the diagonal term, the diagonal entry of A, and the
//                 cell's source terms in application order.
// Only +, -, *, / are used (no fused multiply-add, -ffp-contract=off), so
// A and b equal the reference's loops bit for bit.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "rsgpu_impl.hpp"

struct rsgpu_impes {
    rsgpu_ctx *ctx = nullptr;
    rsgpu_flux *F = nullptr;
    const rsgpu_flux *Fmf = nullptr;
    int ncomp = 0;
    int S = 0;                            // ELL slots of the MD view (0: CSR)
    DevBuf<int> ell_e;
    DevBuf<double> ell_t;
    DevBuf<double> jj, rh, rt;            // per MD entry (here, there) / per connection
    DevBuf<double> bf;                    // per cell body force
    // per-entry contribution lists of the matrix pattern `pattern_id`
    uint64_t pattern_id = 0;
    int map_status = RSGPU_OK;
    DevBuf<int> cptr, clist, diag_pos;
    // the same for the Fourier (MF) lists (temperature system)
    uint64_t mf_pattern_id = 0;
    int mf_map_status = RSGPU_OK;
    DevBuf<int> mf_cptr, mf_clist;
    DevBuf<double> cf, adv;               // component fluxes [ncomp][nconn]; energy flux
    DevBuf<unsigned char> locked;
    DevBuf<int> mm_stat;                  // truncated count, first bad index
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

__global__ void k_impes_bf(ImpesK K, double *__restrict__ bf)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < K.n)
        bf[e] = body_force(K, e);
}

__device__ __forceinline__ double ph_at(const double *a, int n, int ph, int e)
{
    return a[(size_t)ph * n + e];
}

// MD lists as the kernels read them: CSR (md_ptr, md_elem, md_t; any
// length, one entry after another) or slot-major ELL (entry s of
// connection c at s * nconn + c, S slots unrolled, loads coalesced across
// lanes; padding slots hold element `here` and t = 0 and are skipped).
template <int S>
struct MdView {
    static constexpr int slots = S;
    const int *ptr, *elem;
    const double *t;
    int nconn;
    __device__ __forceinline__ int count(int c) const { return ptr[c + 1] - ptr[c]; }
    __device__ __forceinline__ int idx(int c, int s) const
    {
        return S > 0 ? s * nconn + c : ptr[c] + s;
    }
    // the entry's Jacobian slot: the MD index itself (slot-major in ELL, so
    // the connection kernel's stores coalesce; a connection-major order
    // measured slower overall)
    __device__ __forceinline__ int jidx(int c, int s) const { return idx(c, s); }
};

// CS.calculateMassDensity and calculateDarcyFlux for connection c; the three
// phases' potential sums run in one pass over MD (each in MD order).
template <int S>
__global__ void k_impes_cs(ImpesK K, MdView<S> md, const int *__restrict__ here,
                           const int *__restrict__ there, const double *__restrict__ bf,
                           double *__restrict__ rho, int *__restrict__ upstream,
                           double *__restrict__ darcy)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K.nconn)
        return;
    const int n = K.n, nc = K.nconn, h = here[c], t = there[c];
    double r[3];
#pragma unroll
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
    double D[3] = {0, 0, 0};
    const int cnt = md.count(c);
    auto entry = [&](int s) {
        const int q = md.idx(c, s);
        const int ek = md.elem[q];
        const double tk = md.t[q], b = bf[ek];
#pragma unroll
        for (int ph = 0; ph < 3; ++ph) {
            const double psi = ph_at(K.S.phase_pressure, n, ph, ek) + r[ph] * b;
            if (S == 0 || s < cnt)
                D[ph] += tk * psi;
        }
    };
    if (S > 0) {
#pragma unroll
        for (int s = 0; s < S; ++s)
            entry(s);
    } else {
        for (int s = 0; s < cnt; ++s)
            entry(s);
    }
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        const int up = (D[ph] > 0) ? h : t;
        upstream[(size_t)ph * nc + c] = up;
        darcy[(size_t)ph * nc + c] = ph_at(K.S.phase_mobility, n, ph, up) * D[ph];
    }
}

// assemblePressureJacobianFlux for connection c: per MD entry the here/there
// Jacobian contributions (jj at 2 * Jacobian slot + side) and the
// connection's residual contributions.  NC > 0: ncomp ==
// NC with the connection's component mobilities and dRdN in registers.
template <int S, int NC>
__global__ void k_impes_conn(ImpesK K, MdView<S> md, const int *__restrict__ mf_ptr,
                             const int *__restrict__ mf_elem, const double *__restrict__ mf_t,
                             const int *__restrict__ here, const int *__restrict__ there,
                             const double *__restrict__ rho, const int *__restrict__ upstream,
                             const double *__restrict__ bf, double *__restrict__ jj,
                             double *__restrict__ rh, double *__restrict__ rt)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K.nconn)
        return;
    const int n = K.n, nc = K.nconn, ncomp = NC > 0 ? NC : K.ncomp;
    const int i = here[c], j = there[c];
    const bool thermal = K.thermal != 0;
    const double deltaHere = thermal ? K.S.heat_capacity[i] : 0.0;
    const double deltaThere = thermal ? K.S.heat_capacity[j] : 0.0;
    const double dRdT_here = thermal ? K.S.dRdT[i] : 0.0;
    const double dRdT_there = thermal ? K.S.dRdT[j] : 0.0;
    int up[3];
    double r[3], em[3];
    bool pres[3];
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        up[ph] = upstream[(size_t)ph * nc + c];
        r[ph] = rho[(size_t)ph * nc + c];
        pres[ph] = K.S.present[(size_t)ph * n + up[ph]] != 0;
        em[ph] = thermal ? ph_at(K.S.energy_mobility, n, ph, up[ph]) : 0.0;
    }
    constexpr int NR = NC > 0 ? NC : 1;
    double mob[3][NR], dRh[NR], dRt[NR];
    if (NC > 0) {
#pragma unroll
        for (int nu = 0; nu < NR; ++nu) {
            dRh[nu] = K.S.dRdN[(size_t)nu * n + i];
            dRt[nu] = K.S.dRdN[(size_t)nu * n + j];
#pragma unroll
            for (int ph = 0; ph < 3; ++ph)
                mob[ph][nu] = K.S.comp_mobility[((size_t)ph * NR + nu) * n + up[ph]];
        }
    }
    double ri = 0, rj = 0;
    const int cnt = md.count(c);
    auto entry = [&](int s) {
        const int q = md.idx(c, s);
        const int ik = md.elem[q];
        const double tk = md.t[q];
        const double b = bf[ik];
        double psi[3];
#pragma unroll
        for (int ph = 0; ph < 3; ++ph)
            psi[ph] = ph_at(K.S.phase_pressure, n, ph, ik) + r[ph] * b;
        if (S > 0 && s >= cnt)
            return;
        double Ji = 0, Jj = 0;
#pragma unroll
        for (int ph = 0; ph < 3; ++ph) {
            if (!pres[ph])
                continue;
            for (int nu = 0; nu < ncomp; ++nu) {
                const double phaseCompMob =
                    NC > 0 ? mob[ph][nu]
                           : K.S.comp_mobility[((size_t)ph * ncomp + nu) * n + up[ph]];
                const double dRdNhere = NC > 0 ? dRh[nu] : K.S.dRdN[(size_t)nu * n + i];
                const double dRdNthere = NC > 0 ? dRt[nu] : K.S.dRdN[(size_t)nu * n + j];
                const double dNdt = -phaseCompMob * tk * psi[ph];
                const double d2Ndtp = -phaseCompMob * tk;
                ri += dRdNhere * dNdt;
                rj += -dRdNthere * dNdt;
                Ji += dRdNhere * d2Ndtp;
                Jj += -dRdNthere * d2Ndtp;
            }
            if (thermal) {
                const double phaseEnergyMob = em[ph];
                const double dTdt = -phaseEnergyMob * tk * psi[ph];
                const double d2Tdtp = -phaseEnergyMob * tk;
                ri += dRdT_here * dTdt / deltaHere;
                rj += -dRdT_there * dTdt / deltaThere;
                Ji += dRdT_here * d2Tdtp / deltaHere;
                Jj += -dRdT_there * d2Tdtp / deltaThere;
            }
        }
        const size_t w = (size_t)md.jidx(c, s);
        jj[2 * w] = Ji;
        jj[2 * w + 1] = Jj;
    };
    if (S > 0) {
#pragma unroll
        for (int s = 0; s < S; ++s)
            entry(s);
    } else {
        for (int s = 0; s < cnt; ++s)
            entry(s);
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

// slot-major copy of the MD lists (padding: element `here`, t = 0)
template <int S>
__global__ void k_impes_ell(int nconn, const int *__restrict__ md_ptr,
                            const int *__restrict__ md_elem, const double *__restrict__ md_t,
                            const int *__restrict__ here, int *__restrict__ ell_e,
                            double *__restrict__ ell_t)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nconn)
        return;
    const int q0 = md_ptr[c], cnt = md_ptr[c + 1] - q0;
    for (int s = 0; s < S; ++s) {
        const bool v = s < cnt;
        ell_e[(size_t)s * nconn + c] = v ? md_elem[q0 + s] : here[c];
        ell_t[(size_t)s * nconn + c] = v ? md_t[q0 + s] : 0.0;
    }
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

// ---- per-entry contribution lists (once per matrix pattern)
// pos[2 v + side] = the position in A of entry v's element in row here[c]
// (side 0) / there[c] (side 1), v the entry's Jacobian slot; -1 for ELL
// padding
template <int S>
__global__ void k_map_count(MdView<S> md, const int *__restrict__ md_ptr,
                            const int *__restrict__ md_elem, const int *__restrict__ here,
                            const int *__restrict__ there, const int *__restrict__ rowptr,
                            const int *__restrict__ colind, int *__restrict__ pos,
                            int *__restrict__ cnt, int *__restrict__ err)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= md.nconn)
        return;
    const int rows[2] = {here[c], there[c]};
    const int q0 = md_ptr[c], m = md_ptr[c + 1] - q0;
    for (int s = 0; s < (S > 0 ? S : m); ++s) {
        const size_t v = (size_t)md.jidx(c, s);
        if (s >= m) {
            pos[2 * v] = pos[2 * v + 1] = -1;
            continue;
        }
        const int k = md_elem[q0 + s];
        for (int side = 0; side < 2; ++side) {
            const int r = rows[side], r0 = rowptr[r], r1 = rowptr[r + 1];
            const int p = find_col(colind, r0, r1, k);
            if (p >= r1 || colind[p] != k) {
                atomicOr(err, 1);
                pos[2 * v + side] = -1;
                continue;
            }
            pos[2 * v + side] = p;
            atomicAdd(&cnt[p], 1);
        }
    }
}

__global__ void k_map_fill(size_t nitems, const int *__restrict__ pos,
                           const int *__restrict__ cptr, int *__restrict__ fill,
                           int *__restrict__ clist)
{
    const size_t it = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitems)
        return;
    const int p = pos[it];
    if (p < 0)
        return;
    clist[cptr[p] + atomicAdd(&fill[p], 1)] = (int)it;
}

// each entry's list in connection order (the item's connection: item / 2
// in CSR layout, (item / 2) mod nconn in slot-major layout)
__global__ void k_map_sort(int nnz, int ell_nconn, const int *__restrict__ cptr,
                           int *__restrict__ clist)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz)
        return;
    auto key = [&](int it) { return ell_nconn > 0 ? (it >> 1) % ell_nconn : it; };
    const int a = cptr[p], e = cptr[p + 1];
    for (int i = a + 1; i < e; ++i) {
        const int v = clist[i], kv = key(v);
        int j = i - 1;
        while (j >= a && key(clist[j]) > kv) {
            clist[j + 1] = clist[j];
            --j;
        }
        clist[j + 1] = v;
    }
}

__global__ void k_map_diag(int n, const int *__restrict__ rowptr,
                           const int *__restrict__ colind, int *__restrict__ dpos,
                           int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const int r0 = rowptr[r], r1 = rowptr[r + 1];
    const int p = find_col(colind, r0, r1, r);
    if (p >= r1 || colind[p] != r) {
        atomicOr(err, 1);
        dpos[r] = -1;
    } else {
        dpos[r] = p;
    }
}

// ---- assembly
__global__ void k_impes_vals(int nnz, const int *__restrict__ cptr,
                             const int *__restrict__ clist, const double *__restrict__ jj,
                             double *__restrict__ val)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz)
        return;
    double v = 0;
    for (int k = cptr[p]; k < cptr[p + 1]; ++k)
        v += jj[clist[k]];
    val[p] = v;
}

__global__ void k_impes_rows(ImpesK K, const int *__restrict__ cell_ptr,
                             const int *__restrict__ cell_conn,
                             const signed char *__restrict__ cell_sign,
                             const double *__restrict__ rh, const double *__restrict__ rt,
                             const int *__restrict__ dpos, double *__restrict__ val,
                             const int *__restrict__ src_ptr, const int *__restrict__ src_term,
                             const int *__restrict__ src_kind,
                             const double *__restrict__ src_q,
                             const double *__restrict__ src_qe, double *__restrict__ bvec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= K.n)
        return;
    const int n = K.n, ncomp = K.ncomp;
    double acc = 0;
    // flux part: incident connections in ascending order
    for (int a = cell_ptr[r]; a < cell_ptr[r + 1]; ++a) {
        const int c = cell_conn[a];
        acc += cell_sign[a] > 0 ? rh[c] : rt[c];
    }
    // diagonal part
    const double dp = K.S.pressure[r] - K.S.p0[r];
    const double dRdp = K.S.dRdp[r];
    acc += (K.S.R0[r] + dRdp * dp) / K.dt;
    val[dpos[r]] += dRdp / K.dt;
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

// ---- explicit molar-mass update (MolarMassExplicit, Discretisation.java:768-919)
// component fluxes of connection c: sum over present upstream phases of
// (mole fraction x molar density) x Darcy flux, phases in order
__global__ void k_mm_flux(ImpesK K, const double *__restrict__ darcy,
                          const int *__restrict__ upstream, double *__restrict__ cf)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K.nconn)
        return;
    const int n = K.n, nc = K.nconn, ncomp = K.ncomp;
    for (int nu = 0; nu < ncomp; ++nu) {
        double f = 0;
        for (int ph = 0; ph < 3; ++ph) {
            const int up = upstream[(size_t)ph * nc + c];
            if (!K.S.present[(size_t)ph * n + up])
                continue;
            const double D = darcy[(size_t)ph * nc + c];
            const double xi = ph_at(K.S.molar_density, n, ph, up);
            const double C = K.S.mole_fraction[((size_t)ph * ncomp + nu) * n + up];
            f += C * xi * D;
        }
        cf[(size_t)nu * nc + c] = f;
    }
}

// dN of cell r per component: its connections in ascending order (here:
// -= dt f, there: += dt f), then its source terms; locked cells keep N;
// N_new = N_old + dN, negatives truncated or reported (stat[0]: count,
// stat[1]: first cell * ncomp + component)
__global__ void k_mm_rows(ImpesK K, const int *__restrict__ cell_ptr,
                          const int *__restrict__ cell_conn,
                          const signed char *__restrict__ cell_sign,
                          const double *__restrict__ cf, const int *__restrict__ src_ptr,
                          const int *__restrict__ src_term, const int *__restrict__ src_kind,
                          const double *__restrict__ src_q,
                          const unsigned char *__restrict__ locked, int truncate,
                          const double *__restrict__ N_old, double *__restrict__ N_new,
                          int *stat)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= K.n)
        return;
    const int n = K.n, nc = K.nconn, ncomp = K.ncomp;
    const double dt = K.dt;
    for (int nu = 0; nu < ncomp; ++nu) {
        double dN = 0;
        for (int a = cell_ptr[r]; a < cell_ptr[r + 1]; ++a) {
            const double f = cf[(size_t)nu * nc + cell_conn[a]];
            if (cell_sign[a] > 0)
                dN -= dt * f;
            else
                dN += dt * f;
        }
        for (int a = src_ptr[r]; a < src_ptr[r + 1]; ++a) {
            const int t = src_term[a];
            const double *q = src_q + (size_t)t * 3 * ncomp;
            if (src_kind[t] == 0)
                dN += dt * q[nu];
            else
                for (int ph = 0; ph < 3; ++ph)
                    dN += dt * q[ph * ncomp + nu];
        }
        if (locked[r])
            dN = 0;
        double Nn = N_old[(size_t)nu * n + r] + dN;
        if (Nn < 0) {
            if (truncate) {
                Nn = 0;
                atomicAdd(&stat[0], 1);
            } else {
                atomicMin(&stat[1], r * ncomp + nu);
            }
        }
        N_new[(size_t)nu * n + r] = Nn;
    }
}

// ---- temperature system (TemperatureNewtonRaphson, Discretisation.java:577-752)
// advective energy flux of connection c: phases in order, MD entries inside
template <int S>
__global__ void k_temp_adv(ImpesK K, MdView<S> md, const double *__restrict__ rho,
                           const int *__restrict__ upstream, const double *__restrict__ bf,
                           double *__restrict__ adv)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K.nconn)
        return;
    const int n = K.n, nc = K.nconn;
    const int cnt = md.count(c);
    double energyFlux = 0;
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        const int up = upstream[(size_t)ph * nc + c];
        if (!K.S.present[(size_t)ph * n + up])
            continue;
        const double energyMob = ph_at(K.S.energy_mobility, n, ph, up);
        const double r = rho[(size_t)ph * nc + c];
        auto entry = [&](int s) {
            const int q = md.idx(c, s);
            const int ek = md.elem[q];
            const double psi = ph_at(K.S.phase_pressure, n, ph, ek) + r * bf[ek];
            if (S == 0 || s < cnt)
                energyFlux += energyMob * md.t[q] * psi;
        };
        if (S > 0) {
#pragma unroll
            for (int s = 0; s < S; ++s)
                entry(s);
        } else {
            for (int s = 0; s < cnt; ++s)
                entry(s);
        }
    }
    adv[c] = energyFlux;
}

// A's entries from the Fourier lists: +t (here row), -t (there row) in
// connection order
__global__ void k_temp_vals(int nnz, const int *__restrict__ cptr,
                            const int *__restrict__ clist, const double *__restrict__ mf_t,
                            double *__restrict__ val)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz)
        return;
    double v = 0;
    for (int k = cptr[p]; k < cptr[p + 1]; ++k) {
        const int it = clist[k];
        const double t = mf_t[it >> 1];
        v += (it & 1) ? -t : t;
    }
    val[p] = v;
}

// b of cell r: advective fluxes of its connections, then their Fourier
// terms, then the diagonal term (and A's diagonal entry), then -qe of its
// source terms
__global__ void k_temp_rows(ImpesK K, const int *__restrict__ cell_ptr,
                            const int *__restrict__ cell_conn,
                            const signed char *__restrict__ cell_sign,
                            const double *__restrict__ adv, const int *__restrict__ mf_ptr,
                            const int *__restrict__ mf_elem, const double *__restrict__ mf_t,
                            const int *__restrict__ dpos, double *__restrict__ val,
                            const int *__restrict__ src_ptr, const int *__restrict__ src_term,
                            const double *__restrict__ src_qe, double *__restrict__ bvec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= K.n)
        return;
    double acc = 0;
    for (int a = cell_ptr[r]; a < cell_ptr[r + 1]; ++a) {
        const double e = adv[cell_conn[a]];
        acc += cell_sign[a] > 0 ? e : -e;
    }
    for (int a = cell_ptr[r]; a < cell_ptr[r + 1]; ++a) {
        const int c = cell_conn[a];
        const bool h = cell_sign[a] > 0;
        for (int q = mf_ptr[c]; q < mf_ptr[c + 1]; ++q) {
            const double tk = mf_t[q];
            const double T = K.S.temperature[mf_elem[q]];
            acc += h ? tk * T : -tk * T;
        }
    }
    const double delta = K.S.heat_capacity[r];
    const double dT = K.S.temperature[r] - K.S.T0[r];
    acc += delta * dT / K.dt;
    val[dpos[r]] += delta / K.dt;
    for (int a = src_ptr[r]; a < src_ptr[r + 1]; ++a)
        acc += -src_qe[src_term[a]];
    bvec[r] = acc;
}

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

template <int S>
MdView<S> md_view(const rsgpu_impes *E)
{
    const rsgpu_flux *F = E->F;
    MdView<S> v;
    v.ptr = F->md_ptr.p;
    v.elem = S > 0 ? E->ell_e.p : F->md_elem.p;
    v.t = S > 0 ? E->ell_t.p : F->md_t.p;
    v.nconn = F->nconn;
    return v;
}

// f(view) with the view type of E's layout
template <class Fn>
void with_view(const rsgpu_impes *E, Fn f)
{
    if (E->S == 6)
        f(md_view<6>(E));
    else if (E->S == 18)
        f(md_view<18>(E));
    else
        f(md_view<0>(E));
}

// Per-entry contribution lists of A's pattern: for entry p of row r, the
// indices 2v + side of the MD entries (here side 0, there side 1; v the
// entry's Jacobian slot) whose element is p's column, over r's incident
// connections, in connection order.
void build_map(rsgpu_impes *E, const rsgpu_csr *A)
{
    const rsgpu_flux *F = E->F;
    hipStream_t s = E->ctx->stream;
    const int nnz = A->nnz, n = F->ncells;
    const size_t nit = E->jj.n;   // 2 x view entries
    DevBuf<int> pos, cnt, fill;
    pos.alloc(std::max(nit, (size_t)1));
    cnt.alloc((size_t)nnz + 1);
    fill.alloc(std::max(nnz, 1));
    cnt.zero(s);
    fill.zero(s);
    E->cptr.alloc((size_t)nnz + 1);
    E->clist.alloc(std::max(nit, (size_t)1));
    E->diag_pos.alloc(std::max(n, 1));
    hipLaunchKernelGGL(k_impes_clear, dim3(1), dim3(1), 0, s, E->err.p);
    if (F->nconn)
        with_view(E, [&](auto v) {
            hipLaunchKernelGGL(k_map_count<decltype(v)::slots>, dim3((F->nconn + 255) / 256),
                               dim3(256), 0, s, v, F->md_ptr.p, F->md_elem.p, F->here.p,
                               F->there.p, A->rowptr.p, A->colind.p, pos.p, cnt.p, E->err.p);
        });
    RS_CHECK_LAUNCH();
    size_t tmp = 0;
    RS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt.p, E->cptr.p, nnz + 1, s));
    DevBuf<unsigned char> work;
    work.alloc(std::max(tmp, (size_t)1));
    RS_HIP(hipcub::DeviceScan::ExclusiveSum(work.p, tmp, cnt.p, E->cptr.p, nnz + 1, s));
    if (nit)
        hipLaunchKernelGGL(k_map_fill, dim3((unsigned)((nit + 255) / 256)), dim3(256), 0, s,
                           nit, pos.p, E->cptr.p, fill.p, E->clist.p);
    if (nnz)
        hipLaunchKernelGGL(k_map_sort, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz,
                           E->S > 0 ? F->nconn : 0, E->cptr.p, E->clist.p);
    if (n)
        hipLaunchKernelGGL(k_map_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, A->rowptr.p,
                           A->colind.p, E->diag_pos.p, E->err.p);
    RS_CHECK_LAUNCH();
    int h = 0;
    RS_HIP(hipMemcpyAsync(&h, E->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    E->map_status = h ? RSGPU_NOT_IN_PATTERN : RSGPU_OK;
    E->pattern_id = A->pattern_id;
}

// The per-entry lists of the Fourier (MF) lists on A's pattern (CSR view).
void build_map_mf(rsgpu_impes *E, const rsgpu_csr *A)
{
    const rsgpu_flux *F = E->F, *Mf = E->Fmf;
    hipStream_t s = E->ctx->stream;
    const int nnz = A->nnz;
    const size_t nit = 2 * (size_t)Mf->nentries;
    DevBuf<int> pos, cnt, fill;
    pos.alloc(std::max(nit, (size_t)1));
    cnt.alloc((size_t)nnz + 1);
    fill.alloc(std::max(nnz, 1));
    cnt.zero(s);
    fill.zero(s);
    E->mf_cptr.alloc((size_t)nnz + 1);
    E->mf_clist.alloc(std::max(nit, (size_t)1));
    hipLaunchKernelGGL(k_impes_clear, dim3(1), dim3(1), 0, s, E->err.p);
    MdView<0> v;
    v.ptr = Mf->md_ptr.p;
    v.elem = Mf->md_elem.p;
    v.t = Mf->md_t.p;
    v.nconn = Mf->nconn;
    if (Mf->nconn)
        hipLaunchKernelGGL(k_map_count<0>, dim3((Mf->nconn + 255) / 256), dim3(256), 0, s, v,
                           Mf->md_ptr.p, Mf->md_elem.p, F->here.p, F->there.p, A->rowptr.p,
                           A->colind.p, pos.p, cnt.p, E->err.p);
    RS_CHECK_LAUNCH();
    size_t tmp = 0;
    RS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt.p, E->mf_cptr.p, nnz + 1, s));
    DevBuf<unsigned char> work;
    work.alloc(std::max(tmp, (size_t)1));
    RS_HIP(hipcub::DeviceScan::ExclusiveSum(work.p, tmp, cnt.p, E->mf_cptr.p, nnz + 1, s));
    if (nit)
        hipLaunchKernelGGL(k_map_fill, dim3((unsigned)((nit + 255) / 256)), dim3(256), 0, s,
                           nit, pos.p, E->mf_cptr.p, fill.p, E->mf_clist.p);
    if (nnz)
        hipLaunchKernelGGL(k_map_sort, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, 0,
                           E->mf_cptr.p, E->mf_clist.p);
    RS_CHECK_LAUNCH();
    int h = 0;
    RS_HIP(hipMemcpyAsync(&h, E->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    E->mf_map_status = h ? RSGPU_NOT_IN_PATTERN : RSGPU_OK;
    E->mf_pattern_id = A->pattern_id;
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
        // slot-major copy of the MD lists when every list is short
        int maxlen = 0;
        if (F->nconn) {
            std::vector<int> mp(F->nconn + 1);
            RS_HIP(hipMemcpy(mp.data(), F->md_ptr.p, sizeof(int) * mp.size(),
                             hipMemcpyDeviceToHost));
            for (int c = 0; c < F->nconn; ++c)
                maxlen = std::max(maxlen, mp[c + 1] - mp[c]);
        }
        E->S = getenv("RSGPU_IMPES_CSR") ? 0 : maxlen <= 6 ? 6 : maxlen <= 18 ? 18 : 0;
        if (F->nconn == 0)
            E->S = 0;
        if (E->S > 0) {
            E->ell_e.alloc((size_t)E->S * F->nconn);
            E->ell_t.alloc((size_t)E->S * F->nconn);
            hipLaunchKernelGGL(E->S == 6 ? k_impes_ell<6> : k_impes_ell<18>,
                               dim3((F->nconn + 255) / 256), dim3(256), 0, F->ctx->stream,
                               F->nconn, F->md_ptr.p, F->md_elem.p, F->md_t.p, F->here.p,
                               E->ell_e.p, E->ell_t.p);
            RS_CHECK_LAUNCH();
            E->jj.alloc(2 * (size_t)E->S * F->nconn);
        } else {
            E->jj.alloc(2 * (size_t)std::max(F->nentries, 1));
        }
        E->bf.alloc(std::max(F->ncells, 1));
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
            const ImpesK K = make_k(E, P, S);
            hipLaunchKernelGGL(k_impes_bf, dim3((F->ncells + 255) / 256), dim3(256), 0,
                               E->ctx->stream, K, E->bf.p);
            with_view(E, [&](auto v) {
                hipLaunchKernelGGL(k_impes_cs<decltype(v)::slots>, dim3((F->nconn + 255) / 256),
                                   dim3(256), 0, E->ctx->stream, K, v, F->here.p, F->there.p,
                                   E->bf.p, rho, upstream, darcy);
            });
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
        if (E->pattern_id != A->pattern_id)
            build_map(E, A);
        if (E->map_status != RSGPU_OK)
            throw RsError{E->map_status, "Jacobian entry outside the matrix pattern"};
        const ImpesK K = make_k(E, P, S);
        const rsgpu_flux *M = E->Fmf ? E->Fmf : F;   // MF lists read only when thermal
        if (F->ncells) {
            hipLaunchKernelGGL(k_impes_bf, dim3((F->ncells + 255) / 256), dim3(256), 0, s, K,
                               E->bf.p);
            RS_CHECK_LAUNCH();
        }
        if (F->nconn) {
            with_view(E, [&](auto v) {
                constexpr int SL = decltype(v)::slots;
                const dim3 g((F->nconn + 255) / 256), bl(256);
                auto go = [&](auto kern) {
                    hipLaunchKernelGGL(kern, g, bl, 0, s, K, v, M->md_ptr.p, M->md_elem.p,
                                       M->md_t.p, F->here.p, F->there.p, rho, upstream,
                                       E->bf.p, E->jj.p, E->rh.p, E->rt.p);
                };
                switch (E->ncomp) {
                case 1:
                    go(k_impes_conn<SL, 1>);
                    break;
                case 2:
                    go(k_impes_conn<SL, 2>);
                    break;
                case 3:
                    go(k_impes_conn<SL, 3>);
                    break;
                case 4:
                    go(k_impes_conn<SL, 4>);
                    break;
                default:
                    go(k_impes_conn<SL, 0>);
                    break;
                }
            });
            RS_CHECK_LAUNCH();
        }
        if (A->nnz) {
            hipLaunchKernelGGL(k_impes_vals, dim3((A->nnz + 255) / 256), dim3(256), 0, s,
                               A->nnz, E->cptr.p, E->clist.p, E->jj.p, A->val.p);
            RS_CHECK_LAUNCH();
        }
        if (F->ncells) {
            hipLaunchKernelGGL(k_impes_rows, dim3((F->ncells + 255) / 256), dim3(256), 0, s, K,
                               F->cell_ptr.p, F->cell_conn.p, F->cell_sign.p, E->rh.p, E->rt.p,
                               E->diag_pos.p, A->val.p, E->src_ptr.p, E->src_term.p,
                               E->src_kind.p, E->src_q.p, E->src_qe.p, b);
            RS_CHECK_LAUNCH();
        }
        A->values_changed();
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(E->ctx, e);
    }
}

extern "C" int rsgpu_impes_molar_update(rsgpu_impes *E, const rsgpu_impes_params *P,
                                        const rsgpu_impes_state *S, const double *darcy,
                                        const int32_t *upstream, int nlocked,
                                        const int32_t *locked_host, int truncate,
                                        const double *N_old, double *N_new, int *truncated,
                                        int *bad)
{
    if (!E || !P || !S || nlocked < 0 || (nlocked && !locked_host) ||
        (E->F->nconn && (!darcy || !upstream)) || (E->F->ncells && (!N_old || !N_new)))
        return RSGPU_BAD_ARGUMENT;
    try {
        const rsgpu_flux *F = E->F;
        const int n = F->ncells;
        if (P->dt == 0.0)
            throw RsError{RSGPU_BAD_ARGUMENT, "time step must be nonzero"};
        if (n && (!S->present || !S->mole_fraction || !S->molar_density))
            throw RsError{RSGPU_BAD_ARGUMENT, "missing control-volume arrays"};
        hipStream_t s = E->ctx->stream;
        std::vector<unsigned char> lk(std::max(n, 1), 0);
        for (int t = 0; t < nlocked; ++t) {
            if (locked_host[t] < 0 || locked_host[t] >= n)
                throw RsError{RSGPU_BAD_ARGUMENT, "locked cell out of range"};
            lk[locked_host[t]] = 1;
        }
        E->locked.upload(lk, s);
        E->cf.alloc((size_t)std::max(E->ncomp * F->nconn, 1));
        E->mm_stat.alloc(2);
        const int init[2] = {0, 0x7fffffff};
        RS_HIP(hipMemcpyAsync(E->mm_stat.p, init, sizeof(init), hipMemcpyHostToDevice, s));
        const ImpesK K = make_k(E, P, S);
        if (F->nconn)
            hipLaunchKernelGGL(k_mm_flux, dim3((F->nconn + 255) / 256), dim3(256), 0, s, K,
                               darcy, upstream, E->cf.p);
        if (n)
            hipLaunchKernelGGL(k_mm_rows, dim3((n + 255) / 256), dim3(256), 0, s, K,
                               F->cell_ptr.p, F->cell_conn.p, F->cell_sign.p, E->cf.p,
                               E->src_ptr.p, E->src_term.p, E->src_kind.p, E->src_q.p,
                               E->locked.p, truncate ? 1 : 0, N_old, N_new, E->mm_stat.p);
        RS_CHECK_LAUNCH();
        int h[2];
        RS_HIP(hipMemcpyAsync(h, E->mm_stat.p, sizeof(h), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        if (truncated)
            *truncated = h[0];
        if (bad)
            *bad = h[1] == 0x7fffffff ? -1 : h[1];
        if (h[1] != 0x7fffffff) {
            const int e = h[1] / E->ncomp, nu = h[1] % E->ncomp;
            E->ctx->err = "Negative number of moles of component " + std::to_string(nu) +
                          " in element " + std::to_string(e + 1);
            return RSGPU_BAD_ARGUMENT;
        }
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(E->ctx, e);
    }
}

extern "C" int rsgpu_impes_assemble_temperature(rsgpu_impes *E, rsgpu_csr *A,
                                                const rsgpu_impes_params *P,
                                                const rsgpu_impes_state *S, const double *rho,
                                                const int32_t *upstream, double *b)
{
    if (!E || !A || !P || !S || (E->F->nconn && (!rho || !upstream)) || (E->F->ncells && !b))
        return RSGPU_BAD_ARGUMENT;
    try {
        const rsgpu_flux *F = E->F;
        const int n = F->ncells;
        if (!E->Fmf)
            throw RsError{RSGPU_BAD_ARGUMENT, "the temperature system needs the Fourier lists"};
        if (A->n != n || A->ncols != n)
            throw RsError{RSGPU_BAD_ARGUMENT, "matrix size differs from the mesh"};
        if (P->dt == 0.0)
            throw RsError{RSGPU_BAD_ARGUMENT, "time step must be nonzero"};
        if (n && (!S->present || !S->energy_mobility || !S->phase_pressure || !S->temperature ||
                  !S->heat_capacity || !S->T0 || !S->cx || !S->cy || !S->cz))
            throw RsError{RSGPU_BAD_ARGUMENT, "missing control-volume arrays"};
        hipStream_t s = E->ctx->stream;
        if (E->pattern_id != A->pattern_id)
            build_map(E, A);
        if (E->mf_pattern_id != A->pattern_id)
            build_map_mf(E, A);
        if (E->map_status != RSGPU_OK || E->mf_map_status != RSGPU_OK)
            throw RsError{RSGPU_NOT_IN_PATTERN, "Jacobian entry outside the matrix pattern"};
        const ImpesK K = make_k(E, P, S);
        const rsgpu_flux *Mf = E->Fmf;
        E->adv.alloc(std::max(F->nconn, 1));
        if (n) {
            hipLaunchKernelGGL(k_impes_bf, dim3((n + 255) / 256), dim3(256), 0, s, K, E->bf.p);
            RS_CHECK_LAUNCH();
        }
        if (F->nconn)
            with_view(E, [&](auto v) {
                hipLaunchKernelGGL(k_temp_adv<decltype(v)::slots>, dim3((F->nconn + 255) / 256),
                                   dim3(256), 0, s, K, v, rho, upstream, E->bf.p, E->adv.p);
            });
        RS_CHECK_LAUNCH();
        if (A->nnz)
            hipLaunchKernelGGL(k_temp_vals, dim3((A->nnz + 255) / 256), dim3(256), 0, s, A->nnz,
                               E->mf_cptr.p, E->mf_clist.p, Mf->md_t.p, A->val.p);
        if (n)
            hipLaunchKernelGGL(k_temp_rows, dim3((n + 255) / 256), dim3(256), 0, s, K,
                               F->cell_ptr.p, F->cell_conn.p, F->cell_sign.p, E->adv.p,
                               Mf->md_ptr.p, Mf->md_elem.p, Mf->md_t.p, E->diag_pos.p, A->val.p,
                               E->src_ptr.p, E->src_term.p, E->src_qe.p, b);
        RS_CHECK_LAUNCH();
        A->values_changed();
        return RSGPU_OK;
    } catch (const RsError &e) {
        return impes_fail(E->ctx, e);
    }
}
