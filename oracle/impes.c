/*
 * CPU oracle, TEST INFRASTRUCTURE ONLY: the IMPES pressure Jacobian of rs.
 *
 * Restates, statement by statement and in the reference's loop order:
 *   rs/src/no/uib/cipr/rs/field/CS.java:94-170   (control-surface
 *     secondaries: saturation-weighted phase density, Darcy phase flux and
 *     upstream element, body force with gravity and centrifuge terms)
 *   rs/src/no/uib/cipr/rs/numerics/Discretisation.java:284-289 (the Newton
 *     step assembles the flux part, then the diagonal part)
 *   Discretisation.java:365-453 (assemblePressureJacobianDiagonal and the
 *     regular / outlet source terms)
 *   Discretisation.java:459-561 (assemblePressureJacobianFlux)
 *   rs/src/no/uib/cipr/rs/numerics/LinearSolver.java:226-235, 287-288
 *     (addToJacobian = CompRowMatrix.add, addToResidual = DenseVector.add;
 *     A and b are zero when assembly starts)
 * The control-volume secondaries (CV: flash, PVT and rock-fluid results)
 * are inputs, as rs computes them before the assembly.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* CS.bodyForce (CS.java:161-170) */
static double body_force(const or_impes_params *P, const or_impes_state *S, int e)
{
    const double x = S->cx[e], y = S->cy[e], z = S->cz[e];
    const double distance = (P->ax - x) * (P->ax - x) + (P->ay - y) * (P->ay - y);
    return P->g * z - (P->omega * P->omega / 2) * distance;
}

#define PH(a, ph, e) ((a)[(size_t)(ph) * n + (e)])

void or_impes_cs(const or_impes_params *P, const or_impes_state *S, const int *here,
                 const int *there, const int *md_ptr, const int *md_elem, const double *md_t,
                 double *rho, int *upstream, double *darcy)
{
    const int n = P->n, nc = P->nconn;
    for (int c = 0; c < nc; ++c) {
        const int h = here[c], t = there[c];
        /* calculateMassDensity (:107-126) */
        for (int ph = 0; ph < 3; ++ph) {
            rho[(size_t)ph * nc + c] = 0;
            const double Sh = PH(S->saturation, ph, h), St = PH(S->saturation, ph, t);
            const double Ssum = Sh + St;
            const double rh = PH(S->mass_density, ph, h), rt = PH(S->mass_density, ph, t);
            const double av = (Sh * rh + St * rt) / Ssum;
            if (!isnan(av))
                rho[(size_t)ph * nc + c] = av;
        }
        /* calculateDarcyFlux (:131-155) */
        for (int ph = 0; ph < 3; ++ph) {
            double D = 0;
            for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
                const int ek = md_elem[q];
                const double tk = md_t[q];
                const double psi = PH(S->phase_pressure, ph, ek) +
                                   rho[(size_t)ph * nc + c] * body_force(P, S, ek);
                D += tk * psi;
            }
            const int up = (D > 0) ? h : t;
            upstream[(size_t)ph * nc + c] = up;
            darcy[(size_t)ph * nc + c] = PH(S->phase_mobility, ph, up) * D;
        }
    }
}

int or_impes_assemble(const or_impes_params *P, const or_impes_state *S, const int *here,
                      const int *there, const int *md_ptr, const int *md_elem,
                      const double *md_t, const int *mf_ptr, const int *mf_elem,
                      const double *mf_t, const double *rho, const int *upstream, int nterms,
                      const int *src_elem, const int *src_kind, const double *src_q,
                      const double *src_qe, const int *rowptr, const int *colind, double *val,
                      double *b)
{
    const int n = P->n, nc = P->nconn, ncomp = P->ncomp, thermal = P->thermal;
    memset(val, 0, sizeof(double) * (size_t)rowptr[n]);
    memset(b, 0, sizeof(double) * (size_t)n);
    /* assemblePressureJacobianFlux (:459-561) */
    for (int c = 0; c < nc; ++c) {
        const int i = here[c], j = there[c];
        const double deltaHere = thermal ? S->heat_capacity[i] : 0;
        const double deltaThere = thermal ? S->heat_capacity[j] : 0;
        const double dRdT_here = thermal ? S->dRdT[i] : 0;
        const double dRdT_there = thermal ? S->dRdT[j] : 0;
        double ri = 0, rj = 0;
        for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
            const int ik = md_elem[q];
            const double tk = md_t[q];
            const double bf = body_force(P, S, ik);
            double Ji = 0, Jj = 0;
            for (int ph = 0; ph < 3; ++ph) {
                const int up = upstream[(size_t)ph * nc + c];
                if (!PH(S->present, ph, up))
                    continue;
                const double psi = PH(S->phase_pressure, ph, ik) + rho[(size_t)ph * nc + c] * bf;
                for (int nu = 0; nu < ncomp; ++nu) {
                    const double phaseCompMob =
                        S->comp_mobility[((size_t)ph * ncomp + nu) * n + up];
                    const double dRdNhere = S->dRdN[(size_t)nu * n + i];
                    const double dRdNthere = S->dRdN[(size_t)nu * n + j];
                    const double dNdt = -phaseCompMob * tk * psi;
                    const double d2Ndtp = -phaseCompMob * tk;
                    ri += dRdNhere * dNdt;
                    rj += -dRdNthere * dNdt;
                    Ji += dRdNhere * d2Ndtp;
                    Jj += -dRdNthere * d2Ndtp;
                }
                if (thermal) {
                    const double phaseEnergyMob = PH(S->energy_mobility, ph, up);
                    const double dTdt = -phaseEnergyMob * tk * psi;
                    const double d2Tdtp = -phaseEnergyMob * tk;
                    ri += dRdT_here * dTdt / deltaHere;
                    rj += -dRdT_there * dTdt / deltaThere;
                    Ji += dRdT_here * d2Tdtp / deltaHere;
                    Jj += -dRdT_there * d2Tdtp / deltaThere;
                }
            }
            if (or_csr_add(rowptr, colind, val, i, ik, Ji) ||
                or_csr_add(rowptr, colind, val, j, ik, Jj))
                return OR_NOT_IN_PATTERN;
        }
        if (thermal)
            for (int q = mf_ptr[c]; q < mf_ptr[c + 1]; ++q) {
                const int ek = mf_elem[q];
                const double F = mf_t[q] * S->temperature[ek];
                ri += -dRdT_here * F / deltaHere;
                rj += dRdT_there * F / deltaThere;
            }
        b[i] += ri;
        b[j] += rj;
    }
    /* assemblePressureJacobianDiagonal (:365-400) */
    for (int i = 0; i < n; ++i) {
        const double dp = S->pressure[i] - S->p0[i];
        const double dRdp = S->dRdp[i];
        b[i] += (S->R0[i] + dRdp * dp) / P->dt;
        if (or_csr_add(rowptr, colind, val, i, i, dRdp / P->dt))
            return OR_NOT_IN_PATTERN;
    }
    /* sources in field order, each over its elements: regular (:405-424)
     * and outlet (:429-452) */
    for (int s = 0; s < nterms; ++s) {
        const int e = src_elem[s];
        const double *q = src_q + (size_t)s * 3 * ncomp;
        for (int nu = 0; nu < ncomp; ++nu) {
            double qnu;
            if (src_kind[s] == 0) {
                qnu = q[nu];
            } else {
                qnu = 0;
                for (int ph = 0; ph < 3; ++ph)
                    qnu += q[ph * ncomp + nu];
            }
            const double dRdN = S->dRdN[(size_t)nu * n + e];
            b[e] += dRdN * qnu;
        }
        if (thermal) {
            const double delta = S->heat_capacity[e];
            const double dRdT = S->dRdT[e];
            b[e] += dRdT * src_qe[s] / delta;
        }
    }
    return OR_OK;
}

/* MolarMassExplicit.solve (Discretisation.java:768-919): dN from the
 * connection fluxes (calculateMolarMassFlux, :842-879), then the sources
 * (:884-919), zero in locked cells, N_new = N_old + dN with the negative
 * check of updateMolarMass (:803-837).  bad: first element * ncomp +
 * component (elements, then components, ascending) with N_new < 0. */
int or_impes_molar_update(const or_impes_params *P, const or_impes_state *S, const int *here,
                          const int *there, const double *darcy, const int *upstream,
                          int nterms, const int *src_elem, const int *src_kind,
                          const double *src_q, const unsigned char *locked, int truncate,
                          const double *N_old, double *N_new, int *truncated, int *bad)
{
    const int n = P->n, nc = P->nconn, ncomp = P->ncomp;
    const double dt = P->dt;
    double *dN = (double *)calloc((size_t)ncomp * n + 1, sizeof(double));
    double *cf = (double *)calloc((size_t)ncomp + 1, sizeof(double));
    for (int c = 0; c < nc; ++c) {
        const int i = here[c], j = there[c];
        for (int nu = 0; nu < ncomp; ++nu)
            cf[nu] = 0;
        for (int ph = 0; ph < 3; ++ph) {
            const int up = upstream[(size_t)ph * nc + c];
            if (!PH(S->present, ph, up))
                continue;
            const double D = darcy[(size_t)ph * nc + c];
            const double xi = PH(S->molar_density, ph, up);
            for (int nu = 0; nu < ncomp; ++nu) {
                const double C = S->mole_fraction[((size_t)ph * ncomp + nu) * n + up];
                cf[nu] += C * xi * D;
            }
        }
        for (int nu = 0; nu < ncomp; ++nu) {
            dN[(size_t)nu * n + i] -= dt * cf[nu];
            dN[(size_t)nu * n + j] += dt * cf[nu];
        }
    }
    for (int t = 0; t < nterms; ++t) {
        const int e = src_elem[t];
        const double *q = src_q + (size_t)t * 3 * ncomp;
        for (int nu = 0; nu < ncomp; ++nu) {
            if (src_kind[t] == 0)
                dN[(size_t)nu * n + e] += dt * q[nu];
            else
                for (int ph = 0; ph < 3; ++ph)
                    dN[(size_t)nu * n + e] += dt * q[ph * ncomp + nu];
        }
    }
    for (int e = 0; e < n; ++e)
        if (locked && locked[e])
            for (int nu = 0; nu < ncomp; ++nu)
                dN[(size_t)nu * n + e] = 0;
    int st = OR_OK, cnt = 0;
    if (bad)
        *bad = -1;
    for (int e = 0; e < n; ++e)
        for (int nu = 0; nu < ncomp; ++nu) {
            double Nn = N_old[(size_t)nu * n + e] + dN[(size_t)nu * n + e];
            if (Nn < 0) {
                if (truncate) {
                    Nn = 0;
                    cnt++;
                } else if (st == OR_OK) {
                    st = OR_BAD_ARGUMENT;
                    if (bad)
                        *bad = e * ncomp + nu;
                }
            }
            N_new[(size_t)nu * n + e] = Nn;
        }
    if (truncated)
        *truncated = cnt;
    free(dN);
    free(cf);
    return st;
}

/* TemperatureNewtonRaphson.solve (Discretisation.java:577-590): advective
 * flux (:621-660, residual only), conductive flux (:665-691), diagonal part
 * (:696-730) with the sources' energy rates (:735-752), A and b from zero. */
int or_impes_temperature(const or_impes_params *P, const or_impes_state *S, const int *here,
                         const int *there, const int *md_ptr, const int *md_elem,
                         const double *md_t, const int *mf_ptr, const int *mf_elem,
                         const double *mf_t, const double *rho, const int *upstream,
                         int nterms, const double *src_qe, const int *src_elem,
                         const int *rowptr, const int *colind, double *val, double *b)
{
    const int n = P->n, nc = P->nconn;
    memset(val, 0, sizeof(double) * (size_t)rowptr[n]);
    memset(b, 0, sizeof(double) * (size_t)n);
    for (int c = 0; c < nc; ++c) {
        const int i = here[c], j = there[c];
        double energyFlux = 0;
        for (int ph = 0; ph < 3; ++ph) {
            const int up = upstream[(size_t)ph * nc + c];
            if (!PH(S->present, ph, up))
                continue;
            const double energyMob = PH(S->energy_mobility, ph, up);
            for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
                const int ek = md_elem[q];
                const double tk = md_t[q];
                const double bf = body_force(P, S, ek);
                const double psi = PH(S->phase_pressure, ph, ek) + rho[(size_t)ph * nc + c] * bf;
                energyFlux += energyMob * tk * psi;
            }
        }
        b[i] += energyFlux;
        b[j] += -energyFlux;
    }
    for (int c = 0; c < nc; ++c) {
        const int i = here[c], j = there[c];
        for (int q = mf_ptr[c]; q < mf_ptr[c + 1]; ++q) {
            const int ik = mf_elem[q];
            const double tk = mf_t[q];
            const double T = S->temperature[ik];
            b[i] += tk * T;
            b[j] += -tk * T;
            if (or_csr_add(rowptr, colind, val, i, ik, tk) ||
                or_csr_add(rowptr, colind, val, j, ik, -tk))
                return OR_NOT_IN_PATTERN;
        }
    }
    for (int i = 0; i < n; ++i) {
        const double delta = S->heat_capacity[i];
        const double dT = S->temperature[i] - S->T0[i];
        b[i] += delta * dT / P->dt;
        if (or_csr_add(rowptr, colind, val, i, i, delta / P->dt))
            return OR_NOT_IN_PATTERN;
    }
    for (int t = 0; t < nterms; ++t)
        b[src_elem[t]] += -src_qe[t];
    return OR_OK;
}
