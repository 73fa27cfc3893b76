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
