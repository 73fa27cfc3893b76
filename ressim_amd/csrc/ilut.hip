// ILUT: the dual-threshold incomplete LU of MTJ on the device.
//
// Reference: mtj/src/no/uib/cipr/matrix/sparse/ILUT.java:136-272 (setMatrix,
// factor, gather), selected by rs at LinearSolver.java:211-212.  Row i of
// the factor is row i of A scattered into a dense work row, eliminated with
// every earlier factor row k whose multiplier z[k] / U(k,k) exceeds the
// row's threshold taui = tau ||A(i,:)||_2 in magnitude (z -= mult * U(k,:)),
// then gathered back: per side the entries above taui, the largest
// nl + p (lower) and nu + p (upper) of them, where nl / nu count A's row
// entries on that side, plus the diagonal; A's own entries that are dropped
// stay in the row as zeros (SparseVector.zero keeps the structure).
//
// MTJ walks k over all i earlier rows; the multiplier of a zero work-row
// entry is zero and skipped (|0| <= taui), so only the work row's nonzero
// lower positions matter, taken in ascending order.  Each elimination can
// only add positions right of k, so a sorted list of pending lower positions
// (in LDS) merged with the new ones after each elimination visits exactly
// MTJ's sequence.
//
// Rows are factored by persistent waves, one row at a time per wave, rows
// handed out in ascending order by an atomic ticket.  A wave reads factor
// row k only after row k's flag is set (release / acquire at agent scope),
// and a row waits only on rows with smaller tickets, whose waves are running
// and wait only on still smaller rows: the factorisation cannot deadlock,
// and a bounded spin turns any unexpected stall into an error.  Each wave
// owns a dense work row z (n doubles) and a stamp array in HBM; stamps mark
// positions touched by the current row (bit 0: in A's row, bit 1: kept),
// and each row clears what it touched, so nothing is reset between rows.
//
// The factor is written into per-row slots sized by the bound
// |A(i,:)| + min(nl + p, i) + min(nu + p, n - 1 - i) + 1, compacted into
// CSR, and applied by the ILU(0) sweep kernels (its apply and transApply
// are the solves of ILUT.java:320-443, the same statements as ILU's).
//
// Starting structure: every factorisation starts from A's nonzero entries
// (ILUT.setMatrix compacts the copy of A, SparseVector.compact keeping the
// nonzero entries); MTJ's results that depend on the FlexCompRowMatrix's
// array capacities (explicit zeros kept in rows without spare capacity, a
// reused matrix carrying the previous factor's structure) are not
// reproduced.
#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "rsgpu_impl.hpp"

#define ILUT_PCAP 2048   // pending lower positions of a row (LDS)
#define ILUT_RCAP 2048   // entries ranked or sorted at once (LDS)

enum { ILUT_ST_MISSING = 0, ILUT_ST_ZERO = 1, ILUT_ST_OVERFLOW = 2, ILUT_ST_TIMEOUT = 3 };

// Largest dynamic LDS a workgroup may request on this device (opt-in).
static size_t ilut_lds_limit(int device)
{
    int a = 0, b = 0;
    (void)hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    (void)hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, device);
    return (size_t)std::max(a, b);
}

namespace {

struct IlutArgs {
    int n, p;
    double tau;
    const int *arp, *aci;
    const double *aval;
    const int *off;        // [n + 1] factor row slots
    int *ocol;
    double *oval;
    int2 *info;            // per row: diagonal position, length (in its slot)
    int *flag;             // per row: 1 once the row is final
    int *ticket;
    int *status;           // [4]: missing-diagonal row, zero-diagonal row, overflow, timeout
    double *zb;            // per wave: n work-row values
    int *occ, *lb, *ub;    // per wave: n stamps, lower and upper position lists
    unsigned long long spin_ticks;
};

__device__ __forceinline__ int lane_prefix(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Wait until factor row k is final; false if it is not (the wait is
// bounded: past spin_ticks of the wall clock, or once any wave has timed
// out, it ends and the timeout status is set).  Lane 0 spins on the flag;
// the agent-scope acquire then makes row k's entries visible to the wave.
__device__ bool ilut_wait(const IlutArgs &a, int k)
{
    int ok = 1;
    if (threadIdx.x == 0 &&
        __hip_atomic_load(&a.flag[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(&a.flag[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (__hip_atomic_load(&a.status[ILUT_ST_TIMEOUT], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > a.spin_ticks) {
                atomicOr(&a.status[ILUT_ST_TIMEOUT], 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    ok = __shfl(ok, 0);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok != 0;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __syncthreads();
}

// Mark (stamp bit 1) the entries of list[0..cnt) whose |z| > taui that the
// gather keeps: all of them when at most lim, else the lim first in
// descending |z| with ascending position on ties (IntDoubleEntry.compareTo
// under the stable Collections.sort of lists built in ascending position).
__device__ void ilut_keep(const int *list, int cnt, const double *z, int *occ, double taui,
                          long long lim, int *RI, double *RV, bool &overflow)
{
    const int lane = threadIdx.x;
    int ce = 0;
    for (int t0 = 0; t0 < cnt; t0 += 64) {
        const int t = t0 + lane;
        const int k = t < cnt ? list[t] : 0;
        const bool el = t < cnt && fabs(z[k]) > taui;
        ce += __popcll(__ballot(el));
    }
    if (ce <= lim) {
        for (int t = lane; t < cnt; t += 64) {
            const int k = list[t];
            if (fabs(z[k]) > taui)
                occ[k] |= 2;
        }
        return;
    }
    if (ce > ILUT_RCAP) {
        overflow = true;
        return;
    }
    int m = 0;
    for (int t0 = 0; t0 < cnt; t0 += 64) {
        const int t = t0 + lane;
        const int k = t < cnt ? list[t] : 0;
        const double az = t < cnt ? fabs(z[k]) : 0.0;
        const bool el = t < cnt && az > taui;
        const unsigned long long b = __ballot(el);
        if (el) {
            RI[m + lane_prefix(b)] = k;
            RV[m + lane_prefix(b)] = az;
        }
        m += __popcll(b);
    }
    __syncthreads();
    for (int e = lane; e < m; e += 64) {
        const double ve = RV[e];
        const int ie = RI[e];
        int rank = 0;
        for (int f = 0; f < m; ++f) {
            const double vf = RV[f];
            rank += (vf > ve || (vf == ve && RI[f] < ie)) ? 1 : 0;
        }
        if (rank < lim)
            occ[ie] |= 2;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_ilut_factor(IlutArgs a)
{
    __shared__ int Pbuf[2][ILUT_PCAP];
    __shared__ int Qbuf[ILUT_PCAP];
    __shared__ int RI[ILUT_RCAP];
    __shared__ double RV[ILUT_RCAP];
    const int lane = threadIdx.x;
    const int n = a.n;
    double *z = a.zb + (size_t)blockIdx.x * n;
    int *occ = a.occ + (size_t)blockIdx.x * n;
    int *lb = a.lb + (size_t)blockIdx.x * n;
    int *ub = a.ub + (size_t)blockIdx.x * n;
    bool overflow = false;
    for (;;) {
        int i = 0;
        if (lane == 0)
            i = atomicAdd(a.ticket, 1);
        i = __shfl(i, 0);
        if (i >= n)
            break;
        const int SF = 4 * (i + 1), SA = SF + 1;
        const int a0 = a.arp[i], a1 = a.arp[i + 1];
        // scatter A's nonzero entries (ILUT.scatter, :218-225)
        int plen = 0, nU = 0, cur = 0;
        bool hasdiag = false;
        for (int q0 = a0; q0 < a1; q0 += 64) {
            const int q = q0 + lane;
            const bool in = q < a1;
            const int c = in ? a.aci[q] : 0;
            const double v = in ? a.aval[q] : 0.0;
            const bool nz = in && v != 0.0;
            if (nz) {
                z[c] = v;
                occ[c] = SA;
            }
            const unsigned long long ml = __ballot(nz && c < i), mu = __ballot(nz && c > i);
            if (__ballot(nz && c == i))
                hasdiag = true;
            if (nz && c < i) {
                const int pos = plen + lane_prefix(ml);
                if (pos < ILUT_PCAP)
                    Pbuf[0][pos] = c;
            }
            if (nz && c > i)
                ub[nU + lane_prefix(mu)] = c;
            plen += __popcll(ml);
            nU += __popcll(mu);
        }
        const int nlA = plen, nuA = nU;
        if (plen > ILUT_PCAP) {
            overflow = true;
            plen = ILUT_PCAP;
        }
        if (!hasdiag && lane == 0) {
            atomicMin(&a.status[ILUT_ST_MISSING], i);
            z[i] = 0.0;
            occ[i] = SA;
        }
        // taui = rowi.norm(Two) * tau (SparseVector.norm2, :278-283)
        double taui = 0.0;
        if (i > 0 && lane == 0) {
            double s = 0.0;
            for (int q = a0; q < a1; ++q) {
                const double v = a.aval[q];
                s = s + v * v;
            }
            taui = sqrt(s) * a.tau;
        }
        taui = __shfl(taui, 0);
        wave_sync();

        // eliminate the pending lower positions in ascending order (:168-192)
        int nlall = 0, head = 0;
        while (head < plen) {
            const int k = Pbuf[cur][head];
            ++head;
            const bool ready = ilut_wait(a, k);
            const int offk = a.off[k], capk = a.off[k + 1] - offk;
            const int2 ik = ready ? a.info[k] : make_int2(0, 0);
            // a row that never became final (a stalled wait) is not read
            const bool sane = ready && ik.x >= 0 && ik.x < ik.y && ik.y <= capk;
            const double dk = sane ? a.oval[offk + ik.x] : 1.0;
            const double luik = z[k] / dk;
            if (sane && !(fabs(luik) <= taui)) {
                const int u0 = offk + ik.x + 1, u1 = offk + ik.y;
                int q = 0;
                for (int e0 = u0; e0 < u1; e0 += 64) {
                    const int e = e0 + lane;
                    int c = 0;
                    bool fresh = false;
                    if (e < u1) {
                        c = a.ocol[e];
                        const double v = a.oval[e];
                        const double zc = z[c];
                        z[c] = zc - luik * v;
                        fresh = occ[c] < SF;
                        if (fresh)
                            occ[c] = SF;
                    }
                    const unsigned long long ml = __ballot(fresh && c < i),
                                             mu = __ballot(fresh && c > i);
                    if (fresh && c < i) {
                        const int pos = q + lane_prefix(ml);
                        if (pos < ILUT_PCAP)
                            Qbuf[pos] = c;
                    }
                    if (fresh && c > i)
                        ub[nU + lane_prefix(mu)] = c;
                    q += __popcll(ml);
                    nU += __popcll(mu);
                }
                if (lane == 0)
                    z[k] = luik;
                if (q > 0) {
                    // merge the new positions (ascending, all > k) into the
                    // pending list
                    if (q > ILUT_PCAP) {
                        overflow = true;
                        q = ILUT_PCAP;
                    }
                    const int rem = plen - head;
                    __syncthreads();
                    const int *P = Pbuf[cur] + head;
                    int *R = Pbuf[cur ^ 1];
                    for (int t = lane; t < rem; t += 64) {
                        const int x = P[t];
                        int lo = 0, hi = q;
                        while (lo < hi) {
                            const int m = (lo + hi) >> 1;
                            if (Qbuf[m] < x)
                                lo = m + 1;
                            else
                                hi = m;
                        }
                        if (t + lo < ILUT_PCAP)
                            R[t + lo] = x;
                    }
                    for (int t = lane; t < q; t += 64) {
                        const int y = Qbuf[t];
                        int lo = 0, hi = rem;
                        while (lo < hi) {
                            const int m = (lo + hi) >> 1;
                            if (P[m] < y)
                                lo = m + 1;
                            else
                                hi = m;
                        }
                        if (t + lo < ILUT_PCAP)
                            R[t + lo] = y;
                    }
                    __syncthreads();
                    if (rem + q > ILUT_PCAP)
                        overflow = true;
                    cur ^= 1;
                    head = 0;
                    plen = min(rem + q, ILUT_PCAP);
                }
            }
            if (lane == 0)
                lb[nlall] = k;
            ++nlall;
            wave_sync();
        }

        // gather (:231-272): row 0 is never gathered and keeps A's row
        if (i > 0) {
            const long long limL = std::max(0LL, (long long)nlA + a.p);
            const long long limU = std::max(0LL, (long long)nuA + a.p);
            ilut_keep(lb, nlall, z, occ, taui, limL, RI, RV, overflow);
            wave_sync();
            ilut_keep(ub, nU, z, occ, taui, limU, RI, RV, overflow);
            wave_sync();
        }
        // the row: lower entries in ascending order (A's entries, kept ones
        // with their values, the others as zeros), the diagonal, the upper
        // entries sorted by column
        const int off = a.off[i], cap = a.off[i + 1] - off;
        int nLo = 0;
        for (int t0 = 0; t0 < nlall; t0 += 64) {
            const int t = t0 + lane;
            const int k = t < nlall ? lb[t] : 0;
            const int o = t < nlall ? occ[k] : 0;
            const bool out = t < nlall && (o & 3) != 0;
            const unsigned long long b = __ballot(out);
            const int pos = nLo + lane_prefix(b);
            if (out && pos < cap) {
                a.ocol[off + pos] = k;
                a.oval[off + pos] = (o & 2) ? z[k] : 0.0;
            }
            nLo += __popcll(b);
        }
        int nF = 0;
        for (int t0 = 0; t0 < nU; t0 += 64) {
            const int t = t0 + lane;
            const int c = t < nU ? ub[t] : 0;
            const bool out = t < nU && (occ[c] & 3) != 0;
            const unsigned long long b = __ballot(out);
            if (out && nF + lane_prefix(b) < ILUT_RCAP)
                RI[nF + lane_prefix(b)] = c;
            nF += __popcll(b);
        }
        if (nF > ILUT_RCAP) {
            overflow = true;
            nF = ILUT_RCAP;
        }
        if (nLo + 1 + nF > cap) {
            overflow = true;
            nLo = min(nLo, cap - 1);
            nF = cap - 1 - nLo;
        }
        __syncthreads();
        if (lane == 0) {
            a.ocol[off + nLo] = i;
            a.oval[off + nLo] = z[i];
        }
        for (int e = lane; e < nF; e += 64) {
            const int c = RI[e];
            int r = 0;
            for (int f = 0; f < nF; ++f)
                r += RI[f] < c ? 1 : 0;
            const int o = occ[c];
            a.ocol[off + nLo + 1 + r] = c;
            a.oval[off + nLo + 1 + r] = (i == 0 || (o & 2)) ? z[c] : 0.0;
        }
        if (lane == 0) {
            a.info[i] = make_int2(nLo, nLo + 1 + nF);
            // ILUT.factor tests U(k,k) of rows k < n - 1 (:176-178)
            if (i <= n - 2 && z[i] == 0.0)
                atomicMin(&a.status[ILUT_ST_ZERO], i);
        }
        __threadfence();
        if (lane == 0)
            __hip_atomic_store(&a.flag[i], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        // clear what the row touched
        for (int t = lane; t < nlall; t += 64) {
            const int k = lb[t];
            z[k] = 0.0;
            occ[k] = 0;
        }
        for (int t = lane; t < nU; t += 64) {
            const int c = ub[t];
            z[c] = 0.0;
            occ[c] = 0;
        }
        if (lane == 0) {
            z[i] = 0.0;
            occ[i] = 0;
        }
        wave_sync();
    }
    if (overflow && lane == 0)
        atomicOr(&a.status[ILUT_ST_OVERFLOW], 1);
}

// Factor rows from their slots into CSR (rowptr given); diagonal positions
// and the sweeps' per-row ranges (L: [rp, dg), U: [dg + 1, rp + 1)).
__global__ void k_ilut_compact(int n, const int *off, const int2 *info, const int *ocol,
                               const double *oval, const int *rp, int *ci, double *val, int *dg,
                               int2 *bl, int2 *bu)
{
    const int r = blockIdx.x;
    if (r >= n)
        return;
    const int src = off[r], dst = rp[r], len = info[r].y;
    for (int e = threadIdx.x; e < len; e += blockDim.x) {
        ci[dst + e] = ocol[src + e];
        val[dst + e] = oval[src + e];
    }
    if (threadIdx.x == 0) {
        dg[r] = dst + info[r].x;
        bl[r] = make_int2(dst, dst + info[r].x);
        bu[r] = make_int2(dst + info[r].x + 1, dst + len);
    }
}

// ---------------------------------------------------------------- sweeps
// ILUT's fill makes long dependency chains (level counts near n / 2 on grid
// matrices), so its solves run as one wave walking the rows in order with
// the vector in LDS.  Gather rows (L.solve, U.solve): lane t forms the
// product of entry t, and the products are added to sum = 0 in entry
// (column) order by readlane -- the statement order of
// UnitLowerFlexCompRowMatrix.solve / UpperFlexCompRowMatrix.solve
// (ILUT.java:321-344, 390-413); lanes past the row's end add -0.0, an exact
// identity.  Scatter rows (transSolve, :347-370, 417-443): x[i] is final
// when row i is reached; lanes subtract v * x[i] from their distinct
// columns.  A ring of SEQ_D rows of operands (the first 64 entries of each
// row and its right-hand side or diagonal) is in flight ahead of the row
// being solved, and the row ranges SEQ_D rows further ahead.
#define SEQ_D 8

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// s + p(lane 0) + p(lane 1) + ... + p(lane m - 1), left to right
__device__ __forceinline__ double lane_serial_sum(double s, double p, int m)
{
    for (int u0 = 0; u0 < m; u0 += 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            s = s + readlane_f64(p, u0 + k);
    }
    return s;
}

struct SeqSlot {
    int2 b;        // the row's entry range
    int c;         // this lane's first-chunk column
    double v;      // and value
    double aux;    // L: b[i]; U, U': the diagonal
};

// MODE 0: L solve ascending (x[i] = rhs[i] - sum), 1: U solve descending
// (x[i] = (x[i] - sum) / U(i,i)), 2: U' ascending (x[i] /= U(i,i), then
// scatter), 3: L' descending (scatter).  X: the LDS vector, in place, with
// SEQ_D junk rows past n: the steps past the last row run on empty ranges
// (1, 1) and write there, so every step is the same straight-line code and
// the compiler's in-order load counting keeps the ring in flight.  Every row
// holds at most 64 entries on the swept side (checked by the caller).
template <int MODE>
__device__ void seq_sweep(int n, const int2 *__restrict__ bnd, const int *__restrict__ ci,
                          const double *__restrict__ lu, const double *__restrict__ rhs,
                          double *X)
{
    constexpr bool fwd = MODE == 0 || MODE == 2;
    constexpr bool gather = MODE == 0 || MODE == 1;
    const int lane = threadIdx.x;
    auto row_of = [&](int s) { return s < n ? (fwd ? s : n - 1 - s) : s; };
    // per-row data of the SEQ_D steps from base, lane l holding step base + l
    // (vector loads: uniform addresses would become scalar loads, whose
    // counter the LDS waits share).  bnd is padded by 3 SEQ_D empty ranges
    // (1, 1) on both sides, so steps past the last row read padding.
    auto block_bounds = [&](int base) {
        const int st = base + (lane & (SEQ_D - 1));
        return bnd[fwd ? st : n - 1 - st];
    };
    auto block_aux = [&](int base, int2 b) {
        if (MODE == 0)
            return rhs[row_of(min(base + (lane & (SEQ_D - 1)), n - 1))];
        if (MODE == 3)
            return 0.0;
        return lu[b.x - 1];
    };
    auto load_slot = [&](SeqSlot &q, int2 bv, double av, int k) {
        q.b.x = __builtin_amdgcn_readlane(bv.x, k);
        q.b.y = __builtin_amdgcn_readlane(bv.y, k);
        q.aux = readlane_f64(av, k);
        const int t = min(q.b.x + lane, q.b.y - 1);
        const bool in = q.b.x + lane < q.b.y;
        const int c = ci[max(t, 0)];
        const double v = lu[max(t, 0)];
        q.c = in ? c : 0;
        q.v = in ? v : 0.0;
    };
    // loop entry as after an iteration: (B1, A1) and B2 in flight, then the
    // ring's loads last
    SeqSlot ring[SEQ_D];
    const int2 B0 = block_bounds(0);
    const double A0 = block_aux(0, B0);
    int2 B1 = block_bounds(SEQ_D);
    double A1 = block_aux(SEQ_D, B1);
    int2 B2 = block_bounds(2 * SEQ_D);
#pragma unroll
    for (int k = 0; k < SEQ_D; ++k)
        load_slot(ring[k], B0, A0, k);
    int s0 = 0;
    do {
        // ring: steps s0.. ; (B1, A1): steps s0 + D..; B2: steps s0 + 2D..
        const double A2 = block_aux(s0 + 2 * SEQ_D, B2);
        const int2 B3 = block_bounds(s0 + 3 * SEQ_D);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < SEQ_D; ++k) {
            const int s = s0 + k;
            const int i = row_of(s);
            const SeqSlot q = ring[k];
            const int j0 = q.b.x, j1 = q.b.y;
            if (gather) {
                const double p = lane < j1 - j0 ? q.v * X[q.c] : -0.0;
                const double sum = lane_serial_sum(0.0, p, j1 - j0);
                const double xi = MODE == 0 ? q.aux - sum : (X[i] - sum) / q.aux;
                if (lane == 0)
                    X[i] = xi;
            } else {
                double xi = X[i];
                if (MODE == 2) {
                    xi = xi / q.aux;
                    if (lane == 0)
                        X[i] = xi;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (j0 + lane < j1)
                    X[q.c] = X[q.c] - q.v * xi;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // refill with step s + D
            load_slot(ring[k], B1, A1, k);
            __builtin_amdgcn_sched_barrier(0);
        }
        B1 = B2;
        A1 = A2;
        B2 = B3;
        s0 += SEQ_D;
    } while (s0 < n);
}

__device__ __forceinline__ bool seq_skip(const int *sk1, const int *sk2)
{
    return (sk1 && *sk1) || (sk2 && *sk2);
}

// apply: L then U; transApply: U' then L'.  One wave, x in LDS.
template <bool TRANS>
__global__ void __launch_bounds__(64) k_ilut_seq(int n, const int2 *bl, const int2 *bu,
                                                 const int *ci, const double *lu,
                                                 const double *b, double *x, const int *sk1,
                                                 const int *sk2)
{
    if (seq_skip(sk1, sk2))
        return;
    extern __shared__ double X[];
    if (TRANS) {
        for (int r = threadIdx.x; r < n; r += 64)
            X[r] = b[r];
        __syncthreads();
        seq_sweep<2>(n, bu, ci, lu, nullptr, X);
        seq_sweep<3>(n, bl, ci, lu, nullptr, X);
    } else {
        seq_sweep<0>(n, bl, ci, lu, b, X);
        seq_sweep<1>(n, bu, ci, lu, nullptr, X);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += 64)
        x[r] = X[r];
}

} // namespace

struct IlutData {
    uint64_t pattern_id = 0;
    int p = INT_MIN;
    std::vector<int> h_off;
    DevBuf<int> off, ocol, flag, ctl, rp, ci, dg;
    DevBuf<int2> bl, bu;     // per row: L and U entry ranges of the factor
    bool seq = false;        // solves by the one-wave sequential sweeps
    DevBuf<double> oval, val;
    DevBuf<int2> info;
    DevBuf<double> zb;
    DevBuf<int> occ, lb, ub;
    int waves = 0, scratch_n = 0;
    bool sweeps = false;
};

void ilut_destroy(IlutData *d) { delete d; }

int ilut_set_matrix(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    hipStream_t s = M->ctx->stream;
    const int n = A->n;
    M->ready = false;
    if (A->n != A->ncols)
        throw RsError{RSGPU_BAD_ARGUMENT, "ILU only applies to square matrices"};
    if (n >= (1 << 29))
        throw RsError{RSGPU_BAD_ARGUMENT, "ILUT supports fewer than 2^29 rows"};
    if (!M->ilut)
        M->ilut = new IlutData();
    IlutData &D = *M->ilut;
    // factor row slots: |A(i,:)| + min(nl + p, i) + min(nu + p, n - 1 - i) + 1
    if (D.pattern_id != A->pattern_id || D.p != M->ilut_p || D.h_off.size() != (size_t)n + 1) {
        std::vector<int> off(n + 1, 0);
        long long t = 0;
        for (int r = 0; r < n; ++r) {
            long long nl = 0, nu = 0;
            for (int j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; ++j) {
                if (A->h_colind[j] < r)
                    nl++;
                else if (A->h_colind[j] > r)
                    nu++;
            }
            const long long kl = std::min<long long>(std::max(0LL, nl + M->ilut_p), r);
            const long long ku = std::min<long long>(std::max(0LL, nu + M->ilut_p), n - 1 - r);
            t += (A->h_rowptr[r + 1] - A->h_rowptr[r]) + kl + ku + 1;
            if (t > INT32_MAX)
                throw RsError{RSGPU_BAD_ARGUMENT,
                              "ILUT fill bound exceeds 2^31 entries; use a smaller p"};
            off[r + 1] = (int)t;
        }
        D.h_off = off;
        D.off.upload(off, s);
        D.ocol.alloc(std::max(1, off[n]));
        D.oval.alloc(std::max(1, off[n]));
        D.info.alloc(std::max(1, n));
        D.flag.alloc(std::max(1, n));
        D.pattern_id = A->pattern_id;
        D.p = M->ilut_p;
    }
    // per-wave work rows: up to 512 waves within 4 GiB
    const size_t per_wave = (size_t)std::max(n, 1) * 20;
    const int want = (int)std::max<size_t>(
        1, std::min<size_t>({(size_t)512, (size_t)std::max(n, 1), ((size_t)4 << 30) / per_wave}));
    if (D.waves != want || D.scratch_n != n) {
        const size_t m = (size_t)want * std::max(n, 1);
        D.zb.alloc(m);
        D.occ.alloc(m);
        D.lb.alloc(m);
        D.ub.alloc(m);
        D.zb.zero(s);
        D.occ.zero(s);
        D.waves = want;
        D.scratch_n = n;
    }
    int ctl0[5] = {0, INT_MAX, INT_MAX, 0, 0};
    D.ctl.upload(ctl0, 5, s);
    if (n > 0) {
        D.flag.zero(s);
        int khz = 0;
        RS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, M->ctx->device));
        IlutArgs a;
        a.n = n;
        a.p = M->ilut_p;
        a.tau = M->ilut_tau;
        a.arp = A->rowptr.p;
        a.aci = A->colind.p;
        a.aval = A->val.p;
        a.off = D.off.p;
        a.ocol = D.ocol.p;
        a.oval = D.oval.p;
        a.info = D.info.p;
        a.flag = D.flag.p;
        a.ticket = D.ctl.p;
        a.status = D.ctl.p + 1;
        a.zb = D.zb.p;
        a.occ = D.occ.p;
        a.lb = D.lb.p;
        a.ub = D.ub.p;
        // a stalled wait ends after 30 s of the device's wall clock
        a.spin_ticks = (unsigned long long)(khz > 0 ? khz : 100000) * 30000ull;
        hipLaunchKernelGGL(k_ilut_factor, dim3(D.waves), dim3(64), 0, s, a);
        RS_CHECK_LAUNCH();
    }
    int ctl[5];
    RS_HIP(hipMemcpyAsync(ctl, D.ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, s));
    std::vector<int2> info(n);
    if (n)
        RS_HIP(hipMemcpyAsync(info.data(), D.info.p, sizeof(int2) * n, hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    const int *st = ctl + 1;
    if (st[ILUT_ST_OVERFLOW] || st[ILUT_ST_TIMEOUT]) {
        // rows may have left stamps behind
        D.zb.zero(s);
        D.occ.zero(s);
        RS_HIP(hipStreamSynchronize(s));
        if (st[ILUT_ST_TIMEOUT])
            throw RsError{RSGPU_HIP_ERROR, "ILUT factorisation stalled waiting for a row"};
        throw RsError{RSGPU_BAD_ARGUMENT,
                      "ILUT: a row's work lists exceed the device capacity (2048 pending "
                      "lower positions, 2048 ranked or sorted entries)"};
    }
    if (st[ILUT_ST_MISSING] != INT_MAX) {
        if (bad_row)
            *bad_row = st[ILUT_ST_MISSING];
        throw RsError{RSGPU_MISSING_DIAGONAL,
                      "Missing diagonal entry on row " + std::to_string(st[ILUT_ST_MISSING] + 1)};
    }
    if (st[ILUT_ST_ZERO] != INT_MAX) {
        if (bad_row)
            *bad_row = st[ILUT_ST_ZERO];
        M->ctx->err = "Zero diagonal entry on row " + std::to_string(st[ILUT_ST_ZERO] + 1) +
                      " during ILU process";
        return RSGPU_ZERO_PIVOT;
    }
    // CSR of the factor
    std::vector<int> rp(n + 1, 0);
    for (int r = 0; r < n; ++r)
        rp[r + 1] = rp[r] + info[r].y;
    const int nnz = rp[n];
    D.rp.upload(rp, s);
    D.ci.alloc(std::max(1, nnz));
    D.val.alloc(std::max(1, nnz));
    D.dg.alloc(std::max(1, n));
    // sweep ranges padded by 3 SEQ_D empty ranges (1, 1) on both sides
    const size_t pad = 3 * SEQ_D;
    if (D.bl.n != (size_t)n + 2 * pad) {
        std::vector<int2> e((size_t)n + 2 * pad, make_int2(1, 1));
        D.bl.upload(e, s);
        D.bu.upload(e, s);
    }
    if (n)
        hipLaunchKernelGGL(k_ilut_compact, dim3(n), dim3(64), 0, s, n, D.off.p, D.info.p,
                           D.ocol.p, D.oval.p, D.rp.p, D.ci.p, D.val.p, D.dg.p, D.bl.p + pad,
                           D.bu.p + pad);
    RS_CHECK_LAUNCH();
    std::vector<int> ci(nnz);
    if (nnz)
        RS_HIP(hipMemcpyAsync(ci.data(), D.ci.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    // sweep schedules for the factor's pattern, rebuilt when it changed
    if (!D.sweeps || M->h_rowptr != rp || M->h_colind != ci) {
        D.sweeps = false;
        std::vector<int> dg(n);
        for (int r = 0; r < n; ++r)
            dg[r] = rp[r] + info[r].x;
        M->n = n;
        M->nblocks = 1;
        M->h_bstart = {0, n};
        M->bstart.upload(M->h_bstart, s);
        M->max_block_rows = n;
        M->nnz = nnz;
        M->h_rowptr = rp;
        M->h_colind = ci;
        M->src_identity = true;
        M->rowptr.upload(rp, s);
        M->colind.upload(ci, s);
        M->diag.upload(dg, s);
        prec_setup_pattern(M, rp, ci, dg, 1);
        D.sweeps = true;
    }
    if (nnz)
        RS_HIP(hipMemcpyAsync(M->lu64.p, D.val.p, sizeof(double) * nnz,
                              hipMemcpyDeviceToDevice, s));
    prec_refresh_fp64(M);
    RS_HIP(hipStreamSynchronize(s));
    // the sequential sweeps: the vector in LDS, at most 64 entries per row
    // and side (one lane per entry)
    int side = 0;
    for (int r = 0; r < n; ++r)
        side = std::max({side, info[r].x, info[r].y - info[r].x - 1});
    D.seq = n > 0 && side <= 64 &&
            ((size_t)n + SEQ_D) * sizeof(double) <= ilut_lds_limit(M->ctx->device) &&
            getenv("RSGPU_ILUT_LEVELS") == nullptr;
    M->pattern_id = A->pattern_id;
    M->ready = true;
    return RSGPU_OK;
}

// ILUT.apply / transApply by the sequential sweeps; false when they do not
// apply (the vector exceeds LDS): the level-scheduled ILU sweeps run.
bool ilut_apply(rsgpu_prec *M, const double *b, double *x, bool trans, const int *sk1,
                const int *sk2)
{
    IlutData *D = M->ilut;
    if (!D || !D->seq)
        return false;
    const size_t shm = ((size_t)M->n + SEQ_D) * sizeof(double);
    auto k = trans ? k_ilut_seq<true> : k_ilut_seq<false>;
    static size_t granted[2] = {65536, 65536};
    if (shm > granted[trans]) {
        RS_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)shm));
        granted[trans] = shm;
    }
    hipLaunchKernelGGL(k, dim3(1), dim3(64), shm, M->ctx->stream, M->n, D->bl.p + 3 * SEQ_D,
                       D->bu.p + 3 * SEQ_D, M->colind.p, M->lu64.p, b, x, sk1, sk2);
    RS_CHECK_LAUNCH();
    return true;
}
