// Krylov solvers on gfx950 with the monitor and all scalars on the device.
//
// Reference: BiCGstab.solve (mtj/src/no/uib/cipr/matrix/sparse/
// BiCGstab.java:65-116), GMRES.solve (GMRES.java:127-176) with
// GivensRotation (mtj/src/no/uib/cipr/matrix/GivensRotation.java:42-93) and
// the dtrtrs back-solve (AbstractTriangDenseMatrix.java:227-256), IR.solve
// (IR.java:57-70), DefaultIterationMonitor (DefaultIterationMonitor.java:
// 133-156).
//
// Design.  The host enqueues the fixed kernel sequence of one iteration (or
// one GMRES inner step) without waiting; every kernel starts by reading the
// device flags (done / cycle_end) and returns at once when the solve has
// stopped.  Scalar decisions (alpha, beta, omega, Givens rotations,
// convergence, breakdown) are taken by the finishing workgroup of the
// reduction that produces their operands, so no kernel launch exists only
// to do scalar work.  The host reads the flags one step behind (an event
// per step), so the GPU queue never drains while the host decides.
// Each vector update keeps the reference's per-element operation order
// (including DenseVector.add/set's alpha == 0 shortcuts); dot products and
// norms are deterministic tree sums.
#include <algorithm>
#include <cmath>

#include "spmv_core.hpp"

#define RS_VBS 256

struct rsgpu_solver {
    rsgpu_ctx *ctx = nullptr;
    rsgpu_solver_opts o{};
    int n = 0;
    size_t ldv = 0;               // internal vector stride (multiple of 4)
    DevState *st = nullptr;       // device
    DevState *st_in = nullptr;    // inner solve (mixed)
    DevBuf<double> hist, hist_in;
    DevBuf<double> partials;
    DevBuf<unsigned> ticket;
    DevBuf<double> gmres_small;   // H, s, gc, gs for outer and inner
    // FP64 work vectors
    DevBuf<double> w64;
    // FP32 work vectors (inner solve)
    DevBuf<float> w32;
    DevBuf<double> xr;            // mixed: residual / correction in FP64
    int last_hist = 0;
    std::vector<double> h_hist;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int *flags_host = nullptr;    // pinned [2][4]
    int *flags_dev = nullptr;     // the same memory as the device addresses it
    // the matrix whose FP64 sliced copy this solve filled (its SpMVs then
    // stream the sliced copy)
    const rsgpu_csr *sl64 = nullptr;
    int fused_steps = 0;          // fused GMRES steps enqueued by the last solve
    DevBuf<DevState> st_probe;    // state of rsgpu_gmres_fused_step calls
    DevBuf<double> dcgs;          // FP64 DCGS2: Hr ((m+1) x m) and ds (m+2)
};

// vector kernels: at least 4 elements per lane, at most 1024 groups
static int grid_for(int n)
{
    int g = (n + 4 * RS_VBS - 1) / (4 * RS_VBS);
    return std::max(1, std::min(g, 1024));
}

// Both flags are read unconditionally: a short-circuit test makes the
// second read wait for the first, one more memory round trip at the start
// of every solver kernel.
__device__ __forceinline__ bool stopped(const DevState *st, bool inner_step)
{
    const int d = st->done;
    const int c = st->cycle_end;
    return (d != 0) | (inner_step & (c != 0));
}

// block pattern of a 3x3 block matrix (null for a plain CSR) and the
// plan's 16-bit column offsets (null when absent)
struct BMat {
    const int *brp, *bcol;
    const short *coff;
};

// =========================================================== fused SpMV
// y = A x (mult rule), plus up to two partial dot products over the rows,
// finished by Fin.
template <typename T>
struct EpiMultDot {
    T *y;
    const T *z1;    // nullptr -> no first dot
    const T *z2;    // nullptr -> second dot uses y itself
    int nd;
    double a1 = 0, a2 = 0;
    T z1v = T(0), z2v = T(0);
    // the dot operands of row r are read here, before the row's own loads,
    // so their latency overlaps the row's instead of following it
    __device__ T init(int r)
    {
        if (nd >= 1)
            z1v = z1[r];
        if (nd >= 2 && z2)
            z2v = z2[r];
        return T(0);
    }
    __device__ void row(int r, T dot)
    {
        const T v = (dot != T(0)) ? dot : T(0);
        y[r] = v;
        if (nd >= 1)
            a1 += (double)(z1v * v);
        if (nd >= 2)
            a2 += (double)(z2 ? z2v * v : v * v);
    }
};

// residual r = b + (-1)*A x, optional copy, norm^2 partial
template <typename T>
struct EpiResidNorm {
    T *y;
    T *copy;
    const T *b;
    float *y32;     // optional FP32 shadow of the residual (mixed form)
    double a1 = 0, a2 = 0;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot)
    {
        const T v = b[r] + T(-1) * dot;
        y[r] = v;
        if (copy)
            copy[r] = v;
        if (y32)
            y32[r] = (float)v;
        a1 += (double)(v * v);
    }
};

enum {
    FIN_NONE = 0,
    FIN_ALPHA,       // BiCGstab: alpha = rho1 / (rtilde . v)
    FIN_OMEGA,       // BiCGstab: omega = (t . s) / (t . t)
    FIN_INIT_NORM,   // monitor(||r||) at iteration 0 (BiCGstab, IR)
    FIN_GM_NORMR,    // GMRES: normr = ||r||; outer monitor
    FIN_CG_ALPHA,    // CG: alpha = rho / (p . q)
    FIN_RHO_BREAK,   // BiCG, CGS: rho1 = t; breakdown if 0; beta = rho1 / rho2
    FIN_QMR_RHO0,    // QMR: rho = ||y|| before the loop
    FIN_QMR_XI0,     // QMR: xi = ||z|| before the loop
    FIN_QMR_DELTA,   // QMR: delta = z . y; breakdown if 0; p / q scale factors
    FIN_QMR_EP,      // QMR: ep = q . p_tld; beta = ep / delta; breakdowns
    FIN_QMR_RHO,     // QMR: rho_1 = rho; rho = ||y||
    FIN_QMR_XI,      // QMR: xi = ||z||; theta, gamma, eta; d / s scale factor
};

__device__ __forceinline__ void breakdown(DevState *st)
{
    st->status = RSGPU_BREAKDOWN;
    st->done = 1;
}

__device__ void finish_scalar(DevState *st, int fin, double t1, double t2)
{
    switch (fin) {
    case FIN_ALPHA:
        st->alpha = st->rho1 / t1;
        break;
    case FIN_OMEGA:
        st->ts = t1;
        st->tt = t2;
        st->omega = t1 / t2;
        break;
    case FIN_INIT_NORM:
        if (dev_monitor(st, sqrt(t1)))
            st->done = 1;
        break;
    case FIN_GM_NORMR:
        st->normr = sqrt(t1);
        if (dev_monitor(st, st->normr))
            st->done = 1;
        break;
    case FIN_CG_ALPHA:
        st->alpha = st->rho1 / t1;
        break;
    case FIN_RHO_BREAK:
        st->rho1 = t1;
        if (t1 == 0) {
            breakdown(st);
            break;
        }
        if (!st->first)
            st->beta = st->rho1 / st->rho2;
        break;
    case FIN_QMR_RHO0:
        st->q_rho = sqrt(t1);
        break;
    case FIN_QMR_XI0:
        st->q_xi = sqrt(t1);
        break;
    case FIN_QMR_DELTA:
        st->q_delta = t1;
        if (t1 == 0) {
            breakdown(st);
            break;
        }
        if (!st->first) {
            st->q_c1 = -st->q_xi * st->q_delta / st->q_ep;
            st->q_c2 = -st->q_rho * st->q_delta / st->q_ep;
        }
        break;
    case FIN_QMR_EP:
        st->q_ep = t1;
        if (t1 == 0) {
            breakdown(st);
            break;
        }
        st->q_beta = st->q_ep / st->q_delta;
        if (st->q_beta == 0)
            breakdown(st);
        break;
    case FIN_QMR_RHO:
        st->q_rho1 = st->q_rho;
        st->q_rho = sqrt(t1);
        break;
    case FIN_QMR_XI: {
        st->q_xi = sqrt(t1);
        st->q_gamma1 = st->q_gamma;
        st->q_theta1 = st->q_theta;
        st->q_theta = st->q_rho / (st->q_gamma1 * st->q_beta);
        st->q_gamma = 1 / sqrt(1 + st->q_theta * st->q_theta);
        if (st->q_gamma == 0) {
            breakdown(st);
            break;
        }
        st->q_eta = -st->q_eta * st->q_rho1 * st->q_gamma * st->q_gamma /
                    (st->q_beta * st->q_gamma1 * st->q_gamma1);
        st->q_val = st->q_theta1 * st->q_theta1 * st->q_gamma * st->q_gamma;
        break;
    }
    default:
        break;
    }
}

// the block-row core for 3x3 block matrices, the CSR core otherwise
template <typename T, bool B3, class Epi>
__device__ __forceinline__ void spmv_any(const int *wg_rows, const int *rowptr,
                                         const int *colind, const BMat &bm, const T *val,
                                         const T *x, Epi &epi)
{
    if (B3)
        spmv_core_b3<T, RS_SPMV_BS, spmv_nnz<T>()>(wg_rows, rowptr, colind, bm.brp, bm.bcol,
                                                   val, x, epi);
    else
        spmv_core_off<T, RS_SPMV_BS, spmv_nnz<T>()>(wg_rows, rowptr, colind, bm.coff, val, x,
                                                    epi);
}

// SpMV with fused dot products.  Each workgroup computes a contiguous run
// of plan groups (XCD-aware: the runs of one XCD are adjacent) and pays the
// cross-workgroup reduction once, at the end: with one group per workgroup
// the reduction's atomic round trip at every group's end held each
// workgroup slot well beyond its loads.
template <typename T, class Epi, int NR, bool B3>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv_red(const int *__restrict__ wg_rows, int ngroups, const int *__restrict__ rowptr,
               const int *__restrict__ colind, BMat bm, const T *__restrict__ val,
               const T *__restrict__ x, Epi epi, DevState *st, RedSlot slot,
               int fin, int inner_step)
{
    if (stopped(st, inner_step))
        return;
    const int G = (int)gridDim.x, w = xcd_remap((int)blockIdx.x, G);
    const int per = ngroups / G, rem = ngroups % G;
    const int g0 = w * per + min(w, rem), g1 = g0 + per + (w < rem ? 1 : 0);
    for (int g = g0; g < g1; ++g) {
        if (g > g0)
            __syncthreads();   // the previous group's LDS reads are done
        if (B3)
            spmv_core_b3_g<T, RS_SPMV_BS, spmv_nnz<T>()>(g, wg_rows, rowptr, colind, bm.brp,
                                                         bm.bcol, val, x, epi);
        else
            spmv_core_off_g<T, RS_SPMV_BS, spmv_nnz<T>()>(g, wg_rows, rowptr, colind, bm.coff,
                                                          val, x, epi);
    }
    double v[NR], tot[NR];
    v[0] = epi.a1;
    if (NR > 1)
        v[NR - 1] = epi.a2;
    if (block_reduce_finish<RS_SPMV_BS, NR>(slot, v, tot) && threadIdx.x == 0)
        finish_scalar(st, fin, tot[0], NR > 1 ? tot[NR - 1] : 0.0);
}

// k_spmv_red over the sliced copy: each workgroup computes a contiguous run
// of 4-chunk tiles, then the same deterministic reduction
template <typename T, class Epi, int NR, bool B3>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv_red_sl(SlView sv, int ntiles, const T *__restrict__ val, const T *__restrict__ x,
                  Epi epi, DevState *st, RedSlot slot, int fin, int inner_step)
{
    if (stopped(st, inner_step))
        return;
    const int G = (int)gridDim.x, w = xcd_remap((int)blockIdx.x, G);
    const int per = ntiles / G, rem = ntiles % G;
    const int g0 = w * per + min(w, rem), g1 = g0 + per + (w < rem ? 1 : 0);
    for (int g = g0; g < g1; ++g) {
        const int row0 = sliced_wave_row(g);
        if (row0 < sv.n)
            sliced_rows_any<B3, T>(row0, sv, val, x, epi);
    }
    double v[NR], tot[NR];
    v[0] = epi.a1;
    if (NR > 1)
        v[NR - 1] = epi.a2;
    if (block_reduce_finish<RS_SPMV_BS, NR>(slot, v, tot) && threadIdx.x == 0)
        finish_scalar(st, fin, tot[0], NR > 1 ? tot[NR - 1] : 0.0);
}

template <typename T, bool B3>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv_skip_sl(SlView sv, const T *__restrict__ val, const T *__restrict__ x,
                   T *__restrict__ y, const DevState *st, int inner_step)
{
    if (stopped(st, inner_step))
        return;
    const int row0 = sliced_wave_row(xcd_remap((int)blockIdx.x, (int)gridDim.x));
    if (row0 >= sv.n)
        return;
    EpiMult<T> epi{y};
    sliced_rows_any<B3, T>(row0, sv, val, x, epi);
}

// GMRES's scale step fused into the next step's SpMV: v = (1/H(i,i-1)) w is
// formed on the fly for every gathered entry (k_gm_scale's rule, so the
// products are bit-identical to scaling first), and each wave stores its own
// 64 rows of v (the basis vector v_i).  Skipped when the step is.
template <typename T, bool B3>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv_scale_sl(SlView sv, const T *__restrict__ val, const T *__restrict__ w,
                    T *__restrict__ vout, T *__restrict__ y, const DevState *st)
{
    if (stopped(st, true))
        return;
    const int row0 = sliced_wave_row(xcd_remap((int)blockIdx.x, (int)gridDim.x));
    if (row0 >= sv.n)
        return;
    const T sc = (T)(1. / st->hsub);
    const int r = row0 + (int)(threadIdx.x & 63);
    if (r < sv.n)
        vout[r] = (sc == T(0)) ? T(0) : sc * w[r];
    EpiMult<T> epi{y};
    sliced_rows_any<B3, T, EpiMult<T>, true>(row0, sv, val, w, epi, sc);
}

// y = A x without a reduction, skipped when the solve has stopped
template <typename T, bool B3>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv_skip(const int *__restrict__ wg_rows, const int *__restrict__ rowptr,
                const int *__restrict__ colind, BMat bm, const T *__restrict__ val,
                const T *__restrict__ x, T *__restrict__ y, const DevState *st,
                int inner_step)
{
    if (stopped(st, inner_step))
        return;
    EpiMult<T> epi{y};
    spmv_any<T, B3>(wg_rows, rowptr, colind, bm, val, x, epi);
}

// An empty asm statement taking a loaded vector's elements as inputs: the
// loads must complete before it, so they cannot be sunk under a later
// condition.  It emits no instruction.
template <typename T, int N>
__device__ __forceinline__ void pin_vec(const T *e)
{
    if constexpr (N == 4)
        asm volatile("" ::"v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]));
    else
        asm volatile("" ::"v"(e[0]), "v"(e[1]));
}

// 16-byte vector of T for the solver's internal vectors (leading dimension
// padded to a multiple of 4 elements, so every internal vector is 16-byte
// aligned); element-wise arithmetic is unchanged by the vector access.
template <typename T>
struct V16;
template <>
struct V16<double> {
    using type = double2;
    static constexpr int N = 2;
};
template <>
struct V16<float> {
    using type = float4;
    static constexpr int N = 4;
};

// =========================================================== BiCGstab kernels
// rho1 = rtilde . r; breakdown checks; beta
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_bicg_rho(int n, const T *__restrict__ rt, const T *__restrict__ r,
               DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        a += (double)(rt[i] * r[i]);
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->rho1 = tot[0];
        if (st->rho1 == 0 || st->omega == 0) {
            st->status = RSGPU_BREAKDOWN;
            st->done = 1;
            return;
        }
        if (!st->first)
            st->beta = (st->rho1 / st->rho2) * (st->alpha / st->omega);
    }
}

// p = r (first) or temp = (-omega) v + p; p = r + beta*temp
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_bicg_p(int n, const T *__restrict__ r, const T *__restrict__ v,
             T *__restrict__ p, const DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i >= n)
        return;
    if (st->first) {
        p[i] = r[i];
        return;
    }
    const T mo = (T)(-st->omega), beta = (T)st->beta;
    T temp = (mo == T(0)) ? T(0) : mo * v[i];
    temp += p[i];
    T pi = r[i];
    if (beta != T(0))
        pi += beta * temp;
    p[i] = pi;
}

// s = r + (-alpha) v; monitor(||s||)
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_bicg_s(int n, const T *__restrict__ r, const T *__restrict__ v,
             T *__restrict__ s, DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    const T ma = (T)(-st->alpha);
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        T si = r[i];
        if (ma != T(0))
            si += ma * v[i];
        s[i] = si;
        a += (double)(si * si);
    }
    double vv[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, vv, tot) && threadIdx.x == 0) {
        if (dev_monitor(st, sqrt(tot[0]))) {
            if (st->conv)
                st->half = 1;     // return x.add(alpha, phat)
            else
                st->done = 1;
        }
    }
}

// early return after s converged: x += alpha * phat
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_bicg_half(int n, const T *__restrict__ phat, T *__restrict__ x,
                DevState *st, RedSlot slot)
{
    if (st->done || !st->half)
        return;
    const T al = (T)st->alpha;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        if (al != T(0))
            x[i] += al * phat[i];
    double v[1] = {0.0}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0)
        st->done = 1;
}

// x += alpha phat; x += omega shat; r = s + (-omega) t; monitor(||r||)
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_bicg_update(int n, const T *__restrict__ phat, const T *__restrict__ shat,
                  const T *__restrict__ s, const T *__restrict__ t,
                  T *__restrict__ x, T *__restrict__ r, DevState *st,
                  RedSlot slot)
{
    if (st->done)
        return;
    const T al = (T)st->alpha, om = (T)st->omega, mo = (T)(-st->omega);
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        T xi = x[i];
        if (al != T(0))
            xi += al * phat[i];
        if (om != T(0))
            xi += om * shat[i];
        x[i] = xi;
        T ri = s[i];
        if (mo != T(0))
            ri += mo * t[i];
        r[i] = ri;
        a += (double)(ri * ri);
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->rho2 = st->rho1;
        st->first = 0;
        st->iter++;
        if (dev_monitor(st, sqrt(tot[0])))
            st->done = 1;
    }
}

// =========================================================== CG kernels
// rho = r . z; beta = rho / rho_1 after the first iteration (CG.java:71-78)
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_cg_rho(int n, const T *__restrict__ r, const T *__restrict__ z, DevState *st,
             RedSlot slot)
{
    if (st->done)
        return;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        a += (double)(r[i] * z[i]);
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->rho1 = tot[0];
        if (!st->first)
            st->beta = st->rho1 / st->rho2;
    }
}

// first: p.set(z); then p.scale(beta).add(z) (AbstractVector.scale: beta 0
// zeroes, beta 1 keeps; DenseVector.add adds in place)
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_cg_p(int n, const T *__restrict__ z, T *__restrict__ p, const DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i >= n)
        return;
    if (st->first) {
        p[i] = z[i];
        return;
    }
    const T beta = (T)st->beta;
    T pi = beta == T(0) ? T(0) : (beta == T(1) ? p[i] : beta * p[i]);
    pi += z[i];
    p[i] = pi;
}

// x.add(alpha, p); r.add(-alpha, q); monitor(||r||) (CG.java:82-87)
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_cg_update(int n, const T *__restrict__ p, const T *__restrict__ q, T *__restrict__ x,
                T *__restrict__ r, DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    const T al = (T)st->alpha, ma = (T)(-st->alpha);
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        if (al != T(0))
            x[i] += al * p[i];
        T ri = r[i];
        if (ma != T(0))
            ri += ma * q[i];
        r[i] = ri;
        a += (double)(ri * ri);
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->rho2 = st->rho1;
        st->first = 0;
        st->iter++;
        if (dev_monitor(st, sqrt(tot[0])))
            st->done = 1;
    }
}

// =========================================================== BiCG, CGS, QMR kernels
// (BiCG.java:74-121, CGS.java:78-121, QMR.java:136-243).  MTJ's vector
// operations keep their special cases: scale(0) zeroes, scale(1) keeps,
// add(0, y) and set(0, y) are a no-op and a zeroing.

// dot(x, y) (norm: dot(x, x)) reduced and handed to finish_scalar
__global__ void __launch_bounds__(RS_VBS)
    k_dot_fin(int n, const double *__restrict__ x, const double *__restrict__ y, DevState *st,
              RedSlot slot, int fin)
{
    if (st->done)
        return;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        a += x[i] * y[i];
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0)
        finish_scalar(st, fin, tot[0], 0.0);
}

// a.scale(beta).add(b) on two vector pairs, or a = b on the first iteration
__global__ void __launch_bounds__(RS_VBS)
    k_two_scale_add(int n, double *__restrict__ a1, const double *__restrict__ b1,
                    double *__restrict__ a2, const double *__restrict__ b2,
                    const DevState *st, const double *beta1, const double *beta2)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i >= n)
        return;
    if (st->first) {
        a1[i] = b1[i];
        a2[i] = b2[i];
        return;
    }
    const double c1 = *beta1, c2 = *beta2;
    const double x1 = c1 == 0 ? 0.0 : (c1 == 1 ? a1[i] : c1 * a1[i]);
    const double x2 = c2 == 0 ? 0.0 : (c2 == 1 ? a2[i] : c2 * a2[i]);
    a1[i] = x1 + b1[i];
    a2[i] = x2 + b2[i];
}

// BiCG: x += alpha p; r += -alpha q; rtilde += -alpha qtilde; rho2 = rho1;
// monitor(||r||)
__global__ void __launch_bounds__(RS_VBS)
    k_bicg_upd(int n, const double *__restrict__ p, const double *__restrict__ q,
               const double *__restrict__ qt, double *__restrict__ x, double *__restrict__ r,
               double *__restrict__ rt, DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    const double al = st->alpha, ma = -st->alpha;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        if (al != 0)
            x[i] += al * p[i];
        double ri = r[i];
        if (ma != 0) {
            ri += ma * q[i];
            rt[i] += ma * qt[i];
        }
        r[i] = ri;
        a += ri * ri;
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->rho2 = st->rho1;
        st->first = 0;
        st->iter++;
        if (dev_monitor(st, sqrt(tot[0])))
            st->done = 1;
    }
}

// CGS: first u = r, p = u; else u = r + beta q; sum = q + beta p;
// p = u + beta sum (set then add, add(0, .) a no-op)
__global__ void __launch_bounds__(RS_VBS)
    k_cgs_ups(int n, const double *__restrict__ r, const double *__restrict__ q,
              double *__restrict__ u, double *__restrict__ sum, double *__restrict__ p,
              const DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i >= n)
        return;
    if (st->first) {
        u[i] = r[i];
        p[i] = r[i];
        return;
    }
    const double beta = st->beta;
    double ui = r[i], si = q[i], pi = p[i];
    if (beta != 0) {
        ui += beta * q[i];
        si += beta * pi;
    }
    pi = ui;
    if (beta != 0)
        pi += beta * si;
    u[i] = ui;
    sum[i] = si;
    p[i] = pi;
}

// CGS: q = -alpha vhat + u (set(-alpha, vhat).add(u)); sum = u + q
__global__ void __launch_bounds__(RS_VBS)
    k_cgs_qsum(int n, const double *__restrict__ vh, const double *__restrict__ u,
               double *__restrict__ q, double *__restrict__ sum, const DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i >= n)
        return;
    const double ma = -st->alpha;
    double qi = ma == 0 ? 0.0 : ma * vh[i];
    qi += u[i];
    q[i] = qi;
    sum[i] = u[i] + qi;
}

// CGS: x += alpha uhat; r += -alpha qhat; rho2 = rho1; monitor(||r||)
__global__ void __launch_bounds__(RS_VBS)
    k_cgs_upd(int n, const double *__restrict__ uh, const double *__restrict__ qh,
              double *__restrict__ x, double *__restrict__ r, DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    const double al = st->alpha, ma = -st->alpha;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        if (al != 0)
            x[i] += al * uh[i];
        double ri = r[i];
        if (ma != 0)
            ri += ma * qh[i];
        r[i] = ri;
        a += ri * ri;
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->rho2 = st->rho1;
        st->first = 0;
        st->iter++;
        if (dev_monitor(st, sqrt(tot[0])))
            st->done = 1;
    }
}

__global__ void __launch_bounds__(RS_VBS)
    k_copy2(int n, const double *__restrict__ src, double *__restrict__ d1,
            double *__restrict__ d2, const DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i < n) {
        d1[i] = src[i];
        d2[i] = src[i];
    }
}

// QMR loop head: breakdown on rho or xi == 0; v = (1/rho) v_tld;
// y.scale(1/rho); w = (1/xi) w_tld; z.scale(1/xi); delta = z . y
__global__ void __launch_bounds__(RS_VBS)
    k_qmr_head(int n, const double *__restrict__ vt, const double *__restrict__ wt,
               double *__restrict__ v, double *__restrict__ y, double *__restrict__ w,
               double *__restrict__ z, DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    const double rho = st->q_rho, xi = st->q_xi;
    if (rho == 0 || xi == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            breakdown(st);
        return;
    }
    const double ir = 1 / rho, ix = 1 / xi;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        v[i] = ir == 0 ? 0.0 : ir * vt[i];
        const double yi = ir == 0 ? 0.0 : (ir == 1 ? y[i] : ir * y[i]);
        y[i] = yi;
        w[i] = ix == 0 ? 0.0 : ix * wt[i];
        const double zi = ix == 0 ? 0.0 : (ix == 1 ? z[i] : ix * z[i]);
        z[i] = zi;
        a += zi * yi;
    }
    double vv[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, vv, tot) && threadIdx.x == 0)
        finish_scalar(st, FIN_QMR_DELTA, tot[0], 0.0);
}

// QMR: v_tld = -beta v + p_tld (set then add); w_tld = -beta w
__global__ void __launch_bounds__(RS_VBS)
    k_qmr_vw(int n, const double *__restrict__ v, const double *__restrict__ pt,
             const double *__restrict__ w, double *__restrict__ vt, double *__restrict__ wt,
             const DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i >= n)
        return;
    const double mb = -st->q_beta;
    double vi = mb == 0 ? 0.0 : mb * v[i];
    vt[i] = vi + pt[i];
    wt[i] = mb == 0 ? 0.0 : mb * w[i];
}

// QMR: first d = eta p, s = eta p_tld; else d.scale(val).add(eta, p),
// s likewise; x += d; r += -1 s; monitor(||r||)
__global__ void __launch_bounds__(RS_VBS)
    k_qmr_upd(int n, const double *__restrict__ p, const double *__restrict__ pt,
              double *__restrict__ d, double *__restrict__ s, double *__restrict__ x,
              double *__restrict__ r, DevState *st, RedSlot slot)
{
    if (st->done)
        return;
    const double eta = st->q_eta, val = st->q_val;
    const bool first = st->first != 0;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS) {
        double di, si;
        if (first) {
            di = eta == 0 ? 0.0 : eta * p[i];
            si = eta == 0 ? 0.0 : eta * pt[i];
        } else {
            di = val == 0 ? 0.0 : (val == 1 ? d[i] : val * d[i]);
            si = val == 0 ? 0.0 : (val == 1 ? s[i] : val * s[i]);
            if (eta != 0) {
                di += eta * p[i];
                si += eta * pt[i];
            }
        }
        d[i] = di;
        s[i] = si;
        x[i] += di;
        const double ri = r[i] + -1.0 * si;
        r[i] = ri;
        a += ri * ri;
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->first = 0;
        st->iter++;
        if (dev_monitor(st, sqrt(tot[0])))
            st->done = 1;
    }
}

// y = A' x (add = 0) or y = y + A' x (add = 1, transMultAdd with alpha 1) on
// the transposed CSR, skipped when the solve has stopped
template <class Epi>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmvt_skip(const int *__restrict__ wg_rows, const int *__restrict__ rowptr,
                 const int *__restrict__ colind, const double *__restrict__ val,
                 const double *__restrict__ x, Epi epi, const DevState *st)
{
    if (st->done)
        return;
    spmv_core<double, RS_SPMV_BS, RS_SPMV_NNZ>(wg_rows, rowptr, colind, val, x, epi);
}

// =========================================================== GMRES kernels
// v0 = (1/normr) r; s = [normr, 0...]; inner check on |s_0|
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_start(int n, const T *__restrict__ r, T *__restrict__ v0, DevState *st,
               RedSlot slot)
{
    if (st->done)
        return;
    const T sc = (T)(1 / st->normr);
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        v0[i] = (sc == T(0)) ? T(0) : sc * r[i];
    double v[1] = {0.0}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        for (int k = 0; k <= st->m; ++k)
            st->s[k] = 0;
        st->s[0] = st->normr;
        st->gi = 0;
        st->cycle_end = 0;
        if (st->m <= 0 || dev_monitor(st, fabs(st->s[0]))) {
            st->cycle_end = 1;
            if (st->status)
                st->done = 1;
        }
    }
}

// H(i+1,i) = ||w|| from its square, previous Givens rotations on column i,
// the new rotation (GivensRotation(x, y), GivensRotation.java:42-60), the
// rotated rhs, iteration count and the inner monitor on |s_{i+1}|.
__device__ void gm_finish_column(DevState *st, int i, double wnorm2)
{
    const int ld = st->m + 1;
    double *Hc = st->H + i * ld;
    Hc[i + 1] = sqrt(wnorm2);
    st->hsub = Hc[i + 1];
    // previous rotations on column i
    for (int q = 0; q < i; ++q) {
        const double c = st->gc[q], s = st->gs[q];
        const double temp = c * Hc[q] + s * Hc[q + 1];
        Hc[q + 1] = -s * Hc[q] + c * Hc[q + 1];
        Hc[q] = temp;
    }
    // new rotation (GivensRotation(x, y))
    const double x = Hc[i], y = Hc[i + 1];
    const double roe = fabs(x) > fabs(y) ? x : y;
    const double scale = fabs(x) + fabs(y);
    double c, s;
    if (scale != 0) {
        const double xs = x / scale, ys = y / scale;
        double rr = scale * sqrt(xs * xs + ys * ys);
        if (roe < 0)
            rr *= -1;
        c = x / rr;
        s = y / rr;
    } else {
        c = 1;
        s = 0;
    }
    st->gc[i] = c;
    st->gs[i] = s;
    {
        const double temp = c * Hc[i] + s * Hc[i + 1];
        Hc[i + 1] = -s * Hc[i] + c * Hc[i + 1];
        Hc[i] = temp;
    }
    {
        double *sv = st->s;
        const double temp = c * sv[i] + s * sv[i + 1];
        sv[i + 1] = -s * sv[i] + c * sv[i + 1];
        sv[i] = temp;
    }
    st->iter++;
    st->gi = i + 1;
    if (!(i + 1 < st->m) || dev_monitor(st, fabs(st->s[i + 1]))) {
        st->cycle_end = 1;
        if (st->status)
            st->done = 1;
    }
}

// MGS step k of column i: w -= H(k-1,i) v_{k-1} (k > 0), H(k,i) = w . v_k
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_mgs(int n, int i, int k, T *__restrict__ w, const T *__restrict__ vprev,
             const T *__restrict__ vk, DevState *st, RedSlot slot)
{
    if (stopped(st, true))
        return;
    using VT = typename V16<T>::type;
    constexpr int N = V16<T>::N;
    const int ld = st->m + 1;
    const T h = k > 0 ? (T)(-st->H[(k - 1) + i * ld]) : T(0);
    const bool upd = k > 0 && h != T(0);
    double a = 0;
    const int nv = n / N;
    const int stride = gridDim.x * RS_VBS;
    for (int j = blockIdx.x * RS_VBS + threadIdx.x; j < nv; j += stride) {
        VT wv = reinterpret_cast<const VT *>(w)[j];
        const VT kv = reinterpret_cast<const VT *>(vk)[j];
        T *we = reinterpret_cast<T *>(&wv);
        const T *ke = reinterpret_cast<const T *>(&kv);
        if (upd) {
            const VT pv = reinterpret_cast<const VT *>(vprev)[j];
            const T *pe = reinterpret_cast<const T *>(&pv);
#pragma unroll
            for (int c = 0; c < N; ++c)
                we[c] += h * pe[c];
            reinterpret_cast<VT *>(w)[j] = wv;
        }
#pragma unroll
        for (int c = 0; c < N; ++c)
            a += (double)(we[c] * ke[c]);
    }
    for (int j = nv * N + blockIdx.x * RS_VBS + threadIdx.x; j < n; j += stride) {
        T wj = w[j];
        if (upd) {
            wj += h * vprev[j];
            w[j] = wj;
        }
        a += (double)(wj * vk[j]);
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0)
        st->H[k + i * ld] = tot[0];
}

// last MGS update w -= H(i,i) v_i, H(i+1,i) = ||w||, Givens, monitor
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_last(int n, int i, T *__restrict__ w, const T *__restrict__ vi,
              DevState *st, RedSlot slot)
{
    if (stopped(st, true))
        return;
    using VT = typename V16<T>::type;
    constexpr int N = V16<T>::N;
    const int ld = st->m + 1;
    const T h = (T)(-st->H[i + i * ld]);
    const bool upd = h != T(0);
    double a = 0;
    const int nv = n / N;
    const int stride = gridDim.x * RS_VBS;
    for (int j = blockIdx.x * RS_VBS + threadIdx.x; j < nv; j += stride) {
        VT wv = reinterpret_cast<const VT *>(w)[j];
        T *we = reinterpret_cast<T *>(&wv);
        if (upd) {
            const VT iv = reinterpret_cast<const VT *>(vi)[j];
            const T *ie = reinterpret_cast<const T *>(&iv);
#pragma unroll
            for (int c = 0; c < N; ++c)
                we[c] += h * ie[c];
            reinterpret_cast<VT *>(w)[j] = wv;
        }
#pragma unroll
        for (int c = 0; c < N; ++c)
            a += (double)(we[c] * we[c]);
    }
    for (int j = nv * N + blockIdx.x * RS_VBS + threadIdx.x; j < n; j += stride) {
        T wj = w[j];
        if (upd) {
            wj += h * vi[j];
            w[j] = wj;
        }
        a += (double)(wj * wj);
    }
    double v[1] = {a}, tot[1];
    __shared__ double lh[3 * 64];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x < 64) {
        if (i + 1 < 64)
            gm_finish_column_wave(st, i, sqrt(__shfl(tot[0], 0, 64)), lh);
        else if (threadIdx.x == 0)
            gm_finish_column(st, i, tot[0]);
    }
}

// Classical Gram-Schmidt (no reference counterpart; the FP32 inner GMRES of
// the mixed form, and both passes of the FP64 GMRES's CGS2): all
// projections of column i from one pass over w and the basis,
// H(k,i) = w . v_k for k = 0..i.
#define RS_CGS_CHUNK 8
#define RS_CGS_ET 4
// Every thread accumulates up to 8 dot products in registers over its
// elements, each wave reduces them with shuffles, one LDS exchange per 8
// columns gives the group's partials, and the last group to arrive sums
// the partials with one wave per column.  The summation order is fixed by
// the grid, so results are reproducible run to run.
template <typename T, int CC>
__device__ __forceinline__ void cgs_accum(int n, int nv, int stride, const T *__restrict__ w,
                                          const T *__restrict__ V0, size_t ldv,
                                          double (&a)[RS_CGS_CHUNK])
{
    using VT = typename V16<T>::type;
    constexpr int N = V16<T>::N;
    // RS_CGS_ET elements per thread per pass, every load of the pass in
    // flight before the first product (indices past the end clamp to the
    // last element; their products are selected away), elements summed in
    // the same order as one at a time
    constexpr int ET = RS_CGS_ET;
    for (int j0 = blockIdx.x * RS_VBS + threadIdx.x; j0 < nv; j0 += ET * stride) {
        VT wv[ET], kv[ET][CC];
#pragma unroll
        for (int e = 0; e < ET; ++e) {
            const int je = min(j0 + e * stride, nv - 1);
            wv[e] = reinterpret_cast<const VT *>(w)[je];
#pragma unroll
            for (int q = 0; q < CC; ++q)
                kv[e][q] = reinterpret_cast<const VT *>(V0 + (size_t)q * ldv)[je];
        }
#pragma unroll
        for (int e = 0; e < ET; ++e) {
            const bool on = j0 + e * stride < nv;
            const T *we = reinterpret_cast<const T *>(&wv[e]);
#pragma unroll
            for (int q = 0; q < CC; ++q) {
                const T *ke = reinterpret_cast<const T *>(&kv[e][q]);
#pragma unroll
                for (int c = 0; c < N; ++c) {
                    const double u = a[q] + (double)(we[c] * ke[c]);
                    a[q] = on ? u : a[q];
                }
            }
        }
    }
    for (int j = nv * N + blockIdx.x * RS_VBS + threadIdx.x; j < n; j += stride)
#pragma unroll
        for (int q = 0; q < CC; ++q)
            a[q] += (double)(w[j] * V0[(size_t)q * ldv + j]);
}

// second: the projections go to st->h2 (CGS2's second pass) instead of
// column i of H
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_cgs_dots(int n, int i, const T *__restrict__ w, const T *__restrict__ V, size_t ldv,
                  DevState *st, RedSlot slot, int second)
{
    constexpr int N = V16<T>::N;
    constexpr int NW = RS_VBS / 64;
    __shared__ double sh[NW][RS_CGS_CHUNK];
    __shared__ int last;
    // state reads before the stop test, so they share its round trip
    double *Hb = st->H, *h2b = st->h2;
    const int ld = st->m + 1;
    if (stopped(st, true))
        return;
    const int cnt = i + 1;
    const int G = (int)gridDim.x;
    const int nv = n / N;
    const int stride = G * RS_VBS;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c0 = 0; c0 < cnt; c0 += RS_CGS_CHUNK) {
        const int cc = min(RS_CGS_CHUNK, cnt - c0);
        const T *V0 = V + (size_t)c0 * ldv;
        double a[RS_CGS_CHUNK];
#pragma unroll
        for (int q = 0; q < RS_CGS_CHUNK; ++q)
            a[q] = 0;
        // the column count is a compile-time constant in each case, so all
        // of a chunk's loads are unconditional and in flight together
        switch (cc) {
#define RS_CGS_CASE(K)                                                      \
    case K:                                                                 \
        cgs_accum<T, K>(n, nv, stride, w, V0, ldv, a);                       \
        break;
            RS_CGS_CASE(1) RS_CGS_CASE(2) RS_CGS_CASE(3) RS_CGS_CASE(4)
            RS_CGS_CASE(5) RS_CGS_CASE(6) RS_CGS_CASE(7) RS_CGS_CASE(8)
#undef RS_CGS_CASE
        default:
            break;
        }
#pragma unroll
        for (int q = 0; q < RS_CGS_CHUNK; ++q)
            if (q < cc) {
                const double r = wave_sum(a[q]);
                if (lane == 0)
                    sh[wid][q] = r;
            }
        __syncthreads();
        // wave 0 publishes the group's partials (its own store wait below
        // covers them before the arrival ticket)
        if (threadIdx.x < cc) {
            double r = 0;
#pragma unroll
            for (int u = 0; u < NW; ++u)
                r += sh[u][threadIdx.x];
            __hip_atomic_store(&slot.partials[(size_t)(c0 + threadIdx.x) * G + blockIdx.x], r,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
    // arrival ticket (as block_reduce_finish); thread 0 is in wave 0, which
    // made every partial store of this group
    if (threadIdx.x == 0) {
        wait_vm_stores();
        const int ng = G < RS_NGROUP ? G : RS_NGROUP;
        const int g = (int)blockIdx.x % ng;
        const int gsize = (G - g + ng - 1) / ng;
        unsigned *gt = slot.ticket + (1 + g) * RS_TPAD;
        int fin = 0;
        const unsigned t =
            __hip_atomic_fetch_add(gt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int)t == gsize - 1) {
            __hip_atomic_store(gt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned u = __hip_atomic_fetch_add(slot.ticket, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
            if ((int)u == ng - 1) {
                __hip_atomic_store(slot.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fin = 1;
            }
        }
        last = fin;
    }
    __syncthreads();
    if (!last)
        return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // wave wid sums columns wid, wid + NW, ...: per lane the partials
    // b = lane, lane + 64, ... in that order.  All loads of a batch (8
    // columns x 4 partials per lane) are issued before the first add, so the
    // finishing group waits for one memory round trip per batch instead of
    // one per column and partial.
    for (int k0 = wid; k0 < cnt; k0 += 8 * NW) {
        double acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            acc[c] = 0;
        for (int b0 = lane; b0 < G; b0 += 256) {
            double t[8][4];
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    t[c][u] = __hip_atomic_load(
                        &slot.partials[(size_t)min(k0 + c * NW, cnt - 1) * G +
                                       min(b0 + 64 * u, G - 1)],
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                acc[c] += t[c][0];
#pragma unroll
                for (int u = 1; u < 4; ++u) {
                    const double a = acc[c] + t[c][u];
                    acc[c] = b0 + 64 * u < G ? a : acc[c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const double r = wave_sum(acc[c]);
            if (lane == 0 && k0 + c * NW < cnt) {
                if (second)
                    h2b[k0 + c * NW] = r;
                else
                    Hb[k0 + c * NW + i * ld] = r;
            }
        }
    }
}

// sixteen basis vectors of element j from column k0 (indices past i clamp
// to v_i, so every load is valid and none sits under a condition)
template <typename T, typename VT>
__device__ __forceinline__ void cgs_load16(const T *__restrict__ V, size_t ldv, int i, int k0,
                                           int j, VT (&kv)[16])
{
#pragma unroll
    for (int q = 0; q < 16; ++q)
        kv[q] = reinterpret_cast<const VT *>(V + (size_t)min(k0 + q, i) * ldv)[j];
}

// w_j -= h_k v_kj for k = k0 .. min(k0 + 15, i), in k order, skipped where
// h_k == 0 (add's alpha == 0 rule): the loads are pinned before any update,
// so they overlap and none sinks under the skip rule; skipped updates are
// selected away
template <typename T, typename VT, int N>
__device__ __forceinline__ void cgs_apply16(VT &wv, VT (&kv)[16], const double *Hc, int i,
                                            int k0)
{
    T *we = reinterpret_cast<T *>(&wv);
    T hq[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        hq[q] = (T)(-Hc[min(k0 + q, i)]);
#pragma unroll
    for (int q = 0; q < 16; ++q)
        pin_vec<T, N>(reinterpret_cast<T *>(&kv[q]));
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const bool on = k0 + q <= i && hq[q] != T(0);
        const T *ke = reinterpret_cast<const T *>(&kv[q]);
#pragma unroll
        for (int c = 0; c < N; ++c) {
            const T u = we[c] + hq[q] * ke[c];
            we[c] = on ? u : we[c];
        }
    }
}

// w -= sum_k h_k v_k (each element in k order, add's alpha == 0 skip).
// PASS 0 (classical Gram-Schmidt) and PASS 2 (CGS2's second pass): then
// H(i+1,i) = ||w||, Givens, monitor; PASS 2 first adds its coefficients
// (st->h2) into column i of H.  PASS 1 (CGS2's first pass): the update
// alone.  h is column i of H for PASS 0 and 1, st->h2 for PASS 2.
template <typename T, int PASS>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_cgs_last(int n, int i, T *__restrict__ w, const T *__restrict__ V, size_t ldv,
                  DevState *st, RedSlot slot)
{
    using VT = typename V16<T>::type;
    constexpr int N = V16<T>::N;
    const int nv = n / N;
    const int stride = gridDim.x * RS_VBS;
    const int j0 = blockIdx.x * RS_VBS + threadIdx.x;
    // The first element's vector and first sixteen basis vectors are loaded
    // before the stop test, together with the state reads: the kernel then
    // waits for one memory round trip before its first update, not two.
    const int jf = nv > 0 ? min(j0, nv - 1) : 0;
    VT wv{}, kv[16];
    if (nv > 0) {
        wv = reinterpret_cast<const VT *>(w)[jf];
        cgs_load16<T, VT>(V, ldv, i, 0, jf, kv);
    }
    double *Hb = st->H;
    const double *h2b = st->h2;
    const int ld = st->m + 1;
    if (stopped(st, true))
        return;
    const double *Hc = PASS == 2 ? h2b : Hb + i * ld;
    double a = 0;
    for (int j = j0; j < nv; j += stride) {
        if (j != j0) {
            wv = reinterpret_cast<const VT *>(w)[j];
            cgs_load16<T, VT>(V, ldv, i, 0, j, kv);
        }
        cgs_apply16<T, VT, N>(wv, kv, Hc, i, 0);
        for (int k0 = 16; k0 <= i; k0 += 16) {
            cgs_load16<T, VT>(V, ldv, i, k0, j, kv);
            cgs_apply16<T, VT, N>(wv, kv, Hc, i, k0);
        }
        reinterpret_cast<VT *>(w)[j] = wv;
        const T *we = reinterpret_cast<const T *>(&wv);
#pragma unroll
        for (int c = 0; c < N; ++c)
            a += (double)(we[c] * we[c]);
    }
    for (int j = nv * N + blockIdx.x * RS_VBS + threadIdx.x; j < n; j += stride) {
        T wj = w[j];
        for (int k = 0; k <= i; ++k) {
            const T h = (T)(-Hc[k]);
            if (h != T(0))
                wj += h * V[(size_t)k * ldv + j];
        }
        w[j] = wj;
        a += (double)(wj * wj);
    }
    if (PASS == 1)
        return;
    double v[1] = {a}, tot[1];
    __shared__ double lh[3 * 64];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x < 64) {
        if (i + 1 < 64) {
            // H(:, i) = h1 + h2, one entry per lane (each lane reads back
            // its own store in gm_finish_column_wave)
            if (PASS == 2 && (int)threadIdx.x <= i)
                Hb[threadIdx.x + i * ld] += h2b[threadIdx.x];
            gm_finish_column_wave(st, i, sqrt(__shfl(tot[0], 0, 64)), lh);
        } else if (threadIdx.x == 0) {
            if (PASS == 2) {
                double *Hi = Hb + i * ld;
                for (int k = 0; k <= i; ++k)
                    Hi[k] += h2b[k];
            }
            gm_finish_column(st, i, tot[0]);
        }
    }
}

// FP64 CGS2's first update and second projection pass in one sweep over
// the basis: w -= V h1 (k_gm_cgs_last<double, 1>'s update: each element in
// k order, add's alpha == 0 skip), then st->h2 = V' w over the updated
// elements (the second pass of k_gm_cgs_dots, summed per thread in element
// order, then over the group's waves, then over the groups in order).  Each
// basis element is loaded once and used by both, so the step reads the
// basis three times instead of four.  One element per lane: the i + 1 basis
// values of an element stay in registers between the update and the
// products.
#define RS_UD_COLS 32
__global__ void __launch_bounds__(RS_VBS)
    k_gm_cgs_upd_dots(int n, int i, double *__restrict__ w, const double *__restrict__ V,
                      size_t ldv, DevState *st, RedSlot slot)
{
    constexpr int NW = RS_VBS / 64;
    __shared__ double sh[NW][RS_UD_COLS];
    __shared__ double hs[RS_UD_COLS];
    __shared__ int last;
    double *Hb = st->H, *h2b = st->h2;
    const int ld = st->m + 1;
    if (stopped(st, true))
        return;
    const int cnt = i + 1;
    const int G = (int)gridDim.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x < RS_UD_COLS)
        hs[threadIdx.x] = -Hb[min((int)threadIdx.x, i) + i * ld];
    __syncthreads();
    double a[RS_UD_COLS];
#pragma unroll
    for (int q = 0; q < RS_UD_COLS; ++q)
        a[q] = 0;
    for (int j = blockIdx.x * RS_VBS + threadIdx.x; j < n; j += G * RS_VBS) {
        double wj = w[j];
        double kv[RS_UD_COLS];
#pragma unroll
        for (int c0 = 0; c0 < RS_UD_COLS; c0 += 8) {
            if (c0 < cnt) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    kv[c0 + u] = V[(size_t)min(c0 + u, i) * ldv + j];
            }
        }
#pragma unroll
        for (int q = 0; q < RS_UD_COLS; ++q) {
            if (q < cnt) {
                const double h = hs[q];
                const double u = wj + h * kv[q];
                wj = h != 0.0 ? u : wj;
            }
        }
        w[j] = wj;
#pragma unroll
        for (int q = 0; q < RS_UD_COLS; ++q)
            if (q < cnt)
                a[q] += wj * kv[q];
    }
#pragma unroll
    for (int q = 0; q < RS_UD_COLS; ++q)
        if (q < cnt) {
            const double r = wave_sum(a[q]);
            if (lane == 0)
                sh[wid][q] = r;
        }
    __syncthreads();
    if (threadIdx.x < cnt) {
        double r = 0;
#pragma unroll
        for (int u = 0; u < NW; ++u)
            r += sh[u][threadIdx.x];
        __hip_atomic_store(&slot.partials[(size_t)threadIdx.x * G + blockIdx.x], r,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // arrival ticket (as block_reduce_finish); thread 0 is in wave 0, which
    // made every partial store of this group
    if (threadIdx.x == 0) {
        wait_vm_stores();
        const int ng = G < RS_NGROUP ? G : RS_NGROUP;
        const int g = (int)blockIdx.x % ng;
        const int gsize = (G - g + ng - 1) / ng;
        unsigned *gt = slot.ticket + (1 + g) * RS_TPAD;
        int fin = 0;
        const unsigned t =
            __hip_atomic_fetch_add(gt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int)t == gsize - 1) {
            __hip_atomic_store(gt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned u = __hip_atomic_fetch_add(slot.ticket, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
            if ((int)u == ng - 1) {
                __hip_atomic_store(slot.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fin = 1;
            }
        }
        last = fin;
    }
    __syncthreads();
    if (!last)
        return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // wave wid sums columns wid, wid + NW, ... over the groups in order
    for (int k = wid; k < cnt; k += NW) {
        double acc = 0;
        for (int b0 = lane; b0 < G; b0 += 256) {
            double t4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                t4[u] = __hip_atomic_load(&slot.partials[(size_t)k * G + min(b0 + 64 * u, G - 1)],
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            acc += t4[0];
#pragma unroll
            for (int u = 1; u < 4; ++u) {
                const double s_ = acc + t4[u];
                acc = b0 + 64 * u < G ? s_ : acc;
            }
        }
        const double r = wave_sum(acc);
        if (lane == 0)
            h2b[k] = r;
    }
}

// FP64 GMRES with delayed reorthogonalisation (DCGS2; no reference
// counterpart, the orthogonalisation of the FP64 drop-in's fused step): the
// update of step i.  With s = st->ds (the reorthogonalisation coefficients
// of v~_i = V(:, i) on V(:, 0..i-1)), z = H(0..i-1, i) (the projections of
// w on them), rho = st->drho and c = st->dc from the fused step's finishing
// wave:
//   v_i = (v~_i - sum_k s_k v_k) / rho      (stored over v~_i; i > 0)
//   u   = (w - sum_k z_k v_k - c v_i) / rho  (stored over w)
// each element's sums in k order, one pass over the basis; then
// eta = ||u||, the subdiagonal Hr(i+1, i) and the next step's scale.
__global__ void __launch_bounds__(RS_VBS)
    k_gm_dcgs2_update(int n, int i, double *__restrict__ w, double *__restrict__ V, size_t ldv,
                      DevState *st, RedSlot slot)
{
    __shared__ double zs[RS_UD_COLS], ss[RS_UD_COLS];
    if (stopped(st, true))
        return;
    const int ld = st->m + 1;
    const double *Hc = st->H + i * ld;
    const double rho = st->drho, c = st->dc;
    if ((int)threadIdx.x < i) {
        zs[threadIdx.x] = Hc[threadIdx.x];
        ss[threadIdx.x] = st->ds[threadIdx.x];
    }
    __syncthreads();
    double *vi = V + (size_t)i * ldv;
    double a = 0;
    for (int j = blockIdx.x * RS_VBS + threadIdx.x; j < n; j += gridDim.x * RS_VBS) {
        double vt = vi[j], wj = w[j];
        for (int k0 = 0; k0 < i; k0 += 8) {
            double kv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                kv[u] = V[(size_t)min(k0 + u, i - 1) * ldv + j];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                asm volatile("" ::"v"(kv[u]));
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool on = k0 + u < i;
                const double v1 = vt - ss[min(k0 + u, i - 1)] * kv[u];
                const double w1 = wj - zs[min(k0 + u, i - 1)] * kv[u];
                vt = on ? v1 : vt;
                wj = on ? w1 : wj;
            }
        }
        const double vn = i > 0 ? vt / rho : vt;
        const double u = (wj - c * vn) / rho;
        if (i > 0)
            vi[j] = vn;
        w[j] = u;
        a += u * u;
    }
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        const double eta = sqrt(tot[0]);
        st->Hr[i + 1 + i * ld] = eta;
        st->hsub = eta;
    }
}

// DCGS2: the end of a full cycle finishes column m - 1 (the fused step of a
// following column would have): s = V(:, 0..m-1)' v~_m and alpha = v~_m'
// v~_m with v~_m = w / st->hsub formed on the fly as the fused step forms
// its input, then dcgs_finish_column_wave.  Skipped when the cycle ended
// earlier.
__global__ void __launch_bounds__(RS_VBS)
    k_gm_dcgs2_close(int n, int m, const double *__restrict__ w, const double *__restrict__ V,
                     size_t ldv, DevState *st, RedSlot slot)
{
    constexpr int NW = RS_VBS / 64;
    __shared__ double sh[NW][32];
    __shared__ double colsum[32];
    __shared__ double lh[3 * 64];
    __shared__ int last;
    if (stopped(st, true))
        return;
    const int G = (int)gridDim.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const double sc = 1. / st->hsub;
    double a[32];
#pragma unroll
    for (int q = 0; q < 32; ++q)
        a[q] = 0;
    for (int j = blockIdx.x * RS_VBS + threadIdx.x; j < n; j += G * RS_VBS) {
        const double vt = sc == 0.0 ? 0.0 : sc * w[j];
#pragma unroll
        for (int k0 = 0; k0 < 32; k0 += 8) {
            if (k0 <= m) {
                double kv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    kv[u] = V[(size_t)min(k0 + u, m - 1) * ldv + j];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    asm volatile("" ::"v"(kv[u]));
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int k = k0 + u;
                    const double f = k < m ? kv[u] : vt;
                    const double p = a[k] + vt * f;
                    a[k] = k <= m ? p : a[k];
                }
            }
        }
    }
    {
        const double r = wave_sum_cols32(a);
        if ((lane & 1) == 0 && (lane >> 1) <= m)
            sh[wid][lane >> 1] = r;
    }
    __syncthreads();
    if ((int)threadIdx.x <= m) {
        double r = 0;
#pragma unroll
        for (int u = 0; u < NW; ++u)
            r += sh[u][threadIdx.x];
        __hip_atomic_store(&slot.partials[(size_t)threadIdx.x * G + blockIdx.x], r,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) {
        wait_vm_stores();
        const int ng = G < RS_NGROUP ? G : RS_NGROUP;
        const int g = (int)blockIdx.x % ng;
        const int gsize = (G - g + ng - 1) / ng;
        unsigned *gt = slot.ticket + (1 + g) * RS_TPAD;
        int fin = 0;
        const unsigned t =
            __hip_atomic_fetch_add(gt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int)t == gsize - 1) {
            __hip_atomic_store(gt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned u = __hip_atomic_fetch_add(slot.ticket, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
            if ((int)u == ng - 1) {
                __hip_atomic_store(slot.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fin = 1;
            }
        }
        last = fin;
    }
    __syncthreads();
    if (!last)
        return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int k = wid; k <= m; k += NW) {
        double acc = 0;
        for (int b1 = lane; b1 < G; b1 += 256) {
            double t4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                t4[u] = __hip_atomic_load(&slot.partials[(size_t)k * G + min(b1 + 64 * u, G - 1)],
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            acc += t4[0];
#pragma unroll
            for (int u = 1; u < 4; ++u) {
                const double s_ = acc + t4[u];
                acc = b1 + 64 * u < G ? s_ : acc;
            }
        }
        const double r = wave_sum(acc);
        if (lane == 0)
            colsum[k] = r;
    }
    __syncthreads();
    if (wid == 0)
        dcgs_finish_column_wave(st, m - 1, lh, colsum);
}

// v_{i+1} = (1/H(i+1,i)) w
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_scale(int n, const T *__restrict__ w, T *__restrict__ vn,
               const DevState *st)
{
    if (stopped(st, true))
        return;
    const int j = blockIdx.x * RS_VBS + threadIdx.x;
    if (j >= n)
        return;
    const T sc = (T)(1. / st->hsub);
    vn[j] = (sc == T(0)) ? T(0) : sc * w[j];
}

// dtrtrs on the leading gi x gi triangle of H (singular -> status 14)
// One wave.  Up to 64 columns the triangle is staged in LDS with one load
// per lane and column, and lane i carries s[i]: each column's division is
// formed by every lane and the updates of s[0..k-1] run one per lane, each
// s[i] still updated in descending k (the sequential loop's operations and
// order, bit for bit).
__global__ void k_gm_backsolve(DevState *st)
{
    __shared__ double Hs[64 * 64];
    if (st->done)
        return;
    const int ld = st->m + 1, m = st->gi;
    const double *H = st->H;
    double *s = st->s;
    const int lane = threadIdx.x;
    if (m > 64) {
        if (lane != 0)
            return;
        for (int k = 0; k < m; ++k)
            if (H[k + k * ld] == 0) {
                st->status = RSGPU_SINGULAR;
                st->done = 1;
                return;
            }
        for (int k = m - 1; k >= 0; --k)
            if (s[k] != 0) {
                s[k] = s[k] / H[k + k * ld];
                for (int i = 0; i < k; ++i)
                    s[i] = s[i] - s[k] * H[i + k * ld];
            }
        return;
    }
    const bool zero = lane < m && H[lane + lane * ld] == 0;
    if (__any(zero)) {
        if (lane == 0) {
            st->status = RSGPU_SINGULAR;
            st->done = 1;
        }
        return;
    }
    for (int k = 0; k < m; ++k)
        if (lane <= k)
            Hs[k * 64 + lane] = H[lane + k * ld];
    double si = lane < m ? s[lane] : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = m - 1; k >= 0; --k) {
        const double sk = __shfl(si, k, 64);
        if (sk != 0) {
            const double q = sk / Hs[k * 64 + k];
            if (lane == k)
                si = q;
            else if (lane < k)
                si = si - q * Hs[k * 64 + lane];
        }
    }
    if (lane < m)
        s[lane] = si;
}

// x += s_j v_j, j = 0..gi-1, in that order per element
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_xupdate(int n, const T *__restrict__ V, size_t ldv, T *__restrict__ x,
                 const DevState *st)
{
    if (st->done)
        return;
    const int j = blockIdx.x * RS_VBS + threadIdx.x;
    if (j >= n)
        return;
    const int m = st->gi;
    T xj = x[j];
    for (int q = 0; q < m; ++q) {
        const T sq = (T)st->s[q];
        if (sq != T(0))
            xj += sq * V[(size_t)q * ldv + j];
    }
    x[j] = xj;
}

// ||r|| at the end of a cycle: iter.next(), cycle reset, outer monitor
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_gm_outer(int n, const T *__restrict__ r, DevState *st, RedSlot slot,
               int first)
{
    if (st->done)
        return;
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        a += (double)(r[i] * r[i]);
    double v[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, v, tot) && threadIdx.x == 0) {
        st->normr = sqrt(tot[0]);
        if (!first)
            st->iter++;
        st->cycle_end = 0;
        if (dev_monitor(st, st->normr))
            st->done = 1;
    }
}

// IR: x += z (FP64 x, correction in T)
template <typename T>
__global__ void __launch_bounds__(RS_VBS)
    k_ir_add(int n, const T *__restrict__ z, double *__restrict__ x,
             DevState *st)
{
    if (st->done)
        return;
    const int i = blockIdx.x * RS_VBS + threadIdx.x;
    if (i < n)
        x[i] += (double)z[i];
}

__global__ void k_ir_next(DevState *st)
{
    if (threadIdx.x == 0 && !st->done)
        st->iter++;
}

// Cycle-end sequence: the preconditioner apply of the outer residual must
// run although the inner loop has set cycle_end.
__global__ void k_clear_cycle(DevState *st)
{
    if (threadIdx.x == 0 && !st->done)
        st->cycle_end = 0;
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, size_t n, T v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}

// =========================================================== host drivers
namespace {

struct Vecs {
    int n;
};

template <typename T>
const T *csr_values(rsgpu_csr *A);
template <>
const double *csr_values<double>(rsgpu_csr *A) { return A->val.p; }
template <>
const float *csr_values<float>(rsgpu_csr *A) { return A->val32.p; }

void state_init(rsgpu_solver *S, DevState *dst, DevBuf<double> &hist,
                int max_iter, double rtol, double atol, double dtol, int m,
                double *small, double *dcgs = nullptr)
{
    DevState h;
    std::memset(&h, 0, sizeof(h));
    h.rtol = rtol;
    h.atol = atol;
    h.dtol = dtol;
    h.max_iter = max_iter;
    h.first = 1;
    h.rho1 = h.rho2 = h.alpha = h.beta = h.omega = 1;
    h.q_gamma = 1;     // QMR.java:138-139 initial values
    h.q_eta = -1;
    h.hist = hist.p;
    h.hist_cap = (int)hist.n;
    h.m = m;
    if (small) {
        h.H = small;
        h.s = small + (size_t)(m + 1) * m;
        h.gc = h.s + (m + 1);
        h.gs = h.gc + (m + 1);
        h.h2 = h.gs + (m + 1);
    }
    if (dcgs) {
        h.Hr = dcgs;
        h.ds = dcgs + (size_t)(m + 1) * m;
    }
    RS_HIP(hipMemcpyAsync(dst, &h, sizeof(h), hipMemcpyHostToDevice,
                          S->ctx->stream));
}

RedSlot slot_of(rsgpu_solver *S) { return RedSlot{S->partials.p, S->ticket.p}; }

// The solver's two stop flags into the host's pinned buffer: one launch in
// place of two 4-byte copies (each a copy kernel of its own); the host reads
// them after the event recorded behind this kernel.
__global__ void k_post_flags(const DevState *__restrict__ st, int *dst)
{
    if (threadIdx.x == 0) {
        const int d = st->done, c = st->cycle_end;
        __hip_atomic_store(dst, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(dst + 1, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// host view of the flags, read one step behind
struct Poller {
    rsgpu_solver *S;
    DevState *st;
    int k = 0;
    bool pending = false;
    void post()
    {
        hipStream_t s = S->ctx->stream;
        hipLaunchKernelGGL(k_post_flags, dim3(1), dim3(64), 0, s, st,
                           S->flags_dev + 4 * (k & 1));
        RS_CHECK_LAUNCH();
        RS_HIP(hipEventRecord(S->ev[k & 1], s));
        k++;
    }
    // wait for the previous post; returns {done, cycle_end}
    void wait_prev(int &done, int &cend)
    {
        done = 0;
        cend = 0;
        if (k < 2)
            return;
        const int q = (k - 2) & 1;
        RS_HIP(hipEventSynchronize(S->ev[q]));
        done = S->flags_host[4 * q];
        cend = S->flags_host[4 * q + 1];
    }
    void wait_last(int &done, int &cend)
    {
        const int q = (k - 1) & 1;
        RS_HIP(hipEventSynchronize(S->ev[q]));
        done = S->flags_host[4 * q];
        cend = S->flags_host[4 * q + 1];
    }
};

template <typename T>
BMat bmat_of(const rsgpu_csr *A)
{
    return BMat{A->brp.p, A->bcol.p, spmv_plan<T>(A).coff_ptr()};
}

// k_spmv_red's grid: the workgroups the device holds at once (occupancy x
// CUs, queried once per instance), at most one per plan group
template <typename T, class Epi, int NR, bool B3>
static int red_grid(int ngroups)
{
    static int cap = 0;
    if (cap == 0) {
        int per_cu = 0, cus = 0, dev = 0;
        RS_HIP(hipGetDevice(&dev));
        RS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void *)k_spmv_red<T, Epi, NR, B3>, RS_SPMV_BS, 0));
        RS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cap = std::max(1, per_cu) * std::max(1, cus);
    }
    return std::max(1, std::min(ngroups, cap));
}

template <typename T, class Epi, int NR, bool B3>
static int red_grid_sl(int ntiles)
{
    static int cap = 0;
    if (cap == 0) {
        int per_cu = 0, cus = 0, dev = 0;
        RS_HIP(hipGetDevice(&dev));
        RS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void *)k_spmv_red_sl<T, Epi, NR, B3>, RS_SPMV_BS, 0));
        RS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cap = std::max(1, per_cu) * std::max(1, cus);
    }
    return std::max(1, std::min(ntiles, cap));
}

// The solve streams the sliced copy of A in precision T when it holds one:
// FP64 when this solve filled it, FP32 whenever it exists (rsgpu_csr_sync_f32
// fills it with the FP32 CSR copy).
template <typename T>
static const T *sliced_vals(const rsgpu_solver *S, const rsgpu_csr *A)
{
    if (!A->sl.ok())
        return nullptr;
    if (sizeof(T) == 8)
        return S->sl64 == A ? (const T *)A->sl.v64.p : nullptr;
    return (const T *)A->sl.v32.p;
}

template <typename T, class Epi, int NR>
void launch_red(rsgpu_solver *S, rsgpu_csr *A, const T *x, Epi epi, DevState *st, int fin,
                int inner)
{
    hipStream_t s = S->ctx->stream;
    if (const T *sv = sliced_vals<T>(S, A)) {
        const int ntiles = sliced_tiles(A->n);
        if (A->sl.b3)
            hipLaunchKernelGGL((k_spmv_red_sl<T, Epi, NR, true>),
                               dim3(red_grid_sl<T, Epi, NR, true>(ntiles)), dim3(RS_SPMV_BS), 0,
                               s, A->slview(), ntiles, sv, x, epi, st, slot_of(S), fin, inner);
        else
            hipLaunchKernelGGL((k_spmv_red_sl<T, Epi, NR, false>),
                               dim3(red_grid_sl<T, Epi, NR, false>(ntiles)), dim3(RS_SPMV_BS), 0,
                               s, A->slview(), ntiles, sv, x, epi, st, slot_of(S), fin, inner);
        RS_CHECK_LAUNCH();
        return;
    }
    const SpmvPlan &pl = spmv_plan<T>(A);
    const dim3 g(A->bs == 3 ? red_grid<T, Epi, NR, true>(pl.nwg)
                            : red_grid<T, Epi, NR, false>(pl.nwg)),
        b(RS_SPMV_BS);
    if (A->bs == 3)
        hipLaunchKernelGGL((k_spmv_red<T, Epi, NR, true>), g, b, 0, s, pl.wg_rows.p, pl.nwg,
                           A->rowptr.p, A->colind.p, bmat_of<T>(A), csr_values<T>(A), x, epi, st,
                           slot_of(S), fin, inner);
    else
        hipLaunchKernelGGL((k_spmv_red<T, Epi, NR, false>), g, b, 0, s, pl.wg_rows.p, pl.nwg,
                           A->rowptr.p, A->colind.p, bmat_of<T>(A), csr_values<T>(A), x, epi, st,
                           slot_of(S), fin, inner);
    RS_CHECK_LAUNCH();
}

template <typename T>
void launch_spmv_red(rsgpu_solver *S, rsgpu_csr *A, const T *x, EpiMultDot<T> epi,
                     DevState *st, int fin, int inner)
{
    if (epi.nd <= 1)
        launch_red<T, EpiMultDot<T>, 1>(S, A, x, epi, st, fin, inner);
    else
        launch_red<T, EpiMultDot<T>, 2>(S, A, x, epi, st, fin, inner);
}

template <typename T>
void launch_resid(rsgpu_solver *S, rsgpu_csr *A, const T *b, const T *x, T *r,
                  T *copy, float *r32, DevState *st, int fin)
{
    launch_red<T, EpiResidNorm<T>, 1>(S, A, x, EpiResidNorm<T>{r, copy, b, r32}, st, fin, 0);
}

template <typename T>
void launch_spmv_skip(rsgpu_solver *S, rsgpu_csr *A, const T *x, T *y, DevState *st)
{
    hipStream_t s = S->ctx->stream;
    if (const T *sv = sliced_vals<T>(S, A)) {
        hipLaunchKernelGGL((A->sl.b3 ? k_spmv_skip_sl<T, true> : k_spmv_skip_sl<T, false>),
                           dim3(sliced_tiles(A->n)), dim3(RS_SPMV_BS), 0, s, A->slview(), sv, x,
                           y, st, 1);
        RS_CHECK_LAUNCH();
        return;
    }
    const SpmvPlan &pl = spmv_plan<T>(A);
    const dim3 g(pl.nwg), b(RS_SPMV_BS);
    if (A->bs == 3)
        hipLaunchKernelGGL((k_spmv_skip<T, true>), g, b, 0, s, pl.wg_rows.p, A->rowptr.p,
                           A->colind.p, bmat_of<T>(A), csr_values<T>(A), x, y, st, 1);
    else
        hipLaunchKernelGGL((k_spmv_skip<T, false>), g, b, 0, s, pl.wg_rows.p, A->rowptr.p,
                           A->colind.p, bmat_of<T>(A), csr_values<T>(A), x, y, st, 1);
}

// ------------------------------------------------------------ BiCGstab
// M.apply, or M.transApply when the solve is the transposed one
// (rsgpu_solver_opts.prec_transpose: M is the preconditioner of A, the
// matrix is A')
static void papply(rsgpu_solver *S, rsgpu_prec *M, const double *b, double *x,
                   const DevState *st)
{
    if (S->o.prec_transpose)
        prec_trans_apply(M, b, x, st);
    else
        prec_apply(M, b, x, st);
}

static void papply(rsgpu_solver *S, rsgpu_prec *M, const float *b, float *x,
                   const DevState *st)
{
    if (S->o.prec_transpose)
        throw RsError{RSGPU_BAD_ARGUMENT, "the transposed solve runs in FP64"};
    prec_apply(M, b, x, st);
}

// work: p s phat shat t v r rtilde (8 n)
template <typename T>
void run_bicgstab(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const T *b,
                  T *x, T *work, DevState *st)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const size_t L = S->ldv;
    T *p = work, *sv = work + L, *phat = work + 2 * L, *shat = work + 3 * L,
      *t = work + 4 * L, *v = work + 5 * L, *r = work + 6 * L, *rt = work + 7 * L;
    const int G = grid_for(n), E = (n + RS_VBS - 1) / RS_VBS;
    RedSlot sl = slot_of(S);
    // r = b - A x; rtilde = r; monitor(||r||) at iter 0
    launch_resid<T>(S, A, b, x, r, rt, nullptr, st, FIN_INIT_NORM);
    Poller P{S, st};
    int done = 0, cend = 0;
    for (;;) {
        hipLaunchKernelGGL(k_bicg_rho<T>, dim3(G), dim3(RS_VBS), 0, s, n, rt, r,
                           st, sl);
        hipLaunchKernelGGL(k_bicg_p<T>, dim3(E), dim3(RS_VBS), 0, s, n, r, v, p,
                           st);
        papply(S, M, p, phat, st);
        launch_spmv_red<T>(S, A, phat, EpiMultDot<T>{v, rt, nullptr, 1}, st,
                           FIN_ALPHA, 0);
        hipLaunchKernelGGL(k_bicg_s<T>, dim3(G), dim3(RS_VBS), 0, s, n, r, v, sv,
                           st, sl);
        hipLaunchKernelGGL(k_bicg_half<T>, dim3(G), dim3(RS_VBS), 0, s, n, phat,
                           x, st, sl);
        papply(S, M, sv, shat, st);
        launch_spmv_red<T>(S, A, shat, EpiMultDot<T>{t, sv, nullptr, 2}, st,
                           FIN_OMEGA, 0);
        hipLaunchKernelGGL(k_bicg_update<T>, dim3(G), dim3(RS_VBS), 0, s, n,
                           phat, shat, sv, t, x, r, st, sl);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_prev(done, cend);
        if (done)
            break;
    }
    P.wait_last(done, cend);
}

// ------------------------------------------------------------ CG
// work: r z p q (CG.java:51-55)
template <typename T>
void run_cg(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const T *b, T *x, T *work,
            DevState *st)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const size_t L = S->ldv;
    T *r = work, *z = work + L, *p = work + 2 * L, *q = work + 3 * L;
    const int G = grid_for(n), E = (n + RS_VBS - 1) / RS_VBS;
    RedSlot sl = slot_of(S);
    // r = b - A x; monitor(||r||) at iter 0
    launch_resid<T>(S, A, b, x, r, nullptr, nullptr, st, FIN_INIT_NORM);
    Poller P{S, st};
    int done = 0, cend = 0;
    for (;;) {
        papply(S, M, r, z, st);
        hipLaunchKernelGGL(k_cg_rho<T>, dim3(G), dim3(RS_VBS), 0, s, n, r, z, st, sl);
        hipLaunchKernelGGL(k_cg_p<T>, dim3(E), dim3(RS_VBS), 0, s, n, z, p, st);
        launch_spmv_red<T>(S, A, p, EpiMultDot<T>{q, p, nullptr, 1}, st, FIN_CG_ALPHA, 0);
        hipLaunchKernelGGL(k_cg_update<T>, dim3(G), dim3(RS_VBS), 0, s, n, p, q, x, r, st, sl);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_prev(done, cend);
        if (done)
            break;
    }
    P.wait_last(done, cend);
}

// ------------------------------------------------------------ BiCG, CGS, QMR
// M.transApply (the solve's apply when the solve is the transposed one)
static void ptapply(rsgpu_solver *S, rsgpu_prec *M, const double *b, double *x,
                    const DevState *st)
{
    if (S->o.prec_transpose)
        prec_apply(M, b, x, st);
    else
        prec_trans_apply(M, b, x, st);
}

// transposed CSR of A with current values (once per solve)
static void prepare_transpose(rsgpu_csr *A)
{
    csr_build_transpose(A);
    if (A->nnz)
        csr_launch_gather(A->ctx->stream, A->tperm.p, A->val.p, A->tval.p, A->nnz);
}

// y = A' x, or y = y + A' x (transMultAdd, alpha 1)
static void launch_spmvt(rsgpu_solver *S, rsgpu_csr *A, const double *x, double *y, bool add,
                         const DevState *st)
{
    if (A->ncols == 0)
        return;
    hipStream_t s = S->ctx->stream;
    const dim3 g(A->tplan.nwg), b(RS_SPMV_BS);
    if (!add)
        hipLaunchKernelGGL((k_spmvt_skip<EpiPlain<double>>), g, b, 0, s, A->tplan.wg_rows.p,
                           A->trowptr.p, A->tcolind.p, A->tval.p, x, EpiPlain<double>{y}, st);
    else
        hipLaunchKernelGGL((k_spmvt_skip<EpiTransAdd<double>>), g, b, 0, s,
                           A->tplan.wg_rows.p, A->trowptr.p, A->tcolind.p, A->tval.p, x,
                           EpiTransAdd<double>{y, 1.0, 1.0}, st);
    RS_CHECK_LAUNCH();
}

// BiCG.java:74-121; work: r rtilde z ztilde p ptilde q qtilde
void run_bicg(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const double *b, double *x,
              double *work, DevState *st)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const size_t L = S->ldv;
    double *r = work, *rt = work + L, *z = work + 2 * L, *zt = work + 3 * L;
    double *p = work + 4 * L, *pt = work + 5 * L, *q = work + 6 * L, *qt = work + 7 * L;
    const int G = grid_for(n), E = (n + RS_VBS - 1) / RS_VBS;
    RedSlot sl = slot_of(S);
    prepare_transpose(A);
    launch_resid<double>(S, A, b, x, r, rt, nullptr, st, FIN_INIT_NORM);
    Poller P{S, st};
    int done = 0, cend = 0;
    for (;;) {
        papply(S, M, r, z, st);
        ptapply(S, M, rt, zt, st);
        hipLaunchKernelGGL(k_dot_fin, dim3(G), dim3(RS_VBS), 0, s, n, z, rt, st, sl,
                           (int)FIN_RHO_BREAK);
        hipLaunchKernelGGL(k_two_scale_add, dim3(E), dim3(RS_VBS), 0, s, n, p, z, pt, zt, st,
                           &st->beta, &st->beta);
        launch_spmv_red<double>(S, A, p, EpiMultDot<double>{q, pt, nullptr, 1}, st,
                                FIN_CG_ALPHA, 0);
        launch_spmvt(S, A, pt, qt, false, st);
        hipLaunchKernelGGL(k_bicg_upd, dim3(G), dim3(RS_VBS), 0, s, n, p, q, qt, x, r, rt, st,
                           sl);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_prev(done, cend);
        if (done)
            break;
    }
    P.wait_last(done, cend);
}

// CGS.java:78-121; work: r rtilde u p phat vhat q sum uhat qhat
void run_cgs(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const double *b, double *x,
             double *work, DevState *st)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const size_t L = S->ldv;
    double *r = work, *rt = work + L, *u = work + 2 * L, *p = work + 3 * L;
    double *ph = work + 4 * L, *vh = work + 5 * L, *q = work + 6 * L, *sum = work + 7 * L;
    double *uh = work + 8 * L, *qh = work + 9 * L;
    const int G = grid_for(n), E = (n + RS_VBS - 1) / RS_VBS;
    RedSlot sl = slot_of(S);
    launch_resid<double>(S, A, b, x, r, rt, nullptr, st, FIN_INIT_NORM);
    Poller P{S, st};
    int done = 0, cend = 0;
    for (;;) {
        hipLaunchKernelGGL(k_dot_fin, dim3(G), dim3(RS_VBS), 0, s, n, rt, r, st, sl,
                           (int)FIN_RHO_BREAK);
        hipLaunchKernelGGL(k_cgs_ups, dim3(E), dim3(RS_VBS), 0, s, n, r, q, u, sum, p, st);
        papply(S, M, p, ph, st);
        launch_spmv_red<double>(S, A, ph, EpiMultDot<double>{vh, rt, nullptr, 1}, st,
                                FIN_CG_ALPHA, 0);
        hipLaunchKernelGGL(k_cgs_qsum, dim3(E), dim3(RS_VBS), 0, s, n, vh, u, q, sum, st);
        papply(S, M, sum, uh, st);
        launch_spmv_skip<double>(S, A, uh, qh, st);
        hipLaunchKernelGGL(k_cgs_upd, dim3(G), dim3(RS_VBS), 0, s, n, uh, qh, x, r, st, sl);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_prev(done, cend);
        if (done)
            break;
    }
    P.wait_last(done, cend);
}

// QMR.java:136-243 with M1 = M2 = M; work: r y z v w p q d s v_tld w_tld
// y_tld z_tld p_tld
void run_qmr(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const double *b, double *x,
             double *work, DevState *st)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const size_t L = S->ldv;
    double *V[14];
    for (int k = 0; k < 14; ++k)
        V[k] = work + (size_t)k * L;
    double *r = V[0], *y = V[1], *z = V[2], *v = V[3], *w = V[4], *p = V[5], *q = V[6];
    double *d = V[7], *sv = V[8], *vt = V[9], *wt = V[10], *yt = V[11], *zt = V[12],
           *pt = V[13];
    const int G = grid_for(n), E = (n + RS_VBS - 1) / RS_VBS;
    RedSlot sl = slot_of(S);
    prepare_transpose(A);
    launch_resid<double>(S, A, b, x, r, nullptr, nullptr, st, FIN_INIT_NORM);
    hipLaunchKernelGGL(k_copy2, dim3(E), dim3(RS_VBS), 0, s, n, r, vt, wt, st);
    papply(S, M, vt, y, st);
    hipLaunchKernelGGL(k_dot_fin, dim3(G), dim3(RS_VBS), 0, s, n, y, y, st, sl,
                       (int)FIN_QMR_RHO0);
    ptapply(S, M, wt, z, st);
    hipLaunchKernelGGL(k_dot_fin, dim3(G), dim3(RS_VBS), 0, s, n, z, z, st, sl,
                       (int)FIN_QMR_XI0);
    Poller P{S, st};
    int done = 0, cend = 0;
    for (;;) {
        hipLaunchKernelGGL(k_qmr_head, dim3(G), dim3(RS_VBS), 0, s, n, vt, wt, v, y, w, z, st,
                           sl);
        papply(S, M, y, yt, st);
        ptapply(S, M, z, zt, st);
        hipLaunchKernelGGL(k_two_scale_add, dim3(E), dim3(RS_VBS), 0, s, n, p, yt, q, zt, st,
                           &st->q_c1, &st->q_c2);
        launch_spmv_red<double>(S, A, p, EpiMultDot<double>{pt, q, nullptr, 1}, st,
                                FIN_QMR_EP, 0);
        hipLaunchKernelGGL(k_qmr_vw, dim3(E), dim3(RS_VBS), 0, s, n, v, pt, w, vt, wt, st);
        papply(S, M, vt, y, st);
        hipLaunchKernelGGL(k_dot_fin, dim3(G), dim3(RS_VBS), 0, s, n, y, y, st, sl,
                           (int)FIN_QMR_RHO);
        launch_spmvt(S, A, q, wt, true, st);
        ptapply(S, M, wt, z, st);
        hipLaunchKernelGGL(k_dot_fin, dim3(G), dim3(RS_VBS), 0, s, n, z, z, st, sl,
                           (int)FIN_QMR_XI);
        hipLaunchKernelGGL(k_qmr_upd, dim3(G), dim3(RS_VBS), 0, s, n, p, pt, d, sv, x, r, st,
                           sl);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_prev(done, cend);
        if (done)
            break;
    }
    P.wait_last(done, cend);
}

// ------------------------------------------------------------ GMRES(m)
// work: w u r + (m+1) basis vectors
template <typename T>
void run_gmres(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const T *b, T *x,
               T *work, DevState *st, int m)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const size_t L = S->ldv;
    T *w = work, *u = work + L, *r = work + 2 * L, *V = work + 3 * L;
    const int G = grid_for(n), E = (n + RS_VBS - 1) / RS_VBS;
    RedSlot sl = slot_of(S);
    // u = b - A x; r = M u; normr; outer monitor at iter 0
    launch_resid<T>(S, A, b, x, u, nullptr, nullptr, st, FIN_NONE);
    papply(S, M, u, r, st);
    hipLaunchKernelGGL(k_gm_outer<T>, dim3(G), dim3(RS_VBS), 0, s, n, r, st, sl,
                       1);
    Poller P{S, st};
    int done = 0, cend = 0;
    const int every = std::max(1, S->o.check_every);
    for (;;) {
        hipLaunchKernelGGL(k_gm_start<T>, dim3(G), dim3(RS_VBS), 0, s, n, r, V,
                           st, sl);
        const T *slv = sliced_vals<T>(S, A);
        // fused step: SpMV, block ILU apply and the first projection pass in
        // one launch; its input and output vectors alternate between w and u
        const bool fused = slv && S->o.cgs && !S->o.prec_transpose && M &&
                           (size_t)(m + 1) * M->nblocks <= S->partials.n &&
                           gmres_fused_ok(A, M, (int)sizeof(T), m);
        T *wcur = w, *wnext = u;
        // FP64 CGS2: the first update and the second projection share a pass
        // (RSGPU_CGS_UD=0: separate kernels)
        const char *eud = getenv("RSGPU_CGS_UD");
        const bool cgs_ud = sizeof(T) == 8 && m + 1 <= RS_UD_COLS && !(eud && eud[0] == '0');
        // FP64 fused steps: delayed reorthogonalisation (DCGS2) in place of
        // CGS2 (RSGPU_DCGS2=0: CGS2)
        const char *edc = getenv("RSGPU_DCGS2");
        const bool dcgs = fused && sizeof(T) == 8 && S->dcgs.p && !(edc && edc[0] == '0') &&
                          (size_t)2 * (m + 1) * M->nblocks <= S->partials.n;
        for (int i = 0; i < m; ++i) {
            T *vi = V + (size_t)i * L;
            if (fused) {
                launch_gmres_fused<T>(s, A, slv, M, i > 0 ? wcur : vi, vi, i > 0, wnext, V, L,
                                      i, st, sl, nullptr, dcgs);
                ++S->fused_steps;
                if (dcgs) {
                    hipLaunchKernelGGL(k_gm_dcgs2_update, dim3(G), dim3(RS_VBS), 0, s, n, i,
                                       (double *)wnext, (double *)V, L, st, sl);
                } else if (sizeof(T) == 4) {
                    hipLaunchKernelGGL((k_gm_cgs_last<T, 0>), dim3(G), dim3(RS_VBS), 0, s, n, i,
                                       wnext, V, L, st, sl);
                } else {
                    // CGS2: w -= V h1 and h2 = V' w in one pass, then w -= V h2
                    if (cgs_ud)
                        hipLaunchKernelGGL(k_gm_cgs_upd_dots, dim3(G), dim3(RS_VBS), 0, s, n,
                                           i, (double *)wnext, (const double *)V, L, st, sl);
                    else {
                        hipLaunchKernelGGL((k_gm_cgs_last<T, 1>), dim3(G), dim3(RS_VBS), 0, s,
                                           n, i, wnext, V, L, st, sl);
                        hipLaunchKernelGGL(k_gm_cgs_dots<T>, dim3(std::max(1, G / 4)),
                                           dim3(RS_VBS), 0, s, n, i, wnext, V, L, st, sl, 1);
                    }
                    hipLaunchKernelGGL((k_gm_cgs_last<T, 2>), dim3(G), dim3(RS_VBS), 0, s, n, i,
                                       wnext, V, L, st, sl);
                }
                RS_CHECK_LAUNCH();
                std::swap(wcur, wnext);
                if ((i + 1) % every == 0 && i + 1 < m) {
                    P.post();
                    P.wait_prev(done, cend);
                    if (done || cend)
                        break;
                }
                continue;
            }
            if (i > 0 && slv) {
                // v_i = w / H(i,i-1) stored by the SpMV that reads it
                hipLaunchKernelGGL((A->sl.b3 ? k_spmv_scale_sl<T, true>
                                             : k_spmv_scale_sl<T, false>),
                                   dim3(sliced_tiles(n)), dim3(RS_SPMV_BS), 0, s, A->slview(),
                                   slv, w, vi, u, st);
                RS_CHECK_LAUNCH();
            } else {
                launch_spmv_skip<T>(S, A, vi, u, st);
            }
            papply(S, M, u, w, st);
            if (S->o.cgs && sizeof(T) == 4) {
                hipLaunchKernelGGL(k_gm_cgs_dots<T>, dim3(std::max(1, G / 4)), dim3(RS_VBS), 0,
                                   s, n, i, w, V, L, st, sl, 0);
                hipLaunchKernelGGL((k_gm_cgs_last<T, 0>), dim3(G), dim3(RS_VBS), 0, s, n, i, w,
                                   V, L, st, sl);
            } else if (S->o.cgs) {
                // CGS2: two classical passes, H(:,i) = h1 + h2
                hipLaunchKernelGGL(k_gm_cgs_dots<T>, dim3(std::max(1, G / 4)), dim3(RS_VBS), 0,
                                   s, n, i, w, V, L, st, sl, 0);
                if (cgs_ud) {
                    hipLaunchKernelGGL(k_gm_cgs_upd_dots, dim3(G), dim3(RS_VBS), 0, s, n, i,
                                       (double *)w, (const double *)V, L, st, sl);
                } else {
                    hipLaunchKernelGGL((k_gm_cgs_last<T, 1>), dim3(G), dim3(RS_VBS), 0, s, n,
                                       i, w, V, L, st, sl);
                    hipLaunchKernelGGL(k_gm_cgs_dots<T>, dim3(std::max(1, G / 4)),
                                       dim3(RS_VBS), 0, s, n, i, w, V, L, st, sl, 1);
                }
                hipLaunchKernelGGL((k_gm_cgs_last<T, 2>), dim3(G), dim3(RS_VBS), 0, s, n, i, w,
                                   V, L, st, sl);
            } else {
                for (int k = 0; k <= i; ++k)
                    hipLaunchKernelGGL(k_gm_mgs<T>, dim3(G), dim3(RS_VBS), 0, s, n, i, k, w,
                                       k > 0 ? V + (size_t)(k - 1) * L : V, V + (size_t)k * L,
                                       st, sl);
                hipLaunchKernelGGL(k_gm_last<T>, dim3(G), dim3(RS_VBS), 0, s, n, i, w, vi, st,
                                   sl);
            }
            if (!slv)
                hipLaunchKernelGGL(k_gm_scale<T>, dim3(E), dim3(RS_VBS), 0, s, n, w,
                                   V + (size_t)(i + 1) * L, st);
            RS_CHECK_LAUNCH();
            if ((i + 1) % every == 0 && i + 1 < m) {
                P.post();
                P.wait_prev(done, cend);
                if (done || cend)
                    break;
            }
        }
        // DCGS2: a full cycle's last column is finished here
        if (dcgs)
            hipLaunchKernelGGL(k_gm_dcgs2_close, dim3(G), dim3(RS_VBS), 0, s, n, m,
                               (const double *)wcur, (const double *)V, L, st, sl);
        hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), 0, s, st);
        hipLaunchKernelGGL(k_gm_xupdate<T>, dim3(E), dim3(RS_VBS), 0, s, n, V, L, x, st);
        launch_resid<T>(S, A, b, x, u, nullptr, nullptr, st, FIN_NONE);
        // r = M u belongs to the outer loop: clear the inner-loop flag first
        hipLaunchKernelGGL(k_clear_cycle, dim3(1), dim3(64), 0, s, st);
        papply(S, M, u, r, st);
        hipLaunchKernelGGL(k_gm_outer<T>, dim3(G), dim3(RS_VBS), 0, s, n, r, st,
                           sl, 0);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_last(done, cend);
        if (done)
            break;
    }
}

} // namespace

// =========================================================== public API
static int solver_catch(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

extern "C" void rsgpu_solver_opts_default(rsgpu_solver_opts *o)
{
    std::memset(o, 0, sizeof(*o));
    o->kind = RSGPU_BICGSTAB;
    o->restart = 30;
    o->max_iter = 100000;
    o->rtol = 1e-5;
    o->atol = 1e-50;
    o->dtol = 1e5;
    o->mixed = 0;
    o->inner_max_iter = 200;
    o->inner_rtol = 1e-4;
    o->check_every = 4;
    o->prec_transpose = 0;
    o->cgs = 0;
}

extern "C" int rsgpu_solver_create(rsgpu_ctx *ctx, const rsgpu_solver_opts *opts,
                                   int n, rsgpu_solver **out)
{
    if (!ctx || !opts || !out || n < 0)
        return RSGPU_BAD_ARGUMENT;
    if (opts->kind < 0 || opts->kind > RSGPU_QMR)
        return RSGPU_BAD_ARGUMENT;
    if (opts->mixed && opts->kind > RSGPU_CG) {
        ctx->err = "the mixed form's inner solver is BiCGstab, GMRES or CG";
        return RSGPU_BAD_ARGUMENT;
    }
    if (opts->kind == RSGPU_GMRES && opts->restart <= 0) {
        ctx->err = "restart must be a positive integer";
        return RSGPU_BAD_ARGUMENT;
    }
    rsgpu_solver *S = new rsgpu_solver();
    try {
        RS_HIP(hipSetDevice(ctx->device));
        S->ctx = ctx;
        S->o = *opts;
        S->n = n;
        RS_HIP(hipMalloc((void **)&S->st, sizeof(DevState)));
        RS_HIP(hipMalloc((void **)&S->st_in, sizeof(DevState)));
        const int maxwg = std::max({grid_for(n), (n + 7) / 8, 4096});
        S->partials.alloc((size_t)maxwg * 2);
        S->ticket.alloc(2 * RS_TICKET_WORDS);
        S->ticket.zero(ctx->stream);
        const int m = std::max(1, opts->restart);
        S->gmres_small.alloc((size_t)2 * ((m + 1) * m + 4 * (m + 1)));
        if (opts->kind == RSGPU_GMRES && !opts->mixed)
            S->dcgs.alloc((size_t)(m + 1) * m + m + 2);
        size_t hcap = std::min<size_t>((size_t)opts->max_iter * 2 + 8, 1u << 22);
        S->hist.alloc(hcap);
        S->hist_in.alloc(std::min<size_t>((size_t)opts->inner_max_iter * 2 + 8,
                                          1u << 20));
        const int inner_kind = opts->mixed ? opts->kind : opts->kind;
        size_t nvec64 = 0, nvec32 = 0;
        if (!opts->mixed) {
            nvec64 = opts->kind == RSGPU_BICGSTAB ? 8
                     : opts->kind == RSGPU_GMRES  ? 3 + (size_t)m + 1
                     : opts->kind == RSGPU_CG     ? 4
                     : opts->kind == RSGPU_BICG   ? 8
                     : opts->kind == RSGPU_CGS    ? 10
                     : opts->kind == RSGPU_QMR    ? 14
                                                  : 2;
        } else {
            nvec64 = 2;   // r, scratch
            nvec32 = 2 + (inner_kind == RSGPU_GMRES ? 3 + (size_t)m + 1 : 8);
        }
        S->ldv = ((size_t)n + 3) & ~(size_t)3;
        // the solver's vectors start as zero vectors, as the reference's
        // template-built DenseVectors do (a preconditioner such as AMG reads
        // its output vector's old content)
        S->w64.alloc(std::max<size_t>(1, nvec64 * S->ldv));
        S->w64.zero(ctx->stream);
        if (nvec32) {
            S->w32.alloc(std::max<size_t>(1, nvec32 * S->ldv));
            S->w32.zero(ctx->stream);
        }
        RS_HIP(hipEventCreateWithFlags(&S->ev[0], hipEventDisableTiming));
        RS_HIP(hipEventCreateWithFlags(&S->ev[1], hipEventDisableTiming));
        RS_HIP(hipHostMalloc((void **)&S->flags_host, sizeof(int) * 8,
                             hipHostMallocDefault));
        RS_HIP(hipHostGetDevicePointer((void **)&S->flags_dev, S->flags_host, 0));
        RS_HIP(hipStreamSynchronize(ctx->stream));
    } catch (const RsError &e) {
        rsgpu_solver_destroy(S);
        return solver_catch(ctx, e);
    }
    *out = S;
    return RSGPU_OK;
}

extern "C" int rsgpu_solver_destroy(rsgpu_solver *S)
{
    if (!S)
        return RSGPU_OK;
    if (S->st)
        (void)hipFree(S->st);
    if (S->st_in)
        (void)hipFree(S->st_in);
    for (auto &e : S->ev)
        if (e)
            (void)hipEventDestroy(e);
    if (S->flags_host)
        (void)hipHostFree(S->flags_host);
    delete S;
    return RSGPU_OK;
}

namespace {

double dev_norm2(rsgpu_solver *S, const double *v, int n);

template <typename T>
void run_kind(rsgpu_solver *S, int kind, rsgpu_csr *A, rsgpu_prec *M, const T *b,
              T *x, T *work, DevState *st, int m)
{
    if (kind == RSGPU_BICGSTAB)
        run_bicgstab<T>(S, A, M, b, x, work, st);
    else if (kind == RSGPU_CG)
        run_cg<T>(S, A, M, b, x, work, st);
    else if (kind == RSGPU_GMRES)
        run_gmres<T>(S, A, M, b, x, work, st, m);
    else if constexpr (sizeof(T) == 8) {
        if (kind == RSGPU_BICG)
            run_bicg(S, A, M, b, x, work, st);
        else if (kind == RSGPU_CGS)
            run_cgs(S, A, M, b, x, work, st);
        else if (kind == RSGPU_QMR)
            run_qmr(S, A, M, b, x, work, st);
        else
            throw RsError{RSGPU_BAD_ARGUMENT, "unsupported solver"};
    } else
        throw RsError{RSGPU_BAD_ARGUMENT, "unsupported inner solver"};
}

// IR with M = preconditioner (IR.java:57-70), FP64
void run_ir(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const double *b,
            double *x, DevState *st)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    double *r = S->w64.p, *z = S->w64.p + S->ldv;
    const int E = (n + RS_VBS - 1) / RS_VBS;
    Poller P{S, st};
    int done = 0, cend = 0;
    launch_resid<double>(S, A, b, x, r, nullptr, nullptr, st, FIN_INIT_NORM);
    for (;;) {
        papply(S, M, r, z, st);
        hipLaunchKernelGGL(k_ir_add<double>, dim3(E), dim3(RS_VBS), 0, s, n, z, x,
                           st);
        hipLaunchKernelGGL(k_ir_next, dim3(1), dim3(64), 0, s, st);
        launch_resid<double>(S, A, b, x, r, nullptr, nullptr, st, FIN_INIT_NORM);
        RS_CHECK_LAUNCH();
        P.post();
        P.wait_prev(done, cend);
        if (done)
            break;
    }
    P.wait_last(done, cend);
}

// Mixed: FP64 refinement on the true residual around an FP32 Krylov solve
void run_mixed(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const double *b,
               double *x, DevState *st, rsgpu_stats *stats)
{
    hipStream_t s = S->ctx->stream;
    const int n = A->n;
    const int m = std::max(1, S->o.restart);
    double *r = S->w64.p;
    float *r32 = S->w32.p, *z32 = S->w32.p + S->ldv, *work32 = S->w32.p + 2 * S->ldv;
    const int E = (n + RS_VBS - 1) / RS_VBS;
    if (!A->f32_current())
        throw RsError{RSGPU_BAD_ARGUMENT,
                      "mixed solve needs rsgpu_csr_sync_f32 after the last value change"};
    if (M && M->kind != RSGPU_PREC_IDENTITY && M->precision != RSGPU_FP32)
        throw RsError{RSGPU_BAD_ARGUMENT, "mixed solve needs an FP32 preconditioner"};
    int inner_total = 0, outer = 0;
    DevState host;
    double *small_in = S->gmres_small.p + ((size_t)(m + 1) * m + 4 * (m + 1));
    launch_resid<double>(S, A, b, x, r, nullptr, r32, st, FIN_INIT_NORM);
    for (;;) {
        RS_HIP(hipMemcpyAsync(&host, st, sizeof(host), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        if (host.done)
            break;
        // inner FP32 solve: z32 = A^-1 r32 from z32 = 0
        state_init(S, S->st_in, S->hist_in, S->o.inner_max_iter, S->o.inner_rtol,
                   1e-50, 1e5, m, small_in);
        hipLaunchKernelGGL(k_fill<float>, dim3(E), dim3(RS_VBS), 0, s, z32,
                           (size_t)n, 0.0f);
        run_kind<float>(S, S->o.kind, A, M, r32, z32, work32, S->st_in, m);
        DevState hin;
        RS_HIP(hipMemcpyAsync(&hin, S->st_in, sizeof(hin), hipMemcpyDeviceToHost,
                              s));
        RS_HIP(hipStreamSynchronize(s));
        inner_total += hin.iter;
        // a breakdown of the inner solve still yields a usable correction
        // unless it produced non-finite values; the outer monitor decides.
        hipLaunchKernelGGL(k_ir_add<float>, dim3(E), dim3(RS_VBS), 0, s, n, z32, x,
                           st);
        hipLaunchKernelGGL(k_ir_next, dim3(1), dim3(64), 0, s, st);
        launch_resid<double>(S, A, b, x, r, nullptr, r32, st, FIN_INIT_NORM);
        outer++;
    }
    if (stats) {
        stats->inner_iterations = inner_total;
        stats->outer_iterations = outer;
    }
}

} // namespace

__global__ void k_norm2_plain(int n, const double *__restrict__ v, double *out,
                              RedSlot slot)
{
    double a = 0;
    for (int i = blockIdx.x * RS_VBS + threadIdx.x; i < n; i += gridDim.x * RS_VBS)
        a += v[i] * v[i];
    double vv[1] = {a}, tot[1];
    if (block_reduce_finish<RS_VBS, 1>(slot, vv, tot) && threadIdx.x == 0)
        *out = sqrt(tot[0]);
}

namespace {
double dev_norm2(rsgpu_solver *S, const double *v, int n)
{
    hipStream_t s = S->ctx->stream;
    double *out = S->partials.p + S->partials.n - 1;
    // partials are sized for the largest grid; the last slot is free
    RedSlot sl{S->partials.p, S->ticket.p + RS_TICKET_WORDS};
    hipLaunchKernelGGL(k_norm2_plain, dim3(grid_for(n)), dim3(RS_VBS), 0, s, n, v,
                       out, sl);
    RS_CHECK_LAUNCH();
    double h = 0;
    RS_HIP(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    return h;
}
} // namespace

extern "C" int rsgpu_solver_solve(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                                  const double *b_dev, double *x_dev,
                                  rsgpu_stats *stats)
{
    if (!S || !A)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_ctx *ctx = S->ctx;
    if (A->n != A->ncols) {
        ctx->err = "!A.isSquare()";
        return RSGPU_BAD_ARGUMENT;
    }
    if (A->n != S->n) {
        ctx->err = "b.size() != A.numRows()";
        return RSGPU_BAD_ARGUMENT;
    }
    rsgpu_prec ident;
    ident.ctx = ctx;
    ident.kind = RSGPU_PREC_IDENTITY;
    ident.n = A->n;
    ident.ready = true;
    if (!M)
        M = &ident;
    const uint64_t want = S->o.prec_transpose ? A->transpose_of : A->pattern_id;
    if (M->kind != RSGPU_PREC_IDENTITY && (!M->ready || M->pattern_id != want)) {
        ctx->err = S->o.prec_transpose
                       ? "transposed solve: the preconditioner must be set on the matrix "
                         "whose transpose is solved"
                       : "Preconditioner.setMatrix must be called before use";
        return RSGPU_BAD_ARGUMENT;
    }
    if (S->o.prec_transpose && S->o.mixed) {
        ctx->err = "the transposed solve runs in FP64";
        return RSGPU_BAD_ARGUMENT;
    }
    try {
        RS_HIP(hipSetDevice(ctx->device));
        const int m = std::max(1, S->o.restart);
        const size_t need =
            (size_t)2 * std::max({A->plan.nwg, A->plan32.nwg, A->bplan.nwg, A->bplan32.nwg,
                                  grid_for(A->n), sliced_tiles(A->n)}) + 2;
        const size_t need_cgs = (size_t)(m + 2) * grid_for(A->n);
        if (S->partials.n < std::max(need, need_cgs) + 4096)
            S->partials.alloc(std::max(need, need_cgs) + 4096);
        state_init(S, S->st, S->hist, S->o.max_iter, S->o.rtol, S->o.atol,
                   S->o.dtol, m, S->gmres_small.p, S->dcgs.p);
        S->fused_steps = 0;
        if (stats)
            std::memset(stats, 0, sizeof(*stats));
        // FP64 solves stream the sliced copy, refreshed from A's values here
        S->sl64 = nullptr;
        if (!S->o.mixed && A->sl.ok()) {
            sliced_fill<double>(A, ctx->stream);
            S->sl64 = A;
        }
        if (S->o.mixed) {
            run_mixed(S, A, M, b_dev, x_dev, S->st, stats);
        } else if (S->o.kind == RSGPU_IR) {
            run_ir(S, A, M, b_dev, x_dev, S->st);
        } else {
            run_kind<double>(S, S->o.kind, A, M, b_dev, x_dev, S->w64.p, S->st, m);
        }
        DevState h;
        RS_HIP(hipMemcpyAsync(&h, S->st, sizeof(h), hipMemcpyDeviceToHost,
                              ctx->stream));
        RS_HIP(hipStreamSynchronize(ctx->stream));
        S->last_hist = std::min(h.hist_len, h.hist_cap);
        if (stats) {
            stats->iterations = h.iter;
            stats->residual = h.residual;
            stats->status = h.status;
            stats->hist_len = h.hist_len;
            // true residual ||b - A x|| in FP64
            double *r = S->w64.p;
            spmv_launch<double>(ctx->stream, A, A->val.p, x_dev, r, 2, -1.0, b_dev);
            stats->true_residual = dev_norm2(S, r, A->n);
            stats->b_norm = dev_norm2(S, b_dev, A->n);
        }
        return h.status;
    } catch (const RsError &e) {
        return solver_catch(ctx, e);
    }
}

// One fused GMRES step on caller vectors (rsgpu_gmres_fused_step): the
// solver's inner state holds the projections, the sliced FP64 copy is
// refreshed when A's values changed since it was filled.
template <typename T>
static int fused_step_call(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M, const T *V, size_t ldv,
                           int i, T *w, double *proj)
{
    if (!S || !A || !M || !V || !w || !proj)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_ctx *ctx = S->ctx;
    const int m = std::max(1, S->o.restart);
    const bool f32 = sizeof(T) == 4;
    const char *why = nullptr;
    if (A->n != S->n || M->n != A->n)
        why = "fused step: sizes differ";
    else if (i < 0 || i + 1 > m)
        why = "fused step: column i must be below the solver's restart";
    else if (ldv < (size_t)A->n || ldv % 4 != 0 || ((uintptr_t)V & 15) || ((uintptr_t)w & 15))
        why = "fused step: ldv must be >= n and a multiple of 4, V and w 16-byte aligned";
    else if (!M->ready || M->pattern_id != A->pattern_id)
        why = "Preconditioner.setMatrix must be called before use";
    else if (f32 && !A->f32_current())
        why = "fused step: FP32 form needs rsgpu_csr_sync_f32 after the last value change";
    else if (!gmres_fused_ok(A, M, (int)sizeof(T), m))
        why = "fused step: the matrix and preconditioner do not meet its conditions";
    if (why) {
        ctx->err = why;
        return RSGPU_BAD_ARGUMENT;
    }
    try {
        RS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const size_t need = (size_t)(i + 1) * M->nblocks + 4096;
        if (S->partials.n < need)
            S->partials.alloc(need);
        const T *slv;
        if (f32) {
            slv = (const T *)A->sl.v32.p;
        } else {
            if (A->sl.gen64 != A->val_gen)
                sliced_fill<double>(A, s);
            slv = (const T *)A->sl.v64.p;
        }
        // a state of its own, set once: the kernel reads its stop flags and
        // writes the projections straight to proj
        if (!S->st_probe.p) {
            S->st_probe.alloc(1);
            state_init(S, S->st_probe.p, S->hist_in, 1, 0.0, 0.0, 1e300, m, nullptr);
        }
        launch_gmres_fused<T>(s, A, slv, M, V + (size_t)i * ldv, nullptr, false, w, V, ldv, i,
                              S->st_probe.p, slot_of(S), proj);
        return RSGPU_OK;
    } catch (const RsError &e) {
        return solver_catch(ctx, e);
    }
}

extern "C" int rsgpu_gmres_fused_step(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                                      const double *V, size_t ldv, int i, double *w,
                                      double *proj)
{
    if (M && M->precision != RSGPU_FP64) {
        if (S)
            S->ctx->err = "fused step: FP64 form needs an FP64 preconditioner";
        return RSGPU_BAD_ARGUMENT;
    }
    return fused_step_call<double>(S, A, M, V, ldv, i, w, proj);
}

extern "C" int rsgpu_gmres_fused_step_f32(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                                          const float *V, size_t ldv, int i, float *w,
                                          double *proj)
{
    if (M && M->precision != RSGPU_FP32) {
        if (S)
            S->ctx->err = "fused step: FP32 form needs an FP32 preconditioner";
        return RSGPU_BAD_ARGUMENT;
    }
    return fused_step_call<float>(S, A, M, V, ldv, i, w, proj);
}

extern "C" int rsgpu_solver_fused_steps(const rsgpu_solver *S)
{
    return S ? S->fused_steps : RSGPU_BAD_ARGUMENT;
}

extern "C" int rsgpu_solver_history(const rsgpu_solver *S, double *out, int cap)
{
    if (!S || !out)
        return RSGPU_BAD_ARGUMENT;
    const int k = std::min(cap, S->last_hist);
    if (k > 0) {
        if (hipMemcpy(out, S->hist.p, sizeof(double) * k, hipMemcpyDeviceToHost) !=
            hipSuccess)
            return RSGPU_HIP_ERROR;
    }
    return k;
}
