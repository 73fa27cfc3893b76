// Batched ensemble solves: many independent pressure systems of uneven size
// solved together, every kernel launch covering all of them.
//
// Reference: each system is solved as BiCGstab.solve
// (mtj/src/no/uib/cipr/matrix/sparse/BiCGstab.java:65-116) with
// DefaultIterationMonitor (DefaultIterationMonitor.java:133-156); the mixed
// form wraps it in the IR.solve refinement (IR.java:57-70) on the FP64 true
// residual.  The reference solves one system per call; running an ensemble
// in one launch has no counterpart there.
//
// Design.  The systems are the diagonal blocks of one block-diagonal CSR
// (rsgpu_csr_block_diag).  Every workgroup of every kernel belongs to exactly
// one system: the SpMV plans and the vector segments are cut at system
// boundaries.  Each system has its own solver state on the device
// (BatchSys): its BiCGstab scalars and monitor, and in the mixed form an
// FP64 refinement monitor.  Dot products and norms are reduced per system:
// a workgroup publishes its partial and takes a ticket on its system's
// counter, and the system's last arrival sums that system's partials in
// group order and takes the system's scalar decision.  A system that has
// stopped makes its workgroups return at once; the ILU(0) sweeps skip its
// blocks.  In the mixed form each system runs its own refinement state
// machine: when its FP32 inner solve ends, the next iteration's outer step
// (x += z, true FP64 residual, refinement monitor, inner restart) runs for
// that system alone, while the others continue their inner iterations, so no
// system waits for the slowest one between refinement steps.
//
// The fixed kernel sequence of a few iterations is captured once per solve
// into a HIP graph and replayed until every system has stopped; the host
// reads a device counter of stopped systems one replay behind.
//
// SpMV stage size (nonzeros staged in LDS per workgroup, i.e. rows per
// wavefront) is picked per matrix pattern by timing each candidate on the
// batch itself.
#include <algorithm>
#include <cmath>

#include "spmv_core.hpp"

#define RS_BVBS 256        // vector-kernel workgroup size
#define RS_BCHUNK 2048     // vector elements per workgroup (8 per lane)

// ===================================================== block-diagonal matrix
extern "C" int rsgpu_csr_block_diag(rsgpu_ctx *ctx, int count, rsgpu_csr *const *parts,
                                    rsgpu_csr **out)
{
    if (!ctx || !parts || !out || count < 1)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_csr *A = nullptr;
    try {
        long long n = 0, m = 0, nnz = 0;
        for (int k = 0; k < count; ++k) {
            if (!parts[k] || parts[k]->ctx != ctx)
                throw RsError{RSGPU_BAD_ARGUMENT, "block-diagonal parts must share the context"};
            n += parts[k]->n;
            m += parts[k]->ncols;
            nnz += parts[k]->nnz;
        }
        if (n > 0x7fffffffLL || m > 0x7fffffffLL || nnz > 0x7fffffffLL)
            throw RsError{RSGPU_BAD_ARGUMENT, "block-diagonal matrix exceeds 32-bit indices"};
        RS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        A = new rsgpu_csr();
        A->ctx = ctx;
        A->n = (int)n;
        A->ncols = (int)m;
        A->nnz = (int)nnz;
        A->h_rowptr.resize((size_t)n + 1);
        A->h_colind.resize((size_t)nnz);
        A->h_rowptr[0] = 0;
        int r0 = 0, c0 = 0, j0 = 0;
        for (int k = 0; k < count; ++k) {
            const rsgpu_csr *P = parts[k];
            for (int i = 0; i < P->n; ++i)
                A->h_rowptr[r0 + i + 1] = j0 + P->h_rowptr[i + 1];
            for (int j = 0; j < P->nnz; ++j)
                A->h_colind[j0 + j] = c0 + P->h_colind[j];
            r0 += P->n;
            c0 += P->ncols;
            j0 += P->nnz;
        }
        A->rowptr.upload(A->h_rowptr, s);
        A->colind.upload(A->h_colind, s);
        A->val.alloc(std::max<size_t>(nnz, 1));
        size_t off = 0;
        for (int k = 0; k < count; ++k) {
            if (parts[k]->nnz)
                RS_HIP(hipMemcpyAsync(A->val.p + off, parts[k]->val.p,
                                      sizeof(double) * parts[k]->nnz, hipMemcpyDeviceToDevice,
                                      s));
            off += parts[k]->nnz;
        }
        spmv_build_plans(A->h_rowptr, A->n, A, s, false);
        A->pattern_id = next_pattern_id();
        RS_HIP(hipStreamSynchronize(s));
    } catch (const RsError &e) {
        delete A;
        ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        delete A;
        ctx->err = "host out of memory";
        return RSGPU_OUT_OF_MEMORY;
    }
    *out = A;
    return RSGPU_OK;
}

// ===================================================== device state
struct BatchSys {
    DevState in;       // the running BiCGstab (FP64 plain, or the FP32 inner solve)
    DevState out;      // mixed: the FP64 refinement monitor
    int finished;      // the system's solve has ended
    int inner_total;   // mixed: inner iterations summed over refinements
    int pad[2];
};
static_assert(sizeof(BatchSys) % sizeof(int) == 0, "flag stride");

struct Seg {
    int sys, r0, r1, pad;
};

// one SpMV plan of the batch: contiguous row ranges per workgroup that never
// cross a system boundary, each group's system, each system's group range
struct BPlanDev {
    const int *wg_rows;
    const int *wg_sys;
    const int *sys_g0;
    const short *coff;   // 16-bit column offsets from each group's first row, or null
};

struct BPlan {
    int nnz_stage = 0;
    int nwg = 0;
    DevBuf<int> wg_rows, wg_sys, sys_g0;
    DevBuf<short> coff;
    BPlanDev dev() const
    {
        return BPlanDev{wg_rows.p, wg_sys.p, sys_g0.p, coff.n ? coff.p : nullptr};
    }
};

struct RedSeg {
    double *partials;
    int P;               // stride between the reductions' partial arrays
    unsigned *tickets;   // per system: RS_BTICKETS counters, RS_TPAD words apart
    unsigned *ndone;     // systems stopped
};

// ---------------------------------------------------- per-system reduction
// Arrival tickets are counted in two levels per system: group g of a system
// whose groups are g0..g1-1 takes a ticket on sub-counter (g - g0) % RS_BSUB,
// and the last arrival of each sub-counter takes one on the system's top
// counter, so no address sees more than about (g1 - g0) / RS_BSUB atomics.
#define RS_BSUB 16
#define RS_BTICKETS (RS_BSUB + 1)

// Publish this group's partial(s); true in every thread of the system's last
// arriving group, whose thread 0 then holds the system totals in tot[] (the
// partials of groups g0..g1-1 summed in a fixed order).  The store / ticket /
// acquire protocol is block_reduce_finish's, per system.
template <int BS, int NR>
__device__ bool seg_finish(const RedSeg &R, int sys, int g, int g0, int g1,
                           const double (&v)[NR], double (&tot)[NR])
{
    __shared__ double sh[BS / 64 + 1];
    __shared__ int last;
    double r[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q)
        r[q] = block_sum<BS>(v[q], sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
            __hip_atomic_store(&R.partials[(size_t)q * R.P + g], r[q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        wait_vm_stores();
        const int G = g1 - g0;
        const int ns = G < RS_BSUB ? G : RS_BSUB;
        const int j = (g - g0) % ns;
        const int jsize = (G - j + ns - 1) / ns;
        unsigned *base = R.tickets + (size_t)sys * RS_BTICKETS * RS_TPAD;
        unsigned *sub = base + (size_t)(1 + j) * RS_TPAD;
        int fin = 0;
        const unsigned k =
            __hip_atomic_fetch_add(sub, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int)k == jsize - 1) {
            __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned u =
                __hip_atomic_fetch_add(base, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((int)u == ns - 1) {
                __hip_atomic_store(base, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fin = 1;
            }
        }
        last = fin;
    }
    __syncthreads();
    if (!last)
        return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
    for (int q = 0; q < NR; ++q)
        tot[q] = block_sum<BS>(gather_sum<BS>(R.partials + (size_t)q * R.P, g0, g1), sh);
    return true;
}

// The running solve of a system has stopped.  Plain form: the system is
// finished.  Mixed form: the inner solve ended; the next outer step of the
// system takes over.
__device__ void end_inner(BatchSys *S, int mixed, unsigned *ndone)
{
    S->in.done = 1;
    if (mixed) {
        S->inner_total += S->in.iter;
    } else {
        S->finished = 1;
        __hip_atomic_fetch_add(ndone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ bool outer_due(const BatchSys *S)
{
    return S->in.done && !S->finished;
}

// ===================================================== SpMV kernels
enum {
    BF_INIT = 0,   // plain: r = b - A x, rtilde = r, monitor(||r||) at iteration 0
    BF_ALPHA,      // v = A phat, alpha = rho1 / (rtilde . v)
    BF_OMEGA,      // t = A shat, omega = (t . s) / (t . t)
    BF_OUTER,      // mixed: FP64 residual, refinement monitor, inner restart
    BF_FINAL,      // true residual and ||b|| per system (statistics)
};

template <typename T>
struct BEpiDot {   // y = A x (mult rule) with one or two dots
    T *y;
    const T *z1;
    int nd;
    double a1 = 0, a2 = 0;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot)
    {
        const T v = (dot != T(0)) ? dot : T(0);
        y[r] = v;
        a1 += (double)(z1[r] * v);
        if (nd >= 2)
            a2 += (double)(v * v);
    }
};

struct BEpiInit {  // r = b + (-1) A x; rtilde = r
    double *r, *rt;
    const double *b;
    double a1 = 0, a2 = 0;
    __device__ double init(int) const { return 0.0; }
    __device__ void row(int i, double dot)
    {
        const double v = b[i] + (-1.0) * dot;
        r[i] = v;
        rt[i] = v;
        a1 += v * v;
    }
};

struct BEpiOuter {  // FP64 residual; FP32 inner start r = rtilde = (float) res, z = 0
    const double *b;
    float *r32, *rt32, *z32;
    double a1 = 0, a2 = 0;
    __device__ double init(int) const { return 0.0; }
    __device__ void row(int i, double dot)
    {
        const double v = b[i] + (-1.0) * dot;
        const float f = (float)v;
        r32[i] = f;
        rt32[i] = f;
        z32[i] = 0.0f;
        a1 += v * v;
        a2 += (double)(f * f);
    }
};

struct BEpiFinal {  // ||b - A x||^2 and ||b||^2
    const double *b;
    double a1 = 0, a2 = 0;
    __device__ double init(int) const { return 0.0; }
    __device__ void row(int i, double dot)
    {
        const double v = b[i] + (-1.0) * dot;
        a1 += v * v;
        a2 += b[i] * b[i];
    }
};

template <typename T>
struct BEpiMult {
    T *y;
    double a1 = 0, a2 = 0;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot) { y[r] = (dot != T(0)) ? dot : T(0); }
};

__device__ void inner_restart(DevState *in)
{
    in->iter = 0;
    in->status = 0;
    in->conv = 0;
    in->half = 0;
    in->first = 1;
    in->rho1 = in->rho2 = in->alpha = in->beta = in->omega = 1;
}

template <int FIN>
__device__ __forceinline__ bool b_gate(const BatchSys *S)
{
    if (FIN == BF_OUTER)
        return outer_due(S);
    if (FIN == BF_FINAL || FIN < 0)
        return true;
    return !S->in.done;
}

template <int FIN>
__device__ void b_finish(BatchSys *S, const double *tot, int mixed, unsigned *ndone,
                         double *norms, int sys)
{
    switch (FIN) {
    case BF_INIT:
        if (dev_monitor(&S->in, sqrt(tot[0])))
            end_inner(S, mixed, ndone);
        break;
    case BF_ALPHA:
        S->in.alpha = S->in.rho1 / tot[0];
        break;
    case BF_OMEGA:
        S->in.ts = tot[0];
        S->in.tt = tot[1];
        S->in.omega = tot[0] / tot[1];
        break;
    case BF_OUTER: {
        DevState *o = &S->out;
        if (!o->first)
            o->iter++;
        o->first = 0;
        if (dev_monitor(o, sqrt(tot[0]))) {
            S->finished = 1;
            __hip_atomic_fetch_add(ndone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            inner_restart(&S->in);
            if (dev_monitor(&S->in, sqrt(tot[1])))
                S->inner_total += 0;   // the inner solve has nothing to do
            else
                S->in.done = 0;
        }
        break;
    }
    case BF_FINAL:
        norms[2 * sys] = sqrt(tot[0]);
        norms[2 * sys + 1] = sqrt(tot[1]);
        break;
    default:
        break;
    }
}

template <typename T, int NNZ, class Epi, int NR, int FIN>
__global__ void __launch_bounds__(RS_SPMV_BS)
    kb_spmv(BPlanDev pl, const int *__restrict__ rowptr, const int *__restrict__ colind,
            const T *__restrict__ val, const T *__restrict__ x, Epi epi, BatchSys *st,
            RedSeg R, int mixed, double *norms)
{
    const int g = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int sys = pl.wg_sys[g];
    BatchSys *S = st + sys;
    if (!b_gate<FIN>(S))
        return;
    spmv_core_off<T, RS_SPMV_BS, NNZ>(pl.wg_rows, rowptr, colind, pl.coff, val, x, epi);
    if (NR == 0)
        return;
    constexpr int NRR = (NR > 0) ? NR : 1;
    double v[NRR], tot[NRR];
    v[0] = epi.a1;
    if (NRR > 1)
        v[NRR - 1] = epi.a2;
    if (seg_finish<RS_SPMV_BS, NRR>(R, sys, g, pl.sys_g0[sys], pl.sys_g0[sys + 1], v, tot) &&
        threadIdx.x == 0)
        b_finish<FIN>(S, tot, mixed, R.ndone, norms, sys);
}

// ===================================================== vector kernels
#define SEG_LOOP(i) for (int i = sg.r0 + (int)threadIdx.x; i < sg.r1; i += RS_BVBS)

// rho1 = rtilde . r; breakdown checks; beta (BiCGstab.java:74-88)
template <typename T>
__global__ void __launch_bounds__(RS_BVBS)
    kb_rho(const Seg *__restrict__ seg, const int *__restrict__ sys_g0,
           const T *__restrict__ rt, const T *__restrict__ r, BatchSys *st, RedSeg R,
           int mixed)
{
    const Seg sg = seg[blockIdx.x];
    BatchSys *S = st + sg.sys;
    if (S->in.done)
        return;
    double a = 0;
    SEG_LOOP(i) a += (double)(rt[i] * r[i]);
    double v[1] = {a}, tot[1];
    if (seg_finish<RS_BVBS, 1>(R, sg.sys, blockIdx.x, sys_g0[sg.sys], sys_g0[sg.sys + 1], v,
                               tot) &&
        threadIdx.x == 0) {
        DevState *in = &S->in;
        in->rho1 = tot[0];
        if (in->rho1 == 0 || in->omega == 0) {
            in->status = RSGPU_BREAKDOWN;
            end_inner(S, mixed, R.ndone);
            return;
        }
        if (!in->first)
            in->beta = (in->rho1 / in->rho2) * (in->alpha / in->omega);
    }
}

// p = r (first) or temp = (-omega) v + p; p = r + beta*temp
template <typename T>
__global__ void __launch_bounds__(RS_BVBS)
    kb_p(const Seg *__restrict__ seg, const T *__restrict__ r, const T *__restrict__ v,
         T *__restrict__ p, const BatchSys *st)
{
    const Seg sg = seg[blockIdx.x];
    const DevState *in = &st[sg.sys].in;
    if (in->done)
        return;
    if (in->first) {
        SEG_LOOP(i) p[i] = r[i];
        return;
    }
    const T mo = (T)(-in->omega), beta = (T)in->beta;
    SEG_LOOP(i)
    {
        T temp = (mo == T(0)) ? T(0) : mo * v[i];
        temp += p[i];
        T pi = r[i];
        if (beta != T(0))
            pi += beta * temp;
        p[i] = pi;
    }
}

// s = r + (-alpha) v; monitor(||s||)
template <typename T>
__global__ void __launch_bounds__(RS_BVBS)
    kb_s(const Seg *__restrict__ seg, const int *__restrict__ sys_g0, const T *__restrict__ r,
         const T *__restrict__ v, T *__restrict__ s, BatchSys *st, RedSeg R, int mixed)
{
    const Seg sg = seg[blockIdx.x];
    BatchSys *S = st + sg.sys;
    if (S->in.done)
        return;
    const T ma = (T)(-S->in.alpha);
    double a = 0;
    SEG_LOOP(i)
    {
        T si = r[i];
        if (ma != T(0))
            si += ma * v[i];
        s[i] = si;
        a += (double)(si * si);
    }
    double vv[1] = {a}, tot[1];
    if (seg_finish<RS_BVBS, 1>(R, sg.sys, blockIdx.x, sys_g0[sg.sys], sys_g0[sg.sys + 1], vv,
                               tot) &&
        threadIdx.x == 0) {
        if (dev_monitor(&S->in, sqrt(tot[0]))) {
            if (S->in.conv)
                S->in.half = 1;   // return x.add(alpha, phat)
            else
                end_inner(S, mixed, R.ndone);
        }
    }
}

// early return after s converged: x += alpha * phat
template <typename T>
__global__ void __launch_bounds__(RS_BVBS)
    kb_half(const Seg *__restrict__ seg, const int *__restrict__ sys_g0,
            const T *__restrict__ phat, T *__restrict__ x, BatchSys *st, RedSeg R, int mixed)
{
    const Seg sg = seg[blockIdx.x];
    BatchSys *S = st + sg.sys;
    if (S->in.done || !S->in.half)
        return;
    const T al = (T)S->in.alpha;
    SEG_LOOP(i)
    if (al != T(0)) x[i] += al * phat[i];
    double v[1] = {0.0}, tot[1];
    if (seg_finish<RS_BVBS, 1>(R, sg.sys, blockIdx.x, sys_g0[sg.sys], sys_g0[sg.sys + 1], v,
                               tot) &&
        threadIdx.x == 0)
        end_inner(S, mixed, R.ndone);
}

// x += alpha phat; x += omega shat; r = s + (-omega) t; monitor(||r||)
template <typename T>
__global__ void __launch_bounds__(RS_BVBS)
    kb_update(const Seg *__restrict__ seg, const int *__restrict__ sys_g0,
              const T *__restrict__ phat, const T *__restrict__ shat, const T *__restrict__ s,
              const T *__restrict__ t, T *__restrict__ x, T *__restrict__ r, BatchSys *st,
              RedSeg R, int mixed)
{
    const Seg sg = seg[blockIdx.x];
    BatchSys *S = st + sg.sys;
    if (S->in.done)
        return;
    const T al = (T)S->in.alpha, om = (T)S->in.omega, mo = (T)(-S->in.omega);
    double a = 0;
    SEG_LOOP(i)
    {
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
    if (seg_finish<RS_BVBS, 1>(R, sg.sys, blockIdx.x, sys_g0[sg.sys], sys_g0[sg.sys + 1], v,
                               tot) &&
        threadIdx.x == 0) {
        DevState *in = &S->in;
        in->rho2 = in->rho1;
        in->first = 0;
        in->iter++;
        if (dev_monitor(in, sqrt(tot[0])))
            end_inner(S, mixed, R.ndone);
    }
}

// mixed outer step, part 1: x += z (IR.java:66, x.add(z))
__global__ void __launch_bounds__(RS_BVBS)
    kb_outer_x(const Seg *__restrict__ seg, const float *__restrict__ z, double *__restrict__ x,
               const BatchSys *st)
{
    const Seg sg = seg[blockIdx.x];
    if (!outer_due(&st[sg.sys]))
        return;
    SEG_LOOP(i) x[i] += (double)z[i];
}

template <typename T>
__global__ void kb_fill(T *__restrict__ a, size_t n, T v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}

// ===================================================== host side
struct rsgpu_batch {
    rsgpu_ctx *ctx = nullptr;
    rsgpu_solver_opts o{};
    int flags = 0;
    int nsys = 0, n = 0;
    size_t ldv = 0;
    std::vector<int> h_sys;
    // vector segments
    int nvg = 0;
    DevBuf<Seg> vseg;
    DevBuf<int> sys_vg0;
    // SpMV plans of the current pattern (FP64 and FP32 stages)
    uint64_t plan_pattern = 0;
    BPlan p64, p32;
    float tune_us64 = 0, tune_us32 = 0;
    float sweep_us[2] = {0, 0};      // ILU apply: level schedules, line wavefront
    // state
    DevBuf<BatchSys> st;
    std::vector<BatchSys> h_st;
    DevBuf<double> partials;
    DevBuf<unsigned> tickets, ndone;
    DevBuf<double> norms;
    DevBuf<double> w64;
    DevBuf<float> w32;
    DevBuf<int> blk_sys;
    hipEvent_t ev[2] = {nullptr, nullptr};
    unsigned *flags_host = nullptr;
    int last_launches = 0;
};

namespace {

void build_bplan(const std::vector<int> &rp, const std::vector<int> &ci,
                 const std::vector<int> &sys, int nnz_stage, BPlan &pl, hipStream_t s)
{
    const int nsys = (int)sys.size() - 1;
    std::vector<int> starts, wsys, g0(nsys + 1, 0);
    starts.reserve(rp.size() / 16 + nsys + 2);
    for (int k = 0; k < nsys; ++k) {
        g0[k] = (int)starts.size();
        int r = sys[k];
        const int e_sys = sys[k + 1];
        while (r < e_sys) {
            starts.push_back(r);
            wsys.push_back(k);
            const int j0 = rp[r];
            int e = r + 1;
            if (rp[r + 1] - j0 <= nnz_stage)
                while (e < e_sys && e - r < RS_SPMV_BS && rp[e + 1] - j0 <= nnz_stage)
                    ++e;
            r = e;
        }
    }
    g0[nsys] = (int)starts.size();
    starts.push_back(sys[nsys]);
    pl.nnz_stage = nnz_stage;
    pl.nwg = (int)wsys.size();
    pl.wg_rows.upload(starts, s);
    pl.wg_sys.upload(wsys, s);
    pl.sys_g0.upload(g0, s);
    spmv_build_coff(rp, ci, starts, pl.coff, s);
    RS_HIP(hipStreamSynchronize(s));
}

RedSeg red_of(rsgpu_batch *B)
{
    return RedSeg{B->partials.p, (int)(B->partials.n / 2), B->tickets.p, B->ndone.p};
}

template <typename T>
const T *vals(rsgpu_csr *A);
template <>
const double *vals<double>(rsgpu_csr *A) { return A->val.p; }
template <>
const float *vals<float>(rsgpu_csr *A) { return A->val32.p; }

// launch one batched SpMV kernel with the plan's stage size
template <typename T, class Epi, int NR, int FIN>
void launch_bspmv(rsgpu_batch *B, rsgpu_csr *A, const BPlan &pl, const T *x, Epi epi)
{
    hipStream_t s = B->ctx->stream;
    const dim3 g(pl.nwg), b(RS_SPMV_BS);
    const RedSeg R = red_of(B);
    const int mixed = B->o.mixed;
    switch (pl.nnz_stage) {
    case 1024:
        hipLaunchKernelGGL((kb_spmv<T, 1024, Epi, NR, FIN>), g, b, 0, s, pl.dev(), A->rowptr.p,
                           A->colind.p, vals<T>(A), x, epi, B->st.p, R, mixed, B->norms.p);
        break;
    case 2048:
        hipLaunchKernelGGL((kb_spmv<T, 2048, Epi, NR, FIN>), g, b, 0, s, pl.dev(), A->rowptr.p,
                           A->colind.p, vals<T>(A), x, epi, B->st.p, R, mixed, B->norms.p);
        break;
    case 4096:
        hipLaunchKernelGGL((kb_spmv<T, 4096, Epi, NR, FIN>), g, b, 0, s, pl.dev(), A->rowptr.p,
                           A->colind.p, vals<T>(A), x, epi, B->st.p, R, mixed, B->norms.p);
        break;
    default:
        if (sizeof(T) == 8)
            throw RsError{RSGPU_BAD_ARGUMENT, "unsupported FP64 SpMV stage"};
        hipLaunchKernelGGL((kb_spmv<T, sizeof(T) == 4 ? 8192 : 4096, Epi, NR, FIN>), g, b, 0,
                           s, pl.dev(), A->rowptr.p, A->colind.p, vals<T>(A), x, epi, B->st.p,
                           R, mixed, B->norms.p);
        break;
    }
    RS_CHECK_LAUNCH();
}

// Candidate stages: FP64 1024/2048/4096 nonzeros (8-32 KiB of products),
// FP32 2048/4096/8192.  Each is timed as the solver uses it, y = A x fused
// with two per-system dot products (median of 5 after a warm-up; no solver
// state is touched), and the fastest plan is kept.
template <typename T>
float autotune(rsgpu_batch *B, rsgpu_csr *A, BPlan &best)
{
    hipStream_t s = B->ctx->stream;
    const int c64[3] = {1024, 2048, 4096}, c32[3] = {2048, 4096, 8192};
    const int *cand = sizeof(T) == 8 ? c64 : c32;
    const bool tune = !(B->flags & RSGPU_BATCH_NO_AUTOTUNE);
    T *x = (T *)(sizeof(T) == 8 ? (void *)B->w64.p : (void *)B->w32.p);
    T *y = x + B->ldv;
    float best_us = 0;
    int best_stage = sizeof(T) == 8 ? RS_SPMV_NNZ : RS_SPMV_NNZ32;
    if (tune) {
        BPlan pl[3];
        size_t need = 1;
        for (int c = 0; c < 3; ++c) {
            build_bplan(A->h_rowptr, A->h_colind, B->h_sys, cand[c], pl[c], s);
            need = std::max(need, (size_t)pl[c].nwg);
        }
        if (B->partials.n < 2 * need)
            B->partials.alloc(2 * need);
        hipEvent_t e0, e1;
        RS_HIP(hipEventCreate(&e0));
        RS_HIP(hipEventCreate(&e1));
        best_us = 1e30f;
        for (int c = 0; c < 3; ++c) {
            std::vector<float> t;
            for (int rep = 0; rep < 6; ++rep) {
                RS_HIP(hipEventRecord(e0, s));
                launch_bspmv<T, BEpiDot<T>, 2, -1>(B, A, pl[c], x, BEpiDot<T>{y, x, 2});
                RS_HIP(hipEventRecord(e1, s));
                RS_HIP(hipEventSynchronize(e1));
                float ms = 0;
                RS_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep)
                    t.push_back(ms * 1000.f);
            }
            std::sort(t.begin(), t.end());
            if (t[t.size() / 2] < best_us) {
                best_us = t[t.size() / 2];
                best_stage = cand[c];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    build_bplan(A->h_rowptr, A->h_colind, B->h_sys, best_stage, best, s);
    RS_HIP(hipStreamSynchronize(s));
    return best_us;
}

void prepare_plans(rsgpu_batch *B, rsgpu_csr *A)
{
    if (B->plan_pattern == A->pattern_id)
        return;
    B->tune_us64 = autotune<double>(B, A, B->p64);
    if (B->o.mixed)
        B->tune_us32 = autotune<float>(B, A, B->p32);
    const size_t need = std::max({(size_t)B->nvg, (size_t)B->p64.nwg, (size_t)B->p32.nwg, (size_t)1});
    if (B->partials.n < 2 * need)
        B->partials.alloc(2 * need);
    B->plan_pattern = A->pattern_id;
}

void batch_papply(rsgpu_prec *M, const double *b, double *x)
{
    prec_apply(M, b, x, nullptr);
}
void batch_papply(rsgpu_prec *M, const float *b, float *x) { prec_apply(M, b, x, nullptr); }

// ILU sweep kernel: when the preconditioner has line-wavefront records, its
// apply is timed with the wavefront and with the level schedules (median of
// 5 after a warm-up, on the batch's own blocks) and the faster is kept for
// this pattern (rsgpu_prec_sweep_form then reports it).  The two are
// bit-identical, so the choice changes time only.
template <typename T>
void autotune_sweeps(rsgpu_batch *B, rsgpu_prec *M)
{
    if (!M->band || M->band_pick >= 0 || (B->flags & RSGPU_BATCH_NO_AUTOTUNE) ||
        getenv("RSGPU_ILU_BAND_APPLY"))
        return;
    hipStream_t s = B->ctx->stream;
    T *x = (T *)(sizeof(T) == 8 ? (void *)B->w64.p : (void *)B->w32.p);
    T *y = x + B->ldv;
    hipEvent_t e0, e1;
    RS_HIP(hipEventCreate(&e0));
    RS_HIP(hipEventCreate(&e1));
    float us[2];
    for (int pick = 0; pick < 2; ++pick) {
        M->band_pick = pick;
        std::vector<float> t;
        for (int rep = 0; rep < 6; ++rep) {
            RS_HIP(hipEventRecord(e0, s));
            batch_papply(M, x, y);
            RS_HIP(hipEventRecord(e1, s));
            RS_HIP(hipEventSynchronize(e1));
            float ms = 0;
            RS_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep)
                t.push_back(ms * 1000.f);
        }
        std::sort(t.begin(), t.end());
        us[pick] = t[t.size() / 2];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    M->band_pick = us[1] < us[0] ? 1 : 0;
    B->sweep_us[0] = us[0];
    B->sweep_us[1] = us[1];
}

// One BiCGstab iteration of every system that is still running (T = double:
// the plain FP64 solve on x; T = float: the inner solve on z), preceded in
// the mixed form by the outer step of every system whose inner solve ended.
template <typename T>
void enqueue_iteration(rsgpu_batch *B, rsgpu_csr *A, rsgpu_prec *M, const double *b64,
                       double *x64, T *work, T *x)
{
    hipStream_t s = B->ctx->stream;
    const size_t L = B->ldv;
    T *p = work, *sv = work + L, *phat = work + 2 * L, *shat = work + 3 * L, *t = work + 4 * L,
      *v = work + 5 * L, *r = work + 6 * L, *rt = work + 7 * L;
    const RedSeg R = red_of(B);
    const int mixed = B->o.mixed;
    const dim3 G(B->nvg), BS(RS_BVBS);
    const Seg *sg = B->vseg.p;
    const int *g0 = B->sys_vg0.p;
    const BPlan &pl = sizeof(T) == 8 ? B->p64 : B->p32;
    if (mixed) {
        float *z = (float *)x;
        hipLaunchKernelGGL(kb_outer_x, G, BS, 0, s, sg, z, x64, B->st.p);
        launch_bspmv<double, BEpiOuter, 2, BF_OUTER>(
            B, A, B->p64, x64, BEpiOuter{b64, (float *)r, (float *)rt, z});
    }
    hipLaunchKernelGGL(kb_rho<T>, G, BS, 0, s, sg, g0, rt, r, B->st.p, R, mixed);
    hipLaunchKernelGGL(kb_p<T>, G, BS, 0, s, sg, r, v, p, B->st.p);
    batch_papply(M, p, phat);
    launch_bspmv<T, BEpiDot<T>, 1, BF_ALPHA>(B, A, pl, phat, BEpiDot<T>{v, rt, 1});
    hipLaunchKernelGGL(kb_s<T>, G, BS, 0, s, sg, g0, r, v, sv, B->st.p, R, mixed);
    hipLaunchKernelGGL(kb_half<T>, G, BS, 0, s, sg, g0, phat, x, B->st.p, R, mixed);
    batch_papply(M, sv, shat);
    launch_bspmv<T, BEpiDot<T>, 2, BF_OMEGA>(B, A, pl, shat, BEpiDot<T>{t, sv, 2});
    hipLaunchKernelGGL(kb_update<T>, G, BS, 0, s, sg, g0, phat, shat, sv, t, x, r, B->st.p, R,
                       mixed);
    RS_CHECK_LAUNCH();
}

// Per-block owner system for the sweeps' skip test; every system must start
// at a block boundary.
void prepare_skip(rsgpu_batch *B, rsgpu_prec *M)
{
    if (M->kind != RSGPU_PREC_BLOCK_ILU0) {
        M->skip = BlkSkip{};
        return;
    }
    const std::vector<int> &bs = M->h_bstart;
    std::vector<int> owner(M->nblocks);
    for (int k = 0; k < B->nsys; ++k)
        if (!std::binary_search(bs.begin(), bs.end(), B->h_sys[k]))
            throw RsError{RSGPU_BAD_ARGUMENT,
                          "preconditioner blocks must not cross system boundaries"};
    int k = 0;
    for (int b = 0; b < M->nblocks; ++b) {
        while (bs[b] >= B->h_sys[k + 1])
            ++k;
        owner[b] = k;
    }
    B->blk_sys.upload(owner, B->ctx->stream);
    M->skip = BlkSkip{B->blk_sys.p, &B->st.p[0].in.done, (int)(sizeof(BatchSys) / sizeof(int))};
}

template <typename T>
void run_batch(rsgpu_batch *B, rsgpu_csr *A, rsgpu_prec *M, const double *b, double *x)
{
    hipStream_t s = B->ctx->stream;
    T *work, *xi;
    if (B->o.mixed) {
        work = (T *)B->w32.p + B->ldv;     // z, then the BiCGstab vectors
        xi = (T *)B->w32.p;
    } else {
        work = (T *)B->w64.p;
        xi = (T *)x;
    }
    if (!B->o.mixed) {
        // r = b - A x, rtilde = r, monitor at iteration 0 (BiCGstab.java:69-72)
        const size_t L = B->ldv;
        launch_bspmv<double, BEpiInit, 1, BF_INIT>(
            B, A, B->p64, x, BEpiInit{(double *)work + 6 * L, (double *)work + 7 * L, b});
    } else {
        hipLaunchKernelGGL(kb_fill<float>, dim3((B->n + 255) / 256), dim3(256), 0, s,
                           (float *)xi, (size_t)B->n, 0.0f);
    }
    const int per = std::max(1, B->o.check_every);
    const bool graph = !(B->flags & RSGPU_BATCH_NO_GRAPH);
    hipGraphExec_t exec = nullptr;
    if (graph) {
        hipGraph_t g = nullptr;
        RS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        try {
            for (int k = 0; k < per; ++k)
                enqueue_iteration<T>(B, A, M, b, x, work, xi);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        RS_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RS_HIP(e);
    }
    // every system stops within its monitors' iteration limits; the cap only
    // guards against a state machine that would never stop
    const long long inner = B->o.mixed ? (long long)B->o.inner_max_iter + 2 : 1;
    const long long cap = ((long long)B->o.max_iter + 2) * inner / per + 8;
    int k = 0;
    unsigned done = 0;
    try {
        for (long long it = 0;; ++it) {
            if (it > cap)
                throw RsError{RSGPU_HIP_ERROR, "batched solve did not stop"};
            if (graph)
                RS_HIP(hipGraphLaunch(exec, s));
            else
                for (int q = 0; q < per; ++q)
                    enqueue_iteration<T>(B, A, M, b, x, work, xi);
            RS_HIP(hipMemcpyAsync(B->flags_host + (k & 1), B->ndone.p, sizeof(unsigned),
                                  hipMemcpyDeviceToHost, s));
            RS_HIP(hipEventRecord(B->ev[k & 1], s));
            ++k;
            if (k >= 2) {
                RS_HIP(hipEventSynchronize(B->ev[k & 1]));
                done = B->flags_host[k & 1];
                if ((int)done >= B->nsys)
                    break;
            }
        }
        RS_HIP(hipEventSynchronize(B->ev[(k - 1) & 1]));
    } catch (...) {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        throw;
    }
    if (exec)
        RS_HIP(hipGraphExecDestroy(exec));
    B->last_launches = k;
}

} // namespace

// ===================================================== C ABI
static int batch_fail(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

extern "C" int rsgpu_batch_create(rsgpu_ctx *ctx, const rsgpu_solver_opts *opts, int nsys,
                                  const int32_t *sys_start_host, int flags, rsgpu_batch **out)
{
    if (!ctx || !opts || !sys_start_host || !out || nsys < 1)
        return RSGPU_BAD_ARGUMENT;
    if (opts->kind != RSGPU_BICGSTAB) {
        ctx->err = "batched solves run BiCGstab";
        return RSGPU_BAD_ARGUMENT;
    }
    if (opts->prec_transpose) {
        ctx->err = "batched solves are not transposed";
        return RSGPU_BAD_ARGUMENT;
    }
    if (sys_start_host[0] != 0)
        return RSGPU_BAD_ARGUMENT;
    for (int k = 0; k < nsys; ++k)
        if (sys_start_host[k + 1] <= sys_start_host[k]) {
            ctx->err = "system boundaries must rise strictly";
            return RSGPU_BAD_ARGUMENT;
        }
    rsgpu_batch *B = new rsgpu_batch();
    try {
        RS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        B->ctx = ctx;
        B->o = *opts;
        B->flags = flags;
        B->nsys = nsys;
        B->h_sys.assign(sys_start_host, sys_start_host + nsys + 1);
        B->n = B->h_sys[nsys];
        B->ldv = ((size_t)B->n + 3) & ~(size_t)3;
        std::vector<Seg> segs;
        std::vector<int> g0(nsys + 1);
        for (int k = 0; k < nsys; ++k) {
            g0[k] = (int)segs.size();
            for (int r = B->h_sys[k]; r < B->h_sys[k + 1]; r += RS_BCHUNK)
                segs.push_back(Seg{k, r, std::min(r + RS_BCHUNK, B->h_sys[k + 1]), 0});
        }
        g0[nsys] = (int)segs.size();
        B->nvg = (int)segs.size();
        B->vseg.upload(segs, s);
        B->sys_vg0.upload(g0, s);
        B->st.alloc(nsys);
        B->h_st.resize(nsys);
        B->tickets.alloc((size_t)nsys * RS_BTICKETS * RS_TPAD);
        B->tickets.zero(s);
        B->ndone.alloc(1);
        B->norms.alloc(2 * (size_t)nsys);
        B->partials.alloc(2 * (size_t)B->nvg);
        // the work vectors start as zero vectors, as the reference's
        // template-built DenseVectors do
        if (opts->mixed) {
            B->w64.alloc(2 * B->ldv);
            B->w32.alloc(std::max<size_t>(9 * B->ldv, 2 * B->ldv));
            B->w32.zero(s);
        } else {
            B->w64.alloc(8 * B->ldv);
            B->w32.alloc(2 * B->ldv);
        }
        B->w64.zero(s);
        RS_HIP(hipEventCreateWithFlags(&B->ev[0], hipEventDisableTiming));
        RS_HIP(hipEventCreateWithFlags(&B->ev[1], hipEventDisableTiming));
        RS_HIP(hipHostMalloc((void **)&B->flags_host, sizeof(unsigned) * 4,
                             hipHostMallocDefault));
        RS_HIP(hipStreamSynchronize(s));
    } catch (const RsError &e) {
        rsgpu_batch_destroy(B);
        return batch_fail(ctx, e);
    }
    *out = B;
    return RSGPU_OK;
}

extern "C" int rsgpu_batch_destroy(rsgpu_batch *B)
{
    if (!B)
        return RSGPU_OK;
    if (B->ctx)
        (void)hipStreamSynchronize(B->ctx->stream);
    for (auto &e : B->ev)
        if (e)
            (void)hipEventDestroy(e);
    if (B->flags_host)
        (void)hipHostFree(B->flags_host);
    delete B;
    return RSGPU_OK;
}

extern "C" int rsgpu_batch_solve(rsgpu_batch *B, rsgpu_csr *A, rsgpu_prec *M,
                                 const double *b_dev, double *x_dev, rsgpu_stats *stats_host)
{
    if (!B || !A || !b_dev || !x_dev)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_ctx *ctx = B->ctx;
    if (A->n != A->ncols || A->n != B->n) {
        ctx->err = "the matrix must be square with the batch's total size";
        return RSGPU_BAD_ARGUMENT;
    }
    rsgpu_prec ident;
    ident.ctx = ctx;
    ident.kind = RSGPU_PREC_IDENTITY;
    ident.n = A->n;
    ident.ready = true;
    if (!M)
        M = &ident;
    if (M->kind != RSGPU_PREC_IDENTITY && (!M->ready || M->pattern_id != A->pattern_id)) {
        ctx->err = "Preconditioner.setMatrix must be called before use";
        return RSGPU_BAD_ARGUMENT;
    }
    if (M->kind != RSGPU_PREC_IDENTITY && M->kind != RSGPU_PREC_ILU0 &&
        M->kind != RSGPU_PREC_BLOCK_ILU0 && M->kind != RSGPU_PREC_DIAGONAL) {
        ctx->err = "batched solves take an ILU(0), block ILU(0), diagonal or no preconditioner";
        return RSGPU_BAD_ARGUMENT;
    }
    const int want = B->o.mixed ? RSGPU_FP32 : RSGPU_FP64;
    if (M->kind != RSGPU_PREC_IDENTITY && M->precision != want) {
        ctx->err = B->o.mixed ? "mixed batched solve needs an FP32 preconditioner"
                              : "FP64 batched solve needs an FP64 preconditioner";
        return RSGPU_BAD_ARGUMENT;
    }
    if (B->o.mixed && !A->f32_current()) {
        ctx->err = "mixed solve needs rsgpu_csr_sync_f32 after the last value change";
        return RSGPU_BAD_ARGUMENT;
    }
    try {
        RS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        prepare_plans(B, A);
        if (M->kind == RSGPU_PREC_ILU0 || M->kind == RSGPU_PREC_BLOCK_ILU0) {
            if (M->precision == RSGPU_FP32)
                autotune_sweeps<float>(B, M);
            else
                autotune_sweeps<double>(B, M);
        }
        prepare_skip(B, M);
        // per-system states (DefaultIterationMonitor of each solve)
        for (int k = 0; k < B->nsys; ++k) {
            BatchSys &h = B->h_st[k];
            std::memset(&h, 0, sizeof(h));
            DevState *in = &h.in, *o = &h.out;
            if (B->o.mixed) {
                in->rtol = B->o.inner_rtol;
                in->atol = 1e-50;
                in->dtol = 1e5;
                in->max_iter = B->o.inner_max_iter;
                in->done = 1;   // the first outer step starts the inner solve
            } else {
                in->rtol = B->o.rtol;
                in->atol = B->o.atol;
                in->dtol = B->o.dtol;
                in->max_iter = B->o.max_iter;
            }
            in->first = 1;
            in->rho1 = in->rho2 = in->alpha = in->beta = in->omega = 1;
            o->rtol = B->o.rtol;
            o->atol = B->o.atol;
            o->dtol = B->o.dtol;
            o->max_iter = B->o.max_iter;
            o->first = 1;
        }
        RS_HIP(hipMemcpyAsync(B->st.p, B->h_st.data(), sizeof(BatchSys) * B->nsys,
                              hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(kb_fill<unsigned>, dim3(1), dim3(64), 0, s, B->ndone.p, (size_t)1,
                           0u);
        if (B->o.mixed)
            run_batch<float>(B, A, M, b_dev, x_dev);
        else
            run_batch<double>(B, A, M, b_dev, x_dev);
        M->skip = BlkSkip{};
        // true residual ||b - A x|| and ||b|| of every system
        launch_bspmv<double, BEpiFinal, 2, BF_FINAL>(B, A, B->p64, x_dev, BEpiFinal{b_dev});
        std::vector<double> nrm(2 * (size_t)B->nsys);
        RS_HIP(hipMemcpyAsync(B->h_st.data(), B->st.p, sizeof(BatchSys) * B->nsys,
                              hipMemcpyDeviceToHost, s));
        RS_HIP(hipMemcpyAsync(nrm.data(), B->norms.p, sizeof(double) * nrm.size(),
                              hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        int worst = RSGPU_OK;
        for (int k = 0; k < B->nsys; ++k) {
            const BatchSys &h = B->h_st[k];
            const DevState &mon = B->o.mixed ? h.out : h.in;
            if (stats_host) {
                rsgpu_stats &st = stats_host[k];
                std::memset(&st, 0, sizeof(st));
                st.iterations = mon.iter;
                st.residual = mon.residual;
                st.status = mon.status;
                st.hist_len = mon.hist_len;
                st.inner_iterations = B->o.mixed ? h.inner_total : 0;
                st.outer_iterations = B->o.mixed ? h.out.iter : 0;
                st.true_residual = nrm[2 * k];
                st.b_norm = nrm[2 * k + 1];
            }
            if (worst == RSGPU_OK && mon.status != RSGPU_OK)
                worst = mon.status;
        }
        return worst;
    } catch (const RsError &e) {
        M->skip = BlkSkip{};
        return batch_fail(ctx, e);
    }
}

extern "C" int rsgpu_batch_info(const rsgpu_batch *B, int *spmv_stage_fp64,
                                int *spmv_stage_fp32, double *spmv_us_fp64,
                                double *spmv_us_fp32, int *graph_launches)
{
    if (!B)
        return RSGPU_BAD_ARGUMENT;
    if (spmv_stage_fp64)
        *spmv_stage_fp64 = B->p64.nnz_stage;
    if (spmv_stage_fp32)
        *spmv_stage_fp32 = B->p32.nnz_stage;
    if (spmv_us_fp64)
        *spmv_us_fp64 = B->tune_us64;
    if (spmv_us_fp32)
        *spmv_us_fp32 = B->tune_us32;
    if (graph_launches)
        *graph_launches = B->last_launches;
    return RSGPU_OK;
}
