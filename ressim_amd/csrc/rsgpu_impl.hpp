// Internal definitions shared by the rsgpu HIP translation units.
// Target: gfx950 (CDNA4), wave64.  All arithmetic that must agree bit for
// bit with the reference's sequential Java loops is compiled with
// -ffp-contract=off and keeps the reference's per-row summation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rsgpu.h"

#define RS_WAVE 64

// ---------------------------------------------------------------- errors
struct RsError {
    int code;
    std::string msg;
};

#define RS_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            throw RsError{e_ == hipErrorOutOfMemory ? RSGPU_OUT_OF_MEMORY     \
                                                    : RSGPU_HIP_ERROR,        \
                          std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define RS_CHECK_LAUNCH() RS_HIP(hipGetLastError())

// ---------------------------------------------------------------- device buffer
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        if (count == n && p)
            return;
        release();
        if (count) {
            RS_HIP(hipMalloc((void **)&p, sizeof(T) * count));
            n = count;
        }
    }
    void upload(const T *h, size_t count, hipStream_t s)
    {
        alloc(count);
        if (count)
            RS_HIP(hipMemcpyAsync(p, h, sizeof(T) * count,
                                  hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &v, hipStream_t s)
    {
        upload(v.data(), v.size(), s);
    }
    void zero(hipStream_t s)
    {
        if (n)
            RS_HIP(hipMemsetAsync(p, 0, sizeof(T) * n, s));
    }
};

// ---------------------------------------------------------------- handles
struct rsgpu_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    // pinned host staging for small device->host reads
    void *pinned = nullptr;
};

struct SpmvPlan {
    // contiguous row ranges per workgroup; long rows get their own group
    DevBuf<int> wg_rows;
    int nwg = 0;
    // column - (first row of the entry's workgroup) as 16 bits, when every
    // entry fits (stencil matrices): the SpMV then streams 2 index bytes per
    // nonzero instead of 4; empty otherwise
    DevBuf<short> coff;
    const short *coff_ptr() const { return coff.n ? coff.p : nullptr; }
};

// Sliced copy of a CSR pattern for stencil-like matrices: rows in chunks of
// RS_SL_ROWS (one workgroup of four waves), each chunk with one table, the
// ascending union of its rows' column offsets (column - row, shared by
// chunks with equal unions), and per row a bit mask of the table slots the
// row holds.  Values are stored slot-major per chunk (value of chunk c,
// slot s, row c * RS_SL_ROWS + i at (cbase[c] + s) * RS_SL_ROWS + i, zero
// where the row lacks the slot), so a wave reads one slot of its 64 rows as
// one coalesced run and gathers x at row + offset, contiguous across the
// wave.  Built when every chunk's union has at most 32 offsets; absent
// otherwise (nchunks == 0).
//
// Block form (b3, 3x3 block matrices): the same chunks of scalar rows, but a
// table lists block offsets (block column minus block row, ascending), a
// row's slots are (entry k, block column e) at 3k + e, and mask16 holds per
// scalar row the bits of the entries its block row has.  Row 3I+a reads
// slot 3k+e against x[3 (I + d_k) + e], which is its expanded CSR order.
// Built when every chunk's union has at most RS_SL3_MAXB block offsets.
#define RS_SL_ROWS 256
#define RS_SL3_MAXB 10
struct SlicedFmt {
    int nchunks = 0, ntables = 0;
    bool b3 = false;
    long long nslots = 0;              // values per copy (64 per slot row)
    DevBuf<int> ctab, cbase, tptr, toff;
    DevBuf<unsigned> mask;
    DevBuf<unsigned short> mask16;     // block form
    DevBuf<double> v64;                // filled by sliced_fill (per solve)
    DevBuf<float> v32;                 // filled with the FP32 CSR copy
    uint64_t gen64 = 0;                // value generation v64 was filled from
    bool ok() const { return nchunks > 0; }
};

// 256-row tiles (one workgroup each) of the sliced kernels
inline int sliced_tiles(int n) { return (n + 255) / 256; }

// what the sliced kernels read (plain pointers)
struct SlView {
    int n, ncols, nchunks;
    const int *ctab, *cbase, *tptr, *toff;
    const unsigned *mask;
    const unsigned short *mask16;      // block form only
};

struct rsgpu_csr {
    rsgpu_ctx *ctx = nullptr;
    int n = 0, ncols = 0, nnz = 0;
    uint64_t pattern_id = 0;
    std::vector<int> h_rowptr, h_colind;
    DevBuf<int> rowptr, colind;
    DevBuf<double> val;
    DevBuf<float> val32;
    // Value generation: every entry point that changes val counts it up;
    // the FP32 copy and the sliced FP64 copy record the generation they
    // were filled from, and their users refuse a stale copy.  Writes
    // through the rsgpu_csr_values_dev pointer are not seen.
    uint64_t val_gen = 1;
    uint64_t gen32 = 0;
    void values_changed() { ++val_gen; }
    bool f32_current() const { return val32.p && gen32 == val_gen; }
    SpmvPlan plan;     // FP64 kernels: 2048-nonzero stages
    SpmvPlan plan32;   // FP32 kernels: 4096-nonzero stages (same LDS bytes)
    // transposed copy (lazy): rows = columns of A, values gathered via tperm
    bool has_t = false;
    DevBuf<int> trowptr, tcolind, tperm;
    SpmvPlan tplan;
    DevBuf<double> tval;
    // pattern id of the matrix this one is the transpose of (rsgpu_csr_transpose)
    uint64_t transpose_of = 0;
    // 3x3 block rows (rsgpu_csr_create_bsr3): the CSR above is the expanded
    // matrix with each block row's scalar rows interleaved; brp/bcol is the
    // block pattern and bplan/bplan32 group whole block rows
    int bs = 1;
    int nb = 0, nnzb = 0;
    DevBuf<int> brp, bcol;
    SpmvPlan bplan, bplan32;
    SlicedFmt sl;
    SlView slview() const
    {
        return SlView{n, ncols, sl.nchunks, sl.ctab.p, sl.cbase.p, sl.tptr.p, sl.toff.p,
                      sl.mask.p, sl.mask16.p};
    }
};

// Per-connection transmissibility lists of a mesh (flux.hip): rows =
// connections with their MD entries in the reference's order, plus each
// cell's incident connections
struct rsgpu_flux {
    rsgpu_ctx *ctx = nullptr;
    int nconn = 0, ncells = 0;
    DevBuf<int> md_ptr, md_elem;
    DevBuf<double> md_t;
    SpmvPlan plan;
    DevBuf<int> cell_ptr, cell_conn;   // connections of each cell, ascending,
    DevBuf<signed char> cell_sign;     // +1 here, -1 there
    DevBuf<int> here, there;
    int nentries = 0;                  // MD entries (md_ptr[nconn])
};

// Level-scheduled sweep data of one triangular factor (see ilu.hip)
struct SweepData {
    int W = 4;               // padded row width (entries per lane)
    int nlev_total = 0;      // sum over blocks
    int max_levels = 0;      // per block
    DevBuf<int> blk_lev;     // [nblocks+1] level ranges per block
    DevBuf<int> lev_meta;    // per level: count, ell offset, row offset, pad
    DevBuf<int> rows;        // local row per (level,t)
    DevBuf<int> row_glob;    // global row per (level,t) (diagonal gather)
    DevBuf<int> ell_col;     // local column per slot (nb = zero slot)
    DevBuf<int> colb, rowsb; // the same as LDS byte offsets (times element size)
    DevBuf<int> ell_src;     // source position in LU, -1 padding
    size_t ell_size = 0;
    int nrows_lev = 0;
    int max_count = 0, max_width = 0;
    // R = 3: a level lists cells (3x3 block rows) and one lane handles the
    // cell's three rows in order.  `rows` holds three entries per cell (the
    // factorisation); per cell, `colb` holds `irec` ints (the cell's LDS byte
    // offset, then those of its W neighbour blocks in this triangle, the zero
    // slots as padding) and `ell_src` `vrec` positions of the cell record
    // (the three rows' neighbour-block entries, the three intra-cell
    // coefficients, for U the three diagonals)
    int R = 1;
    int irec = 0, vrec = 0;
};

struct AmgData;
void amg_destroy(AmgData *d);
struct IccData;
void icc_destroy(IccData *d);
struct IlutData;
void ilut_destroy(IlutData *d);

// Block b of a block-Jacobi sweep is skipped while
// done[(blk_sys ? blk_sys[b] : b) * stride] != 0 (the done flag of the
// system that owns the block, in an array of per-system states).
struct BlkSkip {
    const int *blk_sys = nullptr;
    const int *done = nullptr;
    int stride = 0;
};

__device__ __forceinline__ bool blk_skipped(const BlkSkip &k, int b)
{
    return k.done && k.done[(size_t)(k.blk_sys ? k.blk_sys[b] : b) * k.stride];
}

struct rsgpu_prec {
    rsgpu_ctx *ctx = nullptr;
    int kind = 0, nblocks = 1, precision = RSGPU_FP64;
    int n = 0;
    uint64_t pattern_id = 0;
    bool ready = false;
    // block structure (user_bstart: caller-given block boundaries, else
    // nblocks equal contiguous ranges)
    std::vector<int> user_bstart;
    std::vector<int> h_bstart;
    DevBuf<int> bstart;
    int max_block_rows = 0;
    // factor pattern (A's pattern or its block-diagonal part)
    int nnz = 0;
    std::vector<int> h_rowptr, h_colind;
    DevBuf<int> rowptr, colind, diag, src;   // src: position in A (block filter)
    bool src_identity = true;
    DevBuf<int> fac_rows;       // rows in factor level order, per block
    DevBuf<int> fac_lev;        // level row ranges
    DevBuf<int> fac_blk_lev;    // per block level ranges
    int fac_levels_max = 0;
    DevBuf<double> lu64;
    DevBuf<float> lu32;
    DevBuf<int> status;          // [2]: code, bad row
    SweepData L, U, Lt, Ut;      // forward/backward and transposed sweeps
    bool has_t = false;
    DevBuf<double> ell64_L, ell64_U, ell64_Lt, ell64_Ut, dgU64, dgUt64;
    DevBuf<float> ell32_L, ell32_U, dgU32;
    // cell sweeps (3x3 block matrices): L and U are cell schedules, their
    // ELL arrays the cell records; cell_factor: the block-slot factorisation
    // with its per-cell records
    bool cell = false, cell_factor = false;
    DevBuf<int> cell_rec;
    DevBuf<double> scratch64;    // global fallback vectors for large blocks
    DevBuf<float> scratch32;
    DevBuf<double> tmp64;        // FP32 apply on FP64 vectors
    DevBuf<float> tmpa32, tmpb32;
    int lds_rows = 0;            // rows that fit the LDS vector
    // AMG hierarchy (amg.hip); SSOR reuses it with a single smoothed level
    rsgpu_amg_params amg_params{};
    AmgData *amg = nullptr;
    // DiagonalPreconditioner (ilu.hip)
    DevBuf<double> invdiag;
    DevBuf<float> invdiag32;
    // per-block skip for the ILU sweeps (batched solves); set only while a
    // batched solve enqueues its applies
    BlkSkip skip{};
    IccData *icc = nullptr;     // ICC: fill-in reach lists and factor levels
    // ILUT: drop tolerance and extra entries per side (ILUT.java:91-118
    // defaults), the device factorisation's buffers
    double ilut_tau = 1e-6;
    int ilut_p = 25;
    IlutData *ilut = nullptr;
    // line-wavefront sweeps (every block a grid of Ly <= 64 lines of Lx rows
    // coupled only among 9-point neighbours): per block (Lx, Ly), per row the
    // factor positions of the 4 L slots, the diagonal and the 4 U slots (-1
    // absent), the slot masks, and the per-row records gathered after each
    // factorisation
    bool band = false;
    // which apply kernel a batched solve's timing picked for this pattern
    // (-1: none, the block-count rule; 0 level schedules; 1 line wavefront)
    int band_pick = -1;
    std::vector<int2> h_band_dims;   // host copy of band_dims
    DevBuf<int2> band_dims;
    DevBuf<int> band_pos;
    DevBuf<unsigned char> band_mask;
    DevBuf<double> band64;       // per row: L[4], U[4] (8 values), diagonals after
    DevBuf<float> band32;
    // FP32: per row 1 / U_ii in FP64 (padded like the records), for the full
    // blocks' U sweep
    DevBuf<double> band_rd;
    // elements of the largest full block's per-line slot layout (Ly lines of
    // round_up(Lx + 2 Ly - 2, 16) slots)
    int band_slots = 0;
    // the line-wavefront factorisation (every block full): its 9-value
    // input and output records per row, padded like the sweep records
    bool band_factor = false;
    DevBuf<double> bf_in64, bf_out64;
    DevBuf<float> bf_in32, bf_out32;
    ~rsgpu_prec()
    {
        amg_destroy(amg);
        icc_destroy(icc);
        ilut_destroy(ilut);
    }
};

// ---------------------------------------------------------------- device state
enum {
    ST_FLAG_DONE = 1,
};

// Scalars of an iterative solve kept on the device so that iterations can
// be enqueued without host round trips.  Layout shared by all solvers.
struct DevState {
    // monitor (DefaultIterationMonitor)
    int iter;
    int status;
    int done;
    int conv;
    int hist_len;
    int hist_cap;
    int max_iter;
    int half;           // BiCGstab: s converged, x += alpha*phat pending
    int first;
    int cycle_end;      // GMRES: inner loop finished
    int gi;             // GMRES: inner index i
    int pad0;
    double initR, rtol, atol, dtol, residual;
    double *hist;
    // BiCGstab
    double rho1, rho2, alpha, beta, omega;
    double ts, tt;
    // GMRES(m): Hessenberg (m+1) x m column-major, rotated rhs s, Givens
    double normr;
    double hsub;        // unrotated H(i+1,i) for the basis scaling
    double *H, *s, *gc, *gs;
    double *h2;         // CGS2: the second pass's projections of column gi
    int m;
    int pad1;
    // FP64 GMRES with delayed reorthogonalisation (DCGS2): the Hessenberg
    // before rotations ((m+1) x m, column-major), the reorthogonalisation
    // projections of the newest basis vector (and its squared norm last),
    // and the step's scalars: rho (norm of the reorthogonalised vector) and
    // c (its projection coefficient)
    double *Hr, *ds;
    double drho, dc;
    // QMR (QMR.java:136-243); c1, c2: the p / q scale factors, val: the d / s
    // scale factor of the coming update
    double q_rho, q_rho1, q_xi, q_gamma, q_gamma1, q_theta, q_theta1, q_eta, q_delta, q_ep,
        q_beta, q_c1, q_c2, q_val;
};

// ---------------------------------------------------------------- reductions
// Deterministic two-level reduction: fixed per-thread strides, wave64 shuffle
// trees, per-workgroup partials, and a finishing workgroup chosen by an
// arrival ticket that sums the partials in a fixed order and resets the
// ticket for the next launch.
struct RedSlot {
    double *partials;   // [max_wg * 2]
    unsigned *ticket;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Sums of a[0..31] over the wave's 64 lanes: lane l returns column l >> 1.
// Each exchange round halves the columns a lane keeps (it sends the half
// its partner keeps), so the 32 sums take 32 shuffles instead of 32 trees
// of 6; the summation order is fixed.
__device__ __forceinline__ double wave_sum_cols32(const double (&a)[32])
{
    const int lane = threadIdx.x & 63;
    double b[16];
    {
        const bool hi = lane & 32;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const double keep = hi ? a[q + 16] : a[q], send = hi ? a[q] : a[q + 16];
            b[q] = keep + __shfl_xor(send, 32, 64);
        }
    }
#pragma unroll
    for (int w = 8, o = 16; w >= 1; w >>= 1, o >>= 1) {
        const bool hi = lane & o;
#pragma unroll
        for (int q = 0; q < w; ++q) {
            const double keep = hi ? b[q + w] : b[q], send = hi ? b[q] : b[q + w];
            b[q] = keep + __shfl_xor(send, o, 64);
        }
    }
    return b[0] + __shfl_xor(b[0], 1, 64);
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (BS > 64) {
        if (lane == 0)
            sh[w] = v;
        __syncthreads();
        v = 0;
        if (threadIdx.x < 64) {
            v = threadIdx.x < BS / 64 ? sh[threadIdx.x] : 0.0;
            v = wave_sum(v);
        }
        __syncthreads();
    }
    return v;   // valid in thread 0
}

// Sum of p[i] for i = i0 + threadIdx.x, i0 + threadIdx.x + BS, ... < i1, in
// that order, read with agent-scope atomic loads (the partials of other
// workgroups).  Eight loads are issued before the first add: a load-add
// loop waits for each coherent load in turn, a memory round trip per
// iteration.  Indices past i1 re-read the last valid partial and are not
// added.
template <int BS>
__device__ __forceinline__ double gather_sum(const double *p, int i0, int i1)
{
    double acc = 0;
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 8 * BS) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            t[u] = __hip_atomic_load(&p[min(i + u * BS, i1 - 1)], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        acc += t[0];
#pragma unroll
        for (int u = 1; u < 8; ++u) {
            const double a = acc + t[u];
            acc = i + u * BS < i1 ? a : acc;
        }
    }
    return acc;
}

// Publish this group's partial(s); returns true in every thread of the
// finishing group, whose thread 0 then holds the totals in tot[].
//
// The partials are written with agent-scope atomic stores, which the L2
// writes through to the device coherence point, and the store is waited for
// (vmcnt(0)) before an arrival ticket is taken.  No group therefore needs a
// device-wide release fence (on a multi-XCD device that writes back the
// whole L2).  Arrivals are counted in two levels -- RS_NGROUP counters on
// separate 128-byte lines, then one global counter taken by the last
// arrival of each group -- so no single address serialises thousands of
// atomics.  Only the finishing group performs an acquire (L2 invalidate)
// before it reads the partials with agent-scope atomic loads.  Every
// counter is reset by the arrival that completes it, ready for the next
// launch.
#define RS_NGROUP 64
#define RS_TPAD 32
#define RS_TICKET_WORDS ((RS_NGROUP + 1) * RS_TPAD)

// Waits until this wave's vector-memory stores have completed.  The compiler
// signal fences on both sides keep the compiler from moving memory accesses
// (the relaxed partial stores before, the ticket atomics after) across the
// wait, which is not a memory operation in its model; they emit nothing.
__device__ __forceinline__ void wait_vm_stores()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int BS, int NR>
__device__ bool block_reduce_finish(RedSlot slot, const double (&v)[NR],
                                    double (&tot)[NR])
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
            __hip_atomic_store(&slot.partials[(size_t)q * gridDim.x + blockIdx.x], r[q],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wait_vm_stores();
        const int G = (int)gridDim.x;
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
                __hip_atomic_store(slot.ticket, 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
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
        tot[q] = block_sum<BS>(
            gather_sum<BS>(slot.partials + (size_t)q * gridDim.x, 0, (int)gridDim.x), sh);
    return true;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  This bijection gives XCD x a contiguous run of logical groups, so
// neighbouring row ranges (which gather overlapping parts of x) share one
// L2 instead of all eight fetching the same lines.
#define RS_NUM_XCD 8
__device__ __forceinline__ int xcd_remap(int bid, int nwg)
{
    const int x = bid % RS_NUM_XCD, k = bid / RS_NUM_XCD;
    const int per = nwg / RS_NUM_XCD, rem = nwg % RS_NUM_XCD;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// java.lang.Math.max (NaN wins)
__device__ __forceinline__ double java_max(double a, double b)
{
    if (a != a)
        return a;
    if (b != b)
        return b;
    return a > b ? a : b;
}

// AbstractIterationMonitor.converged + DefaultIterationMonitor.convergedI.
// Returns true when the solve must stop (converged or a failure status).
__device__ __forceinline__ bool dev_monitor(DevState *st, double r)
{
    if (st->hist && st->hist_len < st->hist_cap)
        st->hist[st->hist_len] = r;
    st->hist_len++;
    st->residual = r;
    if (st->iter == 0)
        st->initR = r;
    if (r < java_max(st->rtol * st->initR, st->atol)) {
        st->conv = 1;
        return true;
    }
    if (r > st->dtol * st->initR) {
        st->status = RSGPU_DIVERGENCE;
        return true;
    }
    if (st->iter >= st->max_iter) {
        st->status = RSGPU_ITERATIONS;
        return true;
    }
    if (r != r) {
        st->status = RSGPU_DIVERGENCE;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- GMRES columns
// GMRES.solve's per-column work on the least-squares system (GMRES.java:
// 155-163, GivensRotation.java:42-93) by one wave of the finishing group
// (all 64 lanes active): column i's entries 0..i come from src (H's own
// column when src is null) and its subdiagonal is hn.  The entries, the
// earlier rotations and s(i), s(i+1) are fetched with one load per lane into
// LDS (lh: 3 x 64 doubles), lane 0 applies the rotations in order, and the
// rotated entries are stored back one per lane: the sequential loop's
// operations and order, bit for bit.  i + 1 < 64.
__device__ inline void gm_finish_column_wave(DevState *st, int i, double hn, double *lh,
                                             const double *src = nullptr)
{
    const int lane = threadIdx.x & 63;
    const int ld = st->m + 1;
    double *Hc = st->H + i * ld;
    const double *Sc = src ? src : Hc;
    double *gcp = st->gc, *gsp = st->gs, *sv = st->s;
    double *lc = lh + 64, *ls = lh + 128;
    if (lane <= i)
        lh[lane] = Sc[lane];
    if (lane < i) {
        lc[lane] = gcp[lane];
        ls[lane] = gsp[lane];
    }
    const double s0 = sv[i], s1 = sv[i + 1];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        st->hsub = hn;
        // previous rotations on column i; cur is the running Hc[q]
        double cur = lh[0];
        for (int q = 0; q < i; ++q) {
            const double c = lc[q], s = ls[q], nx = lh[q + 1];
            const double temp = c * cur + s * nx;
            cur = -s * cur + c * nx;
            lh[q] = temp;
        }
        // new rotation (GivensRotation(x, y))
        const double x = cur, y = hn;
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
        gcp[i] = c;
        gsp[i] = s;
        Hc[i] = c * x + s * y;
        Hc[i + 1] = -s * x + c * y;
        const double t0 = c * s0 + s * s1;
        const double t1 = -s * s0 + c * s1;
        sv[i] = t0;
        sv[i + 1] = t1;
        st->iter++;
        st->gi = i + 1;
        if (!(i + 1 < st->m) || dev_monitor(st, fabs(t1))) {
            st->cycle_end = 1;
            if (st->status)
                st->done = 1;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < i)
        Hc[lane] = lh[lane];
}

// Delayed reorthogonalisation, the step that finishes column c (one wave of
// the finishing group).  The basis vector v~ = u / eta that followed column
// c (eta = st->hsub, its first projection's norm) has projections
// ds[0..c] = V' v~ and squared norm ds[c+1] (ds in LDS or global memory
// written before this wave's group started); its reorthogonalised form is
// v = (v~ - V ds) / rho with rho = sqrt(ds[c+1] - sum ds^2), so the Arnoldi
// relation of column c becomes Hr(0..c, c) += eta ds, Hr(c+1, c) = eta rho.
// Column c is then rotated and monitored (gm_finish_column_wave).  Returns
// rho in every lane.
__device__ inline double dcgs_finish_column_wave(DevState *st, int c, double *lh,
                                                 const double *ds)
{
    const int lane = threadIdx.x & 63;
    const int ld = st->m + 1;
    double *Hr = st->Hr;
    const double sk = lane <= c ? ds[lane] : 0.0;
    const double alpha = ds[c + 1];
    const double ss = wave_sum(sk * sk);
    const double r2 = alpha - ss;
    const double rho = r2 > 0 ? sqrt(r2) : 0.0;
    const double eta = st->hsub;
    if (lane <= c)
        Hr[lane + c * ld] += eta * sk;
    if (lane == 0)
        Hr[c + 1 + c * ld] = eta * rho;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    gm_finish_column_wave(st, c, eta * rho, lh, Hr + c * ld);
    return rho;
}

// ---------------------------------------------------------------- host API
// spmv.hip
void spmv_build_plan(const std::vector<int> &rowptr, int n, SpmvPlan &plan,
                     hipStream_t s, int nnz_wg = 2048,
                     const std::vector<int> *colind = nullptr);
// 16-bit column offsets of a plan (left empty when an entry does not fit)
void spmv_build_coff(const std::vector<int> &rowptr, const std::vector<int> &colind,
                     const std::vector<int> &starts, DevBuf<short> &coff, hipStream_t s);
// both plans of a CSR matrix (FP64 and FP32 stage sizes)
void spmv_build_plans(const std::vector<int> &rowptr, int n, rsgpu_csr *A, hipStream_t s,
                      bool sliced = true);
// sliced values from A->val (T = double: A->sl.v64, float: A->sl.v32)
template <typename T>
void sliced_fill(rsgpu_csr *A, hipStream_t s);
// y = A x over the sliced copy of precision T (mult's zero rule)
template <typename T>
void sliced_mult(hipStream_t s, const rsgpu_csr *A, const T *x, T *y);
// block form of the sliced copy of a 3x3 block matrix (block pattern)
void sliced_build3(rsgpu_csr *A, const std::vector<int> &brp, const std::vector<int> &bcol,
                   hipStream_t s);
// block-row plans of a 3x3 block matrix (scalar row bounds at block rows)
void spmv_build_plan_b3(const std::vector<int> &brp, int nb, SpmvPlan &plan, hipStream_t s,
                        int nnz_wg);
template <typename T>
void spmv_launch(hipStream_t s, const rsgpu_csr *A, const T *val, const T *x,
                 T *y, int mode, double alpha, const T *b);
void csr_build_transpose(rsgpu_csr *A);
void csr_transmult_launch(rsgpu_csr *A, double alpha, const double *x,
                          double *y, bool add);
void csr_launch_gather(hipStream_t s, const int *perm, const double *src,
                       double *dst, int n);
void csr_residual_dd(rsgpu_csr *A, const double *b, const double *x, double *r);
void csr_zero_rows_dev(rsgpu_csr *A, double diagonal, const int *rows_dev,
                       int nrows);
void convert_f64_to_f32(hipStream_t s, const double *a, float *b, size_t n);
void convert_f32_to_f64(hipStream_t s, const float *a, double *b, size_t n);

// ilu.hip
void prec_setup(rsgpu_prec *M, rsgpu_csr *A);
void prec_setup_pattern(rsgpu_prec *M, const std::vector<int> &rp, const std::vector<int> &ci,
                        const std::vector<int> &dg, int bs);
void prec_refresh_fp64(rsgpu_prec *M);
int ilut_set_matrix(rsgpu_prec *M, rsgpu_csr *A, int *bad_row);
void amg_apply32(rsgpu_prec *M, const float *b, float *x, const DevState *st);
bool ilut_apply(rsgpu_prec *M, const double *b, double *x, bool trans, const int *sk1,
                const int *sk2);
int prec_factor(rsgpu_prec *M, rsgpu_csr *A, int *bad_row);
int icc_set_matrix(rsgpu_prec *M, rsgpu_csr *A, int *bad_row);
void prec_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st);
void prec_apply(rsgpu_prec *M, const float *b, float *x, const DevState *st);
void prec_trans_apply(rsgpu_prec *M, const double *b, double *x,
                      const DevState *st = nullptr);
// GMRES inner step fused into one launch (sliced SpMV with the basis scale,
// line-wavefront block ILU(0) apply, first projection pass onto V): whether
// it applies to (A, M) with elements of elem_bytes and restart m, and its
// launch (wout must not alias win; the projections go to column i of the
// state's H, or to hout when that is given)
bool gmres_fused_ok(const rsgpu_csr *A, const rsgpu_prec *M, int elem_bytes, int m);
// whether ILU.apply takes the line wavefront when M has its records (few
// blocks per CU; many take the level schedules)
bool prec_band_apply(const rsgpu_prec *M);
// whether ILU.apply on 3x3 cells streams its records through LDS rings
bool prec_cell_stream(const rsgpu_prec *M);
template <typename T>
void launch_gmres_fused(hipStream_t s, const rsgpu_csr *A, const T *slv, rsgpu_prec *M,
                        const T *win, T *vout, bool scale, T *wout, const T *V, size_t ldv,
                        int i, DevState *st, RedSlot slot, double *hout = nullptr,
                        bool dcgs = false);

// amg.hip
int amg_setup(rsgpu_prec *M, rsgpu_csr *A, int *bad_row);
void amg_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st, bool trans);
int amg_levels(const rsgpu_prec *M, int *sizes, int *nnzs, int cap);
int ssor_setup(rsgpu_prec *M, rsgpu_csr *A, int *bad_row);
void ssor_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st);
int diag_setup(rsgpu_prec *M, rsgpu_csr *A, int *bad_row);
int amg_get_level(const rsgpu_prec *M, int k, int32_t *rowptr, int32_t *colind,
                  double *data, int32_t *pt, int32_t *Prp, int32_t *Pci, double *Pv,
                  int *pnnz);

uint64_t next_pattern_id();
// flux.hip: per-device address of rsgpu_flux_rates_noalloc's static scratch
// array, taken with `device` current (rsgpu_ctx_create)
int rs_rates_scratch_init(int device, std::string &err);
