// CSR SpMV core shared by the plain SpMV entry points and the fused Krylov
// kernels.
//
// Layout: each workgroup owns a contiguous row range whose nonzeros fit the
// LDS stage (plan built on the host, spmv.hip).  Phase 1 streams values,
// column indices and the gathered x coalesced (one nonzero per lane, NNZ_WG
// per group) and stores the products val[j]*x[col[j]] in LDS.  Phase 2 gives
// one row per lane and sums that row's products left to right, which is the
// statement order of CompRowMatrix.mult/multAdd
// (mtj/src/no/uib/cipr/matrix/sparse/CompRowMatrix.java:301-353); with
// -ffp-contract=off the per-row result is bit-identical to the sequential
// loop.  A row longer than the stage is summed chunk by chunk by one lane,
// again in order.
#pragma once

#include "rsgpu_impl.hpp"

#define RS_SPMV_BS 256
#define RS_SPMV_NNZ 2048
#define RS_SPMV_NNZ32 4096

// Stage size and plan by value type: FP32 groups stage twice the nonzeros
// (the same 16 KiB of products), halving the fixed per-group work.
template <typename T>
constexpr int spmv_nnz()
{
    return sizeof(T) == 4 ? RS_SPMV_NNZ32 : RS_SPMV_NNZ;
}

template <typename T>
inline const SpmvPlan &spmv_plan(const rsgpu_csr *A)
{
    if (A->bs == 3)
        return sizeof(T) == 4 ? A->bplan32 : A->bplan;
    return sizeof(T) == 4 ? A->plan32 : A->plan;
}

// coff (may be null): per nonzero, its column minus the first row of its
// workgroup, in 16 bits (SpmvPlan::coff); used instead of colind when given.
// g: the plan group this workgroup computes.
template <typename T, int BS, int NNZ_WG, class Epi>
__device__ __forceinline__ void spmv_core_off_g(int g, const int *__restrict__ wg_rows,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ colind,
                                                const short *__restrict__ coff,
                                                const T *__restrict__ val,
                                                const T *__restrict__ x, Epi &epi)
{
    __shared__ T prod[NNZ_WG];
    const int r0 = wg_rows[g], r1 = wg_rows[g + 1];
    const int j0 = rowptr[r0], j1 = rowptr[r1];
    if (j1 - j0 <= NNZ_WG) {
        // Every lane issues all of its value/column loads, then all of its
        // gathers, before the first product is stored: indices past the
        // group's last nonzero are clamped to it (a valid, already-fetched
        // address) and only the store is predicated.  A guarded load per
        // slot would serialise the slots (one memory latency each).  The
        // value and index streams are read nontemporally: the batched
        // ensembles stream gigabytes through this core, and keeping them
        // out of the last-level cache keeps the vectors in it (ordinary
        // loads made the FP32 batched SpMV 15% slower).
        constexpr int Q = NNZ_WG / BS;
        if (j1 > j0) {
            T v[Q];
            int c[Q];
            if (coff) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int j = min(j0 + (int)threadIdx.x + q * BS, j1 - 1);
                    v[q] = __builtin_nontemporal_load(&val[j]);
                    c[q] = r0 + (int)__builtin_nontemporal_load(&coff[j]);
                }
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int j = min(j0 + (int)threadIdx.x + q * BS, j1 - 1);
                    v[q] = __builtin_nontemporal_load(&val[j]);
                    c[q] = __builtin_nontemporal_load(&colind[j]);
                }
            }
            T xg[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q)
                xg[q] = x[c[q]];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int jl = (int)threadIdx.x + q * BS;
                if (j0 + jl < j1)
                    prod[jl] = v[q] * xg[q];
            }
        }
        __syncthreads();
        for (int r = r0 + (int)threadIdx.x; r < r1; r += BS) {
            const int a = rowptr[r] - j0, e = rowptr[r + 1] - j0;
            T dot = epi.init(r);
            for (int j = a; j < e; ++j)
                dot += prod[j];
            epi.row(r, dot);
        }
    } else {
        // a single row (the plan isolates rows longer than the stage)
        T dot = epi.init(r0);
        for (int c = j0; c < j1; c += NNZ_WG) {
            const int ce = c + NNZ_WG < j1 ? c + NNZ_WG : j1;
            for (int j = c + (int)threadIdx.x; j < ce; j += BS)
                prod[j - c] = val[j] * x[colind[j]];
            __syncthreads();
            if (threadIdx.x == 0)
                for (int j = 0; j < ce - c; ++j)
                    dot += prod[j];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            epi.row(r0, dot);
    }
}

// one group per workgroup, XCD-aware
template <typename T, int BS, int NNZ_WG, class Epi>
__device__ __forceinline__ void spmv_core_off(const int *__restrict__ wg_rows,
                                              const int *__restrict__ rowptr,
                                              const int *__restrict__ colind,
                                              const short *__restrict__ coff,
                                              const T *__restrict__ val,
                                              const T *__restrict__ x, Epi &epi)
{
    spmv_core_off_g<T, BS, NNZ_WG>(xcd_remap((int)blockIdx.x, (int)gridDim.x), wg_rows,
                                   rowptr, colind, coff, val, x, epi);
}

template <typename T, int BS, int NNZ_WG, class Epi>
__device__ __forceinline__ void spmv_core(const int *__restrict__ wg_rows,
                                          const int *__restrict__ rowptr,
                                          const int *__restrict__ colind,
                                          const T *__restrict__ val,
                                          const T *__restrict__ x, Epi &epi)
{
    spmv_core_off<T, BS, NNZ_WG>(wg_rows, rowptr, colind, (const short *)nullptr, val, x,
                                 epi);
}

// 3x3 block rows (row-interleaved layout: block row I is scalar rows 3I..3I+2,
// each listing its blocks in ascending block column, three scalars per
// block, so the value array is the CSR of the expanded matrix and scalar row
// 3I+a starts at 9 brp[I] + 3a (brp[I+1] - brp[I])).  The plan's groups are
// whole block rows.  Phase 1a streams the group's values into LDS
// coalesced, one per lane, as the CSR core does but with no column-index
// stream.  Phase 1b gives one block per lane: one coalesced block column
// index and one 3-element gather of x per block (instead of nine column
// indices and nine gathers), and the block's nine staged values are
// multiplied in place.  Phase 2 is the CSR phase 2: each scalar row summed
// left to right, bit-identical to CompRowMatrix.mult on the expanded matrix.
// The block-column and x loads are issued before the LDS barriers, which
// wait for LDS traffic only, so they overlap the staging.
__device__ __forceinline__ void lds_only_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T, int BS, int NNZ_WG, class Epi>
__device__ __forceinline__ void spmv_core_b3_g(int g, const int *__restrict__ wg_rows,
                                               const int *__restrict__ rowptr,
                                               const int *__restrict__ colind,
                                               const int *__restrict__ brp,
                                               const int *__restrict__ bcol,
                                               const T *__restrict__ val,
                                               const T *__restrict__ x, Epi &epi)
{
    constexpr int MAXB = NNZ_WG / 9;          // blocks per stage
    constexpr int QB = (MAXB + BS - 1) / BS;  // blocks per lane
    constexpr int Q = NNZ_WG / BS;            // staged values per lane
    static_assert(BS / 3 <= 255, "local block rows are numbered in one byte");
    __shared__ int s_brp[BS / 3 + 2];
    __shared__ unsigned char s_brow[MAXB + 1];
    const int r0 = wg_rows[g], r1 = wg_rows[g + 1];
    if ((r0 % 3) || (r1 % 3) || r1 - r0 > BS) {
        // one scalar row of a block row longer than the stage (the plan
        // splits such block rows into their three scalar rows): CSR core
        spmv_core_off_g<T, BS, NNZ_WG>(g, wg_rows, rowptr, colind, (const short *)nullptr, val,
                                       x, epi);
        return;
    }
    const int I0 = r0 / 3, nI = (r1 - r0) / 3;
    const int B0 = brp[I0], nbk = brp[I0 + nI] - B0, nv = 9 * nbk;
    if (nbk > MAXB) {
        spmv_core_off_g<T, BS, NNZ_WG>(g, wg_rows, rowptr, colind, (const short *)nullptr, val,
                                       x, epi);
        return;
    }
    __shared__ T prod[NNZ_WG];
    const int j0 = 9 * B0;   // rowptr[r0]
    if (nbk > 0) {
        T v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q)
            v[q] = __builtin_nontemporal_load(&val[j0 + min((int)threadIdx.x + q * BS, nv - 1)]);
        T xg[QB][3];
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int t = min((int)threadIdx.x + q * BS, nbk - 1);
            const int c = 3 * __builtin_nontemporal_load(&bcol[B0 + t]);
#pragma unroll
            for (int e = 0; e < 3; ++e)
                xg[q][e] = x[c + e];
        }
        for (int t = threadIdx.x; t <= nI; t += BS)
            s_brp[t] = brp[I0 + t];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int jl = (int)threadIdx.x + q * BS;
            if (jl < nv)
                prod[jl] = v[q];
        }
        lds_only_barrier();
        // local block row of every staged block
        for (int r = threadIdx.x; r < nI; r += BS)
            for (int b = s_brp[r]; b < s_brp[r + 1]; ++b)
                s_brow[b - B0] = (unsigned char)r;
        lds_only_barrier();
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int t = (int)threadIdx.x + q * BS;
            if (t < nbk) {
                const int r = s_brow[t];
                const int b0r = s_brp[r], nbr = s_brp[r + 1] - b0r;
                const int pos0 = 9 * (b0r - B0) + 3 * (B0 + t - b0r);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int p = pos0 + a * 3 * nbr + e;
                        prod[p] = prod[p] * xg[q][e];
                    }
            }
        }
    } else {
        for (int t = threadIdx.x; t <= nI; t += BS)
            s_brp[t] = brp[I0 + t];
    }
    __syncthreads();
    for (int r = r0 + (int)threadIdx.x; r < r1; r += BS) {
        const int rr = r - r0, rl = rr / 3, a = rr - 3 * rl;
        const int b0r = s_brp[rl], len = 3 * (s_brp[rl + 1] - b0r);
        const int s0 = 9 * (b0r - B0) + a * len;
        T dot = epi.init(r);
        for (int j = s0; j < s0 + len; ++j)
            dot += prod[j];
        epi.row(r, dot);
    }
}

template <typename T, int BS, int NNZ_WG, class Epi>
__device__ __forceinline__ void spmv_core_b3(const int *__restrict__ wg_rows,
                                             const int *__restrict__ rowptr,
                                             const int *__restrict__ colind,
                                             const int *__restrict__ brp,
                                             const int *__restrict__ bcol,
                                             const T *__restrict__ val,
                                             const T *__restrict__ x, Epi &epi)
{
    spmv_core_b3_g<T, BS, NNZ_WG>(xcd_remap((int)blockIdx.x, (int)gridDim.x), wg_rows, rowptr,
                                  colind, brp, bcol, val, x, epi);
}

// Sliced core (SlicedFmt): chunks of RS_SL_ROWS rows, one wave per 64 of
// them, one row per lane.  A row's present slots are visited in table
// order, which is its CSR order (ascending columns), and each product is
// rounded to T before it is added, as in the CSR core, so every row sum is
// bit-identical to the sequential loop.  Absent slots are skipped rather
// than added as zeros, so non-finite x entries outside a row's pattern never
// reach it.  The chunk's table is read up front as 32 offsets (wide scalar
// loads; the offset array is padded by 32 entries), slots past the table
// reuse its last offset and value slot, and gather columns clamp into
// [0, ncols), so every address is valid and all loads of a batch of 16
// slots are in flight before the first product.  A workgroup's four waves
// read adjacent 64-row pieces of each slot row, so one slot of the group is
// one contiguous run.  Values are read with ordinary loads: the
// nontemporal hint keeps them out of the last-level cache, which the FP32
// inner solve's working set largely fits.
#define RS_SL_BATCH 16
template <typename T, class Epi, bool SCALE = false, int B = RS_SL_BATCH>
__device__ __forceinline__ void sliced_batch(int s0, int W, int r, int cmax, unsigned m,
                                             const int (&offs)[32], const T *vc,
                                             const T *__restrict__ x, T &dot, T sc = T(1))
{
    T v[B], xg[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int sl = min(s0 + k, W - 1);
        const int col = min(max(r + offs[s0 + k], 0), cmax);
        v[k] = vc[(size_t)sl * RS_SL_ROWS];
        xg[k] = x[col];
    }
    // SCALE: x is sc * x[col] (zero when sc is zero), formed exactly as a
    // separate scaling pass would store it
    if (SCALE) {
#pragma unroll
        for (int k = 0; k < B; ++k)
            xg[k] = (sc == T(0)) ? T(0) : sc * xg[k];
    }
    // every product and sum is formed and then selected: a conditional add
    // lets the compiler sink the slot's loads under the condition, where
    // each waits for memory on its own
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const T q = dot + v[k] * xg[k];
        dot = ((m >> (s0 + k)) & 1u) ? q : dot;
    }
}

// the 64 rows from row0 (wave-uniform, row0 < n); B: slots per batch of
// loads in flight (32: every slot of a row in one round trip)
template <typename T, class Epi, bool SCALE = false, int B = RS_SL_BATCH>
__device__ __forceinline__ void sliced_rows(int row0, const SlView &S, const T *__restrict__ val,
                                            const T *__restrict__ x, Epi &epi, T sc = T(1))
{
    const int lane = threadIdx.x & 63;
    const int c = row0 / RS_SL_ROWS, sub = row0 - c * RS_SL_ROWS;
    const int r = row0 + lane;
    const int t = S.ctab[c];
    const int o0 = S.tptr[t], W = S.tptr[t + 1] - o0;
    const T *vc = val + (size_t)S.cbase[c] * RS_SL_ROWS + sub + lane;
    const int *off = S.toff + o0;
    // lanes past the last row compute a copy of the last row and store
    // nothing (no condition on them before the store, which would again
    // let the compiler sink loads under it)
    const bool live = r < S.n;
    const int rr = min(r, S.n - 1);
    const unsigned m = S.mask[rr];
    const int cmax = S.ncols - 1;
    T dot = epi.init(rr);
    // an opaque start value: selecting between a product and a constant
    // start is otherwise turned into a branch around the first slot's loads
    asm volatile("" : "+v"(dot));
    if (W > 0) {
        const int lastoff = off[W - 1];
        int offs[32];
#pragma unroll
        for (int k = 0; k < 32; ++k)
            offs[k] = off[k];
#pragma unroll
        for (int k = 0; k < 32; ++k)
            offs[k] = k < W ? offs[k] : lastoff;
        sliced_batch<T, Epi, SCALE, B>(0, W, rr, cmax, m, offs, vc, x, dot, sc);
        if (B < 32 && W > B)
            sliced_batch<T, Epi, SCALE, B>(B, W, rr, cmax, m, offs, vc, x, dot, sc);
    }
    if (live)
        epi.row(r, dot);
}

// Block form (SlicedFmt::b3): row r = 3I + a of the 64 rows from row0.
// NB table entries are read (NB >= the chunk's entry count W; entries past
// W repeat the last one and their slots repeat the last slot), and the 3 NB
// value loads and gathers are all issued before the first product.  Slots
// are summed in the order 3k + e, each product rounded to T before it is
// added, and absent entries are skipped, as in the scalar form.
template <int NB, typename T, class Epi, bool SCALE>
__device__ __forceinline__ void sliced_rows3_nb(int r, int I, unsigned m, int W, const int *off,
                                                const T *vc, const T *__restrict__ x, int cmax,
                                                T &dot, T sc)
{
    const int lastoff = off[W - 1];
    int offs[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
        offs[k] = off[k];
#pragma unroll
    for (int k = 0; k < NB; ++k)
        offs[k] = k < W ? offs[k] : lastoff;
    const int ns = 3 * W;
    T v[3 * NB], xg[3 * NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int sl = min(3 * k + e, ns - 1);
            const int col = min(max(3 * (I + offs[k]) + e, 0), cmax);
            v[3 * k + e] = vc[(size_t)sl * RS_SL_ROWS];
            xg[3 * k + e] = x[col];
        }
    if (SCALE) {
#pragma unroll
        for (int k = 0; k < 3 * NB; ++k)
            xg[k] = (sc == T(0)) ? T(0) : sc * xg[k];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const T q = dot + v[3 * k + e] * xg[3 * k + e];
            dot = ((m >> k) & 1u) ? q : dot;
        }
}

template <typename T, class Epi, bool SCALE = false>
__device__ __forceinline__ void sliced_rows3(int row0, const SlView &S, const T *__restrict__ val,
                                             const T *__restrict__ x, Epi &epi, T sc = T(1))
{
    const int lane = threadIdx.x & 63;
    const int c = row0 / RS_SL_ROWS, sub = row0 - c * RS_SL_ROWS;
    const int r = row0 + lane;
    const int t = S.ctab[c];
    const int o0 = S.tptr[t], W = S.tptr[t + 1] - o0;
    const T *vc = val + (size_t)S.cbase[c] * RS_SL_ROWS + sub + lane;
    const int *off = S.toff + o0;
    const bool live = r < S.n;
    const int rr = min(r, S.n - 1);
    const int I = rr / 3;
    const unsigned m = S.mask16[rr];
    const int cmax = S.ncols - 1;
    T dot = epi.init(rr);
    asm volatile("" : "+v"(dot));
    // seven entries: the 3-D seven-point block stencil in one round trip
    if (W > 0 && W <= 7)
        sliced_rows3_nb<7, T, Epi, SCALE>(rr, I, m, W, off, vc, x, cmax, dot, sc);
    else if (W > 7)
        sliced_rows3_nb<RS_SL3_MAXB, T, Epi, SCALE>(rr, I, m, W, off, vc, x, cmax, dot, sc);
    if (live)
        epi.row(r, dot);
}

// the scalar or the block form
template <bool B3, typename T, class Epi, bool SCALE = false>
__device__ __forceinline__ void sliced_rows_any(int row0, const SlView &S,
                                                const T *__restrict__ val,
                                                const T *__restrict__ x, Epi &epi, T sc = T(1))
{
    if constexpr (B3)
        sliced_rows3<T, Epi, SCALE>(row0, S, val, x, epi, sc);
    else
        sliced_rows<T, Epi, SCALE>(row0, S, val, x, epi, sc);
}

// first row of the calling wave in 256-row tile `tile` (wave-uniform)
__device__ __forceinline__ int sliced_wave_row(int tile)
{
    return tile * 256 + 64 * __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// y = A x with CompRowMatrix.mult's rule: y[i] = dot only if dot != 0.
template <typename T>
struct EpiMult {
    T *y;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot) { y[r] = (dot != T(0)) ? dot : T(0); }
};

// y += alpha * (A x)   (multAdd)
template <typename T>
struct EpiMultAdd {
    T *y;
    T alpha;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot) { y[r] += alpha * dot; }
};

// y = b + alpha*(A x)  (multAdd on y.set(b): the residual form)
template <typename T>
struct EpiResid {
    T *y;
    const T *b;
    T alpha;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot) { y[r] = b[r] + alpha * dot; }
};

// y = sum (no zero rule; used for the transposed gather)
template <typename T>
struct EpiPlain {
    T *y;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot) { y[r] = dot; }
};

// transMultAdd: y = alpha * (y * (1/alpha) + A'x), the accumulator starts
// at y*(1/alpha) (CompRowMatrix.java:375-394).
template <typename T>
struct EpiTransAdd {
    T *y;
    T alpha, inv;
    __device__ T init(int r) const { return y[r] * inv; }
    __device__ void row(int r, T dot) { y[r] = dot * alpha; }
};
