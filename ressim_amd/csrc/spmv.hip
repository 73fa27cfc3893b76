// CSR SpMV entry kernels, transposed copy, zeroRows, conversions.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <numeric>

#include "spmv_core.hpp"

// Greedy contiguous row partition: a group takes rows until the next one
// would overflow the LDS stage or the lane count; a longer row is alone.
void spmv_build_coff(const std::vector<int> &rowptr, const std::vector<int> &colind,
                     const std::vector<int> &starts, DevBuf<short> &coff, hipStream_t s)
{
    const int ng = (int)starts.size() - 1;
    std::vector<short> off(colind.size());
    for (int g = 0; g < ng; ++g) {
        const int r0 = starts[g];
        for (int j = rowptr[r0]; j < rowptr[starts[g + 1]]; ++j) {
            const int d = colind[j] - r0;
            if (d < -32768 || d > 32767) {
                coff.release();
                return;
            }
            off[j] = (short)d;
        }
    }
    if (off.empty()) {
        coff.release();
        return;
    }
    coff.upload(off, s);
    // the host vector is released on return: wait for the copy
    RS_HIP(hipStreamSynchronize(s));
}

void spmv_build_plan(const std::vector<int> &rowptr, int n, SpmvPlan &plan,
                     hipStream_t s, int nnz_wg, const std::vector<int> *colind)
{
    std::vector<int> starts;
    starts.reserve(n / 16 + 2);
    int r = 0;
    while (r < n) {
        starts.push_back(r);
        const int j0 = rowptr[r];
        int e = r + 1;
        if (rowptr[r + 1] - j0 <= nnz_wg)
            while (e < n && e - r < RS_SPMV_BS &&
                   rowptr[e + 1] - j0 <= nnz_wg)
                ++e;
        r = e;
    }
    starts.push_back(n);
    plan.nwg = (int)starts.size() - 1;
    plan.wg_rows.upload(starts, s);
    if (colind)
        spmv_build_coff(rowptr, *colind, starts, plan.coff, s);
    else
        plan.coff.release();
    RS_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- sliced
// RSGPU_SLICED=0 leaves the sliced copy out (the solvers then run the CSR core)
static bool sliced_enabled()
{
    const char *e = std::getenv("RSGPU_SLICED");
    return !(e && e[0] == '0');
}

// Chunk tables, shared tables and row masks of SlicedFmt from the host
// pattern.  The format is left out when a chunk's union exceeds 32 offsets,
// a row repeats a column, or the slot count passes 2^31 slot rows.
static void sliced_reset(SlicedFmt &F)
{
    F.nchunks = F.ntables = 0;
    F.nslots = 0;
    F.gen64 = 0;
    F.b3 = false;
    F.v64.release();
    F.v32.release();
}

static void sliced_build(rsgpu_csr *A, hipStream_t s)
{
    SlicedFmt &F = A->sl;
    sliced_reset(F);
    const int n = A->n;
    if (n == 0 || A->bs != 1 || !sliced_enabled())
        return;
    const std::vector<int> &rp = A->h_rowptr, &ci = A->h_colind;
    const int nch = (n + RS_SL_ROWS - 1) / RS_SL_ROWS;
    std::vector<int> ctab(nch), cbase(nch), tptr{0}, toff;
    std::vector<unsigned> mask(n, 0u);
    std::map<std::vector<int>, int> tables;
    std::vector<int> u, merged;
    long long rows = 0;
    for (int c = 0; c < nch; ++c) {
        const int r0 = RS_SL_ROWS * c, r1 = std::min(n, r0 + RS_SL_ROWS);
        u.clear();
        for (int r = r0; r < r1; ++r) {
            size_t i = 0;
            int j = rp[r];
            const int e = rp[r + 1];
            merged.clear();
            while (i < u.size() || j < e) {
                if (j > rp[r] && j < e && ci[j] == ci[j - 1])
                    return;
                const int d = j < e ? ci[j] - r : 0;
                if (j == e || (i < u.size() && u[i] < d)) {
                    merged.push_back(u[i++]);
                } else if (i == u.size() || d < u[i]) {
                    merged.push_back(d);
                    ++j;
                } else {
                    merged.push_back(u[i++]);
                    ++j;
                }
            }
            u.swap(merged);
            if (u.size() > 32)
                return;
        }
        for (int r = r0; r < r1; ++r) {
            unsigned mk = 0;
            size_t i = 0;
            for (int j = rp[r]; j < rp[r + 1]; ++j) {
                while (u[i] != ci[j] - r)
                    ++i;
                mk |= 1u << i;
            }
            mask[r] = mk;
        }
        auto it = tables.find(u);
        int t;
        if (it == tables.end()) {
            t = (int)tables.size();
            tables.emplace(u, t);
            toff.insert(toff.end(), u.begin(), u.end());
            tptr.push_back((int)toff.size());
        } else {
            t = it->second;
        }
        ctab[c] = t;
        cbase[c] = (int)rows;
        rows += (long long)u.size();
        if (rows >= (1LL << 31))
            return;
    }
    // the kernels read 32 offsets from every table's start
    toff.insert(toff.end(), 32, 0);
    F.ctab.upload(ctab, s);
    F.cbase.upload(cbase, s);
    F.tptr.upload(tptr, s);
    F.toff.upload(toff, s);
    F.mask.upload(mask, s);
    RS_HIP(hipStreamSynchronize(s));
    F.ntables = (int)tables.size();
    F.nslots = rows * RS_SL_ROWS;
    F.nchunks = nch;
}

// Block form of SlicedFmt for a 3x3 block matrix from its block pattern
// (rows sorted, no repeats): per chunk of scalar rows, the ascending union
// of the block offsets of every block row with a scalar row in the chunk
// (a block row can straddle two chunks; each of its scalar rows takes the
// mask of its own chunk's table).  Left out when a union exceeds
// RS_SL3_MAXB offsets or the slot count passes 2^31 slot rows.
void sliced_build3(rsgpu_csr *A, const std::vector<int> &brp, const std::vector<int> &bcol,
                   hipStream_t s)
{
    SlicedFmt &F = A->sl;
    sliced_reset(F);
    const int n = A->n;
    if (n == 0 || A->bs != 3 || !sliced_enabled())
        return;
    const int nch = (n + RS_SL_ROWS - 1) / RS_SL_ROWS;
    std::vector<int> ctab(nch), cbase(nch), tptr{0}, toff;
    std::vector<unsigned short> mask(n, 0);
    std::map<std::vector<int>, int> tables;
    std::vector<int> u;
    long long rows = 0;
    for (int c = 0; c < nch; ++c) {
        const int r0 = RS_SL_ROWS * c, r1 = std::min(n, r0 + RS_SL_ROWS);
        const int I0 = r0 / 3, I1 = (r1 - 1) / 3;
        u.clear();
        for (int I = I0; I <= I1; ++I)
            for (int b = brp[I]; b < brp[I + 1]; ++b)
                u.push_back(bcol[b] - I);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if (u.size() > RS_SL3_MAXB)
            return;
        for (int r = r0; r < r1; ++r) {
            const int I = r / 3;
            unsigned mk = 0;
            for (int b = brp[I]; b < brp[I + 1]; ++b)
                mk |= 1u << (std::lower_bound(u.begin(), u.end(), bcol[b] - I) - u.begin());
            mask[r] = (unsigned short)mk;
        }
        auto it = tables.find(u);
        int t;
        if (it == tables.end()) {
            t = (int)tables.size();
            tables.emplace(u, t);
            toff.insert(toff.end(), u.begin(), u.end());
            tptr.push_back((int)toff.size());
        } else {
            t = it->second;
        }
        ctab[c] = t;
        cbase[c] = (int)rows;
        rows += 3LL * (long long)u.size();
        if (rows >= (1LL << 31))
            return;
    }
    toff.insert(toff.end(), 32, 0);
    F.ctab.upload(ctab, s);
    F.cbase.upload(cbase, s);
    F.tptr.upload(tptr, s);
    F.toff.upload(toff, s);
    F.mask16.upload(mask, s);
    RS_HIP(hipStreamSynchronize(s));
    F.ntables = (int)tables.size();
    F.nslots = rows * RS_SL_ROWS;
    F.b3 = true;
    F.nchunks = nch;
}

void spmv_build_plans(const std::vector<int> &rowptr, int n, rsgpu_csr *A, hipStream_t s,
                      bool sliced)
{
    spmv_build_plan(rowptr, n, A->plan, s, RS_SPMV_NNZ, &A->h_colind);
    spmv_build_plan(rowptr, n, A->plan32, s, RS_SPMV_NNZ32, &A->h_colind);
    if (sliced)
        sliced_build(A, s);
}

// Sliced values from the CSR values (zero where a row lacks the slot), one
// workgroup per chunk, one wave per 64 rows.  A wave's 64 rows are one
// contiguous run of the CSR values (at most 32 entries per row, 16 KiB):
// the wave stages the run in LDS with coalesced loads, then each lane walks
// its row's slots and stores slot by slot, 64 consecutive values per store
// across the wave.  Scalar form: slot s holds the row's next entry when
// mask bit s is set.  Block form: slots 3k..3k+2 hold the next three
// entries when block bit k is set.
#define RS_FILL_ROW_MAX 32
template <typename T, bool B3>
__global__ void __launch_bounds__(256)
    k_sliced_fill(SlView S, const int *__restrict__ rowptr, const double *__restrict__ val,
                  T *__restrict__ out)
{
    __shared__ double stage[4][64 * RS_FILL_ROW_MAX];
    const int c = (int)blockIdx.x, wave = (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int i = wave * 64 + lane;                   // row within the chunk
    const int r0 = c * RS_SL_ROWS + wave * 64;        // first row of the wave
    const int r = c * RS_SL_ROWS + i;
    double *st = stage[wave];
    if (r0 < S.n) {
        // eight loads per lane in flight per round (indices past the run
        // clamp to its last entry; only the LDS store is predicated)
        const int base = rowptr[r0], len = rowptr[min(r0 + 64, S.n)] - base;
        for (int j0 = 0; j0 < len; j0 += 8 * 64) {
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                v[q] = val[base + min(j0 + q * 64 + lane, len - 1)];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (j0 + q * 64 + lane < len)
                    st[j0 + q * 64 + lane] = v[q];
        }
    }
    __syncthreads();
    const int t = S.ctab[c];
    const int W = S.tptr[t + 1] - S.tptr[t];
    T *o = out + (size_t)S.cbase[c] * RS_SL_ROWS + i;
    int j = 0;
    unsigned m = 0;
    if (r < S.n) {
        j = rowptr[r] - rowptr[r0];
        m = B3 ? (unsigned)S.mask16[r] : S.mask[r];
    }
    if (B3) {
        for (int k = 0; k < W; ++k) {
            const bool has = (m >> k) & 1u;
            for (int e = 0; e < 3; ++e) {
                T v = T(0);
                if (has)
                    v = (T)st[j++];
                o[(size_t)(3 * k + e) * RS_SL_ROWS] = v;
            }
        }
    } else {
        for (int sl = 0; sl < W; ++sl) {
            T v = T(0);
            if ((m >> sl) & 1u)
                v = (T)st[j++];
            o[(size_t)sl * RS_SL_ROWS] = v;
        }
    }
}

template <typename T>
static DevBuf<T> &sliced_buf(rsgpu_csr *A);
template <>
DevBuf<double> &sliced_buf<double>(rsgpu_csr *A) { return A->sl.v64; }
template <>
DevBuf<float> &sliced_buf<float>(rsgpu_csr *A) { return A->sl.v32; }

template <typename T>
void sliced_fill(rsgpu_csr *A, hipStream_t s)
{
    SlicedFmt &F = A->sl;
    if (!F.ok())
        return;
    DevBuf<T> &b = sliced_buf<T>(A);
    b.alloc((size_t)F.nslots);
    hipLaunchKernelGGL((F.b3 ? k_sliced_fill<T, true> : k_sliced_fill<T, false>),
                       dim3((unsigned)F.nchunks), dim3(256), 0, s, A->slview(), A->rowptr.p,
                       A->val.p, b.p);
    RS_CHECK_LAUNCH();
    if (sizeof(T) == 8)
        F.gen64 = A->val_gen;
}
template void sliced_fill<double>(rsgpu_csr *, hipStream_t);
template void sliced_fill<float>(rsgpu_csr *, hipStream_t);

template <typename T, bool B3>
__global__ void __launch_bounds__(256)
    k_spmv_sliced(SlView S, const T *__restrict__ val, const T *__restrict__ x,
                  T *__restrict__ y)
{
    const int row0 = sliced_wave_row(xcd_remap((int)blockIdx.x, (int)gridDim.x));
    if (row0 >= S.n)
        return;
    EpiMult<T> epi{y};
    sliced_rows_any<B3, T>(row0, S, val, x, epi);
}

template <typename T>
void sliced_mult(hipStream_t s, const rsgpu_csr *A, const T *x, T *y)
{
    const T *val = sizeof(T) == 8 ? (const T *)A->sl.v64.p : (const T *)A->sl.v32.p;
    hipLaunchKernelGGL((A->sl.b3 ? k_spmv_sliced<T, true> : k_spmv_sliced<T, false>),
                       dim3(sliced_tiles(A->n)), dim3(256), 0, s, A->slview(), val, x, y);
    RS_CHECK_LAUNCH();
}
template void sliced_mult<double>(hipStream_t, const rsgpu_csr *, const double *, double *);
template void sliced_mult<float>(hipStream_t, const rsgpu_csr *, const float *, float *);

// Block-row plan of a 3x3 block matrix: whole block rows per group (at most
// RS_SPMV_BS / 3 of them and nnz_wg / 9 blocks); a block row with more
// blocks than the stage becomes three single-row groups.
void spmv_build_plan_b3(const std::vector<int> &brp, int nb, SpmvPlan &plan, hipStream_t s,
                        int nnz_wg)
{
    const int maxb = nnz_wg / 9, maxrows = RS_SPMV_BS / 3;
    std::vector<int> starts;
    starts.reserve(nb / 16 + 2);
    int I = 0;
    while (I < nb) {
        const int b0 = brp[I];
        if (brp[I + 1] - b0 > maxb) {
            for (int a = 0; a < 3; ++a)
                starts.push_back(3 * I + a);
            ++I;
            continue;
        }
        starts.push_back(3 * I);
        int e = I + 1;
        while (e < nb && e - I < maxrows && brp[e + 1] - b0 <= maxb)
            ++e;
        I = e;
    }
    starts.push_back(3 * nb);
    plan.nwg = (int)starts.size() - 1;
    plan.wg_rows.upload(starts, s);
}

template <typename T, class Epi, bool B3>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv(const int *__restrict__ wg_rows, const int *__restrict__ rowptr,
           const int *__restrict__ colind, const short *__restrict__ coff,
           const int *__restrict__ brp, const int *__restrict__ bcol, const T *__restrict__ val,
           const T *__restrict__ x, Epi epi)
{
    if (B3)
        spmv_core_b3<T, RS_SPMV_BS, spmv_nnz<T>()>(wg_rows, rowptr, colind, brp, bcol, val, x,
                                                   epi);
    else
        spmv_core_off<T, RS_SPMV_BS, spmv_nnz<T>()>(wg_rows, rowptr, colind, coff, val, x, epi);
}

template <typename T, class Epi>
static void launch_k_spmv(hipStream_t s, const rsgpu_csr *A, const T *val, const T *x, Epi epi)
{
    const SpmvPlan &pl = spmv_plan<T>(A);
    const dim3 g(pl.nwg), bl(RS_SPMV_BS);
    if (A->bs == 3)
        hipLaunchKernelGGL((k_spmv<T, Epi, true>), g, bl, 0, s, pl.wg_rows.p, A->rowptr.p,
                           A->colind.p, nullptr, A->brp.p, A->bcol.p, val, x, epi);
    else
        hipLaunchKernelGGL((k_spmv<T, Epi, false>), g, bl, 0, s, pl.wg_rows.p, A->rowptr.p,
                           A->colind.p, pl.coff_ptr(), nullptr, nullptr, val, x, epi);
}

// mode: 0 mult, 1 multAdd(alpha), 2 residual b + alpha*Ax, 3 plain sum,
// 4 transMultAdd(alpha) accumulator form
template <typename T>
void spmv_launch(hipStream_t s, const rsgpu_csr *A, const T *val, const T *x,
                 T *y, int mode, double alpha, const T *b)
{
    if (A->n == 0)
        return;
    switch (mode) {
    case 0:
        launch_k_spmv<T>(s, A, val, x, EpiMult<T>{y});
        break;
    case 1:
        launch_k_spmv<T>(s, A, val, x, EpiMultAdd<T>{y, (T)alpha});
        break;
    case 2:
        launch_k_spmv<T>(s, A, val, x, EpiResid<T>{y, b, (T)alpha});
        break;
    default:
        launch_k_spmv<T>(s, A, val, x, EpiPlain<T>{y});
        break;
    }
    RS_CHECK_LAUNCH();
}

template void spmv_launch<double>(hipStream_t, const rsgpu_csr *,
                                  const double *, const double *, double *,
                                  int, double, const double *);
template void spmv_launch<float>(hipStream_t, const rsgpu_csr *,
                                 const float *, const float *, float *, int,
                                 double, const float *);

// ---------------------------------------------------------------- transpose
__global__ void k_gather(const int *__restrict__ perm,
                         const double *__restrict__ src,
                         double *__restrict__ dst, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[perm[i]];
}

void csr_launch_gather(hipStream_t s, const int *perm, const double *src,
                       double *dst, int n)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, perm,
                       src, dst, n);
    RS_CHECK_LAUNCH();
}

// Transposed pattern: row c of A' lists the rows i with a(i,c) != 0 in
// ascending i, the order in which transMult's row-by-row scatter
// (CompRowMatrix.java:356-372) adds into y[c].
void csr_build_transpose(rsgpu_csr *A)
{
    if (A->has_t)
        return;
    hipStream_t s = A->ctx->stream;
    const int n = A->n, m = A->ncols, nnz = A->nnz;
    std::vector<int> trp(m + 1, 0), tci(nnz), tpm(nnz);
    for (int j = 0; j < nnz; ++j)
        trp[A->h_colind[j] + 1]++;
    for (int c = 0; c < m; ++c)
        trp[c + 1] += trp[c];
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int j = A->h_rowptr[i]; j < A->h_rowptr[i + 1]; ++j) {
            int c = A->h_colind[j];
            int p = fill[c]++;
            tci[p] = i;
            tpm[p] = j;
        }
    A->trowptr.upload(trp, s);
    A->tcolind.upload(tci, s);
    A->tperm.upload(tpm, s);
    A->tval.alloc(std::max(nnz, 1));
    spmv_build_plan(trp, m, A->tplan, s, RS_SPMV_NNZ);
    RS_HIP(hipStreamSynchronize(s));
    A->has_t = true;
}

template <class Epi>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_spmv_t(const int *__restrict__ wg_rows, const int *__restrict__ rowptr,
             const int *__restrict__ colind, const double *__restrict__ val,
             const double *__restrict__ x, Epi epi)
{
    spmv_core<double, RS_SPMV_BS, RS_SPMV_NNZ>(wg_rows, rowptr, colind, val, x,
                                               epi);
}

void csr_transmult_launch(rsgpu_csr *A, double alpha, const double *x,
                          double *y, bool add)
{
    csr_build_transpose(A);
    hipStream_t s = A->ctx->stream;
    if (A->ncols == 0)
        return;
    csr_launch_gather(s, A->tperm.p, A->val.p, A->tval.p, A->nnz);
    dim3 g(A->tplan.nwg), bl(RS_SPMV_BS);
    if (!add)
        hipLaunchKernelGGL((k_spmv_t<EpiPlain<double>>), g, bl, 0, s,
                           A->tplan.wg_rows.p, A->trowptr.p, A->tcolind.p,
                           A->tval.p, x, EpiPlain<double>{y});
    else
        hipLaunchKernelGGL((k_spmv_t<EpiTransAdd<double>>), g, bl, 0, s,
                           A->tplan.wg_rows.p, A->trowptr.p, A->tcolind.p,
                           A->tval.p, x,
                           EpiTransAdd<double>{y, alpha, 1. / alpha});
    RS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- compensated residual
// r = b - A x by Ogita, Rump and Oishi's Dot2 ("Accurate sum and dot
// product", SIAM J. Sci. Comput. 26, 2005, Algorithm 5.3): every product is
// split exactly (TwoProd through an FMA), the running sum p is carried by
// Knuth's TwoSum, all rounding errors are summed plainly in s, and the
// result is p + s.  Its error is at most eps |r| + gamma_n^2 sum |a_ij x_j|
// (their Proposition 5.5): as if computed in twice the working precision
// and rounded once.  Eight lanes per row, each over a strided share of the
// row, then the lanes' (p, s) pairs are combined by the same TwoSum step
// through lane shuffles.
struct DD {
    double hi, lo;   // running sum p, sum of rounding errors s
};

__device__ __forceinline__ DD dd_add(DD a, double b, double be)
{
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b - bb);
    return DD{s, a.lo + (e + be)};
}

__global__ void __launch_bounds__(256)
    k_residual_dd(int n, const int *__restrict__ rowptr, const int *__restrict__ colind,
                  const double *__restrict__ val, const double *__restrict__ b,
                  const double *__restrict__ x, double *__restrict__ r)
{
    const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 3);
    const int sub = threadIdx.x & 7;
    DD acc{0.0, 0.0};
    if (row < n) {
        if (sub == 0)
            acc.hi = b[row];
        for (int j = rowptr[row] + sub; j < rowptr[row + 1]; j += 8) {
            const double a = val[j], xv = x[colind[j]];
            const double p = a * xv;
            const double pe = __builtin_fma(a, xv, -p);
            acc = dd_add(acc, -p, -pe);
        }
    }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
        const double oh = __shfl_xor(acc.hi, off, 8);
        const double ol = __shfl_xor(acc.lo, off, 8);
        acc = dd_add(acc, oh, ol);
    }
    if (row < n && sub == 0)
        r[row] = acc.hi + acc.lo;
}

void csr_residual_dd(rsgpu_csr *A, const double *b, const double *x, double *r)
{
    if (A->n == 0)
        return;
    const long long threads = (long long)A->n * 8;
    hipLaunchKernelGGL(k_residual_dd, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       A->ctx->stream, A->n, A->rowptr.p, A->colind.p, A->val.p, b, x, r);
    RS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- zeroRows
__global__ void k_zero_rows(const int *__restrict__ rowptr,
                            const int *__restrict__ colind, double *val,
                            const int *__restrict__ rows, int nrows,
                            double diagonal)
{
    // one wave per listed row
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= nrows)
        return;
    const int r = rows[w];
    for (int j = rowptr[r] + lane; j < rowptr[r + 1]; j += 64)
        val[j] = (colind[j] == r) ? diagonal : 0.0;
}

void csr_zero_rows_dev(rsgpu_csr *A, double diagonal, const int *rows_dev,
                       int nrows)
{
    if (nrows <= 0)
        return;
    const int threads = nrows * 64;
    hipLaunchKernelGGL(k_zero_rows, dim3((threads + 255) / 256), dim3(256), 0,
                       A->ctx->stream, A->rowptr.p, A->colind.p, A->val.p,
                       rows_dev, nrows, diagonal);
    RS_CHECK_LAUNCH();
    A->values_changed();
}

// ---------------------------------------------------------------- conversions
__global__ void k_f64_to_f32(const double *__restrict__ a,
                             float *__restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = (float)a[i];
}

__global__ void k_f32_to_f64(const float *__restrict__ a,
                             double *__restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = (double)a[i];
}

void convert_f64_to_f32(hipStream_t s, const double *a, float *b, size_t n)
{
    if (!n)
        return;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, s, a, b, n);
    RS_CHECK_LAUNCH();
}

void convert_f32_to_f64(hipStream_t s, const float *a, double *b, size_t n)
{
    if (!n)
        return;
    hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, s, a, b, n);
    RS_CHECK_LAUNCH();
}
