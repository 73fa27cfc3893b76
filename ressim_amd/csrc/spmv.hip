// CSR SpMV entry kernels, transposed copy, zeroRows, conversions.
#include <algorithm>
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

void spmv_build_plans(const std::vector<int> &rowptr, int n, rsgpu_csr *A, hipStream_t s)
{
    spmv_build_plan(rowptr, n, A->plan, s, RS_SPMV_NNZ, &A->h_colind);
    spmv_build_plan(rowptr, n, A->plan32, s, RS_SPMV_NNZ32, &A->h_colind);
}

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
