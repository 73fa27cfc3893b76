// C-ABI entry points: context, CompRowMatrix, Preconditioner.
// Each function maps C++ exceptions (RsError) to a status code and keeps the
// message for rsgpu_last_error().
#include <algorithm>
#include <atomic>

#include "spmv_core.hpp"

static std::atomic<uint64_t> g_pattern_id{1};
uint64_t next_pattern_id() { return g_pattern_id++; }

static int fail(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

#define RS_TRY(ctx) try {
#define RS_CATCH(ctx)                                                         \
    }                                                                         \
    catch (const RsError &e) { return fail(ctx, e); }                         \
    catch (const std::bad_alloc &) {                                          \
        return fail(ctx, RsError{RSGPU_OUT_OF_MEMORY, "host out of memory"}); \
    }

extern "C" int rsgpu_version(void) { return 1; }

// ---------------------------------------------------------------- context
extern "C" int rsgpu_ctx_create(int device, rsgpu_ctx **out)
{
    if (!out)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_ctx *c = new rsgpu_ctx();
    c->device = device;
    RS_TRY(c)
    RS_HIP(hipSetDevice(device));
    {
        const int rc = rs_rates_scratch_init(device, c->err);
        if (rc != RSGPU_OK)
            throw RsError{rc, c->err};
    }
    RS_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
    c->stream = c->own;
    *out = c;
    return RSGPU_OK;
    RS_CATCH(c)
}

extern "C" int rsgpu_ctx_destroy(rsgpu_ctx *c)
{
    if (!c)
        return RSGPU_OK;
    if (c->own) {
        (void)hipStreamSynchronize(c->own);
        (void)hipStreamDestroy(c->own);
    }
    delete c;
    return RSGPU_OK;
}

extern "C" int rsgpu_ctx_set_stream(rsgpu_ctx *c, void *stream)
{
    if (!c)
        return RSGPU_BAD_ARGUMENT;
    c->stream = stream ? (hipStream_t)stream : c->own;
    return RSGPU_OK;
}

extern "C" int rsgpu_ctx_synchronize(rsgpu_ctx *c)
{
    if (!c)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(c)
    RS_HIP(hipStreamSynchronize(c->stream));
    return RSGPU_OK;
    RS_CATCH(c)
}

extern "C" const char *rsgpu_last_error(rsgpu_ctx *c)
{
    return c ? c->err.c_str() : "no context";
}

// ---------------------------------------------------------------- CSR
extern "C" int rsgpu_csr_create(rsgpu_ctx *ctx, int n, int ncols, int nnz,
                                const int32_t *rowptr, const int32_t *colind,
                                rsgpu_csr **out)
{
    if (!ctx || !out || n < 0 || ncols < 0 || nnz < 0 || !rowptr ||
        (nnz > 0 && !colind))
        return RSGPU_BAD_ARGUMENT;
    // The whole rowptr is checked before any colind entry is read: a row
    // range past nnz would otherwise read past the caller's colind array.
    if (rowptr[0] != 0 || rowptr[n] != nnz) {
        ctx->err = "rowptr must start at 0 and end at nnz";
        return RSGPU_BAD_ARGUMENT;
    }
    for (int i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i] || rowptr[i + 1] > nnz) {
            ctx->err = "rowptr must be non-decreasing and at most nnz (row " +
                       std::to_string(i) + ")";
            return RSGPU_BAD_ARGUMENT;
        }
    rsgpu_csr *A = new rsgpu_csr();
    try {
    RS_HIP(hipSetDevice(ctx->device));
    A->ctx = ctx;
    A->n = n;
    A->ncols = ncols;
    A->nnz = nnz;
    A->h_rowptr.assign(rowptr, rowptr + n + 1);
    A->h_colind.assign(colind, colind + nnz);
    for (int i = 0; i < n; ++i) {
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            if (colind[j] < 0 || colind[j] >= ncols)
                throw RsError{RSGPU_BAD_ARGUMENT,
                              "nz[" + std::to_string(i) + "][" +
                                  std::to_string(j - rowptr[i]) + "]=" +
                                  std::to_string(colind[j]) +
                                  ", which is not a valid column index"};
        // Arrays.sort of each row (CompRowMatrix.java:188)
        std::sort(A->h_colind.begin() + rowptr[i],
                  A->h_colind.begin() + rowptr[i + 1]);
    }
    hipStream_t s = ctx->stream;
    A->rowptr.upload(A->h_rowptr, s);
    A->colind.upload(A->h_colind, s);
    A->val.alloc(std::max(nnz, 1));
    A->val.zero(s);
    spmv_build_plans(A->h_rowptr, n, A, s);
    A->pattern_id = next_pattern_id();
    RS_HIP(hipStreamSynchronize(s));
    *out = A;
    return RSGPU_OK;
    }
    catch (const RsError &e) {
        delete A;
        return fail(ctx, e);
    }
    catch (const std::bad_alloc &) {
        delete A;
        return fail(ctx, RsError{RSGPU_OUT_OF_MEMORY, "host out of memory"});
    }
}

// 3x3 block matrix: the expanded CSR with each block row's three scalar rows
// listing the row's blocks in ascending block column (three scalars per
// block), plus the block pattern for the block-row SpMV core.
extern "C" int rsgpu_csr_create_bsr3(rsgpu_ctx *ctx, int nb, int nnzb, const int32_t *brp,
                                     const int32_t *bcol, rsgpu_csr **out)
{
    if (!ctx || !out || nb < 0 || nnzb < 0 || !brp || (nnzb > 0 && !bcol))
        return RSGPU_BAD_ARGUMENT;
    if ((long long)nnzb * 9 > 0x7fffffffLL || (long long)nb * 3 > 0x7fffffffLL) {
        ctx->err = "block matrix exceeds 32-bit scalar indices";
        return RSGPU_BAD_ARGUMENT;
    }
    // The block rowptr is checked whole before any bcol entry is read.
    if (brp[0] != 0 || brp[nb] != nnzb) {
        ctx->err = "block rowptr must start at 0 and end at nnzb";
        return RSGPU_BAD_ARGUMENT;
    }
    for (int I = 0; I < nb; ++I)
        if (brp[I + 1] < brp[I] || brp[I + 1] > nnzb) {
            ctx->err = "block rowptr must be non-decreasing and at most nnzb (block row " +
                       std::to_string(I) + ")";
            return RSGPU_BAD_ARGUMENT;
        }
    rsgpu_csr *A = new rsgpu_csr();
    try {
    RS_HIP(hipSetDevice(ctx->device));
    std::vector<int> hb(brp, brp + nb + 1), hc(bcol, bcol + nnzb);
    for (int I = 0; I < nb; ++I) {
        for (int b = hb[I]; b < hb[I + 1]; ++b)
            if (hc[b] < 0 || hc[b] >= nb)
                throw RsError{RSGPU_BAD_ARGUMENT, "block column out of range"};
        // Arrays.sort of each row (CompRowMatrix.java:188), at block level
        std::sort(hc.begin() + hb[I], hc.begin() + hb[I + 1]);
        for (int b = hb[I] + 1; b < hb[I + 1]; ++b)
            if (hc[b] == hc[b - 1])
                throw RsError{RSGPU_BAD_ARGUMENT, "duplicate block column"};
    }
    const int n = 3 * nb, nnz = 9 * nnzb;
    A->ctx = ctx;
    A->n = A->ncols = n;
    A->nnz = nnz;
    A->bs = 3;
    A->nb = nb;
    A->nnzb = nnzb;
    A->h_rowptr.resize(n + 1);
    A->h_colind.resize(nnz);
    int p = 0;
    A->h_rowptr[0] = 0;
    for (int I = 0; I < nb; ++I)
        for (int a = 0; a < 3; ++a) {
            for (int b = hb[I]; b < hb[I + 1]; ++b)
                for (int c = 0; c < 3; ++c)
                    A->h_colind[p++] = 3 * hc[b] + c;
            A->h_rowptr[3 * I + a + 1] = p;
        }
    hipStream_t s = ctx->stream;
    A->rowptr.upload(A->h_rowptr, s);
    A->colind.upload(A->h_colind, s);
    A->brp.upload(hb, s);
    A->bcol.upload(hc.empty() ? std::vector<int>(1, 0) : hc, s);
    A->val.alloc(std::max(nnz, 1));
    A->val.zero(s);
    spmv_build_plans(A->h_rowptr, n, A, s, false);
    spmv_build_plan_b3(hb, nb, A->bplan, s, RS_SPMV_NNZ);
    spmv_build_plan_b3(hb, nb, A->bplan32, s, RS_SPMV_NNZ32);
    sliced_build3(A, hb, hc, s);
    A->pattern_id = next_pattern_id();
    RS_HIP(hipStreamSynchronize(s));
    *out = A;
    return RSGPU_OK;
    }
    catch (const RsError &e) {
        delete A;
        return fail(ctx, e);
    }
    catch (const std::bad_alloc &) {
        delete A;
        return fail(ctx, RsError{RSGPU_OUT_OF_MEMORY, "host out of memory"});
    }
}

extern "C" int rsgpu_csr_block_info(const rsgpu_csr *A, int *bs, int *nb, int *nnzb)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    if (bs)
        *bs = A->bs;
    if (nb)
        *nb = A->bs == 3 ? A->nb : A->n;
    if (nnzb)
        *nnzb = A->bs == 3 ? A->nnzb : A->nnz;
    return RSGPU_OK;
}

extern "C" int rsgpu_csr_transpose(const rsgpu_csr *A, rsgpu_csr **out)
{
    if (!A || !out)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_csr *T = new rsgpu_csr();
    RS_TRY(A->ctx)
    RS_HIP(hipSetDevice(A->ctx->device));
    hipStream_t s = A->ctx->stream;
    const int n = A->n, m = A->ncols, nnz = A->nnz;
    std::vector<int> trp(m + 1, 0), tci(std::max(nnz, 0)), tpm(std::max(nnz, 1), 0);
    for (int j = 0; j < nnz; ++j)
        trp[A->h_colind[j] + 1]++;
    for (int c = 0; c < m; ++c)
        trp[c + 1] += trp[c];
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int j = A->h_rowptr[i]; j < A->h_rowptr[i + 1]; ++j) {
            const int p = fill[A->h_colind[j]]++;
            tci[p] = i;
            tpm[p] = j;
        }
    T->ctx = A->ctx;
    T->n = m;
    T->ncols = n;
    T->nnz = nnz;
    T->h_rowptr = trp;
    T->h_colind = tci;
    T->rowptr.upload(trp, s);
    T->colind.upload(tci, s);
    T->val.alloc(std::max(nnz, 1));
    DevBuf<int> perm;
    perm.upload(tpm, s);
    csr_launch_gather(s, perm.p, A->val.p, T->val.p, nnz);
    spmv_build_plans(T->h_rowptr, m, T, s);
    T->pattern_id = next_pattern_id();
    T->transpose_of = A->pattern_id;
    RS_HIP(hipStreamSynchronize(s));
    *out = T;
    return RSGPU_OK;
    }
    catch (const RsError &e) {
        delete T;
        return fail(A->ctx, e);
    }
    catch (const std::bad_alloc &) {
        delete T;
        return fail(A->ctx, RsError{RSGPU_OUT_OF_MEMORY, "host out of memory"});
    }
}

extern "C" int rsgpu_csr_destroy(rsgpu_csr *A)
{
    if (A) {
        if (A->ctx)
            (void)hipStreamSynchronize(A->ctx->stream);
        delete A;
    }
    return RSGPU_OK;
}

extern "C" int rsgpu_csr_get_pattern(const rsgpu_csr *A, int32_t *rowptr,
                                     int32_t *colind)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    if (rowptr)
        std::copy(A->h_rowptr.begin(), A->h_rowptr.end(), rowptr);
    if (colind)
        std::copy(A->h_colind.begin(), A->h_colind.end(), colind);
    return RSGPU_OK;
}

extern "C" int rsgpu_csr_set_values(rsgpu_csr *A, const double *data)
{
    if (!A || (!data && A->nnz))
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    if (A->nnz) {
        RS_HIP(hipMemcpyAsync(A->val.p, data, sizeof(double) * A->nnz,
                              hipMemcpyHostToDevice, A->ctx->stream));
        RS_HIP(hipStreamSynchronize(A->ctx->stream));
    }
    A->values_changed();
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_set_values_dev(rsgpu_csr *A, const double *data)
{
    if (!A || (!data && A->nnz))
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    if (A->nnz)
        RS_HIP(hipMemcpyAsync(A->val.p, data, sizeof(double) * A->nnz,
                              hipMemcpyDeviceToDevice, A->ctx->stream));
    A->values_changed();
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_get_values(const rsgpu_csr *A, double *data)
{
    if (!A || (!data && A->nnz))
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    if (A->nnz) {
        RS_HIP(hipMemcpyAsync(data, A->val.p, sizeof(double) * A->nnz,
                              hipMemcpyDeviceToHost, A->ctx->stream));
        RS_HIP(hipStreamSynchronize(A->ctx->stream));
    }
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" double *rsgpu_csr_values_dev(rsgpu_csr *A) { return A ? A->val.p : nullptr; }

extern "C" int rsgpu_csr_zero(rsgpu_csr *A)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    A->val.zero(A->ctx->stream);
    A->values_changed();
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

// CompRowMatrix.add for a batch, in array order.  Index lookup on the host
// pattern (binary search as getIndex), values accumulated on the device by
// one ordered pass per row so that repeated (row,col) pairs add in order.
__global__ void k_add_ordered(const int *__restrict__ rp, const int *__restrict__ pos,
                              const double *__restrict__ v, double *val, int nrows_touched)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows_touched)
        return;
    for (int q = rp[t]; q < rp[t + 1]; ++q)
        val[pos[q]] += v[q];
}

extern "C" int rsgpu_csr_add_batch(rsgpu_csr *A, int count, const int32_t *rows,
                                   const int32_t *cols, const double *vals)
{
    if (!A || count < 0 || (count && (!rows || !cols || !vals)))
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    if (count == 0)
        return RSGPU_OK;
    std::vector<int> pos(count);
    for (int q = 0; q < count; ++q) {
        const int r = rows[q], c = cols[q];
        if (r < 0 || r >= A->n || c < 0 || c >= A->ncols)
            throw RsError{RSGPU_BAD_ARGUMENT, "index out of range"};
        auto b = A->h_colind.begin() + A->h_rowptr[r],
             e = A->h_colind.begin() + A->h_rowptr[r + 1];
        auto it = std::lower_bound(b, e, c);
        if (it == e || *it != c)
            throw RsError{RSGPU_NOT_IN_PATTERN,
                          "Entry (" + std::to_string(r + 1) + ", " +
                              std::to_string(c + 1) +
                              ") is not in the matrix structure"};
        pos[q] = (int)(it - A->h_colind.begin());
    }
    // group by target position keeping array order: entries for the same
    // position are applied sequentially by one thread
    std::vector<int> order(count);
    for (int q = 0; q < count; ++q)
        order[q] = q;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return pos[a] < pos[b]; });
    std::vector<int> grp_ptr{0}, gpos(count);
    std::vector<double> gval(count);
    for (int q = 0; q < count; ++q) {
        gpos[q] = pos[order[q]];
        gval[q] = vals[order[q]];
        if (q > 0 && gpos[q] != gpos[q - 1])
            grp_ptr.push_back(q);
    }
    grp_ptr.push_back(count);
    hipStream_t s = A->ctx->stream;
    DevBuf<int> dgp, dpos;
    DevBuf<double> dv;
    dgp.upload(grp_ptr, s);
    dpos.upload(gpos, s);
    dv.upload(gval, s);
    const int ng = (int)grp_ptr.size() - 1;
    hipLaunchKernelGGL(k_add_ordered, dim3((ng + 255) / 256), dim3(256), 0, s,
                       dgp.p, dpos.p, dv.p, A->val.p, ng);
    RS_CHECK_LAUNCH();
    A->values_changed();
    RS_HIP(hipStreamSynchronize(s));
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_zero_rows(rsgpu_csr *A, double diagonal,
                                   const int32_t *rows, int nrows)
{
    if (!A || nrows < 0 || (nrows && !rows))
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    for (int q = 0; q < nrows; ++q)
        if (rows[q] < 0 || rows[q] >= A->n)
            throw RsError{RSGPU_BAD_ARGUMENT, "row index out of range"};
    if (diagonal != 0)
        for (int q = 0; q < nrows; ++q) {
            const int r = rows[q];
            if (r >= A->ncols ||
                !std::binary_search(A->h_colind.begin() + A->h_rowptr[r],
                                    A->h_colind.begin() + A->h_rowptr[r + 1], r))
                throw RsError{RSGPU_NOT_IN_PATTERN,
                              "Entry (" + std::to_string(r + 1) + ", " +
                                  std::to_string(r + 1) +
                                  ") is not in the matrix structure"};
        }
    hipStream_t s = A->ctx->stream;
    DevBuf<int> d;
    d.upload(rows, nrows, s);
    csr_zero_rows_dev(A, diagonal, d.p, nrows);
    RS_HIP(hipStreamSynchronize(s));
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_mult(rsgpu_csr *A, const double *x, double *y)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    spmv_launch<double>(A->ctx->stream, A, A->val.p, x, y, 0, 0.0, nullptr);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_multadd(rsgpu_csr *A, double alpha, const double *x,
                                 double *y)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    spmv_launch<double>(A->ctx->stream, A, A->val.p, x, y, 1, alpha, nullptr);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_transmult(rsgpu_csr *A, const double *x, double *y)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    csr_transmult_launch(A, 1.0, x, y, false);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_transmultadd(rsgpu_csr *A, double alpha, const double *x,
                                      double *y)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    csr_transmult_launch(A, alpha, x, y, true);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_residual_dd(rsgpu_csr *A, const double *b, const double *x,
                                     double *r)
{
    if (!A || (A->n && (!b || !x || !r)))
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    csr_residual_dd(A, b, x, r);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_sync_f32(rsgpu_csr *A)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    A->val32.alloc(std::max(A->nnz, 1));
    convert_f64_to_f32(A->ctx->stream, A->val.p, A->val32.p, A->nnz);
    sliced_fill<float>(A, A->ctx->stream);
    A->gen32 = A->val_gen;
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_sliced_info(const rsgpu_csr *A, int *nchunks, int64_t *nslots,
                                     int *ntables)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    if (nchunks)
        *nchunks = A->sl.nchunks;
    if (nslots)
        *nslots = A->sl.nslots;
    if (ntables)
        *ntables = A->sl.ntables;
    return RSGPU_OK;
}

extern "C" int rsgpu_csr_sync_sliced(rsgpu_csr *A)
{
    if (!A)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    if (!A->sl.ok()) {
        A->ctx->err = "the matrix has no sliced copy";
        return RSGPU_BAD_ARGUMENT;
    }
    sliced_fill<double>(A, A->ctx->stream);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_mult_sliced(rsgpu_csr *A, const double *x, double *y)
{
    if (!A || !A->sl.ok() || A->sl.gen64 != A->val_gen) {
        if (A)
            A->ctx->err = "mult_sliced needs rsgpu_csr_sync_sliced after the last value change";
        return RSGPU_BAD_ARGUMENT;
    }
    RS_TRY(A->ctx)
    sliced_mult<double>(A->ctx->stream, A, x, y);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_mult_sliced_f32(rsgpu_csr *A, const float *x, float *y)
{
    if (!A || !A->sl.ok() || !A->sl.v32.p || !A->f32_current()) {
        if (A)
            A->ctx->err = "mult_sliced_f32 needs rsgpu_csr_sync_f32 after the last value change";
        return RSGPU_BAD_ARGUMENT;
    }
    RS_TRY(A->ctx)
    sliced_mult<float>(A->ctx->stream, A, x, y);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_mult_f32(rsgpu_csr *A, const float *x, float *y)
{
    if (!A || !A->f32_current()) {
        if (A)
            A->ctx->err = "mult_f32 needs rsgpu_csr_sync_f32 after the last value change";
        return RSGPU_BAD_ARGUMENT;
    }
    RS_TRY(A->ctx)
    spmv_launch<float>(A->ctx->stream, A, A->val32.p, x, y, 0, 0.0, nullptr);
    return RSGPU_OK;
    RS_CATCH(A->ctx)
}

// inout: the output vector's content is an input too (Preconditioner.apply
// of SSOR/AMG starts from x's old content)
template <class F>
static int host_roundtrip(rsgpu_ctx *ctx, int nin, const double *in_host,
                          int nout, double *out_host, F f, bool inout = false)
{
    hipStream_t s = ctx->stream;
    DevBuf<double> a, b;
    a.upload(in_host, nin, s);
    if (inout)
        b.upload(out_host, nout, s);
    b.alloc(std::max(nout, 1));
    f(a.p, b.p);
    if (nout)
        RS_HIP(hipMemcpyAsync(out_host, b.p, sizeof(double) * nout,
                              hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    return RSGPU_OK;
}

extern "C" int rsgpu_csr_mult_host(rsgpu_csr *A, const double *x, double *y)
{
    if (!A || !x || !y)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    return host_roundtrip(A->ctx, A->ncols, x, A->n, y, [&](double *a, double *b) {
        spmv_launch<double>(A->ctx->stream, A, A->val.p, a, b, 0, 0.0, nullptr);
    });
    RS_CATCH(A->ctx)
}

extern "C" int rsgpu_csr_transmult_host(rsgpu_csr *A, const double *x, double *y)
{
    if (!A || !x || !y)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(A->ctx)
    return host_roundtrip(A->ctx, A->n, x, A->ncols, y, [&](double *a, double *b) {
        csr_transmult_launch(A, 1.0, a, b, false);
    });
    RS_CATCH(A->ctx)
}

// ---------------------------------------------------------------- preconditioner
extern "C" int rsgpu_prec_create(rsgpu_ctx *ctx, int kind, int nblocks,
                                 int precision, rsgpu_prec **out)
{
    if (!ctx || !out)
        return RSGPU_BAD_ARGUMENT;
    if (kind != RSGPU_PREC_IDENTITY && kind != RSGPU_PREC_ILU0 &&
        kind != RSGPU_PREC_BLOCK_ILU0 && kind != RSGPU_PREC_AMG &&
        kind != RSGPU_PREC_DIAGONAL && kind != RSGPU_PREC_SSOR && kind != RSGPU_PREC_ICC &&
        kind != RSGPU_PREC_ILUT) {
        ctx->err = "preconditioner kind not available";
        return RSGPU_BAD_ARGUMENT;
    }
    if (precision != RSGPU_FP64 && precision != RSGPU_FP32)
        return RSGPU_BAD_ARGUMENT;
    if ((kind == RSGPU_PREC_SSOR || kind == RSGPU_PREC_ICC || kind == RSGPU_PREC_ILUT) &&
        precision != RSGPU_FP64) {
        ctx->err = "SSOR, ICC and ILUT run in FP64";
        return RSGPU_BAD_ARGUMENT;
    }
    if (kind == RSGPU_PREC_BLOCK_ILU0 && nblocks < 1)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_prec *M = new rsgpu_prec();
    M->ctx = ctx;
    M->kind = kind;
    M->nblocks = kind == RSGPU_PREC_BLOCK_ILU0 ? nblocks : 1;
    M->precision = precision;
    M->ready = (kind == RSGPU_PREC_IDENTITY);
    rsgpu_amg_params_default(&M->amg_params);
    if (kind == RSGPU_PREC_SSOR) {
        // SSOR(F) defaults: reverse sweep, omega 1 (SSOR.java:103-105)
        M->amg_params.omega_pre_f = 1;
        M->amg_params.omega_pre_r = 1;
        M->amg_params.reverse = 1;
        M->amg_params.block_rows = 0;
        M->amg_params.ordering = 0;
    }
    *out = M;
    return RSGPU_OK;
}

extern "C" int rsgpu_prec_set_ssor(rsgpu_prec *M, double omega_f, double omega_r, int reverse,
                                   int block_rows, int ordering)
{
    if (!M || M->kind != RSGPU_PREC_SSOR)
        return RSGPU_BAD_ARGUMENT;
    // SSOR.setOmega (SSOR.java:122-130)
    if (omega_f < 0 || omega_f > 2 || omega_r < 0 || omega_r > 2 || block_rows < 0 ||
        ordering < 0 || ordering > 1) {
        M->ctx->err = "omegaF and omegaR must be between 0 and 2";
        return RSGPU_BAD_ARGUMENT;
    }
    M->amg_params.omega_pre_f = omega_f;
    M->amg_params.omega_pre_r = omega_r;
    M->amg_params.reverse = reverse ? 1 : 0;
    M->amg_params.block_rows = block_rows;
    M->amg_params.ordering = ordering;
    M->ready = false;
    return RSGPU_OK;
}

extern "C" int rsgpu_prec_set_ilut(rsgpu_prec *M, double tau, int p)
{
    if (!M || M->kind != RSGPU_PREC_ILUT)
        return RSGPU_BAD_ARGUMENT;
    M->ilut_tau = tau;
    M->ilut_p = p;
    M->ready = false;
    return RSGPU_OK;
}

extern "C" int rsgpu_prec_set_blocks(rsgpu_prec *M, int nblocks, const int32_t *bstart_host)
{
    if (!M || M->kind != RSGPU_PREC_BLOCK_ILU0 || nblocks < 1 || !bstart_host)
        return RSGPU_BAD_ARGUMENT;
    if (bstart_host[0] != 0) {
        M->ctx->err = "block boundaries must start at row 0";
        return RSGPU_BAD_ARGUMENT;
    }
    for (int b = 0; b < nblocks; ++b)
        if (bstart_host[b + 1] <= bstart_host[b]) {
            M->ctx->err = "block boundaries must rise strictly";
            return RSGPU_BAD_ARGUMENT;
        }
    M->user_bstart.assign(bstart_host, bstart_host + nblocks + 1);
    M->nblocks = nblocks;
    M->ready = false;
    return RSGPU_OK;
}

extern "C" int rsgpu_prec_set_amg_params(rsgpu_prec *M, const rsgpu_amg_params *p)
{
    if (!M || !p || M->kind != RSGPU_PREC_AMG)
        return RSGPU_BAD_ARGUMENT;
    // SSOR.setOmega range (SSOR.java:122-130)
    if (p->omega_pre_f < 0 || p->omega_pre_f > 2 || p->omega_pre_r < 0 || p->omega_pre_r > 2 ||
        p->omega_post_f < 0 || p->omega_post_f > 2 || p->omega_post_r < 0 ||
        p->omega_post_r > 2 || p->nu1 < 0 || p->nu2 < 0 || p->gamma < 1 || p->min < 0 ||
        p->block_rows < 0 || p->max_coarse < 1 || p->ordering < 0 || p->ordering > 1 ||
        !(p->coarsen_ratio >= 0 && p->coarsen_ratio <= 1)) {
        M->ctx->err = "AMG parameter out of range";
        return RSGPU_BAD_ARGUMENT;
    }
    M->amg_params = *p;
    return RSGPU_OK;
}

extern "C" int rsgpu_amg_levels(const rsgpu_prec *M, int *sizes, int *nnzs, int cap)
{
    if (!M || M->kind != RSGPU_PREC_AMG || !M->ready)
        return 0;
    try {
        return amg_levels(M, sizes, nnzs, cap);
    } catch (...) {
        return 0;
    }
}

extern "C" int rsgpu_amg_get_level(const rsgpu_prec *M, int k, int32_t *rowptr,
                                   int32_t *colind, double *data, int32_t *pt, int32_t *Prp,
                                   int32_t *Pci, double *Pv, int *pnnz)
{
    if (!M || M->kind != RSGPU_PREC_AMG || !M->ready)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    RS_HIP(hipStreamSynchronize(M->ctx->stream));
    return amg_get_level(M, k, rowptr, colind, data, pt, Prp, Pci, Pv, pnnz);
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_destroy(rsgpu_prec *M)
{
    if (M) {
        if (M->ctx)
            (void)hipStreamSynchronize(M->ctx->stream);
        delete M;
    }
    return RSGPU_OK;
}

extern "C" int rsgpu_prec_set_matrix(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    if (!M || !A)
        return RSGPU_BAD_ARGUMENT;
    if (bad_row)
        *bad_row = -1;
    RS_TRY(M->ctx)
    RS_HIP(hipSetDevice(M->ctx->device));
    if (M->kind == RSGPU_PREC_IDENTITY) {
        M->n = A->n;
        M->pattern_id = A->pattern_id;
        M->ready = true;
        return RSGPU_OK;
    }
    if (M->kind == RSGPU_PREC_AMG || M->kind == RSGPU_PREC_SSOR ||
        M->kind == RSGPU_PREC_DIAGONAL) {
        M->ready = false;
        M->n = A->n;
        M->pattern_id = A->pattern_id;
        const int st = M->kind == RSGPU_PREC_AMG    ? amg_setup(M, A, bad_row)
                       : M->kind == RSGPU_PREC_SSOR ? ssor_setup(M, A, bad_row)
                                                    : diag_setup(M, A, bad_row);
        M->ready = st == RSGPU_OK;
        M->nnz = A->nnz;
        return st;
    }
    if (M->kind == RSGPU_PREC_ILUT) {
        try {
            return ilut_set_matrix(M, A, bad_row);
        } catch (const RsError &e) {
            M->ready = false;
            throw;
        }
    }
    if (M->kind == RSGPU_PREC_ICC) {
        int st = RSGPU_OK;
        try {
            st = icc_set_matrix(M, A, bad_row);
        } catch (const RsError &e) {
            if (e.code == RSGPU_MISSING_DIAGONAL && bad_row) {
                auto p = e.msg.rfind(' ');
                *bad_row = std::stoi(e.msg.substr(p + 1)) - 1;
            }
            M->ready = false;
            throw;
        }
        if (st == RSGPU_ZERO_PIVOT)
            M->ctx->err = "Zero diagonal entry encountered on row " +
                          std::to_string((bad_row ? *bad_row : 0) + 1) +
                          " during ICC process";
        return st;
    }
    if (!M->ready || M->pattern_id != A->pattern_id) {
        try {
            prec_setup(M, A);
        } catch (const RsError &e) {
            if (e.code == RSGPU_MISSING_DIAGONAL && bad_row) {
                // message carries "row k+1"
                auto p = e.msg.rfind(' ');
                *bad_row = std::stoi(e.msg.substr(p + 1)) - 1;
            }
            M->ready = false;
            throw;
        }
    }
    int st = prec_factor(M, A, bad_row);
    if (st == RSGPU_ZERO_PIVOT)
        M->ctx->err = "Zero pivot encountered on row " +
                      std::to_string((bad_row ? *bad_row : 0) + 1) +
                      " during ILU process";
    return st;
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_apply(rsgpu_prec *M, const double *b, double *x)
{
    if (!M || !M->ready)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    prec_apply(M, b, x, nullptr);
    return RSGPU_OK;
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_apply_f32(rsgpu_prec *M, const float *b, float *x)
{
    if (!M || !M->ready)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    prec_apply(M, b, x, nullptr);
    return RSGPU_OK;
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_trans_apply(rsgpu_prec *M, const double *b, double *x)
{
    if (!M || !M->ready)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    prec_trans_apply(M, b, x);
    return RSGPU_OK;
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_apply_host(rsgpu_prec *M, const double *b, double *x)
{
    if (!M || !M->ready || !b || !x)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    return host_roundtrip(
        M->ctx, M->n, b, M->n, x, [&](double *a, double *o) { prec_apply(M, a, o, nullptr); },
        true);
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_trans_apply_host(rsgpu_prec *M, const double *b,
                                           double *x)
{
    if (!M || !M->ready || !b || !x)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    return host_roundtrip(
        M->ctx, M->n, b, M->n, x, [&](double *a, double *o) { prec_trans_apply(M, a, o); },
        true);
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_get_factors(const rsgpu_prec *M, int32_t *rowptr,
                                      int32_t *colind, double *lu, int *nnz)
{
    if (!M || !M->ready || M->kind == RSGPU_PREC_IDENTITY)
        return RSGPU_BAD_ARGUMENT;
    RS_TRY(M->ctx)
    if (nnz)
        *nnz = M->nnz;
    if (rowptr)
        std::copy(M->h_rowptr.begin(), M->h_rowptr.end(), rowptr);
    if (colind)
        std::copy(M->h_colind.begin(), M->h_colind.end(), colind);
    if (lu && M->nnz) {
        if (M->precision == RSGPU_FP64) {
            RS_HIP(hipMemcpy(lu, M->lu64.p, sizeof(double) * M->nnz,
                             hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(M->nnz);
            RS_HIP(hipMemcpy(t.data(), M->lu32.p, sizeof(float) * M->nnz,
                             hipMemcpyDeviceToHost));
            for (int i = 0; i < M->nnz; ++i)
                lu[i] = t[i];
        }
    }
    return RSGPU_OK;
    RS_CATCH(M->ctx)
}

extern "C" int rsgpu_prec_sweep_form(const rsgpu_prec *M, int *form)
{
    if (!M || !form)
        return RSGPU_BAD_ARGUMENT;
    *form = M->cell ? (prec_cell_stream(M) ? 3 : 2) : (M->band && prec_band_apply(M)) ? 1 : 0;
    return RSGPU_OK;
}

extern "C" int rsgpu_prec_info(const rsgpu_prec *M, int *nnz, int *nl, int *nu)
{
    if (!M)
        return RSGPU_BAD_ARGUMENT;
    if (nnz)
        *nnz = M->nnz;
    if (M->kind == RSGPU_PREC_AMG) {
        const int m = rsgpu_amg_levels(M, nullptr, nullptr, 0);
        if (nl)
            *nl = m;
        if (nu)
            *nu = m;
        return RSGPU_OK;
    }
    if (nl)
        *nl = M->L.max_levels;
    if (nu)
        *nu = M->U.max_levels;
    return RSGPU_OK;
}

// host-array solve: upload b and x, solve, download x
struct rsgpu_solver;
extern "C" int rsgpu_solver_solve_host(rsgpu_solver *S, rsgpu_csr *A,
                                       rsgpu_prec *M, const double *b,
                                       double *x, rsgpu_stats *st)
{
    if (!S || !A || !b || !x)
        return RSGPU_BAD_ARGUMENT;
    rsgpu_ctx *ctx = A->ctx;
    RS_TRY(ctx)
    hipStream_t s = ctx->stream;
    DevBuf<double> db, dx;
    db.upload(b, A->n, s);
    dx.upload(x, A->n, s);
    int rc = rsgpu_solver_solve(S, A, M, db.p, dx.p, st);
    if (A->n)
        RS_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * A->n,
                              hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    return rc;
    RS_CATCH(ctx)
}
