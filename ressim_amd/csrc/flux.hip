// Connection fluxes and cell rates from MPFA-O transmissibilities.
//
// Reference: PotentialCalculator.calculateFlux
// (rs/src/no/uib/cipr/rs/numerics/PotentialCalculator.java:265-280):
// flux[c] = sum over c.MD, in MD order, of t_k * potential[k].  That is an
// SpMV over the per-connection lists (rows = connections, MD order kept),
// run by the CSR core with its row-sequential sum.  A cell's rate is its
// net outflow: the sum over its connections, in ascending connection
// order, of +flux (cell on the "here" side) or -flux ("there" side), which
// equals the cell's row of the flux-assembled matrix times the potential
// (PotentialCalculator.discretizeFlux, :319-333) up to summation order.
#include <algorithm>
#include <mutex>
#include <string>

#include "spmv_core.hpp"


int flux_md_lists(const rsgpu_flux *F, const int **md_ptr, const int **md_elem, int *nconn,
                  int *ncells)
{
    *md_ptr = F->md_ptr.p;
    *md_elem = F->md_elem.p;
    *nconn = F->nconn;
    *ncells = F->ncells;
    return RSGPU_OK;
}

static int flux_fail(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

__global__ void __launch_bounds__(RS_SPMV_BS)
    k_flux(const int *__restrict__ wg_rows, const int *__restrict__ rowptr,
           const int *__restrict__ colind, const double *__restrict__ val,
           const double *__restrict__ x, EpiPlain<double> epi)
{
    spmv_core<double, RS_SPMV_BS, RS_SPMV_NNZ>(wg_rows, rowptr, colind, val, x, epi);
}

__global__ void k_rates(int nsel, const int *__restrict__ cells, const int *__restrict__ cptr,
                        const int *__restrict__ cconn, const signed char *__restrict__ csign,
                        const double *__restrict__ flux, double *__restrict__ rates)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsel)
        return;
    const int c = cells[t];
    double q = 0;
    for (int e = cptr[c]; e < cptr[c + 1]; ++e) {
        const double f = flux[cconn[e]];
        q += csign[e] > 0 ? f : -f;
    }
    rates[t] = q;
}

// tbar[q] = (w[here(c)] - w[there(c)]) * p[md_elem[q]] for every MD entry q
// of connection c: the gradient with respect to the transmissibilities of
// sum_i w_i (net outflow of cell i) at fixed potential p.
__global__ void k_flux_vjp(int nconn, const int *__restrict__ md_ptr,
                           const int *__restrict__ md_elem, const int *__restrict__ here,
                           const int *__restrict__ there, const double *__restrict__ p,
                           const double *__restrict__ w, double *__restrict__ tbar)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nconn)
        return;
    const double a = w[here[c]] - w[there[c]];
    for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q)
        tbar[q] = a * p[md_elem[q]];
}

// host: cell -> connection lists from here/there, and the SpMV plan
static void flux_finish(rsgpu_flux *F, const std::vector<int> &md_ptr,
                        const std::vector<int> &here, const std::vector<int> &there)
{
    hipStream_t s = F->ctx->stream;
    const int nc = F->nconn, n = F->ncells;
    std::vector<int> cnt(n + 1, 0);
    for (int c = 0; c < nc; ++c) {
        if (here[c] < 0 || here[c] >= n || there[c] < 0 || there[c] >= n)
            throw RsError{RSGPU_BAD_ARGUMENT, "connection cell index out of range"};
        cnt[here[c] + 1]++;
        cnt[there[c] + 1]++;
    }
    for (int i = 0; i < n; ++i)
        cnt[i + 1] += cnt[i];
    std::vector<int> pos(cnt.begin(), cnt.end() - 1), conn(std::max(cnt[n], 1));
    std::vector<signed char> sign(std::max(cnt[n], 1));
    for (int c = 0; c < nc; ++c) {
        conn[pos[here[c]]] = c;
        sign[pos[here[c]]++] = 1;
        conn[pos[there[c]]] = c;
        sign[pos[there[c]]++] = -1;
    }
    F->here.upload(here.data(), std::max(nc, 0), s);
    F->there.upload(there.data(), std::max(nc, 0), s);
    if (nc == 0) {
        F->here.alloc(1);
        F->there.alloc(1);
    }
    F->cell_ptr.upload(cnt, s);
    F->cell_conn.upload(conn, s);
    F->cell_sign.upload(sign, s);
    F->nentries = md_ptr[nc];
    spmv_build_plan(md_ptr, nc, F->plan, s);
    RS_HIP(hipStreamSynchronize(s));
}

extern "C" int rsgpu_flux_create(rsgpu_ctx *ctx, int nconn, int ncells, const int32_t *md_ptr,
                                 const int32_t *md_elem, const double *md_t,
                                 const int32_t *here, const int32_t *there, rsgpu_flux **out)
{
    if (!ctx || !out || nconn < 0 || ncells < 0 || !md_ptr || (nconn && (!here || !there)))
        return RSGPU_BAD_ARGUMENT;
    rsgpu_flux *F = new rsgpu_flux();
    try {
        RS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        F->ctx = ctx;
        F->nconn = nconn;
        F->ncells = ncells;
        std::vector<int> mp(md_ptr, md_ptr + nconn + 1);
        if (mp[0] != 0)
            throw RsError{RSGPU_BAD_ARGUMENT, "md_ptr[0] must be 0"};
        for (int c = 0; c < nconn; ++c)
            if (mp[c + 1] < mp[c])
                throw RsError{RSGPU_BAD_ARGUMENT, "md_ptr must be nondecreasing"};
        const int tot = mp[nconn];
        for (int q = 0; q < tot; ++q)
            if (md_elem[q] < 0 || md_elem[q] >= ncells)
                throw RsError{RSGPU_BAD_ARGUMENT, "MD element index out of range"};
        F->md_ptr.upload(mp, s);
        F->md_elem.upload(md_elem, std::max(tot, 0), s);
        F->md_t.upload(md_t, std::max(tot, 0), s);
        if (tot == 0) {
            F->md_elem.alloc(1);
            F->md_t.alloc(1);
        }
        flux_finish(F, mp, std::vector<int>(here, here + nconn),
                    std::vector<int>(there, there + nconn));
        *out = F;
        return RSGPU_OK;
    } catch (const RsError &e) {
        delete F;
        return flux_fail(ctx, e);
    }
}

// Device-resident variant used by rsgpu_mpfa_o_flux (mpfa.hip)
rsgpu_flux *flux_from_device(rsgpu_ctx *ctx, int nconn, int ncells, const std::vector<int> &mp,
                             DevBuf<int> &md_elem, DevBuf<double> &md_t,
                             const std::vector<int> &here, const std::vector<int> &there)
{
    rsgpu_flux *F = new rsgpu_flux();
    try {
        F->ctx = ctx;
        F->nconn = nconn;
        F->ncells = ncells;
        F->md_ptr.upload(mp, ctx->stream);
        std::swap(F->md_elem.p, md_elem.p);
        std::swap(F->md_elem.n, md_elem.n);
        std::swap(F->md_t.p, md_t.p);
        std::swap(F->md_t.n, md_t.n);
        if (!F->md_elem.p)
            F->md_elem.alloc(1);
        if (!F->md_t.p)
            F->md_t.alloc(1);
        flux_finish(F, mp, here, there);
        return F;
    } catch (...) {
        delete F;
        throw;
    }
}

extern "C" int rsgpu_flux_destroy(rsgpu_flux *F)
{
    if (F) {
        if (F->ctx)
            (void)hipStreamSynchronize(F->ctx->stream);
        delete F;
    }
    return RSGPU_OK;
}

extern "C" int rsgpu_flux_info(const rsgpu_flux *F, int *nconn, int *ncells, int *nentries)
{
    if (!F)
        return RSGPU_BAD_ARGUMENT;
    if (nconn)
        *nconn = F->nconn;
    if (ncells)
        *ncells = F->ncells;
    if (nentries)
        *nentries = F->nentries;
    return RSGPU_OK;
}

extern "C" int rsgpu_flux_apply(rsgpu_flux *F, const double *p_dev, double *flux_dev)
{
    if (!F || (F->nconn && (!p_dev || !flux_dev)))
        return RSGPU_BAD_ARGUMENT;
    try {
        if (F->nconn) {
            hipLaunchKernelGGL(k_flux, dim3(F->plan.nwg), dim3(RS_SPMV_BS), 0, F->ctx->stream,
                               F->plan.wg_rows.p, F->md_ptr.p, F->md_elem.p, F->md_t.p, p_dev,
                               EpiPlain<double>{flux_dev});
            RS_CHECK_LAUNCH();
        }
        return RSGPU_OK;
    } catch (const RsError &e) {
        return flux_fail(F->ctx, e);
    }
}

extern "C" int rsgpu_flux_rates(rsgpu_flux *F, const double *flux_dev, int ncells_sel,
                                const int32_t *cells_host, double *rates_host)
{
    if (!F || ncells_sel < 0 || (ncells_sel && (!cells_host || !rates_host || !flux_dev)))
        return RSGPU_BAD_ARGUMENT;
    try {
        for (int t = 0; t < ncells_sel; ++t)
            if (cells_host[t] < 0 || cells_host[t] >= F->ncells)
                throw RsError{RSGPU_BAD_ARGUMENT, "cell index out of range"};
        if (!ncells_sel)
            return RSGPU_OK;
        hipStream_t s = F->ctx->stream;
        DevBuf<int> cells;
        DevBuf<double> rates;
        cells.upload(cells_host, ncells_sel, s);
        rates.alloc(ncells_sel);
        hipLaunchKernelGGL(k_rates, dim3((ncells_sel + 255) / 256), dim3(256), 0, s, ncells_sel,
                           cells.p, F->cell_ptr.p, F->cell_conn.p, F->cell_sign.p, flux_dev,
                           rates.p);
        RS_CHECK_LAUNCH();
        RS_HIP(hipMemcpyAsync(rates_host, rates.p, sizeof(double) * ncells_sel,
                              hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        return RSGPU_OK;
    } catch (const RsError &e) {
        return flux_fail(F->ctx, e);
    }
}

// ---------------------------------------------------------------- no-allocation rates
// rsgpu_flux_rates_noalloc stages the selected cells and their rates in a
// file-scope device array instead of allocating them per call: rates at
// byte 0 (RS_RATES_NOALLOC_MAX doubles), cells after them.  Each GPU the
// library runs on holds its own copy (768 KiB) until the process ends.  In
// host code the array's name is not a device address: host code uses only
// the per-device address hipGetSymbolAddress returned (g_scratch_ptr), taken
// once per device when the first context on it is created.
#define RS_RATES_NOALLOC_MAX 65536
#define RS_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char
    scratch[(size_t)RS_RATES_NOALLOC_MAX * (sizeof(double) + sizeof(int))];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[RS_SCRATCH_DEVICES];

// Called by rsgpu_ctx_create with `device` current.
int rs_rates_scratch_init(int device, std::string &err)
{
    if (device < 0 || device >= RS_SCRATCH_DEVICES) {
        err = "device ordinal beyond the library's per-device table";
        return RSGPU_BAD_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device])
        return RSGPU_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return RSGPU_HIP_ERROR;
    }
    g_scratch_ptr[device] = p;
    return RSGPU_OK;
}

extern "C" int rsgpu_flux_rates_noalloc(rsgpu_flux *F, const double *flux_dev, int ncells_sel,
                                        const int32_t *cells_host, double *rates_host)
{
    if (!F || ncells_sel < 0 || (ncells_sel && (!cells_host || !rates_host || !flux_dev)))
        return RSGPU_BAD_ARGUMENT;
    try {
        if (ncells_sel > RS_RATES_NOALLOC_MAX)
            throw RsError{RSGPU_BAD_ARGUMENT,
                          "rsgpu_flux_rates_noalloc: " + std::to_string(ncells_sel) +
                              " cells exceed its limit of " +
                              std::to_string(RS_RATES_NOALLOC_MAX) +
                              " (use rsgpu_flux_rates)"};
        for (int t = 0; t < ncells_sel; ++t)
            if (cells_host[t] < 0 || cells_host[t] >= F->ncells)
                throw RsError{RSGPU_BAD_ARGUMENT, "cell index out of range"};
        if (!ncells_sel)
            return RSGPU_OK;
        const int dev = F->ctx->device;
        void *base = nullptr;
        if (dev >= 0 && dev < RS_SCRATCH_DEVICES) {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            base = g_scratch_ptr[dev];
        }
        if (!base)
            throw RsError{RSGPU_HIP_ERROR, "no-allocation scratch not set up on this device"};
        double *rates = static_cast<double *>(base);
        int *cells = reinterpret_cast<int *>(static_cast<char *>(base) +
                                             sizeof(double) * RS_RATES_NOALLOC_MAX);
        hipStream_t s = F->ctx->stream;
        RS_HIP(hipMemcpyAsync(cells, cells_host, sizeof(int) * ncells_sel,
                              hipMemcpyHostToDevice, s));
        // every thread t < ncells_sel writes rates[t]: nothing to clear
        hipLaunchKernelGGL(k_rates, dim3((ncells_sel + 255) / 256), dim3(256), 0, s, ncells_sel,
                           cells, F->cell_ptr.p, F->cell_conn.p, F->cell_sign.p, flux_dev,
                           rates);
        RS_CHECK_LAUNCH();
        RS_HIP(hipMemcpyAsync(rates_host, rates, sizeof(double) * ncells_sel,
                              hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        return RSGPU_OK;
    } catch (const RsError &e) {
        return flux_fail(F->ctx, e);
    }
}

extern "C" int rsgpu_flux_vjp(rsgpu_flux *F, const double *p_dev, const double *w_dev,
                              double *tbar_dev)
{
    if (!F || (F->nconn && (!p_dev || !w_dev || !tbar_dev)))
        return RSGPU_BAD_ARGUMENT;
    try {
        if (F->nconn) {
            hipLaunchKernelGGL(k_flux_vjp, dim3((F->nconn + 255) / 256), dim3(256), 0,
                               F->ctx->stream, F->nconn, F->md_ptr.p, F->md_elem.p, F->here.p,
                               F->there.p, p_dev, w_dev, tbar_dev);
            RS_CHECK_LAUNCH();
        }
        return RSGPU_OK;
    } catch (const RsError &e) {
        return flux_fail(F->ctx, e);
    }
}
