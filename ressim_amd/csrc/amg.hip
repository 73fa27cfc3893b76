// Smoothed-aggregation AMG preconditioner on gfx950.
//
// Reference: mtj/src/no/uib/cipr/matrix/sparse/AMG.java (setMatrix :228-294,
// cycle :302-342, Aggregator :376-642, Interpolator :648-928), SSOR.java
// (apply :143-194), CompColMatrix.multAdd/transMult (:277-318), DenseLU
// (dgetrf/dgetrs with one right-hand side).
//
// The whole setup runs on the device, level by level:
//   strength N(i) = {j : |a_ij| >= eps sqrt(a_ii a_jj)}, eps = 0.08 * 0.5^k;
//   aggregation in three phases.  Phase 1 (disjoint strongly coupled
//   neighbourhoods) and phase 3 (remnants) are greedy loops in row order in
//   the reference; here they are prioritised rounds: a candidate wins when
//   its hashed priority is the largest on every node of its neighbourhood,
//   so winners are disjoint and the result is maximal, as the greedy loop's.
//   Phase 2 (each free node joins the aggregate it overlaps most, the first
//   to reach the largest overlap in HashSet iteration order) is independent
//   per node and follows the reference exactly;
//   the smoothed prolongator P = (I - omega D^-1 A_F) P_tent row by row with
//   the reference's per-entry operation order;
//   P' by a stable radix sort (rows stay ascending in every column);
//   the Galerkin operator as (P' (A P)) with one wavefront per output row
//   and an LDS hash accumulator.  The wavefront walks the left factor's row
//   in order and spreads the matching right row over its lanes, so every
//   entry receives its terms in the order of createGalerkinSlow
//   (AMG.java:883-919) and only nonzero results are stored;
//   the coarsest level is factored by a dense LU (dgetf2 order) in one
//   workgroup.
//
// The smoother is SSOR (forward omega_F, reverse omega_R).  Rows are visited
// in natural order (SSOR.java) or by greedy colour (multicolour SSOR, the
// default: rows of one colour are independent); either way the rows form
// levels whose members do not couple, and a coupling reads this sweep's
// value when its row was visited earlier, the previous value otherwise.
// Optionally the rows are split into contiguous blocks (block-Jacobi:
// couplings that leave a block read the vector from the start of the
// application).  A block that fits LDS is swept by one workgroup with its
// vector in LDS; a single block larger than that is swept in place in
// global memory with one launch per level (colour).
//
// AMG.apply starts the cycle from x's old content (AMG.java:208-216); the
// zero_guess option starts from zero instead, which makes the
// preconditioner a fixed linear operator.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <type_traits>

#include "spmv_core.hpp"

#define RS_AMG_LDS_ROWS 8190

struct Skip {
    const int *a, *b;
    __device__ __forceinline__ bool on() const { return (a && *a) || (b && *b); }
};

// One level (colour) of a smoother sweep: its rows, their first group and
// their first row slot.
struct AmgLevelMeta {
    int count, gbase, roff, pad;
};

// one level of the hierarchy
struct AmgLevel {
    int n = 0, nnz = 0, c = 0;
    // matrix (level 0 borrows the pattern of the caller's matrix)
    DevBuf<int> rp_own, ci_own;
    const int *rp = nullptr, *ci = nullptr;
    DevBuf<double> v;
    DevBuf<int> diag;
    std::vector<int> h_rp, h_ci;
    SpmvPlan plan;
    // aggregation / transfer
    DevBuf<int> pt;
    int pnnz = 0;
    DevBuf<int> Prp, Pci, Trp, Tci;
    DevBuf<double> Pv, Tv;
    SpmvPlan Pplan, Tplan;
    // smoother, SELL-64: per level (colour) the rows sorted by descending
    // off-diagonal count in groups of 64; entry s of lane t of group g at
    // gofs[g] + 64 s + t, in the row's CSR order
    int nb = 1, max_block_rows = 0, nlev_max = 0;
    DevBuf<int> bstart, blk_lev, meta, rows, rlen, row_glob, gofs, ell_col, ell_src;
    DevBuf<double> ell_val, dgl;
    size_t ell_size = 0;
    int nrows_lev = 0;
    uint64_t sched_pattern = 0;
    int nb_rows_param = -1, ordering_param = -1;
    bool global_sweep = false;          // one block, swept level by level
    std::vector<AmgLevelMeta> h_meta;   // host copy (launch sizes)
    // vectors
    DevBuf<double> u, ualt, f, r;
    // FP32 cycle: the level's matrix, transfer and smoother values and vectors
    DevBuf<float> v32, Pv32, Tv32, ell_val32, dgl32, u32, ualt32, f32, r32;
};

struct AmgData {
    rsgpu_amg_params p;
    std::vector<std::unique_ptr<AmgLevel>> L;   // m levels; the last is dense
    int m = 0;
    int nc = 0;
    DevBuf<double> lu;
    DevBuf<int> ipiv, flag;
    DevBuf<double> cinv;       // coarsest inverse (column-major), above 256 rows
    DevBuf<double> cpart;      // its column slices' partial products
    DevBuf<double> cf, cu;     // FP64 coarsest vectors of the FP32 cycle
    bool fp32 = false;         // the cycle runs in FP32 (coarsest solve in FP64)
    DevBuf<char> tmp_bytes;    // hipcub temporary storage
    uint64_t pattern0 = 0;
};

void amg_destroy(AmgData *d) { delete d; }

extern "C" void rsgpu_amg_params_default(rsgpu_amg_params *p)
{
    if (!p)
        return;
    // AMG() (AMG.java:204-206)
    p->omega_pre_f = 1;
    p->omega_pre_r = 1.85;
    p->omega_post_f = 1.85;
    p->omega_post_r = 1;
    p->nu1 = 1;
    p->nu2 = 1;
    p->gamma = 1;
    p->min = 40;
    p->omega = 2. / 3;
    p->reverse = 1;
    p->block_rows = 0;
    p->zero_guess = 0;
    p->max_coarse = 2048;
    p->ordering = 1;
    p->coarsen_ratio = 0;
}

// ---------------------------------------------------------------- helpers
template <typename T>
static void swap_buf(DevBuf<T> &a, DevBuf<T> &b)
{
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

static void scan_exclusive(AmgData *D, hipStream_t s, const int *in, int *out, int n)
{
    size_t bytes = 0;
    RS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
    if (D->tmp_bytes.n < bytes)
        D->tmp_bytes.alloc(bytes);
    RS_HIP(hipcub::DeviceScan::ExclusiveSum(D->tmp_bytes.p, bytes, in, out, n, s));
}

static int read_int(hipStream_t s, const int *p)
{
    int v = 0;
    RS_HIP(hipMemcpyAsync(&v, p, sizeof(int), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    return v;
}

static void set_int(hipStream_t s, int *p, int v)
{
    RS_HIP(hipMemcpyAsync(p, &v, sizeof(int), hipMemcpyHostToDevice, s));
}

static inline int gridn(long long n, int bs = 256)
{
    return (int)std::max<long long>(1, (n + bs - 1) / bs);
}

__device__ __forceinline__ unsigned long long amg_prio(int i)
{
    // MurmurHash3 32-bit finaliser of the index, index in the low word
    unsigned h = (unsigned)i;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return ((unsigned long long)h << 32) | (unsigned)i;
}

// Java 8 HashSet<Integer> iteration key: (bucket, value)
__device__ __forceinline__ unsigned long long hs_key(int v, int cap)
{
    unsigned h = (unsigned)v;
    return ((unsigned long long)((h ^ (h >> 16)) & (unsigned)(cap - 1)) << 32) | h;
}

__device__ __forceinline__ int hs_cap(int size)
{
    int cap = 16;
    while (size > (cap * 3) / 4)
        cap <<= 1;
    return cap;
}

__device__ __forceinline__ int bsearch_dev(const int *a, int key, int lo, int hi)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        int v = a[mid];
        if (v == key)
            return mid;
        if (v < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return -1;
}

// ---------------------------------------------------------------- strength
__global__ void k_amg_diag(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                           int *__restrict__ diag, int *bad)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int d = bsearch_dev(ci, i, rp[i], rp[i + 1]);
    diag[i] = d;
    if (d < 0)
        atomicMin(bad, i + 1);
}

// N(i) count (or fill) and the initial R set (AMG.java:482-532)
__global__ void k_amg_strength(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                               const double *__restrict__ v, const int *__restrict__ diag,
                               double eps, int *__restrict__ cnt, const int *__restrict__ Np,
                               int *__restrict__ Ni, char *__restrict__ R)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double aii = v[diag[i]];
    int q = Np ? Np[i] : 0, c = 0;
    char h = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int col = ci[j];
        const double aij = v[j];
        if (col != i && aij != 0)
            h = 1;
        const double ajj = v[diag[col]];
        if (fabs(aij) >= eps * sqrt(aii * ajj)) {
            if (Ni)
                Ni[q + c] = col;
            ++c;
        }
    }
    if (cnt)
        cnt[i] = c;
    if (R)
        R[i] = h;
}

// ---------------------------------------------------------------- phase 1
__global__ void k_p1_cand(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                          const char *__restrict__ R, char *__restrict__ cand,
                          unsigned long long *__restrict__ m, int *any)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = false;
    if (i < n) {
        ok = R[i] != 0;
        for (int q = Np[i]; q < Np[i + 1] && ok; ++q)
            ok = R[Ni[q]] != 0;
        cand[i] = ok;
        m[i] = 0;
    }
    if (__ballot(ok) && (threadIdx.x & 63) == 0)
        atomicOr(any, 1);
}

__global__ void k_p1_max(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                         const char *__restrict__ cand, unsigned long long *m)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !cand[i])
        return;
    const unsigned long long p = amg_prio(i);
    for (int q = Np[i]; q < Np[i + 1]; ++q)
        atomicMax(&m[Ni[q]], p);
}

__global__ void k_p1_sel(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                         const char *__restrict__ cand, const unsigned long long *__restrict__ m,
                         char *R, int *root_of)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !cand[i])
        return;
    const unsigned long long p = amg_prio(i);
    for (int q = Np[i]; q < Np[i + 1]; ++q)
        if (m[Ni[q]] != p)
            return;
    for (int q = Np[i]; q < Np[i + 1]; ++q) {
        R[Ni[q]] = 0;
        root_of[Ni[q]] = i;
    }
}

__global__ void k_root_flag(int n, const int *__restrict__ root_of, int *__restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flag[i] = root_of[i] == i;
}

__global__ void k_owner(int n, const int *__restrict__ root_of, const int *__restrict__ id,
                        int *__restrict__ owner, int *__restrict__ pt)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    int r = root_of[j];
    int o = r >= 0 ? id[r] : -1;
    owner[j] = o;
    pt[j] = o;
}

// ---------------------------------------------------------------- phase 2
// AMG.java:584-611 for one free node: the aggregate whose overlap count
// first reaches the maximum while N(i) is iterated in HashSet order.
__global__ void k_enlarge(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                          const int *__restrict__ owner, char *R, int *pt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !R[i])
        return;
    const int *N = Ni + Np[i];
    const int len = Np[i + 1] - Np[i];
    const int cap = hs_cap(len);
    int maxv = 0;
    for (int a = 0; a < len; ++a) {
        const int k = owner[N[a]];
        if (k < 0)
            continue;
        int c = 0;
        for (int b = 0; b < len; ++b)
            c += owner[N[b]] == k;
        maxv = c > maxv ? c : maxv;
    }
    if (maxv == 0)
        return;
    int best = 0x7fffffff, win = -1;
    for (int a = 0; a < len; ++a) {
        const int k = owner[N[a]];
        if (k < 0)
            continue;
        const unsigned long long ka = hs_key(N[a], cap);
        int c = 0, occ = 0, rank = 0;
        for (int b = 0; b < len; ++b) {
            const bool lt = hs_key(N[b], cap) < ka;
            const bool same = owner[N[b]] == k;
            c += same;
            occ += same && lt;
            rank += lt;
        }
        if (c == maxv && occ == maxv - 1 && rank < best) {
            best = rank;
            win = k;
        }
    }
    pt[i] = win;
    R[i] = 0;
}

// ---------------------------------------------------------------- phase 3
__global__ void k_p3_cand(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                          const char *__restrict__ R, char *__restrict__ cand,
                          unsigned long long *__restrict__ m, int *any)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = false;
    if (i < n) {
        if (R[i])
            for (int q = Np[i]; q < Np[i + 1] && !ok; ++q)
                ok = R[Ni[q]] != 0;
        cand[i] = ok;
        m[i] = 0;
    }
    if (__ballot(ok) && (threadIdx.x & 63) == 0)
        atomicOr(any, 1);
}

__global__ void k_p3_max(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                         const char *__restrict__ R, const char *__restrict__ cand,
                         unsigned long long *m)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !cand[i])
        return;
    const unsigned long long p = amg_prio(i);
    for (int q = Np[i]; q < Np[i + 1]; ++q)
        if (R[Ni[q]])
            atomicMax(&m[Ni[q]], p);
}

__global__ void k_p3_sel(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                         const char *__restrict__ R, const char *__restrict__ cand,
                         const unsigned long long *__restrict__ m, int *__restrict__ sel)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int s = 0;
    if (cand[i]) {
        const unsigned long long p = amg_prio(i);
        s = 1;
        for (int q = Np[i]; q < Np[i + 1]; ++q)
            if (R[Ni[q]] && m[Ni[q]] != p) {
                s = 0;
                break;
            }
    }
    sel[i] = s;
}

__global__ void k_p3_assign(int n, const int *__restrict__ Np, const int *__restrict__ Ni,
                            const char *__restrict__ R, const int *__restrict__ sel,
                            const int *__restrict__ id, int base, int *pt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !sel[i])
        return;
    const int a = base + id[i];
    for (int q = Np[i]; q < Np[i + 1]; ++q)
        if (R[Ni[q]])
            pt[Ni[q]] = a;
}

__global__ void k_p3_clear(int n, const int *__restrict__ pt, char *R)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && R[j] && pt[j] >= 0)
        R[j] = 0;
}

// ---------------------------------------------------------------- prolongator
// Row i of P = (I - omega D^-1 A_F) P_tent (AMG.java:830-873): aggregates in
// ascending order; each entry accumulates its couplings in column order.
__device__ __forceinline__ bool amg_weak(double a, int col, const int *Ni, int n0, int n1)
{
    return a != 0 && bsearch_dev(Ni, col, n0, n1) < 0;
}

__global__ void k_prolong(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                          const double *__restrict__ v, const int *__restrict__ diag,
                          const int *__restrict__ Np, const int *__restrict__ Ni,
                          const int *__restrict__ pt, double omega, int *__restrict__ cnt,
                          const int *__restrict__ Prp, int *__restrict__ Pci,
                          double *__restrict__ Pv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int own = pt[i];
    if (own == -1) {
        if (cnt)
            cnt[i] = 0;
        return;
    }
    const int r0 = rp[i], r1 = rp[i + 1], n0 = Np[i], n1 = Np[i + 1];
    double weak = 0;
    for (int j = r0; j < r1; ++j) {
        const int col = ci[j];
        if (pt[col] == -1)
            continue;
        if (amg_weak(v[j], col, Ni, n0, n1))
            weak += v[j];
    }
    const double scale = -omega / v[diag[i]];
    int prev = -1, q = 0;
    const int base = Prp ? Prp[i] : 0;
    for (;;) {
        int next = 0x7fffffff;
        for (int j = r0; j < r1; ++j) {
            const int k = pt[ci[j]];
            if (k <= prev || k >= next)
                continue;
            if (amg_weak(v[j], ci[j], Ni, n0, n1))
                continue;
            next = k;
        }
        if (own > prev && own < next)
            next = own;
        if (next == 0x7fffffff)
            break;
        double acc = 0;
        for (int j = r0; j < r1; ++j) {
            if (pt[ci[j]] != next)
                continue;
            if (amg_weak(v[j], ci[j], Ni, n0, n1))
                continue;
            acc += v[j];
        }
        if (next == own)
            acc -= weak;
        acc *= scale;
        if (next == own)
            acc += 1;
        if (acc != 0) {
            if (Pci) {
                Pci[base + q] = next;
                Pv[base + q] = acc;
            }
            ++q;
        }
        prev = next;
    }
    if (cnt)
        cnt[i] = q;
}

__global__ void k_entry_rows(int n, const int *__restrict__ rp, int *__restrict__ erow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
        erow[e] = i;
}

__global__ void k_iota(int n, int *a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = i;
}

__global__ void k_col_count(int nnz, const int *__restrict__ ci, int *cnt)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz)
        atomicAdd(&cnt[ci[e]], 1);
}

__global__ void k_transpose_fill(int nnz, const int *__restrict__ perm,
                                 const int *__restrict__ erow, const double *__restrict__ v,
                                 int *__restrict__ tci, double *__restrict__ tv)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) {
        const int e = perm[p];
        tci[p] = erow[e];
        tv[p] = v[e];
    }
}

// ---------------------------------------------------------------- SpGEMM
// Upper bound of a product row's length: sum of the right rows it touches.
__global__ void k_spgemm_ub(int nrows, const int *__restrict__ Xrp, const int *__restrict__ Xci,
                            const int *__restrict__ Yrp, int *ubmax)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int u = 0;
    if (i < nrows)
        for (int e = Xrp[i]; e < Xrp[i + 1]; ++e) {
            const int j = Xci[e];
            u += Yrp[j + 1] - Yrp[j];
        }
    for (int o = 32; o > 0; o >>= 1)
        u = max(u, __shfl_xor(u, o, 64));
    if ((threadIdx.x & 63) == 0)
        atomicMax(ubmax, u);
}

__device__ __forceinline__ unsigned hslot(int k, int tab)
{
    return ((unsigned)k * 2654435761u) & (unsigned)(tab - 1);
}

// C = X Y, one wavefront per row of X.  The wave walks X's row entries in
// order; lane t handles the t-th entry of the matching Y row (distinct
// columns), so each output entry receives its products in X-row order.
// count pass: cnt[row] = nonzero entries; fill pass: sorted entries at Crp.
template <int TAB>
__global__ void __launch_bounds__(64)
    k_spgemm(int nrows, const int *__restrict__ Xrp, const int *__restrict__ Xci,
             const double *__restrict__ Xv, const int *__restrict__ Yrp,
             const int *__restrict__ Yci, const double *__restrict__ Yv, int *__restrict__ cnt,
             const int *__restrict__ Crp, int *__restrict__ Cci, double *__restrict__ Cv,
             int *err)
{
    __shared__ int keys[TAB];
    __shared__ double vals[TAB];
    __shared__ int ck[TAB];
    __shared__ double cv[TAB];
    const int row = blockIdx.x;
    const int lane = threadIdx.x;
    if (row >= nrows)
        return;
    for (int s = lane; s < TAB; s += 64) {
        keys[s] = -1;
        vals[s] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __syncthreads();
    bool overflow = false;
    for (int e = Xrp[row]; e < Xrp[row + 1]; ++e) {
        const int j = Xci[e];
        const double xv = Xv[e];
        for (int t = Yrp[j] + lane; t < Yrp[j + 1]; t += 64) {
            const int k = Yci[t];
            unsigned s = hslot(k, TAB);
            int probes = 0;
            for (;;) {
                const int old = atomicCAS(&keys[s], -1, k);
                if (old == -1 || old == k)
                    break;
                s = (s + 1) & (TAB - 1);
                if (++probes >= TAB) {
                    overflow = true;
                    break;
                }
            }
            if (!overflow)
                vals[s] += xv * Yv[t];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __syncthreads();
    }
    if (__ballot(overflow)) {
        if (lane == 0)
            atomicOr(err, 1);
        return;
    }
    // compact nonzero entries
    int mine = 0;
    for (int s = lane; s < TAB; s += 64)
        mine += keys[s] != -1 && vals[s] != 0;
    int incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += y;
    }
    const int total = __shfl(incl, 63, 64);
    if (cnt) {
        if (lane == 0)
            cnt[row] = total;
        return;
    }
    int w = incl - mine;
    for (int s = lane; s < TAB; s += 64)
        if (keys[s] != -1 && vals[s] != 0) {
            ck[w] = keys[s];
            cv[w] = vals[s];
            ++w;
        }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __syncthreads();
    const int base = Crp[row];
    for (int a = lane; a < total; a += 64) {
        const int k = ck[a];
        int rank = 0;
        for (int b = 0; b < total; ++b)
            rank += ck[b] < k;
        Cci[base + rank] = k;
        Cv[base + rank] = cv[a];
    }
}

// ---------------------------------------------------------------- dense coarsest
__global__ void k_dense_fill(int nc, const int *__restrict__ rp, const int *__restrict__ ci,
                             const double *__restrict__ v, double *__restrict__ a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    for (int c = 0; c < nc; ++c)
        a[i + (size_t)c * nc] = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j)
        a[i + (size_t)ci[j] * nc] = v[j];
}

// LAPACK dgetf2 (right-looking, partial pivoting) in one workgroup
__global__ void __launch_bounds__(256) k_dense_getf2(int n, double *a, int *ipiv, int *info)
{
    __shared__ int sjp;
    __shared__ int sinfo;
    if (threadIdx.x == 0)
        sinfo = 0;
    for (int j = 0; j < n; ++j) {
        __syncthreads();
        if (threadIdx.x == 0) {
            int jp = j;
            double amax = fabs(a[j + (size_t)j * n]);
            for (int i = j + 1; i < n; ++i)
                if (fabs(a[i + (size_t)j * n]) > amax) {
                    amax = fabs(a[i + (size_t)j * n]);
                    jp = i;
                }
            ipiv[j] = jp;
            sjp = jp;
            if (a[jp + (size_t)j * n] == 0 && sinfo == 0)
                sinfo = j + 1;
        }
        __syncthreads();
        const int jp = sjp;
        const bool nz = a[jp + (size_t)j * n] != 0;
        __syncthreads();
        if (nz) {
            if (jp != j)
                for (int c = threadIdx.x; c < n; c += blockDim.x) {
                    double t = a[j + (size_t)c * n];
                    a[j + (size_t)c * n] = a[jp + (size_t)c * n];
                    a[jp + (size_t)c * n] = t;
                }
            __syncthreads();
            if (j < n - 1) {
                const double r = 1.0 / a[j + (size_t)j * n];
                for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x)
                    a[i + (size_t)j * n] *= r;
            }
        }
        __syncthreads();
        if (j < n - 1) {
            const int w = n - 1 - j;
            for (int q = threadIdx.x; q < w * w; q += blockDim.x) {
                const int i = j + 1 + q % w, c = j + 1 + q / w;
                const double yj = a[j + (size_t)c * n];
                if (yj != 0) {
                    const double temp = -1.0 * yj;
                    a[i + (size_t)c * n] = a[i + (size_t)c * n] + a[i + (size_t)j * n] * temp;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *info = sinfo;
}

// u = f, then dgetrs ('N' or 'T') with one right-hand side
// dgetf2 column j for large matrices, as launches: the pivot (first row of
// largest |a(i,j)|, i >= j, as idamax) and row swap, the column scaling, and
// the trailing rank-1 update over many workgroups; same operations as the
// one-workgroup kernel above
__global__ void __launch_bounds__(1024) k_getf2_pivot(int n, int j, double *a, int *ipiv,
                                                     int *info)
{
    __shared__ double sv[1024];
    __shared__ int si[1024];
    const int t = threadIdx.x;
    double bv = -1.0;
    int bi = n;
    for (int i = j + 1 + t; i < n; i += blockDim.x) {
        const double v = fabs(a[i + (size_t)j * n]);
        if (v > bv) {   // first occurrence of the largest in this thread's rows
            bv = v;
            bi = i;
        }
    }
    sv[t] = bv;
    si[t] = bi;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            const double v2 = sv[t + w];
            const int i2 = si[t + w];
            if (v2 > sv[t] || (v2 == sv[t] && i2 < si[t])) {
                sv[t] = v2;
                si[t] = i2;
            }
        }
        __syncthreads();
    }
    // the sequential scan starts from row j and moves only on a larger value
    const double vj = fabs(a[j + (size_t)j * n]);
    const int jp = si[0] < n && sv[0] > vj ? si[0] : j;
    __syncthreads();
    if (t == 0) {
        ipiv[j] = jp;
        if (a[jp + (size_t)j * n] == 0 && *info == 0)
            *info = j + 1;
    }
    const bool nz = a[jp + (size_t)j * n] != 0;
    __syncthreads();
    if (!nz)
        return;
    if (jp != j)
        for (int c = t; c < n; c += blockDim.x) {
            const double x = a[j + (size_t)c * n];
            a[j + (size_t)c * n] = a[jp + (size_t)c * n];
            a[jp + (size_t)c * n] = x;
        }
    __syncthreads();
    if (j < n - 1) {
        const double r = 1.0 / a[j + (size_t)j * n];
        for (int i = j + 1 + t; i < n; i += blockDim.x)
            a[i + (size_t)j * n] *= r;
    }
}

__global__ void __launch_bounds__(256) k_getf2_update(int n, int j, double *a)
{
    const int w = n - 1 - j;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)w * w)
        return;
    const int i = j + 1 + (int)(q % w), c = j + 1 + (int)(q / w);
    const double yj = a[j + (size_t)c * n];
    if (yj != 0) {
        const double temp = -1.0 * yj;
        a[i + (size_t)c * n] = a[i + (size_t)c * n] + a[i + (size_t)j * n] * temp;
    }
}

// dgetrs with one right-hand side b (in place) in one workgroup
__device__ void dense_getrs_body(int n, const double *__restrict__ a,
                                 const int *__restrict__ ipiv, double *b, int trans)
{
    __shared__ int flag;
    if (!trans) {
        if (threadIdx.x == 0)
            for (int j = 0; j < n; ++j)
                if (ipiv[j] != j) {
                    double t = b[j];
                    b[j] = b[ipiv[j]];
                    b[ipiv[j]] = t;
                }
        __syncthreads();
        for (int k = 0; k < n; ++k) {
            const double bk = b[k];
            __syncthreads();
            if (bk != 0)
                for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x)
                    b[i] = b[i] - bk * a[i + (size_t)k * n];
            __syncthreads();
        }
        for (int k = n - 1; k >= 0; --k) {
            if (threadIdx.x == 0) {
                flag = b[k] != 0;
                if (flag)
                    b[k] = b[k] / a[k + (size_t)k * n];
            }
            __syncthreads();
            if (flag) {
                const double bk = b[k];
                for (int i = threadIdx.x; i < k; i += blockDim.x)
                    b[i] = b[i] - bk * a[i + (size_t)k * n];
            }
            __syncthreads();
        }
    } else if (threadIdx.x == 0) {
        for (int i = 0; i < n; ++i) {
            double t = b[i];
            for (int k = 0; k < i; ++k)
                t -= a[k + (size_t)i * n] * b[k];
            b[i] = t / a[i + (size_t)i * n];
        }
        for (int i = n - 1; i >= 0; --i) {
            double t = b[i];
            for (int k = i + 1; k < n; ++k)
                t -= a[k + (size_t)i * n] * b[k];
            b[i] = t;
        }
        for (int j = n - 1; j >= 0; --j)
            if (ipiv[j] != j) {
                double t = b[j];
                b[j] = b[ipiv[j]];
                b[ipiv[j]] = t;
            }
    }
}

__global__ void __launch_bounds__(256)
    k_dense_getrs(int n, const double *__restrict__ a, const int *__restrict__ ipiv,
                  const double *__restrict__ f, double *b, int trans, Skip sk)
{
    if (sk.on())
        return;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        b[i] = f[i];
    __syncthreads();
    dense_getrs_body(n, a, ipiv, b, trans);
}

// column j = blockIdx.x of the inverse: the LU solve of e_j
__global__ void __launch_bounds__(256)
    k_dense_inv(int n, const double *__restrict__ a, const int *__restrict__ ipiv, double *inv)
{
    double *b = inv + (size_t)blockIdx.x * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        b[i] = i == (int)blockIdx.x ? 1.0 : 0.0;
    __syncthreads();
    dense_getrs_body(n, a, ipiv, b, 0);
}

// dgetrs 'N' for n <= 64 in one wavefront: lane i holds b[i]; the row swaps,
// the unit-lower and the upper column sweeps run in dgetrs's order with the
// pivot values broadcast across lanes.
__global__ void __launch_bounds__(64)
    k_dense_getrs64(int n, const double *__restrict__ a, const int *__restrict__ ipiv,
                    const double *__restrict__ f, double *b, Skip sk)
{
    if (sk.on())
        return;
    const int lane = threadIdx.x;
    double v = lane < n ? f[lane] : 0.0;
    for (int j = 0; j < n; ++j) {
        const int p = ipiv[j];
        if (p != j) {
            const double vj = __shfl(v, j, 64), vp = __shfl(v, p, 64);
            if (lane == j)
                v = vp;
            if (lane == p)
                v = vj;
        }
    }
    for (int k = 0; k < n; ++k) {
        const double bk = __shfl(v, k, 64);
        if (bk != 0 && lane > k && lane < n)
            v = v - bk * a[lane + (size_t)k * n];
    }
    for (int k = n - 1; k >= 0; --k) {
        double bk = __shfl(v, k, 64);
        if (bk != 0) {
            bk = bk / a[k + (size_t)k * n];
            if (lane == k)
                v = bk;
            if (lane < k)
                v = v - bk * a[lane + (size_t)k * n];
        }
    }
    if (lane < n)
        b[lane] = v;
}

// ---------------------------------------------------------------- smoother
template <typename T>
__global__ void k_amg_gather(const int *__restrict__ src, const double *__restrict__ v,
                             T *__restrict__ ell, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int s = src[i];
        ell[i] = s >= 0 ? (T)v[s] : (T)-0.0;
    }
}

template <typename T>
__global__ void k_amg_rowdiag(const int *__restrict__ row_glob, const int *__restrict__ diag,
                              const double *__restrict__ v, T *__restrict__ dgl, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dgl[i] = (T)v[diag[row_glob[i]]];
}

// One group of one level of an SSOR sweep (SSOR.java:143-194): lane t owns
// row t of group g and forms its sum over the row's off-diagonal entries in
// CSR order, then x_r += omega ((f_r - sum) / a_rr - x_r).  BLOCK: the
// vector of the block is in LDS; an entry column < 0 encodes the global
// column -1-c, read from xold (the vector at the start of the application).
template <typename T, bool BLOCK>
__device__ __forceinline__ void amg_ssor_group(const AmgLevelMeta &m, int g,
                                               const int *__restrict__ rows,
                                               const int *__restrict__ rlen,
                                               const int *__restrict__ gofs,
                                               const int *__restrict__ ecol,
                                               const T *__restrict__ eval,
                                               const T *__restrict__ dgl,
                                               const T *__restrict__ f, int b0,
                                               const T *__restrict__ xold, T omega, T *xv)
{
    const int lane = threadIdx.x & 63;
    const int i = (g << 6) + lane;
    if (i >= m.count)
        return;
    const int r = rows[m.roff + i], len = rlen[m.roff + i];
    const int e0 = gofs[m.gbase + g] + lane;
    // eight entries' loads in flight at a time; the sum stays in entry order
    T sigma = 0;
    int q = 0;
    for (; q + 8 <= len; q += 8) {
        int c[8];
        T v[8], xc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            c[u] = ecol[e0 + 64 * (q + u)];
            v[u] = eval[e0 + 64 * (q + u)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            xc[u] = (!BLOCK || c[u] >= 0) ? xv[c[u]] : xold[-1 - c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            sigma += v[u] * xc[u];
    }
    for (; q < len; ++q) {
        const int c = ecol[e0 + 64 * q];
        const T xc = (!BLOCK || c >= 0) ? xv[c] : xold[-1 - c];
        sigma += eval[e0 + 64 * q] * xc;
    }
    sigma = (f[b0 + r] - sigma) / dgl[m.roff + i];
    const T xr = xv[r];
    xv[r] = xr + omega * (sigma - xr);
}

// The same update with eight lanes per row (rows with many entries): lane k
// of the row's unit forms the products of entries k, k + 8, ... of a chunk
// of RS_ROW8_CHUNK into LDS, and the unit's lane 0 adds them in entry order.
// i: the row's slot in the level; P: the unit's RS_ROW8_CHUNK LDS slots.
#define RS_ROW8_CHUNK 128
template <typename T, bool BLOCK>
__device__ __forceinline__ void amg_ssor_row8(const AmgLevelMeta &m, int i, bool live,
                                              const int *__restrict__ rows,
                                              const int *__restrict__ rlen,
                                              const int *__restrict__ gofs,
                                              const int *__restrict__ ecol,
                                              const T *__restrict__ eval,
                                              const T *__restrict__ dgl,
                                              const T *__restrict__ f, int b0,
                                              const T *__restrict__ xold, T omega, T *xv, T *P)
{
    const int k = threadIdx.x & 7;
    const int r = live ? rows[m.roff + i] : 0;
    const int len = live ? rlen[m.roff + i] : 0;
    const int e0 = live ? gofs[m.gbase + (i >> 6)] + (i & 63) : 0;
    // the longest row among the wave's eight units bounds the chunk loop
    int lmax = len;
    for (int o = 8; o < 64; o <<= 1)
        lmax = max(lmax, __shfl_xor(lmax, o, 64));
    T sigma = 0;
    constexpr int PU = RS_ROW8_CHUNK / 8;
    for (int s0 = 0; s0 < lmax; s0 += RS_ROW8_CHUNK) {
        T p[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int sq = s0 + k + 8 * u;
            int c = 0;
            T v = T(-0.0);
            if (sq < len) {
                c = ecol[e0 + 64 * sq];
                v = eval[e0 + 64 * sq];
            }
            const T xc = sq >= len ? T(0) : (!BLOCK || c >= 0) ? xv[c] : xold[-1 - c];
            p[u] = v * xc;
        }
#pragma unroll
        for (int u = 0; u < PU; ++u)
            P[k + 8 * u] = p[u];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (k == 0) {
            // slots past the row's end hold -0.0 * 0 = -0.0, an exact
            // identity: add in blocks of eight with the reads issued first
            const int nq = (min(RS_ROW8_CHUNK, len - s0) + 7) & ~7;
            for (int q = 0; q < nq; q += 8) {
                T t[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    t[u] = P[q + u];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    sigma += t[u];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (k == 0 && live) {
        sigma = (f[b0 + r] - sigma) / dgl[m.roff + i];
        const T xr = xv[r];
        xv[r] = xr + omega * (sigma - xr);
    }
}

template <typename T>
__device__ __forceinline__ void amg_ssor_sweep(int lb, int le, int step,
                                               const AmgLevelMeta *__restrict__ meta,
                                               const int *__restrict__ rows,
                                               const int *__restrict__ rlen,
                                               const int *__restrict__ gofs,
                                               const int *__restrict__ ecol,
                                               const T *__restrict__ eval,
                                               const T *__restrict__ dgl,
                                               const T *__restrict__ f, int b0,
                                               const T *__restrict__ xold, T omega, T *xv)
{
    __shared__ T P8[32 * RS_ROW8_CHUNK];
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int l = lb; l != le; l += step) {
        const AmgLevelMeta m = meta[l];
        if (m.pad) {
            // eight lanes per row: blockDim / 8 rows at a time
            const int unit = threadIdx.x >> 3, nu = blockDim.x >> 3;
            const int nr = (m.count + nu - 1) / nu * nu;
            for (int i = unit; i < nr; i += nu)
                amg_ssor_row8<T, true>(m, i, i < m.count, rows, rlen, gofs, ecol, eval, dgl, f,
                                       b0, xold, omega, xv, P8 + RS_ROW8_CHUNK * unit);
        } else {
            const int ng = (m.count + 63) >> 6;
            for (int g = wave; g < ng; g += nw)
                amg_ssor_group<T, true>(m, g, rows, rlen, gofs, ecol, eval, dgl, f, b0, xold,
                                        omega, xv);
        }
        __syncthreads();
    }
}

// One SSOR application per block, the block's vector in LDS.
template <typename T>
__global__ void __launch_bounds__(256)
    k_amg_ssor(const int *__restrict__ bstart, const int *__restrict__ blk_lev,
               const AmgLevelMeta *__restrict__ meta, const int *__restrict__ rows,
               const int *__restrict__ rlen, const int *__restrict__ gofs,
               const int *__restrict__ ecol, const T *__restrict__ eval,
               const T *__restrict__ dgl, const T *__restrict__ f, const T *__restrict__ xold,
               T *__restrict__ xnew, T omega_f, T omega_r, int reverse, Skip sk)
{
    if (sk.on())
        return;
    __shared__ T xv[RS_AMG_LDS_ROWS + 1];
    const int blk = blockIdx.x;
    const int b0 = bstart[blk], nb = bstart[blk + 1] - b0;
    for (int r = threadIdx.x; r < nb; r += blockDim.x)
        xv[r] = xold[b0 + r];
    if (threadIdx.x == 0)
        xv[nb] = 0;
    __syncthreads();
    const int l0 = blk_lev[blk], l1 = blk_lev[blk + 1];
    amg_ssor_sweep<T>(l0, l1, 1, meta, rows, rlen, gofs, ecol, eval, dgl, f, b0, xold, omega_f,
                      xv);
    if (reverse && l1 > l0)
        amg_ssor_sweep<T>(l1 - 1, l0 - 1, -1, meta, rows, rlen, gofs, ecol, eval, dgl, f, b0,
                          xold, omega_r, xv);
    for (int r = threadIdx.x; r < nb; r += blockDim.x)
        xnew[b0 + r] = xv[r];
}

// One level (colour) of an in-place SSOR sweep over a whole matrix: the
// level's rows do not couple, so every row reads the current vector.  Four
// groups per workgroup.
template <typename T>
__global__ void __launch_bounds__(256)
    k_amg_ssor_level(AmgLevelMeta m, const int *__restrict__ rows, const int *__restrict__ rlen,
                     const int *__restrict__ gofs, const int *__restrict__ ecol,
                     const T *__restrict__ eval, const T *__restrict__ dgl,
                     const T *__restrict__ f, T *x, T omega, Skip sk)
{
    if (sk.on())
        return;
    if (m.pad) {
        __shared__ T P8[32 * RS_ROW8_CHUNK];
        const int unit = threadIdx.x >> 3;
        const int i = blockIdx.x * 32 + unit;
        // a wave leaves together: its eight units share the chunk loop
        if (__ballot(i < m.count) == 0)
            return;
        amg_ssor_row8<T, false>(m, i, i < m.count, rows, rlen, gofs, ecol, eval, dgl, f, 0, x,
                                omega, x, P8 + RS_ROW8_CHUNK * unit);
        return;
    }
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if ((g << 6) >= m.count)
        return;
    amg_ssor_group<T, false>(m, g, rows, rlen, gofs, ecol, eval, dgl, f, 0, x, omega, x);
}

// ---------------------------------------------------------------- cycle kernels
template <typename T, class Epi>
__global__ void __launch_bounds__(RS_SPMV_BS)
    k_amg_spmv(const int *__restrict__ wg_rows, const int *__restrict__ rowptr,
               const int *__restrict__ colind, const T *__restrict__ val,
               const T *__restrict__ x, Epi epi, Skip sk)
{
    if (sk.on())
        return;
    spmv_core<T, RS_SPMV_BS, RS_SPMV_NNZ>(wg_rows, rowptr, colind, val, x, epi);
}

// prolongation u += P uc with CompColMatrix.multAdd's per-row order
template <typename T>
struct EpiSeqAdd {
    T *y;
    __device__ T init(int r) const { return y[r]; }
    __device__ void row(int r, T dot) { y[r] = dot; }
};

template <typename T>
__global__ void k_amg_copy(const T *__restrict__ a, T *__restrict__ b, int n, Skip sk)
{
    if (sk.on())
        return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a ? a[i] : T(0);
}

template <typename S, typename D>
__global__ void k_amg_cvt(const S *__restrict__ a, D *__restrict__ b, size_t n, Skip sk)
{
    if (sk.on())
        return;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = (D)a[i];
}

// coarsest solve by the explicit inverse (column-major): u = C^-1 f as
// RS_CINV_SLICES column slices (partial sums per slice, coalesced column
// reads), then the slices added in order; C^-T f (transposed cycle): one
// wave per row of C^-T, i.e. per column of the array, lanes over its
// entries, a fixed-order lane reduction
#define RS_CINV_SLICES 64
__global__ void __launch_bounds__(256)
    k_amg_cinv_part(int n, const double *__restrict__ inv, const double *__restrict__ f,
                    double *__restrict__ part, Skip sk)
{
    if (sk.on())
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int sl = blockIdx.y;
    const int j0 = (int)(((long long)n * sl) / RS_CINV_SLICES);
    const int j1 = (int)(((long long)n * (sl + 1)) / RS_CINV_SLICES);
    double acc = 0;
    for (int j = j0; j < j1; ++j)
        acc += inv[i + (size_t)j * n] * f[j];
    part[(size_t)sl * n + i] = acc;
}

__global__ void __launch_bounds__(256)
    k_amg_cinv_sum(int n, const double *__restrict__ part, double *__restrict__ u, Skip sk)
{
    if (sk.on())
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double acc = 0;
    for (int sl = 0; sl < RS_CINV_SLICES; ++sl)
        acc += part[(size_t)sl * n + i];
    u[i] = acc;
}

__global__ void __launch_bounds__(256)
    k_amg_cinv_t(int n, const double *__restrict__ inv, const double *__restrict__ f,
                 double *__restrict__ u, Skip sk)
{
    if (sk.on())
        return;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n)
        return;
    double acc = 0;
    for (int j = lane; j < n; j += 64)
        acc += inv[j + (size_t)i * n] * f[j];
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_down(acc, o, 64);
    if (lane == 0)
        u[i] = acc;
}

// ---------------------------------------------------------------- host: level setup
static void plan_from_host(const std::vector<int> &rp, int n, SpmvPlan &plan, hipStream_t s)
{
    spmv_build_plan(rp, n, plan, s);
}

static void download(hipStream_t s, std::vector<int> &h, const int *d, size_t n)
{
    h.resize(n);
    if (n)
        RS_HIP(hipMemcpyAsync(h.data(), d, sizeof(int) * n, hipMemcpyDeviceToHost, s));
}

// Symmetrised in-block level schedule and full-row ELL for SSOR
// Greedy colouring in row order of the symmetrised in-block graph: a row
// takes the smallest colour none of its earlier in-block neighbours has.
static void greedy_colour(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                          const std::vector<int> &blk, std::vector<int> &colour)
{
    const int nnz = rp[n];
    std::vector<int> tp(n + 1, 0), ti(std::max(nnz, 1));
    for (int j = 0; j < nnz; ++j)
        tp[ci[j] + 1]++;
    for (int i = 0; i < n; ++i)
        tp[i + 1] += tp[i];
    std::vector<int> pos(tp.begin(), tp.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int j = rp[i]; j < rp[i + 1]; ++j)
            ti[pos[ci[j]]++] = i;
    std::vector<int> mark(n + 2, -1);
    colour.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            const int c = ci[j];
            if (c < i && blk[c] == blk[i])
                mark[colour[c]] = i;
        }
        for (int j = tp[i]; j < tp[i + 1]; ++j) {
            const int c = ti[j];
            if (c < i && blk[c] == blk[i])
                mark[colour[c]] = i;
        }
        int col = 0;
        while (mark[col] == i)
            ++col;
        colour[i] = col;
    }
}

// RSGPU_AMG_ROW8=0 keeps one lane per row on every level
static bool split_rows()
{
    const char *e = getenv("RSGPU_AMG_ROW8");
    return !(e && e[0] == '0');
}

static void build_ssor_schedule(AmgLevel &L, int block_rows, int ordering, hipStream_t s)
{
    const int n = L.n;
    const auto &rp = L.h_rp, &ci = L.h_ci;
    int nb = 1;
    if (block_rows > 0 && n > 0)
        nb = (n + block_rows - 1) / block_rows;
    std::vector<int> bstart(nb + 1);
    for (int b = 0; b <= nb; ++b)
        bstart[b] = (int)(((long long)n * b) / nb);
    int maxrows = 0;
    for (int b = 0; b < nb; ++b)
        maxrows = std::max(maxrows, bstart[b + 1] - bstart[b]);
    const bool global = maxrows > RS_AMG_LDS_ROWS;
    if (global && nb > 1)
        throw RsError{RSGPU_BAD_ARGUMENT, "AMG smoother block larger than the LDS vector"};
    auto offd = [&](int r) {
        int c = 0;
        for (int j = rp[r]; j < rp[r + 1]; ++j)
            c += ci[j] != r;
        return c;
    };
    std::vector<int> level(n, 0), pushed(n, 0), blk(n, 0), colour;
    for (int b = 0; b < nb; ++b)
        for (int r = bstart[b]; r < bstart[b + 1]; ++r)
            blk[r] = b;
    if (ordering == 1)
        greedy_colour(n, rp, ci, blk, colour);
    std::vector<int> blk_lev(nb + 1), rows, rlen, row_glob, gofs, col, src;
    std::vector<AmgLevelMeta> meta;
    int nlev_max = 0;
    for (int b = 0; b < nb; ++b) {
        const int b0 = bstart[b], b1 = bstart[b + 1];
        int maxl = -1;
        if (ordering == 1) {
            for (int r = b0; r < b1; ++r) {
                level[r] = colour[r];
                maxl = std::max(maxl, level[r]);
            }
        } else {
            for (int r = b0; r < b1; ++r) {
                int l = pushed[r];
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = ci[j];
                    if (c >= b0 && c < r)
                        l = std::max(l, level[c] + 1);
                }
                level[r] = l;
                maxl = std::max(maxl, l);
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = ci[j];
                    if (c > r && c < b1)
                        pushed[c] = std::max(pushed[c], l + 1);
                }
            }
        }
        const int nlev = maxl + 1;
        nlev_max = std::max(nlev_max, nlev);
        std::vector<std::vector<int>> byl(std::max(nlev, 0));
        for (int r = b0; r < b1; ++r)
            byl[level[r]].push_back(r);
        blk_lev[b] = (int)meta.size();
        for (int l = 0; l < nlev; ++l) {
            // rows of one level do not couple: any order within it; longest first
            std::vector<int> &R = byl[l];
            std::vector<int> len(R.size());
            for (size_t t = 0; t < R.size(); ++t)
                len[t] = offd(R[t]);
            std::vector<int> ord(R.size());
            for (size_t t = 0; t < R.size(); ++t)
                ord[t] = (int)t;
            std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return len[x] > len[y]; });
            AmgLevelMeta m;
            m.count = (int)R.size();
            m.gbase = (int)gofs.size();
            m.roff = (int)rows.size();
            // pad: 1 = eight lanes per row (long rows), 0 = one lane per row
            long long tot = 0;
            for (size_t t = 0; t < R.size(); ++t)
                tot += len[t];
            m.pad = (m.count > 0 && tot > 32LL * m.count && split_rows()) ? 1 : 0;
            for (int t0 = 0; t0 < m.count; t0 += 64) {
                const int tn = std::min(64, m.count - t0);
                const int w = len[ord[t0]];
                const size_t base = col.size();
                if (base + (size_t)64 * w > (size_t)INT32_MAX)
                    throw RsError{RSGPU_BAD_ARGUMENT, "AMG smoother arrays exceed 2^31 entries"};
                gofs.push_back((int)base);
                col.resize(base + (size_t)64 * w, b1 - b0);
                src.resize(base + (size_t)64 * w, -1);
                for (int t = 0; t < tn; ++t) {
                    const int r = R[ord[t0 + t]];
                    rows.push_back(r - b0);
                    rlen.push_back(len[ord[t0 + t]]);
                    row_glob.push_back(r);
                    int q = 0;
                    for (int j = rp[r]; j < rp[r + 1]; ++j) {
                        const int c = ci[j];
                        if (c == r)
                            continue;
                        const size_t o = base + (size_t)64 * q + t;
                        col[o] = (c >= b0 && c < b1) ? c - b0 : -1 - c;
                        src[o] = j;
                        ++q;
                    }
                }
            }
            meta.push_back(m);
        }
    }
    blk_lev[nb] = (int)meta.size();
    // sentinels keep every array non-empty
    col.resize(col.size() + 64, 0);
    src.resize(src.size() + 64, -1);
    row_glob.push_back(0);
    rows.push_back(0);
    rlen.push_back(0);
    gofs.push_back(0);
    L.nb = nb;
    L.max_block_rows = maxrows;
    L.nlev_max = nlev_max;
    L.global_sweep = global;
    L.h_meta = meta;
    L.bstart.upload(bstart, s);
    L.blk_lev.upload(blk_lev, s);
    L.meta.upload((const int *)meta.data(), meta.size() * 4, s);
    L.rows.upload(rows, s);
    L.rlen.upload(rlen, s);
    L.row_glob.upload(row_glob, s);
    L.gofs.upload(gofs, s);
    L.ell_col.upload(col, s);
    L.ell_src.upload(src, s);
    L.ell_size = col.size();
    L.nrows_lev = (int)row_glob.size();
    L.ell_val.alloc(L.ell_size);
    L.dgl.alloc(L.nrows_lev);
    RS_HIP(hipStreamSynchronize(s));   // the host vectors are released next
}

static void ssor_values(AmgLevel &L, hipStream_t s, bool fp32 = false)
{
    hipLaunchKernelGGL(k_amg_gather<double>, dim3(gridn((long long)L.ell_size)), dim3(256), 0, s,
                       L.ell_src.p, L.v.p, L.ell_val.p, L.ell_size);
    hipLaunchKernelGGL(k_amg_rowdiag<double>, dim3(gridn(L.nrows_lev)), dim3(256), 0, s,
                       L.row_glob.p, L.diag.p, L.v.p, L.dgl.p, L.nrows_lev);
    if (fp32) {
        L.ell_val32.alloc(L.ell_size);
        L.dgl32.alloc(L.nrows_lev);
        hipLaunchKernelGGL(k_amg_gather<float>, dim3(gridn((long long)L.ell_size)), dim3(256), 0,
                           s, L.ell_src.p, L.v.p, L.ell_val32.p, L.ell_size);
        hipLaunchKernelGGL(k_amg_rowdiag<float>, dim3(gridn(L.nrows_lev)), dim3(256), 0, s,
                           L.row_glob.p, L.diag.p, L.v.p, L.dgl32.p, L.nrows_lev);
    }
    RS_CHECK_LAUNCH();
}

template <int TAB>
static void spgemm_launch(hipStream_t s, int nrows, const int *Xrp, const int *Xci,
                          const double *Xv, const int *Yrp, const int *Yci, const double *Yv,
                          int *cnt, const int *Crp, int *Cci, double *Cv, int *err)
{
    hipLaunchKernelGGL(k_spgemm<TAB>, dim3(nrows), dim3(64), 0, s, nrows, Xrp, Xci, Xv, Yrp,
                       Yci, Yv, cnt, Crp, Cci, Cv, err);
    RS_CHECK_LAUNCH();
}

// C = X Y (CSR, sorted rows, nonzeros only); returns nnz
static int spgemm(AmgData *D, hipStream_t s, int nrows, const int *Xrp, const int *Xci,
                  const double *Xv, const int *Yrp, const int *Yci, const double *Yv,
                  int ycols, DevBuf<int> &Crp, DevBuf<int> &Cci, DevBuf<double> &Cv)
{
    Crp.alloc((size_t)nrows + 1);
    if (nrows == 0) {
        set_int(s, Crp.p, 0);
        Cci.alloc(1);
        Cv.alloc(1);
        return 0;
    }
    DevBuf<int> cnt, flag;
    cnt.alloc((size_t)nrows + 1);
    flag.alloc(2);
    RS_HIP(hipMemsetAsync(flag.p, 0, 2 * sizeof(int), s));
    hipLaunchKernelGGL(k_spgemm_ub, dim3(gridn(nrows)), dim3(256), 0, s, nrows, Xrp, Xci, Yrp,
                       flag.p);
    RS_CHECK_LAUNCH();
    int ub = std::min(read_int(s, flag.p), ycols);
    int tab = 64;
    while (tab < 2 * ub && tab < 2048)
        tab <<= 1;
    RS_HIP(hipMemsetAsync(flag.p, 0, 2 * sizeof(int), s));
    RS_HIP(hipMemsetAsync(cnt.p + nrows, 0, sizeof(int), s));
    auto go = [&](int *c, const int *rp, int *ci, double *cv) {
        switch (tab) {
        case 64:
            spgemm_launch<64>(s, nrows, Xrp, Xci, Xv, Yrp, Yci, Yv, c, rp, ci, cv, flag.p);
            break;
        case 128:
            spgemm_launch<128>(s, nrows, Xrp, Xci, Xv, Yrp, Yci, Yv, c, rp, ci, cv, flag.p);
            break;
        case 256:
            spgemm_launch<256>(s, nrows, Xrp, Xci, Xv, Yrp, Yci, Yv, c, rp, ci, cv, flag.p);
            break;
        case 512:
            spgemm_launch<512>(s, nrows, Xrp, Xci, Xv, Yrp, Yci, Yv, c, rp, ci, cv, flag.p);
            break;
        case 1024:
            spgemm_launch<1024>(s, nrows, Xrp, Xci, Xv, Yrp, Yci, Yv, c, rp, ci, cv, flag.p);
            break;
        default:
            spgemm_launch<2048>(s, nrows, Xrp, Xci, Xv, Yrp, Yci, Yv, c, rp, ci, cv, flag.p);
            break;
        }
    };
    go(cnt.p, nullptr, nullptr, nullptr);
    if (read_int(s, flag.p))
        throw RsError{RSGPU_BAD_ARGUMENT, "AMG Galerkin product row exceeds 2048 entries"};
    scan_exclusive(D, s, cnt.p, Crp.p, nrows + 1);
    const int nnz = read_int(s, Crp.p + nrows);
    Cci.alloc(std::max(nnz, 1));
    Cv.alloc(std::max(nnz, 1));
    go(nullptr, Crp.p, Cci.p, Cv.p);
    RS_HIP(hipStreamSynchronize(s));
    return nnz;
}

// aggregation of level L into L.pt; returns the aggregate count
static int amg_aggregate(AmgData *D, AmgLevel &L, int k, hipStream_t s)
{
    const int n = L.n;
    const int G = gridn(n);
    DevBuf<int> cnt, Np, Ni, root_of, owner, id, sel, flag;
    DevBuf<char> R, cand;
    DevBuf<unsigned long long> m;
    cnt.alloc((size_t)n + 1);
    Np.alloc((size_t)n + 1);
    R.alloc(n);
    cand.alloc(n);
    m.alloc(n);
    root_of.alloc(n);
    owner.alloc(n);
    id.alloc((size_t)n + 1);
    sel.alloc((size_t)n + 1);
    flag.alloc(1);
    L.pt.alloc(n);
    const double eps = 0.08 * std::pow(0.5, k);
    hipLaunchKernelGGL(k_amg_strength, dim3(G), dim3(256), 0, s, n, L.rp, L.ci, L.v.p, L.diag.p,
                       eps, cnt.p, (const int *)nullptr, (int *)nullptr, R.p);
    RS_HIP(hipMemsetAsync(cnt.p + n, 0, sizeof(int), s));
    scan_exclusive(D, s, cnt.p, Np.p, n + 1);
    const int nn = read_int(s, Np.p + n);
    Ni.alloc(std::max(nn, 1));
    hipLaunchKernelGGL(k_amg_strength, dim3(G), dim3(256), 0, s, n, L.rp, L.ci, L.v.p, L.diag.p,
                       eps, (int *)nullptr, (const int *)Np.p, Ni.p, (char *)nullptr);
    RS_HIP(hipMemsetAsync(root_of.p, 0xff, sizeof(int) * n, s));
    // Every round retires at least the candidate of largest priority; a
    // neighbourhood can only be empty with a NaN diagonal, which the round
    // limit turns into an error.
    const int max_rounds = 10000;
    int rounds = 0;
    // phase 1
    for (;;) {
        if (++rounds > max_rounds)
            throw RsError{RSGPU_BAD_ARGUMENT, "AMG aggregation did not terminate (NaN entries?)"};
        set_int(s, flag.p, 0);
        hipLaunchKernelGGL(k_p1_cand, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, R.p, cand.p, m.p,
                           flag.p);
        RS_CHECK_LAUNCH();
        if (!read_int(s, flag.p))
            break;
        hipLaunchKernelGGL(k_p1_max, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, cand.p, m.p);
        hipLaunchKernelGGL(k_p1_sel, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, cand.p, m.p, R.p,
                           root_of.p);
        RS_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_root_flag, dim3(G), dim3(256), 0, s, n, root_of.p, sel.p);
    RS_HIP(hipMemsetAsync(sel.p + n, 0, sizeof(int), s));
    scan_exclusive(D, s, sel.p, id.p, n + 1);
    int c = read_int(s, id.p + n);
    hipLaunchKernelGGL(k_owner, dim3(G), dim3(256), 0, s, n, root_of.p, id.p, owner.p, L.pt.p);
    // phase 2
    hipLaunchKernelGGL(k_enlarge, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, owner.p, R.p,
                       L.pt.p);
    RS_CHECK_LAUNCH();
    // phase 3
    rounds = 0;
    for (;;) {
        if (++rounds > max_rounds)
            throw RsError{RSGPU_BAD_ARGUMENT, "AMG aggregation did not terminate (NaN entries?)"};
        set_int(s, flag.p, 0);
        hipLaunchKernelGGL(k_p3_cand, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, R.p, cand.p, m.p,
                           flag.p);
        RS_CHECK_LAUNCH();
        if (!read_int(s, flag.p))
            break;
        hipLaunchKernelGGL(k_p3_max, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, R.p, cand.p, m.p);
        hipLaunchKernelGGL(k_p3_sel, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, R.p, cand.p, m.p,
                           sel.p);
        RS_HIP(hipMemsetAsync(sel.p + n, 0, sizeof(int), s));
        scan_exclusive(D, s, sel.p, id.p, n + 1);
        hipLaunchKernelGGL(k_p3_assign, dim3(G), dim3(256), 0, s, n, Np.p, Ni.p, R.p, sel.p,
                           id.p, c, L.pt.p);
        hipLaunchKernelGGL(k_p3_clear, dim3(G), dim3(256), 0, s, n, L.pt.p, R.p);
        RS_CHECK_LAUNCH();
        c += read_int(s, id.p + n);
    }
    if (c == 0)
        return 0;
    // prolongator (count, scan, fill)
    L.Prp.alloc((size_t)n + 1);
    hipLaunchKernelGGL(k_prolong, dim3(G), dim3(256), 0, s, n, L.rp, L.ci, L.v.p, L.diag.p,
                       Np.p, Ni.p, L.pt.p, D->p.omega, cnt.p, (const int *)nullptr,
                       (int *)nullptr, (double *)nullptr);
    RS_HIP(hipMemsetAsync(cnt.p + n, 0, sizeof(int), s));
    scan_exclusive(D, s, cnt.p, L.Prp.p, n + 1);
    L.pnnz = read_int(s, L.Prp.p + n);
    L.Pci.alloc(std::max(L.pnnz, 1));
    L.Pv.alloc(std::max(L.pnnz, 1));
    hipLaunchKernelGGL(k_prolong, dim3(G), dim3(256), 0, s, n, L.rp, L.ci, L.v.p, L.diag.p,
                       Np.p, Ni.p, L.pt.p, D->p.omega, (int *)nullptr, (const int *)L.Prp.p,
                       L.Pci.p, L.Pv.p);
    RS_CHECK_LAUNCH();
    return c;
}

// P' as CSR (c rows), rows ascending within each column (stable sort)
static void amg_transpose(AmgData *D, AmgLevel &L, hipStream_t s)
{
    const int n = L.n, c = L.c, nnz = L.pnnz;
    L.Trp.alloc((size_t)c + 1);
    L.Tci.alloc(std::max(nnz, 1));
    L.Tv.alloc(std::max(nnz, 1));
    DevBuf<int> erow, keys_out, vals_in, vals_out, cnt;
    erow.alloc(std::max(nnz, 1));
    keys_out.alloc(std::max(nnz, 1));
    vals_in.alloc(std::max(nnz, 1));
    vals_out.alloc(std::max(nnz, 1));
    cnt.alloc((size_t)c + 1);
    hipLaunchKernelGGL(k_entry_rows, dim3(gridn(n)), dim3(256), 0, s, n, L.Prp.p, erow.p);
    hipLaunchKernelGGL(k_iota, dim3(gridn(nnz)), dim3(256), 0, s, nnz, vals_in.p);
    RS_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * ((size_t)c + 1), s));
    hipLaunchKernelGGL(k_col_count, dim3(gridn(nnz)), dim3(256), 0, s, nnz, L.Pci.p, cnt.p);
    RS_CHECK_LAUNCH();
    scan_exclusive(D, s, cnt.p, L.Trp.p, c + 1);
    int bits = 1;
    while ((1 << bits) <= c && bits < 31)
        ++bits;
    size_t bytes = 0;
    RS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, L.Pci.p, keys_out.p, vals_in.p,
                                              vals_out.p, nnz, 0, bits, s));
    if (D->tmp_bytes.n < bytes)
        D->tmp_bytes.alloc(bytes);
    RS_HIP(hipcub::DeviceRadixSort::SortPairs(D->tmp_bytes.p, bytes, L.Pci.p, keys_out.p,
                                              vals_in.p, vals_out.p, nnz, 0, bits, s));
    hipLaunchKernelGGL(k_transpose_fill, dim3(gridn(nnz)), dim3(256), 0, s, nnz, vals_out.p,
                       erow.p, L.Pv.p, L.Tci.p, L.Tv.p);
    RS_CHECK_LAUNCH();
    RS_HIP(hipStreamSynchronize(s));
}

// u and ualt carry one extra zero entry: the padding column of an in-place
// (global) sweep
static void level_vectors(AmgLevel &L, hipStream_t s, bool fp32 = false)
{
    L.u.alloc((size_t)L.n + 1);
    L.ualt.alloc((size_t)L.n + 1);
    L.u.zero(s);
    L.ualt.zero(s);
    L.f.alloc(std::max(L.n, 1));
    L.r.alloc(std::max(L.n, 1));
    if (fp32) {
        L.u32.alloc((size_t)L.n + 1);
        L.ualt32.alloc((size_t)L.n + 1);
        L.u32.zero(s);
        L.ualt32.zero(s);
        L.f32.alloc(std::max(L.n, 1));
        L.r32.alloc(std::max(L.n, 1));
    }
}

// AMG.setMatrix (AMG.java:228-294).  Level 0 holds its own copy of A
// (new CompRowMatrix(A), AMG.java:232); its pattern, SpMV plan and smoother
// schedule are kept while A's pattern stays the same.
int amg_setup(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    hipStream_t s = M->ctx->stream;
    if (A->n != A->ncols)
        throw RsError{RSGPU_BAD_ARGUMENT, "AMG only applies to square matrices"};
    if (!M->amg)
        M->amg = new AmgData();
    AmgData *D = M->amg;
    D->p = M->amg_params;
    D->fp32 = M->precision == RSGPU_FP32;
    const rsgpu_amg_params &p = D->p;
    const bool same = !D->L.empty() && D->pattern0 == A->pattern_id;
    if (!same) {
        D->L.clear();
        D->L.push_back(std::make_unique<AmgLevel>());
        AmgLevel &L0 = *D->L[0];
        L0.n = A->n;
        L0.nnz = A->nnz;
        L0.rp_own.alloc((size_t)A->n + 1);
        L0.ci_own.alloc(std::max(A->nnz, 1));
        RS_HIP(hipMemcpyAsync(L0.rp_own.p, A->rowptr.p, sizeof(int) * ((size_t)A->n + 1),
                              hipMemcpyDeviceToDevice, s));
        if (A->nnz)
            RS_HIP(hipMemcpyAsync(L0.ci_own.p, A->colind.p, sizeof(int) * A->nnz,
                                  hipMemcpyDeviceToDevice, s));
        L0.rp = L0.rp_own.p;
        L0.ci = L0.ci_own.p;
        L0.h_rp = A->h_rowptr;
        L0.h_ci = A->h_colind;
        plan_from_host(L0.h_rp, L0.n, L0.plan, s);
        D->pattern0 = A->pattern_id;
    } else {
        D->L.resize(1);
    }
    {
        AmgLevel &L0 = *D->L[0];
        L0.v.alloc(std::max(A->nnz, 1));
        if (A->nnz)
            RS_HIP(hipMemcpyAsync(L0.v.p, A->val.p, sizeof(double) * A->nnz,
                                  hipMemcpyDeviceToDevice, s));
    }
    DevBuf<int> bad;
    bad.alloc(1);
    for (int k = 0; D->L[k]->n > p.min; ++k) {
        AmgLevel &L = *D->L[k];
        const int n = L.n;
        L.diag.alloc(n);
        set_int(s, bad.p, 0x7fffffff);
        hipLaunchKernelGGL(k_amg_diag, dim3(gridn(n)), dim3(256), 0, s, n, L.rp, L.ci, L.diag.p,
                           bad.p);
        RS_CHECK_LAUNCH();
        const int b = read_int(s, bad.p);
        if (b != 0x7fffffff) {
            if (bad_row)
                *bad_row = b;
            M->ctx->err = "Matrix is missing a diagonal entry on row " + std::to_string(b);
            D->L.clear();
            return RSGPU_MISSING_DIAGONAL;
        }
        const int c = amg_aggregate(D, L, k, s);
        if (c == 0)
            break;
        // optional stop: a level whose aggregates keep more than the given
        // fraction of its rows (aggregates degenerating towards singletons)
        // becomes the coarsest
        if (p.coarsen_ratio > 0 && (double)c > p.coarsen_ratio * n)
            break;
        L.c = c;
        amg_transpose(D, L, s);
        // AP = A P (n x c), Ac = P' AP (c x c)
        DevBuf<int> APrp, APci;
        DevBuf<double> APv;
        spgemm(D, s, n, L.rp, L.ci, L.v.p, L.Prp.p, L.Pci.p, L.Pv.p, c, APrp, APci, APv);
        D->L.push_back(std::make_unique<AmgLevel>());
        AmgLevel &Lc = *D->L[k + 1];
        Lc.n = c;
        Lc.nnz = spgemm(D, s, c, L.Trp.p, L.Tci.p, L.Tv.p, APrp.p, APci.p, APv.p, c,
                        Lc.rp_own, Lc.ci_own, Lc.v);
        Lc.rp = Lc.rp_own.p;
        Lc.ci = Lc.ci_own.p;
    }
    D->m = (int)D->L.size();
    // smoothers, transfer plans and vectors
    for (int k = 0; k < D->m; ++k) {
        AmgLevel &L = *D->L[k];
        level_vectors(L, s, D->fp32);
        if (k == D->m - 1)
            break;
        if (k > 0) {
            download(s, L.h_rp, L.rp, (size_t)L.n + 1);
            RS_HIP(hipStreamSynchronize(s));
            download(s, L.h_ci, L.ci, (size_t)L.h_rp[L.n]);
        }
        std::vector<int> prp, trp;
        download(s, prp, L.Prp.p, (size_t)L.n + 1);
        download(s, trp, L.Trp.p, (size_t)L.c + 1);
        RS_HIP(hipStreamSynchronize(s));
        if (k > 0)
            plan_from_host(L.h_rp, L.n, L.plan, s);
        plan_from_host(prp, L.n, L.Pplan, s);
        plan_from_host(trp, L.c, L.Tplan, s);
        if (!(k == 0 && L.sched_pattern == A->pattern_id && L.nb_rows_param == p.block_rows &&
              L.ordering_param == p.ordering)) {
            build_ssor_schedule(L, p.block_rows, p.ordering, s);
            L.sched_pattern = k == 0 ? A->pattern_id : 0;
            L.nb_rows_param = p.block_rows;
            L.ordering_param = p.ordering;
        }
        ssor_values(L, s, D->fp32);
        if (D->fp32) {
            // FP32 copies of the level's matrix and transfer operators
            L.v32.alloc(std::max(L.nnz, 1));
            L.Pv32.alloc(std::max(L.pnnz, 1));
            L.Tv32.alloc(std::max(L.pnnz, 1));
            hipLaunchKernelGGL((k_amg_cvt<double, float>), dim3(gridn(L.nnz)), dim3(256), 0, s,
                               L.v.p, L.v32.p, (size_t)L.nnz, Skip{nullptr, nullptr});
            hipLaunchKernelGGL((k_amg_cvt<double, float>), dim3(gridn(L.pnnz)), dim3(256), 0, s,
                               L.Pv.p, L.Pv32.p, (size_t)L.pnnz, Skip{nullptr, nullptr});
            hipLaunchKernelGGL((k_amg_cvt<double, float>), dim3(gridn(L.pnnz)), dim3(256), 0, s,
                               L.Tv.p, L.Tv32.p, (size_t)L.pnnz, Skip{nullptr, nullptr});
            RS_CHECK_LAUNCH();
        }
    }
    // coarsest level: dense LU (AMG.java:268-271)
    AmgLevel &Lc = *D->L[D->m - 1];
    D->nc = Lc.n;
    if (D->nc > p.max_coarse)
        throw RsError{RSGPU_BAD_ARGUMENT,
                      "AMG coarsest level has " + std::to_string(D->nc) +
                          " rows, above max_coarse"};
    D->lu.alloc((size_t)std::max(D->nc, 1) * std::max(D->nc, 1));
    D->ipiv.alloc(std::max(D->nc, 1));
    D->flag.alloc(1);
    if (D->nc > 0) {
        const int *crp = Lc.rp;
        const int *cci = Lc.ci;
        hipLaunchKernelGGL(k_dense_fill, dim3(gridn(D->nc)), dim3(256), 0, s, D->nc, crp, cci,
                           Lc.v.p, D->lu.p);
        if (D->nc <= 256) {
            hipLaunchKernelGGL(k_dense_getf2, dim3(1), dim3(256), 0, s, D->nc, D->lu.p,
                               D->ipiv.p, D->flag.p);
        } else {
            set_int(s, D->flag.p, 0);
            for (int j = 0; j < D->nc; ++j) {
                hipLaunchKernelGGL(k_getf2_pivot, dim3(1), dim3(1024), 0, s, D->nc, j, D->lu.p,
                                   D->ipiv.p, D->flag.p);
                const long long w = D->nc - 1 - j;
                if (w > 0)
                    hipLaunchKernelGGL(k_getf2_update, dim3(gridn(w * w)), dim3(256), 0, s,
                                       D->nc, j, D->lu.p);
            }
        }
        RS_CHECK_LAUNCH();
        if (read_int(s, D->flag.p)) {
            M->ctx->err = "AMG coarsest matrix is singular";
            return RSGPU_SINGULAR;
        }
        // above 256 rows the cycle's coarsest solve multiplies by the
        // explicit inverse (its columns solved from the LU at setup): the
        // one-workgroup LU solve costs two barriers per row
        if (D->nc > 256) {
            D->cinv.alloc((size_t)D->nc * D->nc);
            D->cpart.alloc((size_t)D->nc * RS_CINV_SLICES);
            hipLaunchKernelGGL(k_dense_inv, dim3(D->nc), dim3(256), 0, s, D->nc, D->lu.p,
                               D->ipiv.p, D->cinv.p);
            RS_CHECK_LAUNCH();
        } else {
            D->cinv.release();
        }
    }
    if (D->fp32) {
        D->cf.alloc(std::max(D->nc, 1));
        D->cu.alloc(std::max(D->nc, 1));
    }
    RS_HIP(hipStreamSynchronize(s));
    return RSGPU_OK;
}

// ---------------------------------------------------------------- cycle
// the level's arrays of the cycle's precision
template <typename T>
struct AmgArr;
template <>
struct AmgArr<double> {
    static double *u(AmgLevel &L) { return L.u.p; }
    static double *f(AmgLevel &L) { return L.f.p; }
    static double *r(AmgLevel &L) { return L.r.p; }
    static const double *v(AmgLevel &L) { return L.v.p; }
    static const double *Pv(AmgLevel &L) { return L.Pv.p; }
    static const double *Tv(AmgLevel &L) { return L.Tv.p; }
    static const double *ev(AmgLevel &L) { return L.ell_val.p; }
    static const double *dg(AmgLevel &L) { return L.dgl.p; }
    static double *ualt(AmgLevel &L) { return L.ualt.p; }
    static void swap(AmgLevel &L) { swap_buf(L.u, L.ualt); }
};
template <>
struct AmgArr<float> {
    static float *u(AmgLevel &L) { return L.u32.p; }
    static float *f(AmgLevel &L) { return L.f32.p; }
    static float *r(AmgLevel &L) { return L.r32.p; }
    static const float *v(AmgLevel &L) { return L.v32.p; }
    static const float *Pv(AmgLevel &L) { return L.Pv32.p; }
    static const float *Tv(AmgLevel &L) { return L.Tv32.p; }
    static const float *ev(AmgLevel &L) { return L.ell_val32.p; }
    static const float *dg(AmgLevel &L) { return L.dgl32.p; }
    static float *ualt(AmgLevel &L) { return L.ualt32.p; }
    static void swap(AmgLevel &L) { swap_buf(L.u32, L.ualt32); }
};

// the coarsest solve in FP64: u = C^-1 f (DenseLU.solve, or the inverse)
static void amg_coarse64(AmgData *D, const double *f, double *u, bool trans, hipStream_t s,
                         Skip sk)
{
    if (D->nc == 0)
        return;
    if (D->cinv.p && !trans) {
        hipLaunchKernelGGL(k_amg_cinv_part, dim3(gridn(D->nc), RS_CINV_SLICES), dim3(256), 0, s,
                           D->nc, D->cinv.p, f, D->cpart.p, sk);
        hipLaunchKernelGGL(k_amg_cinv_sum, dim3(gridn(D->nc)), dim3(256), 0, s, D->nc,
                           D->cpart.p, u, sk);
    } else if (D->cinv.p)
        hipLaunchKernelGGL(k_amg_cinv_t, dim3((D->nc + 3) / 4), dim3(256), 0, s, D->nc,
                           D->cinv.p, f, u, sk);
    else if (D->nc <= 64 && !trans)
        hipLaunchKernelGGL(k_dense_getrs64, dim3(1), dim3(64), 0, s, D->nc, D->lu.p, D->ipiv.p,
                           f, u, sk);
    else
        hipLaunchKernelGGL(k_dense_getrs, dim3(1), dim3(256), 0, s, D->nc, D->lu.p, D->ipiv.p,
                           f, u, trans ? 1 : 0, sk);
    RS_CHECK_LAUNCH();
}

// one SSOR application on level L (the block kernel or, for one block past
// the LDS vector, one launch per level / colour)
template <typename T>
static void amg_smooth(AmgLevel &L, const T *f, double wf, double wr, int reverse,
                       hipStream_t s, Skip sk)
{
    using A = AmgArr<T>;
    if (L.global_sweep) {
        const int nl = (int)L.h_meta.size();
        auto launch = [&](int l, double w) {
            const AmgLevelMeta &m = L.h_meta[l];
            const int ng = (m.count + 63) >> 6;
            const int nblk = m.pad ? (m.count + 31) / 32 : (ng + 3) / 4;
            hipLaunchKernelGGL(k_amg_ssor_level<T>, dim3(std::max(1, nblk)), dim3(256),
                               0, s, m, L.rows.p, L.rlen.p, L.gofs.p, L.ell_col.p, A::ev(L),
                               A::dg(L), f, A::u(L), (T)w, sk);
        };
        for (int l = 0; l < nl; ++l)
            launch(l, wf);
        if (reverse)
            for (int l = nl - 1; l >= 0; --l)
                launch(l, wr);
        RS_CHECK_LAUNCH();
        return;
    }
    hipLaunchKernelGGL(k_amg_ssor<T>, dim3(L.nb), dim3(256), 0, s, L.bstart.p, L.blk_lev.p,
                       (const AmgLevelMeta *)L.meta.p, L.rows.p, L.rlen.p, L.gofs.p,
                       L.ell_col.p, A::ev(L), A::dg(L), f, A::u(L), A::ualt(L), (T)wf, (T)wr,
                       reverse, sk);
    RS_CHECK_LAUNCH();
    A::swap(L);
}

// AMG.cycle (AMG.java:302-342) in precision T; the coarsest solve in FP64
template <typename T>
static void amg_cycle(AmgData *D, int k, const T *f, bool trans, hipStream_t s, Skip sk)
{
    using A = AmgArr<T>;
    AmgLevel &L = *D->L[k];
    const rsgpu_amg_params &p = D->p;
    if (k == D->m - 1) {
        if (std::is_same<T, double>::value) {
            amg_coarse64(D, (const double *)f, (double *)A::u(L), trans, s, sk);
        } else if (D->nc > 0) {
            hipLaunchKernelGGL((k_amg_cvt<T, double>), dim3(gridn(D->nc)), dim3(256), 0, s, f,
                               D->cf.p, (size_t)D->nc, sk);
            amg_coarse64(D, D->cf.p, D->cu.p, trans, s, sk);
            hipLaunchKernelGGL((k_amg_cvt<double, T>), dim3(gridn(D->nc)), dim3(256), 0, s,
                               D->cu.p, A::u(L), (size_t)D->nc, sk);
            RS_CHECK_LAUNCH();
        }
        return;
    }
    // SSOR.transApply is apply (the matrix is assumed symmetric)
    for (int i = 0; i < p.nu1; ++i)
        amg_smooth<T>(L, f, p.omega_pre_f, p.omega_pre_r, p.reverse, s, sk);
    AmgLevel &C = *D->L[k + 1];
    hipLaunchKernelGGL(k_amg_copy<T>, dim3(gridn(C.n)), dim3(256), 0, s, (const T *)nullptr,
                       A::u(C), C.n, sk);
    // r = f - A u (r.set(f); A.multAdd(-1, u, r))
    hipLaunchKernelGGL((k_amg_spmv<T, EpiResid<T>>), dim3(L.plan.nwg), dim3(RS_SPMV_BS), 0, s,
                       L.plan.wg_rows.p, L.rp, L.ci, A::v(L), A::u(L),
                       EpiResid<T>{A::r(L), f, T(-1)}, sk);
    // f_c = I' r (CompColMatrix.transMult)
    hipLaunchKernelGGL((k_amg_spmv<T, EpiPlain<T>>), dim3(L.Tplan.nwg), dim3(RS_SPMV_BS), 0, s,
                       L.Tplan.wg_rows.p, L.Trp.p, L.Tci.p, A::Tv(L), A::r(L),
                       EpiPlain<T>{A::f(C)}, sk);
    RS_CHECK_LAUNCH();
    for (int g = 0; g < p.gamma; ++g)
        amg_cycle<T>(D, k + 1, A::f(C), trans, s, sk);
    // u += I u_c (CompColMatrix.multAdd)
    hipLaunchKernelGGL((k_amg_spmv<T, EpiSeqAdd<T>>), dim3(L.Pplan.nwg), dim3(RS_SPMV_BS), 0, s,
                       L.Pplan.wg_rows.p, L.Prp.p, L.Pci.p, A::Pv(L), A::u(C),
                       EpiSeqAdd<T>{A::u(L)}, sk);
    RS_CHECK_LAUNCH();
    for (int i = 0; i < p.nu2; ++i)
        amg_smooth<T>(L, f, p.omega_post_f, p.omega_post_r, p.reverse, s, sk);
}

template <typename T>
static void amg_apply_t(AmgData *D, const T *b, T *x, hipStream_t s, Skip sk, bool trans)
{
    using A = AmgArr<T>;
    AmgLevel &L0 = *D->L[0];
    const int n = L0.n;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_amg_copy<T>, dim3(gridn(n)), dim3(256), 0, s,
                       D->p.zero_guess ? (const T *)nullptr : (const T *)x, A::u(L0), n, sk);
    RS_CHECK_LAUNCH();
    amg_cycle<T>(D, 0, b, trans, s, sk);
    hipLaunchKernelGGL(k_amg_copy<T>, dim3(gridn(n)), dim3(256), 0, s, (const T *)A::u(L0), x, n,
                       sk);
    RS_CHECK_LAUNCH();
}

// AMG.apply / transApply (AMG.java:208-226); an FP32 hierarchy on FP64
// vectors converts at the boundary
void amg_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st, bool trans)
{
    AmgData *D = M->amg;
    hipStream_t s = M->ctx->stream;
    Skip sk{st ? &st->done : nullptr, st ? &st->cycle_end : nullptr};
    const int n = D->L[0]->n;
    if (!D->fp32) {
        amg_apply_t<double>(D, b, x, s, sk, trans);
        return;
    }
    M->tmpa32.alloc(std::max(n, 1));
    M->tmpb32.alloc(std::max(n, 1));
    hipLaunchKernelGGL((k_amg_cvt<double, float>), dim3(gridn(n)), dim3(256), 0, s, b,
                       M->tmpa32.p, (size_t)n, sk);
    hipLaunchKernelGGL((k_amg_cvt<double, float>), dim3(gridn(n)), dim3(256), 0, s,
                       (const double *)x, M->tmpb32.p, (size_t)n, sk);
    amg_apply_t<float>(D, M->tmpa32.p, M->tmpb32.p, s, sk, trans);
    hipLaunchKernelGGL((k_amg_cvt<float, double>), dim3(gridn(n)), dim3(256), 0, s,
                       (const float *)M->tmpb32.p, x, (size_t)n, sk);
    RS_CHECK_LAUNCH();
}

void amg_apply32(rsgpu_prec *M, const float *b, float *x, const DevState *st)
{
    AmgData *D = M->amg;
    Skip sk{st ? &st->done : nullptr, st ? &st->cycle_end : nullptr};
    if (!D->fp32)
        throw RsError{RSGPU_BAD_ARGUMENT, "FP32 apply needs an FP32 AMG"};
    amg_apply_t<float>(D, b, x, M->ctx->stream, sk, false);
}

int amg_levels(const rsgpu_prec *M, int *sizes, int *nnzs, int cap)
{
    const AmgData *D = M->amg;
    if (!D)
        return 0;
    for (int k = 0; k < D->m && k < cap; ++k) {
        if (sizes)
            sizes[k] = D->L[k]->n;
        if (nnzs)
            nnzs[k] = D->L[k]->nnz;
    }
    return D->m;
}

int amg_get_level(const rsgpu_prec *M, int k, int32_t *rowptr, int32_t *colind, double *data,
                  int32_t *pt, int32_t *Prp, int32_t *Pci, double *Pv, int *pnnz)
{
    const AmgData *D = M->amg;
    if (!D || k < 0 || k >= D->m)
        return RSGPU_BAD_ARGUMENT;
    const AmgLevel &L = *D->L[k];
    const int n = L.n, nnz = L.nnz;
    if (rowptr)
        RS_HIP(hipMemcpy(rowptr, L.rp, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost));
    if (colind && nnz)
        RS_HIP(hipMemcpy(colind, L.ci, sizeof(int) * nnz, hipMemcpyDeviceToHost));
    if (data && nnz)
        RS_HIP(hipMemcpy(data, L.v.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    if (pnnz)
        *pnnz = k < D->m - 1 ? L.pnnz : 0;
    if (k < D->m - 1) {
        if (pt && n)
            RS_HIP(hipMemcpy(pt, L.pt.p, sizeof(int) * n, hipMemcpyDeviceToHost));
        if (Prp)
            RS_HIP(hipMemcpy(Prp, L.Prp.p, sizeof(int) * ((size_t)n + 1),
                             hipMemcpyDeviceToHost));
        if (Pci && L.pnnz)
            RS_HIP(hipMemcpy(Pci, L.Pci.p, sizeof(int) * L.pnnz, hipMemcpyDeviceToHost));
        if (Pv && L.pnnz)
            RS_HIP(hipMemcpy(Pv, L.Pv.p, sizeof(double) * L.pnnz, hipMemcpyDeviceToHost));
    }
    return RSGPU_OK;
}

// ---------------------------------------------------------------- SSOR preconditioner
// SSOR.setMatrix (SSOR.java:132-141): copy of the matrix, diagonal indices
// (missing diagonal -> error), and the sweep schedule (reused while the
// pattern is the same).  The parameters live in amg_params (omega_pre_f,
// omega_pre_r, reverse, block_rows, ordering).
int ssor_setup(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    hipStream_t s = M->ctx->stream;
    if (A->n != A->ncols)
        throw RsError{RSGPU_BAD_ARGUMENT, "SSOR only applies to square matrices"};
    if (!M->amg)
        M->amg = new AmgData();
    AmgData *D = M->amg;
    D->p = M->amg_params;
    const rsgpu_amg_params &p = D->p;
    const bool same = !D->L.empty() && D->pattern0 == A->pattern_id;
    if (!same) {
        D->L.clear();
        D->L.push_back(std::make_unique<AmgLevel>());
        AmgLevel &L0 = *D->L[0];
        L0.n = A->n;
        L0.nnz = A->nnz;
        L0.rp = A->rowptr.p;       // pattern borrowed; values copied below
        L0.ci = A->colind.p;
        L0.h_rp = A->h_rowptr;
        L0.h_ci = A->h_colind;
        D->pattern0 = A->pattern_id;
    }
    AmgLevel &L = *D->L[0];
    L.rp = A->rowptr.p;
    L.ci = A->colind.p;
    L.v.alloc(std::max(A->nnz, 1));
    if (A->nnz)
        RS_HIP(hipMemcpyAsync(L.v.p, A->val.p, sizeof(double) * A->nnz,
                              hipMemcpyDeviceToDevice, s));
    L.diag.alloc(std::max(L.n, 1));
    DevBuf<int> bad;
    bad.alloc(1);
    set_int(s, bad.p, 0x7fffffff);
    if (L.n)
        hipLaunchKernelGGL(k_amg_diag, dim3(gridn(L.n)), dim3(256), 0, s, L.n, L.rp, L.ci,
                           L.diag.p, bad.p);
    RS_CHECK_LAUNCH();
    const int b = read_int(s, bad.p);
    if (b != 0x7fffffff) {
        if (bad_row)
            *bad_row = b;
        M->ctx->err = "Missing diagonal on row " + std::to_string(b);
        D->L.clear();
        return RSGPU_MISSING_DIAGONAL;
    }
    if (!(L.sched_pattern == A->pattern_id && L.nb_rows_param == p.block_rows &&
          L.ordering_param == p.ordering)) {
        build_ssor_schedule(L, p.block_rows, p.ordering, s);
        L.sched_pattern = A->pattern_id;
        L.nb_rows_param = p.block_rows;
        L.ordering_param = p.ordering;
        level_vectors(L, s);
    }
    ssor_values(L, s);
    D->m = 1;
    RS_HIP(hipStreamSynchronize(s));
    return RSGPU_OK;
}

// SSOR.apply(b, x): one application starting from x's content
void ssor_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st)
{
    AmgData *D = M->amg;
    hipStream_t s = M->ctx->stream;
    Skip sk{st ? &st->done : nullptr, st ? &st->cycle_end : nullptr};
    AmgLevel &L = *D->L[0];
    const int n = L.n;
    if (n == 0)
        return;
    const rsgpu_amg_params &p = D->p;
    hipLaunchKernelGGL(k_amg_copy<double>, dim3(gridn(n)), dim3(256), 0, s, (const double *)x,
                       L.u.p, n, sk);
    amg_smooth<double>(L, b, p.omega_pre_f, p.omega_pre_r, p.reverse, s, sk);
    hipLaunchKernelGGL(k_amg_copy<double>, dim3(gridn(n)), dim3(256), 0, s,
                       (const double *)L.u.p, x, n, sk);
    RS_CHECK_LAUNCH();
}
