// ILU(0) and block-Jacobi ILU(0) on gfx950.
//
// Reference: ILU.factor (mtj/src/no/uib/cipr/matrix/sparse/ILU.java:86-125),
// ILU.apply -> UnitLowerCompRowMatrix.solve / UpperCompRowMatrix.solve
// (UnitLowerCompRowMatrix.java:49-68, UpperCompRowMatrix.java:49-68),
// ILU.transApply -> the transSolve column sweeps (:70-91, :70-88), and
// block-Jacobi = BlockDiagonalPreconditioner over contiguous row ranges
// (mtj/src/no/uib/cipr/matrix/distributed/BlockDiagonalPreconditioner.java:
// 46-75, DistRowMatrix.java:56-76).
//
// Design.  Natural-order ILU(0) is sequential along its dependency chains,
// so rows are grouped into levels (rows whose L-part only references earlier
// levels) per block.  Each block is owned by one workgroup, which walks its
// levels with workgroup barriers only; blocks never synchronise with each
// other.  Within a level one lane owns one row and performs exactly the
// reference's per-row operation sequence, so factors and sweeps are
// bit-identical to the sequential loops.
//
// Sweep layout.  The block's vector lives in LDS and is initialised with the
// block's right-hand side, so both sweeps run in place (the pipelined kernel
// up to the device's opt-in LDS limit; the generic kernel up to 64 KiB, and
// through a global scratch region beyond).  The factor values of each level
// are stored lane-major: row t of a level owns W consecutive (value, local
// column) slots, W the sweep's largest row padded to a multiple of 4, so a
// lane fetches its operands with a few 16-byte loads.  Padding slots are
// identity updates (value -0.0 for the sum form, +0.0 for the subtract
// form, column = the zero slot), which keeps results bit-identical including
// signed zeros.  The U sweeps' diagonals are stored in level order.  With
// one wave per block, the pipelined apply keeps up to eight levels of
// operands in flight in a register ring (see sweep_pf).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "spmv_core.hpp"

#define RS_LDS_BYTES 65536

// Records are padded by RS_BAND_PAD rows on both sides (zeros) and so is
// the block's LDS vector, so every lane addresses its row at step s as
// base +/- s without clamping: prefetches past either end of a line, a block
// or the matrix read padding or a neighbouring line and are never used.
#define RS_BAND_PAD 64

struct LevelMeta {
    int count, off, roff, pad;
};

// cell factorisation records: at most RS_CELL_NB blocks per block row
#define RS_CELL_NB 8
#define RS_CREC (4 + 4 * RS_CELL_NB)
#define RS_CQ (3 * RS_CELL_NB)

// ---------------------------------------------------------------- host setup
struct HostSweep {
    int W = 4;
    std::vector<int> blk_lev;          // nblocks+1
    std::vector<LevelMeta> meta;
    std::vector<int> rows, row_glob, col, src;
    int max_count = 0, max_width = 0;
    int R = 1;                         // rows per level entry (3: cells)
    int irec = 0, vrec = 0;            // R = 3: record sizes per cell
};

static int lds_limit(int device);
static int ilu_kernel_choice();
static void band_setup(rsgpu_prec *M, const std::vector<int> &rp, const std::vector<int> &ci,
                       const std::vector<int> &dg);
static size_t cell_xv_bytes(int max_block_rows, int es);

static int pad_width(int w)
{
    w = std::max(w, 1);
    return (w + 3) & ~3;
}

// Entries of one triangular part in operation order.  Row r's entries are
// positions j in [begin(r), end(r)), each with a column col(j) and a position
// pos(j) in the factor's CSR values.
struct RangeEntries {   // a contiguous range of each factor row (L or U)
    const int *b0, *e0;  // begin/end per row
    int boff, eoff;      // added to b0[r] / e0[r]
    const int *ci;
    int begin(int r) const { return b0[r] + boff; }
    int end(int r) const { return e0[r] + eoff; }
    int col(int j) const { return ci[j]; }
    int pos(int j) const { return j; }
};
struct FlatEntries {    // an explicit list per row (transposed factors)
    std::vector<int> ptr, cols, poss;
    int begin(int r) const { return ptr[r]; }
    int end(int r) const { return ptr[r + 1]; }
    int col(int j) const { return cols[j]; }
    int pos(int j) const { return poss[j]; }
};

// Levels of one triangular part.  forward: dependencies on smaller rows;
// else larger.
template <class Ent>
static void build_sweep(int nblocks, const std::vector<int> &bstart, const Ent &ent,
                        bool forward, int W, HostSweep &H)
{
    H.W = W;
    H.blk_lev.assign(nblocks + 1, 0);
    H.meta.clear();
    H.rows.clear();
    H.row_glob.clear();
    H.col.clear();
    H.src.clear();
    H.max_count = 0;
    H.max_width = 0;
    const int n = bstart[nblocks];
    std::vector<int> level(n, 0);
    H.rows.reserve(n);
    H.row_glob.reserve(n + 1);
    H.col.reserve((size_t)W * n + W);
    H.src.reserve((size_t)W * n + W);
    std::vector<int> cnt, first, order;
    for (int b = 0; b < nblocks; ++b) {
        const int b0 = bstart[b], b1 = bstart[b + 1], nb = b1 - b0;
        int maxl = -1;
        if (forward) {
            for (int r = b0; r < b1; ++r) {
                int l = 0;
                for (int j = ent.begin(r); j < ent.end(r); ++j)
                    l = std::max(l, level[ent.col(j)] + 1);
                level[r] = l;
                maxl = std::max(maxl, l);
            }
        } else {
            for (int r = b1 - 1; r >= b0; --r) {
                int l = 0;
                for (int j = ent.begin(r); j < ent.end(r); ++j)
                    l = std::max(l, level[ent.col(j)] + 1);
                level[r] = l;
                maxl = std::max(maxl, l);
            }
        }
        const int nlev = maxl + 1;
        // rows by level, ascending within a level (counting sort)
        cnt.assign(nlev + 1, 0);
        for (int r = b0; r < b1; ++r)
            cnt[level[r] + 1]++;
        for (int l = 0; l < nlev; ++l)
            cnt[l + 1] += cnt[l];
        first.assign(cnt.begin(), cnt.end());
        order.resize(nb);
        for (int r = b0; r < b1; ++r)
            order[first[level[r]]++] = r;
        H.blk_lev[b] = (int)H.meta.size();
        for (int l = 0; l < nlev; ++l) {
            LevelMeta m;
            m.count = cnt[l + 1] - cnt[l];
            m.off = (int)H.col.size();
            m.roff = (int)H.rows.size();
            m.pad = 0;
            H.max_count = std::max(H.max_count, m.count);
            for (int t = 0; t < m.count; ++t) {
                const int r = order[cnt[l] + t];
                H.rows.push_back(r - b0);
                H.row_glob.push_back(r);
            }
            H.col.resize(H.col.size() + (size_t)W * m.count, nb);
            H.src.resize(H.src.size() + (size_t)W * m.count, -1);
            for (int t = 0; t < m.count; ++t) {
                const int r = order[cnt[l] + t];
                const int j0 = ent.begin(r), j1 = ent.end(r);
                H.max_width = std::max(H.max_width, j1 - j0);
                for (int j = j0; j < j1; ++j) {
                    H.col[m.off + (size_t)t * W + (j - j0)] = ent.col(j) - b0;
                    H.src[m.off + (size_t)t * W + (j - j0)] = ent.pos(j);
                }
            }
            H.meta.push_back(m);
        }
    }
    H.blk_lev[nblocks] = (int)H.meta.size();
    // trailing dummy row of slots: the arrays are never empty
    H.col.resize(H.col.size() + W, 0);
    H.src.resize(H.src.size() + W, -1);
    H.row_glob.push_back(0);
}

template <class Ent>
static int max_entries(const Ent &ent, int n)
{
    int w = 0;
    for (int r = 0; r < n; ++r)
        w = std::max(w, ent.end(r) - ent.begin(r));
    return w;
}

// es: element size of the apply's vector; the pipelined kernel reads row
// and operand positions as LDS byte offsets (position * es)
static void upload_sweep(const HostSweep &H, SweepData &D, hipStream_t s, int es)
{
    D.W = H.W;
    D.R = H.R;
    D.irec = H.irec;
    D.vrec = H.vrec;
    D.blk_lev.upload(H.blk_lev, s);
    D.lev_meta.upload((const int *)H.meta.data(), H.meta.size() * 4, s);
    D.rows.upload(H.rows, s);
    D.row_glob.upload(H.row_glob, s);
    if (H.R == 3) {
        // cell records: colb already holds LDS byte offsets
        D.colb.upload(H.col, s);
        D.rowsb.upload(std::vector<int>(1, 0), s);
        D.ell_col.release();
    } else {
        D.ell_col.upload(H.col, s);
        std::vector<int> cb(H.col.size()), rb(H.rows.size());
        for (size_t q = 0; q < cb.size(); ++q)
            cb[q] = H.col[q] * es;
        for (size_t q = 0; q < rb.size(); ++q)
            rb[q] = H.rows[q] * es;
        D.colb.upload(cb, s);
        D.rowsb.upload(rb.empty() ? std::vector<int>(1, 0) : rb, s);
    }
    D.ell_src.upload(H.src, s);
    D.ell_size = H.src.size();
    D.nrows_lev = (int)H.row_glob.size();
    D.nlev_total = (int)H.meta.size();
    D.max_levels = 0;
    for (size_t b = 0; b + 1 < H.blk_lev.size(); ++b)
        D.max_levels = std::max(D.max_levels, H.blk_lev[b + 1] - H.blk_lev[b]);
    D.max_count = H.max_count;
    D.max_width = H.max_width;
}

// ---------------------------------------------------------------- cell sweeps
// For 3x3 block matrices (rows 3I..3I+2 of cell I, full 3x3 blocks) the rows
// of a cell depend on each other, so a row schedule needs three levels per
// cell.  A cell schedule puts whole cells on a level; the lane of a cell
// computes its three rows in order with the reference's per-row operation
// sequence: in the L sweep row 3I+a sums its entries in cells before I
// (ascending columns), then its strictly lower diagonal-block entries; in
// the U sweep row 3I+a first sums its strictly upper diagonal-block
// entries, then its entries in later cells.  The three rows of a cell have
// the same neighbour blocks, so the lane gathers each neighbour's three
// vector entries once for all three rows.

static int round4(int x) { return (x + 3) & ~3; }

// position of column c in row r of the factor pattern, or -1
static int find_pos(const std::vector<int> &rp, const std::vector<int> &ci, int r, int c)
{
    auto b = ci.begin() + rp[r], e = ci.begin() + rp[r + 1];
    auto it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? (int)(it - ci.begin()) : -1;
}

// First external entry of row r = 3I+a in one triangle and their count:
// L: [rp[r], dg[r] - a) (the a intra entries 3I..r-1 follow); U:
// [dg[r] + 1 + (2 - a), rp[r+1]) (after the intra entries r+1..3I+2).
static void cell_ext(const std::vector<int> &rp, const std::vector<int> &dg, int r,
                     bool forward, int &j0, int &cnt)
{
    const int a = r % 3;
    if (forward) {
        j0 = rp[r];
        cnt = dg[r] - a - rp[r];
    } else {
        j0 = dg[r] + 3 - a;
        cnt = rp[r + 1] - j0;
    }
}

// true when every cell has its full diagonal block and its three rows have
// the same external columns, whole blocks 3J..3J+2, in both triangles; wb =
// the most neighbour blocks of a cell in either triangle
static bool cell_pattern(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                         const std::vector<int> &dg, int &wb)
{
    wb = 0;
    if (n % 3)
        return false;
    for (int I = 0; I < n / 3; ++I) {
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c)
                if (find_pos(rp, ci, 3 * I + a, 3 * I + c) < 0)
                    return false;
        for (int f = 0; f < 2; ++f) {
            int j00, cnt0;
            cell_ext(rp, dg, 3 * I, f == 0, j00, cnt0);
            if (cnt0 < 0 || cnt0 % 3)
                return false;
            for (int a = 0; a < 3; ++a) {
                int j0, cnt;
                cell_ext(rp, dg, 3 * I + a, f == 0, j0, cnt);
                if (cnt != cnt0)
                    return false;
                for (int q = 0; q < cnt; ++q) {
                    const int J = ci[j00 + q - q % 3] / 3;
                    if (ci[j00 + q - q % 3] % 3 || ci[j0 + q] != 3 * J + q % 3 ||
                        (f == 0 ? J >= I : J <= I))
                        return false;
                }
            }
            wb = std::max(wb, cnt0 / 3);
        }
    }
    return true;
}

// Cell schedule of one triangle with WB neighbour-block slots per cell.
// Level meta: count = cells, off = cells before (record index), roff = rows
// before (3 per cell, the factorisation's row list).
static void build_sweep_cells(int nblocks, const std::vector<int> &bstart,
                              const std::vector<int> &rp, const std::vector<int> &ci,
                              const std::vector<int> &dg, bool forward, int WB, int es,
                              HostSweep &H)
{
    const int irec = round4(1 + WB), vrec = round4(9 * WB + (forward ? 3 : 6));
    H.W = WB;
    H.R = 3;
    H.irec = irec;
    H.vrec = vrec;
    H.blk_lev.assign(nblocks + 1, 0);
    H.meta.clear();
    H.rows.clear();
    H.row_glob.clear();
    H.col.clear();
    H.src.clear();
    H.max_count = 0;
    H.max_width = 0;
    const int n = bstart[nblocks], ncell = n / 3;
    std::vector<int> level(ncell, 0), cnt, first, order;
    int cells = 0;
    for (int b = 0; b < nblocks; ++b) {
        const int c0 = bstart[b] / 3, c1 = bstart[b + 1] / 3, b0 = bstart[b];
        const int zb = (bstart[b + 1] - b0) * es;   // the zero slots
        int maxl = -1;
        for (int q = 0; q < c1 - c0; ++q) {
            const int I = forward ? c0 + q : c1 - 1 - q;
            int j0, ne, l = 0;
            cell_ext(rp, dg, 3 * I, forward, j0, ne);
            for (int k = 0; k < ne; k += 3)
                l = std::max(l, level[ci[j0 + k] / 3] + 1);
            level[I] = l;
            maxl = std::max(maxl, l);
        }
        const int nlev = maxl + 1;
        cnt.assign(nlev + 1, 0);
        for (int I = c0; I < c1; ++I)
            cnt[level[I] + 1]++;
        for (int l = 0; l < nlev; ++l)
            cnt[l + 1] += cnt[l];
        first.assign(cnt.begin(), cnt.end());
        order.resize(c1 - c0);
        for (int I = c0; I < c1; ++I)
            order[first[level[I]]++] = I;
        H.blk_lev[b] = (int)H.meta.size();
        for (int l = 0; l < nlev; ++l) {
            LevelMeta m;
            m.count = cnt[l + 1] - cnt[l];
            m.off = cells;
            m.roff = 3 * cells;
            m.pad = 0;
            H.max_count = std::max(H.max_count, m.count);
            for (int t = 0; t < m.count; ++t) {
                const int I = order[cnt[l] + t];
                int j0, ne;
                cell_ext(rp, dg, 3 * I, forward, j0, ne);
                const int nbr = ne / 3;
                H.max_width = std::max(H.max_width, nbr);
                const size_t ib = H.col.size(), vb = H.src.size();
                H.col.resize(ib + irec, zb);
                H.src.resize(vb + vrec, -1);
                H.col[ib] = (3 * I - b0) * es;
                for (int k = 0; k < nbr; ++k)
                    H.col[ib + 1 + k] = (ci[j0 + 3 * k] - b0) * es;
                for (int a = 0; a < 3; ++a) {
                    const int r = 3 * I + a;
                    H.rows.push_back(r - b0);
                    H.row_glob.push_back(r);
                    int ja, na;
                    cell_ext(rp, dg, r, forward, ja, na);
                    for (int q = 0; q < na; ++q)
                        H.src[vb + 3 * WB * a + q] = ja + q;
                }
                const size_t iv = vb + 9 * WB;
                if (forward) {
                    H.src[iv + 0] = find_pos(rp, ci, 3 * I + 1, 3 * I);
                    H.src[iv + 1] = find_pos(rp, ci, 3 * I + 2, 3 * I);
                    H.src[iv + 2] = find_pos(rp, ci, 3 * I + 2, 3 * I + 1);
                } else {
                    H.src[iv + 0] = find_pos(rp, ci, 3 * I, 3 * I + 1);
                    H.src[iv + 1] = find_pos(rp, ci, 3 * I, 3 * I + 2);
                    H.src[iv + 2] = find_pos(rp, ci, 3 * I + 1, 3 * I + 2);
                    for (int a = 0; a < 3; ++a)
                        H.src[iv + 3 + a] = dg[3 * I + a];
                }
                ++cells;
            }
            H.meta.push_back(m);
        }
    }
    H.blk_lev[nblocks] = (int)H.meta.size();
    // trailing dummy record: the arrays are never empty
    H.col.resize(H.col.size() + irec, 0);
    H.src.resize(H.src.size() + vrec, -1);
    H.row_glob.push_back(0);
}

// Records of the cell factorisation (k_ilu_factor_cells); false when a
// block row has more than RS_CELL_NB blocks.
static bool build_cell_records(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                               const std::vector<int> &dg, std::vector<int> &rec)
{
    const int ncell = n / 3;
    rec.assign((size_t)ncell * RS_CREC, 0);
    std::vector<int> bcI;
    for (int I = 0; I < ncell; ++I) {
        const int r = 3 * I, nb = (rp[r + 1] - rp[r]) / 3, dI = (dg[r] - rp[r]) / 3;
        if (nb > RS_CELL_NB)
            return false;
        int *R = rec.data() + (size_t)RS_CREC * I;
        R[0] = rp[r];
        R[1] = nb;
        R[2] = dI;
        bcI.resize(nb);
        for (int q = 0; q < nb; ++q)
            bcI[q] = ci[rp[r] + 3 * q] / 3;
        for (int q = 0; q < dI; ++q) {
            const int J = bcI[q], rJ = 3 * J;
            const int nbJ = (rp[rJ + 1] - rp[rJ]) / 3, dJ = (dg[rJ] - rp[rJ]) / 3;
            int nm = 0;
            unsigned code[2] = {0, 0};
            for (int sJ = dJ + 1; sJ < nbJ; ++sJ) {
                const int K = ci[rp[rJ] + 3 * sJ] / 3;
                const auto it = std::lower_bound(bcI.begin(), bcI.end(), K);
                if (it == bcI.end() || *it != K)
                    continue;
                const int sI = (int)(it - bcI.begin());
                code[nm >> 2] |= (unsigned)((sI << 4) | sJ) << (8 * (nm & 3));
                ++nm;
            }
            R[4 + 4 * q] = rp[rJ];
            R[5 + 4 * q] = (nbJ << 16) | (dJ << 8) | nm;
            R[6 + 4 * q] = (int)code[0];
            R[7 + 4 * q] = (int)code[1];
        }
    }
    return true;
}

// MTJ's factor loop reads colind[rowptr[k+1]] when row `index` has a column
// beyond row k's last one (ILU.java:113-120) and updates the next row's
// first entry when that column matches.  That cannot happen when the pattern
// is structurally symmetric; other patterns are checked explicitly.
static bool ilu_quirk_active(int n, const std::vector<int> &rp,
                             const std::vector<int> &ci,
                             const std::vector<int> &dg)
{
    bool sym = true;
    for (int i = 0; i < n && sym; ++i)
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            int c = ci[j];
            if (!std::binary_search(ci.begin() + rp[c], ci.begin() + rp[c + 1], i)) {
                sym = false;
                break;
            }
        }
    if (sym)
        return false;
    const int nnz = rp[n];
    for (int k = 1; k < n; ++k)
        for (int i = rp[k]; i < dg[k]; ++i) {
            int index = ci[i];
            for (int j = dg[index] + 1, l = rp[k] + 1; j < rp[index + 1]; ++j) {
                while (l < rp[k + 1] && ci[l] < ci[j])
                    l++;
                if (l == rp[k + 1] && l < nnz && ci[l] == ci[j])
                    return true;
            }
        }
    return false;
}

// Sweeps, factor storage and scratch for the factor pattern (rp, ci) with
// diagonal positions dg, over the blocks in M->h_bstart: everything of the
// setup that follows from the factor's pattern (ILU(0): A's pattern or its
// block-diagonal part; ILUT: the pattern its factorisation produced).
void prec_setup_pattern(rsgpu_prec *M, const std::vector<int> &rp,
                        const std::vector<int> &ci, const std::vector<int> &dg, int bs)
{
    hipStream_t s = M->ctx->stream;
    const int n = M->n, nb = M->nblocks;
    const bool icc = M->kind == RSGPU_PREC_ICC;
    M->has_t = false;
    // L (unit lower, forward), U (upper, backward) in operation order: the
    // factor row's entries left and right of the diagonal
    const int es = M->precision == RSGPU_FP64 ? 8 : 4;
    // 3x3 block matrices with cell-aligned blocks: cell schedules, when the
    // pipelined cell sweep applies (at most 64 cells per level, the block in
    // LDS)
    M->cell = false;
    bool aligned = bs == 3 && !icc && getenv("RSGPU_ILU_CELLS") == nullptr &&
                   ilu_kernel_choice() != 3;
    for (int b = 0; aligned && b <= nb; ++b)
        aligned = M->h_bstart[b] % 3 == 0;
    int wb = 0;
    if (aligned && cell_pattern(n, rp, ci, dg, wb) && wb <= 8) {
        const int WB = wb <= 2 ? 2 : wb <= 4 ? wb : 8;   // the kernel's instances
        HostSweep HL, HU;
        build_sweep_cells(nb, M->h_bstart, rp, ci, dg, true, WB, es, HL);
        build_sweep_cells(nb, M->h_bstart, rp, ci, dg, false, WB, es, HU);
        int maxlev = 0;
        for (int b = 0; b < nb; ++b)
            maxlev = std::max(maxlev, (HL.blk_lev[b + 1] - HL.blk_lev[b]) +
                                          (HU.blk_lev[b + 1] - HU.blk_lev[b]));
        const size_t shm =
            cell_xv_bytes(M->max_block_rows, es) + (size_t)maxlev * sizeof(LevelMeta);
        // 32-bit byte offsets into the record arrays
        const size_t lim = (size_t)INT32_MAX;
        const bool addressable = HL.col.size() * 4 <= lim && HU.col.size() * 4 <= lim &&
                                 HL.src.size() * es <= lim && HU.src.size() * es <= lim;
        if (HL.max_count <= 64 && HU.max_count <= 64 && addressable &&
            shm <= (size_t)lds_limit(M->ctx->device)) {
            upload_sweep(HL, M->L, s, es);
            upload_sweep(HU, M->U, s, es);
            RS_HIP(hipStreamSynchronize(s));
            M->cell = true;
        }
    }
    M->cell_factor = false;
    if (M->cell && getenv("RSGPU_ILU_CELL_FACTOR") == nullptr) {
        std::vector<int> rec;
        if (build_cell_records(n, rp, ci, dg, rec)) {
            M->cell_rec.upload(rec, s);
            RS_HIP(hipStreamSynchronize(s));
            M->cell_factor = true;
        }
    }
    if (!M->cell_factor)
        M->cell_rec.release();
    if (!M->cell) {
        const RangeEntries eL{rp.data(), dg.data(), 0, 0, ci.data()};
        const RangeEntries eU{dg.data(), rp.data() + 1, 1, 0, ci.data()};
        const int W = pad_width(std::max(max_entries(eL, n), max_entries(eU, n)));
        HostSweep H;
        build_sweep(nb, M->h_bstart, eL, true, W, H);
        upload_sweep(H, M->L, s, es);
        build_sweep(nb, M->h_bstart, eU, false, W, H);
        upload_sweep(H, M->U, s, es);
        RS_HIP(hipStreamSynchronize(s));   // H's host arrays are released next
    }
    M->status.alloc(2);
    if (M->precision == RSGPU_FP64) {
        M->lu64.alloc(std::max(M->nnz, 1));
        M->ell64_L.alloc(M->L.ell_size);
        M->ell64_U.alloc(M->U.ell_size);
        if (!M->cell)
            M->dgU64.alloc(M->U.nrows_lev);
    } else {
        M->lu32.alloc(std::max(M->nnz, 1));
        M->ell32_L.alloc(M->L.ell_size);
        M->ell32_U.alloc(M->U.ell_size);
        if (!M->cell)
            M->dgU32.alloc(M->U.nrows_lev);
    }
    band_setup(M, rp, ci, dg);
    const int elem = M->precision == RSGPU_FP64 ? 8 : 4;
    M->lds_rows = RS_LDS_BYTES / elem - 1;
    if (M->max_block_rows > M->lds_rows) {
        if (M->precision == RSGPU_FP64)
            M->scratch64.alloc((size_t)n + nb);
        else
            M->scratch32.alloc((size_t)n + nb);
    }
    RS_HIP(hipStreamSynchronize(s));
    M->ready = true;
}

void prec_setup(rsgpu_prec *M, rsgpu_csr *A)
{
    hipStream_t s = M->ctx->stream;
    const int n = A->n;
    M->n = n;
    M->pattern_id = A->pattern_id;
    M->ready = false;
    M->has_t = false;
    const bool icc = M->kind == RSGPU_PREC_ICC;
    if (A->n != A->ncols)
        throw RsError{RSGPU_BAD_ARGUMENT, icc ? "ICC only applies to square matrices"
                                              : "ILU only applies to square matrices"};
    int nb = (M->kind == RSGPU_PREC_BLOCK_ILU0) ? M->nblocks : 1;
    if (nb < 1 || (n > 0 && nb > n))
        throw RsError{RSGPU_BAD_ARGUMENT, "block count must be in [1, n]"};
    if (n == 0)
        nb = 1;
    M->nblocks = nb;
    M->h_bstart.resize(nb + 1);
    if (!M->user_bstart.empty()) {
        const std::vector<int> &u = M->user_bstart;
        bool ok = (int)u.size() == nb + 1 && u[0] == 0 && u[nb] == n;
        for (int b = 0; ok && b < nb; ++b)
            ok = u[b] < u[b + 1];
        if (!ok)
            throw RsError{RSGPU_BAD_ARGUMENT,
                          "block boundaries must rise strictly from 0 to the matrix size"};
        M->h_bstart = u;
    } else {
        for (int b = 0; b <= nb; ++b)
            M->h_bstart[b] = (int)(((long long)n * b) / nb);
    }
    M->max_block_rows = 0;
    for (int b = 0; b < nb; ++b)
        M->max_block_rows =
            std::max(M->max_block_rows, M->h_bstart[b + 1] - M->h_bstart[b]);
    M->bstart.upload(M->h_bstart, s);

    // factor pattern: the block-diagonal part, in A's order
    std::vector<int> rp(n + 1, 0), ci, src;
    ci.reserve(A->nnz);
    src.reserve(A->nnz);
    bool ident = true;
    for (int b = 0; b < nb; ++b) {
        const int b0 = M->h_bstart[b], b1 = M->h_bstart[b + 1];
        for (int r = b0; r < b1; ++r) {
            for (int j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; ++j) {
                int c = A->h_colind[j];
                if (c >= b0 && c < b1) {
                    ci.push_back(c);
                    src.push_back(j);
                } else {
                    ident = false;
                }
            }
            rp[r + 1] = (int)ci.size();
        }
    }
    M->nnz = (int)ci.size();
    M->h_rowptr = rp;
    M->h_colind = ci;
    M->src_identity = ident;
    // diagonal positions (ILU.findDiagonalIndices, ILU.java:127-140)
    std::vector<int> dg(n, -1);
    for (int k = 0; k < n; ++k) {
        auto it = std::lower_bound(ci.begin() + rp[k], ci.begin() + rp[k + 1], k);
        if (it == ci.begin() + rp[k + 1] || *it != k)
            throw RsError{RSGPU_MISSING_DIAGONAL,
                          "Missing diagonal entry on row " + std::to_string(k + 1)};
        dg[k] = (int)(it - ci.begin());
    }
    if (!icc && ilu_quirk_active(n, rp, ci, dg))
        throw RsError{RSGPU_BAD_ARGUMENT,
                      "pattern makes MTJ's ILU update the next row's first entry; "
                      "not supported"};
    M->rowptr.upload(rp, s);
    M->colind.upload(ci, s);
    M->diag.upload(dg, s);
    if (!ident)
        M->src.upload(src, s);

    prec_setup_pattern(M, rp, ci, dg, A->bs);
}

// Line-wavefront form of a block [b0, b1): Ly lines of Lx rows in natural
// order (row b0 + Lx j + i), every in-block coupling of row (i, j) one of
// L: (i-1, j-1), (i, j-1), (i+1, j-1), (i-1, j) and U: (i+1, j), (i-1, j+1),
// (i, j+1), (i+1, j+1) -- in ascending column order the slots 0..3 of each
// part.  Lx is read off the first row (its first column past b0 + 1); Lx >= 3
// keeps the eight offsets distinct.  pos receives per row the factor
// positions of L0..L3, the diagonal and U0..U3 (-1 absent).
static bool band_block(int b0, int b1, const std::vector<int> &rp, const std::vector<int> &ci,
                       const std::vector<int> &dg, int &Lx, int &Ly, std::vector<int> &pos)
{
    const int nbr = b1 - b0;
    Lx = nbr;
    for (int q = dg[b0] + 1; q < rp[b0 + 1]; ++q)
        if (ci[q] >= b0 + 2) {
            Lx = ci[q] - b0;
            break;
        }
    if (Lx < 3 || nbr % Lx != 0)
        return false;
    Ly = nbr / Lx;
    if (Ly > 64)
        return false;
    for (int r = b0; r < b1; ++r) {
        const int i = (r - b0) % Lx, j = (r - b0) / Lx;
        int *p = &pos[(size_t)9 * r];
        for (int k = 0; k < 9; ++k)
            p[k] = -1;
        p[4] = dg[r];
        for (int q = rp[r]; q < rp[r + 1]; ++q) {
            const int d = ci[q] - r;
            int slot = -1;
            if (d == -Lx - 1 && i > 0 && j > 0)
                slot = 0;
            else if (d == -Lx && j > 0)
                slot = 1;
            else if (d == -Lx + 1 && i < Lx - 1 && j > 0)
                slot = 2;
            else if (d == -1 && i > 0)
                slot = 3;
            else if (d == 0)
                continue;
            else if (d == 1 && i < Lx - 1)
                slot = 5;
            else if (d == Lx - 1 && i > 0 && j < Ly - 1)
                slot = 6;
            else if (d == Lx && j < Ly - 1)
                slot = 7;
            else if (d == Lx + 1 && i < Lx - 1 && j < Ly - 1)
                slot = 8;
            if (slot < 0 || p[slot] >= 0)
                return false;
            p[slot] = q;
        }
    }
    return true;
}

// slots per line in the full blocks' LDS layout: the sweep's step count
// Lx + 2 Ly - 2 rounded up to whole 16-step rounds
__host__ __device__ inline int band_slot_pitch(int Lx, int Ly)
{
    return (Lx + 2 * Ly - 2 + 15) & ~15;
}

// RSGPU_ILU_BAND=0 keeps the level schedules for every block
static void band_setup(rsgpu_prec *M, const std::vector<int> &rp, const std::vector<int> &ci,
                       const std::vector<int> &dg)
{
    M->band = false;
    M->band_factor = false;
    M->band_pick = -1;
    M->band_slots = 0;
    M->h_band_dims.clear();
    M->band_dims.release();
    M->band_pos.release();
    M->band_mask.release();
    M->band64.release();
    M->band32.release();
    M->band_rd.release();
    const char *e = getenv("RSGPU_ILU_BAND");
    if ((e && e[0] == '0') || M->cell || ilu_kernel_choice() == 3 || M->n == 0 ||
        M->kind == RSGPU_PREC_ICC)
        return;
    const int n = M->n, nb = M->nblocks;
    std::vector<int> pos((size_t)9 * n);
    std::vector<int2> dims(nb);
    for (int b = 0; b < nb; ++b) {
        int Lx, Ly;
        if (!band_block(M->h_bstart[b], M->h_bstart[b + 1], rp, ci, dg, Lx, Ly, pos))
            return;
        dims[b] = make_int2(Lx, Ly);
    }
    std::vector<unsigned char> mask(n);
    for (int r = 0; r < n; ++r) {
        unsigned m = 0;
        for (int k = 0; k < 4; ++k) {
            if (pos[(size_t)9 * r + k] >= 0)
                m |= 1u << k;
            if (pos[(size_t)9 * r + 5 + k] >= 0)
                m |= 16u << k;
        }
        mask[r] = (unsigned char)m;
    }
    // blocks whose every row holds all of its 9-point neighbours compute the
    // masks from the row position (marked Ly > 0); others read them (-Ly)
    for (int b = 0; b < nb; ++b) {
        const int b0 = M->h_bstart[b], Lx = dims[b].x, Ly = dims[b].y;
        bool full = true;
        for (int r = b0; full && r < M->h_bstart[b + 1]; ++r) {
            const int i = (r - b0) % Lx, j = (r - b0) / Lx;
            const bool w = i > 0, e = i < Lx - 1, s_ = j > 0, nn = j < Ly - 1;
            const unsigned f = (w && s_ ? 1u : 0u) | (s_ ? 2u : 0u) | (e && s_ ? 4u : 0u) |
                               (w ? 8u : 0u) | (e ? 16u : 0u) | (w && nn ? 32u : 0u) |
                               (nn ? 64u : 0u) | (e && nn ? 128u : 0u);
            full = mask[r] == f;
        }
        if (!full)
            dims[b].y = -Ly;
        else
            M->band_slots = std::max(M->band_slots, Ly * band_slot_pitch(Lx, Ly));
    }
    hipStream_t s = M->ctx->stream;
    M->h_band_dims = dims;
    M->band_dims.upload(dims, s);
    M->band_pos.upload(pos, s);
    // masks and records padded by RS_BAND_PAD zero rows on both sides
    std::vector<unsigned char> mpad(mask.size() + 2 * RS_BAND_PAD, 0);
    std::copy(mask.begin(), mask.end(), mpad.begin() + RS_BAND_PAD);
    M->band_mask.upload(mpad, s);
    const size_t np = (size_t)n + 2 * RS_BAND_PAD;
    if (M->precision == RSGPU_FP64) {
        M->band64.alloc(9 * np);
        M->band64.zero(s);
    } else {
        M->band32.alloc(9 * np);
        M->band32.zero(s);
        M->band_rd.alloc(np);
        M->band_rd.zero(s);
    }
    RS_HIP(hipStreamSynchronize(s));
    M->band = true;
    // the line-wavefront factorisation: every block full (RSGPU_ILU_BAND_FACTOR=0
    // keeps the level kernel)
    const char *ef = getenv("RSGPU_ILU_BAND_FACTOR");
    bool all_full = !(ef && ef[0] == '0') && M->kind != RSGPU_PREC_ILUT;
    for (int b = 0; all_full && b < nb; ++b)
        all_full = dims[b].y > 0;
    M->band_factor = all_full;
}

// Transposed sweeps: U' (forward, subtract in ascending source row, divide)
// then L' (backward, subtract in descending source row).
static void prec_setup_transpose(rsgpu_prec *M, int wmin = 0)
{
    if (M->has_t)
        return;
    hipStream_t s = M->ctx->stream;
    const int n = M->n;
    const auto &rp = M->h_rowptr, &ci = M->h_colind;
    std::vector<int> dg(n);
    for (int k = 0; k < n; ++k)
        dg[k] = (int)(std::lower_bound(ci.begin() + rp[k], ci.begin() + rp[k + 1], k) -
                      ci.begin());
    // U' row c: the rows i < c with U(i, c) != 0, ascending i; L' row c: the
    // rows i > c with L(i, c) != 0, descending i (the column sweeps' order)
    FlatEntries eUt, eLt;
    eUt.ptr.assign(n + 1, 0);
    eLt.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        for (int j = dg[i] + 1; j < rp[i + 1]; ++j)
            eUt.ptr[ci[j] + 1]++;
        for (int j = rp[i]; j < dg[i]; ++j)
            eLt.ptr[ci[j] + 1]++;
    }
    for (int c = 0; c < n; ++c) {
        eUt.ptr[c + 1] += eUt.ptr[c];
        eLt.ptr[c + 1] += eLt.ptr[c];
    }
    eUt.cols.resize(eUt.ptr[n]);
    eUt.poss.resize(eUt.ptr[n]);
    eLt.cols.resize(eLt.ptr[n]);
    eLt.poss.resize(eLt.ptr[n]);
    {
        std::vector<int> fu(eUt.ptr.begin(), eUt.ptr.end() - 1);
        std::vector<int> fl(eLt.ptr.begin(), eLt.ptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int j = dg[i] + 1; j < rp[i + 1]; ++j) {
                const int q = fu[ci[j]]++;
                eUt.cols[q] = i;
                eUt.poss[q] = j;
            }
        for (int i = n - 1; i >= 0; --i)
            for (int j = rp[i]; j < dg[i]; ++j) {
                const int q = fl[ci[j]]++;
                eLt.cols[q] = i;
                eLt.poss[q] = j;
            }
    }
    const int W = pad_width(std::max({max_entries(eUt, n), max_entries(eLt, n), wmin}));
    HostSweep H;
    build_sweep(M->nblocks, M->h_bstart, eUt, true, W, H);
    upload_sweep(H, M->Ut, s, M->precision == RSGPU_FP64 ? 8 : 4);
    build_sweep(M->nblocks, M->h_bstart, eLt, false, W, H);
    upload_sweep(H, M->Lt, s, M->precision == RSGPU_FP64 ? 8 : 4);
    M->ell64_Ut.alloc(M->Ut.ell_size);
    M->ell64_Lt.alloc(M->Lt.ell_size);
    M->dgUt64.alloc(M->Ut.nrows_lev);
    if (!M->scratch64.p && M->max_block_rows > RS_LDS_BYTES / 8 - 1)
        M->scratch64.alloc((size_t)n + M->nblocks);
    RS_HIP(hipStreamSynchronize(s));
    M->has_t = true;
}

// ---------------------------------------------------------------- factor
template <typename T>
__global__ void k_load_values(const double *__restrict__ a, const int *__restrict__ src,
                              T *__restrict__ lu, int nnz)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz)
        lu[i] = (T)(src ? a[src[i]] : a[i]);
}

// One workgroup per block; levels in order; one lane per row performing
// the IKJ update sequence of ILU.java:99-122 for that row.
// With a cell schedule (R = 3) a level entry is a cell and its lane
// factors the cell's three rows in order.
template <typename T>
__global__ void __launch_bounds__(256)
    k_ilu_factor(const int *__restrict__ bstart, const int *__restrict__ blk_lev,
                 const LevelMeta *__restrict__ meta, const int *__restrict__ rows,
                 const int *__restrict__ rowptr, const int *__restrict__ colind,
                 const int *__restrict__ diag, T *lu, int *status, int R)
{
    const int b = blockIdx.x;
    const int b0 = bstart[b];
    for (int l = blk_lev[b]; l < blk_lev[b + 1]; ++l) {
        const LevelMeta m = meta[l];
        for (int tr = threadIdx.x * R; tr < m.count * R; tr += blockDim.x * R)
            for (int t = tr; t < tr + R; ++t) {
            const int k = b0 + rows[m.roff + t];
            const int rk0 = rowptr[k], rk1 = rowptr[k + 1], dk = diag[k];
            for (int i = rk0; i < dk; ++i) {
                const int index = colind[i];
                const T LUii = lu[diag[index]];
                if (LUii == T(0)) {
                    atomicMin(&status[1], i);
                    atomicMax(&status[0], (int)RSGPU_ZERO_PIVOT);
                    break;
                }
                const T LUki = (lu[i] /= LUii);
                const int je = rowptr[index + 1];
                int l2 = rk0 + 1;
                for (int j = diag[index] + 1; j < je; ++j) {
                    const int cj = colind[j];
                    while (l2 < rk1 && colind[l2] < cj)
                        l2++;
                    if (l2 < rk1 && colind[l2] == cj)
                        lu[l2] -= LUki * lu[j];
                }
            }
            }
        __syncthreads();
    }
}

// Cell factorisation (3x3 block matrices with a cell schedule).  The IKJ
// sequence of ILU.java:99-122 for row k = 3I+a visits its L entries in
// column order: the entries of each neighbour block J < I (columns 3J+c,
// c = 0..2), then the diagonal block's (3I+c, c < a).  Entry (k, 3J+c) is
// divided by the pivot of row 3J+c and then updates row k at the columns of
// row 3J+c's U part that row k has: block J's columns 3J+c' (c' > c), then
// every block K > J of row J that is also a block of row I.  With whole
// blocks these positions follow from the block slots, so the lane of cell I
// needs no column search: its three rows sit in LDS, the host lists per
// neighbour block J the slot pairs (slot of K in row I, slot in row J), and
// each J's pivots and U entries are loaded once for the three rows.  Rows
// are independent until the diagonal block, so the lane runs every row's
// external steps first (each row's own order is unchanged: J ascending,
// then c), then the diagonal-block steps of rows 1 and 2, which read rows 0
// and 1 after they are final.  Per target the updates arrive in the
// reference's order, so the factors are bit-identical.
//
// Cell record (ints, RS_CREC): [rp(3I), nb(I), slot of I, -] then per
// neighbour slot s < slot of I: [rp(3J), nb(J) << 16 | slot of J in row J
// << 8 | matches, 4 match bytes, 4 match bytes], a match byte = slot in row
// I << 4 | slot in row J, in ascending slot of row J.

template <typename T>
__global__ void __launch_bounds__(64)
    k_ilu_factor_cells(const int *__restrict__ bstart, const int *__restrict__ blk_lev,
                       const LevelMeta *__restrict__ meta, const int *__restrict__ rows,
                       const int *__restrict__ crec, T *lu, int *status)
{
    __shared__ T own[3 * RS_CQ * 64];
    const int t = threadIdx.x;
#define RS_OWN(a, q) own[((a) * RS_CQ + (q)) * 64 + t]
    const int blk = blockIdx.x;
    const int b0 = bstart[blk];
    for (int l = blk_lev[blk]; l < blk_lev[blk + 1]; ++l) {
        const LevelMeta m = meta[l];
        if (t < m.count) {
            const int I = (b0 + rows[m.roff + 3 * t]) / 3;
            const int *rec = crec + (size_t)RS_CREC * I;
            const int rp3 = rec[0], nb = rec[1], dI = rec[2], n3 = 3 * nb;
            // a row's loads are issued together (indices past the row clamp
            // to its last entry; only the LDS stores are predicated): a
            // guarded load per slot waits for memory on its own
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                T row[RS_CQ];
#pragma unroll
                for (int q = 0; q < RS_CQ; ++q)
                    row[q] = lu[rp3 + a * n3 + min(q, n3 - 1)];
#pragma unroll
                for (int q = 0; q < RS_CQ; ++q)
                    if (q < n3)
                        RS_OWN(a, q) = row[q];
            }
            bool bad[3] = {false, false, false};
            for (int s = 0; s < dI; ++s) {
                const int rpJ = rec[4 + 4 * s], w = rec[5 + 4 * s];
                const int n3J = 3 * (w >> 16), dJ = (w >> 8) & 255, nm = w & 255;
                const int dpos = rpJ + 3 * dJ;
                // pivots and block J's U entries of rows 3J, 3J+1
                T piv[3];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    piv[c] = lu[dpos + c * n3J + c];
                const T u01 = lu[dpos + 1], u02 = lu[dpos + 2], u12 = lu[dpos + n3J + 2];
                T Lv[3][3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        Lv[a][c] = T(0);
                        if (bad[a])
                            continue;
                        const int iq = 3 * s + c;
                        if (piv[c] == T(0)) {
                            atomicMin(&status[1], rp3 + a * n3 + iq);
                            atomicMax(&status[0], (int)RSGPU_ZERO_PIVOT);
                            bad[a] = true;
                            continue;
                        }
                        const T L = RS_OWN(a, iq) / piv[c];
                        RS_OWN(a, iq) = L;
                        Lv[a][c] = L;
                        if (c == 0) {
                            RS_OWN(a, iq + 1) -= L * u01;
                            RS_OWN(a, iq + 2) -= L * u02;
                        } else if (c == 1) {
                            RS_OWN(a, iq + 1) -= L * u12;
                        }
                    }
                }
                // blocks K > J of row J that row I has
                for (int mi = 0; mi < nm; ++mi) {
                    const int code = (rec[6 + 4 * s + (mi >> 2)] >> (8 * (mi & 3))) & 255;
                    const int sI = code >> 4, sJ = code & 15;
                    T src[3][3];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int e = 0; e < 3; ++e)
                            src[c][e] = lu[rpJ + c * n3J + 3 * sJ + e];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        if (bad[a])
                            continue;
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int e = 0; e < 3; ++e)
                                RS_OWN(a, 3 * sI + e) -= Lv[a][c] * src[c][e];
                    }
                }
            }
            // the diagonal block: rows 1 and 2 against the final rows above
#pragma unroll
            for (int a = 1; a < 3; ++a) {
                for (int c = 0; c < a && !bad[a]; ++c) {
                    const int iq = 3 * dI + c;
                    const T pv = RS_OWN(c, iq);
                    if (pv == T(0)) {
                        atomicMin(&status[1], rp3 + a * n3 + iq);
                        atomicMax(&status[0], (int)RSGPU_ZERO_PIVOT);
                        bad[a] = true;
                        break;
                    }
                    const T L = RS_OWN(a, iq) / pv;
                    RS_OWN(a, iq) = L;
                    for (int q = iq + 1; q < n3; ++q)
                        RS_OWN(a, q) -= L * RS_OWN(c, q);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int q = 0; q < RS_CQ; ++q)
                    if (q < n3)
                        lu[rp3 + a * n3 + q] = RS_OWN(a, q);
        }
        __syncthreads();
    }
#undef RS_OWN
}

// Padding slots hold the identity of the sweep's update: -0.0 for the
// sum form (s + (-0)*0 == s for every s, signed zeros included) and +0.0
// for the subtract form (a - (+0)*0 == a).
template <typename T>
__global__ void k_gather_ell(const int *__restrict__ src, const T *__restrict__ lu,
                             T *__restrict__ ell, size_t n, T pad)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        int s = src[i];
        ell[i] = s >= 0 ? lu[s] : pad;
    }
}

// level-ordered diagonal: dgl[p] = LU(r, r) for the p-th level row r
template <typename T>
__global__ void k_gather_rowdiag(const int *__restrict__ row_glob,
                                 const int *__restrict__ diag, const T *__restrict__ lu,
                                 T *__restrict__ dgl, int nrows)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrows)
        dgl[i] = lu[diag[row_glob[i]]];
}

// ---------------------------------------------------------------- sweeps
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

enum { FORM_SUM = 0, FORM_SUBSEQ = 1 };

// one row update: FORM_SUM x_r = (x_r - sum_s v_s x_c) [/ d];
// FORM_SUBSEQ x_r = (((x_r - v_0 x_c0) - v_1 x_c1) ...) [/ d]
template <typename T, int FORM, bool DIV>
__device__ __forceinline__ T row_update(T rhs, const T *v, const int *c, int W, const T *xv,
                                        T d)
{
    T acc;
    if (FORM == FORM_SUM) {
        T sum = 0;
        for (int s = 0; s < W; ++s)
            sum += v[s] * xv[c[s]];
        acc = rhs - sum;
    } else {
        acc = rhs;
        for (int s = 0; s < W; ++s)
            acc -= v[s] * xv[c[s]];
    }
    if (DIV)
        acc = acc / d;
    return acc;
}

// Generic in-place sweep (any W, any level width): LDS or global vector.
template <typename T, int FORM, bool DIV, bool LDS>
__device__ void sweep_generic(int lb, int le, int W, const LevelMeta *__restrict__ meta,
                              const int *__restrict__ rows, const int *__restrict__ ecol,
                              const T *__restrict__ eval, const T *__restrict__ dgl, T *xv)
{
    for (int l = lb; l < le; ++l) {
        const LevelMeta m = meta[l];
        for (int t = threadIdx.x; t < m.count; t += blockDim.x) {
            const int r = rows[m.roff + t];
            const size_t q = (size_t)m.off + (size_t)t * W;
            const T d = DIV ? dgl[m.roff + t] : T(1);
            xv[r] = row_update<T, FORM, DIV>(xv[r], eval + q, ecol + q, W, xv, d);
        }
        __syncthreads();
    }
}

// dst[0..n) = src[0..n) by the workgroup with 16 loads per thread in flight
// per batch: a plain strided loop waits for each load before issuing the
// next, one memory round trip per blockDim.x elements.  Indices past the end
// clamp to the last element (n >= 1); only the stores are predicated.
template <typename E>
__device__ __forceinline__ void wg_copy(E *__restrict__ dst, const E *__restrict__ src, int n)
{
    const int nt = (int)blockDim.x;
    for (int r0 = threadIdx.x; r0 < n; r0 += nt * 16) {
        E v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            v[u] = src[min(r0 + nt * u, n - 1)];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (r0 + nt * u < n)
                dst[r0 + nt * u] = v[u];
    }
}

// MODE 0: ILU apply (L: sum form; U: sum form, divide); 1: ILU transApply
// (U': subtract form, divide; L': subtract form); 2: ICC apply (R':
// subtract form, divide; R: sum form, divide)
#define RS_MODE_ILU 0
#define RS_MODE_ILU_T 1
#define RS_MODE_ICC 2

template <typename T, bool LDS, int MODE>
__global__ void __launch_bounds__(256)
    k_ilu_apply(const int *__restrict__ bstart, int W,
                const int *__restrict__ blkA, const LevelMeta *__restrict__ metaA,
                const int *__restrict__ rowsA, const int *__restrict__ colA,
                const T *__restrict__ valA, const T *__restrict__ dglA,
                const int *__restrict__ blkB, const LevelMeta *__restrict__ metaB,
                const int *__restrict__ rowsB, const int *__restrict__ colB,
                const T *__restrict__ valB, const T *__restrict__ dglB,
                const T *__restrict__ b, T *__restrict__ x, T *scratch,
                const int *skip1, const int *skip2, BlkSkip bskip)
{
    if ((skip1 && *skip1) || (skip2 && *skip2))
        return;
    if (blk_skipped(bskip, blockIdx.x))
        return;
    extern __shared__ unsigned char rs_smem[];
    const int blk = blockIdx.x;
    const int b0 = bstart[blk], b1 = bstart[blk + 1], nb = b1 - b0;
    T *xv = LDS ? (T *)rs_smem : scratch + b0 + blk;
    if (nb > 0)
        wg_copy(xv, b + b0, nb);
    if (threadIdx.x == 0)
        xv[nb] = T(0);
    __syncthreads();
    if (MODE == RS_MODE_ILU) {
        sweep_generic<T, FORM_SUM, false, LDS>(blkA[blk], blkA[blk + 1], W, metaA, rowsA, colA,
                                               valA, dglA, xv);
        sweep_generic<T, FORM_SUM, true, LDS>(blkB[blk], blkB[blk + 1], W, metaB, rowsB, colB,
                                              valB, dglB, xv);
    } else if (MODE == RS_MODE_ILU_T) {
        sweep_generic<T, FORM_SUBSEQ, true, LDS>(blkA[blk], blkA[blk + 1], W, metaA, rowsA,
                                                 colA, valA, dglA, xv);
        sweep_generic<T, FORM_SUBSEQ, false, LDS>(blkB[blk], blkB[blk + 1], W, metaB, rowsB,
                                                  colB, valB, dglB, xv);
    } else {
        sweep_generic<T, FORM_SUBSEQ, true, LDS>(blkA[blk], blkA[blk + 1], W, metaA, rowsA,
                                                 colA, valA, dglA, xv);
        sweep_generic<T, FORM_SUM, true, LDS>(blkB[blk], blkB[blk + 1], W, metaB, rowsB, colB,
                                              valB, dglB, xv);
    }
    for (int r = threadIdx.x; r < nb; r += blockDim.x)
        x[b0 + r] = xv[r];
}

// ---------------------------------------------------------------- pipelined
template <typename T, int W>
struct alignas(16) VRow {
    T v[W];
};
template <int W>
struct alignas(16) CRow {
    int c[W];
};

template <typename T, int W>
struct LevelRegs {
    int r;      // LDS byte offset of the row (inactive lanes: a real row)
    bool act;
    T d;
    VRow<T, W> v;
    CRow<W> c;  // LDS byte offsets of the operands
};

// One sweep as the pipelined kernel reads it: level ranges and metadata,
// plus 32-bit buffer resources over the level-ordered row byte offsets,
// operand byte offsets, values and U diagonals (loads beyond a resource's
// size return zero instead of faulting).
struct PfSweep {
    const int *blk;
    const LevelMeta *meta;
    const int *rowsb;
    const int *colb;
    const void *val;
    const void *dgl;
    int rows_bytes, col_bytes, val_bytes, dgl_bytes;
};

struct PfRsrc {
    __amdgpu_buffer_rsrc_t rows, cols, vals, dgl;
};

// buffer resource word 3 for raw (unformatted, stride 0) buffers on gfx9
#define RS_BUF_WORD3 0x00020000

__device__ __forceinline__ PfRsrc pf_rsrc(const PfSweep &S)
{
    PfRsrc R;
    R.rows = __builtin_amdgcn_make_buffer_rsrc((void *)S.rowsb, (short)0, S.rows_bytes,
                                               RS_BUF_WORD3);
    R.cols = __builtin_amdgcn_make_buffer_rsrc((void *)S.colb, (short)0, S.col_bytes,
                                               RS_BUF_WORD3);
    R.vals = __builtin_amdgcn_make_buffer_rsrc((void *)S.val, (short)0, S.val_bytes,
                                               RS_BUF_WORD3);
    R.dgl = __builtin_amdgcn_make_buffer_rsrc((void *)S.dgl, (short)0, S.dgl_bytes,
                                              RS_BUF_WORD3);
    return R;
}

typedef int rs_v4i __attribute__((ext_vector_type(4)));
typedef int rs_v2i __attribute__((ext_vector_type(2)));

// NB bytes (a multiple of 16) from a raw buffer into dst
template <int NB>
__device__ __forceinline__ void buf_load(void *dst, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    static_assert(NB % 16 == 0, "row segments are whole 16-byte chunks");
#pragma unroll
    for (int k = 0; k < NB / 16; ++k) {
        const rs_v4i q = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16 * k, soff, 0);
        __builtin_memcpy((char *)dst + 16 * k, &q, 16);
    }
}

// All loads of a level are independent of each other and of the vector:
// row offset, the lane's W operand offsets and values, level-ordered
// diagonal (in that order, so one wait covers every gather address).  The
// level metadata is wave-uniform: scalar base offsets, 32-bit lane
// offsets.  Inactive lanes (t >= count, or every lane of a level index past
// the sweep's end) redo a real row of the clamped level and write the
// result to the junk slot, so neither loads nor updates sit behind
// branches (branches make the compiler's wait counting fall back to
// waiting for every load, and let it sink prefetches next to their use).
template <typename T, int W, bool DIV>
__device__ __forceinline__ void level_load(const LevelMeta &m, bool valid, const PfRsrc &rs,
                                           LevelRegs<T, W> &R)
{
    const int count = __builtin_amdgcn_readfirstlane(m.count);
    const int off = __builtin_amdgcn_readfirstlane(m.off);
    const int roff = __builtin_amdgcn_readfirstlane(m.roff);
    const int t = threadIdx.x;
    const int tt = min(t, count - 1);
    // nothing here may depend on a loaded value, or the prefetch would
    // wait for itself
    R.act = valid && t < count;
    R.r = __builtin_amdgcn_raw_buffer_load_b32(rs.rows, tt * 4, roff * 4, 0);
    buf_load<W * 4>(&R.c, rs.cols, tt * (W * 4), off * 4);
    buf_load<W * (int)sizeof(T)>(&R.v, rs.vals, tt * (W * (int)sizeof(T)),
                                 off * (int)sizeof(T));
    if (DIV) {
        if (sizeof(T) == 8) {
            const rs_v2i q = __builtin_amdgcn_raw_buffer_load_b64(rs.dgl, tt * 8, roff * 8, 0);
            __builtin_memcpy(&R.d, &q, 8);
        } else {
            const int q = __builtin_amdgcn_raw_buffer_load_b32(rs.dgl, tt * 4, roff * 4, 0);
            __builtin_memcpy(&R.d, &q, 4);
        }
    }
}

// Every gathered operand of a level passes through one empty asm statement
// as an in/out register: all the level's LDS reads are issued before it and
// one wait covers them, instead of the scheduler interleaving arithmetic
// between batches of reads (one LDS round trip per level instead of two).
// The statement emits no instruction and performs no memory access.
template <typename T, int W>
__device__ __forceinline__ void pin_gathers(T (&x)[W], T &r)
{
#define RS_P(i) "+v"(x[i])
    if constexpr (W == 4)
        asm volatile("" : "+v"(r), RS_P(0), RS_P(1), RS_P(2), RS_P(3));
    else if constexpr (W == 8)
        asm volatile("" : "+v"(r), RS_P(0), RS_P(1), RS_P(2), RS_P(3), RS_P(4), RS_P(5), RS_P(6),
                     RS_P(7));
    else if constexpr (W == 12)
        asm volatile("" : "+v"(r), RS_P(0), RS_P(1), RS_P(2), RS_P(3), RS_P(4), RS_P(5), RS_P(6),
                     RS_P(7), RS_P(8), RS_P(9), RS_P(10), RS_P(11));
    else if constexpr (W == 16)
        asm volatile("" : "+v"(r), RS_P(0), RS_P(1), RS_P(2), RS_P(3), RS_P(4), RS_P(5), RS_P(6),
                     RS_P(7), RS_P(8), RS_P(9), RS_P(10), RS_P(11), RS_P(12), RS_P(13), RS_P(14),
                     RS_P(15));
    else
        static_assert(W == 4, "no pin for this width");
#undef RS_P
}

template <typename T, int W>
struct Gathered {
    T rhs;
    T xs[W];
};

template <typename T>
__device__ __forceinline__ T &lds_at(unsigned char *base, int byte_off)
{
    return *reinterpret_cast<T *>(base + byte_off);
}

// Issue the level's LDS reads: the row's right-hand side and its operands.
template <typename T, int W>
__device__ __forceinline__ void level_gather(const LevelRegs<T, W> &R, unsigned char *xv,
                                             Gathered<T, W> &g)
{
    g.rhs = lds_at<T>(xv, R.r);
#pragma unroll
    for (int s = 0; s < W; ++s)
        g.xs[s] = lds_at<T>(xv, R.c.c[s]);
}

// Wait for the gathers once, then the row's operation sequence and store.
template <typename T, int W, int FORM, bool DIV>
__device__ __forceinline__ void level_finish(const LevelRegs<T, W> &R, Gathered<T, W> &g,
                                             unsigned char *xv, int junkb)
{
    pin_gathers<T, W>(g.xs, g.rhs);
    const int wr = R.act ? R.r : junkb;
    const T rhs = g.rhs;
    T acc;
    if (FORM == FORM_SUM) {
        T sum = 0;
#pragma unroll
        for (int s = 0; s < W; ++s)
            sum += R.v.v[s] * g.xs[s];
        acc = rhs - sum;
    } else {
        acc = rhs;
#pragma unroll
        for (int s = 0; s < W; ++s)
            acc -= R.v.v[s] * g.xs[s];
    }
    if (DIV)
        acc = acc / R.d;
    lds_at<T>(xv, wr) = acc;
}

// Loads in flight per level (row offset, operand offsets, values, U diagonal).
template <typename T, int W, bool DIV>
constexpr int level_loads()
{
    return 1 + (W * 4 + 15) / 16 + (W * (int)sizeof(T) + 15) / 16 + (DIV ? 1 : 0);
}

// Prefetch depth: levels whose loads are in flight while one computes.  A
// memory latency spans several levels of compute, so narrow rows keep up
// to 8 levels in flight; the depth is capped by the 63 loads the wave's
// counter can track and by the registers one level's operands occupy.
template <typename T, int W>
constexpr int pf_depth()
{
    constexpr int loads = level_loads<T, W, true>();
    constexpr int regs = 2 + W + W * (int)sizeof(T) / 4 + (int)sizeof(T) / 4;
    int d = 8;
    while (d > 2 && (d * loads > 60 || d * regs > 210))
        --d;
    return d;
}

// DEPTH register sets in rotation: while level l computes, levels
// l+1..l+DEPTH-1 are loading.  The stage loop is fully unrolled, so each
// set is a fixed group of registers (no copies), and the loop body is one
// basic block: level indices past the end load the last level again and
// compute into the junk slot, so no branch separates a prefetch from the
// stage it belongs to.  The level metadata a prefetch needs is read from
// LDS one stage earlier, so its LDS round trip hides behind the previous
// level's compute.  The kernel is one wave and a wave's LDS operations
// complete in issue order, so levels are separated by a wave-scope fence
// (no wait for the store to land).
template <typename T, int W, int FORM, bool DIV>
__device__ __forceinline__ void sweep_pf(int nl, const LevelMeta *sm, const PfRsrc &rs,
                                         unsigned char *xv, int junkb)
{
    constexpr int D = pf_depth<T, W>();
    if (nl <= 0)
        return;
    const int last = nl - 1;
    LevelRegs<T, W> r[D];
    // prologue in the steady state's issue order (level by level), so the
    // wait counts at the loop head are the same from either predecessor
#pragma unroll
    for (int k = 0; k < D - 1; ++k) {
        level_load<T, W, DIV>(sm[min(k, last)], k <= last, rs, r[k]);
        __builtin_amdgcn_sched_barrier(0);
    }
    LevelMeta mq = sm[min(D - 1, last)];
    for (int l = 0; l < nl; l += D) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            // the current level's gathers go out first; the prefetch of
            // level l+k+D-1 (into the set level l+k-1 used) is issued while
            // they are in flight, then one wait and the row updates.  The
            // scheduling barriers keep the three groups in this order (the
            // scheduler would otherwise sink the prefetch past the compute)
            Gathered<T, W> g;
            level_gather<T, W>(r[k], xv, g);
            __builtin_amdgcn_sched_barrier(0);
            const int lev = l + k + D - 1;
            level_load<T, W, DIV>(mq, lev <= last, rs, r[(k + D - 1) % D]);
            mq = sm[min(lev + 1, last)];
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            level_finish<T, W, FORM, DIV>(r[k], g, xv, junkb);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// One wave per block (every level has at most 64 rows); vector and both
// sweeps' level metadata in LDS.
template <typename T, int W, int MODE>
__global__ void __launch_bounds__(64)
    k_ilu_apply_pf(const int *__restrict__ bstart, PfSweep A, PfSweep B,
                   const T *__restrict__ b, T *__restrict__ x, int xv_bytes,
                   const int *skip1, const int *skip2, BlkSkip bskip)
{
    if ((skip1 && *skip1) || (skip2 && *skip2))
        return;
    if (blk_skipped(bskip, blockIdx.x))
        return;
    extern __shared__ unsigned char rs_smem[];
    const int blk = blockIdx.x;
    const int b0 = bstart[blk], b1 = bstart[blk + 1], nb = b1 - b0;
    T *xv = (T *)rs_smem;
    LevelMeta *sm = (LevelMeta *)(rs_smem + xv_bytes);
    const int la0 = A.blk[blk], nla = A.blk[blk + 1] - la0;
    const int lb0 = B.blk[blk], nlb = B.blk[blk + 1] - lb0;
    if (nla > 0)
        wg_copy(sm, A.meta + la0, nla);
    if (nlb > 0)
        wg_copy(sm + nla, B.meta + lb0, nlb);
    if (nb > 0)
        wg_copy(xv, b + b0, nb);
    if (threadIdx.x == 0)
        xv[nb] = T(0);
    __syncthreads();
    const int junkb = (nb + 1) * (int)sizeof(T);
    const PfRsrc ra = pf_rsrc(A), rb = pf_rsrc(B);
    if (MODE == RS_MODE_ILU) {
        sweep_pf<T, W, FORM_SUM, false>(nla, sm, ra, rs_smem, junkb);
        sweep_pf<T, W, FORM_SUM, true>(nlb, sm + nla, rb, rs_smem, junkb);
    } else if (MODE == RS_MODE_ILU_T) {
        sweep_pf<T, W, FORM_SUBSEQ, true>(nla, sm, ra, rs_smem, junkb);
        sweep_pf<T, W, FORM_SUBSEQ, false>(nlb, sm + nla, rb, rs_smem, junkb);
    } else {
        sweep_pf<T, W, FORM_SUBSEQ, true>(nla, sm, ra, rs_smem, junkb);
        sweep_pf<T, W, FORM_SUM, true>(nlb, sm + nla, rb, rs_smem, junkb);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nb; r += blockDim.x)
        x[b0 + r] = xv[r];
}

// ---------------------------------------------------------------- cell sweeps
// One lane per cell (3 rows).  A level's loads per lane are two records: the
// IREC ints (LDS byte offsets of the cell and of its neighbour blocks) and
// the VREC values (row a's 3*WB neighbour-block entries at [3*WB*a, ...),
// the intra-cell coefficients at 9*WB, for U the diagonals at 9*WB+3).
// Each neighbour block's three vector entries are gathered once for the
// three rows.  Padding neighbour slots point at the three zero slots with
// -0.0 values (s + (-0)*0 == s).
template <int WB>
constexpr int cell_irec()
{
    return (1 + WB + 3) & ~3;
}
template <int WB, bool DIV>
constexpr int cell_vrec()
{
    return (9 * WB + (DIV ? 6 : 3) + 3) & ~3;
}

template <typename T, int WB, bool DIV>
struct CellRegs {
    bool act;
    int idx[cell_irec<WB>()];
    T v[cell_vrec<WB, DIV>()];
};

template <typename T, int WB, bool DIV>
__device__ __forceinline__ void cell_load(const LevelMeta &m, bool valid, const PfRsrc &rs,
                                          CellRegs<T, WB, DIV> &R)
{
    constexpr int IR = cell_irec<WB>(), VR = cell_vrec<WB, DIV>();
    constexpr int es = (int)sizeof(T);
    const int count = __builtin_amdgcn_readfirstlane(m.count);
    const int off = __builtin_amdgcn_readfirstlane(m.off);
    const int t = threadIdx.x;
    const int tt = min(t, count - 1);
    R.act = valid && t < count;
    buf_load<IR * 4>(R.idx, rs.cols, tt * (IR * 4), off * (IR * 4));
    buf_load<VR * es>(R.v, rs.vals, tt * (VR * es), off * (VR * es));
}

template <typename T, int WB>
struct CellGathered {
    T rhs[3];
    T xs[WB][3];
};

template <typename T, int WB, bool DIV>
__device__ __forceinline__ void cell_gather(const CellRegs<T, WB, DIV> &R, unsigned char *xv,
                                            CellGathered<T, WB> &g)
{
    constexpr int es = (int)sizeof(T);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        g.rhs[a] = lds_at<T>(xv, R.idx[0] + a * es);
#pragma unroll
    for (int k = 0; k < WB; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            g.xs[k][c] = lds_at<T>(xv, R.idx[1 + k] + c * es);
}

// an empty asm statement with the three gathered values as in/out registers:
// every gather is issued before the first of these, and they wait in order
template <typename T>
__device__ __forceinline__ void pin3(T (&x)[3])
{
    asm volatile("" : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]));
}

template <typename T, int WB, bool DIV>
__device__ __forceinline__ void cell_finish(const CellRegs<T, WB, DIV> &R,
                                            CellGathered<T, WB> &g, unsigned char *xv,
                                            int junkb)
{
    pin3<T>(g.rhs);
#pragma unroll
    for (int k = 0; k < WB; ++k)
        pin3<T>(g.xs[k]);
    constexpr int es = (int)sizeof(T);
    constexpr int iv = 9 * WB;
    T y[3];
    if (!DIV) {
        // L: y_a = rhs_a - (neighbour entries, then L(a,0) y_0, L(a,1) y_1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T sum = 0;
#pragma unroll
            for (int k = 0; k < WB; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    sum += R.v[3 * WB * a + 3 * k + c] * g.xs[k][c];
            if (a == 1)
                sum += R.v[iv + 0] * y[0];
            if (a == 2) {
                sum += R.v[iv + 1] * y[0];
                sum += R.v[iv + 2] * y[1];
            }
            y[a] = g.rhs[a] - sum;
        }
    } else {
        // U: rows 2, 1, 0; y_a = (rhs_a - (U(a,a+1..2) terms, then neighbour
        // entries)) / d_a
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int a = 2 - q;
            T sum = 0;
            if (a == 1)
                sum += R.v[iv + 2] * y[2];
            if (a == 0) {
                sum += R.v[iv + 0] * y[1];
                sum += R.v[iv + 1] * y[2];
            }
#pragma unroll
            for (int k = 0; k < WB; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    sum += R.v[3 * WB * a + 3 * k + c] * g.xs[k][c];
            y[a] = (g.rhs[a] - sum) / R.v[iv + 3 + a];
        }
    }
    const int wr = R.act ? R.idx[0] : junkb;
    const int st = R.act ? es : 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        lds_at<T>(xv, wr + a * st) = y[a];
}

// Prefetch depth of the cell pipeline: (depth-1) levels of loads in flight
// within the wave's 63-load counter, and the register sets within ~216 VGPRs.
template <typename T, int WB>
constexpr int pf3_depth()
{
    constexpr int es = (int)sizeof(T);
    constexpr int loads = cell_irec<WB>() / 4 + cell_vrec<WB, true>() * es / 16;
    constexpr int regs = 1 + cell_irec<WB>() + cell_vrec<WB, true>() * es / 4;
    int d = 6;
    while (d > 2 && ((d - 1) * loads > 60 || d * regs > 216))
        --d;
    return d;
}

// The sweep_pf pipeline with cell level entries.
template <typename T, int WB, bool DIV>
__device__ __forceinline__ void sweep_pf3(int nl, const LevelMeta *sm, const PfRsrc &rs,
                                          unsigned char *xv, int junkb)
{
    constexpr int D = pf3_depth<T, WB>();
    if (nl <= 0)
        return;
    const int last = nl - 1;
    CellRegs<T, WB, DIV> r[D];
#pragma unroll
    for (int k = 0; k < D - 1; ++k) {
        cell_load<T, WB, DIV>(sm[min(k, last)], k <= last, rs, r[k]);
        __builtin_amdgcn_sched_barrier(0);
    }
    LevelMeta mq = sm[min(D - 1, last)];
    for (int l = 0; l < nl; l += D) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            CellGathered<T, WB> g;
            cell_gather<T, WB, DIV>(r[k], xv, g);
            __builtin_amdgcn_sched_barrier(0);
            const int lev = l + k + D - 1;
            cell_load<T, WB, DIV>(mq, lev <= last, rs, r[(k + D - 1) % D]);
            mq = sm[min(lev + 1, last)];
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            cell_finish<T, WB, DIV>(r[k], g, xv, junkb);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// One wave per block (at most 64 cells per level); L then U cell sweeps.
// LDS: the block's vector, three zero slots, one junk slot, level metadata.
template <typename T, int WB>
__global__ void __launch_bounds__(64)
    k_ilu_apply_pf3(const int *__restrict__ bstart, PfSweep A, PfSweep B,
                    const T *__restrict__ b, T *__restrict__ x, int xv_bytes,
                    const int *skip1, const int *skip2, BlkSkip bskip)
{
    if ((skip1 && *skip1) || (skip2 && *skip2))
        return;
    if (blk_skipped(bskip, blockIdx.x))
        return;
    extern __shared__ unsigned char rs_smem[];
    const int blk = blockIdx.x;
    const int b0 = bstart[blk], b1 = bstart[blk + 1], nb = b1 - b0;
    T *xv = (T *)rs_smem;
    LevelMeta *sm = (LevelMeta *)(rs_smem + xv_bytes);
    const int la0 = A.blk[blk], nla = A.blk[blk + 1] - la0;
    const int lb0 = B.blk[blk], nlb = B.blk[blk + 1] - lb0;
    if (nla > 0)
        wg_copy(sm, A.meta + la0, nla);
    if (nlb > 0)
        wg_copy(sm + nla, B.meta + lb0, nlb);
    if (nb > 0)
        wg_copy(xv, b + b0, nb);
    if (threadIdx.x < 3)
        xv[nb + threadIdx.x] = T(0);
    __syncthreads();
    const int junkb = (nb + 3) * (int)sizeof(T);
    const PfRsrc ra = pf_rsrc(A), rb = pf_rsrc(B);
    sweep_pf3<T, WB, false>(nla, sm, ra, rs_smem, junkb);
    sweep_pf3<T, WB, true>(nlb, sm + nla, rb, rs_smem, junkb);
    __syncthreads();
    for (int r = threadIdx.x; r < nb; r += blockDim.x)
        x[b0 + r] = xv[r];
}

// ---------------------------------------------------------------- streamed cell sweeps
// A block's level records are one contiguous run of each record array
// (levels in order, a level's cells in order).  The wave streams whole
// levels into an LDS ring of st3_depth level slots with direct-to-LDS
// buffer loads (per instruction 64 lanes x 16 consecutive bytes, no
// registers held), st3_depth - 1 levels ahead of the level it computes, and
// its lanes read their records from the ring one level ahead; the gathers
// and the rows' arithmetic are the cell pipeline's (cell_gather,
// cell_finish), so the results are the same bits.  Every level issues the
// same NV + 1 copies (levels past the sweep's end copy its last level again
// into a slot that is never read), so the wait for the next level's slot is
// a fixed count of younger copies.  The copies are issued from inline
// assembly: the compiler's own tracking of direct-to-LDS loads waits for
// every one of them before each LDS read, which leaves one level in flight.
// Buffer range checking turns reads past a record array's end into zeros.
#define RS_ST3_D2_64 10
#define RS_ST3_D2_32 16
#define RS_ST3_MAXNV 4

template <typename T, int NV>
constexpr int st3_depth()
{
    constexpr int d = sizeof(T) == 8 ? RS_ST3_D2_64 : RS_ST3_D2_32;
    constexpr int dv = 63 / (NV + 1) + 1;   // copies in flight within the 6-bit vmcnt
    return d < dv ? d : dv;
}

template <int K>
__device__ __forceinline__ void wait_vm_count()
{
    static_assert(K >= 0 && K <= 63, "6-bit vmcnt");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(K) : "memory");
}

// one direct-to-LDS copy: lane i's NB bytes at byte voff of the buffer go
// to LDS byte lds + NB i
template <int NB>
__device__ __forceinline__ void lds_dma(rs_v4i rsrc, unsigned lds, int voff)
{
    static_assert(NB == 16 || NB == 4, "dword or dwordx4 lanes");
    if constexpr (NB == 16)
        asm volatile("s_mov_b32 m0, %0\n\t"
                     "s_nop 0\n\t"
                     "buffer_load_dwordx4 %1, %2, 0 offen lds" ::"s"(lds),
                     "v"(voff), "s"(rsrc)
                     : "memory", "m0");
    else
        asm volatile("s_mov_b32 m0, %0\n\t"
                     "s_nop 0\n\t"
                     "buffer_load_dword %1, %2, 0 offen lds" ::"s"(lds),
                     "v"(voff), "s"(rsrc)
                     : "memory", "m0");
}

__device__ __forceinline__ rs_v4i raw_rsrc(const void *p, int nbytes)
{
    const unsigned long long a = (unsigned long long)p;
    rs_v4i r;
    r.x = (int)(unsigned)a;
    r.y = (int)(unsigned)((a >> 32) & 0xffffu);   // stride 0
    r.z = nbytes;
    r.w = RS_BUF_WORD3;
    return r;
}

__device__ __forceinline__ unsigned lds_addr(const unsigned char *p)
{
    return (unsigned)(size_t)(const __attribute__((address_space(3))) unsigned char *)p;
}

struct St3Ring {
    rs_v4i vals, cols;          // record arrays
    unsigned char *rv_lds, *rc_lds;   // the rings
    int rv, rc;                 // record bytes (values, indices)
    int sv, sc;                 // slot bytes
    bool c16;                   // index copies of 16 (else 4) bytes per lane
};

// copy the level whose first record is `off` into the slot of level m
template <int NV, int D2>
__device__ __forceinline__ void st3_copy(const St3Ring &S, int off, int m)
{
    off = __builtin_amdgcn_readfirstlane(off);
    const int slot = m % D2;
    const int lane = (int)threadIdx.x;
    const unsigned lv = lds_addr(S.rv_lds + slot * S.sv);
    const int vb = off * S.rv + lane * 16;
#pragma unroll
    for (int q = 0; q < NV; ++q)
        lds_dma<16>(S.vals, lv + q * 1024, vb + q * 1024);
    const unsigned lc = lds_addr(S.rc_lds + slot * S.sc);
    if (S.c16)
        lds_dma<16>(S.cols, lc, off * S.rc + lane * 16);
    else
        lds_dma<4>(S.cols, lc, off * S.rc + lane * 4);
}

template <typename T, int WB, bool DIV>
__device__ __forceinline__ void cell_load_ring(const LevelMeta &m, bool valid, const St3Ring &S,
                                               int slot, CellRegs<T, WB, DIV> &R)
{
    constexpr int IR = cell_irec<WB>(), VR = cell_vrec<WB, DIV>();
    constexpr int es = (int)sizeof(T);
    const int count = __builtin_amdgcn_readfirstlane(m.count);
    const int t = threadIdx.x;
    const int tt = min(t, count - 1);
    R.act = valid && t < count;
    const unsigned char *pc = S.rc_lds + slot * S.sc + tt * (IR * 4);
    const unsigned char *pv = S.rv_lds + slot * S.sv + tt * (VR * es);
#pragma unroll
    for (int k = 0; k < IR / 4; ++k) {
        const rs_v4i q = *(const rs_v4i *)(pc + 16 * k);
        __builtin_memcpy((char *)R.idx + 16 * k, &q, 16);
    }
#pragma unroll
    for (int k = 0; k < VR * es / 16; ++k) {
        const rs_v4i q = *(const rs_v4i *)(pv + 16 * k);
        __builtin_memcpy((char *)R.v + 16 * k, &q, 16);
    }
}

template <typename T, int WB, bool DIV, int NV>
__device__ __forceinline__ void sweep_st3(int nl, const LevelMeta *sm, const St3Ring &S,
                                          unsigned char *xv, int junkb)
{
    constexpr int D2 = st3_depth<T, NV>();
    // copies younger than the next level's: those of the D2 - 2 levels after it
    constexpr int WAIT = (D2 - 2) * (NV + 1);
    if (nl <= 0)
        return;
    const int last = nl - 1;
    for (int k = 0; k < D2 - 1; ++k)
        st3_copy<NV, D2>(S, sm[min(k, last)].off, k);
    // metadata one iteration ahead: the next copy's first record, the next
    // ring read's level
    int offc = sm[min(D2 - 1, last)].off;
    LevelMeta mr = sm[min(1, last)];
    wait_vm_count<WAIT>();
    CellRegs<T, WB, DIV> r[2];
    cell_load_ring<T, WB, DIV>(sm[0], true, S, 0, r[0]);
    for (int l = 0; l < nl; l += 2) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int lev = l + k;
            st3_copy<NV, D2>(S, offc, lev + D2 - 1);
            offc = sm[min(lev + D2, last)].off;
            wait_vm_count<WAIT>();
            CellGathered<T, WB> g;
            cell_gather<T, WB, DIV>(r[k], xv, g);
            __builtin_amdgcn_sched_barrier(0);
            cell_load_ring<T, WB, DIV>(mr, lev + 1 <= last, S, (lev + 1) % D2, r[k ^ 1]);
            mr = sm[min(lev + 2, last)];
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            cell_finish<T, WB, DIV>(r[k], g, xv, junkb);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// One wave per block, L then U cell sweeps streamed through the rings.
// LDS: the block's vector (+ zero and junk slots), level metadata, then the
// value ring (D2 slots of sv bytes) and the index ring (D2 slots of sc).
template <typename T, int WB, int NV>
__global__ void __launch_bounds__(64)
    k_ilu_apply_st3(const int *__restrict__ bstart, PfSweep A, PfSweep B,
                    const T *__restrict__ b, T *__restrict__ x, int xv_bytes, int ring_off,
                    int sv, int sc, int c16, const int *skip1, const int *skip2, BlkSkip bskip)
{
    if ((skip1 && *skip1) || (skip2 && *skip2))
        return;
    if (blk_skipped(bskip, blockIdx.x))
        return;
    constexpr int D2 = st3_depth<T, NV>();
    extern __shared__ unsigned char rs_smem[];
    const int blk = blockIdx.x;
    const int b0 = bstart[blk], b1 = bstart[blk + 1], nb = b1 - b0;
    T *xv = (T *)rs_smem;
    LevelMeta *sm = (LevelMeta *)(rs_smem + xv_bytes);
    const int la0 = A.blk[blk], nla = A.blk[blk + 1] - la0;
    const int lb0 = B.blk[blk], nlb = B.blk[blk + 1] - lb0;
    if (nla > 0)
        wg_copy(sm, A.meta + la0, nla);
    if (nlb > 0)
        wg_copy(sm + nla, B.meta + lb0, nlb);
    if (nb > 0)
        wg_copy(xv, b + b0, nb);
    if (threadIdx.x < 3)
        xv[nb + threadIdx.x] = T(0);
    __syncthreads();
    const int junkb = (nb + 3) * (int)sizeof(T);
    constexpr int IR = cell_irec<WB>();
    St3Ring S;
    S.rv_lds = rs_smem + ring_off;
    S.rc_lds = rs_smem + ring_off + D2 * sv;
    S.rc = IR * 4;
    S.sv = sv;
    S.sc = sc;
    S.c16 = __builtin_amdgcn_readfirstlane(c16) != 0;
    S.vals = raw_rsrc(A.val, A.val_bytes);
    S.cols = raw_rsrc(A.colb, A.col_bytes);
    S.rv = cell_vrec<WB, false>() * (int)sizeof(T);
    sweep_st3<T, WB, false, NV>(nla, sm, S, rs_smem, junkb);
    S.vals = raw_rsrc(B.val, B.val_bytes);
    S.cols = raw_rsrc(B.colb, B.col_bytes);
    S.rv = cell_vrec<WB, true>() * (int)sizeof(T);
    sweep_st3<T, WB, true, NV>(nlb, sm + nla, S, rs_smem, junkb);
    // no copy may still be writing the rings when the workgroup's LDS is
    // released
    wait_vm_count<0>();
    __syncthreads();
    for (int r = threadIdx.x; r < nb; r += blockDim.x)
        x[b0 + r] = xv[r];
}

// ---------------------------------------------------------------- host ops
template <typename T>
static void launch_factor_band(rsgpu_prec *M, T *lu);
// status code of the line-wavefront factorisation on a zero pivot
#define RS_BAND_FACTOR_PIVOT (-99)

// band: the line-wavefront factorisation where it applies (full 9-point
// blocks); its zero-pivot marker makes the caller run the level kernel
template <typename T>
static void launch_factor(rsgpu_prec *M, T *lu, const double *aval, bool band = true)
{
    hipStream_t s = M->ctx->stream;
    const int nnz = M->nnz;
    if (nnz > 0) {
        hipLaunchKernelGGL(k_load_values<T>, dim3((nnz + 255) / 256), dim3(256), 0, s, aval,
                           M->src_identity ? nullptr : M->src.p, lu, nnz);
        RS_CHECK_LAUNCH();
    }
    int init[2] = {0, 0x7fffffff};
    RS_HIP(hipMemcpyAsync(M->status.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    if (band && M->band_factor) {
        launch_factor_band<T>(M, lu);
        return;
    }
    if (M->cell_factor) {
        hipLaunchKernelGGL(k_ilu_factor_cells<T>, dim3(M->nblocks), dim3(64), 0, s,
                           M->bstart.p, M->L.blk_lev.p, (const LevelMeta *)M->L.lev_meta.p,
                           M->L.rows.p, M->cell_rec.p, lu, M->status.p);
        RS_CHECK_LAUNCH();
        return;
    }
    hipLaunchKernelGGL(k_ilu_factor<T>, dim3(M->nblocks), dim3(M->L.max_count > 64 ? 256 : 64),
                       0, s, M->bstart.p, M->L.blk_lev.p, (const LevelMeta *)M->L.lev_meta.p,
                       M->L.rows.p, M->rowptr.p, M->colind.p, M->diag.p, lu, M->status.p,
                       M->L.R);
    RS_CHECK_LAUNCH();
}

template <typename T>
static void launch_ell(hipStream_t s, const SweepData &D, const T *lu, T *ell, bool sum_form)
{
    if (D.ell_size == 0)
        return;
    hipLaunchKernelGGL(k_gather_ell<T>, dim3((unsigned)((D.ell_size + 255) / 256)), dim3(256),
                       0, s, D.ell_src.p, lu, ell, D.ell_size, sum_form ? T(-0.0) : T(0.0));
    RS_CHECK_LAUNCH();
}

template <typename T>
static void launch_rowdiag(hipStream_t s, const SweepData &D, const int *diag, const T *lu,
                           T *dgl)
{
    if (D.nrows_lev <= 0)
        return;
    hipLaunchKernelGGL(k_gather_rowdiag<T>, dim3((D.nrows_lev + 255) / 256), dim3(256), 0, s,
                       D.row_glob.p, diag, lu, dgl, D.nrows_lev);
    RS_CHECK_LAUNCH();
}

// per-row line-wavefront records from the factor: L0..L3, U0..U3 (zero
// where absent) of row r at 8 (r + PAD), diagonals at 8 (n + 2 PAD) + r + PAD
// (the padding rows stay zero)
// rd (FP32 factors only, may be null): 1 / U_ii rounded to FP64 at
// rd[r + PAD]; (float)((double)y * rd) equals y / U_ii rounded to FP32 for
// every y (the product's error is below 2^-51 relative, while a quotient of
// two floats that is not exactly representable lies at least 2^-49 relative
// from every rounding midpoint of FP32), at a third of the division's
// dependent instructions
template <typename T>
__global__ void k_band_gather(int n, const int *__restrict__ pos, const T *__restrict__ lu,
                              T *__restrict__ out, double *__restrict__ rd)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const int *p = pos + (size_t)9 * r;
    T *o = out + (size_t)8 * (r + RS_BAND_PAD);
    // every load issued, absent positions read entry 0 and are then
    // replaced by zero (a guarded load per slot waits for memory on its own)
    int q[9];
    T v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        q[k] = p[k];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        v[k] = lu[max(q[k], 0)];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o[k] = q[k] >= 0 ? v[k] : T(0);
        o[4 + k] = q[5 + k] >= 0 ? v[5 + k] : T(0);
    }
    const T dg = v[4];
    out[(size_t)8 * ((size_t)n + 2 * RS_BAND_PAD) + r + RS_BAND_PAD] = dg;
    if (rd)
        rd[r + RS_BAND_PAD] = 1.0 / (double)dg;
}

template <typename T>
static void launch_band_gather(rsgpu_prec *M, const T *lu, T *out)
{
    if (!M->band)
        return;
    hipLaunchKernelGGL(k_band_gather<T>, dim3((M->n + 255) / 256), dim3(256), 0,
                       M->ctx->stream, M->n, M->band_pos.p, lu, out,
                       sizeof(T) == 4 ? M->band_rd.p : nullptr);
    RS_CHECK_LAUNCH();
}

int prec_factor(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    hipStream_t s = M->ctx->stream;
    int st[2];
    if (M->band_factor) {
        // a zero pivot in the line-wavefront factorisation: the level kernel
        // factors again and reports it as ILU.factor does
        if (M->precision == RSGPU_FP64)
            launch_factor<double>(M, M->lu64.p, A->val.p);
        else
            launch_factor<float>(M, M->lu32.p, A->val.p);
        RS_HIP(hipMemcpyAsync(st, M->status.p, sizeof(st), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        const bool again = st[0] == RS_BAND_FACTOR_PIVOT;
        if (M->precision == RSGPU_FP64) {
            if (again)
                launch_factor<double>(M, M->lu64.p, A->val.p, false);
            prec_refresh_fp64(M);
        } else {
            if (again)
                launch_factor<float>(M, M->lu32.p, A->val.p, false);
            launch_ell<float>(s, M->L, M->lu32.p, M->ell32_L.p, true);
            launch_ell<float>(s, M->U, M->lu32.p, M->ell32_U.p, true);
            if (!M->cell)
                launch_rowdiag<float>(s, M->U, M->diag.p, M->lu32.p, M->dgU32.p);
            launch_band_gather<float>(M, M->lu32.p, M->band32.p);
        }
        RS_HIP(hipMemcpyAsync(st, M->status.p, sizeof(st), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        if (st[0] != 0) {
            if (bad_row)
                *bad_row = st[1];
            return st[0];
        }
        return RSGPU_OK;
    }
    if (M->precision == RSGPU_FP64) {
        launch_factor<double>(M, M->lu64.p, A->val.p);
        launch_ell<double>(s, M->L, M->lu64.p, M->ell64_L.p, true);
        launch_ell<double>(s, M->U, M->lu64.p, M->ell64_U.p, true);
        if (!M->cell)
            launch_rowdiag<double>(s, M->U, M->diag.p, M->lu64.p, M->dgU64.p);
        launch_band_gather<double>(M, M->lu64.p, M->band64.p);
        if (M->has_t) {
            launch_ell<double>(s, M->Ut, M->lu64.p, M->ell64_Ut.p, false);
            launch_ell<double>(s, M->Lt, M->lu64.p, M->ell64_Lt.p, false);
            launch_rowdiag<double>(s, M->Ut, M->diag.p, M->lu64.p, M->dgUt64.p);
        }
    } else {
        launch_factor<float>(M, M->lu32.p, A->val.p);
        launch_ell<float>(s, M->L, M->lu32.p, M->ell32_L.p, true);
        launch_ell<float>(s, M->U, M->lu32.p, M->ell32_U.p, true);
        if (!M->cell)
            launch_rowdiag<float>(s, M->U, M->diag.p, M->lu32.p, M->dgU32.p);
        launch_band_gather<float>(M, M->lu32.p, M->band32.p);
    }
    RS_HIP(hipMemcpyAsync(st, M->status.p, sizeof(st), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    if (st[0] != 0) {
        if (bad_row)
            *bad_row = st[1];
        return st[0];
    }
    return RSGPU_OK;
}

// FP64 sweep arrays from the factor values in M->lu64 (the gathers that
// follow a factorisation), for a factor computed outside prec_factor.
void prec_refresh_fp64(rsgpu_prec *M)
{
    hipStream_t s = M->ctx->stream;
    launch_ell<double>(s, M->L, M->lu64.p, M->ell64_L.p, true);
    launch_ell<double>(s, M->U, M->lu64.p, M->ell64_U.p, true);
    if (!M->cell)
        launch_rowdiag<double>(s, M->U, M->diag.p, M->lu64.p, M->dgU64.p);
    launch_band_gather<double>(M, M->lu64.p, M->band64.p);
    if (M->has_t) {
        launch_ell<double>(s, M->Ut, M->lu64.p, M->ell64_Ut.p, false);
        launch_ell<double>(s, M->Lt, M->lu64.p, M->ell64_Lt.p, false);
        launch_rowdiag<double>(s, M->Ut, M->diag.p, M->lu64.p, M->dgUt64.p);
    }
}

template <typename T>
static PfSweep pf_sweep(const SweepData &D, const T *val, const T *dgl)
{
    PfSweep P;
    P.blk = D.blk_lev.p;
    P.meta = (const LevelMeta *)D.lev_meta.p;
    P.rowsb = D.rowsb.p;
    P.colb = D.colb.p;
    P.val = val;
    P.dgl = dgl;
    P.rows_bytes = (int)(D.rowsb.n * 4);
    P.col_bytes = (int)(D.colb.n * 4);
    P.val_bytes = (int)(D.ell_size * sizeof(T));
    P.dgl_bytes = dgl ? (int)((size_t)D.nrows_lev * sizeof(T)) : 0;
    return P;
}

// the pipelined kernel addresses each sweep array with 32-bit byte offsets
static bool pf_addressable(const SweepData &D, size_t es)
{
    const size_t lim = (size_t)INT32_MAX;
    return D.colb.n * 4 <= lim && D.ell_size * es <= lim && D.rowsb.n * 4 <= lim &&
           (size_t)D.nrows_lev * es <= lim;
}

template <typename T, int TRANS, int W>
static void launch_pf(rsgpu_prec *M, const SweepData &A_, const T *va, const T *da,
                      const SweepData &B_, const T *vb, const T *db, const T *b, T *x,
                      size_t xv_bytes, size_t shm, const int *sk1, const int *sk2)
{
    // dynamic LDS beyond the 64 KiB default needs a per-kernel opt-in
    static size_t granted = 65536;
    if (shm > granted) {
        RS_HIP(hipFuncSetAttribute((const void *)k_ilu_apply_pf<T, W, TRANS>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        granted = shm;
    }
    hipLaunchKernelGGL((k_ilu_apply_pf<T, W, TRANS>), dim3(M->nblocks), dim3(64), shm,
                       M->ctx->stream, M->bstart.p, pf_sweep<T>(A_, va, da),
                       pf_sweep<T>(B_, vb, db), b, x, (int)xv_bytes, sk1, sk2, M->skip);
}

// Largest dynamic LDS a workgroup may request on this device (opt-in).
static int lds_limit(int device)
{
    static int cached[64];
    static bool have[64];
    if (device < 0 || device >= 64)
        return 65536;
    if (!have[device]) {
        int a = 0, b = 0;
        (void)hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
        (void)hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, device);
        cached[device] = std::max(a, b);
        have[device] = true;
    }
    return cached[device];
}

// Sweep kernel choice: RSGPU_ILU_KERNEL = pf | generic forces one (tests
// and tuning); by default the pipelined kernel wherever it applies, else
// the generic one.
static int ilu_kernel_choice()
{
    const char *e = getenv("RSGPU_ILU_KERNEL");
    if (!e)
        return 0;
    if (!strcmp(e, "pf"))
        return 2;
    if (!strcmp(e, "generic"))
        return 3;
    return 0;
}

// ---------------------------------------------------------------- line wavefront
// ILU.apply on blocks of line form (band_block): one wave per block, lane j
// owns line j.  In the L sweep lane j treats row (i, j) at step s = i + 2j,
// so its line j-1 couplings (i-1, j-1), (i, j-1), (i+1, j-1) were computed
// by lane j-1 three, two and one steps earlier and (i-1, j) by itself one
// step earlier: each step shifts lane j-1's latest result in with a DPP
// wave shift and keeps the two before it in registers, and no level ever
// waits for LDS.  The U sweep runs backwards (step s treats
// i = Lx - 1 - s + 2 (Ly - 1 - j)) with the neighbours from lane j+1.  Every
// row forms sum = 0, sum += v * x over its present slots in ascending column
// order (absent slots selected away), then rhs - sum (and / diagonal): the
// statement order of UnitLowerCompRowMatrix.solve and
// UpperCompRowMatrix.solve, bit for bit.  The block's vector lives in LDS
// (right-hand side, then y, then x, each written only by the lane that owns
// the row); each lane's rows are consecutive, so the row records (the
// sweep's 4 values of the row's 8, the diagonal and mask) are prefetched
// RS_BAND_D steps ahead in a register ring.  The L and U values of a row
// share one 32-byte record: the L sweep's fetch of a line brings the U
// halves into the last-level cache for the U sweep.
#define RS_BAND_D 8

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf,
                                                      false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const long long q = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)q, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(q >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
#define RS_DPP_WAVE_SHR1 0x138   // lane k gets lane k-1 (lane 0: 0)
#define RS_DPP_WAVE_SHL1 0x130   // lane k gets lane k+1 (lane 63: 0)

// ---------------------------------------------------------------- line-wavefront factorisation
// ILU.factor (ILU.java:86-125) on a full 9-point block of Ly lines of Lx
// rows (every row holds all of its in-block neighbours), one wave per
// block, lane j owning line j: row (x, j) at step s = x + 2j, as the L
// sweep.  The rows its lower entries point at, (x-1, j-1), (x, j-1),
// (x+1, j-1) and (x-1, j), were finished by lane j-1 three, two and one
// steps earlier and by itself one step earlier; each step publishes its
// row's diagonal and four upper entries, shifted to lane j+1 with a DPP
// wave shift.  Per row the statement order of ILU.factor: for each lower
// entry in ascending column, divide by that row's pivot, then subtract its
// multiple of that row's upper entries from the entries of this row in the
// same columns (ascending), so factors are bit-identical to the level
// kernel's.  Records: 9 values per row (L0..L3, diagonal, U0..U3 in
// ascending column; zero where absent), padded by RS_BAND_PAD rows.
#define RS_BF_D 8

template <typename T>
struct BandIn {
    T v[9];
};

template <typename T>
__global__ void k_band_in(int n, const int *__restrict__ pos, const T *__restrict__ lu,
                          T *__restrict__ rec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const int *p = pos + (size_t)9 * r;
    T *o = rec + (size_t)9 * (r + RS_BAND_PAD);
    // every load issued (absent positions read entry 0), then selected
    int q[9];
    T v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        q[k] = p[k];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        v[k] = lu[max(q[k], 0)];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        o[k] = q[k] >= 0 ? v[k] : T(0);
}

// The factor rows of a workgroup's 256 rows are one contiguous run of lu:
// it is staged in LDS (coalesced reads), the rows' band values are written
// into the stage, and the run is written back coalesced.  A run longer than
// the stage (rows with more than 9 entries) is written in place.
template <typename T>
__global__ void __launch_bounds__(256)
    k_band_out(int n, const int *__restrict__ rowptr, const int *__restrict__ pos,
               const T *__restrict__ rec, T *__restrict__ lu)
{
    __shared__ T seg[256 * 9];
    const int r0 = (int)blockIdx.x * 256, r1 = min(r0 + 256, n);
    const int base = rowptr[r0], len = rowptr[r1] - base;
    const bool staged = len <= 256 * 9;
    const int r = r0 + (int)threadIdx.x;
    if (staged) {
        for (int j = threadIdx.x; j < len; j += 256)
            seg[j] = lu[base + j];
        __syncthreads();
    }
    if (r < n) {
        const int *p = pos + (size_t)9 * r;
        const T *o = rec + (size_t)9 * (r + RS_BAND_PAD);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int q = p[k];
            if (q >= 0) {
                if (staged)
                    seg[q - base] = o[k];
                else
                    lu[q] = o[k];
            }
        }
    }
    if (staged) {
        __syncthreads();
        for (int j = threadIdx.x; j < len; j += 256)
            lu[base + j] = seg[j];
    }
}

template <typename T>
__global__ void __launch_bounds__(64)
    k_ilu_factor_band(const int *__restrict__ bstart, const int2 *__restrict__ dims,
                      const T *__restrict__ rin, T *__restrict__ rout, int *status)
{
    const int blk = blockIdx.x;
    const int b0 = bstart[blk], Lx = dims[blk].x, Ly = dims[blk].y;
    const int j = threadIdx.x;
    const bool line = j < Ly;
    const int jj = min(j, Ly - 1);
    const int S = Lx + 2 * Ly - 2;
    // lane j's row at step s is b0 + jj Lx + s - 2 jj (records are padded,
    // so steps before and after the line read other rows or padding)
    const ptrdiff_t rb = (ptrdiff_t)b0 + (ptrdiff_t)jj * Lx - 2 * jj;
    const T *rp = rin + 9 * (rb + RS_BAND_PAD);
    T *wp = rout + 9 * (rb + RS_BAND_PAD);
    T *junk = rout;                       // the first padding row
    const bool js = jj > 0, jn = jj < Ly - 1;
    BandIn<T> ring[RS_BF_D];
#pragma unroll
    for (int k = 0; k < RS_BF_D; ++k) {
        ring[k] = *reinterpret_cast<const BandIn<T> *>(rp + 9 * k);
        __builtin_amdgcn_sched_barrier(0);
    }
    // published (diagonal, U0..U3) of lane j-1's last three rows and of this
    // lane's previous row
    T h1[5], h2[5], h3[5], pv[5];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        h1[q] = h2[q] = h3[q] = pv[q] = T(0);
    bool bad = false;
    int s0 = 0;
    do {
#pragma unroll
        for (int k = 0; k < RS_BF_D; ++k) {
            const int s = s0 + k;
            const int x = s - 2 * jj;
            const bool act = line && x >= 0 && x < Lx;
            const bool w = x > 0, e = x < Lx - 1;
            T a[9];
#pragma unroll
            for (int q = 0; q < 9; ++q)
                a[q] = ring[k].v[q];
            // lower entry L0 = (x-1, j-1): row h3
            const bool p0 = w && js, p1 = js, p2 = e && js, p3 = w;
            T m;
            m = a[0] / h3[0];
            bad |= act && p0 && h3[0] == T(0);
            a[0] = p0 ? m : a[0];
            {
                const T t1 = a[1] - a[0] * h3[1];   // (x, j-1)   <- U0 of L0
                const T t3 = a[3] - a[0] * h3[3];   // (x-1, j)   <- U2 of L0
                const T t4 = a[4] - a[0] * h3[4];   // (x, j)     <- U3 of L0
                a[1] = p0 && p1 ? t1 : a[1];
                a[3] = p0 && p3 ? t3 : a[3];
                a[4] = p0 ? t4 : a[4];
            }
            // L1 = (x, j-1): row h2
            m = a[1] / h2[0];
            bad |= act && p1 && h2[0] == T(0);
            a[1] = p1 ? m : a[1];
            {
                const T t2 = a[2] - a[1] * h2[1];   // (x+1, j-1) <- U0 of L1
                const T t3 = a[3] - a[1] * h2[2];   // (x-1, j)   <- U1 of L1
                const T t4 = a[4] - a[1] * h2[3];   // (x, j)     <- U2 of L1
                const T t5 = a[5] - a[1] * h2[4];   // (x+1, j)   <- U3 of L1
                a[2] = p1 && p2 ? t2 : a[2];
                a[3] = p1 && p3 ? t3 : a[3];
                a[4] = p1 ? t4 : a[4];
                a[5] = p1 && e ? t5 : a[5];
            }
            // L2 = (x+1, j-1): row h1
            m = a[2] / h1[0];
            bad |= act && p2 && h1[0] == T(0);
            a[2] = p2 ? m : a[2];
            {
                const T t4 = a[4] - a[2] * h1[2];   // (x, j)     <- U1 of L2
                const T t5 = a[5] - a[2] * h1[3];   // (x+1, j)   <- U2 of L2
                a[4] = p2 ? t4 : a[4];
                a[5] = p2 && e ? t5 : a[5];
            }
            // L3 = (x-1, j): this lane's previous row
            m = a[3] / pv[0];
            bad |= act && p3 && pv[0] == T(0);
            a[3] = p3 ? m : a[3];
            {
                const T t4 = a[4] - a[3] * pv[1];   // (x, j)     <- U0 of L3
                const T t6 = a[6] - a[3] * pv[3];   // (x-1, j+1) <- U2 of L3
                const T t7 = a[7] - a[3] * pv[4];   // (x, j+1)   <- U3 of L3
                a[4] = p3 ? t4 : a[4];
                a[6] = p3 && jn ? t6 : a[6];
                a[7] = p3 && jn ? t7 : a[7];
            }
            T *o = act ? wp + 9 * s : junk;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                o[q] = a[q];
            T cur[5] = {a[4], a[5], a[6], a[7], a[8]};
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                cur[q] = act ? cur[q] : T(0);
                h3[q] = h2[q];
                h2[q] = h1[q];
                h1[q] = dpp_mov<RS_DPP_WAVE_SHR1>(cur[q]);
                pv[q] = cur[q];
            }
            ring[k] = *reinterpret_cast<const BandIn<T> *>(rp + 9 * (s + RS_BF_D));
            __builtin_amdgcn_sched_barrier(0);
        }
        s0 += RS_BF_D;
    } while (s0 < S);
    if (__ballot(bad) && threadIdx.x == 0)
        atomicExch(status, RS_BAND_FACTOR_PIVOT);
}

template <typename T>
static void launch_factor_band(rsgpu_prec *M, T *lu)
{
    hipStream_t s = M->ctx->stream;
    const size_t np = (size_t)M->n + 2 * RS_BAND_PAD;
    DevBuf<T> *inb, *outb;
    if constexpr (sizeof(T) == 8) {
        inb = &M->bf_in64;
        outb = &M->bf_out64;
    } else {
        inb = &M->bf_in32;
        outb = &M->bf_out32;
    }
    DevBuf<T> &in = *inb, &out = *outb;
    if (in.n != 9 * np) {
        in.alloc(9 * np);
        in.zero(s);
        out.alloc(9 * np);
        out.zero(s);
    }
    const int g = (M->n + 255) / 256;
    hipLaunchKernelGGL(k_band_in<T>, dim3(g), dim3(256), 0, s, M->n, M->band_pos.p, lu, in.p);
    hipLaunchKernelGGL(k_ilu_factor_band<T>, dim3(M->nblocks), dim3(64), 0, s, M->bstart.p,
                       M->band_dims.p, in.p, out.p, M->status.p);
    hipLaunchKernelGGL(k_band_out<T>, dim3(g), dim3(256), 0, s, M->n, M->rowptr.p,
                       M->band_pos.p, out.p, lu);
    RS_CHECK_LAUNCH();
}

template <typename T>
struct alignas(16) BandRec {
    T v[4];
};


// Progress hooks of the wavefront sweeps.  ready(b): every row (i, j) of
// the block with i + 2 j < b holds its right-hand side (called by L sweeps
// before they read those rows); done(t): the first t steps of a U sweep
// are written (called after each round).  The plain apply needs neither;
// the fused GMRES step has other waves produce and consume the vector.
struct NoHook {
    __device__ __forceinline__ void ready(int) {}
    __device__ __forceinline__ void done(int) {}
};

// Global loads in flight: the ring depth D (steps ahead), within the 63
// loads the wave's counter tracks and about 200 registers of ring.
template <typename T, bool UPPER, bool FULL>
constexpr int band_depth()
{
    constexpr int loads = (int)sizeof(T) / 4 + (UPPER ? 1 : 0) + (FULL ? 0 : 1);
    constexpr int regs =
        4 * (int)sizeof(T) / 4 + (UPPER ? (int)sizeof(T) / 4 : 0) + (FULL ? 0 : 1);
    int d = 16;
    while (d > 4 && (d * loads > 62 || d * regs > 200))
        --d;
    return d;
}

template <typename T, bool UPPER, bool FULL>
struct BandStage {
    BandRec<T> v;
    T d;
    unsigned m;
};

// record of row offset q (in rows) from base pointers at the lane's row
template <typename T, bool UPPER, bool FULL>
__device__ __forceinline__ void band_load(const T *rp, const T *dp, const unsigned char *mp,
                                          int q, BandStage<T, UPPER, FULL> &st)
{
    st.v = *reinterpret_cast<const BandRec<T> *>(rp + 8 * q);
    if (UPPER)
        st.d = dp[q];
    if (!FULL)
        st.m = mp[q];
}

template <typename T, bool UPPER, bool FULL, class Hook = NoHook, int DMAX = 16>
__device__ __forceinline__ void band_sweep(int Lx, int Ly, int g0, const T *rec, const T *dgl,
                                           const unsigned char *mask, T *xv, int junk,
                                           Hook hk = Hook())
{
    constexpr int D0 = band_depth<T, UPPER, FULL>();
    constexpr int D = D0 < DMAX ? D0 : DMAX;
    constexpr int dir = UPPER ? -1 : 1;
    const int j = threadIdx.x, jj = min(j, Ly - 1);
    const bool live = j < Ly;
    const int S = Lx + 2 * Ly - 2;
    // row of the lane at step s: rb + dir * s; line position i0 + dir * s
    const int i0 = UPPER ? Lx - 1 + 2 * (Ly - 1 - jj) : -2 * jj;
    const int rb = jj * Lx + i0;
    const bool js = jj > 0, jn = jj < Ly - 1;
    const T *rp = rec + (size_t)8 * (g0 + rb);
    const T *dp = dgl + (g0 + rb);
    const unsigned char *mp = mask + (g0 + rb);
    const T *xr = xv + rb;
    BandStage<T, UPPER, FULL> ring[D];
    // The ring is a true D-deep prefetch only if the slots are loaded in
    // slot order here and refilled in slot order in the loop: the compiler
    // then waits for D-1 outstanding loads per step.  Any other order (the
    // scheduler reorders free loads; a guard branch before the loop lets
    // them sink past it) makes it wait for all of them at the loop head,
    // one full memory round trip per D steps.  Hence the barriers and the
    // do-while (S >= 1 always).
#pragma unroll
    for (int k = 0; k < D; ++k) {
        band_load<T, UPPER, FULL>(rp, dp, mp, dir * k, ring[k]);
        __builtin_amdgcn_sched_barrier(0);
    }
    T cur = T(0), h1 = T(0), h2 = T(0), h3 = T(0);
    if (!UPPER)
        hk.ready(1);
    T rhs = xr[0];
    int s0 = 0;
    do {
        // step s reads the right-hand side of step s + 1
        if (!UPPER)
            hk.ready(s0 + D + 1);
        // base pointers of this round (every access below has an immediate
        // offset from them)
        const T *rps = rp + (ptrdiff_t)8 * dir * s0;
        const T *dps = dp + dir * s0;
        const unsigned char *mps = mp + dir * s0;
        const T *xrs = xr + dir * s0;
        const int is0 = i0 + dir * s0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            BandStage<T, UPPER, FULL> &st = ring[k];
            const int i = is0 + dir * k;
            const bool act = live && (unsigned)i < (unsigned)Lx;
            const T rhs_s = rhs;
            rhs = xrs[dir * (k + 1)];
            // the neighbour line's latest result and the two before it
            h3 = h2;
            h2 = h1;
            h1 = UPPER ? dpp_mov<RS_DPP_WAVE_SHL1>(cur) : dpp_mov<RS_DPP_WAVE_SHR1>(cur);
            const bool w = i > 0, e = i < Lx - 1;
            bool p0, p1, p2, p3;
            if (FULL) {
                // L: (i-1,j-1), (i,j-1), (i+1,j-1), (i-1,j); U: (i+1,j),
                // (i-1,j+1), (i,j+1), (i+1,j+1)
                p0 = UPPER ? e : (w && js);
                p1 = UPPER ? (w && jn) : js;
                p2 = UPPER ? jn : (e && js);
                p3 = UPPER ? (e && jn) : w;
            } else {
                const unsigned m = UPPER ? st.m >> 4 : st.m;
                p0 = m & 1u;
                p1 = m & 2u;
                p2 = m & 4u;
                p3 = m & 8u;
            }
            const T o0 = UPPER ? cur : h3, o1 = UPPER ? h1 : h2;
            const T o2 = UPPER ? h2 : h1, o3 = UPPER ? h3 : cur;
            T sum = T(0);
            T q = sum + st.v.v[0] * o0;
            sum = p0 ? q : sum;
            q = sum + st.v.v[1] * o1;
            sum = p1 ? q : sum;
            q = sum + st.v.v[2] * o2;
            sum = p2 ? q : sum;
            q = sum + st.v.v[3] * o3;
            sum = p3 ? q : sum;
            T val = rhs_s - sum;
            if (UPPER)
                val = val / st.d;
            // inactive lanes store into a padding slot: a branch here would
            // make the compiler's wait counting drain the prefetch ring
            xv[act ? rb + dir * (s0 + k) : junk] = val;
            cur = act ? val : cur;
            // refill the slot for step s + D
            band_load<T, UPPER, FULL>(rps, dps, mps, dir * (k + D), st);
            // keep the refill in its step: left free, the scheduler gathers
            // a round's refills at the round's end, and the next round then
            // waits for all of them (one memory round trip per D steps)
            __builtin_amdgcn_sched_barrier(0);
        }
        s0 += D;
        if (UPPER)
            hk.done(s0);
    } while (s0 < S);
}

// 16-byte vector of T (LDS accesses of the full-block sweeps)
template <typename T>
struct BandVec;
template <>
struct BandVec<float> {
    using type = float4;
    static constexpr int N = 4;
};
template <>
struct BandVec<double> {
    using type = double2;
    static constexpr int N = 2;
};

// Full blocks (every row holds all of its in-block 9-point neighbours): the
// same wavefront, with the block's vector in LDS by line and step instead of
// by row.  Line j owns SP = band_slot_pitch(Lx, Ly) slots from j SP, and slot
// s holds row i = s - 2 j (the row lane j treats at L step s; slots outside
// 0 <= i < Lx are padding), so the L sweep walks its lane's slots upwards and
// the U sweep downwards (U step t: slot SP - 1 - t) and a round of 16 steps
// covers 16 aligned slots: its right-hand sides come in with 16-byte LDS
// reads issued one round ahead, and its results go out with 16-byte writes at
// the round's end (lanes past Ly write nothing).
// NOPRED: the couplings are summed without the boundary tests.  An absent
// coupling's record value is zero, and the sum, which starts at +0, is never
// -0 before its last term, so adding a zero product changes no bit as long
// as the neighbour value is finite: with finite data the result equals the
// tested form bit for bit (non-finite values may then spread further).
// FP32 U sweeps divide by multiplying with the FP64 reciprocal rd (exact,
// see k_band_gather); FP64 sweeps divide.
template <typename T, bool UPPER>
struct FullStage {
    using DT = typename std::conditional<sizeof(T) == 4, double, T>::type;
    BandRec<T> v;
    DT d;
};

template <typename T, bool UPPER, bool NOPRED, class Hook = NoHook>
__device__ __forceinline__ void band_sweep_full(int Lx, int Ly, int SP, int g0, const T *rec,
                                                const T *dgl, const double *rdl, T *xs,
                                                Hook hk = Hook())
{
    constexpr int D = 16;
    static_assert(band_depth<T, UPPER, true>() == D, "full-block rounds are 16 steps");
    constexpr bool RCP = sizeof(T) == 4;
    constexpr int dir = UPPER ? -1 : 1;
    using VT = typename BandVec<T>::type;
    constexpr int NV = BandVec<T>::N;      // elements per 16-byte LDS access
    constexpr int NB = D / NV;             // accesses per round
    const int j = threadIdx.x, jj = min(j, Ly - 1);
    const bool live = j < Ly;
    const bool js = jj > 0, jn = jj < Ly - 1;
    // line position of step 0 (i = slot - 2 j; U step t uses slot SP - 1 - t)
    const int i0 = UPPER ? SP - 1 - 2 * jj : -2 * jj;
    const int rb = jj * Lx + i0;
    const T *rp = rec + (size_t)8 * (g0 + rb);
    using DT = typename FullStage<T, UPPER>::DT;
    const DT *dp = RCP ? (const DT *)(rdl + (g0 + rb)) : (const DT *)(dgl + (g0 + rb));
    VT *xl = reinterpret_cast<VT *>(xs + (size_t)jj * SP);
    FullStage<T, UPPER> ring[D];
    auto load = [&](const T *rps, const DT *dps, int q, FullStage<T, UPPER> &st) {
        st.v = *reinterpret_cast<const BandRec<T> *>(rps + 8 * q);
        if (UPPER)
            st.d = dps[q];
    };
#pragma unroll
    for (int k = 0; k < D; ++k) {
        load(rp, dp, dir * k, ring[k]);
        __builtin_amdgcn_sched_barrier(0);
    }
    const int rounds = SP / D;
    // round r's slots: L [16 r, 16 r + 16), U [SP - 16 (r + 1), SP - 16 r)
    auto slot0 = [&](int r) { return UPPER ? SP - D * (r + 1) : D * r; };
    VT rin[NB], rnext[NB];
    // L round r reads slots [16 r, 16 r + 16) (rows i + 2 j < 16 (r + 1))
    if (!UPPER)
        hk.ready(D);
#pragma unroll
    for (int u = 0; u < NB; ++u)
        rin[u] = xl[slot0(0) / NV + u];
    T cur = T(0), h1 = T(0), h2 = T(0), h3 = T(0);
    int r = 0;
    do {
        const int rn = min(r + 1, rounds - 1);
        if (!UPPER)
            hk.ready(D * (rn + 1));
#pragma unroll
        for (int u = 0; u < NB; ++u)
            rnext[u] = xl[slot0(rn) / NV + u];
        const T *rps = rp + (ptrdiff_t)8 * dir * (D * r);
        const DT *dps = dp + dir * (D * r);
        const int is0 = i0 + dir * (D * r);
        T outv[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            FullStage<T, UPPER> &st = ring[k];
            const int i = is0 + dir * k;
            const bool act = live && (unsigned)i < (unsigned)Lx;
            // L: slot 16 r + k is element k of the round; U: slot
            // SP - 1 - 16 r - k is element 15 - k
            const int e = UPPER ? D - 1 - k : k;
            const T rhs = reinterpret_cast<const T *>(&rin[e / NV])[e % NV];
            h3 = h2;
            h2 = h1;
            h1 = UPPER ? dpp_mov<RS_DPP_WAVE_SHL1>(cur) : dpp_mov<RS_DPP_WAVE_SHR1>(cur);
            const T o0 = UPPER ? cur : h3, o1 = UPPER ? h1 : h2;
            const T o2 = UPPER ? h2 : h1, o3 = UPPER ? h3 : cur;
            T sum = T(0);
            if (NOPRED) {
                sum = sum + st.v.v[0] * o0;
                sum = sum + st.v.v[1] * o1;
                sum = sum + st.v.v[2] * o2;
                sum = sum + st.v.v[3] * o3;
            } else {
                const bool w = i > 0, ea = i < Lx - 1;
                // L: (i-1,j-1), (i,j-1), (i+1,j-1), (i-1,j); U: (i+1,j),
                // (i-1,j+1), (i,j+1), (i+1,j+1)
                const bool p0 = UPPER ? ea : (w && js);
                const bool p1 = UPPER ? (w && jn) : js;
                const bool p2 = UPPER ? jn : (ea && js);
                const bool p3 = UPPER ? (ea && jn) : w;
                T q = sum + st.v.v[0] * o0;
                sum = p0 ? q : sum;
                q = sum + st.v.v[1] * o1;
                sum = p1 ? q : sum;
                q = sum + st.v.v[2] * o2;
                sum = p2 ? q : sum;
                q = sum + st.v.v[3] * o3;
                sum = p3 ? q : sum;
            }
            T val = rhs - sum;
            if (UPPER) {
                if (RCP)
                    val = (T)((double)val * (double)st.d);
                else
                    val = val / (T)st.d;
            }
            outv[e] = val;
            cur = act ? val : cur;
            load(rps, dps, dir * (k + D), st);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (live) {
#pragma unroll
            for (int u = 0; u < NB; ++u)
                xl[slot0(r) / NV + u] = *reinterpret_cast<const VT *>(&outv[u * NV]);
        }
        if (UPPER)
            hk.done(D * (r + 1));
#pragma unroll
        for (int u = 0; u < NB; ++u)
            rin[u] = rnext[u];
        ++r;
    } while (r < rounds);
}

template <typename T>
__global__ void __launch_bounds__(64)
    k_ilu_apply_band(const int *__restrict__ bstart, const int2 *__restrict__ dims,
                     const unsigned char *__restrict__ mask, const T *__restrict__ recL,
                     const T *__restrict__ recU,
                     const T *__restrict__ dgl, const double *__restrict__ rdl,
                     const T *__restrict__ b, T *__restrict__ x, const int *skip1,
                     const int *skip2, BlkSkip bskip, int slots)
{
    if ((skip1 && *skip1) || (skip2 && *skip2))
        return;
    if (blk_skipped(bskip, blockIdx.x))
        return;
    extern __shared__ unsigned char rs_smem[];
    const int b0 = bstart[blockIdx.x], nb = bstart[blockIdx.x + 1] - b0;
    const int2 d = dims[blockIdx.x];   // Lx, Ly; Ly < 0: the block's pattern is not full
    const int Lx = d.x, Ly = d.y < 0 ? -d.y : d.y;
    if (slots && d.y > 0 && sizeof(T) == 4) {
        // full FP32 block: slot layout (band_sweep_full; in FP64 its larger
        // register set made it slower than the per-step form below)
        T *xs = (T *)rs_smem;
        const int SP = band_slot_pitch(Lx, Ly), ns = Ly * SP;
        // slot s of line j holds row s - 2 j of the line: the padding
        // slots [0, 2 j) and [2 j + Lx, SP) are zeroed, then the rows are
        // copied line by line, sixteen lines per batch, loads in flight
        // together (line indices past Ly clamp, only the stores are
        // predicated)
        for (int j = 0; j < Ly; ++j)
            for (int p = threadIdx.x; p < SP - Lx; p += 64)
                xs[j * SP + (p < 2 * j ? p : p + Lx)] = T(0);
        for (int i = threadIdx.x; i < Lx; i += 64)
            for (int j0 = 0; j0 < Ly; j0 += 16) {
                T v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    v[u] = b[b0 + min(j0 + u, Ly - 1) * Lx + i];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (j0 + u < Ly)
                        xs[(j0 + u) * SP + i + 2 * (j0 + u)] = v[u];
            }
        __syncthreads();
        band_sweep_full<T, false, true>(Lx, Ly, SP, b0, recL, dgl, rdl, xs);
        __syncthreads();
        band_sweep_full<T, true, true>(Lx, Ly, SP, b0, recU, dgl, rdl, xs);
        __syncthreads();
        {
            int jl = 0, il = (int)threadIdx.x;
            while (il >= Lx) {
                il -= Lx;
                ++jl;
            }
            for (int q = threadIdx.x; q < nb; q += 64) {
                x[b0 + q] = xs[jl * SP + il + 2 * jl];
                il += 64;
                while (il >= Lx) {
                    il -= Lx;
                    ++jl;
                }
            }
        }
        return;
    }
    T *xv = (T *)rs_smem + RS_BAND_PAD;
    if (nb > 0)
        wg_copy(xv, b + b0, nb);
    for (int r = threadIdx.x; r < RS_BAND_PAD; r += 64) {
        xv[-1 - r] = T(0);
        xv[nb + r] = T(0);
    }
    __syncthreads();
    // junk slot: the last padding row (read only by prefetches of rows
    // outside the block, whose values are never used)
    const int junk = nb + RS_BAND_PAD - 1;
    if (d.y > 0) {
        band_sweep<T, false, true>(Lx, Ly, b0, recL, dgl, mask, xv, junk);
        __syncthreads();
        band_sweep<T, true, true>(Lx, Ly, b0, recU, dgl, mask, xv, junk);
    } else {
        band_sweep<T, false, false>(Lx, Ly, b0, recL, dgl, mask, xv, junk);
        __syncthreads();
        band_sweep<T, true, false>(Lx, Ly, b0, recU, dgl, mask, xv, junk);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nb; r += 64)
        x[b0 + r] = xv[r];
}

// The line wavefront gives each block one wave with a short dependency
// chain: the faster apply while the blocks are about one per CU (the
// 128x128x64 system's 256).  With many blocks per CU the level-scheduled
// kernel's throughput wins (the 64-system ensemble's 7,060 blocks: 352
// against 508 us per FP32 apply, 592 against 1,193 us per FP64 apply), so
// above RS_BAND_MAX_PER_CU blocks per CU the apply takes the level
// schedules, unless a batched solve timed both on the preconditioner
// (band_pick).  RSGPU_ILU_BAND_APPLY=1 forces the wavefront, =0 the levels.
#define RS_BAND_MAX_PER_CU 4
static bool band_apply_preferred(const rsgpu_prec *M)
{
    const char *e = getenv("RSGPU_ILU_BAND_APPLY");
    if (e && (e[0] == '0' || e[0] == '1'))
        return e[0] == '1';
    if (M->band_pick >= 0)
        return M->band_pick == 1;
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess || cus <= 0)
            cus = 256;
    }
    return M->nblocks <= RS_BAND_MAX_PER_CU * cus;
}

bool prec_band_apply(const rsgpu_prec *M) { return band_apply_preferred(M); }

// Layout of full FP32 blocks: the per-line slot layout; RSGPU_BAND_SLOTS=0
// selects the row layout of the per-step sweep (bit-identical), whose
// smaller LDS vector (Lx Ly + 2 RS_BAND_PAD elements against
// Ly round_up(Lx + 2 Ly - 2, 16)) lets more blocks share a CU but was
// slower on the ensemble's 7,060 blocks too (1.20 against 1.11 s per mixed
// ensemble: three blocks per CU already keep its SIMDs busy).
static bool band_use_slots(const rsgpu_prec *)
{
    const char *e = getenv("RSGPU_BAND_SLOTS");
    return !(e && e[0] == '0');
}

template <typename T>
static void launch_band(rsgpu_prec *M, const T *rec, const T *b, T *x, const int *sk1,
                        const int *sk2)
{
    const bool slots = sizeof(T) == 4 && band_use_slots(M);
    const size_t shm =
        std::max((size_t)(M->max_block_rows + 2 * RS_BAND_PAD),
                 slots ? (size_t)M->band_slots : (size_t)0) *
        sizeof(T);
    static bool attr = false;
    if (!attr && shm > 65536) {
        RS_HIP(hipFuncSetAttribute((const void *)k_ilu_apply_band<T>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   lds_limit(M->ctx->device)));
        attr = true;
    }
    // padded layout: rows -PAD .. n+PAD-1 of 8 values (4 L, then 4 U), then
    // the diagonals
    const size_t np = (size_t)M->n + 2 * RS_BAND_PAD;
    const double *rdl = (sizeof(T) == 4 && M->band_rd.p) ? M->band_rd.p + RS_BAND_PAD : nullptr;
    hipLaunchKernelGGL(k_ilu_apply_band<T>, dim3(M->nblocks), dim3(64), shm, M->ctx->stream,
                       M->bstart.p, M->band_dims.p, M->band_mask.p + RS_BAND_PAD,
                       rec + (size_t)8 * RS_BAND_PAD, rec + (size_t)8 * RS_BAND_PAD + 4,
                       rec + 8 * np + RS_BAND_PAD, rdl, b, x,
                       sk1, sk2, M->skip, slots ? 1 : 0);
    RS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- fused GMRES step
// One inner step of GMRES with block-Jacobi ILU(0) on full line blocks
// (every block Ly lines of Lx rows, Lx a multiple of 64, block starts on
// 64-row boundaries) and a sliced matrix copy, in one launch:
//   u = A v_i (for i > 0, v_i = (1 / H(i,i-1)) w is formed on the fly and
//       stored, the rule of k_spmv_scale_sl),
//   w' = M^-1 u (the line-wavefront L and U sweeps),
//   H(0..i, i) = V' w' (the first projection pass of classical
//       Gram-Schmidt; column i taken from v_i as formed above).
// One workgroup per block.  Wave 0 runs the two sweeps at raised issue
// priority; the producer waves (fuse_producers) compute the block's SpMV
// rows, 64 rows of one line at a time, directly into the sweep's LDS vector,
// in the order the L sweep reaches them (the sweep waits on per-task flags
// only when it catches up), and while the U sweep runs they store the
// finished rows of w' and accumulate their projections on the basis.  The
// kernel is bound by memory bandwidth (SpMV values and factor records, then
// factor records and basis), so the overlap hides the sweeps' dependency
// chains rather than any traffic.  Each SpMV row and each
// sweep result is that of the separate kernels bit for bit (same per-row
// operation order); the projections are summed per lane over its rows, over
// lanes and producer waves in a fixed order, then over blocks in block
// order, so they are reproducible run to run (their rounding differs from
// the unfused kernel's grid-stride order).
// producer waves: 11 in FP32 (168 registers per wave, three waves per SIMD);
// 7 in FP64, whose larger sweep state needs more than 168
template <typename T>
constexpr int fuse_producers()
{
    return sizeof(T) == 4 ? 11 : 7;
}
// record prefetch depth of the FP64 sweeps in the fused step (their own
// depth: 11 FP64 producers with an 8-deep ring fit 168 registers but were
// no faster, the SpMV part being bound by memory bandwidth)
template <typename T>
constexpr int fuse_ring()
{
    return 16;
}
template <typename T>
constexpr int fuse_threads()
{
    return 64 * (1 + fuse_producers<T>());
}
#define RS_FUSE_COLS 32
#define RS_FUSE_SPIN (1 << 24)
// bytes of the fused kernel's static LDS arrays (projection sums of 11
// producers, flags), kept free beside its dynamic LDS
#define RS_FUSE_STATIC 6144

// LDS progress between the waves of a fused step.  Flags are written with
// workgroup-scope release and read with acquire; the U sweep's step count is
// published with a plain store after its round's LDS writes (one wave's LDS
// operations complete in issue order) so the sweep never waits for them.
struct FuseHook {
    int *flags;    // phase-A task t = h Ly + j (rows 64 h .. 64 h + 63 of line j) done
    int *uprog;    // U steps completed
    int *err;      // a bounded wait ran out
    int ntA, Ly;
    __device__ __forceinline__ void ready(int bound)
    {
        const int lane = threadIdx.x & 63;
        for (int spin = 0;
             !__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); ++spin) {
            bool miss = false;
            for (int t = lane; t < ntA; t += 64) {
                const int h = t / Ly, j = t - h * Ly;
                const int f = __hip_atomic_load(&flags[t], __ATOMIC_ACQUIRE,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
                miss |= (64 * h + 2 * j < bound) && f == 0;
            }
            if (!__any(miss))
                return;
            if (spin >= RS_FUSE_SPIN) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __device__ __forceinline__ void done(int steps)
    {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if ((threadIdx.x & 63) == 0)
            __hip_atomic_store(uprog, steps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
};

// keeps a loaded 16-byte vector live at this point (its load is issued
// before any code after it)
__device__ __forceinline__ void fuse_pin(const float4 &v)
{
    asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
}
__device__ __forceinline__ void fuse_pin(const double2 &v)
{
    asm volatile("" ::"v"(v.x), "v"(v.y));
}

// SpMV rows into the sweep's LDS vector (CompRowMatrix.mult's zero rule):
// row r of the matrix at xs[r + base]
template <typename T>
struct EpiLds {
    T *xs;
    int base;
    __device__ T init(int) const { return T(0); }
    __device__ void row(int r, T dot) { xs[r + base] = (dot != T(0)) ? dot : T(0); }
};

// The delayed reorthogonalisation step's column work (one wave of the
// finishing group of step i).  colsum holds z = V(:,0..i-1)' w, beta =
// v~_i' w, then (i > 0) s = V(:,0..i-1)' v~_i and alpha = v~_i' v~_i.  With
// rho = sqrt(alpha - s's) and eta = st->hsub (the norm v~_i was scaled by),
// the reorthogonalised v_i = (v~_i - V s) / rho has M^-1 A v_i =
// (w - V Hr s) / rho, so column i of the unrotated Hessenberg is
// Hr(r, i) = ([z; c]_r - (Hr s)_r) / rho with c = (beta - s'z) / rho, the
// product taken over column i - 1 corrected as in dcgs_finish_column_wave
// (which then finishes column i - 1).  The update kernel forms v_i and
// (w - V [z; c]) / rho from ds, rho and c.
__device__ void dcgs_step_wave(DevState *st, int i, bool reo, const double *colsum, double *lh)
{
    const int lane = threadIdx.x & 63;
    const int ld = st->m + 1, cnt = i + 1;
    double *Hr = st->Hr;
    if (!reo) {
        if (lane == 0) {
            Hr[i * ld] = colsum[0];
            st->dc = colsum[0];
            st->drho = 1.0;
        }
        return;
    }
    const double *sl = colsum + cnt;
    const double sk = lane < i ? sl[lane] : 0.0;
    const double zk = lane < i ? colsum[lane] : 0.0;
    const double alpha = sl[i], beta = colsum[i];
    const double ss = wave_sum(sk * sk);
    const double rho = alpha - ss > 0 ? sqrt(alpha - ss) : 0.0;
    const double sz = wave_sum(sk * zk);
    const double c = (beta - sz) / rho;
    const double eta = st->hsub;
    if (lane <= i) {
        double hs = 0;
        for (int k = max(lane - 1, 0); k < i; ++k) {
            const double hk = k < i - 1 ? Hr[lane + k * ld]
                              : lane < i ? Hr[lane + k * ld] + eta * sl[lane]
                                         : eta * rho;
            hs = hs + hk * sl[k];
        }
        const double zr = lane < i ? zk : c;
        Hr[lane + i * ld] = (zr - hs) / rho;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    dcgs_finish_column_wave(st, i - 1, lh, sl);
    if (lane == 0) {
        st->dc = c;
        st->drho = rho;
    }
}

// DCGS (FP64 only): the delayed reorthogonalisation form.  Phase B also
// projects the step's input vector v~_i (column i as the SpMV formed it) on
// the basis (the reorthogonalisation coefficients of v~_i and its squared
// norm), and the finishing wave finishes column i - 1 with them and forms
// column i of the unrotated Hessenberg (dcgs_step_wave).
template <typename T, bool DCGS>
__global__ void __launch_bounds__(fuse_threads<T>())
    k_gm_fused(SlView sv, const T *__restrict__ val, const T *__restrict__ win,
               T *__restrict__ vout, int scale, const int *__restrict__ bstart,
               const int2 *__restrict__ dims, const T *__restrict__ recL,
               const T *__restrict__ recU,
               const T *__restrict__ dgl, const double *__restrict__ rdl, T *__restrict__ wout,
               const T *__restrict__ V, size_t ldv, int i, DevState *st, RedSlot slot,
               int vec_elems, double *hout, int npre, int pre_off)
{
    using VT = typename BandVec<T>::type;
    constexpr int N = BandVec<T>::N;
    constexpr bool SLOTS = sizeof(T) == 4;   // FP32: slot layout; FP64: row layout
    constexpr int P = fuse_producers<T>(), BS = fuse_threads<T>();
    extern __shared__ unsigned char rs_smem[];
    __shared__ double sh[P][RS_FUSE_COLS];
    __shared__ double sh2[DCGS ? P : 1][RS_FUSE_COLS];
    __shared__ double colsum[DCGS ? 2 * RS_FUSE_COLS : 1];
    __shared__ double lh[DCGS ? 3 * 64 : 1];
    __shared__ int uprog, err, last, precnt;
    double *Hb = st->H;
    const int ld = st->m + 1;
    const double hsub = st->hsub;
    {
        const int d_ = st->done, c_ = st->cycle_end;
        if (d_ | c_)
            return;
    }
    const int G = (int)gridDim.x;
    const int b = xcd_remap((int)blockIdx.x, G);
    const int b0 = bstart[b], nb = bstart[b + 1] - b0;
    const int Lx = dims[b].x, Ly = dims[b].y;
    const int SP = band_slot_pitch(Lx, Ly);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int ntA = nb / 64;
    T *xv = (T *)rs_smem + (SLOTS ? 0 : RS_BAND_PAD);
    int *flags = (int *)((T *)rs_smem + vec_elems);
    const T sc = scale ? (T)(1. / hsub) : T(1);
    // LDS set-up: padding zeroed, flags cleared
    if (SLOTS) {
        for (int j = 0; j < Ly; ++j)
            for (int p = threadIdx.x; p < SP - Lx; p += BS)
                xv[j * SP + (p < 2 * j ? p : p + Lx)] = T(0);
    } else {
        for (int r = threadIdx.x; r < RS_BAND_PAD; r += BS) {
            xv[-1 - r] = T(0);
            xv[nb + r] = T(0);
        }
    }
    for (int t = threadIdx.x; t < ntA; t += BS)
        flags[t] = 0;
    if (threadIdx.x == 0) {
        uprog = 0;
        err = 0;
        precnt = 0;
    }
    __syncthreads();
    FuseHook hk{flags, &uprog, &err, ntA, Ly};
    // basis columns 0..npre-1 of the block's rows, copied to LDS by the
    // producers between their SpMV rows and the U sweep's first rows
    T *vpre = (T *)(rs_smem + pre_off);
    const int cnt = i + 1;
    // projection columns: V' w, then (DCGS, i > 0) V' v~_i
    const bool reo = DCGS && scale;
    const int ncol = reo ? 2 * cnt : cnt;
    if (wid == 0) {
        // ---- the sweeps
        __builtin_amdgcn_s_setprio(3);
        if (SLOTS) {
            band_sweep_full<T, false, true>(Lx, Ly, SP, b0, recL, dgl, rdl, xv, hk);
            band_sweep_full<T, true, true>(Lx, Ly, SP, b0, recU, dgl, rdl, xv, hk);
        } else {
            const int junk = nb + RS_BAND_PAD - 1;
            band_sweep<T, false, true, FuseHook, fuse_ring<T>()>(Lx, Ly, b0, recL, dgl, nullptr,
                                                                 xv, junk, hk);
            band_sweep<T, true, true, FuseHook, fuse_ring<T>()>(Lx, Ly, b0, recU, dgl, nullptr,
                                                                xv, junk, hk);
        }
        __builtin_amdgcn_s_setprio(0);
    } else {
        const int p = wid - 1;
        double a[RS_FUSE_COLS], a2[RS_FUSE_COLS];
#pragma unroll
        for (int q = 0; q < RS_FUSE_COLS; ++q) {
            a[q] = 0;
            a2[q] = 0;
        }
        // ---- phase A: SpMV rows, line segments in the L sweep's order
        for (int t = p; t < ntA; t += P) {
            const int h = t / Ly, j = t - h * Ly;
            const int row0 = b0 + j * Lx + 64 * h;
            // row (x, j) lives at slot x + 2 j of line j (FP32) or at its
            // block row (FP64)
            EpiLds<T> epi{xv, SLOTS ? j * SP + 2 * j - (b0 + j * Lx) : -b0};
            // FP32: all 32 slots of a row in flight at once (the producers
            // are latency-bound: one task per wave at a time)
            constexpr int B = SLOTS ? 32 : RS_SL_BATCH;
            if (scale) {
                const int r = row0 + lane;
                vout[r] = (sc == T(0)) ? T(0) : sc * win[r];
                sliced_rows<T, EpiLds<T>, true, B>(row0, sv, val, win, epi, sc);
            } else {
                sliced_rows<T, EpiLds<T>, false, B>(row0, sv, val, win, epi);
            }
            if (lane == 0)
                __hip_atomic_store(&flags[t], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // ---- phase B: finished rows of w' stored and projected, from the
        // block's end (the U sweep finishes it first); a lane owns N
        // consecutive rows of one line
        const int PT = SLOTS ? SP : Lx + 2 * Ly - 2;   // U steps of a full pass
        const int rowsB = 64 * N;
        const int ntB = (nb + rowsB - 1) / rowsB;
        const int kmax = scale ? i - 1 : i;            // columns read from V
        // While the L sweep finishes its last rows the producers would
        // wait: they copy the first npre basis columns of the block to LDS
        // (every load of a lane's share in flight), so the projections
        // read only the rest from memory while the U sweep streams its
        // records.  The readers wait for all producers' copies.
        if (npre > 0) {
            constexpr int NL = 8;
            const int nvec = nb / N;
            for (int v0 = p * 64 + lane; v0 < npre * nvec; v0 += P * 64 * NL) {
                VT t[NL];
#pragma unroll
                for (int u = 0; u < NL; ++u) {
                    const int id = min(v0 + u * P * 64, npre * nvec - 1);
                    const int c = id / nvec, v = id - c * nvec;
                    t[u] = *reinterpret_cast<const VT *>(V + (size_t)min(c, kmax) * ldv + b0 +
                                                         (size_t)v * N);
                }
#pragma unroll
                for (int u = 0; u < NL; ++u) {
                    const int id = v0 + u * P * 64;
                    if (id < npre * nvec)
                        reinterpret_cast<VT *>(vpre)[id] = t[u];
                }
            }
            if (lane == 0)
                __hip_atomic_fetch_add(&precnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        bool pre_ready = npre == 0;
        for (int kk = p; kk < ntB; kk += P) {
            const int q = (ntB - 1 - kk) * rowsB + N * lane;
            const bool live = q < nb;
            const int qc = live ? q : nb - N;
            const int j = qc / Lx, x0 = qc - j * Lx;
            // the lane's rows are final once the U sweep has passed its
            // first row (the lowest slot x0 + 2 j of them)
            const int need = PT - (x0 + 2 * j);
            for (int spin = 0;
                 !__hip_atomic_load(&err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                 ++spin) {
                const int u = __hip_atomic_load(&uprog, __ATOMIC_ACQUIRE,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
                if (!__any(u < need))
                    break;
                if (spin >= RS_FUSE_SPIN) {
                    __hip_atomic_store(&err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            const int li = SLOTS ? j * SP + x0 + 2 * j : qc;
            VT wv;
            T *we = reinterpret_cast<T *>(&wv);
#pragma unroll
            for (int e = 0; e < N; ++e)
                we[e] = xv[li + e];
            const size_t gr = (size_t)b0 + qc;
            if (live)
                *reinterpret_cast<VT *>(wout + gr) = wv;
            // column i: v_i as the SpMV formed it (win's rows, scaled)
            VT vi = *reinterpret_cast<const VT *>(win + gr);
            {
                T *ve = reinterpret_cast<T *>(&vi);
#pragma unroll
                for (int e = 0; e < N; ++e)
                    ve[e] = (sc == T(0)) ? T(0) : sc * ve[e];
            }
            if (!pre_ready) {
                for (int spin = 0;
                     !__hip_atomic_load(&err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                     ++spin) {
                    if (__hip_atomic_load(&precnt, __ATOMIC_ACQUIRE,
                                          __HIP_MEMORY_SCOPE_WORKGROUP) >= P)
                        break;
                    if (spin >= RS_FUSE_SPIN) {
                        __hip_atomic_store(&err, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                pre_ready = true;
            }
            // basis columns in chunks of FC: every load of a chunk is issued
            // (and pinned) before the first product, and every product is
            // formed and then selected, so no load sinks under a condition;
            // the first chunk comes from LDS when it was copied there
            constexpr int FC = 8;
#pragma unroll
            for (int c0 = 0; c0 < RS_FUSE_COLS; c0 += FC) {
                if (c0 < cnt) {
                    VT kv[FC];
                    if (c0 == 0 && npre > 0) {
#pragma unroll
                        for (int u = 0; u < FC; ++u)
                            kv[u] = reinterpret_cast<const VT *>(vpre + (size_t)u * nb)[qc / N];
                    } else {
#pragma unroll
                        for (int u = 0; u < FC; ++u)
                            kv[u] = *reinterpret_cast<const VT *>(
                                V + (size_t)max(min(c0 + u, kmax), 0) * ldv + gr);
                    }
#pragma unroll
                    for (int u = 0; u < FC; ++u)
                        fuse_pin(kv[u]);
#pragma unroll
                    for (int u = 0; u < FC; ++u) {
                        const int k = c0 + u;
                        const VT kk = (scale && k == i) ? vi : kv[u];
                        const T *ke = reinterpret_cast<const T *>(&kk);
                        double acc = a[k];
#pragma unroll
                        for (int e = 0; e < N; ++e)
                            acc = acc + (double)(we[e] * ke[e]);
                        a[k] = (live && k < cnt) ? acc : a[k];
                        if constexpr (DCGS) {
                            const T *ve = reinterpret_cast<const T *>(&vi);
                            double acc2 = a2[k];
#pragma unroll
                            for (int e = 0; e < N; ++e)
                                acc2 = acc2 + (double)(ve[e] * ke[e]);
                            a2[k] = (reo && live && k < cnt) ? acc2 : a2[k];
                        }
                    }
                }
            }
        }
        // ---- projections: the wave's sums per column
        static_assert(RS_FUSE_COLS == 32, "wave_sum_cols32 sums 32 columns");
        {
            const double r = wave_sum_cols32(a);
            if ((lane & 1) == 0 && (lane >> 1) < cnt)
                sh[p][lane >> 1] = r;
        }
        if constexpr (DCGS) {
            if (reo) {
                const double r = wave_sum_cols32(a2);
                if ((lane & 1) == 0 && (lane >> 1) < cnt)
                    sh2[p][lane >> 1] = r;
            }
        }
    }
    __syncthreads();
    // a wait that ran out (never expected: every producer finishes its
    // tasks unconditionally) stops the solve with an error
    if (err && threadIdx.x == 0) {
        st->status = RSGPU_HIP_ERROR;
        st->done = 1;
    }
    if ((int)threadIdx.x < ncol) {
        const int k = (int)threadIdx.x;
        double r = 0;
#pragma unroll
        for (int u = 0; u < P; ++u)
            r += k < cnt ? sh[u][k] : sh2[DCGS ? u : 0][k - cnt];
        __hip_atomic_store(&slot.partials[(size_t)k * G + b], r, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
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
    // wave wid sums columns wid, wid + 8, ... over the blocks in order; all
    // loads of a batch (4 partials per lane) in flight before the adds
    constexpr int NW = BS / 64;
    for (int k = wid; k < ncol; k += NW) {
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
        if (lane == 0) {
            if constexpr (DCGS) {
                colsum[k] = r;
                if (k < cnt)
                    Hb[k + i * ld] = r;
                else
                    st->ds[k - cnt] = r;
            } else if (hout) {
                hout[k] = r;
            } else {
                Hb[k + i * ld] = r;
            }
        }
    }
    if constexpr (DCGS) {
        __syncthreads();
        if (wid == 0)
            dcgs_step_wave(st, i, reo, colsum, lh);
    }
}

// Whether the fused step applies: block ILU(0) with line-wavefront records
// on full blocks of Lx % 64 == 0 starting on 64-row boundaries, the sliced
// copy, at most RS_FUSE_COLS basis vectors, the LDS vector and flags within
// the device limit.  RSGPU_GMRES_FUSED=0 turns it off.
bool gmres_fused_ok(const rsgpu_csr *A, const rsgpu_prec *M, int elem_bytes, int m)
{
    const char *e = getenv("RSGPU_GMRES_FUSED");
    const bool off = e && e[0] == '0';
    if (off || !M || !A || !A->sl.ok() || A->sl.b3 || !M->band || M->n != A->n || M->n == 0 ||
        m + 1 > RS_FUSE_COLS || M->kind == RSGPU_PREC_ILUT || M->kind == RSGPU_PREC_ICC ||
        M->kind == RSGPU_PREC_AMG || M->kind == RSGPU_PREC_SSOR ||
        M->kind == RSGPU_PREC_DIAGONAL || M->kind == RSGPU_PREC_IDENTITY || M->skip.done ||
        (int)M->h_band_dims.size() != M->nblocks || !band_apply_preferred(M))
        return false;
    if ((elem_bytes == 4) != (M->precision == RSGPU_FP32))
        return false;
    size_t vec = 0;
    for (int b = 0; b < M->nblocks; ++b) {
        const int2 d = M->h_band_dims[b];
        const int b0 = M->h_bstart[b], nb = M->h_bstart[b + 1] - b0;
        if (d.y <= 0 || d.x % 64 != 0 || b0 % 64 != 0 || nb != d.x * d.y)
            return false;
        vec = std::max(vec, elem_bytes == 4 ? (size_t)d.y * band_slot_pitch(d.x, d.y)
                                            : (size_t)nb + 2 * RS_BAND_PAD);
    }
    const size_t shm = vec * elem_bytes + (size_t)(M->max_block_rows / 64 + 1) * 4;
    return shm <= (size_t)lds_limit(M->ctx->device);
}

template <typename T>
void launch_gmres_fused(hipStream_t s, const rsgpu_csr *A, const T *slv, rsgpu_prec *M,
                        const T *win, T *vout, bool scale, T *wout, const T *V, size_t ldv,
                        int i, DevState *st, RedSlot slot, double *hout, bool dcgs)
{
    const bool f32 = sizeof(T) == 4;
    size_t vec = 0;
    for (int b = 0; b < M->nblocks; ++b) {
        const int2 d = M->h_band_dims[b];
        vec = std::max(vec, f32 ? (size_t)d.y * band_slot_pitch(d.x, d.y)
                                : (size_t)(M->h_bstart[b + 1] - M->h_bstart[b]) +
                                      2 * RS_BAND_PAD);
    }
    size_t shm = vec * sizeof(T) + (size_t)(M->max_block_rows / 64 + 1) * 4;
    // the first 8 basis columns in LDS when every one is a column of V
    // (FP32; FP64 blocks do not fit) and the device's LDS holds them beside
    // the sweep vector and the kernel's own arrays
    const size_t pre_off = (shm + 15) & ~(size_t)15;
    const size_t pre_bytes = (size_t)8 * M->max_block_rows * sizeof(T);
    const int kmax = scale ? i - 1 : i;
    const int npre = (f32 && kmax >= 7 &&
                      pre_off + pre_bytes + RS_FUSE_STATIC <= (size_t)lds_limit(M->ctx->device))
                         ? 8
                         : 0;
    if (npre)
        shm = pre_off + pre_bytes;
    // the delayed reorthogonalisation form exists in FP64 only
    const bool dc = dcgs && !f32;
    auto kern = dc ? k_gm_fused<T, !f32> : k_gm_fused<T, false>;
    // the dynamic LDS limit is raised to what the launch needs (the
    // kernel's static arrays take the rest of the device's opt-in maximum)
    static size_t attr[2] = {65536, 65536};
    if (shm > attr[dc]) {
        RS_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)shm));
        attr[dc] = shm;
    }
    const size_t np = (size_t)M->n + 2 * RS_BAND_PAD;
    const T *rec = f32 ? (const T *)M->band32.p : (const T *)M->band64.p;
    const double *rdl = (f32 && M->band_rd.p) ? M->band_rd.p + RS_BAND_PAD : nullptr;
    hipLaunchKernelGGL(kern, dim3(M->nblocks), dim3(fuse_threads<T>()), shm, s, A->slview(),
                       slv, win, vout, scale ? 1 : 0, M->bstart.p, M->band_dims.p,
                       rec + (size_t)8 * RS_BAND_PAD, rec + (size_t)8 * RS_BAND_PAD + 4,
                       rec + 8 * np + RS_BAND_PAD, rdl, wout, V, ldv, i, st, slot, (int)vec,
                       hout, npre, (int)pre_off);
    RS_CHECK_LAUNCH();
}

template void launch_gmres_fused<float>(hipStream_t, const rsgpu_csr *, const float *,
                                        rsgpu_prec *, const float *, float *, bool, float *,
                                        const float *, size_t, int, DevState *, RedSlot,
                                        double *, bool);
template void launch_gmres_fused<double>(hipStream_t, const rsgpu_csr *, const double *,
                                         rsgpu_prec *, const double *, double *, bool, double *,
                                         const double *, size_t, int, DevState *, RedSlot,
                                         double *, bool);

template <typename T, int TRANS>
static void launch_apply(rsgpu_prec *M, const SweepData &A_, const T *va, const T *da,
                         const SweepData &B_, const T *vb, const T *db, const T *b, T *x,
                         T *scratch, const int *sk1, const int *sk2)
{
    hipStream_t s = M->ctx->stream;
    if (M->n == 0)
        return;
    if (TRANS == RS_MODE_ILU && M->band && &A_ == &M->L && band_apply_preferred(M) &&
        std::max((size_t)(M->max_block_rows + 2 * RS_BAND_PAD), (size_t)M->band_slots) *
                sizeof(T) <=
            (size_t)lds_limit(M->ctx->device)) {
        launch_band<T>(M, sizeof(T) == 8 ? (const T *)M->band64.p : (const T *)M->band32.p, b,
                       x, sk1, sk2);
        return;
    }
    const int maxc = std::max(A_.max_count, B_.max_count);
    const int choice = ilu_kernel_choice();
    const bool lds = M->max_block_rows <= (int)(RS_LDS_BYTES / sizeof(T)) - 2;
    const size_t xv_bytes = ((size_t)(M->max_block_rows + 2) * sizeof(T) + 15) & ~(size_t)15;
    const size_t pf_shm =
        xv_bytes + (size_t)(A_.max_levels + B_.max_levels) * sizeof(LevelMeta);
    const int W = A_.W;
    // the pipelined kernel keeps the vector in LDS up to the device's
    // opt-in limit (160 KiB on gfx950: 20k FP64 / 40k FP32 rows per block)
    if (choice != 3 && maxc <= 64 && A_.W == B_.W && W <= 16 &&
        pf_shm <= (size_t)lds_limit(M->ctx->device) && pf_addressable(A_, sizeof(T)) &&
        pf_addressable(B_, sizeof(T))) {
        if (W <= 4)
            launch_pf<T, TRANS, 4>(M, A_, va, da, B_, vb, db, b, x, xv_bytes, pf_shm, sk1, sk2);
        else if (W <= 8)
            launch_pf<T, TRANS, 8>(M, A_, va, da, B_, vb, db, b, x, xv_bytes, pf_shm, sk1, sk2);
        else if (W <= 12)
            launch_pf<T, TRANS, 12>(M, A_, va, da, B_, vb, db, b, x, xv_bytes, pf_shm, sk1,
                                    sk2);
        else
            launch_pf<T, TRANS, 16>(M, A_, va, da, B_, vb, db, b, x, xv_bytes, pf_shm, sk1,
                                    sk2);
        RS_CHECK_LAUNCH();
        return;
    }
    const int bs = maxc > 64 ? 256 : 64;
    const size_t shm = lds ? (size_t)(M->max_block_rows + 1) * sizeof(T) : 0;
    auto kern = lds ? k_ilu_apply<T, true, TRANS> : k_ilu_apply<T, false, TRANS>;
    hipLaunchKernelGGL(kern, dim3(M->nblocks), dim3(bs), shm, s, M->bstart.p, A_.W,
                       A_.blk_lev.p, (const LevelMeta *)A_.lev_meta.p, A_.rows.p,
                       A_.ell_col.p, va, da, B_.blk_lev.p, (const LevelMeta *)B_.lev_meta.p,
                       B_.rows.p, B_.ell_col.p, vb, db, b, x, scratch, sk1, sk2, M->skip);
    RS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- ICC
// ICC.factor (ICC.java:83-118) keeps a dense work row Rk per row k and
// eliminates it against every earlier row i with Rk[i] / R(i,i) != 0, using
// row i's stored entries right of its diagonal; the work row fills in at
// columns left of k that A's row k lacks, and those fill values take part
// in the later eliminations of the same row (only positions of A's pattern
// from the diagonal on are stored).  Structurally, the rows i that touch row
// k form its reach: A's columns of row k left of k, closed under "row i's
// upper columns left of k".  The host lists per row its work slots (reach
// and row k's pattern from the diagonal on), per reach row i its work slot,
// pivot position and (target slot, source position) pairs; updates to
// columns right of k outside the pattern are dropped (never read).  Rows
// are factored level by level (a row after every row of its reach), one
// lane per row over its slice of a work array, in the reference's
// operation order per slot: the same R bit for bit.
#define RS_ICC_MAX_PAIRS (1LL << 28)

struct IccData {
    DevBuf<int> woff;    // work-slot range per row
    DevBuf<int> pslot;   // per pattern position: work slot (-1: left of diag, not reached)
    DevBuf<int> rptr;    // reach-entry range per row
    DevBuf<int4> rent;   // reach entry: work slot of i, pivot position, pair range
    DevBuf<int2> pairs;  // (target slot, source position)
    DevBuf<int> lev_ptr, lev_rows;
    int nlev = 0;
    DevBuf<double> work;
};

void icc_destroy(IccData *d) { delete d; }

static void icc_symbolic(rsgpu_prec *M)
{
    hipStream_t s = M->ctx->stream;
    const int n = M->n;
    const auto &rp = M->h_rowptr, &ci = M->h_colind;
    std::vector<int> dg(n);
    for (int k = 0; k < n; ++k)
        dg[k] = (int)(std::lower_bound(ci.begin() + rp[k], ci.begin() + rp[k + 1], k) -
                      ci.begin());
    std::vector<int> woff(n + 1, 0), pslot(std::max(M->nnz, 1), -1), rptr(n + 1, 0);
    std::vector<int4> rent;
    std::vector<int2> pairs;
    std::vector<int> level(n, 0);
    std::vector<int> mark(n, -1), slot(n, -1), reach, wcols;
    std::vector<int> heap;
    long long npairs = 0;
    for (int k = 0; k < n; ++k) {
        // reach: ascending closure with a min-heap
        reach.clear();
        heap.clear();
        auto push = [&](int c) {
            if (mark[c] != k) {
                mark[c] = k;
                heap.push_back(c);
                std::push_heap(heap.begin(), heap.end(), std::greater<int>());
            }
        };
        for (int p = rp[k]; p < dg[k]; ++p)
            push(ci[p]);
        while (!heap.empty()) {
            std::pop_heap(heap.begin(), heap.end(), std::greater<int>());
            const int i = heap.back();
            heap.pop_back();
            reach.push_back(i);
            for (int j = dg[i] + 1; j < rp[i + 1]; ++j)
                if (ci[j] < k)
                    push(ci[j]);
        }
        // work slots: reach (ascending) then row k's columns from the diagonal
        wcols.assign(reach.begin(), reach.end());
        for (int p = dg[k]; p < rp[k + 1]; ++p)
            wcols.push_back(ci[p]);
        woff[k + 1] = woff[k] + (int)wcols.size();
        for (size_t q = 0; q < wcols.size(); ++q)
            slot[wcols[q]] = (int)q;
        for (int p = rp[k]; p < rp[k + 1]; ++p)
            pslot[p] = slot[ci[p]];
        int lv = 0;
        for (int i : reach) {
            lv = std::max(lv, level[i] + 1);
            int4 e;
            e.x = slot[i];
            e.y = dg[i];
            e.z = (int)pairs.size();
            for (int j = dg[i] + 1; j < rp[i + 1]; ++j) {
                const int c = ci[j];
                const bool in = c < k ? true
                                      : std::binary_search(ci.begin() + dg[k],
                                                           ci.begin() + rp[k + 1], c);
                if (in)
                    pairs.push_back(int2{slot[c], j});
            }
            e.w = (int)pairs.size();
            rent.push_back(e);
            npairs = (long long)pairs.size();
            if (npairs > RS_ICC_MAX_PAIRS)
                throw RsError{RSGPU_BAD_ARGUMENT,
                              "ICC: the fill of this matrix exceeds the factorisation's "
                              "update list limit"};
        }
        rptr[k + 1] = (int)rent.size();
        level[k] = lv;
        for (int c : wcols)
            slot[c] = -1;
    }
    int nlev = 0;
    for (int k = 0; k < n; ++k)
        nlev = std::max(nlev, level[k] + 1);
    std::vector<int> lev_ptr(nlev + 1, 0), lev_rows(std::max(n, 1));
    for (int k = 0; k < n; ++k)
        lev_ptr[level[k] + 1]++;
    for (int l = 0; l < nlev; ++l)
        lev_ptr[l + 1] += lev_ptr[l];
    {
        std::vector<int> f(lev_ptr.begin(), lev_ptr.end() - 1);
        for (int k = 0; k < n; ++k)
            lev_rows[f[level[k]]++] = k;
    }
    if (!M->icc)
        M->icc = new IccData;
    IccData &D = *M->icc;
    D.woff.upload(woff, s);
    D.pslot.upload(pslot, s);
    D.rptr.upload(rptr, s);
    D.rent.upload(rent.empty() ? std::vector<int4>(1, int4{0, 0, 0, 0}) : rent, s);
    D.pairs.upload(pairs.empty() ? std::vector<int2>(1, int2{0, 0}) : pairs, s);
    D.lev_ptr.upload(lev_ptr, s);
    D.lev_rows.upload(lev_rows, s);
    D.nlev = nlev;
    D.work.alloc(std::max(woff[n], 1));
    RS_HIP(hipStreamSynchronize(s));
}

// One workgroup; levels in order; one lane per row (ICC.java:89-117).
// status: [code, first failing row].
__global__ void __launch_bounds__(256)
    k_icc_factor(int nlev, const int *__restrict__ lev_ptr, const int *__restrict__ lev_rows,
                 const int *__restrict__ rowptr, const int *__restrict__ diag,
                 const int *__restrict__ woff, const int *__restrict__ pslot,
                 const int *__restrict__ rptr, const int4 *__restrict__ rent,
                 const int2 *__restrict__ pairs, double *work, double *lu, int *status)
{
    for (int l = 0; l < nlev; ++l) {
        for (int t = lev_ptr[l] + (int)threadIdx.x; t < lev_ptr[l + 1]; t += blockDim.x) {
            const int k = lev_rows[t];
            double *w = work + woff[k];
            for (int q = 0; q < woff[k + 1] - woff[k]; ++q)
                w[q] = 0.0;
            for (int p = rowptr[k]; p < rowptr[k + 1]; ++p)
                if (pslot[p] >= 0)
                    w[pslot[p]] = lu[p];
            for (int e = rptr[k]; e < rptr[k + 1]; ++e) {
                const int4 E = rent[e];
                const double Rki = w[E.x] / lu[E.y];
                if (Rki == 0)
                    continue;
                for (int q = E.z; q < E.w; ++q) {
                    const int2 P = pairs[q];
                    w[P.x] -= Rki * lu[P.y];
                }
            }
            const int d = diag[k];
            const double Rkk = w[pslot[d]];
            if (Rkk == 0) {
                atomicMin(&status[1], k);
                atomicMax(&status[0], (int)RSGPU_ZERO_PIVOT);
                continue;
            }
            const double sqRkk = sqrt(Rkk);
            for (int p = d; p < rowptr[k + 1]; ++p)
                lu[p] = w[pslot[p]] / sqRkk;
        }
        __syncthreads();
    }
}

// ICC.setMatrix: R.set(A) then the factorisation; the R' and R sweeps'
// values.  Returns RSGPU_ZERO_PIVOT with the 0-based row of the zero
// diagonal (the first in row order).
static int icc_factor(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    hipStream_t s = M->ctx->stream;
    IccData &D = *M->icc;
    const int nnz = M->nnz;
    if (nnz > 0) {
        hipLaunchKernelGGL(k_load_values<double>, dim3((nnz + 255) / 256), dim3(256), 0, s,
                           A->val.p, M->src_identity ? nullptr : M->src.p, M->lu64.p, nnz);
        RS_CHECK_LAUNCH();
    }
    int init[2] = {0, 0x7fffffff};
    RS_HIP(hipMemcpyAsync(M->status.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    if (M->n > 0)
        hipLaunchKernelGGL(k_icc_factor, dim3(1), dim3(256), 0, s, D.nlev, D.lev_ptr.p,
                           D.lev_rows.p, M->rowptr.p, M->diag.p, D.woff.p, D.pslot.p, D.rptr.p,
                           D.rent.p, D.pairs.p, D.work.p, M->lu64.p, M->status.p);
    RS_CHECK_LAUNCH();
    launch_ell<double>(s, M->U, M->lu64.p, M->ell64_U.p, true);
    launch_rowdiag<double>(s, M->U, M->diag.p, M->lu64.p, M->dgU64.p);
    launch_ell<double>(s, M->Ut, M->lu64.p, M->ell64_Ut.p, false);
    launch_rowdiag<double>(s, M->Ut, M->diag.p, M->lu64.p, M->dgUt64.p);
    int st[2];
    RS_HIP(hipMemcpyAsync(st, M->status.p, sizeof(st), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    if (st[0] != 0) {
        if (bad_row)
            *bad_row = st[1];
        return st[0];
    }
    return RSGPU_OK;
}

int icc_set_matrix(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    if (!M->ready || M->pattern_id != A->pattern_id || !M->icc) {
        prec_setup(M, A);
        M->ready = false;
        // R' and R sweeps run in one kernel with one slot width
        prec_setup_transpose(M, M->U.W);
        if (M->Ut.W > M->U.W) {
            const auto &rp = M->h_rowptr, &ci = M->h_colind;
            std::vector<int> dg(M->n);
            for (int k = 0; k < M->n; ++k)
                dg[k] = (int)(std::lower_bound(ci.begin() + rp[k], ci.begin() + rp[k + 1], k) -
                              ci.begin());
            const RangeEntries eU{dg.data(), rp.data() + 1, 1, 0, ci.data()};
            HostSweep H;
            build_sweep(M->nblocks, M->h_bstart, eU, false, M->Ut.W, H);
            upload_sweep(H, M->U, M->ctx->stream, 8);
            M->ell64_U.alloc(M->U.ell_size);
            M->dgU64.alloc(M->U.nrows_lev);
            RS_HIP(hipStreamSynchronize(M->ctx->stream));
        }
        icc_symbolic(M);
        M->ready = true;
    }
    return icc_factor(M, A, bad_row);
}

static void icc_apply(rsgpu_prec *M, const double *b, double *x, const int *sk1,
                      const int *sk2)
{
    launch_apply<double, RS_MODE_ICC>(M, M->Ut, M->ell64_Ut.p, M->dgUt64.p, M->U,
                                      M->ell64_U.p, M->dgU64.p, b, x, M->scratch64.p, sk1, sk2);
}

// Cell schedules (3x3 block matrices): the pipelined cell sweeps, forward L
// then backward U; prec_setup only selects them when every level holds at
// most 64 cells, the block fits LDS and the arrays are 32-bit addressable.
template <typename T, int WB>
static void launch_pf3(rsgpu_prec *M, const PfSweep &A, const PfSweep &B, const T *b, T *x,
                       size_t xv_bytes, size_t shm, const int *sk1, const int *sk2)
{
    static size_t granted = 65536;
    if (shm > granted) {
        RS_HIP(hipFuncSetAttribute((const void *)k_ilu_apply_pf3<T, WB>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        granted = shm;
    }
    hipLaunchKernelGGL((k_ilu_apply_pf3<T, WB>), dim3(M->nblocks), dim3(64), shm,
                       M->ctx->stream, M->bstart.p, A, B, b, x, (int)xv_bytes, sk1, sk2,
                       M->skip);
}

static size_t cell_xv_bytes(int max_block_rows, int es)
{
    return ((size_t)(max_block_rows + 4) * es + 15) & ~(size_t)15;
}

template <typename T, int WB, int NV>
static void launch_st3(rsgpu_prec *M, const PfSweep &A, const PfSweep &B, const T *b, T *x,
                       size_t xv_bytes, size_t ring_off, int sv, int sc, bool c16,
                       const int *sk1, const int *sk2)
{
    constexpr int D2 = st3_depth<T, NV>();
    const size_t shm = ring_off + (size_t)D2 * (sv + sc);
    static size_t granted = 65536;
    if (shm > granted) {
        RS_HIP(hipFuncSetAttribute((const void *)k_ilu_apply_st3<T, WB, NV>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        granted = shm;
    }
    hipLaunchKernelGGL((k_ilu_apply_st3<T, WB, NV>), dim3(M->nblocks), dim3(64), shm,
                       M->ctx->stream, M->bstart.p, A, B, b, x, (int)xv_bytes, (int)ring_off,
                       sv, sc, (int)c16, sk1, sk2, M->skip);
}

// The streamed cell sweeps' layout for M in elements of es bytes, or false
// where they do not apply: WB of 2 or 3 neighbour blocks, a level's records
// within RS_ST3_MAXNV KiB (values) and 1 KiB (indices), the rings within the
// LDS limit.  RSGPU_ILU_CELL_STREAM=0 keeps the register pipeline
// (k_ilu_apply_pf3).
struct St3Plan {
    int nv, sv, sc;
    bool c16;
    size_t xv_bytes, meta_bytes, ring_off;
};

static bool st3_plan(const rsgpu_prec *M, int es, St3Plan &P)
{
    const char *e = getenv("RSGPU_ILU_CELL_STREAM");
    if ((e && e[0] == '0') || !M->cell || M->n == 0)
        return false;
    const int WB = M->L.W;
    if (M->U.W != WB || (WB != 2 && WB != 3))
        return false;
    const int maxc = std::max(M->L.max_count, M->U.max_count);
    const int rv = ((9 * WB + 6 + 3) & ~3) * es;   // the U record, the larger
    const int rc = ((1 + WB + 3) & ~3) * 4;
    P.nv = (maxc * rv + 1023) / 1024;
    if (maxc <= 0 || maxc > 64 || P.nv > RS_ST3_MAXNV || maxc * rc > 1024)
        return false;
    P.c16 = maxc * rc > 256;
    P.sc = P.c16 ? 1024 : 256;
    P.sv = P.nv * 1024;
    P.xv_bytes = cell_xv_bytes(M->max_block_rows, es);
    P.meta_bytes = (size_t)(M->L.max_levels + M->U.max_levels) * sizeof(LevelMeta);
    P.ring_off = (P.xv_bytes + P.meta_bytes + 15) & ~(size_t)15;
    const int d = es == 8 ? RS_ST3_D2_64 : RS_ST3_D2_32, dv = 63 / (P.nv + 1) + 1;
    const int d2 = std::min(d, dv);
    return P.ring_off + (size_t)d2 * (P.sv + P.sc) <= (size_t)lds_limit(M->ctx->device);
}

bool prec_cell_stream(const rsgpu_prec *M)
{
    St3Plan P;
    return st3_plan(M, M->precision == RSGPU_FP32 ? 4 : 8, P);
}

template <typename T, int WB>
static bool try_st3(rsgpu_prec *M, const PfSweep &A, const PfSweep &B, const T *b, T *x,
                    const int *sk1, const int *sk2)
{
    St3Plan P;
    if (!st3_plan(M, (int)sizeof(T), P))
        return false;
    const size_t xv_bytes = P.xv_bytes, ring_off = P.ring_off;
    const int sv = P.sv, sc = P.sc, nv = P.nv;
    const bool c16 = P.c16;
    switch (nv) {
    case 1:
        launch_st3<T, WB, 1>(M, A, B, b, x, xv_bytes, ring_off, sv, sc, c16, sk1, sk2);
        break;
    case 2:
        launch_st3<T, WB, 2>(M, A, B, b, x, xv_bytes, ring_off, sv, sc, c16, sk1, sk2);
        break;
    case 3:
        launch_st3<T, WB, 3>(M, A, B, b, x, xv_bytes, ring_off, sv, sc, c16, sk1, sk2);
        break;
    default:
        launch_st3<T, WB, 4>(M, A, B, b, x, xv_bytes, ring_off, sv, sc, c16, sk1, sk2);
        break;
    }
    return true;
}

template <typename T>
static void launch_apply_cells(rsgpu_prec *M, const T *vl, const T *vu, const T *b, T *x,
                               const int *sk1, const int *sk2)
{
    if (M->n == 0)
        return;
    const size_t xv_bytes = cell_xv_bytes(M->max_block_rows, (int)sizeof(T));
    const size_t shm =
        xv_bytes + (size_t)(M->L.max_levels + M->U.max_levels) * sizeof(LevelMeta);
    const PfSweep A = pf_sweep<T>(M->L, vl, nullptr), B = pf_sweep<T>(M->U, vu, nullptr);
    if (M->L.W == 2 && try_st3<T, 2>(M, A, B, b, x, sk1, sk2)) {
        RS_CHECK_LAUNCH();
        return;
    }
    if (M->L.W == 3 && try_st3<T, 3>(M, A, B, b, x, sk1, sk2)) {
        RS_CHECK_LAUNCH();
        return;
    }
    switch (M->L.W) {
    case 2:
        launch_pf3<T, 2>(M, A, B, b, x, xv_bytes, shm, sk1, sk2);
        break;
    case 3:
        launch_pf3<T, 3>(M, A, B, b, x, xv_bytes, shm, sk1, sk2);
        break;
    case 4:
        launch_pf3<T, 4>(M, A, B, b, x, xv_bytes, shm, sk1, sk2);
        break;
    default:
        launch_pf3<T, 8>(M, A, B, b, x, xv_bytes, shm, sk1, sk2);
        break;
    }
    RS_CHECK_LAUNCH();
}

__global__ void k_copy_skip(const double *__restrict__ a, double *__restrict__ b, int n,
                            const int *sk1, const int *sk2)
{
    if ((sk1 && *sk1) || (sk2 && *sk2))
        return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a[i];
}

__global__ void k_copy_skip32(const float *__restrict__ a, float *__restrict__ b, int n,
                              const int *sk1, const int *sk2)
{
    if ((sk1 && *sk1) || (sk2 && *sk2))
        return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a[i];
}

// apply with optional device skip flags (done / cycle_end of a solver)
// ---------------------------------------------------------------- diagonal
// DiagonalPreconditioner.setMatrix (DiagonalPreconditioner.java:67-78):
// invdiag = 1 / A(i,i); an absent or zero diagonal is an error.
__global__ void k_invdiag(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                          const double *__restrict__ v, double *__restrict__ inv,
                          float *__restrict__ inv32, int *bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int lo = rp[i], hi = rp[i + 1];
    double d = 0;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int c = ci[mid];
        if (c == i) {
            d = v[mid];
            break;
        }
        if (c < i)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (d == 0) {
        atomicMin(bad, i + 1);
        return;
    }
    const double r = 1 / d;
    inv[i] = r;
    if (inv32)
        inv32[i] = (float)r;
}

template <typename T>
__global__ void k_diag_apply(int n, const T *__restrict__ inv, const T *__restrict__ b,
                             T *__restrict__ x, const int *sk1, const int *sk2)
{
    if ((sk1 && *sk1) || (sk2 && *sk2))
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] = b[i] * inv[i];
}

int diag_setup(rsgpu_prec *M, rsgpu_csr *A, int *bad_row)
{
    hipStream_t s = M->ctx->stream;
    if (A->n != A->ncols)
        throw RsError{RSGPU_BAD_ARGUMENT, "Matrix size differs from preconditioner size"};
    const int n = A->n;
    M->n = n;
    M->invdiag.alloc(std::max(n, 1));
    if (M->precision == RSGPU_FP32)
        M->invdiag32.alloc(std::max(n, 1));
    DevBuf<int> bad;
    int h = 0x7fffffff;
    bad.upload(&h, 1, s);
    if (n)
        hipLaunchKernelGGL(k_invdiag, dim3((n + 255) / 256), dim3(256), 0, s, n, A->rowptr.p,
                           A->colind.p, A->val.p, M->invdiag.p,
                           M->precision == RSGPU_FP32 ? M->invdiag32.p : nullptr, bad.p);
    RS_CHECK_LAUNCH();
    RS_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    if (h != 0x7fffffff) {
        if (bad_row)
            *bad_row = h - 1;
        M->ctx->err = "Zero diagonal on row " + std::to_string(h);
        return RSGPU_ZERO_PIVOT;
    }
    return RSGPU_OK;
}

void prec_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st)
{
    if (M->kind == RSGPU_PREC_AMG) {
        amg_apply(M, b, x, st, false);
        return;
    }
    if (M->kind == RSGPU_PREC_SSOR) {
        ssor_apply(M, b, x, st);
        return;
    }
    if (M->kind == RSGPU_PREC_DIAGONAL) {
        if (M->n)
            hipLaunchKernelGGL(k_diag_apply<double>, dim3((M->n + 255) / 256), dim3(256), 0,
                               M->ctx->stream, M->n, M->invdiag.p, b, x,
                               st ? &st->done : nullptr, st ? &st->cycle_end : nullptr);
        RS_CHECK_LAUNCH();
        return;
    }
    const int *sk1 = st ? &st->done : nullptr;
    const int *sk2 = st ? &st->cycle_end : nullptr;
    hipStream_t s = M->ctx->stream;
    if (M->kind == RSGPU_PREC_IDENTITY) {
        if (M->n)
            hipLaunchKernelGGL(k_copy_skip, dim3((M->n + 255) / 256), dim3(256), 0, s, b, x,
                               M->n, sk1, sk2);
        RS_CHECK_LAUNCH();
        return;
    }
    if (M->kind == RSGPU_PREC_ICC) {
        if (M->n)
            icc_apply(M, b, x, sk1, sk2);
        return;
    }
    if (M->kind == RSGPU_PREC_ILUT && ilut_apply(M, b, x, false, sk1, sk2))
        return;
    if (M->precision == RSGPU_FP64) {
        if (M->cell)
            launch_apply_cells<double>(M, M->ell64_L.p, M->ell64_U.p, b, x, sk1, sk2);
        else
            launch_apply<double, RS_MODE_ILU>(M, M->L, M->ell64_L.p, nullptr, M->U, M->ell64_U.p,
                                        M->dgU64.p, b, x, M->scratch64.p, sk1, sk2);
        return;
    }
    // FP32 factors on FP64 vectors
    M->tmpa32.alloc(M->n);
    M->tmpb32.alloc(M->n);
    convert_f64_to_f32(s, b, M->tmpa32.p, M->n);
    if (M->cell)
        launch_apply_cells<float>(M, M->ell32_L.p, M->ell32_U.p, M->tmpa32.p, M->tmpb32.p,
                                  sk1, sk2);
    else
        launch_apply<float, RS_MODE_ILU>(M, M->L, M->ell32_L.p, nullptr, M->U, M->ell32_U.p,
                                   M->dgU32.p, M->tmpa32.p, M->tmpb32.p, M->scratch32.p, sk1,
                                   sk2);
    convert_f32_to_f64(s, M->tmpb32.p, x, M->n);
}

void prec_apply(rsgpu_prec *M, const float *b, float *x, const DevState *st)
{
    const int *sk1 = st ? &st->done : nullptr;
    const int *sk2 = st ? &st->cycle_end : nullptr;
    hipStream_t s = M->ctx->stream;
    if (M->kind == RSGPU_PREC_IDENTITY) {
        if (M->n)
            hipLaunchKernelGGL(k_copy_skip32, dim3((M->n + 255) / 256), dim3(256), 0, s, b, x,
                               M->n, sk1, sk2);
        RS_CHECK_LAUNCH();
        return;
    }
    if (M->kind == RSGPU_PREC_DIAGONAL && M->precision == RSGPU_FP32) {
        if (M->n)
            hipLaunchKernelGGL(k_diag_apply<float>, dim3((M->n + 255) / 256), dim3(256), 0, s,
                               M->n, M->invdiag32.p, b, x, sk1, sk2);
        RS_CHECK_LAUNCH();
        return;
    }
    if (M->kind == RSGPU_PREC_AMG && M->precision == RSGPU_FP32) {
        amg_apply32(M, b, x, st);
        return;
    }
    if (M->precision != RSGPU_FP32 || M->kind == RSGPU_PREC_AMG ||
        M->kind == RSGPU_PREC_SSOR)
        throw RsError{RSGPU_BAD_ARGUMENT,
                      "FP32 apply needs an FP32 ILU, AMG or diagonal preconditioner"};
    if (M->cell)
        launch_apply_cells<float>(M, M->ell32_L.p, M->ell32_U.p, b, x, sk1, sk2);
    else
        launch_apply<float, RS_MODE_ILU>(M, M->L, M->ell32_L.p, nullptr, M->U, M->ell32_U.p,
                                   M->dgU32.p, b, x, M->scratch32.p, sk1, sk2);
}

void prec_trans_apply(rsgpu_prec *M, const double *b, double *x, const DevState *st)
{
    if (M->kind == RSGPU_PREC_AMG) {
        amg_apply(M, b, x, st, true);
        return;
    }
    // SSOR, DiagonalPreconditioner and ICC: transApply is apply
    if (M->kind == RSGPU_PREC_SSOR || M->kind == RSGPU_PREC_DIAGONAL ||
        M->kind == RSGPU_PREC_ICC) {
        prec_apply(M, b, x, st);
        return;
    }
    hipStream_t s = M->ctx->stream;
    if (M->kind == RSGPU_PREC_IDENTITY) {
        prec_apply(M, b, x, st);
        return;
    }
    const int *sk1 = st ? &st->done : nullptr;
    const int *sk2 = st ? &st->cycle_end : nullptr;
    if (M->precision != RSGPU_FP64)
        throw RsError{RSGPU_BAD_ARGUMENT, "transApply needs an FP64 preconditioner"};
    if (M->kind == RSGPU_PREC_ILUT && ilut_apply(M, b, x, true, sk1, sk2))
        return;
    if (!M->has_t) {
        prec_setup_transpose(M);
        launch_ell<double>(s, M->Ut, M->lu64.p, M->ell64_Ut.p, false);
        launch_ell<double>(s, M->Lt, M->lu64.p, M->ell64_Lt.p, false);
        launch_rowdiag<double>(s, M->Ut, M->diag.p, M->lu64.p, M->dgUt64.p);
    }
    launch_apply<double, RS_MODE_ILU_T>(M, M->Ut, M->ell64_Ut.p, M->dgUt64.p, M->Lt, M->ell64_Lt.p,
                               nullptr, b, x, M->scratch64.p, sk1, sk2);
}
