"""Device black-oil Jacobian and the 3x3 block matrix path against the oracle.

Assembly: pattern, values and residual bit-exact against oracle/blackoil.c.
Block SpMV: bit-exact against CompRowMatrix.mult on the expanded CSR (the
oracle's or_csr_mult).  ILU(0) on the block matrix: bit-exact against MTJ's
scalar ILU(0) on the expanded CSR (the reference-side oracle for the block
path).  Solves: BiCGstab + block-Jacobi ILU(0), FP64 and with the FP32
preconditioner, against the oracle's solution.
"""
import numpy as np
import pytest

from ressim_amd import problems as P

pytestmark = pytest.mark.gpu


def _device_jacobian(gpu, pr, u=None, u_old=None):
    import torch
    from ressim_amd.blackoil import BlackOilJacobian
    bo = BlackOilJacobian(pr)
    J = bo.matrix()
    ud = torch.tensor(pr.u if u is None else u, device="cuda")
    uo = torch.tensor(pr.u_old if u_old is None else u_old, device="cuda")
    R = torch.empty(3 * pr.n, dtype=torch.float64, device="cuda")
    bo.assemble(ud, uo, J, R)
    return bo, J, R


@pytest.mark.parametrize("shape", [(7, 5, 4), (1, 6, 3), (9, 1, 2)])
def test_assembly_bitexact(gpu, oracle, shape):
    nx, ny, nz = shape
    pr = P.config3(nx=nx, ny=ny, nz=nz)
    rng = np.random.default_rng(1)
    u = pr.u.copy()
    u[0::3] += rng.uniform(-300, 300, pr.n)
    u_old = pr.u.copy()
    u_old[0::3] -= 1e5
    bo, J, R = _device_jacobian(gpu, pr, u, u_old)
    brp, bci = oracle.bo_pattern(nx, ny, nz)
    rp, ci = oracle.bo_expand(brp, bci)
    assert J.block_info() == (3, pr.n, len(bci))
    assert np.array_equal(J.getRowPointers(), rp)
    assert np.array_equal(J.getColumnIndices(), ci)
    val, Ro = oracle.bo_assemble(pr.params, nx, ny, nz, pr.kh, u, u_old)
    assert np.array_equal(J.getData(), val)
    assert np.array_equal(R.cpu().numpy(), Ro)


def test_block_spmv_bitexact(gpu, oracle):
    import torch
    pr = P.config3(nx=12, ny=9, nz=5)
    bo, J, R = _device_jacobian(gpu, pr)
    rp, ci, val = J.getRowPointers(), J.getColumnIndices(), J.getData()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(J.n)
    y = torch.zeros(J.n, dtype=torch.float64, device="cuda")
    J.mult(torch.tensor(x, device="cuda"), y)
    assert np.array_equal(y.cpu().numpy(), oracle.mult(rp, ci, val, x))
    # FP32 values and vectors
    from ressim_amd import _lib
    J.sync_f32()
    x32 = x.astype(np.float32)
    y32 = torch.zeros(J.n, dtype=torch.float32, device="cuda")
    xd = torch.tensor(x32, device="cuda")
    torch.cuda.synchronize()
    assert _lib.load().rsgpu_csr_mult_f32(J.h, xd.data_ptr(), y32.data_ptr()) == 0
    J.ctx.synchronize()
    assert np.array_equal(y32.cpu().numpy(), oracle.mult_f32(rp, ci, val.astype(np.float32),
                                                             x32))


def _block_sliced_check(gpu, oracle, A, rp, ci, data, rng):
    import torch
    n = A.n
    nch, nslots, ntab = A.sliced_info()
    assert nch == (n + 255) // 256 and nslots % 64 == 0 and nslots >= len(ci) and ntab >= 1
    A.sync_sliced()
    A.sync_f32()
    x = rng.standard_normal(n)

    def run(xh):
        xd = torch.tensor(xh, device="cuda")
        yd = torch.empty_like(xd)
        A.mult_sliced(xd, yd)
        return yd.cpu().numpy()
    np.testing.assert_array_equal(run(x), oracle.mult(rp, ci, data, x))
    x32 = x.astype(np.float32)
    np.testing.assert_array_equal(run(x32),
                                  oracle.mult_f32(rp, ci, data.astype(np.float32), x32))
    # the block-row core on the same matrix
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    A.mult(torch.tensor(x, device="cuda"), y)
    np.testing.assert_array_equal(y.cpu().numpy(), run(x))
    # inf and NaN reach exactly the rows whose pattern holds them
    xb = x.copy()
    xb[n // 3] = np.inf
    xb[n // 2 + 1] = np.nan
    yo = oracle.mult(rp, ci, data, xb)
    assert np.isfinite(yo).sum() > n // 2
    np.testing.assert_array_equal(run(xb), yo)
    # new values need a new sync; the copy follows them
    d2 = rng.standard_normal(len(ci))
    A.set_data(d2)
    A.sync_sliced()
    np.testing.assert_array_equal(run(x), oracle.mult(rp, ci, d2, x))


@pytest.mark.parametrize("shape", [(12, 9, 5), (23, 7, 1), (50, 1, 1)])
def test_block_sliced_spmv_bitexact(gpu, oracle, shape):
    """The block form of the sliced copy on black-oil Jacobians (seven-,
    five- and three-entry block stencils; 256-row chunks split block rows,
    a partial last chunk): FP64 and FP32 products bit-identical to the
    oracle's row-sequential sums of the expanded CSR and to the block-row
    core, non-finite x entries reaching only the rows that hold them."""
    nx, ny, nz = shape
    pr = P.config3(nx=nx, ny=ny, nz=nz)
    bo, J, R = _device_jacobian(gpu, pr)
    rp, ci = J.getRowPointers(), J.getColumnIndices()
    ntab_expected = {(12, 9, 5): 7, (23, 7, 1): 5, (50, 1, 1): 3}[shape]
    _block_sliced_check(gpu, oracle, J, rp, ci, J.getData(),
                        np.random.default_rng(8 + nx))
    # no chunk's table is wider than the stencil
    nch, nslots, _ = J.sliced_info()
    assert nslots <= nch * 256 * 3 * ntab_expected


def test_block_sliced_generic_patterns(gpu, oracle):
    """Random block patterns within a band: eight to ten block offsets per
    chunk (the wider load batch), an empty block row, and a pattern whose
    chunks span more than ten block offsets (no sliced copy, refused)."""
    import torch
    rng = np.random.default_rng(9)
    nb = 700
    rows = []
    for I in range(nb):
        if I in (5, 86, 699):
            rows.append(np.zeros(0, np.int32))
            continue
        cand = np.arange(max(0, I - 4), min(nb, I + 5))
        k = int(rng.integers(1, len(cand) + 1))
        rows.append(np.sort(rng.choice(cand, size=k, replace=False)).astype(np.int32))
    brp = np.zeros(nb + 1, np.int32)
    brp[1:] = np.cumsum([len(r) for r in rows])
    bci = np.concatenate(rows)
    A = gpu.CompRowMatrix.from_bsr3(nb, brp, bci)
    rp, ci = oracle.bo_expand(brp, bci)
    data = rng.standard_normal(len(ci))
    A.set_data(data)
    assert A.sliced_info()[0] > 0
    _block_sliced_check(gpu, oracle, A, rp, ci, data, rng)
    # offsets spread over more than ten block columns: block-row core only
    rows = [np.sort(rng.choice(nb, size=4, replace=False)).astype(np.int32) for _ in range(nb)]
    brp[1:] = np.cumsum([len(r) for r in rows])
    B = gpu.CompRowMatrix.from_bsr3(nb, brp, np.concatenate(rows))
    assert B.sliced_info()[0] == 0
    x = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        B.mult_sliced(x, torch.empty_like(x))


def test_generic_bsr3_with_long_and_empty_block_rows(gpu, oracle):
    """A block pattern with an empty block row and one block row longer
    than the SpMV stage (its three scalar rows are summed one by one)."""
    import torch
    rng = np.random.default_rng(4)
    nb = 400
    rows = []
    for I in range(nb):
        if I == 7:
            rows.append(np.zeros(0, np.int32))
        elif I == 11:
            rows.append(np.arange(0, nb, 1, dtype=np.int32)[:330])
        else:
            k = int(rng.integers(1, 9))
            rows.append(np.sort(rng.choice(nb, size=k, replace=False)).astype(np.int32))
    brp = np.zeros(nb + 1, np.int32)
    brp[1:] = np.cumsum([len(r) for r in rows])
    bci = np.concatenate(rows)
    A = gpu.CompRowMatrix.from_bsr3(nb, brp, bci)
    rp, ci = oracle.bo_expand(brp, bci)
    assert np.array_equal(A.getRowPointers(), rp) and np.array_equal(A.getColumnIndices(), ci)
    data = rng.standard_normal(len(ci))
    A.set_data(data)
    x = rng.standard_normal(3 * nb)
    y = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    A.mult(torch.tensor(x, device="cuda"), y)
    assert np.array_equal(y.cpu().numpy(), oracle.mult(rp, ci, data, x))


def test_ilu_on_block_matrix_is_scalar_ilu(gpu, oracle):
    import torch
    pr = P.config3(nx=6, ny=5, nz=3)
    bo, J, R = _device_jacobian(gpu, pr)
    M = gpu.ILU(J)
    M.setMatrix(J)
    rp, ci, lu = M.factors()
    st, luo, diag, _ = oracle.ilu_factor(J.getRowPointers(), J.getColumnIndices(), J.getData())
    assert st == 0 and np.array_equal(lu, luo)
    rng = np.random.default_rng(6)
    b = rng.standard_normal(J.n)
    z = torch.zeros(J.n, dtype=torch.float64, device="cuda")
    M.apply(torch.tensor(b, device="cuda"), z)
    assert np.array_equal(z.cpu().numpy(), oracle.ilu_apply(rp, ci, luo, diag, b))


def _plane_blocks(pr, lines):
    """Block boundaries of whole x-lines of one plane (3 scalar rows per cell)."""
    from ressim_amd.ensemble import grid_blocks
    return 3 * grid_blocks([pr], lines * pr.nx)


@pytest.mark.parametrize("mixed", [False, True])
def test_bicgstab_block_ilu_solves_newton_system(gpu, oracle, mixed):
    """FP64 BiCGstab + FP64 block ILU(0); and the mixed form: FP64
    refinement on the true residual around FP32 BiCGstab with the FP32
    matrix and FP32 block ILU(0)."""
    import torch
    from ressim_amd import _lib
    pr = P.config3(nx=16, ny=12, nz=6)
    bo, J, R = _device_jacobian(gpu, pr)
    b = -R
    M = gpu.BlockILU(1, precision=_lib.FP32 if mixed else _lib.FP64)
    M.setBlocks(_plane_blocks(pr, 6))
    if mixed:
        J.sync_f32()
    M.setMatrix(J)
    if mixed:
        s = gpu.MixedPrecisionSolver(b, inner="bicgstab", inner_rtol=1e-3, inner_max_iter=1000)
        s.setIterationMonitor(gpu.DefaultIterationMonitor(50, 1e-10, 1e-300, 1e5))
    else:
        s = gpu.BiCGstab(b)
        s.setIterationMonitor(gpu.DefaultIterationMonitor(2000, 1e-10, 1e-300, 1e5))
    s.setPreconditioner(M)
    x = torch.zeros(J.n, dtype=torch.float64, device="cuda")
    s.solve(J, b, x)
    st = s.last_stats
    assert st.true_residual <= 1e-9 * st.b_norm
    xo, so, _ = oracle.solve("bicgstab", J.getRowPointers(), J.getColumnIndices(), J.getData(),
                             b.cpu().numpy(), rtol=1e-12, atol=1e-300, prec="ilu")
    assert so.status == 0
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    assert err < 1e-7, err


def test_gmres_on_block_matrix(gpu, oracle):
    """GMRES(20) + ILU(0) on the block matrix (the block-row SpMV inside the
    inner steps) against the oracle's GMRES on the expanded CSR."""
    import torch
    pr = P.config3(nx=10, ny=8, nz=4)
    bo, J, R = _device_jacobian(gpu, pr)
    b = -R
    M = gpu.ILU(J)
    M.setMatrix(J)
    s = gpu.GMRES(b, 20)
    s.setPreconditioner(M)
    s.setIterationMonitor(gpu.DefaultIterationMonitor(3000, 1e-11, 1e-300, 1e5))
    x = torch.zeros(J.n, dtype=torch.float64, device="cuda")
    s.solve(J, b, x)
    st = s.last_stats
    assert st.true_residual <= 1e-10 * st.b_norm
    xo, so, _ = oracle.solve("gmres", J.getRowPointers(), J.getColumnIndices(), J.getData(),
                             b.cpu().numpy(), rtol=1e-11, atol=1e-300, restart=20, prec="ilu")
    assert so.status == 0
    assert abs(st.iterations - so.iterations) <= max(3, so.iterations // 10)
    assert np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo) < 1e-8


def _filter_blocks(rp, ci, val, bstart):
    """Entries of A whose row and column lie in the same block (the
    block-Jacobi pattern), as CSR."""
    n = len(rp) - 1
    blk = np.searchsorted(bstart, np.arange(n), side="right") - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = blk[rows] == blk[ci]
    frp = np.zeros(n + 1, np.int32)
    frp[1:] = np.cumsum(np.bincount(rows[keep], minlength=n))
    return frp, ci[keep].astype(np.int32), val[keep]


@pytest.mark.parametrize("stream", ["streamed", "registers"])
@pytest.mark.parametrize("blocks", ["whole", "lines2", "lines6", "lines16", "planes"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_cell_ilu_bitexact(gpu, oracle, monkeypatch, blocks, prec, stream):
    """ILU(0) on a 3x3 block matrix with cell-aligned blocks runs cell
    schedules (three rows per lane, one level per cell wavefront): factors
    and apply bit-exact against the oracle's scalar ILU(0) of the
    block-filtered matrix, in FP64 and FP32, with the level records
    streamed through LDS rings and with the register pipeline; and with
    fewer levels than the row schedule of the same blocks.  Line blocks (two
    neighbour blocks per triangle, levels of up to 16 cells) and the 2-plane
    blocks (three, up to 14 cells: 4,032 B of FP64 U records) stream; the
    whole 11x7x4 block has levels of 28 cells, 8,064 B in FP64 (past the
    rings' 4 KiB: registers) and 4,032 B in FP32 (streamed, with 16-byte
    index copies)."""
    import torch
    from ressim_amd import _lib
    if stream == "registers":
        monkeypatch.setenv("RSGPU_ILU_CELL_STREAM", "0")
    pr = P.config3(nx=11, ny=7, nz=4) if blocks != "lines16" else P.config3(nx=21, ny=16,
                                                                            nz=3)
    rng = np.random.default_rng(12)
    u = pr.u.copy()
    u[0::3] += rng.uniform(-300, 300, pr.n)
    bo, J, R = _device_jacobian(gpu, pr, u)
    n = J.n
    if blocks == "whole":
        bstart = np.array([0, n], np.int32)
    elif blocks == "planes":
        bstart = np.arange(0, n + 1, 3 * pr.nx * pr.ny * 2, dtype=np.int32)
    else:
        bstart = _plane_blocks(pr, int(blocks[5:])).astype(np.int32)
    print(f"cell ILU blocks {blocks} ({len(bstart) - 1}) prec {prec}", flush=True)
    precision = _lib.FP32 if prec == "f32" else _lib.FP64
    if prec == "f32":
        J.sync_f32()

    def make():
        M = gpu.BlockILU(1, precision=precision)
        M.setBlocks(bstart)
        M.setMatrix(J)
        return M

    M = make()
    frp, fci, fval = _filter_blocks(J.getRowPointers(), J.getColumnIndices(), J.getData(),
                                    bstart)
    rp_g, ci_g, lu_g = M.factors()
    assert np.array_equal(rp_g, frp) and np.array_equal(ci_g, fci)
    b = rng.standard_normal(n)
    if prec == "f64":
        st, lu, dg, _ = oracle.ilu_factor(frp, fci, fval)
        assert st == 0 and np.array_equal(lu_g, lu)
        z = torch.zeros(n, dtype=torch.float64, device="cuda")
        M.apply(torch.tensor(b, device="cuda"), z)
        assert np.array_equal(z.cpu().numpy(), oracle.ilu_apply(frp, fci, lu, dg, b))
    else:
        st, lu, dg, _ = oracle.ilu_factor_f32(frp, fci, fval.astype(np.float32))
        assert st == 0 and np.array_equal(lu_g.astype(np.float32), lu)
        b32 = b.astype(np.float32)
        bd = torch.tensor(b32, device="cuda")
        z = torch.zeros(n, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        assert _lib.load().rsgpu_prec_apply_f32(M.h, bd.data_ptr(), z.data_ptr()) == 0
        M.ctx.synchronize()
        assert np.array_equal(z.cpu().numpy(), oracle.ilu_apply_f32(frp, fci, lu, dg, b32))
    # which sweep kernel ran
    form = M.info()["sweep_form"]
    streams = stream == "streamed" and not (blocks == "whole" and prec == "f64")
    assert form == ("cells_streamed" if streams else "cells"), form
    # the cell schedule has about a third of the row schedule's levels
    cells = M.info()
    monkeypatch.setenv("RSGPU_ILU_CELLS", "0")
    rows = make().info()
    assert 2 * cells["levels_lower"] < rows["levels_lower"]
    assert 2 * cells["levels_upper"] < rows["levels_upper"]


def test_cell_ilu_transapply(gpu, oracle):
    """transApply of a cell-scheduled ILU keeps the row sweeps of the
    transposed factors: bit-exact against the oracle."""
    import torch
    pr = P.config3(nx=9, ny=6, nz=3)
    bo, J, R = _device_jacobian(gpu, pr)
    M = gpu.BlockILU(1)
    M.setBlocks(_plane_blocks(pr, 3))
    M.setMatrix(J)
    frp, fci, fval = _filter_blocks(J.getRowPointers(), J.getColumnIndices(), J.getData(),
                                    _plane_blocks(pr, 3).astype(np.int32))
    st, lu, dg, _ = oracle.ilu_factor(frp, fci, fval)
    b = np.random.default_rng(3).standard_normal(J.n)
    xt = torch.zeros(J.n, dtype=torch.float64, device="cuda")
    M.transApply(torch.tensor(b, device="cuda"), xt)
    assert np.array_equal(xt.cpu().numpy(), oracle.ilu_transapply(frp, fci, lu, dg, b))


@pytest.mark.parametrize("factor", ["cells", "rows"])
def test_cell_ilu_zero_pivot(gpu, monkeypatch, factor):
    """A zero diagonal in the first cell: both factorisations of the cell
    schedule stop with "Zero pivot" at the first entry the reference divides
    by it, (row 1, column 0)."""
    if factor == "rows":
        monkeypatch.setenv("RSGPU_ILU_CELL_FACTOR", "0")
    print(f"cell ILU zero pivot, factor {factor}", flush=True)
    pr = P.config3(nx=5, ny=4, nz=2)
    bo, J, R = _device_jacobian(gpu, pr)
    rp, ci, val = J.getRowPointers(), J.getColumnIndices(), J.getData().copy()
    val[rp[0] + list(ci[rp[0]:rp[1]]).index(0)] = 0.0
    J.set_data(val)
    M = gpu.ILU(J)
    with pytest.raises(RuntimeError, match="Zero pivot"):
        M.setMatrix(J)
    assert M.bad_row == rp[1] + list(ci[rp[1]:rp[2]]).index(0)
