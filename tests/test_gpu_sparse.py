"""GPU parity: CompRowMatrix SpMV family and ILU(0) through the C ABI,
bit-exact against the CPU oracle (row sums and triangular sweeps keep the
reference's left-to-right order)."""
import numpy as np
import pytest

from ressim_amd import problems as P
from tests.helpers import oracle_system, pentadiag, random_csr, scipy_csr, tridiag

pytestmark = pytest.mark.gpu


def _mat(gpu, rp, ci, d, n, m):
    return gpu.CompRowMatrix.from_csr(n, m, rp, ci, d)


@pytest.fixture(scope="module")
def systems(oracle):
    return {"c1": oracle_system(oracle, P.config1()),
            "s3": oracle_system(oracle, P.small3d(nx=12, ny=10, nz=7))}


def test_spmv_bitexact_mpfa(gpu, oracle, systems):
    rng = np.random.default_rng(0)
    for s in systems.values():
        rp, ci, d = s["rowptr"], s["colind"], s["data"]
        n = len(rp) - 1
        A = _mat(gpu, rp, ci, d, n, n)
        x = rng.standard_normal(n)
        y = np.empty(n)
        A.mult(x, y)
        np.testing.assert_array_equal(y, oracle.mult(rp, ci, d, x))
        yt = np.empty(n)
        A.transMult(x, yt)
        np.testing.assert_array_equal(yt, oracle.transmult(rp, ci, d, x))


def test_spmv_long_empty_rows_and_device_vectors(gpu, oracle):
    import torch
    rng = np.random.default_rng(1)
    n, m = 700, 9000
    rp, ci, d = random_csr(rng, n, m, 0.003, long_rows={3, 400}, empty_rows={0, 10, 699})
    A = _mat(gpu, rp, ci, d, n, m)
    x = rng.standard_normal(m)
    y = np.empty(n)
    A.mult(x, y)
    np.testing.assert_array_equal(y, oracle.mult(rp, ci, d, x))
    xd = torch.tensor(x, device="cuda")
    y0 = rng.standard_normal(n)
    yd = torch.tensor(y0, device="cuda")
    A.multAdd(-0.75, xd, yd)
    np.testing.assert_array_equal(yd.cpu().numpy(), oracle.multadd(rp, ci, d, -0.75, x, y0))
    xt = rng.standard_normal(n)
    y1 = rng.standard_normal(m)
    ytd = torch.tensor(y1, device="cuda")
    A.transMultAdd(1.3, torch.tensor(xt, device="cuda"), ytd)
    np.testing.assert_array_equal(ytd.cpu().numpy(), oracle.transmultadd(rp, ci, d, 1.3, xt, y1))


@pytest.mark.parametrize("far", [False, True])
def test_spmv_column_offsets_fit_or_not(gpu, oracle, far):
    """70,000 rows: a banded pattern (every column within 16 bits of its
    workgroup's first row: the SpMV reads 16-bit column offsets) and the same
    band plus a column 50,000 away (it does not fit: 32-bit column indices)."""
    import torch
    rng = np.random.default_rng(11)
    n = 70000
    rows = []
    for i in range(n):
        cols = {max(0, i - 300), i, min(n - 1, i + 7)}
        if far:
            cols.add((i + 50000) % n)
        rows.append(sorted(cols))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.asarray(r, np.int32) for r in rows])
    d = rng.standard_normal(len(ci))
    A = _mat(gpu, rp, ci, d, n, n)
    x = rng.standard_normal(n)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    A.mult(torch.tensor(x, device="cuda"), y)
    np.testing.assert_array_equal(y.cpu().numpy(), oracle.mult(rp, ci, d, x))
    A.sync_f32()
    from ressim_amd import _lib
    x32 = torch.tensor(x.astype(np.float32), device="cuda")
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert _lib.load().rsgpu_csr_mult_f32(A.h, x32.data_ptr(), y32.data_ptr()) == 0
    A.ctx.synchronize()
    np.testing.assert_array_equal(y32.cpu().numpy(),
                                  oracle.mult_f32(rp, ci, d.astype(np.float32),
                                                  x.astype(np.float32)))


def test_spmv_f32_bitexact(gpu, oracle, systems):
    import torch
    s = systems["s3"]
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    A.sync_f32()
    x = np.random.default_rng(2).standard_normal(n).astype(np.float32)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty(n, dtype=torch.float32, device="cuda")
    from ressim_amd import _lib
    torch.cuda.synchronize()
    assert _lib.load().rsgpu_csr_mult_f32(A.h, xd.data_ptr(), yd.data_ptr()) == 0
    A.ctx.synchronize()
    np.testing.assert_array_equal(yd.cpu().numpy(), oracle.mult_f32(rp, ci, d.astype(np.float32),
                                                                   x))


def _banded(rng, n, half, empty=()):
    """Random pattern inside a band (every 64-row chunk's offsets fit 32
    slots), rows of varying subsets, some empty rows."""
    rows = []
    for i in range(n):
        if i in empty:
            rows.append(np.zeros(0, np.int32))
            continue
        lo, hi = max(0, i - half), min(n, i + half + 1)
        k = rng.integers(1, hi - lo + 1)
        rows.append(np.sort(rng.choice(np.arange(lo, hi), size=k, replace=False)).astype(np.int32))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return rp, np.concatenate(rows).astype(np.int32), rng.standard_normal(int(rp[-1]))


def five_point(nx, ny, rng):
    """A diagonally dominant 5-point matrix on an nx x ny grid (natural order)."""
    rows = []
    for j in range(ny):
        for i in range(nx):
            r = i + nx * j
            cols = [c for c, ok in ((r - nx, j > 0), (r - 1, i > 0), (r, True),
                                    (r + 1, i < nx - 1), (r + nx, j < ny - 1)) if ok]
            rows.append(cols)
    rp = np.zeros(nx * ny + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32)
    d = -rng.uniform(0.5, 1.5, len(ci))
    for r in range(nx * ny):
        for q in range(rp[r], rp[r + 1]):
            if ci[q] == r:
                d[q] = 6.0 + rng.uniform()
    return rp, ci, d


def _sliced_products(gpu, A, x):
    import torch
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    A.mult_sliced(xd, yd)
    return yd.cpu().numpy()


@pytest.mark.parametrize("case", ["c1", "s3", "banded"])
def test_sliced_spmv_bitexact(gpu, oracle, systems, case):
    """The sliced copy (64-row chunks sharing column-offset tables, per-row
    slot masks): FP64 and FP32 products bit-identical to the oracle's
    row-sequential sums, including rows next to non-finite x entries they do
    not reference, a partial last chunk and empty rows."""
    rng = np.random.default_rng(5)
    if case == "banded":
        n = 1000
        rp, ci, d = _banded(rng, n, 12, empty={0, 63, 64, 999})
    else:
        s = systems[case]
        rp, ci, d = s["rowptr"], s["colind"], s["data"]
        n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    nch, nslots, ntab = A.sliced_info()
    assert nch == (n + 255) // 256 and nslots % 64 == 0 and nslots >= len(ci) and ntab >= 1
    A.sync_sliced()
    A.sync_f32()
    x = rng.standard_normal(n)
    np.testing.assert_array_equal(_sliced_products(gpu, A, x), oracle.mult(rp, ci, d, x))
    x32 = x.astype(np.float32)
    np.testing.assert_array_equal(_sliced_products(gpu, A, x32),
                                  oracle.mult_f32(rp, ci, d.astype(np.float32), x32))
    # inf and NaN in x reach exactly the rows whose pattern holds them
    xb = x.copy()
    xb[n // 3] = np.inf
    xb[n // 2] = np.nan
    yo = oracle.mult(rp, ci, d, xb)
    assert np.isfinite(yo).sum() > n // 2
    np.testing.assert_array_equal(_sliced_products(gpu, A, xb), yo)
    # new values need a new sync; the copy follows them
    d2 = rng.standard_normal(len(d))
    A.set_data(d2)
    A.sync_sliced()
    np.testing.assert_array_equal(_sliced_products(gpu, A, x), oracle.mult(rp, ci, d2, x))


def test_sliced_absent_when_offsets_spread(gpu, oracle):
    """Chunks whose rows span more than 32 column offsets keep the CSR path
    only: no sliced copy, and its entry points refuse."""
    import torch
    rng = np.random.default_rng(6)
    n = 300
    rp, ci, d = random_csr(rng, n, n, 0.05)
    A = _mat(gpu, rp, ci, d, n, n)
    assert A.sliced_info()[0] == 0
    with pytest.raises(ValueError):
        A.sync_sliced()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        A.mult_sliced(x, torch.empty_like(x))


def test_sliced_env_switch_and_solve(gpu, oracle, systems, monkeypatch):
    """RSGPU_SLICED=0 builds no sliced copy; FP64 GMRES + ILU over the
    sliced copy and over CSR reach the same solution."""
    s = systems["s3"]
    rp, ci, d, b = s["rowptr"], s["colind"], s["data"], s["b"]
    n = len(rp) - 1
    monkeypatch.setenv("RSGPU_SLICED", "0")
    A0 = _mat(gpu, rp, ci, d, n, n)
    monkeypatch.delenv("RSGPU_SLICED")
    A1 = _mat(gpu, rp, ci, d, n, n)
    assert A0.sliced_info()[0] == 0 and A1.sliced_info()[0] > 0
    xs = []
    for A in (A0, A1):
        solver = gpu.GMRES(b, 30)
        solver.setIterationMonitor(gpu.DefaultIterationMonitor(100000, 1e-12, 1e-50, 1e5))
        M = gpu.ILU()
        M.setMatrix(A)
        solver.setPreconditioner(M)
        x = np.zeros(n)
        solver.solve(A, b, x)
        xs.append(x)
    assert np.linalg.norm(xs[0] - xs[1]) / np.linalg.norm(xs[0]) < 1e-9


def test_add_zero_rows_patterns(gpu, oracle):
    rng = np.random.default_rng(3)
    rp, ci, d = random_csr(rng, 50, 50, 0.1, diag=4.0)
    A = _mat(gpu, rp, ci, d, 50, 50)
    rows = np.array([3, 3, 7, 3], np.int32)
    cols = np.array([ci[rp[3]], ci[rp[3]], ci[rp[7]], ci[rp[3] + 1 if rp[4] > rp[3] + 1
                                                       else rp[3]]], np.int32)
    vals = np.array([0.1, 0.2, -1.0, 1e-3])
    A.add(rows, cols, vals)
    ref = d.copy()
    for r, c, v in zip(rows, cols, vals):
        assert oracle.csr_add(rp, ci, ref, int(r), int(c), float(v)) == 0
    np.testing.assert_array_equal(A.getData(), ref)
    missing = [c for c in range(50) if c not in ci[rp[5]:rp[6]]][0]
    with pytest.raises(IndexError):
        A.add(5, missing, 1.0)
    gpu.Matrices.zeroRows(A, 2.0, [9, 1])
    assert oracle.zero_rows(rp, ci, ref, 2.0, [9, 1]) == 0
    np.testing.assert_array_equal(A.getData(), ref)
    # CompRowMatrix(int,int,int[][]) sorts rows
    B = gpu.CompRowMatrix(3, 4, [[3, 0], [], [2, 1, 0]])
    assert B.getColumnIndices().tolist() == [0, 3, 0, 1, 2]
    with pytest.raises(ValueError):
        gpu.CompRowMatrix(2, 2, [[0, 2], [1]])


@pytest.mark.parametrize("n", range(1, 11))
def test_ilu_known_answers_gpu(gpu, oracle, n):
    """ILUTest.java:34-44 on the GPU: ILU(0) is exact on these bands."""
    for rp, ci, d in (tridiag(n), pentadiag(n)):
        A = _mat(gpu, rp, ci, d, n, n)
        M = gpu.ILU(A)
        M.setMatrix(A)
        b = scipy_csr(rp, ci, d) @ np.ones(n)
        x = np.empty(n)
        M.apply(b, x)
        assert np.linalg.norm(b - scipy_csr(rp, ci, d) @ x) < 1e-5
        st, lu, dg, _ = oracle.ilu_factor(rp, ci, d)
        np.testing.assert_array_equal(x, oracle.ilu_apply(rp, ci, lu, dg, b))


def test_ilu_factor_apply_bitexact(gpu, oracle, systems):
    rng = np.random.default_rng(4)
    for s in systems.values():
        rp, ci, d = s["rowptr"], s["colind"], s["data"]
        n = len(rp) - 1
        A = _mat(gpu, rp, ci, d, n, n)
        M = gpu.ILU(A)
        M.setMatrix(A)
        st, lu, dg, _ = oracle.ilu_factor(rp, ci, d)
        _, _, lu_gpu = M.factors()
        np.testing.assert_array_equal(lu_gpu, lu)
        b = rng.standard_normal(n)
        x = np.empty(n)
        M.apply(b, x)
        np.testing.assert_array_equal(x, oracle.ilu_apply(rp, ci, lu, dg, b))
        xt = np.empty(n)
        M.transApply(b, xt)
        np.testing.assert_array_equal(xt, oracle.ilu_transapply(rp, ci, lu, dg, b))


@pytest.mark.parametrize("nblocks", [1, 3, 16, 57])
def test_block_ilu_bitexact(gpu, oracle, systems, nblocks):
    s = systems["s3"]
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    M = gpu.BlockILU(nblocks)
    M.setMatrix(A)
    brp, bci, bd, _ = oracle.block_filter(rp, ci, d, nblocks)
    st, lu, dg, _ = oracle.ilu_factor(brp, bci, bd)
    rp_g, ci_g, lu_g = M.factors()
    np.testing.assert_array_equal(rp_g, brp)
    np.testing.assert_array_equal(ci_g, bci)
    np.testing.assert_array_equal(lu_g, lu)
    b = np.random.default_rng(nblocks).standard_normal(n)
    x = np.empty(n)
    M.apply(b, x)
    np.testing.assert_array_equal(x, oracle.ilu_apply(brp, bci, lu, dg, b))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(64, 64, 1, "lines"), (64, 64, 16, "lines"),
                                   (37, 50, 5, "lines"), (128, 64, 2, "lines"),
                                   (20, 130, 2, "levels"), (9, 7, 1, "lines"),
                                   (31, 33, 4, "levels"), (40, 48, -3, "lines"),
                                   (64, 64, -16, "lines")])
def test_block_ilu_line_wavefront(gpu, oracle, monkeypatch, precision, shape):
    """Block ILU(0) apply on 2-D 9-point MPFA-O matrices: blocks of whole
    lines (at most 64 of them) take the line-wavefront kernel, others the
    level schedules; either way the result is the oracle's, bit for bit,
    FP64 and FP32 (FP32 full blocks in both LDS layouts), and the same bits
    as with the wavefront switched off."""
    import torch
    from ressim_amd import _lib
    nx, ny, nb, form = shape
    if nb < 0:
        # a 5-point pattern (no diagonal neighbours): masks read per row
        nb = -nb
        rp, ci, d = five_point(nx, ny, np.random.default_rng(nx))
    else:
        s = oracle_system(oracle, P.config1(nx=nx, ny=ny))
        rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    brp, bci, bd, _ = oracle.block_filter(rp, ci, d, nb)
    b = np.random.default_rng(nx * ny + nb).standard_normal(n)
    b[n // 2] = np.inf          # propagates as in the sequential sweeps
    out = []
    # FP32 full blocks: both LDS layouts of the wavefront (per-line slots,
    # rows), then the level schedules
    runs = (("1", "1"), ("1", "0"), ("0", "1")) if precision == "fp32" else (("1", "1"),
                                                                          ("0", "1"))
    for band, slots in runs:
        monkeypatch.setenv("RSGPU_ILU_BAND", band)
        monkeypatch.setenv("RSGPU_BAND_SLOTS", slots)
        A = _mat(gpu, rp, ci, d, n, n)
        M = gpu.BlockILU(nb, precision=_lib.FP64 if precision == "fp64" else _lib.FP32)
        M.setMatrix(A)
        assert M.info()["sweep_form"] == (form if band == "1" else "levels")
        if precision == "fp64":
            x = np.empty(n)
            M.apply(b, x)
        else:
            bd_ = torch.tensor(b.astype(np.float32), device="cuda")
            xd = torch.empty(n, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            assert _lib.load().rsgpu_prec_apply_f32(M.h, bd_.data_ptr(), xd.data_ptr()) == 0
            M.ctx.synchronize()
            x = xd.cpu().numpy()
        out.append(x)
    if precision == "fp64":
        _, lu, dg, _ = oracle.ilu_factor(brp, bci, bd)
        want = oracle.ilu_apply(brp, bci, lu, dg, b)
    else:
        _, lu, dg, _ = oracle.ilu_factor_f32(brp, bci, bd.astype(np.float32))
        want = oracle.ilu_apply_f32(brp, bci, lu, dg, b.astype(np.float32))
    for x in out:
        np.testing.assert_array_equal(x, want)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(64, 64, 1), (64, 64, 16), (37, 50, 5), (128, 64, 2),
                                   (9, 7, 1), (3, 2, 1)])
def test_block_ilu_line_wavefront_factor(gpu, oracle, monkeypatch, precision, shape):
    """ILU(0) factorisation of full 9-point blocks by the line wavefront:
    the factor is the oracle's (ILU.java:86-125), bit for bit, FP64 and
    FP32, and the same bits as the level-scheduled factorisation
    (RSGPU_ILU_BAND_FACTOR=0)."""
    from ressim_amd import _lib
    nx, ny, nb = shape
    s = oracle_system(oracle, P.config1(nx=nx, ny=ny))
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    brp, bci, bd, _ = oracle.block_filter(rp, ci, d, nb)
    if precision == "fp64":
        st, want, _, _ = oracle.ilu_factor(brp, bci, bd)
    else:
        st, want, _, _ = oracle.ilu_factor_f32(brp, bci, bd.astype(np.float32))
        want = want.astype(np.float64)
    assert st == 0
    print(f"band factor {precision} {nx}x{ny} blocks={nb}", flush=True)
    for flag in ("1", "0"):
        monkeypatch.setenv("RSGPU_ILU_BAND_FACTOR", flag)
        A = _mat(gpu, rp, ci, d, n, n)
        M = gpu.BlockILU(nb, precision=_lib.FP64 if precision == "fp64" else _lib.FP32)
        M.setMatrix(A)
        assert M.info()["sweep_form"] == "lines"
        rp_g, ci_g, lu_g = M.factors()
        np.testing.assert_array_equal(rp_g, brp)
        np.testing.assert_array_equal(ci_g, bci)
        assert np.array_equal(lu_g.view(np.int64), want.view(np.int64)), flag
        # a second factorisation on new values (same pattern) reuses the plan
        d2 = d * (1.0 + 0.25 * np.random.default_rng(nx + nb).standard_normal(d.size))
        A2 = _mat(gpu, rp, ci, d2, n, n)
        M.setMatrix(A2)
        _, _, bd2, _ = oracle.block_filter(rp, ci, d2, nb)
        if precision == "fp64":
            _, want2, _, _ = oracle.ilu_factor(brp, bci, bd2)
        else:
            _, want2, _, _ = oracle.ilu_factor_f32(brp, bci, bd2.astype(np.float32))
            want2 = want2.astype(np.float64)
        assert np.array_equal(M.factors()[2].view(np.int64), want2.view(np.int64)), flag


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_block_ilu_line_wavefront_zero_pivot(gpu, oracle, monkeypatch, precision):
    """A zero pivot inside a full 9-point block: the line-wavefront
    factorisation hands the block set back to the level kernel, which
    reports the reference's error on the reference's row."""
    from ressim_amd import _lib
    s = oracle_system(oracle, P.config1(nx=16, ny=12))
    rp, ci, d = s["rowptr"], s["colind"], s["data"].copy()
    n = len(rp) - 1
    r = 16 * 5 + 7
    # row r's lower entries and diagonal set to zero: its elimination
    # subtracts exact zeros, so U(r,r) = 0 and row r + 1 meets a zero pivot
    lo = rp[r] + np.nonzero(ci[rp[r]:rp[r + 1]] <= r)[0]
    d[lo] = 0.0
    st, _, _, _ = oracle.ilu_factor(rp, ci, d)
    assert st != 0
    msgs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("RSGPU_ILU_BAND_FACTOR", flag)
        A = _mat(gpu, rp, ci, d, n, n)
        M = gpu.BlockILU(1, precision=_lib.FP64 if precision == "fp64" else _lib.FP32)
        with pytest.raises(RuntimeError, match="Zero pivot") as ei:
            M.setMatrix(A)
        msgs.append(str(ei.value))
    assert msgs[0] == msgs[1]


@pytest.mark.parametrize("kernel", ["pf", "generic"])
@pytest.mark.parametrize("nblocks", [1, 3, 16])
def test_block_ilu_sweep_kernels(gpu, oracle, systems, monkeypatch, kernel, nblocks):
    """Both sweep kernels (pipelined, generic) give the oracle's bits for
    apply and transApply, FP64 and FP32, on a 3-D 27-point pattern (rows of
    up to 13 entries per triangle: 16-slot lanes)."""
    import torch
    from ressim_amd import _lib
    monkeypatch.setenv("RSGPU_ILU_KERNEL", kernel)
    s = systems["s3"]
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    print(f"sweep kernel {kernel} nblocks {nblocks} n {n}", flush=True)
    A = _mat(gpu, rp, ci, d, n, n)
    brp, bci, bd, _ = oracle.block_filter(rp, ci, d, nblocks)
    st, lu, dg, _ = oracle.ilu_factor(brp, bci, bd)
    rng = np.random.default_rng(nblocks + 100)
    M = gpu.BlockILU(nblocks)
    M.setMatrix(A)
    b = rng.standard_normal(n)
    x = np.empty(n)
    M.apply(b, x)
    np.testing.assert_array_equal(x, oracle.ilu_apply(brp, bci, lu, dg, b))
    xt = np.empty(n)
    M.transApply(b, xt)
    np.testing.assert_array_equal(xt, oracle.ilu_transapply(brp, bci, lu, dg, b))
    M32 = gpu.BlockILU(nblocks, precision=_lib.FP32)
    M32.setMatrix(A)
    st, lu32, dg32, _ = oracle.ilu_factor_f32(brp, bci, bd.astype(np.float32))
    b32 = b.astype(np.float32)
    bd32 = torch.tensor(b32, device="cuda")
    xd32 = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert _lib.load().rsgpu_prec_apply_f32(M32.h, bd32.data_ptr(), xd32.data_ptr()) == 0
    M32.ctx.synchronize()
    np.testing.assert_array_equal(xd32.cpu().numpy(),
                                  oracle.ilu_apply_f32(brp, bci, lu32, dg32, b32))


@pytest.mark.parametrize("kernel", ["default", "generic"])
def test_ilu_global_scratch_path(gpu, oracle, monkeypatch, kernel):
    """A 9000-row block: the pipelined kernel keeps its FP64 vector in LDS
    beyond the 64 KiB default allocation (opt-in); the generic kernel runs
    the sweeps through the global scratch region.  Both bit-exact."""
    if kernel != "default":
        monkeypatch.setenv("RSGPU_ILU_KERNEL", kernel)
    print(f"9000-row ILU block, kernel {kernel}", flush=True)
    s = oracle_system(oracle, P.config1(nx=100, ny=90))
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    M = gpu.ILU(A)
    M.setMatrix(A)
    st, lu, dg, _ = oracle.ilu_factor(rp, ci, d)
    b = np.random.default_rng(9).standard_normal(n)
    x = np.empty(n)
    M.apply(b, x)
    np.testing.assert_array_equal(x, oracle.ilu_apply(rp, ci, lu, dg, b))


def test_ilu_f32_bitexact(gpu, oracle, systems):
    import torch
    s = systems["s3"]
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    from ressim_amd import _lib
    M = gpu.ILU(A, precision=_lib.FP32)
    M.setMatrix(A)
    st, lu, dg, _ = oracle.ilu_factor_f32(rp, ci, d.astype(np.float32))
    b = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    bd = torch.tensor(b, device="cuda")
    xd = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert _lib.load().rsgpu_prec_apply_f32(M.h, bd.data_ptr(), xd.data_ptr()) == 0
    M.ctx.synchronize()
    np.testing.assert_array_equal(xd.cpu().numpy(), oracle.ilu_apply_f32(rp, ci, lu, dg, b))


def test_ilu_errors_gpu(gpu):
    rp, ci, d = tridiag(5)
    d = d.copy()
    d[rp[0]] = 0.0                # row 1 is eliminated with a zero pivot
    A = _mat(gpu, rp, ci, d, 5, 5)
    M = gpu.ILU(A)
    with pytest.raises(RuntimeError, match="Zero pivot"):
        M.setMatrix(A)
    B = gpu.CompRowMatrix(3, 3, [[0, 1], [0], [1, 2]])
    M2 = gpu.ILU(B)
    with pytest.raises(RuntimeError, match="Missing diagonal entry on row 2"):
        M2.setMatrix(B)


def test_matrix_market_round_trip_through_device(gpu, oracle, systems):
    """A CompRowMatrix written in Matrix Market form (17 digits) and read back
    (CompRowMatrix(MatrixVectorReader)) holds the same pattern and values, and
    its device SpMV equals the oracle's."""
    import io
    s = systems["c1"]
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    B = gpu.CompRowMatrix.from_matrix_market(io.StringIO(A.to_matrix_market(digits=17)))
    nz = d != 0
    assert B.nnz == int(nz.sum())
    x = np.random.default_rng(9).standard_normal(n)
    y = np.empty(n)
    B.mult(x, y)
    np.testing.assert_array_equal(y, oracle.mult(rp, ci, d, x))


def test_create_rejects_rowptr_past_nnz(gpu):
    """rowptr = [0, 3, 1] with nnz = 1: row 0 claims entries past the colind
    array; rejected before any column index is read (scalar and 3x3 block
    patterns)."""
    import ctypes as C
    from ressim_amd import _lib as L
    lib = L.load()
    ctx = gpu.Context.default()
    rp = np.array([0, 3, 1], np.int32)
    ci = np.array([0], np.int32)
    h = C.c_void_p()
    rc = lib.rsgpu_csr_create(ctx.h, 2, 2, 1, rp.ctypes.data_as(C.c_void_p),
                              ci.ctypes.data_as(C.c_void_p), C.byref(h))
    assert rc == L.BAD_ARGUMENT and not h.value
    assert b"rowptr" in lib.rsgpu_last_error(ctx.h)
    rc = lib.rsgpu_csr_create_bsr3(ctx.h, 2, 1, rp.ctypes.data_as(C.c_void_p),
                                   ci.ctypes.data_as(C.c_void_p), C.byref(h))
    assert rc == L.BAD_ARGUMENT and not h.value
    with pytest.raises(ValueError):
        gpu.CompRowMatrix.from_csr(3, 3, np.array([0, 1, 2], np.int32), np.zeros(2, np.int32),
                                   np.zeros(2))


def test_stale_copies_are_refused(gpu, oracle, systems):
    """A value change after sync_f32 / sync_sliced makes the FP32 and sliced
    copies stale: the calls that read them refuse until the next sync."""
    import torch
    from ressim_amd import _lib as L
    lib = L.load()
    s = systems["s3"]
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = _mat(gpu, rp, ci, d, n, n)
    x = torch.tensor(np.random.default_rng(5).standard_normal(n), device="cuda")
    x32, y32 = x.float(), torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    A.sync_f32()
    assert lib.rsgpu_csr_mult_f32(A.h, x32.data_ptr(), y32.data_ptr()) == L.OK
    A.add(np.array([0], np.int32), np.array([0], np.int32), np.array([1.0]))
    assert lib.rsgpu_csr_mult_f32(A.h, x32.data_ptr(), y32.data_ptr()) == L.BAD_ARGUMENT
    assert b"sync_f32" in lib.rsgpu_last_error(A.ctx.h)
    A.sync_f32()
    assert lib.rsgpu_csr_mult_f32(A.h, x32.data_ptr(), y32.data_ptr()) == L.OK
    nch, _, _ = A.sliced_info()
    assert nch > 0
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    A.sync_sliced()
    A.mult_sliced(x, y)
    A.mult_sliced(x32, y32)
    gpu.Matrices.zeroRows(A, 1.0, [1])
    with pytest.raises(ValueError, match="sync_sliced"):
        A.mult_sliced(x, y)
    with pytest.raises(ValueError, match="sync_f32"):
        A.mult_sliced(x32, y32)
    A.sync_sliced()
    A.mult_sliced(x, y)
    ref = torch.empty_like(y)
    A.mult(x, ref)
    assert torch.equal(y, ref)
