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
