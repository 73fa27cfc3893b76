"""GPU parity: smoothed-aggregation AMG (mtj/.../sparse/AMG.java) through the
C ABI against the oracle's restatement with the same aggregation variant
(prioritised parallel rounds) and smoother blocks.

The hierarchy (aggregates, prolongator, Galerkin operators) and one cycle are
compared bit for bit: every entry and every row sum is evaluated in the
reference's order on both sides.  Solves are compared to round-off (the
Krylov dot products are tree sums on the GPU)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from ressim_amd import problems as P
from tests.helpers import oracle_system, scipy_csr

pytestmark = pytest.mark.gpu


def lap1d(n):
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    A.sort_indices()
    return A


def aniso2d(n1, eps=0.01):
    I = sp.identity(n1)
    A = (sp.kron(I, lap1d(n1)) + eps * sp.kron(lap1d(n1), I)).tocsr()
    A.sort_indices()
    return A


def _sys(oracle, name):
    if name == "lap1d":
        A = lap1d(300)
        return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    if name == "aniso":
        A = aniso2d(30)
        return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    pr = {"c1": P.config1(nx=40, ny=36), "s3": P.small3d(nx=14, ny=12, nz=9),
          "c1big": P.config1(nx=96, ny=96), "c2m": P.config2(nx=32, ny=32, nz=16)}[name]
    d = oracle_system(oracle, pr)
    return d["rowptr"], d["colind"], d["data"]


CASES = ["lap1d", "aniso", "c1", "s3"]


def _pair(gpu, oracle, rp, ci, v, **kw):
    n = len(rp) - 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    M = gpu.AMG(**kw)
    M.setMatrix(A)
    okw = dict(kw)
    okw.setdefault("block_rows", 0)
    okw.setdefault("ordering", 1)
    okw["mode"] = 1
    okw.setdefault("max_coarse", 2048)
    H = oracle.AMG(rp, ci, v, **okw)
    return A, M, H


@pytest.mark.parametrize("name", CASES)
def test_hierarchy_bit_exact(gpu, oracle, name):
    rp, ci, v = _sys(oracle, name)
    A, M, H = _pair(gpu, oracle, rp, ci, v, min=20)
    lg, lo = M.levels(), H.levels()
    print(name, "levels", lg, flush=True)
    assert lg == lo and len(lg) >= 2
    for k in range(len(lg)):
        g, o = M.level(k), H.level(k)
        for a, b in zip(g, o):
            if b is None:
                assert a is None
                continue
            assert a.dtype == b.dtype and a.shape == b.shape
            assert np.array_equal(a, b), f"level {k}"


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("zero_guess", [0, 1])
@pytest.mark.parametrize("block_rows,ordering", [(0, 1), (0, 0), (64, 0), (64, 1)])
def test_cycle_bit_exact(gpu, oracle, name, zero_guess, block_rows, ordering):
    rp, ci, v = _sys(oracle, name)
    n = len(rp) - 1
    A, M, H = _pair(gpu, oracle, rp, ci, v, min=20, zero_guess=zero_guess,
                    block_rows=block_rows, ordering=ordering)
    rng = np.random.default_rng(11)
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    xg = M.apply(b, x0.copy())
    xo = H.apply(b, x0)
    assert np.array_equal(xg, xo)
    xg = M.transApply(b, x0.copy())
    xo = H.apply(b, x0, transpose=True)
    assert np.array_equal(xg, xo)


@pytest.mark.parametrize("ordering", [1, 0])
def test_cycle_global_sweep_bit_exact(gpu, oracle, ordering):
    """A level above the LDS vector (9216 rows) is swept in place, one launch
    per level of its schedule."""
    rp, ci, v = _sys(oracle, "c1big")
    n = len(rp) - 1
    A, M, H = _pair(gpu, oracle, rp, ci, v, zero_guess=0, block_rows=0, ordering=ordering)
    rng = np.random.default_rng(4)
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    assert np.array_equal(M.apply(b, x0.copy()), H.apply(b, x0))


def test_setmatrix_again_same_pattern(gpu, oracle):
    """setMatrix rebuilds the hierarchy from the new values (the smoother's
    level-0 schedule is kept for the unchanged pattern)."""
    rp, ci, v = _sys(oracle, "c1")
    n = len(rp) - 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    M = gpu.AMG(min=20, zero_guess=1, block_rows=128)
    M.setMatrix(A)
    v2 = v * 1.5
    v2[::7] *= 1.01
    A.set_data(v2)
    M.setMatrix(A)
    H = oracle.AMG(rp, ci, v2, min=20, zero_guess=1, block_rows=128, mode=1, ordering=1)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(n)
    assert np.array_equal(M.apply(b, np.zeros(n)), H.apply(b, np.zeros(n)))


@pytest.mark.parametrize("name", ["c1", "s3"])
def test_gmres_amg_solve(gpu, oracle, name):
    rp, ci, v = _sys(oracle, name)
    n = len(rp) - 1
    d = oracle_system(oracle, {"c1": P.config1(nx=40, ny=36),
                               "s3": P.small3d(nx=14, ny=12, nz=9)}[name])
    b = d["b"]
    xo, so, _ = oracle.solve("gmres", rp, ci, v, b, rtol=1e-10, prec="amg",
                             amg=dict(zero_guess=1, mode=1, block_rows=0, ordering=1),
                             max_iter=500)
    assert so.status == 0
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    M = gpu.AMG(zero_guess=1)
    M.setMatrix(A)
    s = gpu.GMRES(b, 30)
    s.setIterationMonitor(gpu.DefaultIterationMonitor(500, 1e-10, 1e-50, 1e5))
    s.setPreconditioner(M)
    x = np.zeros(n)
    s.solve(A, b, x)
    As = scipy_csr(rp, ci, v)
    assert np.linalg.norm(b - As @ x) / np.linalg.norm(b) < 1e-9
    xe = spl.spsolve(As.tocsc(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8
    assert abs(s.getIterationMonitor().iterations() - so.iterations) <= 2


def test_missing_diagonal(gpu):
    A = sp.csr_matrix(np.array([[2.0, -1, 0], [-1, 0, -1], [0, -1, 2]]))
    rp, ci = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    keep = ~((np.repeat(np.arange(3), np.diff(rp)) == 1) & (ci == 1))
    ci2 = ci[keep]
    v2 = A.data[keep]
    rp2 = np.array([0, 2, 4, 6], np.int32)
    Ag = gpu.CompRowMatrix.from_csr(3, 3, rp2, ci2, v2)
    M = gpu.AMG(min=1)
    with pytest.raises(RuntimeError, match="diagonal"):
        M.setMatrix(Ag)


# ---------------------------------------------------------------- FP32 cycle
def test_fp32_cycle_against_fp64(gpu, oracle):
    """The FP32 AMG builds the same hierarchy and runs the cycle (SSOR
    smoothers, residual, transfers) in FP32 with the coarsest solve in FP64:
    one cycle agrees with the FP64 cycle to FP32 round-off."""
    rp, ci, v = _sys(oracle, "s3")
    n = len(rp) - 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    M64 = gpu.AMG(zero_guess=1)
    M64.setMatrix(A)
    M32 = gpu.AMG(zero_guess=1, precision=gpu.L.FP32)
    M32.setMatrix(A)
    assert M32.levels() == M64.levels()
    b = np.random.default_rng(3).standard_normal(n)
    x64 = np.zeros(n)
    x32 = np.zeros(n)
    M64.apply(b, x64)
    M32.apply(b, x32)
    err = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    print(f"FP32 cycle vs FP64: {err:.2e}", flush=True)
    assert 1e-9 < err < 1e-5


def test_mixed_gmres_with_fp32_amg(gpu, oracle):
    """FP64 refinement around FP32 GMRES(30) preconditioned by the FP32 AMG
    reaches a true relative residual of 1e-10; the solution agrees with the
    oracle's FP64 GMRES + AMG solve."""
    rp, ci, v = _sys(oracle, "s3")
    n = len(rp) - 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    A.sync_f32()
    b = np.random.default_rng(5).standard_normal(n)
    M = gpu.AMG(zero_guess=1, precision=gpu.L.FP32)
    M.setMatrix(A)
    S = gpu.MixedPrecisionSolver(b, inner="gmres", restart=30, inner_rtol=1e-3,
                                 inner_max_iter=300)
    S.setIterationMonitor(gpu.DefaultIterationMonitor(100, 1e-10, 1e-50, 1e5))
    S.setPreconditioner(M)
    x = np.zeros(n)
    S.solve(A, b, x)
    As = scipy_csr(rp, ci, v, n, n)
    assert np.linalg.norm(b - As @ x) / np.linalg.norm(b) <= 1e-10
    xo, so, _ = oracle.solve("gmres", rp, ci, v, b, rtol=1e-12, prec="amg",
                             amg={"zero_guess": 1, "block_rows": 0, "ordering": 1, "mode": 1})
    assert so.status == 0
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_coarsen_ratio_stops_early(gpu, oracle):
    """coarsen_ratio: the hierarchy is the full one cut at the first level
    whose aggregates keep more than that fraction of its rows (levels above
    it bit-identical), and the cut level becomes the coarsest."""
    rp, ci, v = _sys(oracle, "c2m")     # level ratios 0.19, 0.13, 0.17, 0.32
    n = len(rp) - 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    full = gpu.AMG(zero_guess=1)
    full.setMatrix(A)
    lv = full.levels()
    ratios = [lv[k + 1][0] / lv[k][0] for k in range(len(lv) - 1)]
    # the first level past the top one whose ratio exceeds every earlier ratio
    k_stop = next(k for k in range(1, len(ratios)) if ratios[k] > max(ratios[:k]))
    cut = 0.5 * (max(ratios[:k_stop]) + ratios[k_stop])
    M = gpu.AMG(zero_guess=1, coarsen_ratio=cut, max_coarse=8192)
    M.setMatrix(A)
    assert M.levels() == lv[:k_stop + 1]
    for k in range(k_stop):
        a, b_ = full.level(k), M.level(k)
        for u, w in zip(a, b_):
            assert np.array_equal(u, w)


def test_coarsest_inverse_solve(gpu):
    """A coarsest level above 256 rows is solved with its explicit inverse:
    with the stop at level 0 the whole preconditioner is that solve, equal
    to a direct solve to round-off."""
    A_ = aniso2d(24)
    n = A_.shape[0]
    rp, ci, v = A_.indptr.astype(np.int32), A_.indices.astype(np.int32), A_.data.copy()
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, v)
    M = gpu.AMG(zero_guess=1, coarsen_ratio=1e-9, max_coarse=1024)
    M.setMatrix(A)
    assert len(M.levels()) == 1 and n > 256
    b = np.random.default_rng(8).standard_normal(n)
    x = np.zeros(n)
    M.apply(b, x)
    np.testing.assert_allclose(x, spl.spsolve(A_.tocsc(), b), rtol=1e-10, atol=1e-12)
    xt = np.zeros(n)
    M.transApply(b, xt)
    np.testing.assert_allclose(xt, spl.spsolve(A_.T.tocsc(), b), rtol=1e-10, atol=1e-12)
