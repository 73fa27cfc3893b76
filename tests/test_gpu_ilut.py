"""ILUT on the device (csrc/ilut.hip) against the oracle (oracle/ilut.c,
ILUT.java:136-272): the factor's pattern and values bit-exact, apply and
transApply bit-exact against the oracle's sweeps on the oracle's factor,
the reference's error cases, and the reference's solver test
(BiCGstabILUTTest: a preconditioned solve recovers a known solution)."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import oracle_system
from ressim_amd import problems as P

pytestmark = pytest.mark.gpu


def _lap2d(m, seed, shift=0.0):
    rng = np.random.default_rng(seed)
    T = sp.diags([-1, 4 + shift, -1], [-1, 0, 1], shape=(m, m))
    A = sp.kron(sp.eye(m), T) + sp.kron(sp.diags([-1, -1], [-1, 1], shape=(m, m)), sp.eye(m))
    A = A.tocsr()
    A.data = A.data * (1 + 0.3 * rng.standard_normal(A.nnz))
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()


def _random(n, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.nnz)
    A = (A + sp.diags(rng.standard_normal(n) * 0.3 + 3.0)).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()


def _lap3d(m, seed):
    e = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(m, m))
    I = sp.eye(m)
    A = (sp.kron(sp.kron(I, I), e) + sp.kron(sp.kron(I, e), I) + sp.kron(sp.kron(e, I), I))
    A = A.tocsr()
    A.data = A.data * (1 + 0.2 * np.random.default_rng(seed).standard_normal(A.nnz))
    A = (A + sp.eye(A.shape[0]) * 0.1).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()


def _check_factor(gpu, oracle, rp, ci, d, tau, p):
    n = len(rp) - 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, d)
    M = gpu.ILUT(A, tau, p)
    M.setMatrix(A)
    st, orp, oci, ova, odg, _ = oracle.ilut_factor(rp, ci, d, tau, p)
    assert st == 0
    grp, gci, glu = M.factors()
    np.testing.assert_array_equal(grp, orp)
    np.testing.assert_array_equal(gci, oci)
    assert np.array_equal(glu.view(np.int64), ova.view(np.int64))
    return A, M, (orp, oci, ova, odg)


def _check_apply(M, oracle, fac, seed=5):
    orp, oci, ova, odg = fac
    n = len(orp) - 1
    b = np.random.default_rng(seed).standard_normal(n)
    x = np.empty(n)
    M.apply(b, x)
    np.testing.assert_array_equal(x, oracle.ilu_apply(orp, oci, ova, odg, b))
    xt = np.empty(n)
    M.transApply(b, xt)
    np.testing.assert_array_equal(xt, oracle.ilu_transapply(orp, oci, ova, odg, b))


@pytest.mark.parametrize("tau,p", [(1e-6, 25), (0.0, 0), (1e-2, 2), (0.1, 1), (0.0, 400)])
def test_factor_and_sweeps_lap2d(gpu, oracle, tau, p):
    rp, ci, d = _lap2d(20, 1)
    print(f"ILUT lap2d 20x20 tau={tau} p={p}", flush=True)
    A, M, fac = _check_factor(gpu, oracle, rp, ci, d, tau, p)
    _check_apply(M, oracle, fac)


@pytest.mark.parametrize("n,dens,tau,p", [(200, 0.03, 1e-6, 25), (300, 0.02, 1e-3, 3),
                                          (150, 0.05, 0.0, 150), (400, 0.01, 0.05, 0)])
def test_factor_random_nonsymmetric(gpu, oracle, n, dens, tau, p):
    rp, ci, d = _random(n, dens, n)
    print(f"ILUT random n={n} tau={tau} p={p}", flush=True)
    A, M, fac = _check_factor(gpu, oracle, rp, ci, d, tau, p)
    _check_apply(M, oracle, fac, seed=n)


def test_factor_lap3d_and_refactor(gpu, oracle):
    """A 3-D 7-point matrix (12^3); then new values on the same pattern,
    the factor pattern changes, and the device sweeps follow it."""
    rp, ci, d = _lap3d(12, 2)
    n = len(rp) - 1
    print(f"ILUT lap3d n={n}", flush=True)
    A, M, fac = _check_factor(gpu, oracle, rp, ci, d, 1e-4, 10)
    _check_apply(M, oracle, fac)
    d2 = d * (1 + 0.5 * np.random.default_rng(9).standard_normal(d.size))
    A2 = gpu.CompRowMatrix.from_csr(n, n, rp, ci, d2)
    M.setMatrix(A2)
    st, orp, oci, ova, odg, _ = oracle.ilut_factor(rp, ci, d2, 1e-4, 10)
    assert st == 0
    grp, gci, glu = M.factors()
    np.testing.assert_array_equal(grp, orp)
    np.testing.assert_array_equal(gci, oci)
    assert np.array_equal(glu.view(np.int64), ova.view(np.int64))
    _check_apply(M, oracle, (orp, oci, ova, odg), seed=3)


def test_c1_mpfa_with_locked_rows(gpu, oracle):
    """The C1 MPFA-O pressure matrix (64 x 64) whose locked rows hold
    explicit zeros (Matrices.zeroRows): the factorisation starts from the
    nonzero entries."""
    s = oracle_system(oracle, P.config1())
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    assert np.count_nonzero(d == 0) > 0
    print(f"ILUT C1 n={len(rp) - 1}", flush=True)
    A, M, fac = _check_factor(gpu, oracle, rp, ci, d, 1e-6, 25)
    _check_apply(M, oracle, fac)


def test_errors(gpu, oracle):
    # missing diagonal (zero value): row 2, reported before any elimination
    rp = np.array([0, 2, 5, 7], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    d = np.array([2.0, 1, 1, 0.0, 1, 1, 2])
    A = gpu.CompRowMatrix.from_csr(3, 3, rp, ci, d)
    M = gpu.ILUT(A)
    with pytest.raises(RuntimeError, match="Missing diagonal entry on row 2"):
        M.setMatrix(A)
    assert M.bad_row == 1
    # U(1,1) = 0 after elimination: "Zero diagonal entry on row 2"
    d = np.array([1.0, 1, 1, 1, 1, 1, 1])
    A = gpu.CompRowMatrix.from_csr(3, 3, rp, ci, d)
    M = gpu.ILUT(A, 0.0, 5)
    with pytest.raises(RuntimeError, match="Zero diagonal entry on row 2 during ILU process"):
        M.setMatrix(A)
    assert oracle.ilut_factor(rp, ci, d, 0.0, 5)[0] == 10
    # the last row's diagonal is not tested
    A = gpu.CompRowMatrix.from_csr(2, 2, np.array([0, 2, 4], np.int32),
                                   np.array([0, 1, 0, 1], np.int32), np.array([1.0, 1, 1, 1]))
    M = gpu.ILUT(A, 0.0, 5)
    M.setMatrix(A)
    _, _, lu = M.factors()
    assert lu[3] == 0.0
    with pytest.raises(ValueError):
        gpu.ILUT(gpu.CompRowMatrix.from_csr(2, 3, np.array([0, 1, 2], np.int32),
                                            np.array([0, 1], np.int32), np.ones(2)))


@pytest.mark.parametrize("solver", ["BiCGstab", "GMRES"])
def test_solver_with_ilut_recovers_solution(gpu, solver):
    """BiCGstabILUTTest / IterativeSolverTest.assertSolved: b = A x, a
    preconditioned solve from a random start recovers x."""
    rp, ci, d = _random(500, 0.01, 21)
    n = len(rp) - 1
    rng = np.random.default_rng(4)
    xd = rng.standard_normal(n)
    A = gpu.CompRowMatrix.from_csr(n, n, rp, ci, d)
    b = sp.csr_matrix((d, ci, rp), shape=(n, n)) @ xd
    S = getattr(gpu, solver)(b)
    S.setIterationMonitor(gpu.DefaultIterationMonitor(2000, 1e-12, 1e-50, 1e5))
    M = gpu.ILUT(A)
    M.setMatrix(A)
    S.setPreconditioner(M)
    x = rng.standard_normal(n)
    S.solve(A, b, x)
    np.testing.assert_allclose(x, xd, rtol=0, atol=1e-8)
    assert S.getIterationMonitor().iterations() < 60


def test_larger_default_parameters(gpu, oracle):
    """MTJ's default ILUT(1e-6, 25) on a 2-D 128 x 128 matrix (16,384 rows,
    fill across the 128-wide band): bit-exact factors and sweeps."""
    rp, ci, d = _lap2d(128, 7)
    print(f"ILUT lap2d 128x128 n={len(rp) - 1}", flush=True)
    A, M, fac = _check_factor(gpu, oracle, rp, ci, d, 1e-6, 25)
    _check_apply(M, oracle, fac)


def test_level_scheduled_sweeps(gpu, oracle, monkeypatch):
    """Past the LDS capacity of the sequential sweeps the factor is applied
    by the level-scheduled ILU sweeps (forced here): bit-exact too."""
    monkeypatch.setenv("RSGPU_ILUT_LEVELS", "1")
    rp, ci, d = _lap2d(24, 3)
    A, M, fac = _check_factor(gpu, oracle, rp, ci, d, 1e-4, 8)
    _check_apply(M, oracle, fac, seed=8)
