"""Pins of the CPU oracle's MTJ restatement (CSR, BLAS-1, ILU(0), Krylov).

Independent checks: scipy/numpy dense algebra, the reference's own ILU
known answers (IncompleteFactorizationTest.java:40-76 / ILUTest.java:34-44),
the defining property of ILU(0) (L U equals A on the pattern), brute-force
dense solves for the Krylov methods (IterativeSolverTest.java:93-176
protocol), and the failure reasons of DefaultIterationMonitor and BiCGstab.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401

from ressim_amd import problems as P
from tests.helpers import oracle_system, pentadiag, random_csr, scipy_csr, tridiag


def test_construct_sorts_and_checks(oracle):
    st, rp, ci = oracle.csr_construct(3, 5, [[4, 0, 2], [], [3, 1]])
    assert st == 0
    assert rp.tolist() == [0, 3, 3, 5] and ci.tolist() == [0, 2, 4, 1, 3]
    st, _, _ = oracle.csr_construct(2, 3, [[0, 3], [1]])
    assert st == oracle.BAD_ARGUMENT


def test_add_not_in_pattern_and_zero_rows(oracle):
    rng = np.random.default_rng(1)
    rp, ci, d = random_csr(rng, 30, 30, 0.2, diag=5.0)
    A = scipy_csr(rp, ci, d).toarray()
    d2 = d.copy()
    j = int(ci[rp[3]])
    assert oracle.csr_add(rp, ci, d2, 3, j, 2.5) == 0
    assert d2[rp[3]] == d[rp[3]] + 2.5
    missing = [c for c in range(30) if c not in ci[rp[4]:rp[5]]][0]
    assert oracle.csr_add(rp, ci, d2, 4, missing, 1.0) == oracle.NOT_IN_PATTERN
    rows = [7, 2, 7]
    d3 = d.copy()
    assert oracle.zero_rows(rp, ci, d3, 3.0, rows) == 0
    B = A.copy()
    for r in set(rows):
        B[r, :] = 0
        B[r, r] = 3.0
    np.testing.assert_array_equal(scipy_csr(rp, ci, d3).toarray(), B)


def test_spmv_family_matches_scipy(oracle):
    rng = np.random.default_rng(2)
    rp, ci, d = random_csr(rng, 60, 45, 0.15, empty_rows={5, 6})
    A = scipy_csr(rp, ci, d, 60, 45)
    x = rng.standard_normal(45)
    xt = rng.standard_normal(60)
    y0 = rng.standard_normal(60)
    np.testing.assert_allclose(oracle.mult(rp, ci, d, x), A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.multadd(rp, ci, d, -0.7, x, y0), y0 - 0.7 * (A @ x),
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.transmult(rp, ci, d, xt, 45), A.T @ xt, rtol=1e-13,
                               atol=1e-13)
    y1 = rng.standard_normal(45)
    np.testing.assert_allclose(oracle.transmultadd(rp, ci, d, 1.7, xt, y1), y1 + 1.7 * (A.T @ xt),
                               rtol=1e-12, atol=1e-12)
    # mult's zero rule: a row that sums to exactly zero gives +0.0
    y = oracle.mult(np.array([0, 2], np.int32), np.array([0, 1], np.int32),
                    np.array([1.0, -1.0]), np.array([-0.0, -0.0]))
    assert y[0] == 0.0 and not np.signbit(y[0])


def test_blas1_exact_on_integers(oracle):
    x = np.arange(1, 101, dtype=np.float64)
    assert oracle.dot(x, x) == float(np.sum(np.arange(1, 101) ** 2))
    assert oracle.norm2(np.array([3.0, 4.0])) == 5.0


@pytest.mark.parametrize("n", range(1, 11))
def test_ilu_known_answers(oracle, n):
    """ILUTest: tri- and penta-diagonal, x = 1, b = A x, ILU^-1 b == x."""
    for rp, ci, d in (tridiag(n), pentadiag(n)):
        A = scipy_csr(rp, ci, d)
        x = np.ones(n)
        b = A @ x
        st, lu, dg, _ = oracle.ilu_factor(rp, ci, d)
        assert st == 0
        xs = oracle.ilu_apply(rp, ci, lu, dg, b)
        assert np.linalg.norm(b - A @ xs) < 1e-12


def _dense_ilu0(A, mask):
    """Textbook ILU(0) by masked Gaussian elimination (IKJ)."""
    n = A.shape[0]
    LU = A.copy()
    for i in range(1, n):
        for k in range(i):
            if not mask[i, k]:
                continue
            LU[i, k] /= LU[k, k]
            for j in range(k + 1, n):
                if mask[i, j]:
                    LU[i, j] -= LU[i, k] * LU[k, j]
    return LU


def test_ilu_matches_masked_elimination(oracle):
    s = oracle_system(oracle, P.config1(nx=10, ny=9))
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    A = scipy_csr(rp, ci, d).toarray()
    mask = scipy_csr(rp, ci, np.ones_like(d)).toarray() != 0
    st, lu, dg, _ = oracle.ilu_factor(rp, ci, d)
    assert st == 0
    LU = scipy_csr(rp, ci, lu).toarray()
    np.testing.assert_allclose(LU[mask], _dense_ilu0(A, mask)[mask], rtol=1e-12, atol=1e-13)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    b = np.random.default_rng(0).standard_normal(n)
    np.testing.assert_allclose(oracle.ilu_apply(rp, ci, lu, dg, b),
                               np.linalg.solve(L @ U, b), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(oracle.ilu_transapply(rp, ci, lu, dg, b),
                               np.linalg.solve((L @ U).T, b), rtol=1e-10, atol=1e-12)


def test_ilu_errors(oracle):
    rp, ci, d = tridiag(5)
    d0 = d.copy()
    d0[rp[0]] = 0.0                      # zero pivot on row 1
    st, _, _, bad = oracle.ilu_factor(rp, ci, d0)
    assert st == oracle.ZERO_PIVOT
    # row 2 without its diagonal
    rows = [[0, 1], [0], [1, 2]]
    st, rp2, ci2 = oracle.csr_construct(3, 3, rows)
    st, _, _, bad = oracle.ilu_factor(rp2, ci2, np.ones(len(ci2)))
    assert st == oracle.MISSING_DIAGONAL and bad == 1


def test_block_filter(oracle):
    s = oracle_system(oracle, P.config1(nx=8, ny=8))
    rp, ci, d = s["rowptr"], s["colind"], s["data"]
    n = len(rp) - 1
    brp, bci, bd, src = oracle.block_filter(rp, ci, d, 4)
    bs = oracle.block_starts(n, 4)
    A = scipy_csr(rp, ci, d).toarray()
    M = np.zeros_like(A)
    for b in range(4):
        M[bs[b]:bs[b + 1], bs[b]:bs[b + 1]] = A[bs[b]:bs[b + 1], bs[b]:bs[b + 1]]
    np.testing.assert_array_equal(scipy_csr(brp, bci, bd).toarray(), M)
    np.testing.assert_array_equal(d[src], bd)


@pytest.mark.parametrize("kind,prec", [("bicgstab", "ilu"), ("gmres", "ilu"),
                                       ("gmres", "block_ilu"), ("bicgstab", "none")])
def test_krylov_solves_c1(oracle, kind, prec):
    pr = P.config1(nx=24, ny=20)
    s = oracle_system(oracle, pr)
    rp, ci, d, b = s["rowptr"], s["colind"], s["data"], s["b"]
    A = scipy_csr(rp, ci, d)
    xs = sp.linalg.spsolve(A.tocsc(), b)
    x, st, hist = oracle.solve(kind, rp, ci, d, b, rtol=1e-12, prec=prec, nblocks=4,
                               max_iter=20000)
    assert st.status == 0
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    assert len(hist) == st.hist_len


def test_iterative_solver_protocol_random(oracle):
    """IterativeSolverTest: random banded matrix, shifted diagonal, x0 random."""
    rng = np.random.default_rng(5)
    n = 50
    rp, ci, d = random_csr(rng, n, n, 0.1, diag=0.0)
    A = scipy_csr(rp, ci, d).toarray() + 0.0
    A += 100 * np.eye(n)
    Asp = sp.csr_matrix(A)
    Asp.sort_indices()
    xd = rng.standard_normal(n)
    b = A @ xd
    for kind in ("bicgstab", "gmres", "ir"):
        x, st, _ = oracle.solve(kind, Asp.indptr, Asp.indices, Asp.data, b,
                                x0=rng.standard_normal(n), rtol=1e-12, prec="ilu")
        assert st.status == 0
        np.testing.assert_allclose(x, xd, atol=1e-4 * 1e-3)


def test_gmres_counts_and_finite_termination(oracle):
    """Unpreconditioned GMRES(m >= n) terminates within n inner steps (plus
    the cycle's own count), as in exact arithmetic."""
    rng = np.random.default_rng(7)
    n = 12
    A = rng.standard_normal((n, n)) + 5 * np.eye(n)
    Asp = sp.csr_matrix(A)
    b = rng.standard_normal(n)
    x, st, hist = oracle.solve("gmres", Asp.indptr, Asp.indices, Asp.data, b, rtol=1e-10,
                               prec="none", restart=20)
    assert st.status == 0 and st.iterations <= n + 2
    np.testing.assert_allclose(A @ x, b, atol=1e-8)


def test_bicgstab_initial_residual_quirk(oracle):
    """BiCGstab's first converged(s) call happens at iter 0, so the monitor
    re-bases initR to ||s0|| (DefaultIterationMonitor.java:136-137)."""
    pr = P.config1(nx=16, ny=16)
    s = oracle_system(oracle, pr)
    x, st, hist = oracle.solve("bicgstab", s["rowptr"], s["colind"], s["data"], s["b"],
                               rtol=1e-6, prec="ilu")
    assert st.status == 0
    assert hist[0] == pytest.approx(np.linalg.norm(s["b"]))
    assert hist[-1] < 1e-6 * hist[1]
    assert not hist[-1] < 1e-6 * hist[0] or hist[1] <= hist[0]


def test_monitor_failure_reasons(oracle):
    pr = P.config1(nx=12, ny=12)
    s = oracle_system(oracle, pr)
    args = (s["rowptr"], s["colind"], s["data"])
    _, st, _ = oracle.solve("bicgstab", *args, s["b"], rtol=1e-12, max_iter=1)
    assert st.status == oracle.ITERATIONS
    _, st, _ = oracle.solve("gmres", *args, s["b"], rtol=1e-12, dtol=1e-30)
    assert st.status == oracle.DIVERGENCE
    bn = s["b"].copy()
    bn[3] = np.nan
    _, st, _ = oracle.solve("bicgstab", *args, bn, rtol=1e-12)
    assert st.status == oracle.DIVERGENCE
    # Breakdown in exact (integer) arithmetic: A = [[-2,0],[-1,1]], b = (1,1):
    # r0 = (1,1), v = A r0 = (-2,0), alpha = 2/-2 = -1, s = r0 + v... = (-1,1),
    # t = A s = (2,2), omega = t.s/t.t = 0, r1 = s, rho = r0.r1 = 0.
    A = sp.csr_matrix(np.array([[-2.0, 0.0], [-1.0, 1.0]]))
    A.sort_indices()
    _, st, hist = oracle.solve("bicgstab", A.indptr, A.indices, A.data, np.array([1.0, 1.0]),
                               rtol=1e-12, prec="none")
    assert st.status == oracle.BREAKDOWN and st.iterations == 1
    np.testing.assert_allclose(hist, [np.sqrt(2)] * 3)


# ---------------------------------------------------------------- CG, diagonal, SSOR
def _spd(n1=20, eps=0.3):
    """Anisotropic 2-D Laplacian with a badly scaled diagonal (D A D)."""
    I = sp.identity(n1)
    T = sp.diags([-np.ones(n1 - 1), 2 * np.ones(n1), -np.ones(n1 - 1)], [-1, 0, 1])
    D = sp.diags(np.exp(np.random.default_rng(7).standard_normal(n1 * n1)))
    A = (D @ (sp.kron(I, T) + eps * sp.kron(T, I)) @ D).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()


@pytest.mark.parametrize("prec", ["none", "diagonal", "ilu"])
def test_cg_solves_spd(oracle, prec):
    """CG.java on an SPD matrix: solution equals spsolve; preconditioned CG
    takes fewer iterations than plain CG."""
    rp, ci, v = _spd()
    n = len(rp) - 1
    b = np.random.default_rng(3).standard_normal(n)
    x, st, h = oracle.solve("cg", rp, ci, v, b, rtol=1e-12, prec=prec, max_iter=5000)
    assert st.status == 0
    xe = sp.linalg.spsolve(sp.csr_matrix((v, ci, rp)).tocsc(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-9
    if prec != "none":
        _, st0, _ = oracle.solve("cg", rp, ci, v, b, rtol=1e-12, prec="none", max_iter=5000)
        assert st.iterations < st0.iterations


def test_cg_finite_termination(oracle):
    """A diagonal SPD matrix with k distinct eigenvalues: CG converges in
    k iterations."""
    d = np.repeat([1.0, 2.0, 5.0], 10)
    A = sp.diags(d).tocsr()
    rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    x, st, _ = oracle.solve("cg", rp, ci, v, np.ones(30), rtol=1e-12, prec="none")
    assert st.status == 0 and st.iterations == 3
    np.testing.assert_allclose(x, 1 / d, rtol=1e-12)


def test_diagonal_preconditioner_zero_diagonal(oracle):
    A = sp.csr_matrix(np.array([[2.0, 1.0], [1.0, 0.0]]))
    rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    x, st, _ = oracle.solve("cg", rp, ci, v, np.ones(2), prec="diagonal")
    assert st.status == oracle.ZERO_PIVOT and st.bad_row == 1


@pytest.mark.parametrize("kind,prec", [(k, p) for k in ("bicg", "cgs", "qmr")
                                       for p in ("none", "ilu", "block_ilu", "diagonal")
                                       if (k, p) != ("cgs", "none")])
def test_bicg_cgs_qmr_solve_c1(oracle, kind, prec):
    """BiCG, CGS and QMR (with M.transApply in BiCG and QMR) reach the
    direct solution of a non-symmetric MPFA-O system (divergence tolerance
    raised: CGS's squared residual polynomial swings far above the initial
    residual; unpreconditioned CGS does not converge on this system at all,
    its residual history equals a direct numpy restatement of CGS.java)."""
    pr = P.config1(nx=24, ny=20)
    s = oracle_system(oracle, pr)
    rp, ci, d, b = s["rowptr"], s["colind"], s["data"], s["b"]
    A = scipy_csr(rp, ci, d)
    assert abs(A - A.T).max() > 0
    xs = sp.linalg.spsolve(A.tocsc(), b)
    x, st, hist = oracle.solve(kind, rp, ci, d, b, rtol=1e-12, prec=prec, nblocks=4,
                               max_iter=20000, dtol=1e30)
    assert st.status == 0, (kind, prec, st.status)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    # the monitor tracks the recurrence residual; the true one stays close
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b)
    assert len(hist) == st.hist_len == st.iterations + 1


def test_bicg_is_cg_on_spd(oracle):
    """On a symmetric matrix without preconditioning BiCG's shadow sequence
    equals its primary one (rtilde = r, ptilde = p, A'p = A p), so its
    residual history is CG's up to round-off (which this ill-conditioned
    matrix amplifies after a few dozen steps)."""
    rp, ci, v = _spd()
    b = np.random.default_rng(4).standard_normal(len(rp) - 1)
    _, s1, h1 = oracle.solve("bicg", rp, ci, v, b, rtol=1e-10, prec="none", max_iter=5000)
    _, s2, h2 = oracle.solve("cg", rp, ci, v, b, rtol=1e-10, prec="none", max_iter=5000)
    assert s1.status == s2.status == 0
    assert abs(s1.iterations - s2.iterations) <= 0.1 * s2.iterations
    np.testing.assert_allclose(h1[:20], h2[:20], rtol=1e-10)


def test_qmr_and_cgs_first_steps(oracle):
    """Closed forms of the first iteration without preconditioning: CGS's
    first residual is r1 = r0 - 2 a A r0 + a^2 A^2 r0 (a = r0.r0 / r0.A r0),
    QMR's first iterate is x1 = (r0.A r0 / |A r0|^2)-free BiCG step: both
    checked through the residual after one iteration."""
    rng = np.random.default_rng(9)
    n = 40
    rp, ci, d = random_csr(rng, n, n, 0.15, diag=0.0)
    A = scipy_csr(rp, ci, d).toarray() + 6 * np.eye(n)
    Asp = sp.csr_matrix(A)
    Asp.sort_indices()
    b = rng.standard_normal(n)
    r0 = b.copy()
    Ar = A @ r0
    a = (r0 @ r0) / (r0 @ Ar)
    r1_cgs = r0 - 2 * a * Ar + a * a * (A @ Ar)
    # BiCG (and QMR's Lanczos step with M = I) first residual: r0 - a A r0
    r1_bicg = r0 - a * Ar
    for kind, want in (("cgs", r1_cgs), ("bicg", r1_bicg)):
        _, st, h = oracle.solve(kind, Asp.indptr, Asp.indices, Asp.data, b, rtol=1e-300,
                                max_iter=1, prec="none")
        assert st.status == 3      # iterations exhausted after one step
        assert h[1] == pytest.approx(np.linalg.norm(want), rel=1e-12)
    # QMR after one step: x1 = eta p with p = r0/|r0|, minimising the
    # quasi-residual; its true residual is b - A x1 with x1 = (r0.r0 / r0.A r0)
    # r0 only when the quasi-minimal and Galerkin steps coincide, so check
    # the defining property instead: |r1| <= |r0| and r1 = b - A x1 exactly
    x1, st, h = oracle.solve("qmr", Asp.indptr, Asp.indices, Asp.data, b, rtol=1e-300,
                             max_iter=1, prec="none")
    assert h[1] == pytest.approx(np.linalg.norm(b - A @ x1), rel=1e-10)
    assert h[1] < h[0]


def _upper_dense(rp, ci, R, n):
    U = np.zeros((n, n))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            if ci[j] >= i:
                U[i, ci[j]] = R[j]
    return U


@pytest.mark.parametrize("case", ["tridiagonal", "dense", "arrow"])
def test_icc_known_answers(oracle, case):
    """ICC's factor equals the exact Cholesky factor where the pattern
    holds it: a tridiagonal, a full matrix and an arrow matrix with its
    dense row and column last (none of them fills in)."""
    rng = np.random.default_rng(3)
    n = 24
    if case == "tridiagonal":
        A = sp.diags([-np.ones(n - 1), 4 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).toarray()
    elif case == "dense":
        B = rng.standard_normal((n, n))
        A = B @ B.T + n * np.eye(n)
    else:
        A = np.eye(n) * 10.0
        A[-1, :] = A[:, -1] = rng.uniform(0.1, 1.0, n)
        A[-1, -1] = 50.0
    Asp = sp.csr_matrix(A)
    Asp.sort_indices()
    st, R, dg, bad = oracle.icc_factor(Asp.indptr, Asp.indices, Asp.data)
    assert st == 0
    L = np.linalg.cholesky(A)
    np.testing.assert_allclose(_upper_dense(Asp.indptr, Asp.indices, R, n), L.T, rtol=1e-13,
                               atol=1e-13)
    # entries left of the diagonal keep A's values
    for i in range(n):
        for j in range(Asp.indptr[i], Asp.indptr[i + 1]):
            if Asp.indices[j] < i:
                assert R[j] == Asp.data[j]
    b = rng.standard_normal(n)
    x = oracle.icc_apply(Asp.indptr, Asp.indices, R, dg, b)
    np.testing.assert_allclose(A @ x, b, rtol=1e-10, atol=1e-12)


def test_icc_zero_diagonal_and_cg(oracle):
    rp2, ci2 = np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32)
    st, _, _, bad = oracle.icc_factor(rp2, ci2, np.array([0.0, 1.0, 1.0, 2.0]))
    assert st == 10 and bad == 1          # "Zero diagonal entry ... on row 1"
    st, _, _, bad = oracle.icc_factor(np.array([0, 1, 2], np.int32),
                                      np.array([1, 1], np.int32), np.ones(2))
    assert st == 11 and bad == 1          # "Missing diagonal entry on row 1"
    rp, ci, v = _spd()
    b = np.random.default_rng(8).standard_normal(len(rp) - 1)
    x, s, _ = oracle.solve("cg", rp, ci, v, b, rtol=1e-12, prec="icc", max_iter=5000)
    _, s0, _ = oracle.solve("cg", rp, ci, v, b, rtol=1e-12, prec="none", max_iter=5000)
    assert s.status == 0 and s.iterations < s0.iterations / 3
    Asp = scipy_csr(rp, ci, v)
    assert np.linalg.norm(Asp @ x - b) < 1e-10 * np.linalg.norm(b)


def test_gmres_bjilu_omp_equals_sequential_oracle(oracle):
    """The all-cores CPU baseline (or_gmres_bjilu_omp): with one thread every
    operation is the sequential oracle's, so iterate, iteration count and
    residual are bit-identical to or_gmres + block ILU(0); with several
    threads only the dot products' summation order changes."""
    s = oracle_system(oracle, P.small3d(nx=14, ny=12, nz=9))
    args = (s["rowptr"], s["colind"], s["data"], s["b"])
    xo, so, _ = oracle.solve("gmres", *args, rtol=1e-11, prec="block_ilu", nblocks=7,
                             restart=20)
    assert so.status == 0
    x1, s1 = oracle.gmres_bjilu_omp(*args, nthreads=1, rtol=1e-11, nblocks=7, restart=20)
    assert s1.status == 0 and s1.iterations == so.iterations
    assert s1.residual == so.residual
    np.testing.assert_array_equal(x1, xo)
    x4, s4 = oracle.gmres_bjilu_omp(*args, nthreads=4, rtol=1e-11, nblocks=7, restart=20)
    assert s4.status == 0 and abs(s4.iterations - so.iterations) <= 2
    assert np.linalg.norm(x4 - xo) <= 1e-9 * np.linalg.norm(xo)
    # failure reasons follow DefaultIterationMonitor
    _, s5 = oracle.gmres_bjilu_omp(*args, nthreads=2, rtol=1e-11, nblocks=7, max_iter=3)
    assert s5.status == 3
