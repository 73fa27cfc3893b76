"""Pins of the oracle's ILUT (oracle/ilut.c, ILUT.java:136-272).

* Brute force on tiny inputs: MTJ's algorithm written out in Python over
  dict rows and Python's stable sort (no shared code with the C oracle),
  bit-exact, over random sparse matrices, thresholds and fill counts.
* Closed forms: tau = 0 with p >= n drops nothing, so the factor is the
  LU decomposition without pivoting (scipy's LU of a column-dominant
  matrix, where partial pivoting swaps nothing); a huge tau drops every
  elimination and keeps only the diagonal (the apply is b / diag(A)); a
  tridiagonal matrix has no fill, so ILUT with any small tau is exact.
* The reference's own test (BiCGstabILUTTest / IterativeSolverTest: a
  preconditioned solve recovers a known solution) as a property of the
  factor: with tau = 0, p >= n, one apply solves A x = b.
"""
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

from oracle import pyoracle as O


def _mtj_ilut(n, rowptr, colind, val, tau, p):
    """ILUT.factor over rows {col: value} (structure = keys).  Returns the
    rows (sorted dicts) or ("error", code, row)."""
    rows = []
    for r in range(n):
        d = {}
        for j in range(rowptr[r], rowptr[r + 1]):
            if val[j] != 0.0:
                d[int(colind[j])] = float(val[j])
        rows.append(d)
    for k in range(n):
        if k not in rows[k]:
            return ("error", 11, k + 1)
    for i in range(1, n):
        rowi = rows[i]
        cols = sorted(rowi)
        nrm = 0.0
        for c in cols:
            nrm += rowi[c] * rowi[c]
        taui = math.sqrt(nrm) * tau
        z = [0.0] * n
        for c in cols:
            z[c] = rowi[c]
        for k in range(i):
            rowk = rows[k]
            dk = rowk[k]
            if dk == 0:
                return ("error", 10, k + 1)
            luik = z[k] / dk
            if abs(luik) <= taui:
                continue
            for c in sorted(rowk):
                if c > k:
                    z[c] -= luik * rowk[c]
            z[k] = luik
        nl = sum(1 for c in cols if c < i)
        nu = sum(1 for c in cols if c > i)
        for c in cols:
            rowi[c] = 0.0
        lower = [(c, z[c]) for c in range(i) if abs(z[c]) > taui]
        upper = [(c, z[c]) for c in range(i + 1, n) if abs(z[c]) > taui]
        lower = sorted(lower, key=lambda e: -abs(e[1]))
        upper = sorted(upper, key=lambda e: -abs(e[1]))
        rowi[i] = z[i]
        for c, v in lower[:max(0, min(nl + p, len(lower)))]:
            rowi[c] = v
        for c, v in upper[:max(0, min(nu + p, len(upper)))]:
            rowi[c] = v
    return rows


def _to_csr(rows):
    rp, ci, va = [0], [], []
    for d in rows:
        for c in sorted(d):
            ci.append(c)
            va.append(d[c])
        rp.append(len(ci))
    return np.array(rp), np.array(ci), np.array(va)


def _random(n, density, seed, zeros=0, dominant=False):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="lil")
    A = A.tocsr()
    A.data = rng.standard_normal(A.nnz)
    A = A + sp.diags(rng.standard_normal(n) * 0.3 + (3.0 + (n * density if dominant else 0)))
    A = A.tocsr()
    A.sort_indices()
    if zeros:
        idx = rng.choice(A.nnz, zeros, replace=False)
        for q in idx:
            r = np.searchsorted(A.indptr, q, side="right") - 1
            if A.indices[q] != r:
                A.data[q] = 0.0
    return A


CASES = [(n, dens, tau, p, seed) for seed, (n, dens) in enumerate([(5, 0.5), (9, 0.3), (17, 0.2),
                                                                   (30, 0.12)])
         for tau, p in [(1e-6, 25), (0.0, 0), (0.05, 1), (0.3, 0), (0.0, 100), (1e-3, 2)]]


@pytest.mark.parametrize("n,dens,tau,p,seed", CASES)
def test_bitexact_vs_brute_force(n, dens, tau, p, seed):
    A = _random(n, dens, seed, zeros=2)
    got = O.ilut_factor(A.indptr, A.indices, A.data, tau, p)
    ref = _mtj_ilut(n, A.indptr, A.indices, A.data, tau, p)
    assert got[0] == 0 and isinstance(ref, list)
    rp, ci, va = _to_csr(ref)
    np.testing.assert_array_equal(got[1], rp)
    np.testing.assert_array_equal(got[2], ci)
    assert np.array_equal(got[3].view(np.int64), va.view(np.int64))
    for r in range(n):
        assert got[2][got[4][r]] == r


def test_fill_is_dropped_and_kept_by_size():
    """A 2-D 5-point Laplacian (8 x 8) produces fill; with p = 0 no row keeps
    more nonzero entries on a side than A has there, with p = 3 some row
    keeps more, and a larger p never keeps fewer."""
    m = 8
    T = sp.diags([-1, 4, -1], [-1, 0, 1], shape=(m, m))
    A = (sp.kron(sp.eye(m), T) + sp.kron(sp.diags([-1, -1], [-1, 1], shape=(m, m)), sp.eye(m)))
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    counts = {}
    for p in (0, 3, 40):
        st, rp, ci, va, dg, _ = O.ilut_factor(A.indptr, A.indices, A.data, 0.0, p)
        assert st == 0
        nzl = [int(np.count_nonzero(va[rp[r]:dg[r]])) for r in range(n)]
        counts[p] = nzl
    al = [int(np.count_nonzero(A.indices[A.indptr[r]:A.indptr[r + 1]] < r)) for r in range(n)]
    assert all(c <= a for c, a in zip(counts[0], al))
    assert any(c > a for c, a in zip(counts[3], al))
    assert all(c3 <= c40 for c3, c40 in zip(counts[3], counts[40]))


@pytest.mark.parametrize("n", [12, 40])
def test_no_drop_is_lu(n):
    A = _random(n, 0.25, 100 + n, dominant=True).toarray()
    # column dominance: partial pivoting keeps the natural order
    A += np.diag(np.abs(A).sum(axis=0) + 1)
    S = sp.csr_matrix(A)
    S.sort_indices()
    st, rp, ci, va, dg, _ = O.ilut_factor(S.indptr, S.indices, S.data, 0.0, n)
    assert st == 0
    F = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    P, Ls, Us = scipy.linalg.lu(A)
    assert np.array_equal(P, np.eye(n))
    np.testing.assert_allclose(L, Ls, rtol=0, atol=1e-12)
    np.testing.assert_allclose(U, Us, rtol=1e-12, atol=1e-12 * np.abs(Us).max())
    # the stored pattern is the nonzero pattern of the full factors
    assert set(zip(*np.nonzero(F))) == set(zip(*np.nonzero(np.abs(L - np.eye(n)) + U)))
    # one apply solves the system
    x = np.random.default_rng(n).standard_normal(n)
    b = A @ x
    y = np.zeros(n)
    z = np.zeros(n)
    O.lib().or_ilu_apply(n, O._i32(rp), O._i32(ci), va, O._i32(dg), b, z, y)
    np.testing.assert_allclose(z, x, rtol=1e-10, atol=1e-10)


def test_huge_tau_keeps_diagonal_only():
    A = _random(25, 0.2, 7)
    st, rp, ci, va, dg, _ = O.ilut_factor(A.indptr, A.indices, A.data, 1e300, 5)
    assert st == 0
    n = 25
    d = A.diagonal()
    # row 0 is never gathered (the loop starts at row 1): it keeps A's row
    np.testing.assert_array_equal(va[rp[0]:rp[1]], A.data[A.indptr[0]:A.indptr[1]])
    for r in range(1, n):
        row = va[rp[r]:rp[r + 1]]
        assert row[dg[r] - rp[r]] == d[r]
        assert np.count_nonzero(row) == 1
        # the structure keeps A's nonzero entries, as zeros
        assert set(ci[rp[r]:rp[r + 1]]) == set(A.indices[A.indptr[r]:A.indptr[r + 1]][
            A.data[A.indptr[r]:A.indptr[r + 1]] != 0])
    b = np.random.default_rng(1).standard_normal(n)
    x = np.zeros(n)
    y = np.zeros(n)
    O.lib().or_ilu_apply(n, O._i32(rp), O._i32(ci), va, O._i32(dg), b, x, y)
    # rows >= 1: L = I; U = diag except row 0, so x = b / d below row 0
    # wherever row 0's couplings do not reach (the backward sweep)
    assert np.array_equal(x[1:], b[1:] / d[1:])


def test_tridiagonal_is_exact():
    n = 200
    rng = np.random.default_rng(3)
    lo, di, up = rng.standard_normal(n - 1), rng.standard_normal(n) + 4, rng.standard_normal(n - 1)
    A = sp.diags([lo, di, up], [-1, 0, 1]).tocsr()
    A.sort_indices()
    st, rp, ci, va, dg, _ = O.ilut_factor(A.indptr, A.indices, A.data, 1e-12, 0)
    assert st == 0
    np.testing.assert_array_equal(rp, A.indptr)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    y = np.zeros(n)
    O.lib().or_ilu_apply(n, O._i32(rp), O._i32(ci), va, O._i32(dg), b, x, y)
    ab = np.zeros((3, n))
    ab[0, 1:], ab[1], ab[2, :-1] = up, di, lo
    np.testing.assert_allclose(x, scipy.linalg.solve_banded((1, 1), ab, b), rtol=1e-12,
                               atol=1e-12)


def test_errors():
    # missing diagonal (a zero diagonal value counts as missing: the
    # starting structure holds the nonzero entries)
    A = sp.csr_matrix(np.array([[2.0, 1, 0], [1, 0.0, 1], [0, 1, 2]]))
    A.sort_indices()
    st, *_, bad = O.ilut_factor(A.indptr, A.indices, A.data)
    assert (st, bad) == (11, 2)
    # zero diagonal after elimination on row 2 (1-based), reported while row
    # 3 is factored
    A = sp.csr_matrix(np.array([[1.0, 1, 0], [1, 1, 1], [0, 1, 1]]))
    A.sort_indices()
    st, *_, bad = O.ilut_factor(A.indptr, A.indices, A.data, 0.0, 5)
    assert (st, bad) == (10, 2)
    # the last row's diagonal is never tested
    A = sp.csr_matrix(np.array([[1.0, 1], [1, 1]]))
    A.sort_indices()
    st, rp, ci, va, dg, _ = O.ilut_factor(A.indptr, A.indices, A.data, 0.0, 5)
    assert st == 0 and va[dg[1]] == 0.0
