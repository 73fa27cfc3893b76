"""Shared builders for the parity tests (oracle side)."""
import numpy as np
import scipy.sparse as sp

from ressim_amd import problems as P


def oracle_system(po, pr: P.PressureProblem, with_diag=True):
    """MPFA-O transmissibilities, pattern, flux assembly and locked rows on
    the CPU oracle: returns dict(rowptr, colind, data, b, md_ptr, md_elem,
    md_t, here, there, fails)."""
    md_ptr, md_el, md_t, fails = po.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz,
                                           pr.K6)
    here, there = po.mesh_connections(pr.dim, pr.nx, pr.ny, pr.nz)
    rp, ci = po.pattern(pr.n, here, there, md_ptr, md_el, with_diag)
    A = po.assemble_flux(pr.n, here, there, md_ptr, md_el, md_t, rp, ci)
    flux_only = A.copy()
    if pr.accumulation:
        # CompRowMatrix.add(i, i, acc) on every row (or_csr_add): one IEEE
        # addition into each diagonal entry, vectorised (each pattern row
        # holds its diagonal once)
        rows = np.repeat(np.arange(pr.n, dtype=np.int32), np.diff(rp))
        dpos = np.flatnonzero(ci == rows)
        assert dpos.size == pr.n
        A[dpos] += pr.accumulation
        if pr.n <= 4096:
            chk = flux_only.copy()
            for i in range(pr.n):
                assert po.csr_add(rp, ci, chk, i, i, pr.accumulation) == 0
            assert np.array_equal(chk, A)
    A_unpinned = A.copy()
    st = po.zero_rows(rp, ci, A, 1.0, pr.pinned)
    assert st == 0
    return dict(rowptr=rp, colind=ci, data=A, data_unpinned=A_unpinned, flux=flux_only,
                b=pr.rhs(),
                md_ptr=md_ptr, md_elem=md_el, md_t=md_t, here=here, there=there,
                fails=fails)


def scipy_csr(rp, ci, data, n=None, m=None):
    n = len(rp) - 1 if n is None else n
    m = n if m is None else m
    return sp.csr_matrix((data, ci, rp), shape=(n, m))


def random_csr(rng, n, m, density, long_rows=(), empty_rows=(), diag=None):
    """Random CSR with optional very long and empty rows (sorted columns)."""
    rows = []
    for i in range(n):
        if i in empty_rows:
            rows.append(np.zeros(0, np.int32))
            continue
        k = max(1, rng.binomial(m, density))
        if i in long_rows:
            k = min(m, 5000)
        cols = np.sort(rng.choice(m, size=k, replace=False)).astype(np.int32)
        if diag is not None and i < m and i not in cols:
            cols = np.sort(np.append(cols, i)).astype(np.int32)
        rows.append(cols)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    data = rng.standard_normal(int(rp[-1]))
    if diag is not None:
        for i in range(min(n, m)):
            j = rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i)
            if j < rp[i + 1] and ci[j] == i:
                data[j] = diag
    return rp, ci, data


def band_csr(n, diag, offs):
    """Symmetric band CSR: diagonal `diag`, -1 on the given off-diagonals."""
    rows = [sorted({i} | {i + o for o in offs if 0 <= i + o < n}
                   | {i - o for o in offs if 0 <= i - o < n}) for i in range(n)]
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.asarray(r, np.int32) for r in rows])
    data = np.array([diag if j == i else -1.0 for i, r in enumerate(rows) for j in r])
    return rp, ci, data


def tridiag(n):
    """IncompleteFactorizationTest.testTriDiagonal (2 / -1)."""
    return band_csr(n, 2.0, (1,))


def pentadiag(n):
    """IncompleteFactorizationTest.testPentaDiagonal (4 / -1 on +-1, +-2)."""
    return band_csr(n, 4.0, (1, 2))


def oracle_well_mismatch(po, pr, m, q_obs, need_grad=True):
    """CPU FP64 forward + adjoint of J(m) = sum_w (q_w - q_obs_w)^2 on the
    oracle (K_e = exp(m_e) K0_e with K0 = the problem's tensor shape
    K6 / K6_xx, wells = pr.pinned): (J, q, dJ/dm)."""
    n = pr.n
    K0 = pr.K6 / pr.K6[:, :1]
    K6 = np.exp(m)[:, None] * K0
    md_ptr, md_el, md_t, fails = po.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, K6)
    here, there = po.mesh_connections(pr.dim, pr.nx, pr.ny, pr.nz)
    rp, ci = po.pattern(n, here, there, md_ptr, md_el)
    Af = po.assemble_flux(n, here, there, md_ptr, md_el, md_t, rp, ci)
    A = Af.copy()
    wells = np.asarray(pr.pinned, np.int32)
    po.zero_rows(rp, ci, A, 1.0, wells)
    b = pr.rhs()
    p, st, _ = po.solve("bicgstab", rp, ci, A, b, rtol=1e-13, atol=1e-300, prec="ilu")
    assert st.status == 0
    flux = po.connection_flux(md_ptr, md_el, md_t, p)
    q = po.cell_rates(here, there, flux, wells)
    r = q - q_obs
    J = float(r @ r)
    if not need_grad:
        return J, q, None
    rho = np.zeros(n)
    rho[wells] = 2 * r
    g = scipy_csr(rp, ci, Af).T @ rho
    At = scipy_csr(rp, ci, A).T.tocsr()
    At.sort_indices()
    lam, st2, _ = po.solve("bicgstab", At.indptr.astype(np.int32), At.indices.astype(np.int32),
                           At.data.copy(), g, rtol=1e-13, atol=1e-300, prec="ilu")
    assert st2.status == 0
    lam[wells] = 0
    w = rho - lam
    conn = np.repeat(np.arange(len(here)), np.diff(md_ptr))
    tbar = (w[here] - w[there])[conn] * p[md_el]
    mbar = po.mpfa_o_vjp(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, K6, md_ptr, md_el,
                         tbar)
    return J, q, mbar
