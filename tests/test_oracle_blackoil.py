"""The oracle's fully-implicit black-oil Jacobian, pinned without the GPU.

The reference has no fully-implicit Jacobian, so the oracle (oracle/
blackoil.c) is pinned by what the mathematics fixes: the Jacobian equals
central finite differences of the oracle's own residual; face fluxes are
antisymmetric (the residual sums to the accumulation change); a single cell
has only the accumulation closed form; a two-cell column of one phase gives
the TPFA / upwind / Corey closed form (kr = ke Se^n, Se = (S - Sr) / (1 -
sum Sr), rs/src/no/uib/cipr/rs/rock/CoreyRockFluid.java:129-152); the block
pattern has the 7-point size and its expansion is the Kronecker structure.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from ressim_amd import problems as P


def _jac(po, pr, u, u_old):
    brp, bci = po.bo_pattern(pr.nx, pr.ny, pr.nz)
    rp, ci = po.bo_expand(brp, bci)
    val, R = po.bo_assemble(pr.params, pr.nx, pr.ny, pr.nz, pr.kh, u, u_old)
    n3 = 3 * pr.n
    return sp.csr_matrix((val, ci, rp), shape=(n3, n3)), R


def _noisy(pr, seed=0):
    """The problem's state with pressures off the hydrostatic line (no
    face sits exactly at an upwind switch) and an older state 1 bar lower."""
    rng = np.random.default_rng(seed)
    u = pr.u.copy()
    u[0::3] += rng.uniform(-500.0, 500.0, pr.n)
    u_old = pr.u.copy()
    u_old[0::3] -= 1e5
    return u, u_old


def test_pattern_and_expansion(oracle):
    for nx, ny, nz in ((5, 4, 3), (1, 6, 2), (7, 1, 1)):
        pr = P.config3(nx=nx, ny=ny, nz=nz)
        brp, bci = oracle.bo_pattern(nx, ny, nz)
        assert len(bci) == pr.nnzb
        B = sp.csr_matrix((np.ones(len(bci)), bci, brp), shape=(pr.n, pr.n))
        assert (B != B.T).nnz == 0 and B.diagonal().min() == 1
        for i in range(pr.n):
            assert np.all(np.diff(bci[brp[i]:brp[i + 1]]) > 0)
        rp, ci = oracle.bo_expand(brp, bci)
        E = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(3 * pr.n, 3 * pr.n))
        K = sp.kron(B, np.ones((3, 3))).tocsr()
        assert (E != K).nnz == 0
        for r in range(3 * pr.n):
            assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)


@pytest.mark.parametrize("expo", [(2, 2, 2), (1, 3, 2)])
def test_jacobian_equals_central_differences(oracle, expo):
    pr = P.config3(nx=5, ny=4, nz=3, nw=expo[0], no=expo[1], ng=expo[2])
    u, u_old = _noisy(pr)
    J, R = _jac(oracle, pr, u, u_old)
    Jd = J.toarray()
    worst = 0.0
    for k in range(3 * pr.n):
        h = 1.0 if k % 3 == 0 else 1e-6
        up, um = u.copy(), u.copy()
        up[k] += h
        um[k] -= h
        _, Rp = oracle.bo_assemble(pr.params, pr.nx, pr.ny, pr.nz, pr.kh, up, u_old, False)
        _, Rm = oracle.bo_assemble(pr.params, pr.nx, pr.ny, pr.nz, pr.kh, um, u_old, False)
        fd = (Rp - Rm) / (2 * h)
        worst = max(worst, np.abs(fd - Jd[:, k]).max() / np.abs(Jd[:, k]).max())
        # nothing outside the pattern depends on u_k
        outside = np.setdiff1d(np.arange(3 * pr.n), J[:, k].nonzero()[0])
        assert np.all(fd[outside] == 0.0)
    assert worst < 1e-6, worst


def test_a_dropped_term_fails_the_difference_check(oracle):
    """The check above catches a plausible mistake: zeroing the
    dissolved-gas pressure derivative (rs0 -> 0 only in the Jacobian is
    emulated by comparing against the residual with dissolved gas)."""
    pr = P.config3(nx=4, ny=3, nz=2)
    u, u_old = _noisy(pr, 1)
    pr0 = P.config3(nx=4, ny=3, nz=2, rs0=0.0)
    J0, _ = _jac(oracle, pr0, u, u_old)
    k = 3 * 5
    up, um = u.copy(), u.copy()
    up[k] += 1.0
    um[k] -= 1.0
    _, Rp = oracle.bo_assemble(pr.params, pr.nx, pr.ny, pr.nz, pr.kh, up, u_old, False)
    _, Rm = oracle.bo_assemble(pr.params, pr.nx, pr.ny, pr.nz, pr.kh, um, u_old, False)
    fd = (Rp - Rm) / 2.0
    col = J0.toarray()[:, k]
    assert np.abs(fd - col).max() / np.abs(col).max() > 1e-3


def test_fluxes_cancel(oracle):
    pr = P.config3(nx=6, ny=5, nz=4)
    u, _ = _noisy(pr, 2)
    _, R = oracle.bo_assemble(pr.params, pr.nx, pr.ny, pr.nz, pr.kh, u, u, False)
    for c in range(3):
        assert abs(R[c::3].sum()) <= 1e-12 * np.abs(R[c::3]).sum()


def _props(p, sw, sg, q):
    so = 1.0 - sw - sg
    bw = q["bw0"] * (1 + q["cw"] * (p - q["pref"]))
    bo = q["bo0"] * (1 + q["co"] * (p - q["pref"]))
    bg = q["bg0"] * p / q["pref"]
    rs = q["rs0"] * p / q["pref"]
    phi = q["phi0"] * (1 + q["cr"] * (p - q["pref"]))
    return so, bw, bo, bg, rs, phi


def test_single_cell_accumulation_closed_form(oracle):
    pr = P.config3(nx=1, ny=1, nz=1)
    q = pr.params
    u = np.array([210e5, 0.25, 0.15])
    u_old = np.array([205e5, 0.2, 0.3])
    _, R = oracle.bo_assemble(q, 1, 1, 1, pr.kh, u, u_old, False)
    vdt = q["dx"] * q["dy"] * q["dz"] / q["dt"]

    def acc(v):
        so, bw, bo, bg, rs, phi = _props(v[0], v[1], v[2], q)
        return np.array([phi * v[1] * bw, phi * so * bo, phi * (v[2] * bg + rs * so * bo)])

    assert np.allclose(R, vdt * (acc(u) - acc(u_old)), rtol=1e-13, atol=0)


def test_two_cell_column_upwind_closed_form(oracle):
    """Two cells stacked in z with equal saturations: the water flux is
    T kr_w/mu_w b_w,up (p_i - p_j - rho_avg g (z_i - z_j)), upstream by
    potential, T = dx dy / (dz/2/kz_i + dz/2/kz_j), Corey kr = Se^2."""
    pr = P.config3(nx=1, ny=1, nz=2)
    q = pr.params
    kh = np.array([2e-13, 5e-14])
    u = np.array([200e5, 0.3, 0.2, 200e5 + 5e4, 0.3, 0.2])
    _, R = oracle.bo_assemble(q, 1, 1, 2, kh, u, u, False)
    kz = q["kz_ratio"] * kh
    T = q["dx"] * q["dy"] / (0.5 * q["dz"] / kz[0] + 0.5 * q["dz"] / kz[1])
    p = u[0::3]
    so, bw, bo, bg, rs, phi = _props(p, 0.3, 0.2, q)
    rho_w = q["rhow_s"] * bw
    dphi = (p[0] - p[1]) - 0.5 * (rho_w[0] + rho_w[1]) * q["g"] * (-q["dz"])
    up = 0 if dphi >= 0 else 1
    se = (0.3 - q["swr"]) / (1 - q["swr"] - q["sor"] - q["sgr"])
    Fw = T * (q["kw"] * se ** 2 / q["muw"]) * bw[up] * dphi
    assert np.isclose(R[0], Fw, rtol=1e-12) and np.isclose(R[3], -Fw, rtol=1e-12)


def test_newton_system_solves_with_ilu(oracle):
    pr = P.config3(nx=8, ny=6, nz=4)
    J, R = _jac(oracle, pr, pr.u, pr.u_old)
    x, st, _ = oracle.solve("bicgstab", J.indptr.astype(np.int32), J.indices.astype(np.int32),
                            J.data.copy(), -R, rtol=1e-10, prec="ilu")
    assert st.status == 0
    assert np.linalg.norm(J @ x + R) <= 1e-9 * np.linalg.norm(R)
