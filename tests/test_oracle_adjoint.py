"""Oracle adjoint of the MPFA-O transmissibility map (oracle/mpfa.c
or_mpfa_o_vjp), pinned by central finite differences: with K_e = exp(m_e)
K0_e and phi(m) = sum_q tbar_q t_q(m), the returned mbar must equal
d phi / d m.  rs has no gradients; this is the new capability's own check."""
import numpy as np
import pytest

from ressim_amd import problems as P


def _phi(oracle, pr, K6, tbar):
    md_ptr, md_el, md_t, f = oracle.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, K6)
    return float(np.dot(tbar, md_t)), md_ptr, md_el


@pytest.mark.parametrize("pr", [P.config1(nx=7, ny=6), P.small3d(nx=5, ny=4, nz=4),
                                P.config5(nx=6, ny=5, nz=4)], ids=lambda p: p.name)
def test_mpfa_vjp_directional_fd(oracle, pr):
    rng = np.random.default_rng(3)
    md_ptr, md_el, md_t, _ = oracle.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz,
                                           pr.K6)
    tbar = rng.standard_normal(md_t.size)
    mbar = oracle.mpfa_o_vjp(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                             md_ptr, md_el, tbar)
    h = 1e-5
    for trial in range(3):
        dm = rng.standard_normal(pr.n)
        Kp = pr.K6 * np.exp(h * dm)[:, None]
        Km = pr.K6 * np.exp(-h * dm)[:, None]
        fp, p1, e1 = _phi(oracle, pr, Kp, tbar)
        fm, p2, e2 = _phi(oracle, pr, Km, tbar)
        # the MD structure is unchanged by the perturbation (full tensor)
        assert np.array_equal(p1, md_ptr) and np.array_equal(e1, md_el)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - mbar @ dm) <= 1e-7 * max(1.0, abs(fd))


def test_mpfa_vjp_single_cells(oracle):
    pr = P.small3d(nx=4, ny=4, nz=3)
    md_ptr, md_el, md_t, _ = oracle.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz,
                                           pr.K6)
    tbar = np.random.default_rng(8).standard_normal(md_t.size)
    mbar = oracle.mpfa_o_vjp(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                             md_ptr, md_el, tbar)
    h = 1e-6
    for e in (0, 5, 17, pr.n - 1):
        Kp, Km = pr.K6.copy(), pr.K6.copy()
        Kp[e] *= np.exp(h)
        Km[e] *= np.exp(-h)
        fd = (_phi(oracle, pr, Kp, tbar)[0] - _phi(oracle, pr, Km, tbar)[0]) / (2 * h)
        assert abs(fd - mbar[e]) <= 1e-6 * max(1.0, abs(fd))


def test_mpfa_vjp_uniform_scaling(oracle):
    """T is homogeneous of degree one in K: scaling every tensor by e^s
    scales every transmissibility by e^s, so sum_e mbar_e = tbar . t."""
    pr = P.small3d(nx=5, ny=4, nz=3)
    md_ptr, md_el, md_t, _ = oracle.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz,
                                           pr.K6)
    tbar = np.random.default_rng(2).standard_normal(md_t.size)
    mbar = oracle.mpfa_o_vjp(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                             md_ptr, md_el, tbar)
    assert abs(mbar.sum() - tbar @ md_t) <= 1e-12 * max(1.0, abs(tbar) @ abs(md_t))
