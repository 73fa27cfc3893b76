"""The IMPES pressure-Jacobian oracle (oracle/impes.c) pinned by what the
mathematics fixes, independently of retyping its formulas:

* single phase, single component, unit mobility and unit dRdN, no gravity:
  the flux part of A is minus PotentialCalculator.discretizeFlux's matrix
  (or_assemble_flux, pinned in test_oracle_mpfa.py) bit for bit, and b is
  minus the cells' net outflow (or_connection_flux + or_cell_rates);
* with frozen secondaries (upstream, density, mobilities) the flux residual
  is linear in the phase pressures and A is its derivative: finite
  differences of b in each cell's pressure reproduce A's columns, for three
  phases, several components, gravity, centrifuge and thermal convection;
* the control-surface density is the saturation-weighted average (0 where
  both saturations vanish), the upstream side follows the sign of the
  potential sum, and the body force is g z - omega^2/2 r^2;
* with every phase absent the flux part vanishes and A's diagonal is
  dRdp/dt; a regular source with unit dRdN adds its total component rate
  to b, an outlet source its rate summed over phases and components.
"""
import numpy as np
import pytest

from ressim_amd import problems as P

GRID = dict(dim=3, nx=5, ny=4, nz=3, dx=10.0, dy=12.0, dz=3.0)


@pytest.fixture(scope="module")
def mesh(oracle):
    g = GRID
    n = g["nx"] * g["ny"] * g["nz"]
    logk = P.normal(3, 0, n)
    K6 = P.full_tensor(1e-13 * np.exp(logk), 1e-13 * np.exp(logk), 1e-14 * np.exp(logk))
    md_ptr, md_elem, md_t, fails = oracle.mpfa_o(g["dim"], g["nx"], g["ny"], g["nz"], g["dx"],
                                                 g["dy"], g["dz"], K6)
    here, there = oracle.mesh_connections(g["dim"], g["nx"], g["ny"], g["nz"])
    rp, ci = oracle.pattern(n, here, there, md_ptr, md_elem)
    return dict(n=n, K6=K6, md_ptr=md_ptr, md_elem=md_elem, md_t=md_t, here=here,
                there=there, rowptr=rp, colind=ci)


def _params(mesh, ncomp, thermal=0, **kw):
    p = dict(n=mesh["n"], nconn=len(mesh["here"]), ncomp=ncomp, thermal=thermal, dt=86400.0,
             g=9.81, omega=0.0, ax=0.0, ay=0.0)
    p.update(kw)
    return p


def _state(ncomp, thermal=False, seed=1):
    g = GRID
    return P.impes_state(seed, g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["dz"],
                         ncomp=ncomp, thermal=thermal)


def _run(oracle, mesh, params, st, mf=None, sources=None):
    m = mesh
    rho, up, u = oracle.impes_cs(params, st, m["here"], m["there"], m["md_ptr"], m["md_elem"],
                                 m["md_t"])
    code, val, b = oracle.impes_assemble(params, st, m["here"], m["there"], m["md_ptr"],
                                         m["md_elem"], m["md_t"], rho, up, m["rowptr"],
                                         m["colind"], mf=mf, sources=sources)
    assert code == 0
    return rho, up, u, val, b


def test_single_phase_reduces_to_potential_assembly(oracle, mesh):
    n = mesh["n"]
    st = _state(1)
    st["present"] = np.concatenate([np.ones(n), np.zeros(2 * n)]).astype(np.uint8)
    st["comp_mobility"] = np.concatenate([np.ones(n), np.zeros(2 * n)])
    st["dRdN"] = np.ones(n)
    st["dRdp"] = np.zeros(n)
    st["R0"] = np.zeros(n)
    st["p0"] = st["pressure"].copy()
    params = _params(mesh, 1, g=0.0)
    rho, up, u, val, b = _run(oracle, mesh, params, st)
    A = oracle.assemble_flux(n, mesh["here"], mesh["there"], mesh["md_ptr"], mesh["md_elem"],
                             mesh["md_t"], mesh["rowptr"], mesh["colind"])
    assert np.array_equal(val, -A)
    pw = st["phase_pressure"][:n]
    flux = oracle.connection_flux(mesh["md_ptr"], mesh["md_elem"], mesh["md_t"], pw)
    rates = oracle.cell_rates(mesh["here"], mesh["there"], flux, np.arange(n))
    assert np.array_equal(b, -rates)


def _dense(rp, ci, val, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return A


@pytest.mark.parametrize("thermal", [0, 1])
@pytest.mark.parametrize("omega", [0.0, 0.3])
def test_jacobian_is_the_derivative_of_the_flux_residual(oracle, mesh, thermal, omega):
    n, ncomp = mesh["n"], 3
    st = _state(ncomp, thermal=bool(thermal))
    st["dRdp"] = np.zeros(n)          # flux part only
    st["R0"] = np.zeros(n)
    st["p0"] = st["pressure"].copy()
    params = _params(mesh, ncomp, thermal=thermal, omega=omega, ax=7.0, ay=-3.0)
    mf = (mesh["md_ptr"], mesh["md_elem"], 0.5 * mesh["md_t"]) if thermal else None
    rho, up, u, val, b0 = _run(oracle, mesh, params, st, mf=mf)
    A = _dense(mesh["rowptr"], mesh["colind"], val, n)
    h = 50.0
    for k in range(0, n, 7):
        s2 = dict(st)
        pp = st["phase_pressure"].reshape(3, n).copy()
        pp[:, k] += h
        s2["phase_pressure"] = pp.reshape(-1)
        # secondaries frozen: the base state's rho and upstream
        code, _, b1 = oracle.impes_assemble(params, s2, mesh["here"], mesh["there"],
                                            mesh["md_ptr"], mesh["md_elem"], mesh["md_t"], rho,
                                            up, mesh["rowptr"], mesh["colind"], mf=mf)
        assert code == 0
        fd = (b1 - b0) / h
        scale = np.abs(A[:, k]).max()
        assert scale > 0
        assert np.abs(fd - A[:, k]).max() <= 1e-6 * scale, k


def test_control_surfaces(oracle):
    # two cells, one connection, MD = [(0, +t), (1, -t)]
    t = 2.0e-12
    md_ptr = np.array([0, 2], np.int32)
    md_elem = np.array([0, 1], np.int32)
    md_t = np.array([t, -t])
    here, there = np.array([0], np.int32), np.array([1], np.int32)
    S = np.array([[0.5, 0.25], [0.5, 0.75], [0.0, 0.0]])      # gas absent on both sides
    rhoc = np.array([[1000.0, 1010.0], [800.0, 820.0], [100.0, 110.0]])
    pp = np.array([[2.0e7, 2.01e7], [2.0e7 + 10, 1.9e7], [2.0e7, 2.0e7]])
    mob = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])
    cx, cy, cz = np.array([1.0, 3.0]), np.array([0.5, 0.5]), np.array([100.0, 104.0])
    st = dict(saturation=S.reshape(-1), mass_density=rhoc.reshape(-1),
              phase_pressure=pp.reshape(-1), phase_mobility=mob.reshape(-1),
              cx=cx, cy=cy, cz=cz)
    g, om, ax, ay = 9.81, 0.2, 2.0, 1.0
    params = dict(n=2, nconn=1, ncomp=1, thermal=0, dt=1.0, g=g, omega=om, ax=ax, ay=ay)
    rho, up, u = oracle.impes_cs(params, st, here, there, md_ptr, md_elem, md_t)
    bf = g * cz - om * om / 2 * ((ax - cx) ** 2 + (ay - cy) ** 2)
    for ph in range(3):
        s = S[ph].sum()
        want = (S[ph] * rhoc[ph]).sum() / s if s > 0 else 0.0
        assert rho[ph] == pytest.approx(want, rel=1e-15, abs=0)
        D = t * (pp[ph, 0] + want * bf[0]) - t * (pp[ph, 1] + want * bf[1])
        assert up[ph] == (0 if D > 0 else 1)
        assert u[ph] == pytest.approx(mob[ph, up[ph]] * D, rel=1e-12)
    assert up[0] == 1 and up[1] == 0       # water flows there->here, oil here->there


def test_diagonal_part_and_sources(oracle, mesh):
    n, ncomp = mesh["n"], 2
    st = _state(ncomp)
    st["present"] = np.zeros(3 * n, np.uint8)       # no phase flows: no flux part
    params = _params(mesh, ncomp)
    elem, kind, q, qe = P.impes_sources(5, n, ncomp)
    _, _, _, val, b0 = _run(oracle, mesh, params, st)
    A = _dense(mesh["rowptr"], mesh["colind"], val, n)
    assert np.array_equal(np.diag(A), st["dRdp"] / params["dt"])
    assert np.count_nonzero(A - np.diag(np.diag(A))) == 0
    dp = st["pressure"] - st["p0"]
    assert np.allclose(b0, (st["R0"] + st["dRdp"] * dp) / params["dt"], rtol=1e-14, atol=0)
    st["dRdN"] = np.ones(ncomp * n)
    _, _, _, _, b0 = _run(oracle, mesh, params, st)
    _, _, _, _, b1 = _run(oracle, mesh, params, st, sources=(elem, kind, q, qe))
    want = np.zeros(n)
    for t in range(len(elem)):
        want[elem[t]] += q[t, 0].sum() if kind[t] == 0 else q[t].sum()
    assert np.allclose(b1 - b0, want, rtol=1e-9, atol=1e-18)


def test_missing_pattern_entry(oracle, mesh):
    n = mesh["n"]
    rp, ci = mesh["rowptr"].copy(), mesh["colind"].copy()
    # drop row 0's last entry from the pattern
    ci2 = np.delete(ci, rp[1] - 1)
    rp2 = rp.copy()
    rp2[1:] -= 1
    st = _state(2)
    params = _params(mesh, 2)
    rho, up, u = oracle.impes_cs(params, st, mesh["here"], mesh["there"], mesh["md_ptr"],
                                 mesh["md_elem"], mesh["md_t"])
    code, _, _ = oracle.impes_assemble(params, st, mesh["here"], mesh["there"], mesh["md_ptr"],
                                       mesh["md_elem"], mesh["md_t"], rho, up, rp2, ci2)
    assert code == 12


def test_molar_update_conserves_and_reduces(oracle, mesh):
    """Without sources and locks the connection fluxes move moles between
    cells only (the total of each component is unchanged up to round-off);
    with one phase, one component, unit mole fraction and molar density, dN
    is dt times the cells' net Darcy inflow (the negated cell rates of the
    control surfaces' Darcy fluxes)."""
    n, nconn, ncomp = mesh["n"], len(mesh["here"]), 2
    st = _state(ncomp)
    params = _params(mesh, ncomp)
    rho, up, u = oracle.impes_cs(params, st, mesh["here"], mesh["there"], mesh["md_ptr"],
                                 mesh["md_elem"], mesh["md_t"])
    N_old = 1.0e10 * (1 + np.arange(ncomp * n) % 7)
    code, N_new, tr, bad = oracle.impes_molar_update(params, st, mesh["here"], mesh["there"],
                                                     u, up, N_old)
    assert code == 0
    dN = (N_new - N_old).reshape(ncomp, n)
    for nu in range(ncomp):
        assert abs(dN[nu].sum()) <= 1e-9 * np.abs(dN[nu]).sum()
    # one phase, one component
    st1 = dict(st)
    st1["present"] = np.concatenate([np.ones(n), np.zeros(2 * n)]).astype(np.uint8)
    st1["mole_fraction"] = np.concatenate([np.ones(n), np.zeros(2 * n)])
    st1["molar_density"] = np.ones(3 * n)
    p1 = _params(mesh, 1)
    rho1, up1, u1 = oracle.impes_cs(p1, st1, mesh["here"], mesh["there"], mesh["md_ptr"],
                                    mesh["md_elem"], mesh["md_t"])
    # N_old = 0: N_new is dN itself (negative entries are reported, not truncated)
    code, N1, _, _ = oracle.impes_molar_update(p1, st1, mesh["here"], mesh["there"], u1, up1,
                                               np.zeros(n))
    assert code in (0, 13)
    rates = oracle.cell_rates(mesh["here"], mesh["there"], u1[:nconn], np.arange(n))
    np.testing.assert_allclose(N1, -p1["dt"] * rates, rtol=1e-12,
                               atol=1e-14 * p1["dt"] * np.abs(rates).max())


def test_molar_update_sources_locks_and_negatives(oracle, mesh):
    n, ncomp = mesh["n"], 2
    st = _state(ncomp)
    params = _params(mesh, ncomp)
    rho, up, u = oracle.impes_cs(params, st, mesh["here"], mesh["there"], mesh["md_ptr"],
                                 mesh["md_elem"], mesh["md_t"])
    elem, kind, q, qe = P.impes_sources(5, n, ncomp)
    N_old = np.full(ncomp * n, 1.0e10)
    _, Na, _, _ = oracle.impes_molar_update(params, st, mesh["here"], mesh["there"], u, up,
                                            N_old)
    _, Nb, _, _ = oracle.impes_molar_update(params, st, mesh["here"], mesh["there"], u, up,
                                            N_old, sources=(elem, kind, q, qe))
    want = np.zeros((ncomp, n))
    for t in range(len(elem)):
        want[:, elem[t]] += params["dt"] * (q[t, 0] if kind[t] == 0 else q[t].sum(axis=0))
    np.testing.assert_allclose((Nb - Na).reshape(ncomp, n), want, rtol=1e-6,
                               atol=1e-9 * params["dt"])
    locked = [3, 17]
    _, Nl, _, _ = oracle.impes_molar_update(params, st, mesh["here"], mesh["there"], u, up,
                                            N_old, locked=locked)
    for c in locked:
        assert np.all(Nl.reshape(ncomp, n)[:, c] == 1.0e10)
    # tiny old mole numbers: negatives truncated (counted) or reported
    N_small = np.full(ncomp * n, 1.0)
    code, Nt, tr, _ = oracle.impes_molar_update(params, st, mesh["here"], mesh["there"], u,
                                                up, N_small, truncate=True)
    assert code == 0 and tr > 0 and Nt.min() == 0.0
    code, _, _, bad = oracle.impes_molar_update(params, st, mesh["here"], mesh["there"], u,
                                                up, N_small)
    assert code == 13
    neg = [e * ncomp + nu for e in range(n) for nu in range(ncomp) if Nt[nu * n + e] == 0.0]
    assert bad == min(neg) and tr == len(neg)


def test_temperature_system(oracle, mesh):
    """With every phase absent the advective part vanishes and the
    temperature system's A is PotentialCalculator's flux assembly of the
    Fourier lists plus delta/dt on the diagonal, b the cells' net Fourier
    outflow of T plus delta dT/dt; the advective energy flux of a connection
    is the upstream energy mobility times the control surface's potential
    sum."""
    n, ncomp = mesh["n"], 2
    st = _state(ncomp, thermal=True)
    params = _params(mesh, ncomp, thermal=1)
    mf = (mesh["md_ptr"], mesh["md_elem"], 0.3 * mesh["md_t"])
    rho, up, u = oracle.impes_cs(params, st, mesh["here"], mesh["there"], mesh["md_ptr"],
                                 mesh["md_elem"], mesh["md_t"])
    s0 = dict(st)
    s0["present"] = np.zeros(3 * n, np.uint8)
    code, val, b = oracle.impes_temperature(params, s0, mesh["here"], mesh["there"],
                                            mesh["md_ptr"], mesh["md_elem"], mesh["md_t"], mf,
                                            rho, up, mesh["rowptr"], mesh["colind"])
    assert code == 0
    Af = oracle.assemble_flux(n, mesh["here"], mesh["there"], mf[0], mf[1], mf[2],
                              mesh["rowptr"], mesh["colind"])
    A = _dense(mesh["rowptr"], mesh["colind"], val, n)
    Ad = _dense(mesh["rowptr"], mesh["colind"], Af, n) + np.diag(st["heat_capacity"] /
                                                               params["dt"])
    np.testing.assert_allclose(A, Ad, rtol=1e-14, atol=0)
    fl = oracle.connection_flux(mf[0], mf[1], mf[2], st["temperature"])
    rates = oracle.cell_rates(mesh["here"], mesh["there"], fl, np.arange(n))
    diag = st["heat_capacity"] * (st["temperature"] - st["T0"]) / params["dt"]
    np.testing.assert_allclose(b, rates + diag, rtol=1e-10,
                               atol=1e-12 * np.abs(rates).max())
    # advective part: one connection's contribution against the CS flux
    code, _, b1 = oracle.impes_temperature(params, st, mesh["here"], mesh["there"],
                                           mesh["md_ptr"], mesh["md_elem"], mesh["md_t"],
                                           (mesh["md_ptr"], mesh["md_elem"],
                                            0.0 * mesh["md_t"]), rho, up, mesh["rowptr"],
                                           mesh["colind"])
    nconn = len(mesh["here"])
    adv = np.zeros(nconn)
    for ph in range(3):
        upc = up[ph * nconn:(ph + 1) * nconn]
        pres = st["present"][ph * n + upc] != 0
        D = u[ph * nconn:(ph + 1) * nconn] / st["phase_mobility"][ph * n + upc].clip(1e-300)
        adv += np.where(pres, st["energy_mobility"][ph * n + upc] * D, 0.0)
    want = oracle.cell_rates(mesh["here"], mesh["there"], adv, np.arange(n)) + diag
    np.testing.assert_allclose(b1, want, rtol=1e-8, atol=1e-10 * np.abs(want).max())
