"""Device IMPES pressure Jacobian against the oracle (oracle/impes.c):
control-surface secondaries, A's values and b bit-exact, isothermal and
thermal (Fourier lists), with regular and outlet sources (one cell in two
sources), and the pattern error."""
import numpy as np
import pytest

from ressim_amd import problems as P

pytestmark = pytest.mark.gpu

GRID = dict(dim=3, nx=9, ny=7, nz=4, dx=10.0, dy=12.0, dz=3.0)


@pytest.fixture(scope="module")
def mesh(oracle):
    g = GRID
    n = g["nx"] * g["ny"] * g["nz"]
    logk = P.normal(9, 0, n)
    K6 = P.full_tensor(1e-13 * np.exp(logk), 1e-13 * np.exp(logk), 1e-14 * np.exp(logk))
    md_ptr, md_elem, md_t, _ = oracle.mpfa_o(g["dim"], g["nx"], g["ny"], g["nz"], g["dx"],
                                             g["dy"], g["dz"], K6)
    here, there = oracle.mesh_connections(g["dim"], g["nx"], g["ny"], g["nz"])
    rp, ci = oracle.pattern(n, here, there, md_ptr, md_elem)
    return dict(n=n, md_ptr=md_ptr, md_elem=md_elem, md_t=md_t, here=here, there=there,
                rowptr=rp, colind=ci)


def _device(st):
    import torch
    return {k: torch.tensor(v, device="cuda") for k, v in st.items()}


@pytest.mark.parametrize("case", ["isothermal", "thermal", "centrifuge", "csr-lists",
                                  "five-components"])
def test_impes_assembly_bitexact(gpu, oracle, mesh, monkeypatch, case):
    """Slot-major MD lists (the default for MPFA-O lists) and the CSR lists
    (RSGPU_IMPES_CSR), component counts with register-resident mobilities
    (3) and without (5)."""
    import torch
    from ressim_amd.impes import PressureJacobian
    from ressim_amd.potential import FluxOperator
    if case == "csr-lists":
        monkeypatch.setenv("RSGPU_IMPES_CSR", "1")
    m, n = mesh, mesh["n"]
    ncomp = 5 if case == "five-components" else 3
    thermal = case == "thermal"
    print(f"IMPES assembly {case}", flush=True)
    g = GRID
    st = P.impes_state(11, g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["dz"], ncomp=ncomp,
                       thermal=thermal)
    params = dict(P.IMPES_PARAMS, thermal=int(thermal))
    if case == "centrifuge":
        params.update(omega=0.4, ax=20.0, ay=-5.0)
    F = FluxOperator.from_lists(n, m["md_ptr"], m["md_elem"], m["md_t"], m["here"], m["there"])
    mf_t = 0.37 * m["md_t"]
    Fmf = FluxOperator.from_lists(n, m["md_ptr"], m["md_elem"], mf_t, m["here"],
                                  m["there"]) if thermal else None
    J = PressureJacobian(F, ncomp, fourier=Fmf)
    elem, kind, q, qe = P.impes_sources(3, n, ncomp)
    J.setSources(elem, kind, q, qe)
    A = gpu.CompRowMatrix.from_csr(n, n, m["rowptr"], m["colind"])
    sd = _device(st)
    rho, up, u = J.controlSurfaces(params, sd)
    op = dict(params, n=n, nconn=len(m["here"]), ncomp=ncomp)
    rho_o, up_o, u_o = oracle.impes_cs(op, st, m["here"], m["there"], m["md_ptr"],
                                       m["md_elem"], m["md_t"])
    assert np.array_equal(rho.cpu().numpy(), rho_o)
    assert np.array_equal(up.cpu().numpy(), up_o)
    assert np.array_equal(u.cpu().numpy(), u_o)
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    J.assemble(A, params, sd, rho, up, b)
    code, val_o, b_o = oracle.impes_assemble(
        op, st, m["here"], m["there"], m["md_ptr"], m["md_elem"], m["md_t"], rho_o, up_o,
        m["rowptr"], m["colind"], mf=(m["md_ptr"], m["md_elem"], mf_t) if thermal else None,
        sources=(elem, kind, q, qe))
    assert code == 0
    assert np.array_equal(A.getData(), val_o)
    assert np.array_equal(b.cpu().numpy(), b_o)
    # a second Newton iteration on changed secondaries reuses A and b
    st2 = dict(st)
    st2["phase_pressure"] = st["phase_pressure"] + 1.0e3
    sd2 = _device(st2)
    rho2, up2, _ = J.controlSurfaces(params, sd2)
    J.assemble(A, params, sd2, rho2, up2, b)
    rho2o, up2o, _ = oracle.impes_cs(op, st2, m["here"], m["there"], m["md_ptr"],
                                     m["md_elem"], m["md_t"])
    _, val2, b2 = oracle.impes_assemble(
        op, st2, m["here"], m["there"], m["md_ptr"], m["md_elem"], m["md_t"], rho2o, up2o,
        m["rowptr"], m["colind"], mf=(m["md_ptr"], m["md_elem"], mf_t) if thermal else None,
        sources=(elem, kind, q, qe))
    assert np.array_equal(A.getData(), val2)
    assert np.array_equal(b.cpu().numpy(), b2)


def test_impes_errors(gpu, oracle, mesh):
    import torch
    from ressim_amd.impes import PressureJacobian
    from ressim_amd.potential import FluxOperator
    m, n = mesh, mesh["n"]
    g = GRID
    st = P.impes_state(2, g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["dz"], ncomp=2)
    F = FluxOperator.from_lists(n, m["md_ptr"], m["md_elem"], m["md_t"], m["here"], m["there"])
    J = PressureJacobian(F, 2)
    sd = _device(st)
    rho, up, _ = J.controlSurfaces(P.IMPES_PARAMS, sd)
    # a pattern without row 0's last entry
    rp, ci = m["rowptr"].copy(), m["colind"].copy()
    ci2 = np.delete(ci, rp[1] - 1)
    rp2 = rp.copy()
    rp2[1:] -= 1
    A = gpu.CompRowMatrix.from_csr(n, n, rp2, ci2)
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    with pytest.raises(IndexError, match="outside the matrix pattern"):
        J.assemble(A, P.IMPES_PARAMS, sd, rho, up, b)
    # thermal without Fourier lists
    A = gpu.CompRowMatrix.from_csr(n, n, m["rowptr"], m["colind"])
    with pytest.raises(ValueError):
        J.assemble(A, dict(P.IMPES_PARAMS, thermal=1), sd, rho, up, b)
    with pytest.raises(ValueError):
        J.setSources([n], [0], np.zeros((1, 3, 2)), [0.0])


@pytest.mark.parametrize("truncate", [False, True])
def test_molar_update_bitexact(gpu, oracle, mesh, truncate):
    """MolarMassExplicit.solve on the device: N_new bit-exact against the
    oracle with sources and locked cells; negative mole numbers truncated
    (count) or reported (first element and component)."""
    import torch
    from ressim_amd.impes import PressureJacobian
    from ressim_amd.potential import FluxOperator
    m, n, ncomp = mesh, mesh["n"], 3
    g = GRID
    print(f"molar update truncate={truncate}", flush=True)
    st = P.impes_state(13, g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["dz"], ncomp=ncomp)
    F = FluxOperator.from_lists(n, m["md_ptr"], m["md_elem"], m["md_t"], m["here"], m["there"])
    J = PressureJacobian(F, ncomp)
    elem, kind, q, qe = P.impes_sources(7, n, ncomp)
    J.setSources(elem, kind, q, qe)
    params = dict(P.IMPES_PARAMS)
    sd = _device(st)
    rho, up, u = J.controlSurfaces(params, sd)
    op = dict(params, n=n, nconn=len(m["here"]), ncomp=ncomp)
    _, up_o, u_o = oracle.impes_cs(op, st, m["here"], m["there"], m["md_ptr"], m["md_elem"],
                                   m["md_t"])
    locked = [5, 40, 41]
    N_old = np.where(np.arange(ncomp * n) % 5 == 0, 1.0, 1.0e10)
    code_o, N_o, tr_o, bad_o = oracle.impes_molar_update(
        op, st, m["here"], m["there"], u_o, up_o, N_old, sources=(elem, kind, q, qe),
        locked=locked, truncate=truncate)
    Nd = torch.tensor(N_old, device="cuda")
    Nn = torch.empty_like(Nd)
    if truncate or code_o == 0:
        tr = J.molarUpdate(params, sd, u, up, Nd, Nn, locked=locked, truncate=truncate)
        assert tr == tr_o
    else:
        with pytest.raises(ValueError, match="Negative number of moles"):
            J.molarUpdate(params, sd, u, up, Nd, Nn, locked=locked, truncate=truncate)
        assert J.bad == bad_o
    assert np.array_equal(Nn.cpu().numpy(), N_o)


def test_temperature_system_bitexact(gpu, oracle, mesh):
    """TemperatureNewtonRaphson's A and b on the device bit-exact against
    the oracle (advective flux, Fourier flux, diagonal part, sources)."""
    import torch
    from ressim_amd.impes import PressureJacobian
    from ressim_amd.potential import FluxOperator
    m, n, ncomp = mesh, mesh["n"], 2
    g = GRID
    st = P.impes_state(17, g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["dz"], ncomp=ncomp,
                       thermal=True)
    F = FluxOperator.from_lists(n, m["md_ptr"], m["md_elem"], m["md_t"], m["here"], m["there"])
    mf_t = 0.21 * m["md_t"]
    Fmf = FluxOperator.from_lists(n, m["md_ptr"], m["md_elem"], mf_t, m["here"], m["there"])
    J = PressureJacobian(F, ncomp, fourier=Fmf)
    elem, kind, q, qe = P.impes_sources(9, n, ncomp)
    J.setSources(elem, kind, q, qe)
    params = dict(P.IMPES_PARAMS, thermal=1, omega=0.2, ax=3.0, ay=4.0)
    sd = _device(st)
    rho, up, u = J.controlSurfaces(params, sd)
    A = gpu.CompRowMatrix.from_csr(n, n, m["rowptr"], m["colind"])
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    J.assembleTemperature(A, params, sd, rho, up, b)
    op = dict(params, n=n, nconn=len(m["here"]), ncomp=ncomp)
    rho_o, up_o, _ = oracle.impes_cs(op, st, m["here"], m["there"], m["md_ptr"], m["md_elem"],
                                     m["md_t"])
    code, val_o, b_o = oracle.impes_temperature(op, st, m["here"], m["there"], m["md_ptr"],
                                                m["md_elem"], m["md_t"],
                                                (m["md_ptr"], m["md_elem"], mf_t), rho_o, up_o,
                                                m["rowptr"], m["colind"],
                                                sources=(elem, kind, q, qe))
    assert code == 0
    assert np.array_equal(A.getData(), val_o)
    assert np.array_equal(b.cpu().numpy(), b_o)
