"""Seeded synthetic inputs for the five benchmark/parity workloads.

Random numbers come from a counter-based SplitMix64 stream (Steele, Lea and
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) turned
into standard normals with the Box-Muller transform.  Everything here is
input generation on the host; the physics and the solve run on the GPU.

Permeability tensors are stored per cell as (xx, yy, zz, xy, xz, yz), the
component order of rs ``Tensor3D(xx,yy,zz,xy,xz,yz)``
(rs/src/no/uib/cipr/rs/geometry/Tensor3D.java:37-60).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """Counter-based SplitMix64: output i = mix(base + (offset+i+1)*golden)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0xC2B2AE3D27D4EB4F) % (1 << 64))
        ctr = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = base + ctr * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """Doubles in [0,1) from the top 53 bits."""
    z = splitmix64(seed, stream, count, offset)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller (cosine branch), two uniforms each."""
    u = uniform01(seed, stream, 2 * count)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def full_tensor(kxx, kyy, kzz, off=0.1):
    """K6 with off-diagonals off*sqrt(k_ii*k_jj) (SPD for off < 0.5)."""
    K = np.empty((kxx.size, 6))
    K[:, 0], K[:, 1], K[:, 2] = kxx, kyy, kzz
    K[:, 3] = off * np.sqrt(kxx * kyy)
    K[:, 4] = off * np.sqrt(kxx * kzz)
    K[:, 5] = off * np.sqrt(kyy * kzz)
    return K


@dataclass
class PressureProblem:
    """Single-phase incompressible pressure system on a Cartesian grid."""
    name: str
    dim: int
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    K6: np.ndarray
    sources: dict          # cell -> rate (right-hand side entry)
    pinned: list = field(default_factory=list)   # cells locked by zeroRows
    seed: int = 0
    # accumulation term dR/dp/dt added to each diagonal after the flux part,
    # as Discretisation.assemblePressureJacobianDiagonal does
    # (rs/src/no/uib/cipr/rs/numerics/Discretisation.java:365-376); 0 for
    # the incompressible (potential) systems
    accumulation: float = 0.0

    @property
    def n(self) -> int:
        return self.nx * self.ny * (self.nz if self.dim == 3 else 1)

    @property
    def centre(self) -> int:
        k = (self.nz // 2) if self.dim == 3 else 0
        return (self.nx // 2) + (self.ny // 2) * self.nx + k * self.nx * self.ny

    def rhs(self) -> np.ndarray:
        """Sources, with locked cells set to zero as rs LinearSolver.solve
        does after Matrices.zeroRows (LinearSolver.java:256-260)."""
        b = np.zeros(self.n)
        for c, q in self.sources.items():
            b[c] = q
        for p in self.pinned:
            b[p] = 0.0
        return b


def _pm1(n: int, dim: int, nx: int, ny: int, nz: int):
    """+1 at the first cell, -1 at the opposite corner; the pure-Neumann
    system is made nonsingular by locking the centre cell (p = 0)."""
    k = nz // 2 if dim == 3 else 0
    centre = nx // 2 + (ny // 2) * nx + k * nx * ny
    return {0: 1.0, n - 1: -1.0}, [centre]


def config1(seed: int = 1, nx: int = 64, ny: int = 64) -> PressureProblem:
    """2-D 64x64, dx = dy = 1/64, k_xx, k_yy = exp(g) independent (sigma 1),
    k_xy = 0.1 sqrt(k_xx k_yy); +1 / -1 in opposite corners, centre locked."""
    n = nx * ny
    g = normal(seed, 0, 2 * n)
    kxx, kyy = np.exp(g[:n]), np.exp(g[n:])
    K = full_tensor(kxx, kyy, np.ones(n))
    K[:, 4] = 0.0
    K[:, 5] = 0.0
    src, pin = _pm1(n, 2, nx, ny, 0)
    return PressureProblem("C1-2d-%dx%d" % (nx, ny), 2, nx, ny, 0, 1.0 / nx, 1.0 / ny, -1.0,
                           K, sources=src, pinned=pin, seed=seed)


def config2(seed: int = 2, nx: int = 128, ny: int = 128, nz: int = 64,
            sigma: float = 2.0, kv_ratio: float = 0.1,
            accumulation: float = 1e-3) -> PressureProblem:
    """3-D 128x128x64 pressure Newton Jacobian: MPFA-O flux part with
    k_h = exp(sigma g), k_v = 0.1 k_h, off-diagonals 0.1 sqrt(k_ii k_jj) on
    unit cells (dz = -1: depth grows with k), plus the accumulation diagonal
    (pore volume x compressibility / dt = 1e-3 of the median
    transmissibility); +1 / -1 sources in opposite corners, no locked cell."""
    n = nx * ny * nz
    kh = np.exp(sigma * normal(seed, 0, n))
    K = full_tensor(kh, kh, kv_ratio * kh)
    return PressureProblem("C2-3d-%dx%dx%d" % (nx, ny, nz), 3, nx, ny, nz, 1.0, 1.0, -1.0, K,
                           sources={0: 1.0, n - 1: -1.0}, pinned=[], seed=seed,
                           accumulation=accumulation)


def config4(seed: int = 4, count: int = 64, emin: int = 32, emax: int = 96,
            sigma: float = 1.0, accumulation: float = 1e-3) -> list:
    """Ensemble of ``count`` single-phase pressure systems on cubic grids,
    edge drawn uniformly from {emin..emax}; system k has its own log-normal
    permeability (sigma 1, normal stream k+1), isotropic with the
    0.1 sqrt(k_ii k_jj) off-diagonals; +1 / -1 in opposite corners.  With
    accumulation = 0 the systems are incompressible and the centre cell is
    locked, as the 2-D potential problem; otherwise each is the pressure
    Newton Jacobian with that accumulation diagonal and no locked cell, as
    the 128x128x64 system."""
    u = uniform01(seed, 0, count)
    edges = emin + np.floor(u * (emax - emin + 1)).astype(np.int64)
    out = []
    for k, e in enumerate(edges.tolist()):
        n = e * e * e
        kk = np.exp(sigma * normal(seed, 1 + k, n))
        K = full_tensor(kk, kk, kk)
        src, pin = _pm1(n, 3, e, e, e)
        if accumulation:
            pin = []
        out.append(PressureProblem("C4-%d-%dx%dx%d" % (k, e, e, e), 3, e, e, e, 1.0, 1.0, -1.0,
                                   K, sources=src, pinned=pin, seed=seed,
                                   accumulation=accumulation))
    return out


# Black-oil fluid, rock and grid data of the fully-implicit Jacobian (SI
# units); field names as rsgpu_blackoil_params.  Corey residuals and
# exponents as CoreyRockFluid's Swi, Sor, Sgr, nw, no, ng.
BLACKOIL_DEFAULTS = dict(
    dx=20.0, dy=20.0, dz=4.0, dt=86400.0, g=9.81,
    phi0=0.2, cr=1e-9, pref=200e5, kz_ratio=0.1,
    bw0=1.0 / 1.01, cw=4.5e-10, bo0=1.0 / 1.2, co=1.5e-9, bg0=200.0, rs0=100.0,
    muw=5e-4, muo=2e-3, mug=2e-5,
    rhow_s=1000.0, rhoo_s=800.0, rhog_s=1.0,
    swr=0.05, sor=0.1, sgr=0.05, kw=1.0, ko=1.0, kg=1.0, nw=2, no=2, ng=2, pad=0)


@dataclass
class BlackOilProblem:
    """One Newton step of a fully-implicit black-oil time step: state u =
    (p, Sw, Sg) per cell, previous state u_old, horizontal permeability kh
    per cell (k_z = kz_ratio kh)."""
    name: str
    nx: int
    ny: int
    nz: int
    kh: np.ndarray
    u: np.ndarray
    u_old: np.ndarray
    params: dict

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def nnzb(self) -> int:
        nx, ny, nz = self.nx, self.ny, self.nz
        return nx * ny * nz + 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1))


def config3(seed: int = 3, nx: int = 100, ny: int = 100, nz: int = 50, sigma: float = 1.5,
            **overrides) -> BlackOilProblem:
    """100x100x50 black-oil Newton step: p = 200 bar + the water hydrostatic
    gradient (rho_w,s g z at the cell-centre depth), Sw and Sg uniform in
    [0.1, 0.4] (seeded), log-normal permeability (sigma 1.5, mean log
    100 mD), Corey exponents 2; the previous state equals the current one
    (first Newton iterate of a step), so the residual is the flux
    imbalance."""
    n = nx * ny * nz
    params = dict(BLACKOIL_DEFAULTS, **overrides)
    kh = 1e-13 * np.exp(sigma * normal(seed, 0, n))
    z = ((np.arange(n) // (nx * ny)) + 0.5) * params["dz"]
    u = np.empty(3 * n)
    u[0::3] = params["pref"] + params["rhow_s"] * params["g"] * z
    s = 0.1 + 0.3 * uniform01(seed, 1, 2 * n)
    u[1::3] = s[:n]
    u[2::3] = s[n:]
    return BlackOilProblem("C3-blackoil-%dx%dx%d" % (nx, ny, nz), nx, ny, nz, kh, u, u.copy(),
                           params)


def config5(seed: int = 5, nx: int = 64, ny: int = 64, nz: int = 32,
            sigma: float = 1.0) -> PressureProblem:
    """3-D 64x64x32 with scalar log-permeability (sigma 1) times a fixed
    full-tensor shape; injector in the centre column top, producers in the
    four corner columns at the bottom (pinned)."""
    n = nx * ny * nz
    logk = sigma * normal(seed, 0, n)
    k = np.exp(logk)
    K = full_tensor(k, k, k)
    layer = nx * ny * (nz - 1)
    prods = [layer + 0, layer + nx - 1, layer + nx * (ny - 1), layer + nx * ny - 1]
    inj = (nx // 2) + (ny // 2) * nx
    return PressureProblem("C5-3d-%dx%dx%d" % (nx, ny, nz), 3, nx, ny, nz, 1.0, 1.0, -1.0, K,
                           sources={inj: 1.0}, pinned=prods, seed=seed)


def small3d(seed: int = 7, nx: int = 8, ny: int = 7, nz: int = 5, sigma: float = 1.0):
    """A small 3-D problem for parity tests."""
    n = nx * ny * nz
    kh = np.exp(sigma * normal(seed, 0, n))
    K = full_tensor(kh, kh, 0.3 * kh)
    src, pin = _pm1(n, 3, nx, ny, nz)
    return PressureProblem("small3d", 3, nx, ny, nz, 1.0, 1.0, -1.0, K, sources=src,
                           pinned=pin, seed=seed)


def stencil_nnz(dim: int, nx: int, ny: int, nz: int = 1) -> int:
    """Nonzeros of the full-tensor MPFA-O pattern: prod(3 n_d - 2)."""
    if dim == 2:
        return (3 * nx - 2) * (3 * ny - 2)
    return (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)


# ---------------------------------------------------------------- IMPES state
IMPES_PARAMS = dict(dt=86400.0, g=9.81, omega=0.0, ax=0.0, ay=0.0, thermal=0)


def impes_state(seed: int, nx: int, ny: int, nz: int, dx: float, dy: float, dz: float,
                ncomp: int = 3, thermal: bool = False, absent: float = 0.15) -> dict:
    """Seeded control-volume secondaries of a three-phase, ncomp-component
    IMPES step on a Cartesian grid, as numpy arrays keyed like
    rsgpu_impes_state (per-phase arrays phase-major).  About ``absent`` of
    the (cell, phase) pairs have zero saturation and the phase absent, so
    the density average meets 0/0 and the upstream presence test skips
    phases.  Synthetic values of the right sign and scale (no PVT model)."""
    n = nx * ny * nz
    u = lambda s, m: uniform01(seed, s, m)          # noqa: E731
    i = np.arange(n)
    cx = (i % nx + 0.5) * dx
    cy = ((i // nx) % ny + 0.5) * dy
    cz = (i // (nx * ny) + 0.5) * dz
    p = 2.0e7 + 9.81 * 1000.0 * cz + 2.0e5 * (u(1, n) - 0.5)
    S = u(2, 3 * n).reshape(3, n) + 0.05
    S[u(3, 3 * n).reshape(3, n) < absent] = 0.0
    S[0, S.sum(axis=0) == 0] = 1.0
    S = S / S.sum(axis=0)
    present = (S > 0).astype(np.uint8)
    pc = np.stack([-1.0e4 * u(4, n), np.zeros(n), 5.0e3 * u(5, n)])
    rho = np.stack([1000.0 + 20 * u(6, n), 800.0 + 40 * u(7, n), 120.0 + 30 * u(8, n)])
    mob = np.stack([S[0] ** 2 / 5e-4, S[1] ** 2 / 2e-3, S[2] ** 2 / 2e-5])
    frac = u(9, 3 * ncomp * n).reshape(3, ncomp, n) + 0.01
    frac = frac / frac.sum(axis=1, keepdims=True)
    dens = np.array([55.0e3, 8.0e3, 4.0e3])[:, None, None]          # mol/m^3 per phase
    comp_mob = mob[:, None, :] * frac * dens
    st = dict(
        pressure=p, phase_pressure=p[None, :] + pc, saturation=S, mass_density=rho,
        phase_mobility=mob, present=present, comp_mobility=comp_mob,
        dRdN=(1.0e-5 * (0.5 + u(10, ncomp * n))).reshape(ncomp, n),
        dRdp=-1.0e-9 * (0.5 + u(11, n)) * dx * dy * dz,
        p0=p - 1.0e4 * (u(12, n) - 0.3), R0=1.0e-3 * (u(13, n) - 0.5) * dx * dy * dz,
        cx=cx, cy=cy, cz=cz, mole_fraction=frac,
        molar_density=dens[:, 0, :] * (1 + 0.05 * u(18, 3 * n).reshape(3, n)))
    if thermal:
        st.update(heat_capacity=2.5e6 * (1 + u(14, n)) * dx * dy * dz,
                  dRdT=1.0e-4 * (u(15, n) - 0.3) * dx * dy * dz,
                  energy_mobility=mob * (4.0e6 + 1.0e6 * u(16, 3 * n).reshape(3, n)),
                  temperature=350.0 + 10 * u(17, n),
                  T0=350.0 + 10 * u(17, n) - 0.5 * u(19, n))
    return {k: np.ascontiguousarray(v.reshape(-1)) for k, v in st.items()}


def impes_sources(seed: int, n: int, ncomp: int, count: int = 6):
    """(elem, kind, q [count][3][ncomp], qe): alternating regular and outlet
    terms on seeded cells, with repeats (one cell in two sources)."""
    elem = (uniform01(seed, 20, count) * n).astype(np.int32)
    elem[-1] = elem[0]
    kind = (np.arange(count) % 2).astype(np.int32)
    q = (uniform01(seed, 21, 3 * ncomp * count) - 0.5).reshape(count, 3, ncomp) * 1.0e-2
    q[kind == 0, 1:, :] = 0.0
    qe = (uniform01(seed, 22, count) - 0.5) * 1.0e3
    return elem, kind, q, qe
