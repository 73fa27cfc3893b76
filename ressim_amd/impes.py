"""IMPES pressure Jacobian of rs on the device.

Mirrors the pressure Newton step of
rs/src/no/uib/cipr/rs/numerics/Discretisation.java: the control-surface
secondaries of ``CS.calculateSecondaries`` (field/CS.java:94-170), then
``assemblePressureJacobianFlux`` (:459-561) and
``assemblePressureJacobianDiagonal`` with the regular and outlet source
terms (:365-453), added into A's pattern and b in the order
``LinearSolver.addToJacobian`` / ``addToResidual`` receive them (A and b
start from zero).  The control-volume secondaries (flash, PVT, rock-fluid
results: ``CV`` getters) are inputs: a dict of device tensors keyed by
``IMPES_FIELDS`` (per-phase arrays phase-major, water, oil, gas).

    J = PressureJacobian(flux, ncomp=2)           # flux: FluxOperator (MD lists)
    rho, up, u = J.controlSurfaces(params, state)
    J.setSources(elem, kind, q, qe)                # optional, rs source order
    J.assemble(A, params, state, rho, up, b)       # A: CompRowMatrix, b: tensor
    J.assembleTemperature(A, params, state, rho, up, b)   # thermal runs
    J.molarUpdate(params, state, u, up, N_old, N_new)     # after the solves

``params``: dict with dt, g, omega, ax, ay, thermal.  The temperature
system is TemperatureNewtonRaphson's (:577-752), the molar-mass update
MolarMassExplicit.solve (:768-919).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .mtj import Context, _dptr, _hptr, _raise, _torch_sync
from .potential import FluxOperator

IMPES_FIELDS = L.IMPES_FIELDS
PHASES = ("water", "oil", "gas")


def _params(p: dict) -> L.ImpesParams:
    P = L.ImpesParams()
    P.thermal = int(p.get("thermal", 0))
    for k in ("dt", "g", "omega", "ax", "ay"):
        setattr(P, k, float(p.get(k, 0.0)))
    return P


def _state(s: dict) -> L.ImpesState:
    S = L.ImpesState()
    for f in IMPES_FIELDS:
        a = s.get(f)
        if a is not None:
            want = "uint8" if f == "present" else "float64"
            if str(a.dtype).split(".")[-1] != want or not a.is_cuda or not a.is_contiguous():
                raise TypeError(f"state field {f} must be a contiguous {want} CUDA tensor")
            setattr(S, f, a.data_ptr())
    return S


class PressureJacobian:
    """Device assembly of rs's IMPES pressure Jacobian on a mesh's MD lists
    (``flux``) and, for thermal runs, its Fourier lists (``fourier``)."""

    def __init__(self, flux: FluxOperator, ncomp: int, fourier: Optional[FluxOperator] = None):
        self.flux, self.fourier, self.ncomp = flux, fourier, ncomp
        self.ctx: Context = flux.ctx
        self.n, self.nconn = flux.ncells, flux.nconn
        h = C.c_void_p()
        _raise(L.load().rsgpu_impes_create(flux.h, fourier.h if fourier else None, ncomp,
                                           C.byref(h)), self.ctx.h)
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_impes_destroy(self.h)
        except Exception:
            pass

    def setSources(self, elem=None, kind=None, q=None, qe=None):
        """Source terms in rs's application order; kind 0 regular (q[t][0][:]
        the component rates), 1 outlet (q[t][phase][:], summed over phases)."""
        if elem is None or len(elem) == 0:
            _raise(L.load().rsgpu_impes_set_sources(self.h, 0, None, None, None, None),
                   self.ctx.h)
            return
        e = np.ascontiguousarray(elem, np.int32)
        k = np.ascontiguousarray(kind, np.int32)
        qq = np.ascontiguousarray(q, np.float64).reshape(len(e), 3, self.ncomp)
        qqe = np.ascontiguousarray(np.zeros(len(e)) if qe is None else qe, np.float64)
        _raise(L.load().rsgpu_impes_set_sources(self.h, len(e), _hptr(e), _hptr(k), _hptr(qq),
                                                _hptr(qqe)), self.ctx.h)

    def controlSurfaces(self, params: dict, state: dict, rho=None, upstream=None, darcy=None):
        """CS secondaries: (rho [3][nconn], upstream [3][nconn] int32,
        darcy [3][nconn]) as device tensors."""
        import torch
        dev = state["phase_pressure"].device
        rho = rho if rho is not None else torch.empty(3 * self.nconn, dtype=torch.float64,
                                                      device=dev)
        upstream = upstream if upstream is not None else torch.empty(
            3 * self.nconn, dtype=torch.int32, device=dev)
        darcy = darcy if darcy is not None else torch.empty(3 * self.nconn,
                                                            dtype=torch.float64, device=dev)
        P, S = _params(params), _state(state)
        _torch_sync(rho, upstream)
        _raise(L.load().rsgpu_impes_control_surfaces(self.h, C.byref(P), C.byref(S), _dptr(rho),
                                                     _dptr(upstream), _dptr(darcy)), self.ctx.h)
        self.ctx.synchronize()
        return rho, upstream, darcy

    def assemble(self, A, params: dict, state: dict, rho, upstream, b):
        """A's values and b (device tensor) for one Newton iteration."""
        P, S = _params(params), _state(state)
        _torch_sync(rho, b)
        _raise(L.load().rsgpu_impes_assemble(self.h, A.h, C.byref(P), C.byref(S), _dptr(rho),
                                             _dptr(upstream), _dptr(b)), self.ctx.h)
        return A, b

    def assembleTemperature(self, A, params: dict, state: dict, rho, upstream, b):
        """The temperature system's A and b (TemperatureNewtonRaphson):
        advective and Fourier fluxes, diagonal part, energy sources."""
        P, S = _params(params), _state(state)
        _torch_sync(rho, b)
        _raise(L.load().rsgpu_impes_assemble_temperature(
            self.h, A.h, C.byref(P), C.byref(S), _dptr(rho), _dptr(upstream), _dptr(b)),
            self.ctx.h)
        return A, b

    def molarUpdate(self, params: dict, state: dict, darcy, upstream, N_old, N_new,
                    locked=(), truncate=False):
        """MolarMassExplicit.solve: N_new = N_old + dN ([ncomp][n] device
        tensors); returns the number of truncated negative mole numbers
        (truncate=True) or raises on the first negative one."""
        P, S = _params(params), _state(state)
        lk = np.ascontiguousarray(np.asarray(locked, np.int32).reshape(-1))
        tr, bad = C.c_int(0), C.c_int(-1)
        _torch_sync(N_old, N_new)
        code = L.load().rsgpu_impes_molar_update(
            self.h, C.byref(P), C.byref(S), _dptr(darcy), _dptr(upstream), lk.size,
            _hptr(lk) if lk.size else None, int(truncate), _dptr(N_old), _dptr(N_new),
            C.byref(tr), C.byref(bad))
        self.bad = bad.value   # first element * ncomp + component with N < 0, or -1
        _raise(code, self.ctx.h)
        return tr.value
