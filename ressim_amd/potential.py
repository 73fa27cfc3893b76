"""Fluxes and well rates of a solved potential field on the device.

Mirrors rs/src/no/uib/cipr/rs/numerics/PotentialCalculator.java:
``calculateFlux(mesh, potential)`` (:265-280) returns, per connection, the
sum over its MD transmissibilities (in MD order) of t_k * potential[k].
``rates(cells)`` gives the net outflow of cells (the well rate of a locked
cell, field/Source.java FIXED semantics).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .mtj import Context, _dptr, _hptr, _is_device, _raise, _torch_sync


class FluxOperator:
    """Device flux operator built from MPFA-O transmissibility lists."""

    def __init__(self, h, ctx: Context):
        self.h, self.ctx = h, ctx
        a, b, e = C.c_int(), C.c_int(), C.c_int()
        L.load().rsgpu_flux_info(self.h, C.byref(a), C.byref(b), C.byref(e))
        self.nconn, self.ncells, self.md_entries = a.value, b.value, e.value

    @classmethod
    def from_lists(cls, ncells, md_ptr, md_elem, md_t, here, there,
                   ctx: Optional[Context] = None) -> "FluxOperator":
        ctx = ctx or Context.default()
        arrs = [np.ascontiguousarray(a, t) for a, t in
                ((md_ptr, np.int32), (md_elem, np.int32), (md_t, np.float64),
                 (here, np.int32), (there, np.int32))]
        h = C.c_void_p()
        _raise(L.load().rsgpu_flux_create(ctx.h, len(arrs[0]) - 1, ncells,
                                          *[_hptr(a) for a in arrs], C.byref(h)), ctx.h)
        return cls(h, ctx)

    @classmethod
    def mpfa_o(cls, dim, nx, ny, nz, dx, dy, dz, K6, continuity=1.0,
               ctx: Optional[Context] = None):
        """MPFA-O on the device straight into a flux operator; returns
        (operator, Connection.check failures)."""
        ctx = ctx or Context.default()
        K = np.ascontiguousarray(K6, np.float64)
        h = C.c_void_p()
        fails = C.c_int(0)
        _raise(L.load().rsgpu_mpfa_o_flux(ctx.h, dim, nx, ny, nz, dx, dy, dz, _hptr(K),
                                          continuity, C.byref(h), C.byref(fails)), ctx.h)
        return cls(h, ctx), fails.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_flux_destroy(self.h)
        except Exception:
            pass

    def calculateFlux(self, potential):
        """Per-connection fluxes; torch cuda tensor in -> tensor out, numpy
        in -> numpy out."""
        lib = L.load()
        if _is_device(potential):
            import torch
            out = torch.empty(self.nconn, dtype=torch.float64, device=potential.device)
            _torch_sync(potential)
            _raise(lib.rsgpu_flux_apply(self.h, _dptr(potential), _dptr(out)), self.ctx.h)
            self.ctx.synchronize()
            return out
        import torch
        p = torch.tensor(np.ascontiguousarray(potential, np.float64), device="cuda")
        return self.calculateFlux(p).cpu().numpy()

    def vjp(self, p, w):
        """tbar[q] = (w[here] - w[there]) * p[md_elem[q]] (device tensors)."""
        import torch
        tbar = torch.empty(max(self.md_entries, 1), dtype=torch.float64, device=p.device)
        _torch_sync(p, w)
        _raise(L.load().rsgpu_flux_vjp(self.h, _dptr(p), _dptr(w), _dptr(tbar)), self.ctx.h)
        self.ctx.synchronize()
        return tbar[: self.md_entries]

    def rates(self, flux, cells):
        """Net outflow of ``cells`` from a device flux vector."""
        import torch
        if not _is_device(flux):
            flux = torch.tensor(np.ascontiguousarray(flux, np.float64), device="cuda")
        cells = np.ascontiguousarray(cells, np.int32)
        out = np.zeros(max(cells.size, 1))
        _torch_sync(flux)
        _raise(L.load().rsgpu_flux_rates(self.h, _dptr(flux), cells.size, _hptr(cells),
                                         _hptr(out)), self.ctx.h)
        return out[: cells.size]
