"""Fully-implicit black-oil Newton Jacobian on the device (3x3 blocks).

The reference simulator is IMPES and has no fully-implicit Jacobian
(rs/src/no/uib/cipr/rs/numerics/Discretisation.java:208-225); this module
assembles the Jacobian of a three-phase black-oil time step with dissolved
gas (Corey relative permeabilities in the form of
rs/src/no/uib/cipr/rs/rock/CoreyRockFluid.java:129-152) straight into a
3x3 block ``CompRowMatrix`` that every solver and preconditioner accepts:

    bo = BlackOilJacobian(problem)            # problems.config3()
    J = bo.matrix()
    u = torch.tensor(problem.u, device="cuda"); R = torch.empty_like(u)
    bo.assemble(u, u_old, J, R)               # J = dR/du, R(u)
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .mtj import CompRowMatrix, Context, _dptr, _hptr, _raise, _torch_sync


def params_struct(params: dict) -> L.BlackOilParams:
    p = L.BlackOilParams()
    for k, v in params.items():
        setattr(p, k, v)
    return p


class BlackOilJacobian:
    def __init__(self, problem, ctx: Optional[Context] = None):
        self.ctx = ctx or Context.default()
        self.problem = problem
        self.P = params_struct(problem.params)
        kh = np.ascontiguousarray(problem.kh, np.float64)
        h = C.c_void_p()
        _raise(L.load().rsgpu_blackoil_create(self.ctx.h, C.byref(self.P), problem.nx,
                                              problem.ny, problem.nz, _hptr(kh), C.byref(h)),
               self.ctx.h)
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_blackoil_destroy(self.h)
        except Exception:
            pass

    def matrix(self) -> CompRowMatrix:
        h = C.c_void_p()
        _raise(L.load().rsgpu_blackoil_matrix(self.h, C.byref(h)), self.ctx.h)
        n = 3 * self.problem.n
        return CompRowMatrix._wrap(h, n, n, 9 * self.problem.nnzb, self.ctx)

    def assemble(self, u, u_old, J: Optional[CompRowMatrix] = None, R=None):
        """J = dR/du and R(u) (torch CUDA float64 vectors of 3 n)."""
        _torch_sync(u, u_old)
        _raise(L.load().rsgpu_blackoil_assemble(self.h, _dptr(u), _dptr(u_old),
                                                J.h if J is not None else None,
                                                _dptr(R) if R is not None else None),
               self.ctx.h)
        self.ctx.synchronize()
