"""Adjoint gradient of a well-rate mismatch with respect to log-permeability.

The objective (no reference counterpart; rs has no gradients):
    J(m) = sum_w (q_w(m) - q_obs_w)^2,
where K_e = exp(m_e) K0_e per cell, p solves the MPFA-O pressure system
A(m) p = b with the well cells locked (Matrices.zeroRows, as
PotentialCalculator locks FIXED sources), and q_w is the net outflow of
well cell w (rsgpu_flux_rates, from PotentialCalculator.calculateFlux).

Gradient by the adjoint method, all on the device:
    rho = 2 (q - q_obs) on the wells;   g = A_flux' rho  (dJ/dp)
    A' lambda = g   (transposed SpMV + ILU(A).transApply, FP64)
    w = rho - lambda~  (lambda~ = lambda with the locked rows zeroed: their
                        rows of A do not depend on the transmissibilities)
    tbar = d/dt [sum_i w_i outflow_i]   (rsgpu_flux_vjp)
    dJ/dm = MPFA-O vector-Jacobian product of tbar (rsgpu_mpfa_o_vjp).

``WellRateMismatch`` is a forward/backward pair in the manner of a
``torch.autograd.Function``: ``forward(m)`` returns J and keeps what the
backward pass needs, ``backward()`` returns dJ/dm.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from . import mtj
from .mtj import Context, _dptr, _hptr, _raise
from .potential import FluxOperator


class WellRateMismatch:
    def __init__(self, dim, nx, ny, nz, dx, dy, dz, K0, sources: dict, wells: Sequence[int],
                 q_obs=None, rtol: float = 1e-10, nblocks: int = 64,
                 max_refine: int = 6, ctx: Optional[Context] = None):
        self.dim, self.nx, self.ny, self.nz = dim, nx, ny, nz
        self.dx, self.dy, self.dz = dx, dy, dz
        self.K0 = np.ascontiguousarray(K0, np.float64)
        self.n = self.K0.shape[0]
        self.sources = dict(sources)
        self.wells = np.ascontiguousarray(wells, np.int32)
        self.q_obs = None if q_obs is None else np.asarray(q_obs, np.float64)
        self.rtol = rtol
        self.max_refine = max(1, max_refine)
        self.last_refinements = 0
        # block-Jacobi ILU(0) (nblocks = 1: MTJ's ILU; one workgroup sweeps
        # the whole matrix, so large grids want blocks)
        self.nblocks = max(1, min(nblocks, self.n))
        self.ctx = ctx or Context.default()
        self.last_iterations = (0, 0)

    # ------------------------------------------------------------ pieces
    def _rhs(self):
        import torch
        b = np.zeros(self.n)
        for c, v in self.sources.items():
            b[c] = v
        b[self.wells] = 0.0
        return torch.tensor(b, device="cuda")

    def _solve(self, A, M, b, transpose=False):
        """A x = b (A is the transposed matrix and M is applied through
        transApply when ``transpose``), solved to the FP64-attainable
        solution: BiCGstab + (block) ILU(0) to ``rtol`` on the current
        residual, refined on the residual formed in twice the working
        precision (rsgpu_csr_residual_dd) until the correction is within a
        few units in the last place of x or stops shrinking.  The computed
        x is then a smooth function of the matrix to rounding level, which
        is what central differences of J with a step of 1e-6 need (a single
        solve stopped at a tolerance changes its iteration count with m)."""
        import torch
        s = mtj.BiCGstab(b)
        s.setIterationMonitor(mtj.DefaultIterationMonitor(20000, self.rtol, 1e-300, 1e10))
        s.setPreconditioner(M)
        s.prec_transpose = transpose
        x = torch.zeros(self.n, dtype=torch.float64, device="cuda")
        dx = torch.zeros_like(x)
        r = b.clone()
        iters, prev = 0, float("inf")
        for k in range(self.max_refine):
            dx.zero_()
            try:
                s.solve(A, r, dx)
            except mtj.IterativeSolverNotConvergedException:
                if k == 0:
                    raise
                break                      # the residual is at rounding level
            iters += s.last_stats.iterations
            x += dx
            A.residual_dd(b, x, r)
            d = float(dx.abs().max())
            if d <= 4 * np.finfo(np.float64).eps * float(x.abs().max()) or d >= prev:
                break
            prev = d
        self.last_refinements = k + 1
        return x, iters

    def rates(self, m):
        """Well rates q(m) (forward model only)."""
        return self._forward(m)[0]

    def _forward(self, m):
        K6 = np.ascontiguousarray(np.exp(np.asarray(m, np.float64))[:, None] * self.K0)
        g = (self.dim, self.nx, self.ny, self.nz, self.dx, self.dy, self.dz, K6)
        F, fails = FluxOperator.mpfa_o(*g, ctx=self.ctx)
        A, fails2 = mtj.mpfa_o_assemble(*g, ctx=self.ctx)
        At_flux = A.transpose()
        mtj.Matrices.zeroRows(A, 1.0, self.wells)
        M = mtj.ILU(A) if self.nblocks == 1 else mtj.BlockILU(self.nblocks, ctx=self.ctx)
        M.setMatrix(A)
        p, it = self._solve(A, M, self._rhs())
        flux = F.calculateFlux(p)
        q = F.rates(flux, self.wells)
        return q, dict(K6=K6, F=F, A=A, At_flux=At_flux, M=M, p=p, fwd_iters=it)

    # ------------------------------------------------------------ forward / backward
    def forward(self, m) -> float:
        q, self._saved = self._forward(m)
        if self.q_obs is None:
            raise ValueError("q_obs not set")
        self._r = q - self.q_obs
        self.q = q
        return float(self._r @ self._r)

    def backward(self) -> np.ndarray:
        import torch
        sv = self._saved
        rho = torch.zeros(self.n, dtype=torch.float64, device="cuda")
        rho[torch.as_tensor(self.wells, device="cuda", dtype=torch.long)] = torch.as_tensor(
            2.0 * self._r, device="cuda")
        g = torch.zeros_like(rho)
        sv["At_flux"].mult(rho, g)
        At = sv["A"].transpose()
        lam, it = self._solve(At, sv["M"], g, transpose=True)
        self.last_iterations = (sv["fwd_iters"], it)
        wl = torch.as_tensor(self.wells, device="cuda", dtype=torch.long)
        lam[wl] = 0.0
        w = rho - lam
        F = sv["F"]
        tbar = F.vjp(sv["p"], w)
        lib = L.load()
        mbar = torch.empty(self.n, dtype=torch.float64, device="cuda")
        K6 = sv["K6"]
        _raise(lib.rsgpu_mpfa_o_vjp(self.ctx.h, self.dim, self.nx, self.ny, self.nz, self.dx,
                                    self.dy, self.dz, _hptr(K6), 1.0, F.h, _dptr(tbar),
                                    _dptr(mbar)), self.ctx.h)
        self.ctx.synchronize()
        return mbar.cpu().numpy()


def make_function(problem: "WellRateMismatch"):
    """A torch.autograd.Function bound to ``problem``: J = f(m) with
    m a float64 tensor; J.backward() fills m.grad via the adjoint."""
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, m):
            J = problem.forward(m.detach().cpu().numpy())
            return torch.tensor(J, dtype=torch.float64, device=m.device)

        @staticmethod
        def backward(ctx, gout):
            g = torch.as_tensor(problem.backward(), device=gout.device)
            return gout * g

    return _Fn
