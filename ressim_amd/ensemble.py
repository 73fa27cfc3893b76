"""Batched ensemble solves over the rsgpu C ABI.

The reference solves one pressure system per ``BiCGstab.solve`` call
(mtj/src/no/uib/cipr/matrix/sparse/BiCGstab.java:65-116).  An ensemble of
permeability realisations is solved here in one batch: the systems are the
diagonal blocks of one block-diagonal ``CompRowMatrix``, and each system
keeps its own BiCGstab scalars and ``DefaultIterationMonitor``, so it stops
on its own.  The mixed form runs, per system, FP64 refinement on the true
residual (IR.java:57-70 form) around an FP32 BiCGstab.

    A, sys_start, b = assemble_ensemble(problems)      # device MPFA-O per system
    M = BlockILU(1, precision=FP32); M.setBlocks(grid_blocks(problems, 4096))
    A.sync_f32(); M.setMatrix(A)
    solver = EnsembleSolver(sys_start, DefaultIterationMonitor(1000, 1e-8), mixed=True)
    stats = solver.solve(A, M, b, x)                   # one record per system
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .mtj import (CompRowMatrix, Context, DefaultIterationMonitor,
                  IterativeSolverNotConvergedException, Matrices, NotConvergedReason,
                  Preconditioner, _dptr, _hptr, _is_device, _raise, _torch_sync,
                  mpfa_o_assemble)


def block_diag(mats: Sequence[CompRowMatrix], ctx: Optional[Context] = None):
    """diag(mats[0], ..., mats[-1]) on the device -> (CompRowMatrix, sys_start)."""
    ctx = ctx or mats[0].ctx
    arr = (C.c_void_p * len(mats))(*[m.h.value for m in mats])
    h = C.c_void_p()
    _raise(L.load().rsgpu_csr_block_diag(ctx.h, len(mats), arr, C.byref(h)), ctx.h)
    n = sum(m.n for m in mats)
    ncols = sum(m.m for m in mats)
    nnz = sum(m.nnz for m in mats)
    A = CompRowMatrix._wrap(h, n, ncols, nnz, ctx)
    sys_start = np.zeros(len(mats) + 1, np.int32)
    sys_start[1:] = np.cumsum([m.n for m in mats])
    return A, sys_start


def assemble_ensemble(problems, ctx: Optional[Context] = None):
    """Per problem: device MPFA-O + pattern + flux assembly, the
    accumulation diagonal and the locked rows (Matrices.zeroRows) as for a
    single solve; then the block-diagonal batch.  Returns (A, sys_start, b,
    Connection.check failures)."""
    ctx = ctx or Context.default()
    parts, fails = [], 0
    for pr in problems:
        Ak, f = mpfa_o_assemble(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                                ctx=ctx)
        fails += f
        if pr.accumulation:
            idx = np.arange(pr.n, dtype=np.int32)
            Ak.add(idx, idx, np.full(pr.n, pr.accumulation))
        if pr.pinned:
            Matrices.zeroRows(Ak, 1.0, *pr.pinned)
        parts.append(Ak)
    A, sys_start = block_diag(parts, ctx)
    del parts
    b = np.concatenate([pr.rhs() for pr in problems])
    return A, sys_start, b, fails


def system_blocks(sys_start, rows_per_block: int) -> np.ndarray:
    """Block boundaries that split every system into near-equal contiguous
    blocks of at most ``rows_per_block`` rows (blocks never cross systems).
    rows_per_block <= 0: one block per system (the exact ILU(0) of each)."""
    out = [0]
    for k in range(len(sys_start) - 1):
        a, e = int(sys_start[k]), int(sys_start[k + 1])
        nb = 1 if rows_per_block <= 0 else max(1, -(-(e - a) // rows_per_block))
        out.extend(a + ((e - a) * q) // nb for q in range(1, nb + 1))
    return np.asarray(out, np.int32)


def grid_blocks(problems, max_rows: int) -> np.ndarray:
    """Block boundaries of whole x-lines inside one xy-plane of one system,
    at most ``max_rows`` rows each (at least one line): the planes of each
    grid split into near-equal runs of lines.  Rows are numbered i + j nx +
    k nx ny (NaturalIJKOrdering), so such a block is contiguous and its
    ILU(0) levels are at most about min(lines, nx / 2) rows wide."""
    out = [0]
    base = 0
    for pr in problems:
        nx, ny = pr.nx, pr.ny
        nz = pr.nz if getattr(pr, "dim", 3) == 3 else 1
        lines = max(1, min(ny, max_rows // nx))
        nb = -(-ny // lines)
        cuts = [(ny * q) // nb for q in range(1, nb + 1)]
        for k in range(nz):
            p0 = base + k * nx * ny
            out.extend(p0 + c * nx for c in cuts)
        base += pr.n
    return np.asarray(out, np.int32)


class EnsembleSolver:
    """BiCGstab over every system of a block-diagonal batch at once.

    ``monitor`` is each system's DefaultIterationMonitor (max iterations and
    tolerances).  mixed: FP64 refinement per system around FP32 BiCGstab
    (inner_rtol, inner_max_iter) with an FP32 preconditioner.  check_every:
    iterations per captured HIP graph.  graph=False enqueues the iterations
    directly; autotune=False keeps the default SpMV stage sizes."""

    def __init__(self, sys_start, monitor: Optional[DefaultIterationMonitor] = None,
                 mixed: bool = False, inner_rtol: float = 1e-4, inner_max_iter: int = 200,
                 check_every: int = 4, graph: bool = True, autotune: bool = True,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or Context.default()
        self.sys_start = np.ascontiguousarray(sys_start, np.int32)
        self.nsys = len(self.sys_start) - 1
        self.iter = monitor or DefaultIterationMonitor()
        lib = L.load()
        o = L.SolverOpts()
        lib.rsgpu_solver_opts_default(C.byref(o))
        o.kind = L.BICGSTAB
        o.max_iter, o.rtol, o.atol, o.dtol = (self.iter.maxIter, self.iter.rtol,
                                              self.iter.atol, self.iter.dtol)
        o.mixed = int(mixed)
        o.inner_rtol, o.inner_max_iter = inner_rtol, inner_max_iter
        o.check_every = check_every
        flags = (0 if graph else L.BATCH_NO_GRAPH) | (0 if autotune else L.BATCH_NO_AUTOTUNE)
        h = C.c_void_p()
        _raise(lib.rsgpu_batch_create(self.ctx.h, C.byref(o), self.nsys,
                                      _hptr(self.sys_start), flags, C.byref(h)), self.ctx.h)
        self.h = h
        self.mixed = mixed
        self.last_stats = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_batch_destroy(self.h)
        except Exception:
            pass

    def solve(self, A: CompRowMatrix, M: Optional[Preconditioner], b, x, check=True):
        """Solve every system; x (torch CUDA float64) is the initial guess
        and the result.  Returns the per-system rsgpu_stats records.  With
        check, a system that did not converge raises
        IterativeSolverNotConvergedException (its index in ``.system``)."""
        if not (_is_device(b) and _is_device(x)):
            raise TypeError("batched solves take device vectors")
        if b.numel() != A.n or x.numel() != A.n or A.n != int(self.sys_start[-1]):
            raise ValueError("vector sizes must match the batch")
        st = (L.Stats * self.nsys)()
        _torch_sync(b, x)
        code = L.load().rsgpu_batch_solve(self.h, A.h, M.h if M is not None else None,
                                         _dptr(b), _dptr(x), st)
        self.last_stats = list(st)
        if code in (L.BREAKDOWN, L.DIVERGENCE, L.ITERATIONS):
            if check:
                k = next(i for i, s in enumerate(st) if s.status != L.OK)
                e = IterativeSolverNotConvergedException(NotConvergedReason(code),
                                                         st[k].residual, st[k].iterations)
                e.system = k
                raise e
        else:
            _raise(code, self.ctx.h)
        return self.last_stats

    def info(self):
        a, b, g = C.c_int(), C.c_int(), C.c_int()
        ua, ub = C.c_double(), C.c_double()
        _raise(L.load().rsgpu_batch_info(self.h, C.byref(a), C.byref(b), C.byref(ua),
                                         C.byref(ub), C.byref(g)), self.ctx.h)
        return dict(spmv_stage_fp64=a.value, spmv_stage_fp32=b.value, spmv_us_fp64=ua.value,
                    spmv_us_fp32=ub.value, graph_launches=g.value)
