"""Host-side mirror of MTJ's sparse solver interfaces over the rsgpu C ABI.

Names, argument meaning and error behaviour follow
mtj/src/no/uib/cipr/matrix/sparse/{IterativeSolver,Preconditioner,
CompRowMatrix,ILU,BiCGstab,GMRES,IR,DefaultIterationMonitor,
IterativeSolverNotConvergedException}.java and
mtj/src/no/uib/cipr/matrix/Matrices.java (zeroRows), so code written against
MTJ reads the same:

    A = CompRowMatrix(n, n, nz)            # int[][] nz, rows sorted as MTJ does
    A.set_data(values)                     # or A.add(i, j, v) batches
    M = ILU(A)                             # Preconditioner
    solver = BiCGstab(b)                   # template vector
    solver.setPreconditioner(M)
    solver.setIterationMonitor(DefaultIterationMonitor(5000, 1e-10, 1e-50, 1e5))
    M.setMatrix(A)
    solver.solve(A, b, x)                  # x: initial guess and result

Vectors are numpy float64 arrays (host; copied through the *_host entry
points, as a JNI binding would pass Java double[]) or torch CUDA float64
tensors (device; passed by pointer).  Every computation runs in librsgpu.so
on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


class NotConvergedReason(enum.Enum):
    """NotConvergedException.Reason (mtj/.../NotConvergedException.java:38-48)."""
    Iterations = L.ITERATIONS
    Divergence = L.DIVERGENCE
    Breakdown = L.BREAKDOWN


class IterativeSolverNotConvergedException(Exception):
    """mtj/.../sparse/IterativeSolverNotConvergedException.java:54-89."""

    def __init__(self, reason: NotConvergedReason, residual: float, iterations: int,
                 message: str = ""):
        super().__init__(message or f"{reason.name} (residual {residual:g}, "
                                    f"{iterations} iterations)")
        self.reason = reason
        self.residual = residual
        self.iterations = iterations

    def getReason(self):
        return self.reason

    def getResidual(self):
        return self.residual

    def getIterations(self):
        return self.iterations


class MatrixSingularException(RuntimeError):
    """mtj/.../MatrixSingularException.java (GMRES back-solve)."""


class RsgpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rsgpu status {code}: {msg}")
        self.code = code


def _raise(code: int, ctx_handle, what: str = ""):
    if code == L.OK:
        return
    msg = L.load().rsgpu_last_error(ctx_handle).decode(errors="replace") if ctx_handle else ""
    msg = msg or what
    if code == L.NOT_IN_PATTERN:
        raise IndexError(msg)
    if code in (L.ZERO_PIVOT, L.MISSING_DIAGONAL):
        raise RuntimeError(msg)
    if code == L.BAD_ARGUMENT:
        raise ValueError(msg)
    if code == L.SINGULAR:
        raise MatrixSingularException(msg)
    raise RsgpuError(code, msg)


# ------------------------------------------------------------------ context
class Context:
    """One HIP device + stream.  ``Context.default()`` is shared."""
    _default = {}

    def __init__(self, device: int = 0):
        lib = L.load()
        h = C.c_void_p()
        _raise(lib.rsgpu_ctx_create(device, C.byref(h)), None, "ctx_create failed")
        self.h = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def synchronize(self):
        _raise(L.load().rsgpu_ctx_synchronize(self.h), self.h)

    def set_stream(self, stream_handle: Optional[int]):
        _raise(L.load().rsgpu_ctx_set_stream(self.h, stream_handle), self.h)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_ctx_destroy(self.h)
        except Exception:
            pass


def _is_device(v) -> bool:
    return hasattr(v, "is_cuda") and v.is_cuda


def _dptr(v):
    return C.c_void_p(v.data_ptr())


def _hptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _torch_sync(*vs):
    for v in vs:
        if _is_device(v):
            import torch
            torch.cuda.synchronize(v.device)
            return


# ------------------------------------------------------------------ CompRowMatrix
class CompRowMatrix:
    """mtj/.../sparse/CompRowMatrix.java on the GPU.

    ``CompRowMatrix(n, m, nz)`` with nz a list of per-row column lists
    (:160-190), or ``CompRowMatrix.from_csr(n, m, rowptr, colind)``.
    """

    def __init__(self, numRows: int, numColumns: int, nz: Sequence[Sequence[int]],
                 ctx: Optional[Context] = None):
        if len(nz) != numRows:
            raise ValueError("nz.length != numRows")
        rowptr = np.zeros(numRows + 1, np.int32)
        rowptr[1:] = np.cumsum([len(r) for r in nz])
        colind = (np.concatenate([np.asarray(r, np.int32) for r in nz])
                  if rowptr[-1] else np.zeros(0, np.int32))
        self._create(numRows, numColumns, rowptr, colind, ctx)

    @classmethod
    def from_csr(cls, numRows, numColumns, rowptr, colind, data=None,
                 ctx: Optional[Context] = None) -> "CompRowMatrix":
        self = cls.__new__(cls)
        self._create(numRows, numColumns, np.ascontiguousarray(rowptr, np.int32),
                     np.ascontiguousarray(colind, np.int32), ctx)
        if data is not None:
            self.set_data(data)
        return self

    @classmethod
    def from_bsr3(cls, nb, browptr, bcolind, data=None,
                  ctx: Optional[Context] = None) -> "CompRowMatrix":
        """A matrix of 3x3 blocks on the block pattern (browptr, bcolind): the
        expanded scalar matrix with each block row's three rows interleaved
        (rsgpu_csr_create_bsr3); getData() is in that layout."""
        self = cls.__new__(cls)
        self.ctx = ctx or Context.default()
        brp = np.ascontiguousarray(browptr, np.int32)
        bci = np.ascontiguousarray(bcolind, np.int32)
        if brp.ndim != 1 or brp.size != nb + 1:
            raise ValueError(f"browptr must hold nb + 1 = {nb + 1} entries, got {brp.size}")
        if bci.size < int(brp[-1]):
            raise ValueError(f"bcolind holds {bci.size} entries, browptr[-1] = {int(brp[-1])}")
        h = C.c_void_p()
        _raise(L.load().rsgpu_csr_create_bsr3(self.ctx.h, nb, int(brp[-1]), _hptr(brp),
                                              _hptr(bci) if bci.size else None, C.byref(h)),
               self.ctx.h)
        self.h = h
        self.n = self.m = 3 * nb
        self.nnz = 9 * int(brp[-1])
        if data is not None:
            self.set_data(data)
        return self

    @classmethod
    def from_matrix_market(cls, src, ctx: Optional[Context] = None) -> "CompRowMatrix":
        """CompRowMatrix(MatrixVectorReader) (CompRowMatrix.java:68-148)."""
        from .mmio import read_matrix
        n, m, rp, ci, d = read_matrix(src)
        return cls.from_csr(n, m, rp, ci, d, ctx=ctx)

    def to_matrix_market(self, digits: int = 12) -> str:
        """The matrix in Matrix Market coordinate form (AbstractMatrix.toString
        with MatrixVectorWriter's header)."""
        from .mmio import write_matrix
        return write_matrix(self.getRowPointers(), self.getColumnIndices(), self.getData(),
                            self.n, self.m, digits=digits)

    def block_info(self):
        """(block size, block rows, block nonzeros)."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _raise(L.load().rsgpu_csr_block_info(self.h, C.byref(a), C.byref(b), C.byref(c)),
               self.ctx.h)
        return a.value, b.value, c.value

    @classmethod
    def _wrap(cls, handle, n, m, nnz, ctx) -> "CompRowMatrix":
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = handle
        self.n, self.m, self.nnz = n, m, nnz
        return self

    def _create(self, n, m, rowptr, colind, ctx):
        self.ctx = ctx or Context.default()
        lib = L.load()
        h = C.c_void_p()
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        colind = np.ascontiguousarray(colind, np.int32)
        if rowptr.ndim != 1 or rowptr.size != n + 1:
            raise ValueError(f"rowptr must hold n + 1 = {n + 1} entries, got {rowptr.size}")
        nnz = int(rowptr[-1])
        if colind.size < nnz:
            raise ValueError(f"colind holds {colind.size} entries, rowptr[-1] = {nnz}")
        _raise(lib.rsgpu_csr_create(self.ctx.h, n, m, nnz, _hptr(rowptr),
                                    _hptr(colind) if nnz else None, C.byref(h)),
               self.ctx.h)
        self.h = h
        self.n, self.m, self.nnz = n, m, nnz

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_csr_destroy(self.h)
        except Exception:
            pass

    def numRows(self):
        return self.n

    def numColumns(self):
        return self.m

    def isSquare(self):
        return self.n == self.m

    def getRowPointers(self) -> np.ndarray:
        rp = np.zeros(self.n + 1, np.int32)
        L.load().rsgpu_csr_get_pattern(self.h, _hptr(rp), None)
        return rp

    def getColumnIndices(self) -> np.ndarray:
        ci = np.zeros(max(self.nnz, 1), np.int32)
        L.load().rsgpu_csr_get_pattern(self.h, None, _hptr(ci))
        return ci[: self.nnz]

    def getData(self) -> np.ndarray:
        d = np.zeros(max(self.nnz, 1))
        _raise(L.load().rsgpu_csr_get_values(self.h, _hptr(d)), self.ctx.h)
        return d[: self.nnz]

    def set_data(self, data):
        """Replace the values (getData() array) wholesale."""
        if _is_device(data):
            _torch_sync(data)
            _raise(L.load().rsgpu_csr_set_values_dev(self.h, _dptr(data)), self.ctx.h)
            self.ctx.synchronize()
            return
        d = np.ascontiguousarray(data, np.float64)
        if d.size != self.nnz:
            raise ValueError("data length != nnz")
        _raise(L.load().rsgpu_csr_set_values(self.h, _hptr(d)), self.ctx.h)

    def zero(self):
        _raise(L.load().rsgpu_csr_zero(self.h), self.ctx.h)
        return self

    def add(self, row, column, value):
        """add(row, column, value); arrays apply a batch in array order."""
        r = np.atleast_1d(np.asarray(row, np.int32))
        c = np.atleast_1d(np.asarray(column, np.int32))
        v = np.atleast_1d(np.asarray(value, np.float64))
        _raise(L.load().rsgpu_csr_add_batch(self.h, r.size, _hptr(r), _hptr(c), _hptr(v)),
               self.ctx.h)

    def transpose(self) -> "CompRowMatrix":
        """A new matrix holding A' (values copied now)."""
        h = C.c_void_p()
        _raise(L.load().rsgpu_csr_transpose(self.h, C.byref(h)), self.ctx.h)
        return CompRowMatrix._wrap(h, self.m, self.n, self.nnz, self.ctx)

    def sync_f32(self):
        _raise(L.load().rsgpu_csr_sync_f32(self.h), self.ctx.h)

    def sliced_info(self):
        """(chunks, value slots, tables) of the sliced copy; chunks 0 when the
        pattern has none (a 64-row chunk spans more than 32 column offsets)."""
        a, b, c = C.c_int(), C.c_int64(), C.c_int()
        _raise(L.load().rsgpu_csr_sliced_info(self.h, C.byref(a), C.byref(b), C.byref(c)),
               self.ctx.h)
        return a.value, b.value, c.value

    def sync_sliced(self):
        """Refresh the FP64 sliced copy from the current values."""
        _raise(L.load().rsgpu_csr_sync_sliced(self.h), self.ctx.h)

    def mult_sliced(self, x, y):
        """y = A x over the sliced copy (device tensors, FP64 after
        sync_sliced or FP32 after sync_f32)."""
        _torch_sync(x, y)
        f = (L.load().rsgpu_csr_mult_sliced_f32 if str(x.dtype) == "torch.float32"
             else L.load().rsgpu_csr_mult_sliced)
        _raise(f(self.h, _dptr(x), _dptr(y)), self.ctx.h)
        self.ctx.synchronize()
        return y

    def mult(self, x, y):
        """y = A x (CompRowMatrix.java:301-333)."""
        lib = L.load()
        if _is_device(x):
            _torch_sync(x, y)
            _raise(lib.rsgpu_csr_mult(self.h, _dptr(x), _dptr(y)), self.ctx.h)
            self.ctx.synchronize()
            return y
        xx = np.ascontiguousarray(x, np.float64)
        _raise(lib.rsgpu_csr_mult_host(self.h, _hptr(xx), _hptr(y)), self.ctx.h)
        return y

    def multAdd(self, alpha, x, y):
        """y += alpha A x (device tensors)."""
        _torch_sync(x, y)
        _raise(L.load().rsgpu_csr_multadd(self.h, alpha, _dptr(x), _dptr(y)), self.ctx.h)
        self.ctx.synchronize()
        return y

    def transMult(self, x, y):
        lib = L.load()
        if _is_device(x):
            _torch_sync(x, y)
            _raise(lib.rsgpu_csr_transmult(self.h, _dptr(x), _dptr(y)), self.ctx.h)
            self.ctx.synchronize()
            return y
        xx = np.ascontiguousarray(x, np.float64)
        _raise(lib.rsgpu_csr_transmult_host(self.h, _hptr(xx), _hptr(y)), self.ctx.h)
        return y

    def transMultAdd(self, alpha, x, y):
        _torch_sync(x, y)
        _raise(L.load().rsgpu_csr_transmultadd(self.h, alpha, _dptr(x), _dptr(y)),
               self.ctx.h)
        self.ctx.synchronize()
        return y

    def residual_dd(self, b, x, r):
        """r = b - A x in twice the working precision, rounded once (device
        tensors; rsgpu_csr_residual_dd)."""
        _torch_sync(b, x, r)
        _raise(L.load().rsgpu_csr_residual_dd(self.h, _dptr(b), _dptr(x), _dptr(r)),
               self.ctx.h)
        self.ctx.synchronize()
        return r


class Matrices:
    @staticmethod
    def zeroRows(A: CompRowMatrix, diagonal: float, *rows):
        """mtj/.../Matrices.java:631-652."""
        if len(rows) == 1 and not np.isscalar(rows[0]):
            rows = rows[0]
        r = np.ascontiguousarray(np.asarray(rows, np.int32).reshape(-1))
        _raise(L.load().rsgpu_csr_zero_rows(A.h, diagonal, _hptr(r) if r.size else None,
                                            r.size), A.ctx.h)


# ------------------------------------------------------------------ preconditioners
class Preconditioner:
    """mtj/.../sparse/Preconditioner.java:42-63."""
    kind = L.PREC_IDENTITY

    def __init__(self, nblocks: int = 1, precision: int = L.FP64,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or Context.default()
        h = C.c_void_p()
        _raise(L.load().rsgpu_prec_create(self.ctx.h, self.kind, nblocks, precision,
                                          C.byref(h)), self.ctx.h)
        self.h = h
        self.precision = precision
        self.n = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                L.load().rsgpu_prec_destroy(self.h)
        except Exception:
            pass

    def setMatrix(self, A: CompRowMatrix):
        bad = C.c_int(-1)
        code = L.load().rsgpu_prec_set_matrix(self.h, A.h, C.byref(bad))
        self.bad_row = bad.value
        _raise(code, self.ctx.h)
        self.n = A.n

    def apply(self, b, x):
        lib = L.load()
        if _is_device(b):
            _torch_sync(b, x)
            _raise(lib.rsgpu_prec_apply(self.h, _dptr(b), _dptr(x)), self.ctx.h)
            self.ctx.synchronize()
            return x
        bb = np.ascontiguousarray(b, np.float64)
        _raise(lib.rsgpu_prec_apply_host(self.h, _hptr(bb), _hptr(x)), self.ctx.h)
        return x

    def transApply(self, b, x):
        lib = L.load()
        if _is_device(b):
            _torch_sync(b, x)
            _raise(lib.rsgpu_prec_trans_apply(self.h, _dptr(b), _dptr(x)), self.ctx.h)
            self.ctx.synchronize()
            return x
        bb = np.ascontiguousarray(b, np.float64)
        _raise(lib.rsgpu_prec_trans_apply_host(self.h, _hptr(bb), _hptr(x)), self.ctx.h)
        return x

    def factors(self):
        """(rowptr, colind, lu) of the factored matrix (ILU's internal LU)."""
        nnz = C.c_int(0)
        lib = L.load()
        _raise(lib.rsgpu_prec_get_factors(self.h, None, None, None, C.byref(nnz)), self.ctx.h)
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(nnz.value, 1), np.int32)
        lu = np.zeros(max(nnz.value, 1))
        _raise(lib.rsgpu_prec_get_factors(self.h, _hptr(rp), _hptr(ci), _hptr(lu),
                                          C.byref(nnz)), self.ctx.h)
        return rp, ci[: nnz.value], lu[: nnz.value]

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        L.load().rsgpu_prec_info(self.h, C.byref(a), C.byref(b), C.byref(c))
        f = C.c_int(0)
        L.load().rsgpu_prec_sweep_form(self.h, C.byref(f))
        return dict(nnz=a.value, levels_lower=b.value, levels_upper=c.value,
                    sweep_form=("levels", "lines", "cells", "cells_streamed")[f.value])


class IdentityPreconditioner(Preconditioner):
    kind = L.PREC_IDENTITY


class ILU(Preconditioner):
    """ILU(0) (mtj/.../sparse/ILU.java).  ``ILU(A)`` accepts the matrix the
    reference passes as the factor storage; only its pattern matters here."""
    kind = L.PREC_ILU0

    def __init__(self, A: Optional[CompRowMatrix] = None, precision: int = L.FP64,
                 ctx: Optional[Context] = None):
        if A is not None and not A.isSquare():
            raise ValueError("ILU only applies to square matrices")
        super().__init__(1, precision, ctx or (A.ctx if A is not None else None))


class BlockILU(Preconditioner):
    """Block-Jacobi ILU(0): BlockDiagonalPreconditioner(new ILU(...)) over
    ``nblocks`` contiguous row ranges (distributed/BlockDiagonalPreconditioner
    .java:46-75 with DistRowMatrix ownership)."""
    kind = L.PREC_BLOCK_ILU0

    def __init__(self, nblocks: int, precision: int = L.FP64,
                 ctx: Optional[Context] = None):
        super().__init__(nblocks, precision, ctx)

    def setBlocks(self, bstart):
        """Caller-given block boundaries (0 = bstart[0] < ... < bstart[-1] =
        n), as a DistRowMatrix whose ranks own unequal row ranges; takes
        effect at the next setMatrix."""
        bs = np.ascontiguousarray(bstart, np.int32)
        _raise(L.load().rsgpu_prec_set_blocks(self.h, len(bs) - 1, _hptr(bs)), self.ctx.h)


class DiagonalPreconditioner(Preconditioner):
    """mtj/.../sparse/DiagonalPreconditioner.java: x = b / diag(A)."""
    kind = L.PREC_DIAGONAL

    def __init__(self, n: Optional[int] = None, precision: int = L.FP64,
                 ctx: Optional[Context] = None):
        super().__init__(1, precision, ctx)
        self.size = n

    def setMatrix(self, A: CompRowMatrix):
        if self.size is not None and A.n != self.size:
            raise ValueError("Matrix size differs from preconditioner size")
        super().setMatrix(A)


class ICC(Preconditioner):
    """mtj/.../sparse/ICC.java: incomplete Cholesky on A's pattern (FP64);
    apply = transApply = R' then R solves.  ``ICC(R)`` accepts the matrix the
    reference passes as the factor storage; only its pattern matters here."""
    kind = L.PREC_ICC

    def __init__(self, R: Optional[CompRowMatrix] = None, ctx: Optional[Context] = None):
        if R is not None and not R.isSquare():
            raise ValueError("ICC only applies to square matrices")
        super().__init__(1, L.FP64, ctx or (R.ctx if R is not None else None))


class ILUT(Preconditioner):
    """mtj/.../sparse/ILUT.java: ILU with the dual-threshold dropping of
    Saad, ``ILUT(LU, tau, p)`` (defaults tau = 1e-6, p = 25, ILUT.java:116-118),
    factored on the device at setMatrix; FP64.  ``ILUT(LU)`` accepts the
    matrix the reference passes as its work storage; only its shape is
    checked."""
    kind = L.PREC_ILUT

    def __init__(self, LU: Optional[CompRowMatrix] = None, tau: float = 1e-6, p: int = 25,
                 ctx: Optional[Context] = None):
        if LU is not None and not LU.isSquare():
            raise ValueError("ILU only applies to square matrices")
        super().__init__(1, L.FP64, ctx or (LU.ctx if LU is not None else None))
        self.tau, self.p = float(tau), int(p)
        _raise(L.load().rsgpu_prec_set_ilut(self.h, self.tau, self.p), self.ctx.h)


class SSOR(Preconditioner):
    """mtj/.../sparse/SSOR.java: SSOR(F, reverse, omegaF, omegaR); ``SSOR()``
    is SSOR(F) = reverse sweep, omega 1.  apply(b, x) starts from x's old
    content.  block_rows / ordering: as the AMG smoother (0, 0 = SSOR.java)."""
    kind = L.PREC_SSOR

    def __init__(self, F: Optional[CompRowMatrix] = None, reverse: bool = True,
                 omegaF: float = 1.0, omegaR: float = 1.0, block_rows: int = 0,
                 ordering: int = 0, ctx: Optional[Context] = None):
        if F is not None and not F.isSquare():
            raise ValueError("SSOR only applies to square matrices")
        super().__init__(1, L.FP64, ctx or (F.ctx if F is not None else None))
        self.setOmega(omegaF, omegaR, reverse, block_rows, ordering)

    def setOmega(self, omegaF, omegaR, reverse=True, block_rows=0, ordering=0):
        _raise(L.load().rsgpu_prec_set_ssor(self.h, omegaF, omegaR, int(bool(reverse)),
                                            block_rows, ordering), self.ctx.h)


class AMG(Preconditioner):
    """Smoothed-aggregation AMG (mtj/.../sparse/AMG.java).  The positional
    arguments are AMG(omegaPreF, omegaPreR, omegaPostF, omegaPostR, nu1, nu2,
    gamma, min, omega) (AMG.java:140-156); ``AMG()`` uses the defaults of
    AMG.java:204-206.  Keywords: block_rows (SSOR smoother blocks, 0 = one
    block = SSOR.java), zero_guess (start each cycle from zero instead of x's
    old content), max_coarse.  Aggregation runs as prioritised parallel
    rounds on the device (phases 1 and 3); see the header of amg.hip."""
    kind = L.PREC_AMG
    _ORDER = ("omega_pre_f", "omega_pre_r", "omega_post_f", "omega_post_r", "nu1", "nu2",
              "gamma", "min", "omega")

    def __init__(self, *args, ctx: Optional[Context] = None, precision: int = L.FP64, **kw):
        super().__init__(1, precision, ctx)
        p = L.AmgParams()
        L.load().rsgpu_amg_params_default(C.byref(p))
        if len(args) == 7:
            # AMG(omegaPre, omegaPost, nu1, nu2, gamma, min, omega): SOR smoothing
            # (AMG.java:179-194)
            args = (args[0], args[0], args[1], args[1]) + tuple(args[2:])
            kw.setdefault("reverse", 0)
        elif args and len(args) != 9:
            raise TypeError("AMG takes 0, 7 or 9 positional arguments")
        for name, v in zip(self._ORDER, args):
            setattr(p, name, v)
        for name, v in kw.items():
            if not hasattr(p, name):
                raise TypeError(f"unknown AMG parameter {name}")
            setattr(p, name, v)
        self.params = p
        _raise(L.load().rsgpu_prec_set_amg_params(self.h, C.byref(p)), self.ctx.h)

    def levels(self):
        sz = np.zeros(64, np.int32)
        nz = np.zeros(64, np.int32)
        m = L.load().rsgpu_amg_levels(self.h, _hptr(sz), _hptr(nz), 64)
        return [(int(sz[k]), int(nz[k])) for k in range(m)]

    def level(self, k):
        """(rowptr, colind, data, aggregate, P rowptr, P colind, P data) of
        level k; the last four are None on the coarsest level."""
        lv = self.levels()
        n, nnz = lv[k]
        lib = L.load()
        pn = C.c_int(0)
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        _raise(lib.rsgpu_amg_get_level(self.h, k, _hptr(rp), _hptr(ci), _hptr(v), None, None,
                                       None, None, C.byref(pn)), self.ctx.h)
        if k == len(lv) - 1:
            return rp, ci[:nnz], v[:nnz], None, None, None, None
        pt = np.zeros(max(n, 1), np.int32)
        prp = np.zeros(n + 1, np.int32)
        pci = np.zeros(max(pn.value, 1), np.int32)
        pv = np.zeros(max(pn.value, 1))
        _raise(lib.rsgpu_amg_get_level(self.h, k, None, None, None, _hptr(pt), _hptr(prp),
                                       _hptr(pci), _hptr(pv), C.byref(pn)), self.ctx.h)
        return rp, ci[:nnz], v[:nnz], pt[:n], prp, pci[: pn.value], pv[: pn.value]


# ------------------------------------------------------------------ monitor and solvers
class DefaultIterationMonitor:
    """mtj/.../sparse/DefaultIterationMonitor.java (defaults :85-90)."""

    def __init__(self, maxIter: int = 100000, rtol: float = 1e-5, atol: float = 1e-50,
                 dtol: float = 1e5):
        self.maxIter, self.rtol, self.atol, self.dtol = maxIter, rtol, atol, dtol
        self._iterations = 0
        self._residual = 0.0
        self.history = np.zeros(0)

    def setMaxIterations(self, v):
        self.maxIter = v

    def setRelativeTolerance(self, v):
        self.rtol = v

    def setAbsoluteTolerance(self, v):
        self.atol = v

    def setDivergenceTolerance(self, v):
        self.dtol = v

    def iterations(self):
        return self._iterations

    def residual(self):
        return self._residual


class AbstractIterativeSolver:
    kind = L.BICGSTAB

    def __init__(self, template, restart: int = 30):
        self.n = len(template) if not _is_device(template) else template.numel()
        self.M: Preconditioner = None
        self.iter = DefaultIterationMonitor()
        self.restart = restart
        self.mixed = False
        self.inner_rtol = 1e-4
        self.inner_max_iter = 200
        self.check_every = 4
        # transposed solve: A' x = b with the preconditioner of A applied by
        # transApply (the adjoint's solve)
        self.prec_transpose = False
        # mixed form: classical Gram-Schmidt in the FP32 inner GMRES
        self.cgs = False
        self._h = None
        self._key = None
        self.last_stats = None
        self.fused_steps = 0

    def setPreconditioner(self, M: Preconditioner):
        self.M = M

    def getPreconditioner(self):
        return self.M

    def setIterationMonitor(self, it: DefaultIterationMonitor):
        self.iter = it

    def getIterationMonitor(self):
        return self.iter

    def _handle(self, ctx: Context):
        key = (self.kind, self.restart, self.iter.maxIter, self.iter.rtol, self.iter.atol,
               self.iter.dtol, self.mixed, self.inner_rtol, self.inner_max_iter,
               self.check_every, self.prec_transpose, self.cgs, id(ctx))
        if self._h is not None and key == self._key:
            return self._h
        lib = L.load()
        if self._h is not None:
            lib.rsgpu_solver_destroy(self._h)
            self._h = None
        o = L.SolverOpts()
        lib.rsgpu_solver_opts_default(C.byref(o))
        o.kind, o.restart, o.max_iter = self.kind, self.restart, self.iter.maxIter
        o.rtol, o.atol, o.dtol = self.iter.rtol, self.iter.atol, self.iter.dtol
        o.mixed = int(self.mixed)
        o.inner_rtol, o.inner_max_iter = self.inner_rtol, self.inner_max_iter
        o.check_every = self.check_every
        o.prec_transpose = int(self.prec_transpose)
        o.cgs = int(self.cgs)
        h = C.c_void_p()
        _raise(lib.rsgpu_solver_create(ctx.h, C.byref(o), self.n, C.byref(h)), ctx.h)
        self._h, self._key = h, key
        return h

    def __del__(self):
        try:
            if self._h is not None:
                L.load().rsgpu_solver_destroy(self._h)
        except Exception:
            pass

    def solve(self, A: CompRowMatrix, b, x):
        """solve(A, b, x): x is the initial guess and the result."""
        if not A.isSquare():
            raise ValueError("!A.isSquare()")
        if len(b) != A.n:
            raise ValueError("b.size() != A.numRows()")
        if len(x) != len(b):
            raise ValueError("b.size() != x.size()")
        lib = L.load()
        h = self._handle(A.ctx)
        st = L.Stats()
        Mh = self.M.h if self.M is not None else None
        if _is_device(b):
            _torch_sync(b, x)
            code = lib.rsgpu_solver_solve(h, A.h, Mh, _dptr(b), _dptr(x), C.byref(st))
        else:
            if not (isinstance(x, np.ndarray) and x.dtype == np.float64
                    and x.flags.c_contiguous):
                raise TypeError("x must be a contiguous float64 numpy array")
            bb = np.ascontiguousarray(b, np.float64)
            code = lib.rsgpu_solver_solve_host(h, A.h, Mh, _hptr(bb), _hptr(x), C.byref(st))
        self.last_stats = st
        # GMRES steps that ran as one fused launch (rsgpu_solver_fused_steps)
        self.fused_steps = lib.rsgpu_solver_fused_steps(h)
        self.iter._iterations = st.iterations
        self.iter._residual = st.residual
        hl = max(0, min(st.hist_len, 1 << 22))
        hist = np.zeros(max(hl, 1))
        k = lib.rsgpu_solver_history(h, _hptr(hist), hl)
        self.iter.history = hist[: max(k, 0)]
        if code in (L.BREAKDOWN, L.DIVERGENCE, L.ITERATIONS):
            raise IterativeSolverNotConvergedException(
                NotConvergedReason(code), st.residual, st.iterations)
        _raise(code, A.ctx.h)
        return x


class BiCGstab(AbstractIterativeSolver):
    """mtj/.../sparse/BiCGstab.java."""
    kind = L.BICGSTAB


class CG(AbstractIterativeSolver):
    """mtj/.../sparse/CG.java."""
    kind = L.CG


class BiCG(AbstractIterativeSolver):
    """mtj/.../sparse/BiCG.java (M.apply and M.transApply, A.transMult)."""
    kind = L.BICG


class CGS(AbstractIterativeSolver):
    """mtj/.../sparse/CGS.java."""
    kind = L.CGS


class QMR(AbstractIterativeSolver):
    """mtj/.../sparse/QMR.java with M1 = M2 = the preconditioner
    (QMR.setPreconditioner)."""
    kind = L.QMR


class GMRES(AbstractIterativeSolver):
    """mtj/.../sparse/GMRES.java; restart 30 by default (:84-86)."""
    kind = L.GMRES

    def __init__(self, template, restart: int = 30):
        if restart <= 0:
            raise ValueError("restart must be a positive integer")
        super().__init__(template, restart)

    def setRestart(self, restart: int):
        if restart <= 0:
            raise ValueError("restart must be a positive integer")
        self.restart = restart

    def fused_step(self, A: "CompRowMatrix", M: "Preconditioner", V, i: int, w, proj):
        """One fused inner step on device tensors (rsgpu_gmres_fused_step):
        w = M^-1 (A V[:, i]) and proj[k] = V[:, k]' w for k <= i.  V is a
        column-major (ldv x cols) view: a torch tensor of shape (cols, ldv)
        whose rows are the basis vectors (float64, or float32 for the FP32
        form)."""
        lib = L.load()
        h = self._handle(A.ctx)
        _torch_sync(V, w)
        f = lib.rsgpu_gmres_fused_step_f32 if V.dtype.is_floating_point and \
            V.element_size() == 4 else lib.rsgpu_gmres_fused_step
        code = f(h, A.h, M.h, V.data_ptr(), V.stride(0), i, w.data_ptr(), proj.data_ptr())
        _raise(code, A.ctx.h)


class IR(AbstractIterativeSolver):
    """mtj/.../sparse/IR.java."""
    kind = L.IR


class MixedPrecisionSolver(AbstractIterativeSolver):
    """FP64 iterative refinement on the true residual around an FP32 inner
    solve of the given kind (BiCGstab or GMRES) with an FP32 preconditioner.
    No reference counterpart; the outer loop has IR.java's form."""

    def __init__(self, template, inner: str = "gmres", restart: int = 30,
                 inner_rtol: float = 1e-4, inner_max_iter: int = 200, cgs: bool = False):
        super().__init__(template, restart)
        self.kind = {"gmres": L.GMRES, "bicgstab": L.BICGSTAB}[inner]
        self.mixed = True
        self.inner_rtol = inner_rtol
        self.inner_max_iter = inner_max_iter
        self.cgs = cgs


# ------------------------------------------------------------------ MPFA-O
def mpfa_o(dim, nx, ny, nz, dx, dy, dz, K6, continuity=1.0, ctx: Optional[Context] = None):
    """Device O_MPFA transmissibilities -> (md_ptr, md_elem, md_t, check_failures)."""
    ctx = ctx or Context.default()
    if dim == 3:
        nconn = 3 * nx * ny * nz - nx * ny - nx * nz - ny * nz
        cap = 18
    else:
        nconn = 2 * nx * ny - nx - ny
        cap = 6
    md_ptr = np.zeros(nconn + 1, np.int32)
    md_elem = np.zeros(max(nconn * cap, 1), np.int32)
    md_t = np.zeros(max(nconn * cap, 1))
    fails = C.c_int(0)
    K = np.ascontiguousarray(K6, np.float64).reshape(-1)
    _raise(L.load().rsgpu_mpfa_o(ctx.h, dim, nx, ny, nz, dx, dy, dz, _hptr(K), continuity,
                                 _hptr(md_ptr), _hptr(md_elem), _hptr(md_t),
                                 C.byref(fails)), ctx.h)
    m = int(md_ptr[-1])
    return md_ptr, md_elem[:m].copy(), md_t[:m].copy(), fails.value


def mpfa_o_assemble(dim, nx, ny, nz, dx, dy, dz, K6, continuity=1.0,
                    ctx: Optional[Context] = None):
    """Device MPFA-O + LinearSolver pattern + flux assembly -> CompRowMatrix."""
    ctx = ctx or Context.default()
    h = C.c_void_p()
    fails = C.c_int(0)
    K = np.ascontiguousarray(K6, np.float64).reshape(-1)
    _raise(L.load().rsgpu_mpfa_o_assemble(ctx.h, dim, nx, ny, nz, dx, dy, dz, _hptr(K),
                                          continuity, C.byref(h), C.byref(fails)), ctx.h)
    n = nx * ny * (nz if dim == 3 else 1)
    A = CompRowMatrix._wrap(h, n, n, 0, ctx)
    A.nnz = int(A.getRowPointers()[-1])
    return A, fails.value
