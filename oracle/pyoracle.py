"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke``
and by the ``cpu_baseline`` leg of ``bench.py`` -- never by the shipped
package.  The C sources restate MTJ and rs (see oracle/oracle.h for the
reference file list and the statement-order contract).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

OK, BREAKDOWN, DIVERGENCE, ITERATIONS = 0, 1, 2, 3
ZERO_PIVOT, MISSING_DIAGONAL, NOT_IN_PATTERN, BAD_ARGUMENT, SINGULAR = 10, 11, 12, 13, 14

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class AmgParams(C.Structure):
    _fields_ = [("omega_pre_f", C.c_double), ("omega_pre_r", C.c_double),
                ("omega_post_f", C.c_double), ("omega_post_r", C.c_double),
                ("nu1", C.c_int), ("nu2", C.c_int), ("gamma", C.c_int), ("min", C.c_int),
                ("omega", C.c_double), ("reverse", C.c_int), ("mode", C.c_int),
                ("block_rows", C.c_int), ("max_coarse", C.c_int), ("zero_guess", C.c_int),
                ("ordering", C.c_int)]


def amg_params(**kw):
    """AMG() defaults (AMG.java:204-206) overridden by keyword."""
    p = AmgParams()
    lib().or_amg_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown AMG parameter {k}")
        setattr(p, k, v)
    return p


class Opts(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("rtol", C.c_double), ("atol", C.c_double),
                ("dtol", C.c_double), ("restart", C.c_int), ("prec_kind", C.c_int),
                ("nblocks", C.c_int), ("amg", AmgParams)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("status", C.c_int),
                ("bad_row", C.c_int), ("hist_len", C.c_int)]


class BoParams(C.Structure):
    """or_bo_params (oracle.h); field order identical to rsgpu_blackoil_params."""
    _fields_ = [(f, C.c_double) for f in (
        "dx", "dy", "dz", "dt", "g", "phi0", "cr", "pref", "kz_ratio", "bw0", "cw", "bo0", "co",
        "bg0", "rs0", "muw", "muo", "mug", "rhow_s", "rhoo_s", "rhog_s", "swr", "sor", "sgr",
        "kw", "ko", "kg")] + [(f, C.c_int) for f in ("nw", "no", "ng", "pad")]


IMPES_FIELDS = ("pressure", "phase_pressure", "saturation", "mass_density", "phase_mobility",
                "present", "comp_mobility", "dRdN", "dRdp", "heat_capacity", "dRdT",
                "energy_mobility", "temperature", "p0", "R0", "cx", "cy", "cz",
                "mole_fraction", "molar_density", "T0")


class ImpesParams(C.Structure):
    """or_impes_params (oracle.h)."""
    _fields_ = [(f, C.c_int) for f in ("n", "nconn", "ncomp", "thermal")] + \
               [(f, C.c_double) for f in ("dt", "g", "omega", "ax", "ay")]


class ImpesState(C.Structure):
    """or_impes_state (oracle.h): one pointer per field of IMPES_FIELDS."""
    _fields_ = [(f, C.c_void_p) for f in IMPES_FIELDS]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_csr_construct.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p]
        L.or_csr_add.argtypes = [_i32p, _i32p, _f64p, C.c_int, C.c_int, C.c_double]
        L.or_zero_rows.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_double, _i32p, C.c_int]
        L.or_csr_mult.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_csr_mult_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.or_csr_multadd.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_double, _f64p, _f64p]
        L.or_csr_transmult.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_csr_transmultadd.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_double,
                                          _f64p, _f64p]
        L.or_dot.argtypes = [C.c_int, _f64p, _f64p]
        L.or_dot.restype = C.c_double
        L.or_norm2.argtypes = [C.c_int, _f64p]
        L.or_norm2.restype = C.c_double
        L.or_ilu_factor.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, C.POINTER(C.c_int)]
        L.or_ilu_factor_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _i32p, C.POINTER(C.c_int)]
        L.or_ilu_apply.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _f64p, _f64p, _f64p]
        L.or_ilu_apply_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _i32p, _f32p, _f32p, _f32p]
        L.or_ilu_transapply.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _f64p, _f64p,
                                        _f64p]
        L.or_block_filter_count.argtypes = [C.c_int, _i32p, _i32p, C.c_int, _i32p]
        L.or_block_filter.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p,
                                      _i32p, _f64p, _i32p]
        L.or_block_starts.argtypes = [C.c_int, C.c_int, _i32p]
        for f in (L.or_bicgstab, L.or_gmres, L.or_ir, L.or_cg, L.or_bicg, L.or_cgs, L.or_qmr):
            f.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.POINTER(Opts),
                          C.POINTER(Stats), _f64p, C.c_int]
        L.or_gmres_bjilu_omp.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                         C.POINTER(Opts), C.POINTER(Stats), C.c_int]
        L.or_icc_factor.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, C.c_void_p]
        L.or_icc_apply.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _f64p, _f64p, _f64p]
        L.or_icc_apply.restype = None
        L.or_dgesv.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, _i32p, _f64p, C.c_int]
        L.or_ilut_factor.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_double, C.c_int, _i32p,
                                     _i32p, _f64p, _i32p, C.c_void_p]
        L.or_ilut_bound.argtypes = [C.c_int, _i32p, _i32p, C.c_int]
        L.or_ilut_bound.restype = C.c_longlong
        L.or_amg_params_default.argtypes = [C.c_void_p]
        L.or_amg_setup.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        L.or_amg_apply.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int]
        L.or_amg_apply.restype = None
        L.or_amg_free.argtypes = [C.c_void_p]
        L.or_amg_free.restype = None
        L.or_amg_levels.argtypes = [C.c_void_p, _i32p, _i32p, C.c_int]
        L.or_amg_get_level.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 7
        L.or_amg_p_nnz.argtypes = [C.c_void_p, C.c_int]
        L.or_amg_aggregate.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_double, C.c_int,
                                       _i32p, C.c_void_p, C.c_void_p]
        L.or_hashset_order.argtypes = [_i32p, C.c_int, _i32p]
        L.or_hashset_order.restype = None
        L.or_amg_priority.argtypes = [C.c_int]
        L.or_amg_priority.restype = C.c_uint64
        L.or_ssor_apply.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_int, _f64p, _f64p, _i32p]
        L.or_mesh_counts.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_int)] * 3
        L.or_mesh_connections.argtypes = [C.c_int] * 4 + [_i32p, _i32p]
        L.or_mpfa_o.argtypes = [C.c_int] * 4 + [C.c_double] * 3 + [_f64p, C.c_double, _i32p,
                                                                    _i32p, _f64p,
                                                                    C.POINTER(C.c_int)]
        L.or_mpfa_corner.argtypes = ([C.c_int] * 4 + [C.c_double] * 3 +
                                     [_f64p, C.c_double, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)] + [_f64p] * 5 + [_i32p])
        L.or_element_centers.argtypes = [C.c_int] * 4 + [C.c_double] * 3 + [_f64p]
        L.or_interface_geometry.argtypes = [C.c_int] * 4 + [C.c_double] * 3 + [_f64p, _f64p]
        L.or_pattern_count.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int,
                                       _i32p]
        L.or_pattern.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, _i32p,
                                 _i32p]
        L.or_assemble_flux.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _f64p,
                                       _i32p, _i32p, _f64p]
        L.or_connection_flux.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_cell_rates.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _f64p]
        L.or_mpfa_o_vjp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, _f64p, C.c_double, _i32p, _i32p, _f64p, _f64p]
        L.or_cell_rates.restype = None
        L.or_bo_pattern.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.or_bo_assemble.argtypes = [C.POINTER(BoParams), C.c_int, C.c_int, C.c_int, _f64p,
                                     _f64p, _f64p, C.c_void_p, C.c_void_p]
        L.or_bo_assemble.restype = None
        L.or_bo_expand.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p]
        L.or_impes_cs.argtypes = [C.POINTER(ImpesParams), C.POINTER(ImpesState), _i32p, _i32p,
                                  _i32p, _i32p, _f64p, _f64p, _i32p, _f64p]
        L.or_impes_cs.restype = None
        L.or_impes_molar_update.argtypes = [C.POINTER(ImpesParams), C.POINTER(ImpesState),
                                            _i32p, _i32p, _f64p, _i32p, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _f64p,
                                            _f64p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_impes_temperature.argtypes = [C.POINTER(ImpesParams), C.POINTER(ImpesState),
                                           _i32p, _i32p, _i32p, _i32p, _f64p, _i32p, _i32p,
                                           _f64p, _f64p, _i32p, C.c_int, C.c_void_p,
                                           C.c_void_p, _i32p, _i32p, _f64p, _f64p]
        L.or_impes_assemble.argtypes = [C.POINTER(ImpesParams), C.POINTER(ImpesState), _i32p,
                                        _i32p, _i32p, _i32p, _f64p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, _f64p, _i32p, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, _i32p, _i32p,
                                        _f64p, _f64p]
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- CSR
def csr_construct(n, ncols, rows):
    """CompRowMatrix(int,int,int[][]): rows = list of column lists."""
    rowlen = _i32([len(r) for r in rows])
    cols = _i32(np.concatenate([np.asarray(r, dtype=np.int32) for r in rows])
                if rows and sum(len(r) for r in rows) else np.zeros(0, np.int32))
    rowptr = np.zeros(n + 1, np.int32)
    colind = np.zeros(max(int(rowlen.sum()), 1), np.int32)
    st = lib().or_csr_construct(n, ncols, rowlen, cols if cols.size else np.zeros(1, np.int32),
                                rowptr, colind)
    return st, rowptr, colind[: int(rowlen.sum())]


def csr_add(rowptr, colind, data, row, col, value):
    return lib().or_csr_add(rowptr, colind, data, row, col, value)


def zero_rows(rowptr, colind, data, diagonal, rows):
    rows = _i32(rows)
    return lib().or_zero_rows(len(rowptr) - 1, rowptr, colind, data, diagonal,
                              rows if rows.size else np.zeros(1, np.int32), rows.size)


def mult(rowptr, colind, data, x):
    n = len(rowptr) - 1
    y = np.empty(n)
    lib().or_csr_mult(n, _i32(rowptr), _i32(colind), _f64(data), _f64(x), y)
    return y


def mult_f32(rowptr, colind, data, x):
    n = len(rowptr) - 1
    y = np.empty(n, np.float32)
    lib().or_csr_mult_f32(n, _i32(rowptr), _i32(colind),
                          np.ascontiguousarray(data, np.float32),
                          np.ascontiguousarray(x, np.float32), y)
    return y


def multadd(rowptr, colind, data, alpha, x, y):
    y = _f64(y).copy()
    lib().or_csr_multadd(len(rowptr) - 1, _i32(rowptr), _i32(colind), _f64(data), alpha,
                         _f64(x), y)
    return y


def transmult(rowptr, colind, data, x, ncols=None):
    n = len(rowptr) - 1
    ncols = n if ncols is None else ncols
    y = np.empty(ncols)
    lib().or_csr_transmult(n, ncols, _i32(rowptr), _i32(colind), _f64(data), _f64(x), y)
    return y


def transmultadd(rowptr, colind, data, alpha, x, y):
    n = len(rowptr) - 1
    y = _f64(y).copy()
    lib().or_csr_transmultadd(n, y.size, _i32(rowptr), _i32(colind), _f64(data), alpha,
                              _f64(x), y)
    return y


def dot(x, y):
    return lib().or_dot(len(x), _f64(x), _f64(y))


def norm2(x):
    return lib().or_norm2(len(x), _f64(x))


# ---------------------------------------------------------------- ILU
def ilu_factor(rowptr, colind, data):
    n = len(rowptr) - 1
    lu = _f64(data).copy()
    diag = np.zeros(max(n, 1), np.int32)
    bad = C.c_int(-1)
    st = lib().or_ilu_factor(n, _i32(rowptr), _i32(colind), lu, diag, C.byref(bad))
    return st, lu, diag[:n], bad.value


def ilu_factor_f32(rowptr, colind, data):
    n = len(rowptr) - 1
    lu = np.ascontiguousarray(data, np.float32).copy()
    diag = np.zeros(max(n, 1), np.int32)
    bad = C.c_int(-1)
    st = lib().or_ilu_factor_f32(n, _i32(rowptr), _i32(colind), lu, diag, C.byref(bad))
    return st, lu, diag[:n], bad.value


def ilu_apply(rowptr, colind, lu, diag, b):
    n = len(rowptr) - 1
    x = np.empty(n)
    y = np.empty(n)
    lib().or_ilu_apply(n, _i32(rowptr), _i32(colind), _f64(lu), _i32(diag), _f64(b), x, y)
    return x


def ilu_apply_f32(rowptr, colind, lu, diag, b):
    n = len(rowptr) - 1
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    lib().or_ilu_apply_f32(n, _i32(rowptr), _i32(colind), np.ascontiguousarray(lu, np.float32),
                           _i32(diag), np.ascontiguousarray(b, np.float32), x, y)
    return x


def ilu_transapply(rowptr, colind, lu, diag, b):
    n = len(rowptr) - 1
    x = np.empty(n)
    y = np.empty(n)
    lib().or_ilu_transapply(n, _i32(rowptr), _i32(colind), _f64(lu), _i32(diag), _f64(b), x, y)
    return x


def icc_factor(rowptr, colind, data):
    """ICC.setMatrix: (status, R values on the same pattern, diagind, bad_row
    1-based)."""
    n = len(rowptr) - 1
    R = _f64(data).copy()
    dg = np.zeros(max(n, 1), np.int32)
    bad = C.c_int(-1)
    st = lib().or_icc_factor(n, _i32(rowptr), _i32(colind), R, dg, C.byref(bad))
    return st, R, dg[:n], bad.value


def icc_apply(rowptr, colind, R, diag, b):
    n = len(rowptr) - 1
    x = np.zeros(n)
    y = np.zeros(n)
    lib().or_icc_apply(n, _i32(rowptr), _i32(colind), _f64(R), _i32(diag), _f64(b), x, y)
    return x


def ilut_factor(rowptr, colind, data, tau=1e-6, p=25):
    """ILUT.setMatrix (ILUT.java:136-208): (status, rowptr, colind, values,
    diag positions, bad_row 1-based) of the factor."""
    n = len(rowptr) - 1
    rp, ci = _i32(rowptr), _i32(colind)
    cap = int(lib().or_ilut_bound(n, rp, ci, int(p)))
    orp = np.zeros(n + 1, np.int32)
    oci = np.zeros(max(cap, 1), np.int32)
    ova = np.zeros(max(cap, 1))
    dg = np.zeros(max(n, 1), np.int32)
    bad = C.c_int(-1)
    st = lib().or_ilut_factor(n, rp, ci, _f64(data), float(tau), int(p), orp, oci, ova, dg,
                              C.byref(bad))
    nnz = int(orp[n]) if st == 0 else 0
    return st, orp, oci[:nnz].copy(), ova[:nnz].copy(), dg[:n].copy(), bad.value


def block_starts(n, nblocks):
    bs = np.zeros(nblocks + 1, np.int32)
    st = lib().or_block_starts(n, nblocks, bs)
    if st:
        raise ValueError("bad block count")
    return bs


def block_filter(rowptr, colind, data, nblocks):
    n = len(rowptr) - 1
    bs = block_starts(n, nblocks)
    cnt = lib().or_block_filter_count(n, _i32(rowptr), _i32(colind), nblocks, bs)
    browptr = np.zeros(n + 1, np.int32)
    bcol = np.zeros(max(cnt, 1), np.int32)
    bdat = np.zeros(max(cnt, 1))
    src = np.zeros(max(cnt, 1), np.int32)
    lib().or_block_filter(n, _i32(rowptr), _i32(colind), _f64(data), nblocks, bs, browptr,
                          bcol, bdat, src)
    return browptr, bcol[:cnt], bdat[:cnt], src[:cnt]


# ---------------------------------------------------------------- solvers
PREC = {"none": 0, "identity": 0, "ilu": 1, "block_ilu": 2, "amg": 3, "diagonal": 4,
        "ssor": 5, "icc": 6}


def solve(kind, rowptr, colind, data, b, x0=None, *, max_iter=100000, rtol=1e-5, atol=1e-50,
          dtol=1e5, restart=30, prec="ilu", nblocks=1, hist_cap=None, amg=None):
    """Run or_bicgstab / or_gmres / or_ir / or_cg / or_bicg / or_cgs / or_qmr.
    Returns (x, Stats, history).
    amg: AmgParams (or dict of overrides) when prec == "amg"."""
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    if amg is None or isinstance(amg, dict):
        amg = amg_params(**(amg or {}))
    o = Opts(max_iter, rtol, atol, dtol, restart, PREC[prec], nblocks, amg)
    s = Stats()
    cap = (max_iter + 2) * 2 if hist_cap is None else hist_cap
    cap = max(1, min(cap, 10_000_000))
    hist = np.zeros(cap)
    fn = {"bicgstab": lib().or_bicgstab, "gmres": lib().or_gmres, "ir": lib().or_ir,
          "cg": lib().or_cg, "bicg": lib().or_bicg, "cgs": lib().or_cgs,
          "qmr": lib().or_qmr}[kind]
    fn(n, _i32(rowptr), _i32(colind), _f64(data), _f64(b), x, C.byref(o), C.byref(s), hist, cap)
    return x, s, hist[: min(s.hist_len, cap)]


def gmres_bjilu_omp(rowptr, colind, data, b, x0=None, *, nthreads=0, max_iter=100000,
                    rtol=1e-5, atol=1e-50, dtol=1e5, restart=30, nblocks=1):
    """or_gmres_bjilu_omp: GMRES(restart) + block-Jacobi ILU(0) over nblocks
    equal row ranges on nthreads OpenMP threads (the CPU baseline's
    all-cores leg).  Returns (x, Stats)."""
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    o = Opts(max_iter, rtol, atol, dtol, restart, PREC["block_ilu"], nblocks, amg_params())
    s = Stats()
    lib().or_gmres_bjilu_omp(n, _i32(rowptr), _i32(colind), _f64(data), _f64(b), x,
                             C.byref(o), C.byref(s), int(nthreads))
    return x, s


class AMG:
    """Oracle AMG hierarchy (or_amg_*): setMatrix on construction."""

    def __init__(self, rowptr, colind, data, **params):
        self.p = amg_params(**params)
        self.n = len(rowptr) - 1
        h = C.c_void_p()
        bad = C.c_int(-1)
        st = lib().or_amg_setup(self.n, _i32(rowptr), _i32(colind), _f64(data),
                                C.byref(self.p), C.byref(h), C.byref(bad))
        if st:
            raise RuntimeError(f"or_amg_setup status {st} (row {bad.value})")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_amg_free(self.h)
            self.h = None

    def levels(self):
        sz = np.zeros(64, np.int32)
        nz = np.zeros(64, np.int32)
        m = lib().or_amg_levels(self.h, sz, nz, 64)
        return [(int(sz[k]), int(nz[k])) for k in range(m)]

    def level(self, k):
        """(rowptr, colind, data, pt, P rowptr, P colind, P data); the last
        four are None on the coarsest level."""
        lv = self.levels()
        n, nnz = lv[k]
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        if k < len(lv) - 1:
            pn = lib().or_amg_p_nnz(self.h, k)
            pt = np.zeros(max(n, 1), np.int32)
            prp = np.zeros(n + 1, np.int32)
            pci = np.zeros(max(pn, 1), np.int32)
            pv = np.zeros(max(pn, 1))
            lib().or_amg_get_level(self.h, k, *[a.ctypes.data for a in
                                                  (rp, ci, v, pt, prp, pci, pv)])
            return rp, ci[:nnz], v[:nnz], pt[:n], prp, pci[:pn], pv[:pn]
        lib().or_amg_get_level(self.h, k, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                               None, None, None, None)
        return rp, ci[:nnz], v[:nnz], None, None, None, None

    def apply(self, b, x, transpose=False):
        x = _f64(x).copy()
        lib().or_amg_apply(self.h, _f64(b), x, int(bool(transpose)))
        return x


def amg_aggregate(rowptr, colind, data, eps, mode=0):
    n = len(rowptr) - 1
    pt = np.zeros(max(n, 1), np.int32)
    nagg = C.c_int()
    nroots = C.c_int()
    st = lib().or_amg_aggregate(n, _i32(rowptr), _i32(colind), _f64(data), eps, mode, pt,
                                C.byref(nagg), C.byref(nroots))
    if st:
        raise RuntimeError(f"or_amg_aggregate status {st}")
    return pt[:n], nagg.value, nroots.value


def hashset_order(cols):
    cols = _i32(cols)
    out = np.zeros(max(len(cols), 1), np.int32)
    lib().or_hashset_order(cols, len(cols), out)
    return out[: len(cols)]


def ssor_apply(rowptr, colind, data, b, x, *, omega_f=1.0, omega_r=1.0, reverse=True,
               block_rows=0, ordering=0, return_order=False):
    n = len(rowptr) - 1
    x = _f64(x).copy()
    order = np.zeros(max(n, 1), np.int32)
    st = lib().or_ssor_apply(n, _i32(rowptr), _i32(colind), _f64(data), block_rows, ordering,
                             omega_f, omega_r, int(reverse), _f64(b), x, order)
    if st:
        raise RuntimeError(f"or_ssor_apply status {st}")
    return (x, order[:n]) if return_order else x


def dgesv(a, b):
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    n = a.shape[0]
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    af = np.array(a.T, dtype=np.float64, order="C").reshape(-1)  # column-major copy
    bf = np.array(b.reshape(n, nrhs).T, dtype=np.float64, order="C").reshape(-1)
    piv = np.zeros(n, np.int32)
    st = lib().or_dgesv(n, nrhs, af, n, piv, bf, n)
    return st, bf.reshape(nrhs, n).T


# ---------------------------------------------------------------- mesh / MPFA
def mesh_counts(dim, nx, ny, nz=0):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    st = lib().or_mesh_counts(dim, nx, ny, nz, C.byref(a), C.byref(b), C.byref(c))
    if st:
        raise ValueError("bad mesh")
    return a.value, b.value, c.value


def mesh_connections(dim, nx, ny, nz=0):
    _, nconn, _ = mesh_counts(dim, nx, ny, nz)
    here = np.zeros(max(nconn, 1), np.int32)
    there = np.zeros(max(nconn, 1), np.int32)
    lib().or_mesh_connections(dim, nx, ny, nz, here, there)
    return here[:nconn], there[:nconn]


def mpfa_o(dim, nx, ny, nz, dx, dy, dz, K6, continuity=1.0):
    """O_MPFA transmissibilities: (md_ptr, md_elem, md_t, n_check_failures)."""
    _, nconn, _ = mesh_counts(dim, nx, ny, nz)
    cap = 18 if dim == 3 else 6
    md_ptr = np.zeros(nconn + 1, np.int32)
    md_elem = np.zeros(max(nconn * cap, 1), np.int32)
    md_t = np.zeros(max(nconn * cap, 1))
    fails = C.c_int(0)
    st = lib().or_mpfa_o(dim, nx, ny, nz, dx, dy, dz, _f64(K6).reshape(-1), continuity,
                         md_ptr, md_elem, md_t, C.byref(fails))
    if st:
        raise RuntimeError(f"or_mpfa_o failed with status {st}")
    m = int(md_ptr[-1])
    return md_ptr, md_elem[:m].copy(), md_t[:m].copy(), fails.value


def mpfa_corner(dim, nx, ny, nz, dx, dy, dz, K6, point, continuity=1.0):
    A = np.zeros(144)
    B = np.zeros(96)
    Cm = np.zeros(144)
    D = np.zeros(96)
    T = np.zeros(96)
    el = np.zeros(8, np.int32)
    nc, ne = C.c_int(), C.c_int()
    st = lib().or_mpfa_corner(dim, nx, ny, nz, dx, dy, dz, _f64(K6).reshape(-1), continuity,
                              point, C.byref(nc), C.byref(ne), A, B, Cm, D, T, el)
    if st:
        raise RuntimeError(f"or_mpfa_corner failed with status {st}")
    c, e = nc.value, ne.value

    def cm(a, r, k):
        return a[: r * k].reshape(k, r).T.copy()
    return dict(A=cm(A, c, c), B=cm(B, c, e), C=cm(Cm, c, c), D=cm(D, c, e), T=cm(T, c, e),
                elements=el[:e].copy())


def element_centers(dim, nx, ny, nz, dx, dy, dz):
    n, _, _ = mesh_counts(dim, nx, ny, nz)
    out = np.zeros(3 * n)
    lib().or_element_centers(dim, nx, ny, nz, dx, dy, dz, out)
    return out.reshape(n, 3)


def interface_geometry(dim, nx, ny, nz, dx, dy, dz):
    _, nconn, _ = mesh_counts(dim, nx, ny, nz)
    na = np.zeros(4 * max(nconn, 1))
    cc = np.zeros(3 * max(nconn, 1))
    lib().or_interface_geometry(dim, nx, ny, nz, dx, dy, dz, na, cc)
    return na[: 4 * nconn].reshape(nconn, 4), cc[: 3 * nconn].reshape(nconn, 3)


def pattern(n, here, there, md_ptr, md_elem, with_diag=True):
    nconn = len(here)
    rowlen = np.zeros(n, np.int32)
    args = (n, nconn, _i32(here), _i32(there), _i32(md_ptr), _i32(md_elem), int(with_diag))
    st = lib().or_pattern_count(*args, rowlen)
    if st:
        raise RuntimeError("pattern overflow")
    rowptr = np.zeros(n + 1, np.int32)
    colind = np.zeros(max(int(rowlen.sum()), 1), np.int32)
    lib().or_pattern(*args, rowptr, colind)
    return rowptr, colind[: int(rowlen.sum())]


def assemble_flux(n, here, there, md_ptr, md_elem, md_t, rowptr, colind):
    data = np.zeros(len(colind))
    st = lib().or_assemble_flux(n, len(here), _i32(here), _i32(there), _i32(md_ptr),
                                _i32(md_elem), _f64(md_t), _i32(rowptr), _i32(colind), data)
    if st:
        raise RuntimeError(f"assembly failed with status {st}")
    return data


def mpfa_o_vjp(dim, nx, ny, nz, dx, dy, dz, K6, md_ptr, md_elem, tbar, continuity=1.0):
    """d(tbar . t)/d m for K_e = exp(m_e) K0_e, per cell."""
    n = nx * ny * (nz if dim == 3 else 1)
    out = np.zeros(max(n, 1))
    st = lib().or_mpfa_o_vjp(dim, nx, ny, nz if dim == 3 else 0, dx, dy, dz,
                             _f64(np.ascontiguousarray(K6).reshape(-1)), continuity,
                             _i32(md_ptr), _i32(md_elem), _f64(tbar), out)
    if st:
        raise RuntimeError(f"or_mpfa_o_vjp status {st}")
    return out[:n]


def cell_rates(here, there, flux, cells):
    cells = _i32(cells)
    r = np.zeros(max(len(cells), 1))
    lib().or_cell_rates(len(here), _i32(here), _i32(there), _f64(flux), len(cells), cells, r)
    return r[: len(cells)]


def connection_flux(md_ptr, md_elem, md_t, p):
    nconn = len(md_ptr) - 1
    f = np.zeros(max(nconn, 1))
    lib().or_connection_flux(nconn, _i32(md_ptr), _i32(md_elem), _f64(md_t), _f64(p), f)
    return f[:nconn]


# ---------------------------------------------------------------- black-oil
def _bo_params(params: dict) -> BoParams:
    p = BoParams()
    for k, v in params.items():
        setattr(p, k, v)
    return p


def bo_pattern(nx, ny, nz):
    """7-point block pattern (browptr, bcolind)."""
    n = nx * ny * nz
    nnzb = lib().or_bo_pattern(nx, ny, nz, None, None)
    brp = np.zeros(n + 1, np.int32)
    bci = np.zeros(nnzb, np.int32)
    lib().or_bo_pattern(nx, ny, nz, brp.ctypes.data, bci.ctypes.data)
    return brp, bci


def bo_expand(brp, bci):
    """Scalar CSR pattern of a 3x3 block pattern (block rows interleaved)."""
    nb = len(brp) - 1
    rp = np.zeros(3 * nb + 1, np.int32)
    ci = np.zeros(9 * len(bci), np.int32)
    lib().or_bo_expand(nb, _i32(brp), _i32(bci), rp, ci)
    return rp, ci


def bo_assemble(params: dict, nx, ny, nz, kh, u, u_old, need_jac=True):
    """(values in the row-interleaved block layout or None, residual)."""
    n = nx * ny * nz
    nnzb = lib().or_bo_pattern(nx, ny, nz, None, None)
    val = np.zeros(9 * nnzb) if need_jac else None
    R = np.zeros(3 * n)
    P = _bo_params(params)
    lib().or_bo_assemble(C.byref(P), nx, ny, nz, _f64(kh), _f64(u), _f64(u_old),
                         val.ctypes.data if need_jac else None, R.ctypes.data)
    return val, R


# ---------------------------------------------------------------- IMPES
def _impes_structs(params: dict, state: dict):
    """(ImpesParams, ImpesState, keep-alive list) from a parameter dict (n,
    nconn, ncomp, thermal, dt, g, omega, ax, ay) and a dict of numpy arrays
    keyed by IMPES_FIELDS (thermal fields may be absent)."""
    P = ImpesParams()
    for k, v in params.items():
        setattr(P, k, v)
    S = ImpesState()
    keep = []
    for f in IMPES_FIELDS:
        a = state.get(f)
        if a is None:
            continue
        a = np.ascontiguousarray(a, dtype=np.uint8 if f == "present" else np.float64)
        keep.append(a)
        setattr(S, f, a.ctypes.data)
    return P, S, keep


def impes_cs(params, state, here, there, md_ptr, md_elem, md_t):
    """Control-surface secondaries: (rho [3][nconn], upstream [3][nconn],
    darcy [3][nconn])."""
    P, S, keep = _impes_structs(params, state)
    nc = params["nconn"]
    rho = np.zeros(3 * nc)
    up = np.zeros(3 * nc, np.int32)
    u = np.zeros(3 * nc)
    lib().or_impes_cs(C.byref(P), C.byref(S), _i32(here), _i32(there), _i32(md_ptr),
                      _i32(md_elem), _f64(md_t), rho, up, u)
    return rho, up, u


def impes_assemble(params, state, here, there, md_ptr, md_elem, md_t, rho, upstream,
                   rowptr, colind, mf=None, sources=None):
    """(status, values on the (rowptr, colind) pattern, residual b).  mf =
    (mf_ptr, mf_elem, mf_t) for the thermal Fourier term; sources = (elem,
    kind, q [nterms][3][ncomp], qe) in application order."""
    P, S, keep = _impes_structs(params, state)
    n = params["n"]
    val = np.zeros(int(rowptr[n]))
    b = np.zeros(n)
    if mf is not None:
        mfp, mfe, mft = _i32(mf[0]), _i32(mf[1]), _f64(mf[2])
        keep += [mfp, mfe, mft]
        mfa = (mfp.ctypes.data, mfe.ctypes.data, mft.ctypes.data)
    else:
        mfa = (None, None, None)
    if sources is not None:
        se, sk, sq, sqe = (_i32(sources[0]), _i32(sources[1]), _f64(sources[2]),
                           _f64(sources[3]))
        keep += [se, sk, sq, sqe]
        nt = len(se)
        sa = (se.ctypes.data, sk.ctypes.data, sq.ctypes.data, sqe.ctypes.data)
    else:
        nt, sa = 0, (None, None, None, None)
    st = lib().or_impes_assemble(C.byref(P), C.byref(S), _i32(here), _i32(there),
                                 _i32(md_ptr), _i32(md_elem), _f64(md_t), *mfa, _f64(rho),
                                 _i32(upstream), nt, *sa, _i32(rowptr), _i32(colind), val, b)
    return st, val, b


def impes_molar_update(params, state, here, there, darcy, upstream, N_old, sources=None,
                       locked=None, truncate=False):
    """(status, N_new [ncomp][n], truncated count, bad index)."""
    P, S, keep = _impes_structs(params, state)
    n = params["n"]
    N_new = np.zeros_like(_f64(N_old))
    if sources is not None:
        se, sk, sq = _i32(sources[0]), _i32(sources[1]), _f64(sources[2])
        keep += [se, sk, sq]
        nt, sa = len(se), (se.ctypes.data, sk.ctypes.data, sq.ctypes.data)
    else:
        nt, sa = 0, (None, None, None)
    lk = None
    if locked is not None:
        lk = np.zeros(n, np.uint8)
        lk[np.asarray(locked, np.int64)] = 1
        keep.append(lk)
    tr, bad = C.c_int(0), C.c_int(-1)
    st = lib().or_impes_molar_update(C.byref(P), C.byref(S), _i32(here), _i32(there),
                                     _f64(darcy), _i32(upstream), nt, *sa,
                                     lk.ctypes.data if lk is not None else None,
                                     int(truncate), _f64(N_old), N_new, C.byref(tr),
                                     C.byref(bad))
    return st, N_new, tr.value, bad.value


def impes_temperature(params, state, here, there, md_ptr, md_elem, md_t, mf, rho, upstream,
                      rowptr, colind, sources=None):
    """(status, A values on the pattern, b) of the temperature system."""
    P, S, keep = _impes_structs(params, state)
    n = params["n"]
    val = np.zeros(int(rowptr[n]))
    b = np.zeros(n)
    if sources is not None:
        se, sqe = _i32(sources[0]), _f64(sources[3])
        keep += [se, sqe]
        nt, sa = len(se), (sqe.ctypes.data, se.ctypes.data)
    else:
        nt, sa = 0, (None, None)
    st = lib().or_impes_temperature(C.byref(P), C.byref(S), _i32(here), _i32(there),
                                    _i32(md_ptr), _i32(md_elem), _f64(md_t), _i32(mf[0]),
                                    _i32(mf[1]), _f64(mf[2]), _f64(rho), _i32(upstream), nt,
                                    *sa, _i32(rowptr), _i32(colind), val, b)
    return st, val, b
