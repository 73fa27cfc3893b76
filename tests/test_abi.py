"""The C-ABI library loads and exports every entry point include/rsgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os

from ressim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = _lib.declared_symbols(os.path.join(ROOT, "include", "rsgpu.h"))
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES)


def test_version_and_defaults():
    lib = _lib.load()
    assert lib.rsgpu_version() >= 1
    o = _lib.SolverOpts()
    lib.rsgpu_solver_opts_default(ctypes.byref(o))
    # DefaultIterationMonitor() defaults (DefaultIterationMonitor.java:85-90)
    assert (o.max_iter, o.rtol, o.atol, o.dtol, o.restart) == (100000, 1e-5, 1e-50, 1e5, 30)


def test_null_handles_are_rejected():
    lib = _lib.load()
    assert lib.rsgpu_csr_destroy(None) == 0
    assert lib.rsgpu_csr_mult(None, None, None) == _lib.BAD_ARGUMENT
    assert lib.rsgpu_solver_solve(None, None, None, None, None, None) == _lib.BAD_ARGUMENT
