"""rsgpu_flux_rates_noalloc (the well-rate entry point that allocates
nothing: its staging lives in the library's static device array) against
rsgpu_flux_rates, bit for bit, through the C ABI: two calls in a row, the
second with fewer cells, then the size limit."""
import ctypes as C

import numpy as np
import pytest

from ressim_amd import problems as P

pytestmark = pytest.mark.gpu

LIMIT = 65536


def _rates(lib, fn, F, flux, cells):
    out = np.full(max(cells.size, 1), np.nan)
    rc = fn(F.h, flux.data_ptr(), cells.size, cells.ctypes.data_as(C.c_void_p),
            out.ctypes.data_as(C.c_void_p))
    return rc, out[: cells.size]


def test_noalloc_rates_match_allocating_entry_point(gpu):
    import torch
    from ressim_amd import _lib as L
    from ressim_amd.potential import FluxOperator
    lib = L.load()
    pr = P.config5(nx=48, ny=40, nz=36)          # 69,120 cells: room for the full limit
    F, fails = FluxOperator.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6)
    assert fails == 0
    rng = np.random.default_rng(11)
    p = torch.tensor(rng.standard_normal(pr.n), device="cuda")
    flux = F.calculateFlux(p)
    torch.cuda.synchronize()
    for count in (LIMIT, 37):                    # second call smaller: stale tail unused
        cells = np.ascontiguousarray(rng.choice(pr.n, count, replace=False), np.int32)
        print(f"no-alloc rates: {count} cells of {pr.n}", flush=True)
        rc0, ref = _rates(lib, lib.rsgpu_flux_rates, F, flux, cells)
        rc1, got = _rates(lib, lib.rsgpu_flux_rates_noalloc, F, flux, cells)
        assert rc0 == L.OK and rc1 == L.OK
        np.testing.assert_array_equal(got, ref)
    cells = np.zeros(LIMIT + 1, np.int32)
    rc, _ = _rates(lib, lib.rsgpu_flux_rates_noalloc, F, flux, cells)
    assert rc == L.BAD_ARGUMENT
    assert b"limit" in lib.rsgpu_last_error(F.ctx.h)
    rc, _ = _rates(lib, lib.rsgpu_flux_rates_noalloc, F, flux, np.array([pr.n], np.int32))
    assert rc == L.BAD_ARGUMENT
    rc, out = _rates(lib, lib.rsgpu_flux_rates_noalloc, F, flux, np.zeros(0, np.int32))
    assert rc == L.OK and out.size == 0
