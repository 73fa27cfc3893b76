"""Benchmark: 1M-cell 3-D MPFA-O pressure solve on one MI355X.

Workload (the largest single-GPU configuration of the benchmark list): the
MPFA-O pressure system on 128 x 128 x 64 cells (n = 1,048,576, 27-point,
27,725,560 nonzeros), anisotropic log-normal permeability (sigma 2,
k_v/k_h = 0.1, seeded), +1/-1 sources in opposite corners, the
accumulation diagonal of the pressure Newton Jacobian (no locked cell),
GMRES(30) + block-Jacobi ILU(0) to true relative residual 1e-8:
  * FP64: the drop-in MTJ GMRES(30) form with FP64 block ILU(0) (re-entered
    from its own iterate until the true residual reaches 1e-8);
  * mixed: FP64 iterative refinement on the true residual around FP32
    GMRES(30) + FP32 block ILU(0).
A step = setMatrix (factorisation, plus the FP32 value copy in mixed mode)
+ solve from x0 = 0, on inputs already resident in HBM.  ``value`` is mixed
solves/s.  The SpMV roofline is timed with HIP events on the stream the
kernel runs on.  The CPU baseline is the oracle (scalar C restatement of
MTJ's loops) on a bounded sample of the same FP64 solve.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, s_val, s_vec):
    return nnz * (s_val + 4) + (n + 1) * 4 + n * s_vec + n * s_vec


# column of the fused GMRES step timed by the bench: 15 basis vectors, the
# mean of a GMRES(30) cycle's 1..30
FUSED_I = 14


def fused_bytes(n, nnz, nnz_lu, s, i):
    """Algorithmic bytes of the fused GMRES step at column i: the CSR SpMV
    and the ILU apply (the per-unit figures above) without the round trip
    of their intermediate u = A v (never stored), plus the i + 1 basis
    vectors read for the projections."""
    return spmv_bytes(n, nnz, s, s) + ilu_bytes(n, nnz_lu, s) - 2 * n * s + (i + 1) * n * s


def sliced_bytes(n, nchunks, nslots, s_val):
    """Bytes the sliced SpMV must move: every value slot (zeros where a row
    lacks the slot), the 4-byte row masks, two ints per chunk, x read once
    and y written once (the offset tables are a few hundred bytes)."""
    return nslots * s_val + n * 4 + nchunks * 8 + n * s_val + n * s_val


def ilu_bytes(n, nnz_lu, s):
    """ILU(0) apply (L and U sweeps): factor values and int32 columns once,
    two row-pointer arrays, the diagonal index, b read, y written and read,
    x written (algorithmic bytes per application)."""
    return nnz_lu * (s + 4) + 2 * (n + 1) * 4 + 4 * n + 4 * n * s


# HBM bytes per launch from the committed rocprofv3 counter passes
# (FETCH_SIZE x factor + WRITE_SIZE x factor, the factors from a 512 MiB
# copy in the same process); valid for the 128x128x64 workload they were
# taken on.
PMC_PROFILE = os.path.join(ROOT, "profiles", "k_spmv_mpfa3d_128x128x64_pmc.md")
ILU_PMC_PROFILE = os.path.join(ROOT, "profiles", "k_ilu_apply_mpfa3d_128x128x64_pmc.md")
ILU32_KERNEL = "k_ilu_apply_band<float>"
FUSED_PMC_PROFILE = os.path.join(ROOT, "profiles", "k_gm_fused_mpfa3d_128x128x64_pmc.md")
FUSED_KERNEL = "k_gm_fused<float>"
# the FP64 drop-in's orthogonalisation on the fused step
ORTHO64 = "DCGS2 (classical Gram-Schmidt, reorthogonalisation delayed by one step)"


BO_PMC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                              "k_spmv_b3_blackoil_100x100x50_pmc.md")


def measured_traffic(pr, kernel, profile=PMC_PROFILE, shape=(128, 128, 64)):
    """The kernel's `traffic` column (scaled FETCH_SIZE + WRITE_SIZE bytes
    per dispatch) from a committed counter profile's table, for the grid
    `shape` that profile was taken on."""
    if (pr.nx, pr.ny, pr.nz) != shape or not os.path.exists(profile):
        return None
    col = None
    for line in open(profile):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 2 and cells[1] == "kernel":
            col = cells.index("traffic") if "traffic" in cells else None
        elif col is not None and len(cells) > col and cells[1] == f"`{kernel}`" \
                and cells[col].endswith(" B"):
            return int(cells[col][:-2].replace(",", ""))
    return None


def host_cpu():
    """(threads for the all-cores leg, CPU model): the threads OpenMP is
    given (OMP_NUM_THREADS, else the CPUs this process may run on) and the
    lscpu model name."""
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        threads = 0
    if threads <= 0:
        threads = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return threads, model


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--nblocks", type=int, default=256)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--inner-rtol", type=float, default=1e-3)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-amg", action="store_true",
                    help="skip the GMRES(30) + AMG legs")
    ap.add_argument("--amg-ratio", type=float, default=0.5,
                    help="AMG coarsening stops at a level whose aggregates keep more than this "
                         "fraction of its rows (0: MTJ's rule)")
    ap.add_argument("--cpu-iters", type=int, default=300)
    ap.add_argument("--adjoint", action="store_true",
                    help="measure the adjoint gradient workload (64x64x32 well-rate mismatch) "
                         "instead of the pressure solve and print its own JSON line")
    ap.add_argument("--ensemble", action="store_true",
                    help="measure the batched ensemble workload (64 cubic MPFA-O systems, "
                         "edges 32..96) instead of the pressure solve and print its own JSON "
                         "line")
    ap.add_argument("--ens-count", type=int, default=64)
    ap.add_argument("--ens-emin", type=int, default=32)
    ap.add_argument("--ens-emax", type=int, default=96)
    ap.add_argument("--ens-rows", type=int, default=4096,
                    help="largest block-Jacobi ILU(0) block: whole x-lines inside one "
                         "xy-plane of one system")
    ap.add_argument("--ens-inner-rtol", type=float, default=1e-3)
    ap.add_argument("--ens-accumulation", type=float, default=1e-3,
                    help="accumulation diagonal of each system (0: incompressible, centre "
                         "cell locked)")
    ap.add_argument("--ens-exact", action="store_true",
                    help="also run the FP64 solve with the exact ILU(0) of every system")
    ap.add_argument("--blackoil", action="store_true",
                    help="measure the fully-implicit black-oil Newton step (100x100x50, 3x3 "
                         "blocks) instead of the pressure solve and print its own JSON line")
    ap.add_argument("--bo-lines", type=int, default=16,
                    help="x-lines per block-Jacobi ILU(0) block of the black-oil Jacobian")
    ap.add_argument("--bo-nz", type=int, default=50)
    ap.add_argument("--c2-locked", action="store_true",
                    help="the 128x128x64 system in its incompressible form (no accumulation, "
                         "centre cell locked by zeroRows): FP64 and mixed solve times")
    ap.add_argument("--locked-refinements", type=int, default=60)
    ap.add_argument("--impes", action="store_true",
                    help="measure the IMPES pressure-Jacobian assembly (control surfaces, "
                         "flux and diagonal parts, sources) on the 128x128x64 MPFA-O mesh "
                         "instead of the pressure solve and print its own JSON line")
    ap.add_argument("--impes-ncomp", type=int, default=3)
    ap.add_argument("--impes-thermal", action="store_true")
    ap.add_argument("--restart", type=int, default=30)
    ap.add_argument("--fp64-mgs", action="store_true",
                    help="FP64 GMRES with MTJ's modified Gram-Schmidt (default: classical "
                         "Gram-Schmidt with reorthogonalisation, delayed by one step on the "
                         "fused step)")
    ap.add_argument("--mgs", action="store_true",
                    help="modified Gram-Schmidt in the FP32 inner GMRES (default: classical)")
    ap.add_argument("--probe", choices=["spmv", "ilu", "ilu32", "solve", "mixed", "amg", "amg32",
                                        "fused"],
                    default=None,
                    help="profiling probe: run only the named kernel(s) on the workload "
                         "(plus a 512 MiB device copy as the traffic reference) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each box has one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from ressim_amd import _lib, mtj, problems as P

    if a.adjoint:
        return bench_adjoint(a)
    if a.ensemble:
        return bench_ensemble(a)
    if a.blackoil:
        return bench_blackoil(a)
    if a.impes:
        return bench_impes(a)
    if a.c2_locked:
        return bench_c2_locked(a)
    lib = _lib.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    pr = P.config2(nx=a.nx, ny=a.ny, nz=a.nz)
    n = pr.n
    log(f"workload {pr.name} n={n}")
    ctx = mtj.Context.default()
    stream = torch.cuda.Stream()          # default priority
    ctx.set_stream(stream.cuda_stream)

    t0 = time.time()
    A, fails = mtj.mpfa_o_assemble(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                                   ctx=ctx)
    # accumulation diagonal after the flux part (Discretisation.java:286-290)
    if pr.accumulation:
        idx = np.arange(n, dtype=np.int32)
        A.add(idx, idx, np.full(n, pr.accumulation))
    if pr.pinned:
        mtj.Matrices.zeroRows(A, 1.0, pr.pinned)
    nnz = A.nnz
    log(f"assembled on GPU in {time.time() - t0:.2f}s nnz={nnz} check_failures={fails}")
    b = torch.tensor(pr.rhs(), device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    if a.probe:
        return probe(a, A, n, nnz, lib, mtj, _lib, ctx, stream, dev)

    # ---------------------------------------------------------- solvers
    M64 = mtj.BlockILU(a.nblocks, precision=_lib.FP64, ctx=ctx)
    M32 = mtj.BlockILU(a.nblocks, precision=_lib.FP32, ctx=ctx)
    g64 = mtj.GMRES(b, 30)
    g64.cgs = not a.fp64_mgs           # DCGS2 / CGS2; MTJ's MGS measured beside it
    g64.setIterationMonitor(mtj.DefaultIterationMonitor(20000, a.rtol, 1e-50, 1e5))
    g64.setPreconditioner(M64)
    gmx = mtj.MixedPrecisionSolver(b, inner="gmres", restart=a.restart,
                                   inner_rtol=a.inner_rtol, inner_max_iter=600,
                                   cgs=not a.mgs)
    gmx.setIterationMonitor(mtj.DefaultIterationMonitor(200, a.rtol, 1e-50, 1e5))
    gmx.setPreconditioner(M32)

    def step_fp64():
        M64.setMatrix(A)
        x.zero_()
        g64.solve(A, b, x)
        return g64.last_stats

    # MTJ's GMRES monitors the preconditioned residual ||M^-1 r||: calibrate
    # its relative tolerance once so that the true residual ||b - Ax||/||b||
    # reaches the target (checked again after every timed solve).
    mon_rtol = a.rtol
    for _ in range(4):
        st = step_fp64()
        ratio = (st.true_residual / st.b_norm) / a.rtol
        if ratio <= 1.0:
            break
        mon_rtol = mon_rtol / (1.25 * ratio)
        g64.setIterationMonitor(mtj.DefaultIterationMonitor(20000, mon_rtol, 1e-50, 1e5))
    log(f"fp64 monitor rtol {mon_rtol:.3e} (true/target {ratio:.2f})")

    def step_mixed():
        A.sync_f32()
        M32.setMatrix(A)
        x.zero_()
        gmx.solve(A, b, x)
        return gmx.last_stats

    def timed(fn, steps, warmup, label):
        for _ in range(warmup):
            st = fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            st = fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        if st.true_residual > a.rtol * st.b_norm:
            log(f"WARNING {label}: true residual above target")
        log(f"{label}: {dt * 1e3:.2f} ms/solve iters={st.iterations} "
            f"inner={st.inner_iterations} outer={st.outer_iterations} "
            f"true_rel_res={st.true_residual / st.b_norm:.3e}")
        return dt, st

    dt64, st64 = timed(step_fp64, max(1, a.steps // 2), a.warmup,
                       f"fp64 GMRES(30)+BJ-ILU0 ({'MGS' if a.fp64_mgs else ORTHO64})")
    x64 = x.clone()
    mgs = None
    if not a.fp64_mgs:
        # MTJ's own orthogonalisation (modified Gram-Schmidt), same monitor
        g64.cgs = False
        dtm, stm = timed(step_fp64, 1, 1, "fp64 GMRES(30)+BJ-ILU0 (MGS)")
        mgs = {"ms_per_solve": dtm * 1e3, "iterations": stm.iterations,
               "true_rel_residual": stm.true_residual / stm.b_norm}
        g64.cgs = True
    dtmx, stmx = timed(step_mixed, a.steps, a.warmup, "mixed FP32-GMRES/FP64-IR")
    xmx = x.clone()
    err_mx = float(torch.linalg.norm(xmx - x64) / torch.linalg.norm(x64))
    log(f"mixed vs fp64 relative solution difference {err_mx:.3e}")

    # ---------------------------------------------------------- kernel timings
    # Each kernel is timed with HIP events on the stream it runs on over many
    # back-to-back launches: warm (the same operands every launch, so up to
    # 256 MB of them stay in the last-level cache) and cold (the launches
    # cycle over C copies of the operands, C x the kernel's bytes > 256 MB,
    # so each launch reads its operands from HBM, as inside a solve whose
    # working set far exceeds that cache).  roofline.achieved uses the cold
    # time.
    rp_h, ci_h, d_h = A.getRowPointers(), A.getColumnIndices(), A.getData()
    NCOPY = 4

    def time_launches(launchers, reps):
        for f in launchers:
            f()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for r in range(reps):
            launchers[r % len(launchers)]()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    mats = [A] + [mtj.CompRowMatrix.from_csr(n, n, rp_h, ci_h, d_h, ctx=ctx)
                  for _ in range(NCOPY - 1)]
    xs64 = [torch.randn(n, dtype=torch.float64, device=dev) for _ in range(NCOPY)]
    ys64 = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(NCOPY)]
    xs32 = [v.float() for v in xs64]
    ys32 = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(NCOPY)]
    for Mt in mats:
        Mt.sync_f32()
    M64.setMatrix(A)
    M32.setMatrix(A)
    pre64 = [M64] + [mtj.BlockILU(a.nblocks, precision=_lib.FP64, ctx=ctx)
                     for _ in range(NCOPY - 1)]
    pre32 = [M32] + [mtj.BlockILU(a.nblocks, precision=_lib.FP32, ctx=ctx)
                     for _ in range(NCOPY - 1)]
    for c in range(1, NCOPY):
        pre64[c].setMatrix(mats[c])
        pre32[c].setMatrix(mats[c])
    torch.cuda.synchronize()
    R = a.spmv_reps

    def both(make, reps):
        fs = [make(c) for c in range(NCOPY)]
        return {"warm_us": time_launches(fs[:1], reps) * 1e6,
                "cold_us": time_launches(fs, reps) * 1e6}

    b64, b32 = spmv_bytes(n, nnz, 8, 8), spmv_bytes(n, nnz, 4, 4)
    kt = {}
    kt["spmv_csr_fp64"] = both(lambda c: lambda: lib.rsgpu_csr_mult(
        mats[c].h, xs64[c].data_ptr(), ys64[c].data_ptr()), R)
    kt["spmv_csr_fp32"] = both(lambda c: lambda: lib.rsgpu_csr_mult_f32(
        mats[c].h, xs32[c].data_ptr(), ys32[c].data_ptr()), R)
    nch, nslots, ntab = A.sliced_info()
    if nch:
        for Mt in mats:
            assert lib.rsgpu_csr_sync_sliced(Mt.h) == 0
        kt["spmv_sliced_fp64"] = both(lambda c: lambda: lib.rsgpu_csr_mult_sliced(
            mats[c].h, xs64[c].data_ptr(), ys64[c].data_ptr()), R)
        kt["spmv_sliced_fp32"] = both(lambda c: lambda: lib.rsgpu_csr_mult_sliced_f32(
            mats[c].h, xs32[c].data_ptr(), ys32[c].data_ptr()), R)
    info = M64.info()
    nnz_lu = info["nnz"]
    kt["ilu_apply_fp64"] = both(lambda c: lambda: lib.rsgpu_prec_apply(
        pre64[c].h, xs64[c].data_ptr(), ys64[c].data_ptr()), max(10, R // 2))
    kt["ilu_apply_fp32"] = both(lambda c: lambda: lib.rsgpu_prec_apply_f32(
        pre32[c].h, xs32[c].data_ptr(), ys32[c].data_ptr()), max(10, R // 2))
    # the fused GMRES step (the mixed solve's dominant kernel, k_gm_fused):
    # one launch at column FUSED_I on copy c's matrix, preconditioner and
    # basis (rsgpu_gmres_fused_step_f32)
    gf = mtj.GMRES(xs64[0], a.restart)
    gf.cgs = True
    hf = gf._handle(ctx)
    ldv = (n + 3) // 4 * 4
    Vs = [torch.randn(FUSED_I + 1, ldv, dtype=torch.float32, device=dev) for _ in range(NCOPY)]
    wsf = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(NCOPY)]
    prj = [torch.empty(FUSED_I + 1, dtype=torch.float64, device=dev) for _ in range(NCOPY)]
    fused_ok = nch and lib.rsgpu_gmres_fused_step_f32(
        hf, A.h, M32.h, Vs[0].data_ptr(), ldv, FUSED_I, wsf[0].data_ptr(),
        prj[0].data_ptr()) == 0
    for c in range(1, NCOPY if fused_ok else 1):
        fused_ok = lib.rsgpu_gmres_fused_step_f32(
            hf, mats[c].h, pre32[c].h, Vs[c].data_ptr(), ldv, FUSED_I, wsf[c].data_ptr(),
            prj[c].data_ptr()) == 0
    if fused_ok:
        kt["gmres_fused_fp32"] = both(lambda c: lambda: lib.rsgpu_gmres_fused_step_f32(
            hf, mats[c].h, pre32[c].h, Vs[c].data_ptr(), ldv, FUSED_I, wsf[c].data_ptr(),
            prj[c].data_ptr()), max(10, R // 2))
    else:
        log("fused GMRES step not timed: " + _lib.load().rsgpu_last_error(ctx.h).decode())
    ctx.set_stream(None)
    del mats[1:], pre64[1:], pre32[1:], Vs, wsf, prj
    ilu64_bytes, ilu32_bytes = ilu_bytes(n, nnz_lu, 8), ilu_bytes(n, nnz_lu, 4)
    fz_bytes = fused_bytes(n, nnz, nnz_lu, 4, FUSED_I)
    for k, byt in (("spmv_csr_fp64", b64), ("spmv_csr_fp32", b32), ("spmv_sliced_fp64", b64),
                   ("spmv_sliced_fp32", b32), ("ilu_apply_fp64", ilu64_bytes),
                   ("ilu_apply_fp32", ilu32_bytes), ("gmres_fused_fp32", fz_bytes)):
        if k in kt:
            v = kt[k]
            v["algorithmic_bytes"] = byt
            v["cold_GBps"] = byt / v["cold_us"] / 1e3
            v["warm_GBps"] = byt / v["warm_us"] / 1e3
            v["cold_frac"] = v["cold_GBps"] / PEAK_HBM_GBS
            log(f"{k}: warm {v['warm_us']:.1f} us, cold {v['cold_us']:.1f} us = "
                f"{v['cold_GBps']:.0f} GB/s ({v['cold_frac']:.1%} of 8 TB/s) of "
                f"{byt / 1e6:.1f} MB algorithmic")
    if nch:
        kt["spmv_sliced_fp64"]["format_bytes"] = sliced_bytes(n, nch, nslots, 8)
        kt["spmv_sliced_fp32"]["format_bytes"] = sliced_bytes(n, nch, nslots, 4)
        kt["sliced_format"] = {"chunks": nch, "slots": nslots, "tables": ntab}
    kt["cold_copies"] = NCOPY

    # ---------------------------------------------------------- AMG
    amg, xs_amg = (None, {}) if a.no_amg else bench_amg(a, A, b, n, dev)

    # ---------------------------------------------------------- CPU baseline
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(A, pr, a, mon_rtol, st64.iterations, x64, xmx, xs_amg)
        except Exception as e:  # the baseline must not hide the GPU result
            log(f"cpu baseline failed: {e!r}")
        if amg and cpu:
            for k in xs_amg:
                amg[k]["rel_err_vs_fp64_oracle"] = cpu.pop(f"rel_err_amg_{k}", None)

    spmv_key = "spmv_sliced_fp64" if nch else "spmv_csr_fp64"
    il = kt["ilu_apply_fp32"]
    sp = kt[spmv_key]
    if "gmres_fused_fp32" in kt:
        fz = kt["gmres_fused_fp32"]
        fz["column"] = FUSED_I
        roof = {"bound": "hbm", "achieved": fz["cold_GBps"], "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": fz["cold_frac"],
                "traffic": measured_traffic(pr, FUSED_KERNEL, FUSED_PMC_PROFILE),
                "kernel": f"{FUSED_KERNEL} (the FP32 inner GMRES step: sliced SpMV, "
                          "block-Jacobi ILU(0) apply and the first projection pass in one "
                          f"launch, at column {FUSED_I}; the mixed solve's largest kernel "
                          "time)",
                "algorithmic_bytes": fz_bytes, "cold_us": fz["cold_us"],
                "warm_us": fz["warm_us"]}
    else:
        roof = {"bound": "hbm", "achieved": il["cold_GBps"], "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": il["cold_frac"],
                "traffic": measured_traffic(pr, ILU32_KERNEL, ILU_PMC_PROFILE),
                "kernel": f"{ILU32_KERNEL} (block-Jacobi ILU(0) apply of the FP32 inner "
                          "solve)",
                "algorithmic_bytes": ilu32_bytes, "cold_us": il["cold_us"],
                "warm_us": il["warm_us"]}
    out = {
        "metric": "solves/s + SpMV HBM GB/s, 1M-cell MPFA, mixed vs FP64 at 1e-8 residual",
        "value": 1.0 / dtmx,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dtmx * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement",
        "data": "synthetic (seeded log-normal permeability, MPFA-O assembled on the GPU)",
        "config": {"workload": f"3-D MPFA-O pressure {pr.nx}x{pr.ny}x{pr.nz}, n={n}, "
                               f"nnz={nnz}, GMRES(30) + block-Jacobi ILU(0) x{a.nblocks}, "
                               f"rel residual {a.rtol:g}",
                   "n": n, "nnz": nnz, "nblocks": a.nblocks},
        "fp64": {"solves_per_s": 1.0 / dt64, "ms_per_solve": dt64 * 1e3,
                 "orthogonalisation": "MGS" if a.fp64_mgs else ORTHO64,
                 "mgs": mgs,
                 "iterations": st64.iterations,
                 "true_rel_residual": st64.true_residual / st64.b_norm,
                 "rel_err_vs_fp64_oracle": cpu and cpu.get("rel_err_fp64")},
        "mixed": {"solves_per_s": 1.0 / dtmx, "ms_per_solve": dtmx * 1e3,
                  "outer_iterations": stmx.outer_iterations,
                  "inner_iterations": stmx.inner_iterations,
                  "true_rel_residual": stmx.true_residual / stmx.b_norm,
                  "rel_err_vs_fp64_oracle": cpu and cpu.get("rel_err_mixed"),
                  "rel_err_vs_gpu_fp64": err_mx},
        "amg": amg,
        "kernels": kt,
        "spmv": {"kernel": spmv_key, "cold_GBps": sp["cold_GBps"], "cold_us": sp["cold_us"],
                 "warm_us": sp["warm_us"], "frac": sp["cold_frac"],
                 "algorithmic_bytes": b64},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(a):
    """Gradient of J(m) = sum_w (q_w - q_obs_w)^2 with respect to all
    131,072 log-permeabilities (64x64x32, 4 producers locked, 1 injector;
    q_obs from a second seeded field): forward and adjoint times, central
    differences (step 1e-6) on 32 random cells whose gradient exceeds 1e-3
    of the largest, and the oracle's CPU FP64 adjoint."""
    import torch
    from ressim_amd import problems as P
    from ressim_amd.adjoint import WellRateMismatch
    nx, ny, nz = 64, 64, 32
    pr = P.config5(nx=nx, ny=ny, nz=nz)
    obs = P.config5(seed=6, nx=nx, ny=ny, nz=nz)
    m = np.log(pr.K6[:, 0])
    W0 = WellRateMismatch(3, nx, ny, nz, 1.0, 1.0, -1.0, obs.K6 / obs.K6[:, :1], obs.sources,
                          obs.pinned)
    q_obs = W0.rates(np.log(obs.K6[:, 0]))
    W = WellRateMismatch(3, nx, ny, nz, 1.0, 1.0, -1.0, pr.K6 / pr.K6[:, :1], pr.sources,
                         pr.pinned, q_obs=q_obs)
    W.forward(m)
    W.backward()
    tf, tb = [], []
    for _ in range(max(1, a.steps)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        J = W.forward(m)
        torch.cuda.synchronize()
        tf.append(time.perf_counter() - t)
        t = time.perf_counter()
        g = W.backward()
        torch.cuda.synchronize()
        tb.append(time.perf_counter() - t)
    rng = np.random.default_rng(2024)
    big = np.flatnonzero(np.abs(g) >= 1e-3 * np.abs(g).max())
    cells = rng.choice(big, min(32, big.size), replace=False)
    errs = {}
    for h in (1e-6, 1e-4):
        e = []
        for c in cells:
            mp, mm = m.copy(), m.copy()
            mp[c] += h
            mm[c] -= h
            fd = (W.forward(mp) - W.forward(mm)) / (2 * h)
            e.append(abs(fd - g[c]) / abs(fd))
        errs[h] = e
    log(f"adjoint: J={J:.6e} forward {np.median(tf) * 1e3:.1f} ms backward "
        f"{np.median(tb) * 1e3:.1f} ms iterations {W.last_iterations}")
    cpu = None
    if not a.no_cpu:
        from oracle import pyoracle as po
        from tests.helpers import oracle_well_mismatch
        t = time.perf_counter()
        Jo, qo, go = oracle_well_mismatch(po, pr, m, q_obs)
        cpu = {"seconds": time.perf_counter() - t, "cores": 1, "kind": "port",
               "rel_J_diff": abs(J - Jo) / abs(Jo),
               "rel_grad_diff": float(np.linalg.norm(g - go) / np.linalg.norm(go))}
    out = {"metric": "adjoint gradient of well-rate mismatch w.r.t. log-permeability",
           "value": 1.0 / (np.median(tf) + np.median(tb)), "unit": "gradients/s",
           "n_gpus": 1, "steps": a.steps, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic (seeded log-normal permeability, two fields)",
           "config": {"workload": f"MPFA-O 3-D {nx}x{ny}x{nz}, 4 producers + 1 injector, "
                                  f"BiCGstab + block ILU(0) x{W.nblocks} to 1e-13 forward and "
                                  f"adjoint", "n": pr.n},
           "forward_ms": float(np.median(tf) * 1e3), "backward_ms": float(np.median(tb) * 1e3),
           "fd_cells": int(cells.size),
           "fd_rel_err_h1e-6": {"max": max(errs[1e-6]), "median": float(np.median(errs[1e-6]))},
           "fd_rel_err_h1e-4": {"max": max(errs[1e-4]), "median": float(np.median(errs[1e-4]))},
           "cpu_adjoint": cpu}
    print(json.dumps(out), flush=True)


def bench_impes(a):
    """IMPES pressure-Jacobian assembly (Discretisation.java:284-289 and
    :365-561 with CS.java:94-170) on the 128x128x64 MPFA-O mesh of the
    pressure solve: one step = the control-surface secondaries plus the
    flux part, the diagonal part and six source terms, into A's pattern and
    b, from seeded control-volume secondaries of three phases and
    --impes-ncomp components resident in HBM.  value = assemblies/s."""
    import torch
    from ressim_amd import mtj, problems as P
    from ressim_amd.impes import PressureJacobian
    from ressim_amd.potential import FluxOperator
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    ctx = mtj.Context.default()
    stream = torch.cuda.Stream()          # default priority
    ctx.set_stream(stream.cuda_stream)
    pr = P.config2(nx=a.nx, ny=a.ny, nz=a.nz)
    n, ncomp, thermal = pr.n, a.impes_ncomp, a.impes_thermal
    t0 = time.time()
    F, fails = FluxOperator.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                                   ctx=ctx)
    Fmf = None
    if thermal:
        # Fourier lists: MPFA-O of an isotropic conductivity field
        Fmf, _ = FluxOperator.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz,
                                     np.tile([2.0, 2.0, 2.0, 0.0, 0.0, 0.0], (n, 1)), ctx=ctx)
    A, _ = mtj.mpfa_o_assemble(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                               ctx=ctx)
    nnz, nconn, nent = A.nnz, F.nconn, F.md_entries
    J = PressureJacobian(F, ncomp, fourier=Fmf)
    st = P.impes_state(21, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, ncomp=ncomp,
                       thermal=thermal)
    elem, kind, q, qe = P.impes_sources(4, n, ncomp)
    J.setSources(elem, kind, q, qe)
    sd = {k: torch.tensor(v, device=dev) for k, v in st.items()}
    params = dict(P.IMPES_PARAMS, thermal=int(thermal))
    rho = torch.empty(3 * nconn, dtype=torch.float64, device=dev)
    up = torch.empty(3 * nconn, dtype=torch.int32, device=dev)
    u = torch.empty(3 * nconn, dtype=torch.float64, device=dev)
    b = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log(f"IMPES assembly {pr.name}: n={n} connections={nconn} MD entries={nent} nnz={nnz} "
        f"ncomp={ncomp} thermal={int(thermal)} set up in {time.time() - t0:.1f}s")

    def step(ev=None):
        if ev:
            ev[0].record(stream)
        J.controlSurfaces(params, sd, rho, up, u)
        if ev:
            ev[1].record(stream)
        J.assemble(A, params, sd, rho, up, b)
        if ev:
            ev[2].record(stream)

    for _ in range(max(1, a.warmup)):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
    t = time.perf_counter()
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t) / a.steps
    t_cs = float(np.median([e[0].elapsed_time(e[1]) for e in evs])) * 1e-3
    t_asm = float(np.median([e[1].elapsed_time(e[2]) for e in evs])) * 1e-3
    # the rest of the IMPES step on the same state: the explicit molar-mass
    # update and (thermal runs) the temperature system
    N_old = torch.full((ncomp * n,), 1.0e12, dtype=torch.float64, device=dev)
    N_new = torch.empty_like(N_old)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    J.molarUpdate(params, sd, u, up, N_old, N_new, truncate=True)
    e0.record(stream)
    for _ in range(a.steps):
        J.molarUpdate(params, sd, u, up, N_old, N_new, truncate=True)
    e1.record(stream)
    e1.synchronize()
    t_mm = e0.elapsed_time(e1) / a.steps * 1e-3
    t_temp = None
    if thermal:
        At, _ = mtj.mpfa_o_assemble(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                                    ctx=ctx)
        bt = torch.empty(n, dtype=torch.float64, device=dev)
        J.assembleTemperature(At, params, sd, rho, up, bt)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(a.steps):
            J.assembleTemperature(At, params, sd, rho, up, bt)
        e1.record(stream)
        e1.synchronize()
        t_temp = e0.elapsed_time(e1) / a.steps * 1e-3
    log(f"molar-mass update {t_mm * 1e6:.0f} us per call (host syncs for the status)"
        + (f"; temperature system {t_temp * 1e6:.0f} us" if t_temp else ""))
    # algorithmic bytes: each API call's inputs read once and outputs written once
    cells_cs = 8 * (3 * 4 + 3) * n                       # 4 phase arrays + centres
    cells_asm = (4 * 8 + 24 + 3 + 24 * ncomp + 8 * ncomp + 24) * n
    if thermal:
        cells_asm += (8 + 8 + 24 + 8) * n
    md = 12 * nent
    bytes_cs = cells_cs + 12 * nconn + md + 60 * nconn
    bytes_asm = (cells_asm + 48 * nconn + md + 4 * n + 10 * nconn + 4 * (n + 1) + 4 * nnz
                 + 8 * nnz + 8 * n)
    if thermal:
        bytes_asm += md + 4 * nconn
    achieved = (bytes_cs + bytes_asm) / (t_cs + t_asm) / 1e9
    log(f"control surfaces {t_cs * 1e6:.0f} us, assembly {t_asm * 1e6:.0f} us (kernels + "
        f"one 4-byte status read), wall {wall * 1e3:.2f} ms per step; {achieved:.0f} GB/s of "
        f"algorithmic bytes")
    cpu = None
    if not a.no_cpu:
        from oracle import pyoracle as po
        md_ptr, md_elem, md_t, _ = mtj.mpfa_o(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz,
                                              pr.K6, ctx=ctx)
        here, there = po.mesh_connections(pr.dim, pr.nx, pr.ny, pr.nz)
        rp, ci = A.getRowPointers(), A.getColumnIndices()
        op = dict(params, n=n, nconn=nconn, ncomp=ncomp)
        mf = (md_ptr, md_elem, md_t) if thermal else None
        reps, tc = 0, time.perf_counter()
        while True:
            rho_o, up_o, _ = po.impes_cs(op, st, here, there, md_ptr, md_elem, md_t)
            code, val_o, b_o = po.impes_assemble(op, st, here, there, md_ptr, md_elem, md_t,
                                                 rho_o, up_o, rp, ci, sources=(elem, kind, q, qe),
                                                 mf=mf)
            reps += 1
            if time.perf_counter() - tc > 10 or reps >= 20:
                break
        t_cpu = (time.perf_counter() - tc) / reps
        same = (not thermal) and bool(np.array_equal(A.getData(), val_o) and
                                      np.array_equal(b.cpu().numpy(), b_o))
        log(f"CPU oracle: {t_cpu * 1e3:.1f} ms per assembly (1 core, {reps} reps); GPU result "
            f"bit-identical: {same}")
        cpu = {"value": 1.0 / t_cpu, "unit": "assemblies/s", "cores": 1, "kind": "port",
               "sample": f"the full {n}-cell assembly, {reps} repetitions",
               "bitwise_equal": same}
    value = 1.0 / (t_cs + t_asm)
    out = {"metric": "IMPES pressure-Jacobian assemblies/s (control surfaces + flux + diagonal "
                     "+ sources), 128x128x64 MPFA-O mesh",
           "value": value, "unit": "assemblies/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": (t_cs + t_asm) * 1e3, "wall_ms_per_step":
           wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded control-volume secondaries, MPFA-O "
                                   "transmissibilities of the pressure-solve permeability)",
           "config": {"workload": f"IMPES pressure Jacobian {pr.nx}x{pr.ny}x{pr.nz}, "
                                  f"{ncomp} components, thermal {int(thermal)}",
                      "n": n, "connections": nconn, "md_entries": nent, "nnz": nnz},
           "control_surfaces_us": t_cs * 1e6, "assembly_us": t_asm * 1e6,
           "molar_update_us": t_mm * 1e6,
           "temperature_system_us": t_temp * 1e6 if t_temp else None,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                        "frac": achieved / 8000.0, "traffic": None,
                        "kernel": "k_impes_cs + k_impes_conn + k_impes_rows",
                        "algorithmic_bytes": bytes_cs + bytes_asm},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_ensemble(a):
    """Batched ensemble: ``count`` MPFA-O pressure systems on cubic grids
    (edges uniform in [emin, emax], own log-normal field each, sigma 1,
    centre locked), solved together to true relative residual 1e-8 per
    system.  A step = setMatrix (FP32 values + factorisation) + the batched
    solve from x0 = 0.  value = ensembles solved per second in the mixed form
    (per system: FP64 refinement around FP32 BiCGstab + FP32 block-Jacobi
    ILU(0), blocks of --ens-rows rows inside each system); the FP64 batched
    BiCGstab with the same blocks is reported beside it, and with
    --ens-exact the FP64 BiCGstab with each system's exact ILU(0)."""
    import torch
    from ressim_amd import _lib, ensemble as E, mtj, problems as P
    lib = _lib.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    ctx = mtj.Context.default()
    stream = torch.cuda.Stream()          # default priority
    ctx.set_stream(stream.cuda_stream)
    prs = P.config4(count=a.ens_count, emin=a.ens_emin, emax=a.ens_emax,
                    accumulation=a.ens_accumulation)
    t0 = time.time()
    A, sys_start, b_np, fails = E.assemble_ensemble(prs, ctx)
    n, nnz, nsys = A.n, A.nnz, len(prs)
    log(f"ensemble: {nsys} systems n={n} nnz={nnz} assembled in {time.time() - t0:.1f}s "
        f"check_failures={fails}")
    b = torch.tensor(b_np, device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    bs = E.grid_blocks(prs, a.ens_rows)
    mon = lambda: mtj.DefaultIterationMonitor(5000, a.rtol, 1e-50, 1e5)  # noqa: E731

    def make(prec, blocks, mixed):
        M = mtj.BlockILU(1, precision=prec, ctx=ctx)
        M.setBlocks(blocks)
        if mixed:
            A.sync_f32()
        t = time.time()
        M.setMatrix(A)
        log(f"  pattern analysis + factor ({'fp32' if mixed else 'fp64'}, "
            f"{len(blocks) - 1} blocks) {time.time() - t:.1f}s")
        S = E.EnsembleSolver(sys_start, mon(), mixed=mixed, inner_rtol=a.ens_inner_rtol,
                             inner_max_iter=2000, ctx=ctx)
        return M, S

    def step(M, S, mixed):
        if mixed:
            A.sync_f32()
        M.setMatrix(A)
        x.zero_()
        return S.solve(A, M, b, x)

    def timed(M, S, mixed, steps, warmup, label):
        for _ in range(warmup):
            st = step(M, S, mixed)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            st = step(M, S, mixed)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        rel = max(s.true_residual / s.b_norm for s in st)
        its = [s.iterations for s in st]
        inner = [s.inner_iterations for s in st]
        log(f"{label}: {dt * 1e3:.1f} ms/ensemble ({nsys / dt:.1f} systems/s) iterations "
            f"min/max {min(its)}/{max(its)} inner max {max(inner)} max true_rel_res {rel:.2e} "
            f"graph replays {S.info()['graph_launches']}")
        if rel > a.rtol:
            log(f"WARNING {label}: a true residual is above the target")
        return dt, st

    M32, Smx = make(_lib.FP32, bs, True)
    dtmx, stmx = timed(M32, Smx, True, a.steps, a.warmup, "mixed FP32-BiCGstab/FP64-IR")
    xmx = x.clone()
    sweep32 = M32.info()["sweep_form"]        # the autotuner's ILU apply kernel
    del M32
    M64, S64 = make(_lib.FP64, bs, False)
    dt64, st64 = timed(M64, S64, False, max(1, a.steps // 2), a.warmup, "fp64 BiCGstab")
    x64 = x.clone()
    sweep64 = M64.info()["sweep_form"]
    del M64
    ss = sys_start.astype(np.int64)
    errs = [float(torch.linalg.norm(xmx[ss[k]:ss[k + 1]] - x64[ss[k]:ss[k + 1]]) /
                  torch.linalg.norm(x64[ss[k]:ss[k + 1]])) for k in range(nsys)]
    log(f"mixed vs fp64 per-system relative solution difference: max {max(errs):.2e}")
    exact = None
    if a.ens_exact:
        Mx, Sx = make(_lib.FP64, E.system_blocks(sys_start, 0), False)
        dtx, stx = timed(Mx, Sx, False, 1, 0, "fp64 BiCGstab + exact ILU(0) per system")
        exact = {"ms_per_ensemble": dtx * 1e3, "iterations": [s.iterations for s in stx],
                 "max_true_rel_residual": max(s.true_residual / s.b_norm for s in stx)}
        del Mx

    # batched SpMV (CompRowMatrix.mult over the whole block-diagonal matrix)
    xin = torch.randn(n, dtype=torch.float64, device=dev)
    yout = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        lib.rsgpu_csr_mult(A.h, xin.data_ptr(), yout.data_ptr())
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        lib.rsgpu_csr_mult(A.h, xin.data_ptr(), yout.data_ptr())
    e1.record(stream)
    e1.synchronize()
    t_spmv = e0.elapsed_time(e1) / reps * 1e-3
    b64 = spmv_bytes(n, nnz, 8, 8)
    gbs = b64 / t_spmv / 1e9
    info = Smx.info()
    info["ilu_sweep_fp32"], info["ilu_sweep_fp64"] = sweep32, sweep64
    log(f"batched SpMV fp64 {t_spmv * 1e6:.0f} us {gbs:.0f} GB/s ({gbs / PEAK_HBM_GBS:.1%}); "
        f"autotuned stages fp64 {info['spmv_stage_fp64']} fp32 {info['spmv_stage_fp32']}, "
        f"ILU apply kernels fp32 {sweep32} fp64 {sweep64}")
    ctx.set_stream(None)

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline_ensemble(A, prs, sys_start, stx if exact else st64,
                                        exact is not None)
        except Exception as e:  # the baseline must not hide the GPU result
            log(f"cpu baseline failed: {e!r}")
    out = {
        "metric": "ensemble solves/s, 64 uneven MPFA-O pressure systems, BiCGstab + ILU(0) "
                  "to 1e-8, FP32 inner / FP64 refinement",
        "value": 1.0 / dtmx, "unit": "ensembles/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dtmx * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement",
        "data": "synthetic (seeded edges and log-normal permeability, MPFA-O assembled on the "
                "GPU)",
        "config": {"workload": f"{nsys} cubic MPFA-O systems, edges {a.ens_emin}..{a.ens_emax}, "
                               f"n={n}, nnz={nnz}, BiCGstab + block ILU(0) of {a.ens_rows} rows, "
                               f"rel residual {a.rtol:g} per system",
                   "n": n, "nnz": nnz, "systems": nsys, "rows_per_block": a.ens_rows},
        "systems_per_s": nsys / dtmx,
        "mixed": {"ms_per_ensemble": dtmx * 1e3,
                  "inner_iterations_max": max(s.inner_iterations for s in stmx),
                  "outer_iterations_max": max(s.outer_iterations for s in stmx),
                  "max_true_rel_residual": max(s.true_residual / s.b_norm for s in stmx),
                  "max_rel_err_vs_fp64": max(errs)},
        "fp64": {"ms_per_ensemble": dt64 * 1e3, "ensembles_per_s": 1.0 / dt64,
                 "iterations_max": max(s.iterations for s in st64),
                 "max_true_rel_residual": max(s.true_residual / s.b_norm for s in st64)},
        "fp64_exact_ilu": exact,
        "autotune": info,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": gbs / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "k_spmv<double,EpiMult> over the block-diagonal batch",
                     "algorithmic_bytes": b64},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_c2_locked(a):
    """The 128x128x64 pressure system in its incompressible form: no
    accumulation diagonal, the centre cell locked by Matrices.zeroRows
    (LinearSolver.java:256-260).  It is near-singular (the oracle needs
    18,228 GMRES(30) iterations to 1e-10).  One FP64 GMRES(30) + block
    ILU(0) x256 solve to true relative residual --rtol (monitor calibrated
    as in the main bench), and one mixed solve (FP32 GMRES(30) inner to
    --inner-rtol, FP64 refinement on the true residual, at most
    --locked-refinements refinements).  Prints one JSON line."""
    import torch
    from ressim_amd import _lib, mtj, problems as P
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    pr = P.config2(nx=a.nx, ny=a.ny, nz=a.nz, accumulation=0.0)
    pr.pinned = [pr.centre]
    n = pr.n
    ctx = mtj.Context.default()
    stream = torch.cuda.Stream()          # default priority
    ctx.set_stream(stream.cuda_stream)
    A, fails = mtj.mpfa_o_assemble(pr.dim, pr.nx, pr.ny, pr.nz, pr.dx, pr.dy, pr.dz, pr.K6,
                                   ctx=ctx)
    mtj.Matrices.zeroRows(A, 1.0, pr.pinned)
    b = torch.tensor(pr.rhs(), device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    log(f"locked-cell {pr.name}: n={n} nnz={A.nnz}, centre cell {pr.centre} locked")
    M64 = mtj.BlockILU(a.nblocks, precision=_lib.FP64, ctx=ctx)
    g64 = mtj.GMRES(b, 30)
    g64.cgs = True
    g64.setPreconditioner(M64)

    def fp64(rtol):
        g64.setIterationMonitor(mtj.DefaultIterationMonitor(200000, rtol, 1e-50, 1e5))
        torch.cuda.synchronize()
        t = time.perf_counter()
        M64.setMatrix(A)
        x.zero_()
        g64.solve(A, b, x)
        torch.cuda.synchronize()
        return time.perf_counter() - t, g64.last_stats

    mon = a.rtol
    for _ in range(4):
        dt64, st64 = fp64(mon)
        ratio = (st64.true_residual / st64.b_norm) / a.rtol
        log(f"fp64 GMRES(30)+BJ-ILU0 (monitor rtol {mon:.2e}): {dt64:.2f} s, "
            f"{st64.iterations} iterations, true rel res "
            f"{st64.true_residual / st64.b_norm:.3e}")
        if ratio <= 1.0:
            break
        mon /= 1.25 * ratio
    x64 = x.clone()
    M32 = mtj.BlockILU(a.nblocks, precision=_lib.FP32, ctx=ctx)
    gmx = mtj.MixedPrecisionSolver(b, inner="gmres", restart=30, inner_rtol=a.inner_rtol,
                                   inner_max_iter=600, cgs=True)
    gmx.setIterationMonitor(mtj.DefaultIterationMonitor(a.locked_refinements, a.rtol, 1e-50,
                                                        1e5))
    gmx.setPreconditioner(M32)
    torch.cuda.synchronize()
    t = time.perf_counter()
    A.sync_f32()
    M32.setMatrix(A)
    x.zero_()
    converged = True
    try:
        gmx.solve(A, b, x)
    except mtj.IterativeSolverNotConvergedException as e:
        converged = False
        log(f"mixed: not converged ({e})")
    torch.cuda.synchronize()
    dtmx = time.perf_counter() - t
    stmx = gmx.last_stats
    err = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
    log(f"mixed FP32-GMRES/FP64-IR: {dtmx:.2f} s, {stmx.outer_iterations} refinements, "
        f"{stmx.inner_iterations} inner iterations, true rel res "
        f"{stmx.true_residual / stmx.b_norm:.3e}, vs fp64 {err:.2e}")
    out = {
        "metric": "128x128x64 incompressible (locked centre cell) pressure solves/s, "
                  "GMRES(30) + block ILU(0) x256 to 1e-8",
        "value": 1.0 / dt64, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": dt64 * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded log-normal permeability, assembled on the GPU)",
        "config": {"workload": f"C2 {pr.nx}x{pr.ny}x{pr.nz} MPFA-O, no accumulation, centre "
                               f"cell locked, GMRES(30) + block ILU(0) x{a.nblocks}",
                   "n": n, "nnz": A.nnz},
        "fp64": {"s_per_solve": dt64, "iterations": st64.iterations,
                 "true_rel_residual": st64.true_residual / st64.b_norm,
                 "monitor_rtol": mon},
        "mixed": {"s_per_solve": dtmx, "converged": converged,
                  "refinements": stmx.outer_iterations,
                  "inner_iterations": stmx.inner_iterations,
                  "true_rel_residual": stmx.true_residual / stmx.b_norm,
                  "rel_err_vs_fp64": err},
    }
    print(json.dumps(out), flush=True)


def bench_blackoil(a):
    """Fully-implicit black-oil Newton step on 100 x 100 x 50 cells (3
    unknowns per cell, 3x3 blocks on the 7-point pattern, 3,460,000 blocks):
    the Jacobian and residual assembled on the GPU at p = 200 bar + water
    hydrostatic, Sw, Sg uniform in [0.1, 0.4], log-normal permeability
    (sigma 1.5), Corey exponents 2.  A step = setMatrix (block-Jacobi ILU(0)
    factorisation over --bo-lines x-lines of one plane) + BiCGstab from
    x0 = 0 to true relative residual 1e-8.  value = solves/s of the mixed
    form: FP64 refinement on the true residual around FP32 BiCGstab with the
    FP32 matrix and FP32 block ILU(0) (an FP32 preconditioner inside FP64
    BiCGstab stalls: the rounding of the FP32 sweeps makes the operator
    slightly non-linear, which BiCGstab's recurrences do not tolerate); the
    all-FP64 solve is reported beside it with the error between the two."""
    import torch
    from ressim_amd import _lib, mtj, problems as P
    from ressim_amd.blackoil import BlackOilJacobian
    from ressim_amd.ensemble import grid_blocks
    lib = _lib.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    ctx = mtj.Context.default()
    stream = torch.cuda.Stream()          # default priority
    ctx.set_stream(stream.cuda_stream)
    pr = P.config3(nz=a.bo_nz)
    t0 = time.time()
    bo = BlackOilJacobian(pr, ctx)
    J = bo.matrix()
    u = torch.tensor(pr.u, device=dev)
    u_old = torch.tensor(pr.u_old, device=dev)
    R = torch.empty(3 * pr.n, dtype=torch.float64, device=dev)
    bo.assemble(u, u_old, J, R)
    b = -R
    n, nnzb = J.n, pr.nnzb
    log(f"black-oil {pr.name}: {pr.n} cells, n={n}, block nnz={nnzb}, set up in "
        f"{time.time() - t0:.1f}s")
    if a.probe == "spmv":
        # kernel-only run for rocprofv3: block SpMV FP64 and FP32, then the
        # 512 MiB copy that calibrates FETCH_SIZE / WRITE_SIZE
        xin = torch.randn(n, dtype=torch.float64, device=dev)
        y = torch.empty(n, dtype=torch.float64, device=dev)
        J.sync_f32()
        x32 = xin.float()
        y32 = torch.empty(n, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        for _ in range(a.spmv_reps):
            lib.rsgpu_csr_mult(J.h, xin.data_ptr(), y.data_ptr())
        for _ in range(a.spmv_reps):
            lib.rsgpu_csr_mult_f32(J.h, x32.data_ptr(), y32.data_ptr())
        J.sync_sliced()
        for _ in range(a.spmv_reps):
            lib.rsgpu_csr_mult_sliced(J.h, xin.data_ptr(), y.data_ptr())
        for _ in range(a.spmv_reps):
            lib.rsgpu_csr_mult_sliced_f32(J.h, x32.data_ptr(), y32.data_ptr())
        torch.cuda.synchronize()
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)     # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ctx.set_stream(None)
        log(f"probe spmv: block rows {pr.n}, block nnz {nnzb}, reps {a.spmv_reps} each of the "
            f"block-row core and the sliced block form, FP64 then FP32; copy 512 MiB x3")
        return
    if a.probe == "ilu":
        # kernel-only run for rocprofv3: the block-Jacobi ILU(0) apply (16
        # x-lines of one plane per block) in FP64 and FP32, then the 512 MiB
        # copy that calibrates FETCH_SIZE / WRITE_SIZE
        bs = 3 * grid_blocks([pr], a.bo_lines * pr.nx)
        J.sync_f32()
        zin = torch.randn(n, dtype=torch.float64, device=dev)
        z = torch.empty_like(zin)
        z32in, z32 = zin.float(), torch.empty(n, dtype=torch.float32, device=dev)
        for prec in (_lib.FP64, _lib.FP32):
            M = mtj.BlockILU(1, precision=prec, ctx=ctx)
            M.setBlocks(bs)
            M.setMatrix(J)
            torch.cuda.synchronize()
            for _ in range(a.spmv_reps):
                if prec == _lib.FP64:
                    lib.rsgpu_prec_apply(M.h, zin.data_ptr(), z.data_ptr())
                else:
                    lib.rsgpu_prec_apply_f32(M.h, z32in.data_ptr(), z32.data_ptr())
            torch.cuda.synchronize()
            log(f"probe ilu: {'fp64' if prec == _lib.FP64 else 'fp32'} sweep form "
                f"{M.info()['sweep_form']}, {len(bs) - 1} blocks, reps {a.spmv_reps}")
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)     # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ctx.set_stream(None)
        return
    # assembly kernel time
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        lib.rsgpu_blackoil_assemble(bo.h, u.data_ptr(), u_old.data_ptr(), J.h, R.data_ptr())
    e1.record(stream)
    e1.synchronize()
    t_asm = e0.elapsed_time(e1) / 5 * 1e-3
    bs = 3 * grid_blocks([pr], a.bo_lines * pr.nx)
    x = torch.zeros(n, dtype=torch.float64, device=dev)

    def make(prec):
        M = mtj.BlockILU(1, precision=prec, ctx=ctx)
        M.setBlocks(bs)
        if prec == _lib.FP32:
            J.sync_f32()
            s = mtj.MixedPrecisionSolver(b, inner="bicgstab", inner_rtol=a.inner_rtol,
                                         inner_max_iter=2000)
            s.setIterationMonitor(mtj.DefaultIterationMonitor(200, a.rtol, 1e-300, 1e5))
        else:
            s = mtj.BiCGstab(b)
            s.setIterationMonitor(mtj.DefaultIterationMonitor(20000, a.rtol, 1e-300, 1e5))
        M.setMatrix(J)
        s.setPreconditioner(M)
        return M, s

    def step(M, s):
        if M.precision == _lib.FP32:
            J.sync_f32()
        M.setMatrix(J)
        x.zero_()
        s.solve(J, b, x)
        return s.last_stats

    def timed(M, s, steps, warmup, label):
        for _ in range(warmup):
            st = step(M, s)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            st = step(M, s)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        log(f"{label}: {dt * 1e3:.2f} ms/solve iters={st.iterations} "
            f"inner={st.inner_iterations} true_rel_res={st.true_residual / st.b_norm:.3e}")
        return dt, st

    M64, s64 = make(_lib.FP64)
    dt64, st64 = timed(M64, s64, max(1, a.steps), a.warmup, "fp64 BiCGstab + block ILU(0)")
    x64 = x.clone()
    M32, s32 = make(_lib.FP32)
    dtmx, stmx = timed(M32, s32, a.steps, a.warmup,
                       "fp64 refinement / fp32 BiCGstab + fp32 block ILU(0)")
    err = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
    log(f"mixed vs fp64 relative solution difference {err:.3e}")

    # Block SpMV kernels on the same matrix, each timed warm (the same
    # operands every launch) and cold (launches cycle over J and two more
    # copies with their own vectors: 3 x 152 MB in FP32 and 3 x 289 MB in
    # FP64 exceed the 256 MB last-level cache): the block-row core
    # (rsgpu_csr_mult), the block form of the sliced copy
    # (rsgpu_csr_mult_sliced, what the solves stream) and the CSR kernel on
    # the expanded matrix.
    NC = 3
    Js = [J]
    for _ in range(NC - 1):
        boc = BlackOilJacobian(pr, ctx)
        Jn = boc.matrix()
        boc.assemble(u, u_old, Jn, torch.empty_like(R))
        Js.append(Jn)
    Jcs = [mtj.CompRowMatrix.from_csr(n, n, J.getRowPointers(), J.getColumnIndices(),
                                      J.getData(), ctx=ctx) for _ in range(NC)]
    for Am in Js + Jcs:
        Am.sync_f32()
    for Am in Js:
        Am.sync_sliced()
    xs = [torch.randn(n, dtype=torch.float64, device=dev) for _ in range(NC)]
    ys = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(NC)]
    xs32 = [v.float() for v in xs]
    ys32 = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(NC)]
    torch.cuda.synchronize()

    def launches(fs, reps=60):
        for f in fs:
            f()
        e0.record(stream)
        for r in range(reps):
            fs[r % len(fs)]()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    kern = {"block": (lib.rsgpu_csr_mult, lib.rsgpu_csr_mult_f32, Js),
            "sliced": (lib.rsgpu_csr_mult_sliced, lib.rsgpu_csr_mult_sliced_f32, Js),
            "csr": (lib.rsgpu_csr_mult, lib.rsgpu_csr_mult_f32, Jcs)}
    tk = {}
    for k, (f64, f32, mats) in kern.items():
        for prec, fn, xv, yv in (("fp64", f64, xs, ys), ("fp32", f32, xs32, ys32)):
            fs = [(lambda Am, xx, yy: lambda: fn(Am.h, xx.data_ptr(), yy.data_ptr()))(
                mats[c], xv[c], yv[c]) for c in range(NC)]
            tk[f"{k}_{prec}"] = {"warm_us": launches(fs[:1]) * 1e6,
                                 "cold_us": launches(fs) * 1e6}
    # results of the three kernels on the same input
    outs = []
    for fn, Am in ((lib.rsgpu_csr_mult, J), (lib.rsgpu_csr_mult_sliced, J),
                   (lib.rsgpu_csr_mult, Jcs[0])):
        fn(Am.h, xs[0].data_ptr(), ys[0].data_ptr())
        torch.cuda.synchronize()
        outs.append(ys[0].clone())
    same = bool(torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2]))
    bsr_bytes = nnzb * (72 + 4) + (pr.n + 1) * 4 + 2 * n * 8
    bsr32_bytes = nnzb * (36 + 4) + (pr.n + 1) * 4 + 2 * n * 4
    nch, nslots, _ = J.sliced_info()
    sl_bytes = nslots * 8 + 2 * n + 2 * n * 8
    csr_bytes = spmv_bytes(n, 9 * nnzb, 8, 8)
    for k, v in tk.items():
        byt = bsr_bytes if k.endswith("fp64") else bsr32_bytes
        v["cold_GBps"] = byt / v["cold_us"] / 1e3
        v["cold_frac"] = v["cold_GBps"] / PEAK_HBM_GBS
        log(f"{k}: warm {v['warm_us']:.1f} us, cold {v['cold_us']:.1f} us = "
            f"{v['cold_GBps']:.0f} GB/s ({v['cold_frac']:.1%}) of the BSR bytes")
    best = min(("block_fp64", "sliced_fp64"), key=lambda k: tk[k]["cold_us"])
    t_b = tk[best]["cold_us"] * 1e-6
    gbs = tk[best]["cold_GBps"]
    log(f"block SpMV ({best}) cold {t_b * 1e6:.1f} us {gbs:.0f} GB/s "
        f"({gbs / PEAK_HBM_GBS:.1%}) of {bsr_bytes} BSR bytes (sliced copy's own bytes "
        f"{sl_bytes}); CSR kernel cold {tk['csr_fp64']['cold_us']:.1f} us "
        f"({csr_bytes / tk['csr_fp64']['cold_us'] / 1e3:.0f} GB/s of its bytes); results "
        f"identical: {same}; assembly {t_asm * 1e3:.2f} ms")
    ctx.set_stream(None)
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline_blackoil(J, b, st64.iterations)
        except Exception as e:  # the baseline must not hide the GPU result
            log(f"cpu baseline failed: {e!r}")
    out = {
        "metric": "black-oil Newton-step solves/s, 100x100x50 3x3 BSR, BiCGstab + block "
                  "ILU(0) to 1e-8, FP64 vs FP32 inner with FP64 refinement",
        "value": 1.0 / dtmx, "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dtmx * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement",
        "data": "synthetic (seeded saturations and log-normal permeability, Jacobian "
                "assembled on the GPU)",
        "config": {"workload": f"fully-implicit black-oil Jacobian {pr.nx}x{pr.ny}x{pr.nz}, "
                               f"3x3 blocks, n={n}, block nnz={nnzb}, BiCGstab + block-Jacobi "
                               f"ILU(0) of {a.bo_lines} x-lines, rel residual {a.rtol:g}",
                   "n": n, "nnzb": nnzb, "ilu_blocks": len(bs) - 1},
        "fp64": {"solves_per_s": 1.0 / dt64, "ms_per_solve": dt64 * 1e3,
                 "iterations": st64.iterations,
                 "true_rel_residual": st64.true_residual / st64.b_norm},
        "mixed": {"solves_per_s": 1.0 / dtmx, "ms_per_solve": dtmx * 1e3,
                  "outer_iterations": stmx.outer_iterations,
                  "inner_iterations": stmx.inner_iterations,
                  "true_rel_residual": stmx.true_residual / stmx.b_norm,
                  "rel_err_vs_fp64": err},
        "assembly_ms": t_asm * 1e3,
        "spmv": {"kernels": tk, "bitwise_equal": same, "bsr_bytes": bsr_bytes,
                 "bsr32_bytes": bsr32_bytes, "sliced_bytes": sl_bytes, "csr_bytes": csr_bytes},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": gbs / PEAK_HBM_GBS,
                     "traffic": measured_traffic(
                         pr, ("k_spmv_sliced<double, true>" if best.startswith("sliced")
                              else "k_spmv<double, EpiMult, B3=true>"),
                         BO_PMC_PROFILE, shape=(100, 100, 50)),
                     "kernel": ("k_spmv_sliced<double,true> (block form of the sliced copy)"
                                if best.startswith("sliced")
                                else "k_spmv<double,EpiMult,B3> (3x3 block rows)"),
                     "algorithmic_bytes": bsr_bytes, "cold_us": t_b * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_blackoil(J, b, gpu_iters):
    """Oracle BiCGstab + block-Jacobi ILU(0) on the expanded CSR (one
    thread), setup + 2 and setup + 12 iterations, extrapolated to the GPU
    FP64 iteration count."""
    from oracle import pyoracle as po
    rp, ci, val = J.getRowPointers(), J.getColumnIndices(), J.getData()
    bh = b.cpu().numpy()
    nbl = max(1, J.n // 9600)
    ts = []
    for k in (2, 12):
        t = time.perf_counter()
        po.solve("bicgstab", rp, ci, val, bh, rtol=1e-30, atol=0.0, max_iter=k,
                 prec="block_ilu", nblocks=nbl)
        ts.append(time.perf_counter() - t)
    per = (ts[1] - ts[0]) / 10
    setup = max(0.0, ts[0] - 2 * per)
    total = setup + per * gpu_iters
    log(f"cpu oracle: setup {setup:.2f}s, {per * 1e3:.1f} ms/iteration, {gpu_iters} iterations "
        f"-> {total:.1f}s per solve")
    return {"value": 1.0 / total, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle BiCGstab + block ILU(0) x{nbl} on the same expanded matrix: setup"
                      f" + 2 and setup + 12 iterations timed, extrapolated to the GPU FP64 "
                      f"iteration count {gpu_iters}"}


def cpu_baseline_ensemble(A, prs, sys_start, st_gpu, exact):
    """Oracle BiCGstab + ILU(0) (MTJ's per-system solve, one thread) on a
    bounded sample: the two smallest systems solved completely, and setup + 3
    iterations on the largest.  Ensemble time = sum over systems of setup +
    iterations x per-iteration time, both scaled by nonzeros from the sample,
    with each system's iteration count taken from the GPU FP64 batch (exact
    per-system ILU(0) when it ran)."""
    from oracle import pyoracle as po
    from ressim_amd import ensemble as E
    order = np.argsort([pr.n for pr in prs])
    t_setup_nnz, t_iter_nnz, full = [], [], []
    for k in list(order[:2]) + [order[-1]]:
        pr = prs[k]
        # the system's matrix as the batch holds it (device MPFA-O assembly,
        # bit-identical to the oracle's)
        Ak, _, bk, _ = E.assemble_ensemble([pr], A.ctx)
        S = {"rowptr": Ak.getRowPointers(), "colind": Ak.getColumnIndices(),
             "data": Ak.getData(), "b": bk}
        del Ak
        nz = len(S["colind"])
        t = time.perf_counter()
        po.solve("bicgstab", S["rowptr"], S["colind"], S["data"], S["b"], rtol=1e-8,
                 max_iter=0, prec="ilu")
        t0 = time.perf_counter() - t
        if k == order[-1]:
            t = time.perf_counter()
            _, so, _ = po.solve("bicgstab", S["rowptr"], S["colind"], S["data"], S["b"],
                                rtol=1e-8, max_iter=3, prec="ilu")
            t3 = time.perf_counter() - t
            t_iter_nnz.append((t3 - t0) / 3 / nz)
        else:
            t = time.perf_counter()
            _, so, _ = po.solve("bicgstab", S["rowptr"], S["colind"], S["data"], S["b"],
                                rtol=1e-8, prec="ilu")
            tf = time.perf_counter() - t
            full.append((int(k), so.iterations, tf))
            t_iter_nnz.append((tf - t0) / max(1, so.iterations) / nz)
        t_setup_nnz.append(t0 / nz)
    cs, ci = float(np.mean(t_setup_nnz)), float(np.median(t_iter_nnz))
    rp = A.getRowPointers()
    total = 0.0
    for k, pr in enumerate(prs):
        nz = int(rp[sys_start[k + 1]] - rp[sys_start[k]])
        total += cs * nz + ci * nz * st_gpu[k].iterations
    log(f"cpu oracle: full solves {full}; per-nonzero setup {cs * 1e9:.1f} ns, "
        f"per-iteration {ci * 1e9:.1f} ns; ensemble estimate {total:.1f}s")
    return {"value": 1.0 / total, "unit": "ensembles/s", "cores": 1, "kind": "port",
            "sample": f"oracle BiCGstab + ILU(0) per system: systems {full[0][0]} and "
                      f"{full[1][0]} solved completely, setup + 3 iterations on the largest; "
                      f"ensemble time {total:.1f}s extrapolated by nonzeros with the GPU "
                      f"{'exact-ILU(0)' if exact else 'block-ILU(0)'} FP64 per-system "
                      f"iteration counts"}


def probe(a, A, n, nnz, lib, mtj, _lib, ctx, stream, dev):
    """Kernel-only runs for rocprofv3: the named kernel on the workload, then
    a plain device copy of 512 MiB (read 512 MiB + write 512 MiB) whose
    FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    import torch
    xin = torch.randn(n, dtype=torch.float64, device=dev)
    y = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    reps = a.spmv_reps
    if a.probe == "spmv":
        for _ in range(reps):
            lib.rsgpu_csr_mult(A.h, xin.data_ptr(), y.data_ptr())
        A.sync_f32()
        x32 = xin.float()
        y32 = torch.empty(n, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        for _ in range(reps):
            lib.rsgpu_csr_mult_f32(A.h, x32.data_ptr(), y32.data_ptr())
        if A.sliced_info()[0]:
            assert lib.rsgpu_csr_sync_sliced(A.h) == 0
            for _ in range(reps):
                lib.rsgpu_csr_mult_sliced(A.h, xin.data_ptr(), y.data_ptr())
            for _ in range(reps):
                lib.rsgpu_csr_mult_sliced_f32(A.h, x32.data_ptr(), y32.data_ptr())
    elif a.probe == "mixed":
        A.sync_f32()
        M = mtj.BlockILU(a.nblocks, precision=_lib.FP32, ctx=ctx)
        M.setMatrix(A)
        g = mtj.MixedPrecisionSolver(xin, inner="gmres", restart=a.restart,
                                     inner_rtol=a.inner_rtol, inner_max_iter=600,
                                     cgs=not a.mgs)
        g.setIterationMonitor(mtj.DefaultIterationMonitor(200, a.rtol, 1e-50, 1e5))
        g.setPreconditioner(M)
        bb = torch.ones(n, dtype=torch.float64, device=dev)
        y.zero_()
        g.solve(A, bb, y)
        log(f"probe mixed: inner iterations {g.last_stats.inner_iterations}")
    elif a.probe == "ilu":
        M = mtj.BlockILU(a.nblocks, precision=_lib.FP64, ctx=ctx)
        M.setMatrix(A)
        for _ in range(max(5, reps // 5)):
            lib.rsgpu_prec_apply(M.h, xin.data_ptr(), y.data_ptr())
    elif a.probe in ("amg", "amg32"):
        prec = _lib.FP32 if a.probe == "amg32" else _lib.FP64
        M = mtj.AMG(precision=prec, zero_guess=1, coarsen_ratio=a.amg_ratio, max_coarse=8192,
                    ctx=ctx)
        M.setMatrix(A)
        torch.cuda.synchronize()
        if prec == _lib.FP32:
            x32 = xin.float()
            y32 = torch.zeros(n, dtype=torch.float32, device=dev)
            for _ in range(5):
                lib.rsgpu_prec_apply_f32(M.h, x32.data_ptr(), y32.data_ptr())
        else:
            for _ in range(5):
                lib.rsgpu_prec_apply(M.h, xin.data_ptr(), y.data_ptr())
        log(f"probe {a.probe}: levels {M.levels()}")
    elif a.probe == "fused":
        # the fused GMRES step at the bench's column, FP32 (the mixed inner
        # solve's) and FP64
        A.sync_f32()
        ldv = (n + 3) // 4 * 4
        for prec, dt, f in ((_lib.FP32, torch.float32, lib.rsgpu_gmres_fused_step_f32),
                            (_lib.FP64, torch.float64, lib.rsgpu_gmres_fused_step)):
            M = mtj.BlockILU(a.nblocks, precision=prec, ctx=ctx)
            M.setMatrix(A)
            g = mtj.GMRES(xin, a.restart)
            g.cgs = True
            h = g._handle(ctx)
            V = torch.randn(FUSED_I + 1, ldv, dtype=dt, device=dev)
            w = torch.empty(n, dtype=dt, device=dev)
            pj = torch.empty(FUSED_I + 1, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            for _ in range(max(5, reps // 5)):
                assert f(h, A.h, M.h, V.data_ptr(), ldv, FUSED_I, w.data_ptr(),
                         pj.data_ptr()) == 0
            torch.cuda.synchronize()
    elif a.probe == "ilu32":
        M = mtj.BlockILU(a.nblocks, precision=_lib.FP32, ctx=ctx)
        M.setMatrix(A)
        x32 = xin.float()
        y32 = torch.empty(n, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        for _ in range(max(5, reps // 5)):
            lib.rsgpu_prec_apply_f32(M.h, x32.data_ptr(), y32.data_ptr())
    else:
        M = mtj.BlockILU(a.nblocks, precision=_lib.FP64, ctx=ctx)
        g = mtj.GMRES(xin, 30)
        g.cgs = not a.fp64_mgs
        g.setIterationMonitor(mtj.DefaultIterationMonitor(60, 1e-30, 1e-50, 1e5))
        g.setPreconditioner(M)
        M.setMatrix(A)
        bb = torch.ones(n, dtype=torch.float64, device=dev)
        y.zero_()
        try:
            g.solve(A, bb, y)
        except mtj.IterativeSolverNotConvergedException:
            pass
    torch.cuda.synchronize()
    src = torch.ones(64 << 20, dtype=torch.float64, device=dev)     # 512 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    log(f"probe {a.probe}: n={n} nnz={nnz} reps={reps}; copy 512 MiB x3")


def bench_amg(a, A, b, n, dev):
    """GMRES(30) + smoothed-aggregation AMG (AMG.java with the multicolour
    SSOR smoother, zero initial guess per cycle) on the same system, FP64
    and mixed (FP32 cycle with the FP64 coarsest solve inside FP64
    refinement), with coarsening stopped where aggregates keep more than
    --amg-ratio of a level's rows; and the cycle time of MTJ's own full
    hierarchy.  Returns (results, GPU solutions by leg)."""
    import torch
    from ressim_amd import _lib, mtj
    res, xs = {}, {}
    p = dict(zero_guess=1, coarsen_ratio=a.amg_ratio, max_coarse=8192)

    def cycle_ms(M, reps=5):
        xa = torch.zeros(n, dtype=torch.float64, device=dev)
        M.apply(b, xa)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            M.apply(b, xa)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps * 1e3

    for leg, prec in (("fp64", _lib.FP64), ("mixed", _lib.FP32)):
        M = mtj.AMG(precision=prec, **p)
        M.setMatrix(A)
        torch.cuda.synchronize()
        t = time.perf_counter()
        M.setMatrix(A)
        torch.cuda.synchronize()
        setup = (time.perf_counter() - t) * 1e3
        cyc = cycle_ms(M)
        if prec == _lib.FP64:
            S = mtj.GMRES(b, 30)
            S.cgs = True
            S.setIterationMonitor(mtj.DefaultIterationMonitor(2000, a.rtol, 1e-50, 1e5))
        else:
            S = mtj.MixedPrecisionSolver(b, inner="gmres", restart=30, inner_rtol=a.inner_rtol,
                                         inner_max_iter=600, cgs=True)
            S.setIterationMonitor(mtj.DefaultIterationMonitor(200, a.rtol, 1e-50, 1e5))
        S.setPreconditioner(M)
        xa = torch.zeros(n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t = time.perf_counter()
        S.solve(A, b, xa)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        ax = torch.empty_like(xa)
        A.mult(xa, ax)
        rel = float(torch.linalg.norm(b - ax) / torch.linalg.norm(b))
        st = getattr(S, "last_stats", None)
        lv = M.levels()
        res[leg] = {"ms_per_solve": dt * 1e3, "setup_ms": setup, "cycle_ms": cyc,
                    "levels": len(lv), "level_rows": [l[0] for l in lv],
                    "iterations": S.getIterationMonitor().iterations(),
                    "true_rel_residual": rel}
        if st is not None and leg == "mixed":
            res[leg]["inner_iterations"] = st.inner_iterations
        xs[leg] = xa
        log(f"AMG {leg}: {len(lv)} levels {[l[0] for l in lv]}, setup {setup:.1f} ms, cycle "
            f"{cyc:.2f} ms, GMRES(30)+AMG {res[leg]['iterations']} iterations "
            f"{dt * 1e3:.1f} ms")
        del S, M
    M = mtj.AMG(zero_guess=1)
    M.setMatrix(A)
    lv = M.levels()
    res["mtj_hierarchy"] = {"levels": len(lv), "cycle_ms": cycle_ms(M, 3),
                            "level_rows": [l[0] for l in lv]}
    log(f"AMG with MTJ's full hierarchy: {len(lv)} levels, cycle "
        f"{res['mtj_hierarchy']['cycle_ms']:.2f} ms")
    res["params"] = dict(p, smoother="multicolour SSOR (omegas of AMG())")
    return res, xs


def cpu_baseline(A, pr, a, mon_rtol, gpu_iters, x64, xmx, extra=None):
    """The oracle's FP64 GMRES(30) + block-Jacobi ILU(0) on the same matrix
    and right-hand side, with the GPU FP64 solve's monitor tolerance, solved
    completely twice on the host cores: one thread (MTJ's concurrency; the
    sequential restatement) and all cores (or_gmres_bjilu_omp: OpenMP SpMV,
    vector updates and per-block ILU(0)).  The one-thread solution is the
    FP64 oracle the GPU results' errors are measured against."""
    from oracle import pyoracle as po
    rp = A.getRowPointers()
    ci = A.getColumnIndices()
    d = A.getData()
    b = pr.rhs()
    threads, model = host_cpu()
    log(f"cpu: {threads} threads for the all-cores leg, nproc {os.cpu_count()}, model {model}")
    t = time.perf_counter()
    x1, s1, _ = po.solve("gmres", rp, ci, d, b, rtol=mon_rtol, prec="block_ilu",
                         nblocks=a.nblocks, max_iter=20000, hist_cap=16)
    t1 = time.perf_counter() - t
    t = time.perf_counter()
    xn, sn = po.gmres_bjilu_omp(rp, ci, d, b, nthreads=threads, rtol=mon_rtol,
                                nblocks=a.nblocks, max_iter=20000)
    tn = time.perf_counter() - t
    r1 = np.linalg.norm(b - po.mult(rp, ci, d, x1)) / np.linalg.norm(b)
    rn = np.linalg.norm(b - po.mult(rp, ci, d, xn)) / np.linalg.norm(b)
    nx1 = np.linalg.norm(x1)
    e64 = float(np.linalg.norm(x64.cpu().numpy() - x1) / nx1)
    emx = float(np.linalg.norm(xmx.cpu().numpy() - x1) / nx1)
    log(f"cpu oracle 1 thread: {t1:.2f}s, {s1.iterations} iterations, true rel residual "
        f"{r1:.2e}; {threads} threads: {tn:.2f}s, {sn.iterations} iterations, {rn:.2e}")
    log(f"GPU vs FP64 oracle: FP64 {e64:.2e}, mixed {emx:.2e}")
    errs = {f"rel_err_amg_{k}": float(np.linalg.norm(v.cpu().numpy() - x1) / nx1)
            for k, v in (extra or {}).items()}
    return {**errs,"value": 1.0 / tn, "unit": "solves/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(),
            "sample": f"one complete FP64 GMRES(30) + block ILU(0) x{a.nblocks} solve of the "
                      f"same system (monitor rtol {mon_rtol:.3e}, as the GPU FP64 solve) by the "
                      f"oracle's OpenMP form on {threads} threads: {tn:.2f}s, {sn.iterations} "
                      f"iterations (GPU {gpu_iters}), true rel residual {rn:.2e}",
            "one_thread": {"value": 1.0 / t1, "unit": "solves/s", "cores": 1,
                           "seconds": t1, "iterations": s1.iterations,
                           "true_rel_residual": r1},
            "all_cores_seconds": tn,
            "rel_err_fp64": e64, "rel_err_mixed": emx}


if __name__ == "__main__":
    main()
