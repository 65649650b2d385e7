#!/usr/bin/env python3
"""Benchmark of the external-mode solve at scale (CALPUV2C + CONGRAD).

Workload (one "step" = one CALPUV2C call: explicit stage, assembly, open
boundaries, min-diagonal scaling, CONGRAD to a relative residual of 1e-10,
face transports, depths, HU/HV): an irregular lake on a 1024x1024 lattice
(~696k wet cells, depths 2-40 m, dt = 10 s, a random free-surface
perturbation with sigma = 0.05 m), time-stepped so that every solve starts
from the previous step's P like the model does.

The headline method is the FP32-refined CG (FP32 coefficients and SpMV,
FP64 dot products, alpha/beta and residual refinement); the plain FP64
Jacobi-PCG is timed beside it.  Inputs are resident in HBM before the
timed region.  Accuracy: after the same number of steps both GPU paths
are compared with the FP64 CPU oracle, which is also the CPU baseline
(timed on a bounded sample of the same steps).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "efdcplus8.5_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per wet cell moved by one launch of each kernel
# (every array element the kernel must read or write counted once;
# neighbour values are reuses of data already counted).
BYTES_PER_CELL = {
    # FP64 Jacobi-PCG, compact symmetric operator
    "cg_spmv_pupdate": 72,   # r,CCCI,p_old,CCC,CCS,CCW (6x8) + LS,LN (2x4) + p,Ap (2x8)
    "cg_axpy_dots": 56,      # P,p,r,Ap,CCCI (5x8) + P,r (2x8)
    "cg_init": 64,           # FPTMP,CCC,CCS,CCW,CCCI,P (6x8) + LS,LN + r
    # FP32 inner CG on the unit-diagonal operator
    "mx_spmv_pupdate32": 32,  # r,p_old,as,aw (4x4) + LS,LN (2x4) + p,Ap (2x4)
    "mx_axpy_dots32": 24,     # y,p,r,Ap (4x4) + y,r (2x4)
    "mx_residual64": 64,      # FPTMP,CCC,CCS,CCW,P,D^-1/2 (6x8) + LS,LN + r',y (2x4)
    "mx_apply": 28,           # P, D^-1/2 (2x8), y (4), P (8)
    "mx_prep": 52,            # CCC,CCS,CCW,FPTMP (4x8) + LS + D^-1/2 (8) + as,aw (2x4)
}
CG_KERNELS_FP64 = ("cg_init", "cg_spmv_pupdate", "cg_axpy_dots",
                   "cg_persistent64")
CG_KERNELS_MIXED = ("mx_prep", "mx_residual64", "mx_spmv_pupdate32",
                    "mx_axpy_dots32", "mx_apply", "mx_persistent")


def persistent_bytes(name, iters, outer, nwet):
    """Algorithmic bytes of one persistent solve: the same per-cell traffic
    as the multi-kernel formulation of the iterations it performed (set-up,
    every iteration, and for the mixed solve every refinement step)."""
    b = BYTES_PER_CELL
    if name == "cg_persistent64":
        per = b["cg_init"] + iters * (b["cg_spmv_pupdate"] + b["cg_axpy_dots"])
    else:
        per = (b["mx_prep"] + (outer + 1) * b["mx_residual64"] + outer * b["mx_apply"] +
               iters * (b["mx_spmv_pupdate32"] + b["mx_axpy_dots32"]))
    return per * nwet


def make_ctx(case, mixed, cg_kernel=0):
    from efdc_amd import lib
    ctx = lib.Context(case.grid, cg_mode=lib.CG_MIXED if mixed else lib.CG_FP64,
                      cg_rel_tol=1e-10, RSQM=0.0, ITERM=1000, cg_kernel=cg_kernel)
    ctx.upload_case(case)
    return ctx


def run_steps(ctx, case, n0, k, stats=None):
    iters = []
    for n in range(n0, n0 + k):
        st = ctx.calpuv2c(case.dt, n)
        iters.append(st["iters"])
        if stats is not None:
            stats.append(st)
    return iters


def timed(ctx, case, n0, k):
    ctx.synchronize()
    t0 = time.perf_counter()
    iters = run_steps(ctx, case, n0, k)
    ctx.synchronize()
    return time.perf_counter() - t0, iters


# Algorithmic bytes of the reference method per wet cell (the plain 3-pass
# Jacobi-PCG on the 5-point operator with separate coefficients): FP64
# 160 B per iteration; the FP32 inner solve 88 B per iteration, an FP64
# residual with the FP64 operator 80 B per refinement step (one more than
# the inner solves), the FP32 down-conversion of the coefficients 60 B per
# solve.  The library's compact symmetric operator moves fewer bytes
# (BYTES_PER_CELL); the roofline is quoted on the method's bytes and the
# compact figure is reported beside it.
ALG_FP64_ITER, ALG_FP32_ITER, ALG_RESID64, ALG_DOWNCONV = 160, 88, 80, 60


def alg_solve_bytes(st, nwet, mixed):
    if mixed:
        return nwet * (ALG_FP32_ITER * st["iters"] + ALG_RESID64 * (st["outer"] + 1) +
                       ALG_DOWNCONV)
    return nwet * ALG_FP64_ITER * st["iters"]


def solve_roofline(ctx, case, n0, k, mixed, nwet, workload):
    """Roofline of the CG solve over k further steps: the method's bytes
    per solve (alg_solve_bytes) over the time of the solve's kernels per
    solve (HIP events on the library's stream).  One launch per solve for
    the persistent solve; the sum of the multi-kernel schedule's launches
    otherwise.  ``traffic`` = the HBM bytes per solve from the committed
    rocprofv3 counter summaries of those kernels on this workload (None
    where a summary is missing); ``hbm_gbs`` = traffic / time."""
    kernels = CG_KERNELS_MIXED if mixed else CG_KERNELS_FP64
    ctx.timers(True)
    stats = []
    run_steps(ctx, case, n0, k, stats)
    t = ctx.read_timers()
    ctx.timers(False)
    per = {name: dict(launches=cnt, avg_us=1e3 * ms / cnt, total_ms=ms)
           for name, (cnt, ms) in t.items()}
    cg = {k_: per[k_] for k_ in kernels if k_ in per}
    nsolve = len(stats)
    cg_s = sum(v["total_ms"] for v in cg.values()) * 1e-3
    alg = sum(alg_solve_bytes(st, nwet, mixed) for st in stats)
    # the compact operator's bytes, iterations performed only
    n_it = per.get("cg_iterations", {}).get("launches", 0)
    compact = 0.0
    for name, v in cg.items():
        if name in ("cg_persistent64", "mx_persistent"):
            compact += sum(persistent_bytes(name, st["iters"], st["outer"], nwet) for st in stats)
        else:
            it_k = name in ("cg_spmv_pupdate", "cg_axpy_dots", "mx_spmv_pupdate32",
                            "mx_axpy_dots32")
            compact += BYTES_PER_CELL[name] * nwet * (n_it if (it_k and n_it) else v["launches"])
    traffic = 0.0
    for name, v in cg.items():
        tr = measured_traffic(name, workload)
        if tr is None:
            traffic = None
            break
        traffic += tr * v["launches"] / nsolve
    solve_us = cg_s * 1e6 / nsolve
    names = list(cg)
    return dict(kernel=names[0] if len(names) == 1 else "CG solve: " + "+".join(names),
                achieved=alg / cg_s / 1e9, avg_us=solve_us, bytes_per_launch=alg / nsolve,
                compact_gbs=compact / cg_s / 1e9, compact_bytes_per_solve=compact / nsolve,
                traffic=traffic,
                hbm_gbs=None if traffic is None else traffic / (solve_us * 1e-6) / 1e9,
                iters_per_solve=sum(st["iters"] for st in stats) / nsolve,
                outer_per_solve=sum(st["outer"] for st in stats) / nsolve,
                kernels=per)


def host_cpus():
    """CPUs of this host as lscpu counts them (logical CPUs, cores)."""
    import subprocess
    out = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        txt = ""
    for line in txt.splitlines():
        if ":" in line:
            k_, v = line.split(":", 1)
            out[k_.strip()] = v.strip()
    try:
        cores = int(out["Core(s) per socket"]) * int(out["Socket(s)"])
    except (KeyError, ValueError):
        cores = None
    return dict(logical=int(out.get("CPU(s)", os.cpu_count() or 0)), cores=cores,
                model=out.get("Model name"),
                usable=len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None)


def transport_bench(args):
    """CALCONC at scale (--workload transport): the 2048x1024 all-wet grid,
    KC=20 sigma layers, 4 constituents (salinity, temperature, dye with a
    point source, sediment), FP32 storage with the FP64 tridiagonal solve,
    dt = 5 s.  One step = one CALCONC call (sources, upwind advection,
    2TL update, open boundaries, implicit vertical diffusion, dye decay) for
    all constituents; value = column-steps/s.  The FP64 oracle runs the same
    steps (bounded sample) for the error and the CPU baseline."""
    import numpy as np
    from efdc_amd import grids, lib

    nx, ny, kc = args.tnx, args.tny, args.tkc
    case = grids.transport_case(nx, ny, kc)
    g = case.grid
    nc = len(case.consts)
    ref_case = case.copy() if not args.no_cpu else None
    ctx = lib.Context(g, store=lib.STORE_FP32)
    ctx.upload_case(case)
    cpu = None
    ncmp = 1
    if ref_case is not None:
        from oracle import oracle as orc
        o = orc.Oracle(ref_case, nthreads=args.cpu_threads)
        t_cpu, done = 0.0, 0
        while done < 1 or (t_cpu < args.cpu_seconds and done < 20):
            t0 = time.perf_counter()
            o.calconc(niter=done + 1)
            t_cpu += time.perf_counter() - t0
            done += 1
        ncmp = done
        cpu = dict(value=g.nwet * done / t_cpu, unit="column-steps/s",
                   cores=min(args.cpu_threads, 8), host_cpus=host_cpus(), kind="port",
                   sample=f"{done} CALCONC steps of the same {nx}x{ny}x{kc} grid, "
                          f"{nc} constituents, by the FP64 C restatement (one "
                          f"thread per constituent for advection, column-parallel "
                          f"vertical diffusion), {t_cpu:.1f} s")
    for n in range(1, ncmp + 1):
        ctx.calconc(case.dt, n)
    errs = None
    if ref_case is not None:
        errs = []
        for c in range(nc):
            con, _ = ctx.get_constituent(c)
            want = ref_case.consts[c]["con"][:kc]
            errs.append(float(np.max(np.abs(con[:kc] - want)) /
                              max(1e-30, float(np.max(np.abs(want))))))
    n0 = ncmp + 1
    ctx.autotune(True)          # the first warm-up step times the launch shapes
    for n in range(n0, n0 + max(args.warmup, 1)):
        ctx.calconc(case.dt, n)
    n0 += max(args.warmup, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for n in range(n0, n0 + args.steps):
        ctx.calconc(case.dt, n)
    ctx.synchronize()
    t = time.perf_counter() - t0
    n0 += args.steps
    vol, mass = ctx.balance()
    ctx.timers(True)
    for n in range(n0, n0 + args.steps):
        ctx.calconc(case.dt, n)
    tm = ctx.read_timers()
    ctx.timers(False)
    cnt, ms = tm["tran_upwind_vdiff"]
    s_ = 4
    per_col = kc * s_ * (6 + 2 * nc) + 40
    avg_s = ms / cnt * 1e-3
    ach = per_col * g.nwet / avg_s / 1e9
    out = {
        "metric": "column-steps/s (CALCONC transport), FP32 storage",
        "value": g.nwet * args.steps / t,
        "unit": "column-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"3D transport at scale: {nx}x{ny} all wet, KC={kc}, "
                               f"{nc} constituents, dt={case.dt} s", "columns": g.nwet},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic("tran_upwind_vdiff", f"transport{nx}x{ny}x{kc}"),
                     "kernel": "tran_upwind_vdiff", "avg_us": ms / cnt * 1e3,
                     "bytes_per_launch": per_col * g.nwet},
        "cpu_baseline": cpu,
        "detail": {"autotune": ctx.tune_report(), "stream_copy_gbs": stream_copy_gbs(ctx),
                   "rel_err_vs_fp64_oracle": errs, "after_steps": ncmp,
                   "fp64_balance": {"volume": vol, "mass": [float(m) for m in mass]},
                   "kernels": {k: dict(launches=c_, avg_us=1e3 * m_ / c_)
                               for k, (c_, m_) in tm.items()}},
    }
    print(json.dumps(out))
    ctx.close()


def uvw_bytes(kc, s):
    """Algorithmic bytes per water column of the internal-mode kernels
    (every array element a kernel must read or write counted once, FP64
    internal arrays, s bytes per stored transport value, the internal
    shears formed on the device): k_uvw_col reads U1/V1 (KC planes),
    AVUI/AVVI (KS), SUB3D/SVB3D and DZC (KS, stored) and 26 (LCM) values,
    writes UHDYF/VHDXF/UHDY/VHDX/U/V (KC) and UHDYEK/VHDXEK (KS); k_uvw_w
    reads the current and old layer flows, transports and velocities (12 KC
    planes), the four split planes (KS), QSUM and DZC (KS, stored), W1 (KC),
    writes W (KS), UHDYF2/VHDXF2 (KC), the five stored averages (KC) and
    the column sums; k_uvw_hp reads DZC, writes HPK/HPKI (KC, stored) plus
    11 (LCM) values; k_uvw_uv 13 (LCM) values."""
    ks = kc - 1
    return {"uvw_columns": 64 * kc + 32 * ks + 3 * s * ks + 224,
            "uvw_w_averages": 120 * kc + 40 * ks + 2 * s * ks + 5 * s * kc + 41,
            "uvw_depths": 3 * s * kc + 88,
            "uvw_uv": 104}


def hydro3d_bench(args):
    """The full 3-D hot-path step at scale (--workload hydro3d): a tidal
    basin of 2048x1024 columns (all wet, west open boundary with an M2
    tide, depths 5-20 m, Manning drag), KC=20 sigma layers, salinity and a
    dye source, FP32 storage.  One step = the external solve (CALPUV2C with
    the FP32-refined CG to relative residual 1e-10), the internal mode
    (CALUVW with the internal shears formed on the device, then U/V
    interpolation) and CALCONC; value = column-steps/s.  The FP64 oracle
    runs the same sequence on a 256x128 tile of the same generator for the
    error and the CPU baseline."""
    import numpy as np
    from efdc_amd import grids, lib

    nx, ny, kc = args.tnx, args.tny, args.tkc

    def make(nx_, ny_):
        c = grids.tidal_basin(nx_, ny_, kc, seed=1002, dt=5.0)
        return grids.add_internal_mode(c, av=1e-3)

    def context(case):
        ctx = lib.Context(case.grid, store=lib.STORE_FP32, cg_mode=lib.CG_MIXED,
                          cg_rel_tol=1e-10, RSQM=0.0, ITERM=1000)
        return ctx.upload_case(case)

    # accuracy and CPU baseline on the bounded sample
    cpu = errs = None
    if not args.no_cpu:
        from oracle import oracle as orc
        small = make(256, 128)
        ref = small.copy()
        o = orc.Oracle(ref, rsqm=0.0, iterm=3000, rel_tol=1e-12, nthreads=args.cpu_threads)
        sctx = context(small)
        t_cpu, done = 0.0, 0
        while done < 3 or (t_cpu < args.cpu_seconds and done < 400):
            t0 = time.perf_counter()
            o.step(done + 1)
            t_cpu += time.perf_counter() - t0
            done += 1
            sctx.step(small.dt, done, internal=True)
        L = np.arange(2, small.grid.LA + 1) - 1
        eta = float(np.max(np.abs(sctx.get("P")[L] - ref.h["P"][L]))) / 9.81
        errs = dict(eta_m=eta, after_steps=done)
        for name in ("UHDY2", "VHDX2"):
            b = ref.k[name][:kc, L]
            errs[name] = float(np.max(np.abs(sctx.get(name)[:kc, L] - b)) /
                               max(1e-30, float(np.max(np.abs(b)))))
        errs["W2"] = float(np.max(np.abs(sctx.get("W2")[:kc + 1, L] - ref.w["W2"][:kc + 1, L])))
        errs["con"] = []
        for j, c in enumerate(ref.consts):
            con, _ = sctx.get_constituent(j)
            b = c["con"][:kc, L]
            errs["con"].append(float(np.max(np.abs(con[:kc, L] - b)) /
                                     max(1e-30, float(np.max(np.abs(b))))))
        sctx.close()
        cpu = dict(value=small.grid.nwet * done / t_cpu, unit="column-steps/s",
                   cores=min(args.cpu_threads, 8), host_cpus=host_cpus(), kind="port",
                   sample=f"{done} full steps (CALPUV2C with the FP64 CG, CALUVW, CALCONC) "
                          f"of a 256x128x{kc} tile of the same generator by the FP64 C "
                          f"restatement, {t_cpu:.1f} s (the CALUVW restatement runs on "
                          f"one thread)")
    case = make(nx, ny)
    g = case.grid
    nc = len(case.consts)
    ctx = context(case)
    n0 = 1
    ctx.autotune(True)
    for n in range(n0, n0 + max(args.warmup, 1)):
        ctx.step(case.dt, n, internal=True)
    n0 += max(args.warmup, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for n in range(n0, n0 + args.steps):
        ctx.step(case.dt, n, internal=True)
    ctx.synchronize()
    t = time.perf_counter() - t0
    n0 += args.steps
    ctx.timers(True)
    stats = []
    for n in range(n0, n0 + args.steps):
        stats.append(ctx.step(case.dt, n, internal=True))
    tm = ctx.read_timers()
    ctx.timers(False)
    nstep = args.steps
    per = {k: dict(launches=c_, avg_us=1e3 * m_ / c_, total_ms=m_) for k, (c_, m_) in tm.items()}
    ub = uvw_bytes(kc, 4)
    cand = {}
    for k_, b in ub.items():
        if k_ in per:
            cand[k_] = (b * g.nwet, per[k_]["avg_us"] * 1e-6, per[k_]["launches"])
    if "tran_upwind_vdiff" in per:
        b = (kc * 4 * (6 + 2 * nc) + 40) * g.nwet
        cand["tran_upwind_vdiff"] = (b, per["tran_upwind_vdiff"]["avg_us"] * 1e-6,
                                     per["tran_upwind_vdiff"]["launches"])
    cg_names = [k_ for k_ in CG_KERNELS_MIXED if k_ in per]
    if cg_names:
        cg_s = sum(per[k_]["total_ms"] for k_ in cg_names) * 1e-3 / nstep
        alg = sum(alg_solve_bytes(st, g.nwet, True) for st in stats) / nstep
        cand["CG solve: " + "+".join(cg_names)] = (alg, cg_s, nstep)
    dom = max(cand, key=lambda k_: cand[k_][1] * cand[k_][2])
    b, s_, _ = cand[dom]
    ach = b / s_ / 1e9
    out = {
        "metric": "column-steps/s (CALPUV2C + CALUVW + CALCONC), FP32-refined CG, FP32 storage",
        "value": g.nwet * args.steps / t,
        "unit": "column-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"full 3-D step at scale: {nx}x{ny} tidal basin, KC={kc}, "
                               f"internal mode, {nc} constituents, dt={case.dt} s",
                   "columns": g.nwet},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, f"hydro3d{nx}x{ny}x{kc}"),
                     "kernel": dom, "avg_us": s_ * 1e6, "bytes_per_launch": b},
        "cpu_baseline": cpu,
        "detail": {"stream_copy_gbs": stream_copy_gbs(ctx), "errors_vs_fp64_oracle": errs,
                   "per_kernel_gbs": {k_: v[0] / v[1] / 1e9 for k_, v in cand.items()},
                   "cg_iters_per_step": sum(st["iters"] for st in stats) / nstep,
                   "kernels": per},
    }
    print(json.dumps(out))
    ctx.close()


def restart_bench(args):
    """RESTART.OUT from the device (--workload restart): a tidal basin of
    --rnx x --rny columns, KC = --tkc, internal mode, salinity and a dye,
    FP64 storage, after a few full steps; one "step" = one RESTOUT of the
    whole state (efdc_restart_text: every value formatted E17.8 on the
    device into the file image, then copied to host memory).  value = bytes
    of restart text the device formats per second; the roofline counts
    the FP64 value read and the text written.  The CPU baseline is the
    oracle's writer (the C library's correctly rounded conversion, one
    thread) on a 128x64 tile."""
    import ctypes as C
    import numpy as np
    from efdc_amd import formats, grids, lib

    nx, ny, kc = args.rnx, args.rny, args.tkc

    def make(nx_, ny_):
        c = grids.tidal_basin(nx_, ny_, kc, seed=77, dt=5.0)
        return grids.add_internal_mode(c, av=1e-3)

    def host_arrays(case):
        rng = np.random.default_rng(3)
        lcm = case.grid.LCM
        hp = case.h["HP"]
        return dict(hwq=hp.copy(), h2wq=hp.copy(),
                    qq=rng.uniform(1e-8, 1e-3, (kc + 1, lcm)),
                    qq1=rng.uniform(1e-8, 1e-3, (kc + 1, lcm)),
                    qql=rng.uniform(1e-9, 1e-4, (kc + 1, lcm)),
                    qql1=rng.uniform(1e-9, 1e-4, (kc + 1, lcm)),
                    dml=rng.uniform(0.0, 1.0, (kc + 1, lcm)))

    def opts(case):
        mo = [c["mo"] for c in case.consts]
        return dict(isco=[0, 1, 0, 1, 0, 0, 0, 0, 0], istran=[0, 1, 0, 1, 0, 0, 0, 0, 0],
                    slots=(mo.index(1), -1, mo.index(3)), **host_arrays(case))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        small = make(128, 64)
        o = orc.Oracle(small, rsqm=1e-20, iterm=3000)
        for n in range(1, 4):
            o.step(n)
        kw = opts(small)
        t_cpu, done, nbytes = 0.0, 0, 0
        while done < 2 or (t_cpu < args.cpu_seconds and done < 50):
            t0 = time.perf_counter()
            nbytes += len(o.restart_text(3, 0.5, keep_nlo=True, **kw))
            t_cpu += time.perf_counter() - t0
            done += 1
        cpu = dict(value=nbytes / t_cpu, unit="restart bytes/s", cores=1, host_cpus=host_cpus(),
                   kind="port",
                   sample=f"{done} RESTOUTs of a 128x64x{kc} tile by the C restatement "
                          f"(snprintf's exact conversion), {t_cpu:.1f} s")
    case = make(nx, ny)
    g = case.grid
    ctx = lib.Context(g, RSQM=1e-20, ITERM=3000).upload_case(case)
    for n in range(1, 4):
        ctx.step(case.dt, n, internal=True)
    kw = opts(case)
    r, keep = formats.restart_opts(3, 0.5, keep_nlo=True, **kw)
    nb = C.c_int64(0)
    ctx._check(ctx.lib.efdc_restart_text(ctx.h, C.byref(r), None, 0, C.byref(nb)))
    buf = C.create_string_buffer(nb.value)
    for _ in range(max(args.warmup, 1)):
        ctx._check(ctx.lib.efdc_restart_text(ctx.h, C.byref(r), buf, nb.value, C.byref(nb)))
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx._check(ctx.lib.efdc_restart_text(ctx.h, C.byref(r), buf, nb.value, C.byref(nb)))
    t = time.perf_counter() - t0
    ctx.timers(True)
    for _ in range(args.steps):
        ctx._check(ctx.lib.efdc_restart_text(ctx.h, C.byref(r), buf, nb.value, C.byref(nb)))
    tm = ctx.read_timers()
    ctx.timers(False)
    launches, ms = tm["restart_format"]
    avg = ms * 1e-3 / launches
    text = nb.value
    nc = len(case.consts)
    nbc = sum(len(s["L"]) for s in case.conc_bc["side"] if s is not None)
    nvals = g.nwet * (5 + 4 + 4 * kc + 5 * (kc + 1) + 2 * 2 * kc + 1 + kc) + nbc * 2 * 2 * kc
    alg = text + 8 * nvals
    ach = alg / avg / 1e9
    head = bytes(buf.raw[:200]).split(b"\n")[0]
    out = {
        "metric": "RESTART.OUT bytes formatted per second on the device (E17.8, exact)",
        "value": text / avg,
        "unit": "bytes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * avg, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64->text", "data": "synthetic",
        "config": {"workload": f"RESTOUT of a {nx}x{ny} tidal basin, KC={kc}, internal mode, "
                               f"{nc} constituents", "columns": g.nwet, "text_bytes": text,
                   "values": nvals},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic("restart_format", f"restart{nx}x{ny}x{kc}"),
                     "kernel": "restart_format", "avg_us": avg * 1e6,
                     "bytes_per_launch": alg},
        "cpu_baseline": cpu,
        "detail": {"stream_copy_gbs": stream_copy_gbs(ctx),
                   "end_to_end_host_bytes_per_s": text * args.steps / t,
                   "end_to_end_ms": 1e3 * t / args.steps, "first_line": head.decode()},
    }
    print(json.dumps(out))
    del keep
    ctx.close()


def drying_bench(args):
    """Wetting and drying at scale (--workload drying): the tidal drying
    basin (bottom rising from 2 m below the datum at the open west boundary
    to 0.3 m above it at the east end, a shoal at 45% of the length) at
    1024x512, KC=4, salinity and a dye source, ISDRY=99 (drying and
    rewetting, the setting of the sample lake decks; --isdry 1 for the
    drying-only scheme), HDRY=0.2 m,
    dt = 4 s.  One step = the full hot path (closure, CALPUV2C with its
    drying re-solves, each an FP64 multi-kernel CG to a relative residual of
    1e-10, closure, CALCONC with the dry columns left out); value =
    cell-steps/s.  (Without drying the same basin runs into negative depths
    at its dry east end, so there is no ISDRY=0 run of it to compare with;
    the kernel statistics show what the re-solves cost.)  The FP64 oracle
    runs a bounded sample: the CPU baseline and the check of the dry flags
    and depths step by step."""
    import numpy as np
    from efdc_amd import grids, lib

    nx, ny, kc = args.dnx, args.dny, 4
    case = grids.drying_basin(nx, ny, kc)
    g = case.grid
    L = np.arange(2, g.LA + 1) - 1
    prm = dict(cg_rel_tol=1e-10, RSQM=0.0, ITERM=4000, cg_mode=lib.CG_FP64, cg_kernel=1,
               HDRY=0.2)
    ref_case = case.copy() if not args.no_cpu else None
    isdry = args.isdry
    ctx = lib.Context(g, ISDRY=isdry, **prm).upload_case(case)
    cpu, check = None, None
    n0 = 1
    if ref_case is not None:
        from oracle import oracle as orc
        o = orc.Oracle(ref_case, rsqm=0.0, iterm=4000, nthreads=args.cpu_threads,
                       rel_tol=1e-10, isdry=isdry, hdry=0.2)
        t_cpu, done, same_flags = 0.0, 0, True
        while done < 1 or (t_cpu < args.cpu_seconds and done < 40):
            t0 = time.perf_counter()
            st = o.step(done + 1)
            t_cpu += time.perf_counter() - t0
            if st["status"] != 0:
                raise RuntimeError(f"oracle status {st}")
            ctx.step(case.dt, done + 1)
            m, d, k = ctx.get_dry_state()
            same_flags &= bool(np.array_equal(m[L], o.lmaskdry[L]) and
                               np.array_equal(d[L], o.iscdry[L]) and k == o.o.ncordry)
            done += 1
        cpu = dict(value=g.nwet * done / t_cpu, unit="cell-steps/s",
                   cores=min(args.cpu_threads, 8), host_cpus=host_cpus(), kind="port",
                   sample=f"{done} full steps of the same {nx}x{ny}x{kc} drying basin "
                          f"by the FP64 C restatement, {t_cpu:.1f} s")
        check = dict(after_steps=done, same_dry_flags_every_step=same_flags,
                     eta_err_m=float(np.max(np.abs(ctx.get("P")[L] - ref_case.h["P"][L])) / 9.81),
                     hp_err_m=float(np.max(np.abs(ctx.get("HP")[L] - ref_case.h["HP"][L]))))
        n0 = done + 1
    for n in range(n0, n0 + args.warmup):
        ctx.step(case.dt, n)
    n0 += args.warmup
    ncor = []

    def run(c, k, count):
        c.synchronize()
        t0 = time.perf_counter()
        for n in range(n0, n0 + k):
            st = c.step(case.dt, n)
            if st["status"] != 0:
                raise RuntimeError(f"step status {st}")
            if count:
                ncor.append(c.get_dry_state(arrays=False)[2])
        c.synchronize()
        return time.perf_counter() - t0

    t = run(ctx, args.steps, True)
    m, _, _ = ctx.get_dry_state()
    n0 += args.steps
    ctx.timers(True)
    for n in range(n0, n0 + args.steps):
        ctx.step(case.dt, n)
    tm = ctx.read_timers()
    ctx.timers(False)
    dom = max((k_ for k_ in tm if k_ in BYTES_PER_CELL), key=lambda k_: tm[k_][1])
    cnt, ms = tm[dom]
    if dom in ("cg_spmv_pupdate", "cg_axpy_dots") and tm.get("cg_iterations", (0, 0))[0] > 0:
        # launches after a solve has converged return at once; count the
        # iterations the solves performed (their time includes those exits)
        cnt = tm["cg_iterations"][0]
    avg_s = ms / cnt * 1e-3
    ach = BYTES_PER_CELL[dom] * g.nwet / avg_s / 1e9
    out = {
        "metric": f"cell-steps/s (full step with wetting and drying, ISDRY={isdry})",
        "value": g.nwet * args.steps / t,
        "unit": "cell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"tidal drying basin {nx}x{ny}, KC={kc}, 2 constituents, "
                               f"ISDRY={isdry}, HDRY=0.2 m, dt={case.dt} s", "cells": g.nwet},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, f"drying{nx}x{ny}"), "kernel": dom,
                     "avg_us": ms / cnt * 1e3, "bytes_per_launch": BYTES_PER_CELL[dom] * g.nwet},
        "cpu_baseline": cpu,
        "detail": {"resolves_per_step": float(np.mean(ncor)), "resolves_max": int(max(ncor)),
                   "masked_cells": int(np.sum(m[L] == 0)), "check_vs_oracle": check,
                   "kernels": {k: dict(launches=c_, avg_us=1e3 * m_ / c_)
                               for k, (c_, m_) in tm.items()}},
    }
    print(json.dumps(out))
    ctx.close()


def ensemble_bench(args):
    """Batched ensemble (--workload ensemble): 256 independent basins,
    (nx, ny) log-uniform in 64x64 .. 512x256, irregular lake masks, depths
    2-30 m, Manning n 0.015-0.04, tidal amplitude 0.1-1.0 m on a west
    boundary segment, KC=8, salinity and a dye source each, concatenated in
    one context with a basin table.  One step = the full hot path for every
    basin (closure, CALPUV2C with a per-basin FP64 CG to a relative residual
    of 1e-10, closure, CALCONC); value = column-steps/s over all basins.
    Beside it: the same basins as separate contexts stepped one after the
    other (each with its best single-model CG), on a bounded sample of steps,
    and the FP64 oracle on the first basins (CPU baseline)."""
    import numpy as np
    from efdc_amd import grids, lib

    kc = args.ekc
    t0 = time.perf_counter()
    case, first, members = grids.config3(args.nbasins, kc)
    t_gen = time.perf_counter() - t0
    g = case.grid
    nwet = g.nwet
    nc = len(case.consts)
    prm = dict(cg_rel_tol=1e-10, RSQM=0.0, ITERM=2000, cg_mode=lib.CG_FP64)
    ctx = lib.Context(g, **prm).upload_case(case)
    ctx.set_segments(first)

    # ---- sequential single-basin contexts on a bounded sample of steps
    nseq = min(args.seq_steps, args.steps)
    singles = [lib.Context(m.grid, **prm).upload_case(m) for m in members]
    ctx.synchronize()
    t0 = time.perf_counter()
    for s_, m in zip(singles, members):
        for n in range(1, nseq + 1):
            s_.step(m.dt, n)
    for s_ in singles:
        s_.synchronize()
    t_seq = time.perf_counter() - t0
    for n in range(1, nseq + 1):
        ctx.step(case.dt, n)
    p = ctx.get("P")
    d_seq = 0.0
    for b, (s_, m) in enumerate(zip(singles, members)):
        sl = grids.basin_slice(case, first, b)
        d_seq = max(d_seq, float(np.max(np.abs(p[sl] - s_.get("P")[1:1 + m.grid.nwet]))) / 9.81)
        s_.close()
    del singles

    # ---- CPU baseline: the FP64 oracle on the first basins
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        t_cpu, cols, nb_done, st_done = 0.0, 0, 0, 0
        for m in members:
            o = orc.Oracle(m.copy(), rsqm=0.0, iterm=2000, nthreads=args.cpu_threads,
                           rel_tol=1e-10)
            for n in range(1, 4):
                t1 = time.perf_counter()
                o.step(n)
                t_cpu += time.perf_counter() - t1
                cols += m.grid.nwet
                st_done += 1
            nb_done += 1
            if t_cpu >= args.cpu_seconds:
                break
        cpu = dict(value=cols / t_cpu, unit="column-steps/s", cores=min(args.cpu_threads, 8), host_cpus=host_cpus(),
                   kind="port",
                   sample=f"3 full steps of each of the first {nb_done} basins by the FP64 C "
                          f"restatement (OpenMP), {t_cpu:.1f} s")

    # ---- timed region (the first warm-up step runs the autotuner, the CG
    # tile trials among them)
    n0 = nseq + 1
    ctx.autotune(True)
    ctx.set_segment_tile(0)
    for n in range(n0, n0 + max(args.warmup, 1)):
        ctx.step(case.dt, n)
    n0 += max(args.warmup, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    iters = []
    for n in range(n0, n0 + args.steps):
        iters.append(ctx.step(case.dt, n)["iters"])
    ctx.synchronize()
    t = time.perf_counter() - t0
    n0 += args.steps

    # ---- per-kernel timing over a bounded number of further steps
    nk = min(args.steps, 50)
    ctx.timers(True)
    cell_iters = 0
    sizes = np.array([m.grid.nwet for m in members], dtype=np.int64)
    for n in range(n0, n0 + nk):
        ctx.step(case.dt, n)
        it_b, _ = ctx.segment_stats()
        cell_iters += int(np.sum(sizes * it_b))
    tm = ctx.read_timers()
    ctx.timers(False)
    s_ = 8
    per_col = kc * s_ * (6 + 2 * nc) + 40
    total_bytes = {
        "cg_spmv_pupdate": BYTES_PER_CELL["cg_spmv_pupdate"] * cell_iters,
        "cg_axpy_dots": BYTES_PER_CELL["cg_axpy_dots"] * cell_iters,
        "tran_upwind_vdiff": per_col * nwet * nk,
    }
    cand = {k: tm[k] for k in total_bytes if k in tm}
    dom = max(cand, key=lambda k: cand[k][1])
    cnt, ms = cand[dom]
    ach = total_bytes[dom] / (ms * 1e-3) / 1e9
    vol, mass = ctx.balance_segments()
    out = {
        "metric": "column-steps/s over all basins (batched ensemble)",
        "value": nwet * args.steps / t,
        "unit": "column-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched ensemble: {args.nbasins} basins 64x64..512x256, "
                               f"KC={kc}, {nc} constituents, {nwet} wet columns, dt={case.dt} s",
                   "basins": args.nbasins, "columns": nwet},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, f"ensemble{args.nbasins}"), "kernel": dom,
                     "avg_us": ms / cnt * 1e3, "bytes_per_launch": total_bytes[dom] / cnt},
        "cpu_baseline": cpu,
        "detail": {
            "autotune": ctx.tune_report(), "cg_tile_cells": ctx.segment_tile(),
            "stream_copy_gbs": stream_copy_gbs(ctx),
            "sequential_value": int(sizes.sum()) * nseq / t_seq,
            "sequential_sample": f"{nseq} steps of each basin in its own context, one "
                                 f"basin after the other, {t_seq:.2f} s",
            "eta_diff_vs_sequential_m": d_seq,
            "cg_iters_per_step_max_basin": float(np.mean(iters)),
            "cell_iters_per_step": cell_iters / nk,
            "generate_s": t_gen,
            "fp64_balance_first_basins": {"volume": [float(v) for v in vol[:4]],
                                          "mass": [[float(x) for x in r] for r in mass[:4]]},
            "kernels": {k: dict(launches=c_, avg_us=1e3 * m_ / c_, total_ms=m_)
                        for k, (c_, m_) in tm.items()},
        },
    }
    print(json.dumps(out))
    ctx.close()


def adjoint_bench(args):
    """Adjoint sensitivity (--workload adjoint): a 512x512 basin with KC=8
    (built like the config-0 basin: west M2 tide, salinity, n = 0.025), 16
    dye sources whose loads are 200-step series in U(0,5), 50 observed wet
    cells every 10 steps, data from a seeded true forcing plus 1 % noise.
    One gradient = forward run with a checkpoint every 20 steps, then the
    checkpointed reverse sweep (segment recompute, step repeat, adjoint
    step) giving dJ/dloads (200 x 16) and dJ/dn (per cell, coarsened to
    64 x 64).  FP32 storage forward, FP64 adjoint.  value = column-steps/s
    of forward+adjoint (wet columns x steps / gradient time).  Checked
    against central differences of the forward model (FP64 storage).  The
    observed cells are drawn among the cells within 3 cells of a source:
    in 1000 s the dye does not reach cells farther away."""
    import numpy as np
    from efdc_amd import adjoint, grids, lib

    n, kc, nsteps = args.an, 8, args.asteps
    case, obs = grids.adjoint_case(n, kc, 16, 50)
    g = case.grid
    rng = np.random.default_rng(1004)
    store = lib.STORE_FP32 if args.astore == "fp32" else lib.STORE_FP64
    prob = adjoint.DyeMisfit(case, nsteps=nsteps, obs_L=obs, obs_every=10, ckpt=20,
                             store=store)
    true = rng.uniform(0.0, 5.0, (nsteps, 16))
    y = prob.observations(true)
    prob.data = y * (1.0 + 0.01 * rng.standard_normal(y.shape))
    loads0 = rng.uniform(0.0, 5.0, (nsteps, 16))
    cidx = adjoint.coarse_index(g, 64, 64)
    # warm-up gradient, then the timed ones
    prob.gradient(loads0)
    prob.ctx.synchronize()
    t0 = time.perf_counter()
    reps = max(1, args.steps // nsteps) if args.steps >= nsteps else 1
    for _ in range(reps):
        J, gl, gn, st = prob.gradient(loads0)
    prob.ctx.synchronize()
    t = (time.perf_counter() - t0) / reps
    # forward alone, for the overhead ratio
    t0 = time.perf_counter()
    prob.misfit(loads0)
    prob.ctx.synchronize()
    t_fwd = time.perf_counter() - t0
    # per-kernel time of one gradient
    prob.ctx.timers(True)
    prob.gradient(loads0)
    tm = prob.ctx.read_timers()
    prob.ctx.timers(False)
    # gradient checks: central differences of the forward model with FP64
    # storage (the FP32-storage forward rounds the concentrations to ~1e-7,
    # which a difference quotient of a 1e-3 Manning step cannot resolve)
    ref = prob if store == lib.STORE_FP64 else adjoint.DyeMisfit(
        case, nsteps=nsteps, obs_L=obs, obs_every=10, ckpt=20, store=lib.STORE_FP64)
    ref.data = prob.data
    d = rng.standard_normal(loads0.shape)
    fd = (ref.misfit(loads0 + d) - ref.misfit(loads0 - d)) / 2.0
    ad = float(np.sum(gl * d))
    m = cidx >= 0
    dn = rng.uniform(-1.0, 1.0, 64 * 64)
    n0 = case.manning_n.copy()
    h = 1e-3
    pert = np.zeros(g.LCM)
    pert[m] = dn[cidx[m]] * n0[m]
    fdn = (ref.misfit(loads0, n0 + h * pert) - ref.misfit(loads0, n0 - h * pert)) / (2 * h)
    adn = float(np.sum(gn * pert))
    if ref is not prob:
        ref.close()
    # bytes of the dominant adjoint kernel: the transport column pass reads
    # the flows, depths, layer constants and old level (8 storage-type
    # layer arrays), lambda of CON, writes lcd/own and 5 FP64 scratch
    # planes, and reads the scratch planes back (KC*(8*s + 8 + 2*8 + 10*8))
    s_ = 4 if store == lib.STORE_FP32 else 8
    per_col = {"adj_tran_col": kc * (8 * s_ + 8 + 16 + 80) + 64,
               "adj_tran_gather": kc * (4 * s_ + 3 * 8 + 8) + 64}
    cand = {k: tm[k] for k in per_col if k in tm}
    dom = max(cand, key=lambda k: cand[k][1])
    cnt, ms = cand[dom]
    ach = per_col[dom] * g.nwet * cnt / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        ref = case.copy()
        o = orc.Oracle(ref, rsqm=0.0, iterm=4000, nthreads=args.cpu_threads, rel_tol=1e-13)
        t_cpu, done = 0.0, 0
        while done < 2 or (t_cpu < args.cpu_seconds and done < 10):
            t1 = time.perf_counter()
            o.step(done + 1)
            t_cpu += time.perf_counter() - t1
            done += 1
        cpu = dict(value=g.nwet * done / t_cpu, unit="column-steps/s", cores=min(args.cpu_threads, 8), host_cpus=host_cpus(),
                   kind="port",
                   sample=f"{done} forward steps of the same {n}x{n}x{kc} basin by the FP64 C "
                          f"restatement, {t_cpu:.1f} s (the CPU path has no adjoint: forward only)")
    out = {
        "metric": "column-steps/s of forward+adjoint (dye-misfit gradient w.r.t. loads and Manning n)",
        "value": g.nwet * nsteps / t,
        "unit": "column-steps/s",
        "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": 1e3 * t / nsteps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if s_ == 4 else "f64",
        "data": "synthetic",
        "config": {"workload": f"adjoint sensitivity: {n}x{n}x{kc} basin, 16 dye sources x "
                               f"{nsteps} steps, 50 observed cells every 10 steps, "
                               "checkpoint every 20 steps", "columns": g.nwet},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, f"adjoint{n}x{n}x{kc}"), "kernel": dom,
                     "avg_us": ms / cnt * 1e3, "bytes_per_launch": per_col[dom] * g.nwet},
        "cpu_baseline": cpu,
        "detail": {
            "J": J, "gradient_s": t, "forward_s": t_fwd, "gradient_over_forward": t / t_fwd,
            "load_grad_vs_central_difference": {"adjoint": ad, "fd": fd,
                                                "rel": abs(ad - fd) / abs(fd)},
            "manning_grad_vs_central_difference": {"adjoint": adn, "fd": fdn,
                                                   "rel": abs(adn - fdn) / abs(fdn), "h": h},
            "adj_cg_iters_mean": float(np.mean(st["adj_cg_iters"])),
            "kernels": {k: dict(launches=c_, avg_us=1e3 * m_ / c_, total_ms=m_)
                        for k, (c_, m_) in tm.items()},
        },
    }
    prob.close()
    print(json.dumps(out))


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a plain copy of known size), or
    None when no summary for this kernel and workload exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith((" ", "#")):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    need = ("fetch_kib_per_dispatch", "write_kib_per_dispatch", "fetch_scale",
            "write_scale")
    if not all(k in vals for k in need):
        return None
    return 1024.0 * (vals["fetch_kib_per_dispatch"] * vals["fetch_scale"] +
                     vals["write_kib_per_dispatch"] * vals["write_scale"])


def stream_copy_gbs(ctx, mib=1024, reps=10):
    """Achievable HBM bandwidth: the library's plain streaming copy kernel
    (16-byte loads and stores) of mib MiB, best of reps, timed with HIP
    events on the context's stream."""
    return ctx.stream_copy_gbs(mib, reps)


def copy_probe(mib):
    """One device-to-device copy of exactly ``mib`` MiB (read once, written
    once): the known-size kernel that scales the HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024, help="lattice size")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("lake1024", "lake4096", "transport", "ensemble",
                                           "adjoint", "drying", "hydro3d", "restart"),
                    default="lake1024",
                    help="lake1024: the external solve at scale (the headline; --n sets "
                         "the lattice); lake4096: the same on a 4096x4096 lattice (~11 M "
                         "cells, past the MALL); transport: CALCONC at scale; ensemble: "
                         "256 basins in one context; hydro3d: the full 3-D step "
                         "with the internal mode")
    ap.add_argument("--nbasins", type=int, default=256)
    ap.add_argument("--an", type=int, default=512, help="adjoint: basin size")
    ap.add_argument("--asteps", type=int, default=200, help="adjoint: steps")
    ap.add_argument("--astore", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--ekc", type=int, default=8)
    ap.add_argument("--seq-steps", type=int, default=10,
                    help="ensemble: steps of the sequential single-basin comparison")
    ap.add_argument("--tnx", type=int, default=2048)
    ap.add_argument("--tny", type=int, default=1024)
    ap.add_argument("--tkc", type=int, default=20)
    ap.add_argument("--rnx", type=int, default=1024, help="restart: basin length")
    ap.add_argument("--rny", type=int, default=512, help="restart: basin width")
    ap.add_argument("--dnx", type=int, default=1024, help="drying: basin length")
    ap.add_argument("--dny", type=int, default=512, help="drying: basin width")
    ap.add_argument("--isdry", type=int, default=99, choices=(1, 99),
                    help="drying: 99 drying and rewetting, 1 drying only")
    ap.add_argument("--cg-kernel", choices=("auto", "multi", "persistent"),
                    default="auto", help="launch strategy of the CG solve")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only launch one plain device copy of MIB MiB and "
                         "exit (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(collectives and sharding are out of scope); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "transport":
        transport_bench(args)
        return
    if args.workload == "ensemble":
        ensemble_bench(args)
        return
    if args.workload == "adjoint":
        adjoint_bench(args)
        return
    if args.workload == "drying":
        drying_bench(args)
        return
    if args.workload == "hydro3d":
        hydro3d_bench(args)
        return
    if args.workload == "restart":
        restart_bench(args)
        return

    lake_bench(args)


def lake_bench(args):
    """The headline: CALPUV2C steps on the irregular lake at scale,
    FP32-refined and FP64 solves, inputs resident in HBM."""
    import numpy as np
    from efdc_amd import grids

    n = 4096 if args.workload == "lake4096" else args.n
    workload = f"lake{n}"
    case = grids.config1(n=n)
    nwet = case.grid.nwet
    ref_case = case.copy() if not args.no_cpu else None
    ck = {"auto": 0, "multi": 1, "persistent": 2}[args.cg_kernel]
    ctx_m = make_ctx(case, mixed=True, cg_kernel=ck)
    ctx_d = make_ctx(case, mixed=False, cg_kernel=ck)
    # the FP64 solve in the pipelined form (one exchange per iteration), a
    # side line: the reference's recurrences are the FP64 default
    ctx_p = make_ctx(case, mixed=False, cg_kernel=ck)
    ctx_p.set_cg64_pipeline(True)
    i = np.arange(2, case.grid.LA + 1) - 1

    # ---- accuracy + CPU baseline on a bounded sample of the same steps
    cpu = None
    ncmp = 2
    if not args.no_cpu:
        from oracle import oracle as orc
        o = orc.Oracle(ref_case, rsqm=0.0, iterm=4000, nthreads=args.cpu_threads,
                       rel_tol=1e-10)
        t_cpu = 0.0
        done = 0
        while done < 1 or (t_cpu < args.cpu_seconds and done < 50):
            t0 = time.perf_counter()
            st = o.calpuv2c(n=done + 1)
            t_cpu += time.perf_counter() - t0
            done += 1
            if st["status"] != 0:
                raise RuntimeError(f"oracle status {st}")
        ncmp = done
        cpu = dict(value=nwet * done / t_cpu, unit="cell-steps/s",
                   cores=min(args.cpu_threads, 8), host_cpus=host_cpus(), kind="port",
                   sample=f"{done} CALPUV2C steps of the same {n}x{n} "
                          f"lake by the FP64 C restatement (OpenMP, <=8-way "
                          f"L-domains, serial dot products), {t_cpu:.1f} s")
        del o
    run_steps(ctx_m, case, 1, ncmp)
    run_steps(ctx_d, case, 1, ncmp)
    run_steps(ctx_p, case, 1, ncmp)
    err_m = err_d = err_p = None
    if ref_case is not None:
        pr = ref_case.h["P"][i]
        err_m = float(np.max(np.abs(ctx_m.get("P")[i] - pr))) / 9.81
        err_d = float(np.max(np.abs(ctx_d.get("P")[i] - pr))) / 9.81
        err_p = float(np.max(np.abs(ctx_p.get("P")[i] - pr))) / 9.81
        del ref_case, pr

    # ---- timed region: FP32-refined, then FP64 (the first warm-up step
    # of the FP32-refined solve runs the autotuner)
    n0 = ncmp + 1
    ctx_m.autotune(True)
    run_steps(ctx_m, case, n0, max(args.warmup, 1))
    wu = max(args.warmup, 1)
    t_m, it_m = timed(ctx_m, case, n0 + wu, args.steps)
    run_steps(ctx_d, case, n0, wu)
    t_d, it_d = timed(ctx_d, case, n0 + wu, args.steps)
    run_steps(ctx_p, case, n0, wu)
    t_p, it_p = timed(ctx_p, case, n0 + wu, args.steps)
    n1 = n0 + wu + args.steps
    rf_m = solve_roofline(ctx_m, case, n1, args.steps, True, nwet, workload)
    rf_d = solve_roofline(ctx_d, case, n1, args.steps, False, nwet, workload)
    rf_p = solve_roofline(ctx_p, case, n1, args.steps, False, nwet, workload)
    ctx_p.close()

    value = nwet * args.steps / t_m
    copy_gbs = stream_copy_gbs(ctx_m)
    tune = ctx_m.tune_report()
    launch_m = ctx_m.cg_launch()
    ctx_m.close()
    ctx_d.close()
    out = {
        "metric": "cell-steps/s + CG HBM GB/s (% peak), FP32-refined vs FP64 at <=1e-6 m error",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"external-mode solve at scale: {n}x{n} irregular "
                        f"lake, {nwet} wet cells, dt=10 s, CALPUV2C per step, "
                        "CONGRAD to relative residual 1e-10",
            "wet_cells": nwet,
            "cg": "FP32 SpMV + FP64 dots + FP64 refinement",
            "cg_kernel": args.cg_kernel,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": rf_m["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": rf_m["achieved"] / HBM_PEAK_GBS,
            "traffic": rf_m["traffic"],
            "kernel": rf_m["kernel"],
            "avg_us": rf_m["avg_us"],
            "bytes_per_launch": rf_m["bytes_per_launch"],
            "basis": f"method bytes per solve: {ALG_FP32_ITER} B x inner iterations + "
                     f"{ALG_RESID64} B x (refinement steps + 1) + {ALG_DOWNCONV} B, per wet cell",
            "hbm_gbs": rf_m["hbm_gbs"],
            "compact_operator_gbs": rf_m["compact_gbs"],
            "compact_bytes_per_launch": rf_m["compact_bytes_per_solve"],
        },
        "cpu_baseline": cpu,
        "detail": {
            "autotune": tune,
            # persistent solves: 0 = plain launch of the co-resident grid;
            # solves repeated cooperatively because a wait gave up
            "cg_launch": {"mode": launch_m[0], "repeated_solves": launch_m[1]},
            "stream_copy_gbs": copy_gbs,
            "solve_frac_of_copy": rf_m["achieved"] / copy_gbs,
            "inner_iters_per_solve_mixed": rf_m["iters_per_solve"],
            "refinements_per_solve_mixed": rf_m["outer_per_solve"],
            "fp64_value": nwet * args.steps / t_d,
            "fp64_ms_per_step": 1e3 * t_d / args.steps,
            "fp64_roofline": {k_: rf_d[k_] for k_ in ("kernel", "achieved", "avg_us",
                                                     "bytes_per_launch", "traffic", "hbm_gbs",
                                                     "compact_gbs", "iters_per_solve")},
            "fp64_frac": rf_d["achieved"] / HBM_PEAK_GBS,
            "fp64_pipelined": {
                "ms_per_step": 1e3 * t_p / args.steps,
                "value": nwet * args.steps / t_p,
                "solve_avg_us": rf_p["avg_us"],
                "iters_per_solve": rf_p["iters_per_solve"],
                "eta_err_vs_oracle_m": err_p,
            },
            "cg_iters_per_step_mixed": float(np.mean(it_m)),
            "cg_iters_per_step_fp64": float(np.mean(it_d)),
            "eta_err_vs_oracle_m": {"mixed": err_m, "fp64": err_d,
                                    "after_steps": ncmp},
            "kernels_mixed": rf_m["kernels"],
            "kernels_fp64": rf_d["kernels"],
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
