#!/usr/bin/env python3
"""Benchmark of the SL-VOF transport + pressure multigrid hot path on one MI355X.

Headline (`value`): SL-VOF cell-updates/s of the 3D deformation at 128^3 (sphere r=0.15 at
(0.35,0.35,0.35), period T=3, CFL 0.5, dt = 1/512): one step = compute_dt, defineVelocity!,
transport! (ELVIRA normals, back-trace, tet cutting, VF update, PLIC refresh) and
computeBand!; value = 128^3 x steps / wall time, inputs already on the device.
Beside it (`mg`): pressure multigrid solve time to max|b - A p| <= 1e-10 on the 128^3
variable-density problem (density ratio 1000, exact p = cos cos cos), FP32 V-cycles in FP64
defect correction vs FP64 V-cycles.
`roofline`: the ELVIRA kernel (the dominant kernel of the step) in FP64 FLOP/s from HIP
events on the library's stream over the timed steps; `cpu_baseline`: the oracle on the
host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def elvira_flops_per_cell(N):
    """Counted FP64 operations per interface cell of one ELVIRA pass, from the oracle's
    operation counter on the workload's initial state at a reduced size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ns_multiphase_amd import workloads as W
    import numpy as np
    param, ic = W.deformation3d(N)
    o = pyoracle.OracleContext(param)
    o.init_vof(*ic)
    o.neumann("VF")
    pyoracle.count_flops(True)
    o.compute_normal()
    fl, sp = pyoracle.flops()
    pyoracle.count_flops(False)
    VF = W.interior(o.mesh, o.get("VF"), vof=True)
    nif = int(((VF >= param.VFlo) & (VF <= param.VFhi)).sum())
    return fl / max(nif, 1), sp / max(nif, 1), nif


def run_vof(args, torch):
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    N = args.N
    param, ic = W.deformation3d(N)
    nsteps_period = int(round(param.tFinal / param.max_dt))
    steps = args.steps if args.steps is not None else nsteps_period
    warm = args.warmup if args.warmup is not None else 10
    # warm-up on its own context (first launches, allocation of scratch)
    with solver.Context(param) as w:
        w.init_vof(*ic)
        w.vof_init()
        t = 0.0
        for _ in range(warm):
            t, _ = w.vof_step(t, param.tFinal)
    ctx = solver.Context(param)
    ctx.init_vof(*ic)
    ctx.vof_init()
    ctx.profile(True)
    V0 = ctx.reduce("sum", "VF")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = 0.0
    for _ in range(steps):
        t, dt = ctx.vof_step(t, param.tFinal)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    V1 = ctx.reduce("sum", "VF")
    return ctx, param, steps, warm, elapsed, prof, (V0, V1, t)


def cpu_baseline(args, param, ic):
    """Oracle step time on the host cores, on device states sampled across the period."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ns_multiphase_amd import solver
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    threads = min(threads, 16)
    pyoracle.set_threads(threads)
    nsteps_period = int(round(param.tFinal / param.max_dt))
    samples = [0, nsteps_period // 4, nsteps_period // 2]
    g = solver.Context(param)
    g.init_vof(*ic)
    g.vof_init()
    o = pyoracle.OracleContext(param)
    t = 0.0
    done = 0
    cpu_time = 0.0
    for s in samples:
        while done < s:
            t, _ = g.vof_step(t, param.tFinal)
            done += 1
        for name in ("VF", "nx", "ny", "nz", "D", "band", "u", "v", "w"):
            o.set(name, g.get(name))
        t0 = time.perf_counter()
        o.vof_step(t, param.tFinal)
        cpu_time += time.perf_counter() - t0
        log(f"  cpu sample at step {s}: {time.perf_counter() - t0:.2f} s")
    g.close()
    cells = param.Nx * param.Ny * param.Nz
    return {"value": cells * len(samples) / cpu_time, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"one oracle step (C++ FP64 restatement, OpenMP) from the device state at steps "
                      f"{samples} of the {param.Nx}^3 deformation, {cpu_time:.1f} s total"}


def run_mg(args, torch):
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    import numpy as np
    N = args.mg_N
    dt = 1e-3
    p = W.mms_params(N)
    out = {}
    for prec in ("fp64", "mixed"):
        ctx = solver.Context(p)
        W.mms_setup(ctx, dt)
        P0 = ctx.get("P")
        ctx.pressure_solve(dt, 1e-10, prec)  # warm: builds the hierarchy
        times = []
        for rep in range(args.mg_reps):
            ctx.set("P", P0)
            if rep == args.mg_reps - 1:
                ctx.profile(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cyc, res = ctx.pressure_solve(dt, 1e-10, prec)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        prof = ctx.profile_read()
        P = W.interior(ctx.mesh, ctx.get("P"))
        err = float(np.abs(P - P.mean() - W.mms_exact(ctx.mesh)).max())
        out[prec] = {"ms": 1e3 * min(times), "cycles": cyc, "resid": res, "linf_err": err, "prof": prof}
        ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: one period, 1536 at 128^3)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--mg-N", type=int, default=128)
    ap.add_argument("--mg-reps", type=int, default=3)
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the hot path runs on a "
              "single MI355X; multi-GPU sharding is a design only)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(3)
    from ns_multiphase_amd import workloads as W

    log(f"[bench] SL-VOF deformation3d N={args.N}")
    ctx, param, steps, warm, elapsed, prof, (V0, V1, tend) = run_vof(args, torch)
    cells = param.Nx * param.Ny * param.Nz
    value = cells * steps / elapsed
    log(f"[bench] {steps} steps in {elapsed:.3f} s -> {value:.3e} cell-updates/s (t_end={tend:.4f})")

    # roofline of the dominant kernel (ELVIRA, FP64)
    fl_cell, sp_cell, _ = elvira_flops_per_cell(min(args.N, 64))
    e_ms, e_n, e_units = prof["elvira"]
    roof = None
    if e_n > 0 and e_ms > 0:
        flops_per_launch = fl_cell * e_units / e_n
        avg_s = e_ms / e_n / 1e3
        achieved = flops_per_launch / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "kernel": "k_elvira", "avg_launch_us": 1e6 * avg_s, "launches": e_n,
                "interface_cells_per_launch": e_units / e_n, "flops_per_interface_cell": fl_cell,
                "note": "FP64 vector peak of MI355X (78.6 TFLOP/s); no FMA by construction (-ffp-contract=off)"}
    kernels = {k: {"ms": v[0], "launches": v[1], "avg_us": (1e3 * v[0] / v[1]) if v[1] else None,
                   "units_per_launch": v[2] / v[1] if v[1] else None}
               for k, v in prof.items() if v[1]}
    ctx.close()

    mg = None
    if not args.no_mg:
        log(f"[bench] multigrid MMS N={args.mg_N}")
        mg_raw = run_mg(args, torch)
        mg = {}
        for prec, r in mg_raw.items():
            rb_ms, rb_n, rb_units = r["prof"]["mg_rbgs"]
            s = 8 if prec == "fp64" else 4
            entry = {"solve_ms": r["ms"], "cycles": r["cycles"], "final_resid": r["resid"],
                     "linf_err_vs_exact": r["linf_err"]}
            if rb_n:
                avg = rb_ms / rb_n / 1e3
                bytes_launch = 6.0 * s * rb_units / rb_n  # p read + write, b, 3 face coefficients
                entry["rbgs_fine"] = {"avg_launch_us": 1e6 * avg, "achieved_GBs": bytes_launch / avg / 1e9,
                                      "frac_hbm": bytes_launch / avg / 1e9 / PEAK_HBM_GBS,
                                      "bytes_per_launch": bytes_launch}
            mg[prec] = entry
        mg["speedup_mixed_vs_fp64"] = mg["fp64"]["solve_ms"] / mg["mixed"]["solve_ms"]
        log(f"[bench] mg fp64 {mg['fp64']['solve_ms']:.2f} ms, mixed {mg['mixed']['solve_ms']:.2f} ms")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle)")
        param_c, ic = W.deformation3d(args.N)
        cpu = cpu_baseline(args, param_c, ic)

    line = {
        "metric": "SL-VOF cell-updates/s at 128^3 + MG solve time, mixed vs FP64, at 1e-10 resid",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D deformation {args.N}^3, sphere r=0.15, T=3, CFL 0.5, dt=1/512, RK4, 5 tets",
                   "steps_timed_from_t0": steps, "volume_rel_change": (V1 - V0) / V0,
                   "mg_workload": f"variable-density MMS {args.mg_N}^3, rho 1000/1, tol 1e-10"},
        "roofline": roof, "cpu_baseline": cpu, "mg": mg, "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
