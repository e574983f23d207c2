#!/usr/bin/env python3
"""Benchmark of the SL-VOF transport + pressure multigrid hot path on one MI355X.

Headline (`value`): SL-VOF cell-updates/s of the 3D deformation at 128^3 (sphere r=0.15 at
(0.35,0.35,0.35), period T=3, CFL 0.5, dt = 1/512): one step = compute_dt, defineVelocity!,
transport! (ELVIRA normals, back-trace, tet cutting, VF update, PLIC refresh) and
computeBand!; value = 128^3 x steps / wall time, inputs already on the device.
`rising_bubble`: full Navier-Stokes steps of the 128^3 rising bubble (50 steps) with both
multigrid precisions, and the oracle's step time on the host cores beside them.
`adjoint`: the loss gradient through 64 checkpointed SL-VOF steps at 256^2 (16 stream-function
amplitudes) against the forward's cost and central differences. `large`: the 256^3 pressure
MMS in both precisions and 20 deformation steps at 256^3.
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


def counted_flops(N):
    """Counted FP64 operations (+ - * / = 1 each) of the oracle, on one thread, over one step of
    the workload from its initial state at a reduced size: per interface cell of one ELVIRA
    pass, and per band cell of the SL cut (mesh-plane and PLIC cuts of cutTet, sub-tet
    volumes, the vol / vLiq sums)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ns_multiphase_amd import workloads as W
    param, ic = W.deformation3d(N)
    o = pyoracle.OracleContext(param)
    o.init_vof(*ic)
    o.vof_init()
    threads = pyoracle.get_threads()
    pyoracle.set_threads(1)  # the counters are per thread

    def count(f):
        pyoracle.flops()
        pyoracle.count_flops(True)
        start = pyoracle.flops()[0]
        f()
        n = pyoracle.flops()[0] - start
        pyoracle.count_flops(False)
        return n

    VF = W.interior(o.mesh, o.get("VF"), vof=True)
    nif = int(((VF >= param.VFlo) & (VF <= param.VFhi)).sum())
    band = W.interior(o.mesh, o.get("band"))
    nband = int((abs(band) <= param.nband).sum())
    dt = o.compute_dt()
    o.define_velocity(dt)
    f_elv0 = count(o.compute_normal)          # the transport's first ELVIRA, same state
    f_tr = count(lambda: o.transport(dt, dt))
    f_elv1 = count(o.compute_normal)          # its last ELVIRA and PLIC, same state
    f_plic = count(o.compute_plic)
    pyoracle.set_threads(threads)
    return {"elvira_per_interface_cell": f_elv0 / max(nif, 1),
            "cut_per_band_cell": (f_tr - f_elv0 - f_elv1 - f_plic) / max(nband, 1),
            "N": N, "interface_cells": nif, "band_cells": nband}


# HBM bytes per unit (interface cell for ELVIRA's two kernels together, band cell for the cut
# pass) from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch on 20 steps of this workload,
# FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0 (what the counters read for a plain 256 MiB
# device copy); see profiles/pmc_traffic_deformation3d_128.md.
TRAFFIC_PER_UNIT = {"k_elvira_dist + k_elvira": 1725.0, "k_sl_tree": 847.0}


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
    m = ctx.mesh
    o = m.nghost + 3
    inner = (slice(o, o + m.Nx), slice(o, o + m.Ny), slice(o, o + m.Nz))
    VF0 = ctx.get("VF")[inner].copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = 0.0
    for _ in range(steps):
        t, dt = ctx.vof_step(t, param.tFinal)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    V1 = ctx.reduce("sum", "VF")
    # shape error sum |VF(t) - VF(0)| / N^3 (3D form of test/deformation_error.jl:22); the
    # deformation returns the sphere at t = T
    shape_err = float(abs(ctx.get("VF")[inner] - VF0).sum()) / (m.Nx * m.Ny * m.Nz)
    return ctx, param, steps, warm, elapsed, prof, (V0, V1, t, shape_err)


def cpu_baseline(args, param, ic):
    """Oracle step time on the host cores, on device states sampled across the period."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ns_multiphase_amd import solver
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    threads = min(threads, 16)
    pyoracle.set_threads(threads)
    nsteps_period = int(round(param.tFinal / param.max_dt))
    samples = list(range(0, nsteps_period, max(1, nsteps_period // 64)))  # 64 steps across the period
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
        for name in ("VF", "band", "u", "v", "w"):  # the step recomputes normals, PLIC and velocity
            o.set(name, g.get(name))
        t0 = time.perf_counter()
        o.vof_step(t, param.tFinal)
        cpu_time += time.perf_counter() - t0
    log(f"  cpu: {len(samples)} oracle steps in {cpu_time:.1f} s")
    # the same restatement on one core: two steps from the last sampled state
    pyoracle.set_threads(1)
    one_time = 0.0
    for _ in range(2):
        for name in ("VF", "band", "u", "v", "w"):
            o.set(name, g.get(name))
        t0 = time.perf_counter()
        o.vof_step(t, param.tFinal)
        one_time += time.perf_counter() - t0
    pyoracle.set_threads(threads)
    log(f"  cpu: 2 oracle steps on one core in {one_time:.1f} s")
    g.close()
    cells = param.Nx * param.Ny * param.Nz
    return {"value": cells * len(samples) / cpu_time, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(samples)} oracle steps (C++ FP64 restatement, OpenMP), each from the device "
                      f"state at steps 0, {samples[1] if len(samples) > 1 else 0}, ... of the {param.Nx}^3 "
                      f"deformation period, {cpu_time:.1f} s of CPU time",
            "single_core": {"value": cells * 2 / one_time, "unit": "cell-updates/s", "cores": 1,
                            "sample": f"2 oracle steps on one thread from the device state at step {samples[-1]}, "
                                      f"{one_time:.1f} s"}}


def copy_bandwidth(torch, nbytes, reps=20):
    """Device-to-device copy rate in GB/s (bytes read + written) of an nbytes buffer."""
    a = torch.empty(nbytes // 4, device="cuda", dtype=torch.float32).fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    return 2.0 * nbytes / sec / 1e9


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


def run_bubble(args, torch):
    """Rising bubble, full run_solver steps with solveNS (transport with momentum predictor and
    correction tets, band, props, pressure solve to 1e-10, corrector, interpolateCenter!,
    divergence!), FP32-in-FP64 and FP64 multigrid."""
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    param, ic = W.rising_bubble(args.bubble_N, stepMax=args.bubble_steps)
    out = {}
    for prec in ("mixed", "fp64"):
        with solver.Context(param) as w:  # warm-up: first launches, hierarchy, graphs
            w.init_vof(*ic)
            w.ns_init()
            t = 0.0
            for _ in range(2):
                t, _, _, _ = w.ns_step(t, param.tol, prec)
        ctx = solver.Context(param)
        ctx.init_vof(*ic)
        ctx.ns_init()
        torch.cuda.synchronize()
        t = 0.0
        cycles = 0
        worst = 0.0
        t0 = time.perf_counter()
        for _ in range(args.bubble_steps):
            t, dt, cyc, res = ctx.ns_step(t, param.tol, prec)
            cycles += cyc
            worst = max(worst, res)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        wmax = ctx.reduce("max", "w")
        # per-kernel breakdown on a few more steps with event timing on (the multigrid then
        # runs its iterations eagerly instead of as graphs)
        ctx.profile(True)
        nprof = 5
        for _ in range(nprof):
            t, dt, cyc, res = ctx.ns_step(t, param.tol, prec)
        prof = ctx.profile_read()
        ctx.close()
        out[prec] = {"ms_per_step": 1e3 * elapsed / args.bubble_steps, "steps": args.bubble_steps,
                     "mg_cycles_per_step": cycles / args.bubble_steps, "max_final_resid": worst,
                     "max_w_at_end": wmax,
                     "profiled_ms_per_step": {k: v[0] / nprof for k, v in prof.items() if v[1]}}
        log(f"[bench] bubble {prec}: {out[prec]['ms_per_step']:.2f} ms/step, "
            f"{out[prec]['mg_cycles_per_step']:.1f} MG cycles/step")
    out["speedup_mixed_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    return param, ic, out


def cpu_bubble(args, param, ic, nsamples=2):
    """Oracle rising-bubble step time on the host cores, from device states of the run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ns_multiphase_amd import solver
    from ns_multiphase_amd.params import FIELD
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    threads = min(threads, 16)
    pyoracle.set_threads(threads)
    g = solver.Context(param)
    g.init_vof(*ic)
    g.ns_init()
    o = pyoracle.OracleContext(param)
    t = 0.0
    at = [0] + [args.bubble_steps // 2] * (nsamples - 1)
    done = 0
    cpu_time = 0.0
    for s in at:
        while done < s:
            t, _, _, _ = g.ns_step(t, param.tol, "fp64")
            done += 1
        for name in FIELD:
            if not name.endswith("_bar"):  # the adjoint's fields have no oracle counterpart
                o.set(name, g.get(name))
        t0 = time.perf_counter()
        o.ns_step(t, param.tol, "fp64")
        cpu_time += time.perf_counter() - t0
    g.close()
    log(f"  cpu bubble: {len(at)} oracle steps in {cpu_time:.1f} s")
    return {"ms_per_step": 1e3 * cpu_time / len(at), "cores": threads, "kind": "port",
            "sample": f"{len(at)} oracle steps of the {param.Nx}^3 rising bubble (C++ FP64 restatement, "
                      f"FP64 multigrid to the same 1e-10 residual), from the device state at steps {at}"}


def run_adjoint(args, torch):
    """Gradient of the SL-VOF loss with respect to 16 stream-function amplitudes: 256^2, 64
    steps, the checkpoint interval chosen by an 8 GiB device-memory budget (every state at this
    size); timed against the plain forward loss, and checked per amplitude against central
    differences (h = 1e-6) of the forward loss. The loss and gradient with checkpoints every 8
    steps (each segment run forward again in the reverse sweep) is timed beside it."""
    import numpy as np
    from ns_multiphase_amd import adjoint as A
    from ns_multiphase_amd import solver
    N, nsteps = args.adj_N, 64
    p = A.adjoint_params(N)
    ctx = solver.Context(p)
    ctx.init_vof("circle", 0.2, 0.5, 0.5)
    vf0 = ctx.get("VF")
    basis = A.mode_faces(ctx.mesh)
    a1 = np.random.default_rng(1).normal(0.0, 0.1, 16)
    a2 = np.random.default_rng(2).normal(0.0, 0.1, 16)
    dt = min(A.stable_dt(ctx.mesh, basis, a1), A.stable_dt(ctx.mesh, basis, a2))
    run = A.SLVOFRun(ctx, vf0, basis, dt, nsteps, None, mem_budget=8 << 30)
    every = run.every
    run8 = A.SLVOFRun(ctx, vf0, basis, dt, nsteps, 8)
    target = run.forward(a2, checkpoints=False)
    run.loss_and_grad(a1, target)  # warm-up
    run8.loss_and_grad(a1, target)
    reps = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        L = run.loss(a1, target)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        L2, g = run.loss_and_grad(a1, target)
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        _, g8 = run8.loss_and_grad(a1, target)
    torch.cuda.synchronize()
    t_grad8 = (time.perf_counter() - t0) / reps
    # the reverse sweep alone (recomputing each segment from its checkpoint, when the interval
    # is above 1, then walking it back), after a forward that stored the checkpoints
    m = ctx.mesh
    o = m.nghost + 3
    sl = (slice(o, o + m.Nx), slice(o, o + m.Ny), slice(o, o + m.Nz))
    t_bwd = 0.0
    for _ in range(reps):
        VF = run.forward(a1)
        diff = np.zeros_like(VF)
        diff[sl] = VF[sl] - target[sl]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.backward(diff)
        torch.cuda.synchronize()
        t_bwd += (time.perf_counter() - t0) / reps
    # central differences at h = 1e-6 and 1e-7: the map is piecewise smooth, and a kink (an
    # ELVIRA choice or a cut case switching) within h of the point makes a direction disagree
    checks = []
    for h in (1e-6, 1e-7):
        fd = np.array([(run.loss(a1 + h * e, target) - run.loss(a1 - h * e, target)) / (2 * h) for e in np.eye(16)])
        rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
        checks.append({"h": h, "within_1e-6": int((rel <= 1e-6).sum()), "directions": 16,
                       "max_rel": float(rel.max()), "median_rel": float(np.median(rel))})
    ctx.close()
    out = {"workload": f"{N}^2 x 1, circle r=0.2 at (0.5,0.5), 16-mode stream function (seed 1 control, seed 2 "
                       f"target), {nsteps} steps, dt={dt:.3e}, checkpoints every {every}",
           "checkpoint_every": every,
           "checkpoint_MB": (run.nseg + every - 1) * A.checkpoint_bytes(m) / 1e6,
           "loss": L, "forward_ms": 1e3 * t_fwd, "loss_and_gradient_ms": 1e3 * t_grad,
           "gradient_cost_vs_forward": t_grad / t_fwd, "reverse_sweep_ms": 1e3 * t_bwd,
           "reverse_sweep_vs_forward": t_bwd / t_fwd,
           "loss_and_gradient_ms_every8": 1e3 * t_grad8, "gradient_cost_vs_forward_every8": t_grad8 / t_fwd,
           "every8_vs_every_max_rel": float(np.abs(g8 - g).max() / np.abs(g).max()),
           "fd_check": checks[0], "fd_check_h1e-7": checks[1]}
    log(f"[bench] adjoint: checkpoints every {every} ({out['checkpoint_MB']:.0f} MB), forward {out['forward_ms']:.1f} ms, "
        f"loss+gradient {out['loss_and_gradient_ms']:.1f} ms "
        f"({out['gradient_cost_vs_forward']:.2f}x; {out['loss_and_gradient_ms_every8']:.1f} ms = "
        f"{out['gradient_cost_vs_forward_every8']:.2f}x with checkpoints every 8), reverse sweep {out['reverse_sweep_ms']:.1f} ms "
        f"({out['reverse_sweep_vs_forward']:.2f}x); FD within 1e-6: {checks[0]['within_1e-6']}/16 at h=1e-6, "
        f"{checks[1]['within_1e-6']}/16 at h=1e-7")
    return out


def run_large(args, torch):
    """256^3: the variable-density pressure MMS (mixed vs FP64 to 1e-10, L-inf error against
    the exact pressure), then 20 SL-VOF steps of the 3D deformation (dt = 1/1024)."""
    import numpy as np
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    N = args.large_N
    out = {"workload": f"{N}^3 MMS (rho 1000/1, tol 1e-10) + 20 steps of 3D deformation {N}^3"}
    dt = 1e-3
    p = W.mms_params(N)
    for prec in ("fp64", "mixed"):
        ctx = solver.Context(p)
        W.mms_setup(ctx, dt)
        P0 = ctx.get("P")
        ctx.pressure_solve(dt, 1e-10, prec)  # warm: hierarchy and graphs
        ctx.set("P", P0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cyc, res = ctx.pressure_solve(dt, 1e-10, prec)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0)
        Pn = W.interior(ctx.mesh, ctx.get("P"))
        err = float(np.abs(Pn - Pn.mean() - W.mms_exact(ctx.mesh)).max())
        out[f"mg_{prec}"] = {"solve_ms": ms, "cycles": cyc, "final_resid": res, "linf_err_vs_exact": err}
        ctx.close()
        log(f"[bench] {N}^3 MMS {prec}: {ms:.1f} ms, {cyc} cycles, L-inf {err:.2e}")
    out["mg_speedup_mixed_vs_fp64"] = out["mg_fp64"]["solve_ms"] / out["mg_mixed"]["solve_ms"]
    param, ic = W.deformation3d(N)
    ctx = solver.Context(param)
    ctx.init_vof(*ic)
    ctx.vof_init()
    t = 0.0
    for _ in range(2):
        t, _ = ctx.vof_step(t, param.tFinal)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        t, _ = ctx.vof_step(t, param.tFinal)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out["deformation_20_steps"] = {"ms_per_step": 1e3 * el / 20, "cell_updates_per_s": N ** 3 * 20 / el,
                                   "dt": param.max_dt}
    ctx.close()
    log(f"[bench] {N}^3 deformation: {1e3 * el / 20:.2f} ms/step")
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
    ap.add_argument("--bubble-N", type=int, default=128)
    ap.add_argument("--bubble-steps", type=int, default=50)
    ap.add_argument("--no-bubble", action="store_true")
    ap.add_argument("--adj-N", type=int, default=256)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--large-N", type=int, default=256)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB on the device 3 times (the plain copy that scales the PMC "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 64 * 1024 * 1024
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": 4 * n, "copies": 3}), flush=True)
        return
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
    ctx, param, steps, warm, elapsed, prof, (V0, V1, tend, shape_err) = run_vof(args, torch)
    cells = param.Nx * param.Ny * param.Nz
    value = cells * steps / elapsed
    log(f"[bench] {steps} steps in {elapsed:.3f} s -> {value:.3e} cell-updates/s (t_end={tend:.4f})")

    # rooflines of the two FP64 kernels of the step; the dominant one is `roofline`
    fc = counted_flops(min(args.N, 64))
    roofs = {}
    for key, kern, per, unit_name in (("elvira", "k_elvira_dist + k_elvira", fc["elvira_per_interface_cell"],
                                       "interface_cells"),
                                      ("sl_cut", "k_sl_tree", fc["cut_per_band_cell"], "band_cells")):
        ms, n, units = prof[key]
        if n <= 0 or ms <= 0:
            continue
        avg_s = ms / n / 1e3
        achieved = per * units / n / avg_s / 1e12
        roofs[key] = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                      "frac": achieved / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_PER_UNIT[kern] * units / n,
                      "kernel": kern, "avg_launch_us": 1e6 * avg_s, "launches": n, "total_ms": ms,
                      f"{unit_name}_per_launch": units / n, f"flops_per_{unit_name[:-1]}": per,
                      "note": "FP64 vector peak of MI355X (78.6 TFLOP/s, FMA = 2); the kernels form no FMA "
                              "(-ffp-contract=off); flops counted by the oracle at N=%d" % fc["N"]}
    roof = max(roofs.values(), key=lambda r: r["total_ms"]) if roofs else None
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
        # measured copy bandwidths beside the HBM peak: 2 x 64 MiB stays in the 256 MB Infinity
        # Cache (where the 128^3 levels live), 2 x 2 GiB does not
        mg["copy_GBs"] = {"cache_resident_2x64MiB": copy_bandwidth(torch, 64 << 20),
                          "hbm_2x2GiB": copy_bandwidth(torch, 2 << 30)}
        for prec in ("fp64", "mixed"):
            if "rbgs_fine" in mg[prec]:
                rf = mg[prec]["rbgs_fine"]
                rf["frac_cache_copy"] = rf["achieved_GBs"] / mg["copy_GBs"]["cache_resident_2x64MiB"]
        log(f"[bench] mg fp64 {mg['fp64']['solve_ms']:.2f} ms, mixed {mg['mixed']['solve_ms']:.2f} ms")

    bubble = None
    if not args.no_bubble:
        log(f"[bench] rising bubble N={args.bubble_N}, {args.bubble_steps} steps")
        bparam, bic, bubble = run_bubble(args, torch)
        if not args.no_cpu_baseline:
            bubble["cpu_baseline"] = cpu_bubble(args, bparam, bic)
            bubble["speedup_vs_cpu"] = {prec: bubble["cpu_baseline"]["ms_per_step"] / bubble[prec]["ms_per_step"]
                                        for prec in ("mixed", "fp64")}
        bubble["workload"] = (f"rising bubble {args.bubble_N}^3, r=0.125 at (0.5,0.5,0.3), rho 1000/1, "
                              f"mu 1e-3/1e-5, sigma 0.07, g=-9.81 z, no-slip walls, dt=1e-3, CFL 0.5, "
                              f"finite-difference scheme with correction tets, MG tol 1e-10")

    adjoint = None
    if not args.no_adjoint:
        log(f"[bench] adjoint {args.adj_N}^2")
        adjoint = run_adjoint(args, torch)
    large = None
    if not args.no_large:
        log(f"[bench] large {args.large_N}^3")
        large = run_large(args, torch)

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
                   "t_end": tend, "shape_error": shape_err,
                   "mg_workload": f"variable-density MMS {args.mg_N}^3, rho 1000/1, tol 1e-10"},
        "roofline": roof, "cpu_baseline": cpu, "mg": mg, "rising_bubble": bubble, "adjoint": adjoint,
        "large": large, "kernels": kernels,
        "rooflines": roofs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
