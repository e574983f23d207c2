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
`roofline`: the dominant FP64 kernel of the step by time (the cut pass; `rooflines` holds
ELVIRA beside it) in FP64 FLOP/s from HIP events on the library's stream over the timed steps; `cpu_baseline`: the oracle on the
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


def cpu_threads():
    """Host threads for the oracle legs: OMP_NUM_THREADS (16 on a one-GPU machine's share), at most 16."""
    return min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1, 16)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# HBM bytes per unit (interface cell for ELVIRA's two kernels together, band cell for the cut
# pass) from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch on the timed steps of this
# workload, FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0 (what the counters read for a plain
# 256 MiB device copy); see profiles/pmc_traffic_deformation3d_128.md.
TRAFFIC_PER_UNIT = {"k_elvira_dist + k_elvira": 3190.0, "k_sl_tree": 720.7}


def _sample_steps(nper, k):
    """k step indices spread evenly over a period of nper steps, starting at 0."""
    return sorted({int(round(q * nper / k)) for q in range(k)})


def run_vof(args, torch):
    """The 128^3 deformation. An untimed pass walks the whole period once, saving a device
    checkpoint at each sampled state (steps spread evenly over the period, one per timed step)
    and giving the shape error at t = T. The timed region then runs one step from each sampled
    state (the checkpoint restore, one copy kernel, is inside it). Without --steps the timed
    region is the whole period from t = 0 instead."""
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    N = args.N
    param, ic = W.deformation3d(N)
    nper = int(round(param.tFinal / param.max_dt))
    warm = args.warmup if args.warmup is not None else 10
    ctx = solver.Context(param)
    ctx.init_vof(*ic)
    ctx.vof_init()
    m = ctx.mesh
    o = m.nghost + 3
    inner = (slice(o, o + m.Nx), slice(o, o + m.Ny), slice(o, o + m.Nz))
    VF0 = ctx.get("VF")[inner].copy()
    V0 = ctx.reduce("sum", "VF")
    sampled = args.steps is not None
    samples = _sample_steps(nper, args.steps if sampled else 64)
    t = 0.0
    done = 0
    times = {}
    if sampled:
        for slot, s in enumerate(samples):
            while done < s:
                t, _ = ctx.vof_step(t, param.tFinal)
                done += 1
            ctx.checkpoint_save(slot)
            times[slot] = t
    while t < param.tFinal:
        t, _ = ctx.vof_step(t, param.tFinal)
        done += 1
    tend = t
    V1 = ctx.reduce("sum", "VF")
    # shape error sum |VF(T) - VF(0)| / N^3 (3D form of test/deformation_error.jl:22); the
    # deformation returns the sphere at t = T
    shape_err = float(abs(ctx.get("VF")[inner] - VF0).sum()) / (m.Nx * m.Ny * m.Nz)
    log(f"[bench] full period: {done} steps to t={tend:.4f}, shape error {shape_err:.3e}")
    if sampled:
        steps = len(samples)
        for q in range(warm):  # warm-up steps from sampled states
            slot = q % steps
            ctx.checkpoint_restore(slot)
            ctx.vof_step(times[slot], param.tFinal)
        ctx.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for slot in range(steps):
            ctx.checkpoint_restore(slot)
            ctx.vof_step(times[slot], param.tFinal)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    else:
        steps = nper
        ctx.close()
        with solver.Context(param) as w:
            w.init_vof(*ic)
            w.vof_init()
            tw = 0.0
            for _ in range(warm):
                tw, _ = w.vof_step(tw, param.tFinal)
        ctx = solver.Context(param)
        ctx.init_vof(*ic)
        ctx.vof_init()
        ctx.profile(True)
        ctx.set_deferred_checks(True)  # each step waits for its dt only, not for the previous step
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tt = 0.0
        for _ in range(steps):
            tt, _ = ctx.vof_step(tt, param.tFinal)
        ctx.set_deferred_checks(False)  # reports a cut-walk overflow of any step
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    info = {"V0": V0, "V1": V1, "t_end": tend, "shape_err": shape_err, "samples": samples, "sampled": sampled,
            "times": times}
    return ctx, param, ic, steps, warm, elapsed, prof, info


def cpu_baseline(args, param, ic, info, time_cpu=True):
    """The oracle on the host cores, one step from each sampled device state (the timed states
    of the headline when --steps is given, else 64 states across the period), with the device's
    own step from the same state beside it: max |dVF|, band equality, total-volume difference,
    normals and PLIC constants per step. ELVIRA ties (the stated sphere sits on the field's
    y = z symmetry) take the device's choice where the oracle's errors tie with it to rounding.
    On 4 of the states the oracle runs on one thread with its operation counter: the FLOPs per
    interface cell of ELVIRA and per band cell of the cut at this size, and the one-core time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from ns_multiphase_amd import solver
    threads = cpu_threads()
    pyoracle.set_threads(threads)
    count_at = set(info["samples"][:: max(1, len(info["samples"]) // 4)][:4])
    samples = info["samples"] if time_cpu else sorted(count_at)
    g = solver.Context(param)
    g.init_vof(*ic)
    g.vof_init()
    o = pyoracle.OracleContext(param)
    m = g.mesh
    ng = m.nghost
    cin = (slice(ng, ng + m.Nx), slice(ng, ng + m.Ny), slice(ng, ng + m.Nz))
    vin = (slice(ng + 3, ng + 3 + m.Nx), slice(ng + 3, ng + 3 + m.Ny), slice(ng + 3, ng + 3 + m.Nz))
    t = 0.0
    done = 0
    cpu_time, ncpu, one_time, none = 0.0, 0, 0.0, 0
    err = {"steps_compared": 0, "max_abs_dVF": 0.0, "band_equal": True, "max_rel_dsumVF": 0.0,
           "max_abs_dnormal": 0.0, "max_abs_dD": 0.0}
    fl = {"elvira": 0, "interface_cells": 0, "cut": 0, "band_cells": 0}

    def counted(f):
        pyoracle.flops()
        pyoracle.count_flops(True)
        c0 = pyoracle.flops()[0]
        f()
        n = pyoracle.flops()[0] - c0
        pyoracle.count_flops(False)
        return n

    for s in samples:
        while done < s:
            t, _ = g.vof_step(t, param.tFinal)
            done += 1
        # the oracle starts from the device's state: VF, velocities, and the normals, PLIC
        # constants and band it computes from them
        h0 = g.elvira_choice()
        for name in ("VF", "u", "v", "w"):
            o.set(name, g.get(name))
        o.push_hints("elvira", h0)
        o.compute_normal()
        o.compute_plic()
        o.compute_band()
        t1, dt = g.vof_step(t, param.tFinal)
        done += 1
        h1 = g.elvira_choice()
        if s in count_at:
            VF = o.get("VF")[vin]
            nif = int(((VF >= param.VFlo) & (VF <= param.VFhi)).sum())
            nb = int((np.abs(o.get("band")[cin]) <= param.nband).sum())
            pyoracle.set_threads(1)
            o.push_hints("elvira", h0)
            f_a = counted(o.compute_normal)             # the step's first ELVIRA
            o.define_velocity(t1)
            o.push_hints("elvira", h0)
            o.push_hints("elvira", h1)
            t0 = time.perf_counter()
            f_b = counted(lambda: o.transport(dt, t1))  # ELVIRA, cut, VF update, ELVIRA, PLIC
            one_time += time.perf_counter() - t0
            none += 1
            o.push_hints("elvira", h1)
            f_c = counted(o.compute_normal)             # the step's last ELVIRA
            f_d = counted(o.compute_plic)
            o.compute_band()
            pyoracle.set_threads(threads)
            fl["elvira"] += f_a
            fl["interface_cells"] += nif
            fl["cut"] += f_b - f_a - f_c - f_d
            fl["band_cells"] += nb
        else:
            o.push_hints("elvira", h0)
            o.push_hints("elvira", h1)
            t0 = time.perf_counter()
            o.vof_step(t, param.tFinal)
            cpu_time += time.perf_counter() - t0
            ncpu += 1
        vg, vo = g.get("VF"), o.get("VF")
        err["steps_compared"] += 1
        err["max_abs_dVF"] = max(err["max_abs_dVF"], float(np.abs(vg - vo).max()))
        err["band_equal"] = err["band_equal"] and bool(np.array_equal(g.get("band"), o.get("band")))
        sg, so = g.reduce("sum", "VF"), o.reduce("sum", "VF")
        err["max_rel_dsumVF"] = max(err["max_rel_dsumVF"], abs(sg - so) / abs(so))
        for n in ("nx", "ny", "nz"):
            err["max_abs_dnormal"] = max(err["max_abs_dnormal"], float(np.abs(g.get(n) - o.get(n)).max()))
        err["max_abs_dD"] = max(err["max_abs_dD"], float(np.abs(g.get("D") - o.get("D")).max()))
        t = t1
    st = o.tie_stats("elvira")
    err["elvira_ties_adopted"] = st["adopted"]
    err["elvira_hints_refused"] = st["refused"]
    err["elvira_max_tie_gap"] = st["gap"]  # largest relative error gap of an adopted tie
    g.close()
    log(f"  cpu: {ncpu} oracle steps in {cpu_time:.1f} s on {threads} threads, {none} on one thread in "
        f"{one_time:.1f} s; per step vs device: max|dVF| {err['max_abs_dVF']:.2e}, band equal {err['band_equal']}, "
        f"ties adopted {st['adopted']}, refused {st['refused']}")
    cells = param.Nx * param.Ny * param.Nz
    flops = {"elvira_per_interface_cell": fl["elvira"] / max(fl["interface_cells"], 1),
             "cut_per_band_cell": fl["cut"] / max(fl["band_cells"], 1),
             "states": sorted(count_at), "N": param.Nx}
    if not time_cpu:
        return None, err, flops
    cpu = {"value": cells * ncpu / cpu_time, "unit": "cell-updates/s", "cores": threads, "kind": "port",
           "sample": f"{ncpu} oracle steps (C++ FP64 restatement, OpenMP), each from a device state of the "
                     f"{param.Nx}^3 deformation period (steps {samples[0]}, {samples[1] if len(samples) > 1 else ''}, ... "
                     f"{samples[-1]}: the timed states), {cpu_time:.1f} s of CPU time",
           "single_core": {"value": cells * none / one_time, "unit": "cell-updates/s", "cores": 1,
                           "sample": f"{none} oracle transport! calls on one thread (operation counter on) from the "
                                     f"device states at steps {sorted(count_at)}, {one_time:.1f} s"}}
    return cpu, err, flops


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
        out[prec] = {"ms": 1e3 * min(times), "cycles": cyc, "resid": res, "linf_err": err, "prof": prof,
                     "P": P - P.mean()}
        ctx.close()
    return out


def oracle_mms(N, dt):
    """The oracle's FP64 multigrid on the same MMS to the same residual (1e-10), on the host
    cores: the pressure the device's solves are compared with."""
    import numpy as np
    from ns_multiphase_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = cpu_threads()
    pyoracle.set_threads(threads)
    o = pyoracle.OracleContext(W.mms_params(N))
    W.mms_setup(o, dt)
    t0 = time.perf_counter()
    cyc, res = o.pressure_solve(dt, 1e-10, "fp64")
    sec = time.perf_counter() - t0
    P = W.interior(o.mesh, o.get("P"))
    o.close()
    return P - P.mean(), {"cycles": cyc, "final_resid": res, "seconds": sec, "cores": threads}


def mms_vs_oracle(P_dev, P_ora, info):
    """max|dp| between a device solve and the oracle's, both to max|r| <= 1e-10 with zero mean:
    the two iterates differ by at most that residual through the inverse operator."""
    import numpy as np
    return {"max_abs_dP": float(np.abs(P_dev - P_ora).max()), "max_abs_P": float(np.abs(P_ora).max()),
            "oracle": info}


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


def _directions(ctx):
    """Height-function direction compute_curvature! takes per interior cell from ctx's normals
    (0 x, 1 y, 2 z, -1 none), indexed [i, j, k]: the tie hints of the curvature."""
    import numpy as np
    m = ctx.mesh
    g = m.nghost
    a, b, c = (np.abs(ctx.get(n)[g:g + m.Nx, g:g + m.Ny, g:g + m.Nz]) for n in ("nx", "ny", "nz"))
    d = np.full(a.shape, -1, np.int32)
    d[(a > b) & (a > c)] = 0
    d[(b > a) & (b > c)] = 1
    d[(c > b) & (c > a)] = 2
    return d


def run_bubble_sl(args, torch):
    """Rising bubble with the semi-Lagrangian ("SL-SL") scheme: transport without correction
    tets and the nonlinear SL pressure equation solved by FAS cycles with res_iteration to the
    same 1e-10 (src/Multigrid.jl:375-500, src/Pressure.jl:777-884). Smaller than the FD bubble:
    each cycle rebuilds the diagonal Jacobian by central differences on every level."""
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    N = args.sl_N
    param, ic = W.rising_bubble(N, pressure_scheme="semi-lagrangian", pressureSolver="res_iteration", mg_lvl=2)
    ctx = solver.Context(param)
    ctx.init_vof(*ic)
    ctx.ns_init()
    t = 0.0
    rows = []
    o = None
    if not args.no_oracle and not args.no_cpu_baseline:  # the oracle's first step from the same state
        import numpy as np
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        from ns_multiphase_amd.params import FIELD
        pyoracle.set_threads(cpu_threads())
        o = pyoracle.OracleContext(param)
        for name in FIELD:
            if not name.endswith("_bar"):
                o.set(name, ctx.get(name))
        h0 = ctx.elvira_choice()
    failed = None
    for step in range(args.sl_steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            t1, dt, cyc, res = ctx.ns_step(t, param.tol)
        except solver.MistError as e:  # the FAS cycles stalled above tol (mg_max_cycles): reported, not fatal
            failed = {"step": step, "error": str(e), "ms": 1e3 * (time.perf_counter() - t0)}
            log(f"[bench] SL bubble {N}^3 step {step}: {e}")
            break
        torch.cuda.synchronize()
        rows.append({"ms": 1e3 * (time.perf_counter() - t0), "fas_cycles": cyc, "final_resid": res})
        log(f"[bench] SL bubble {N}^3: {rows[-1]['ms']:.0f} ms, {cyc} FAS cycles, resid {res:.2e}")
        if step == 0 and o is not None:
            o.push_hints("elvira", h0)
            o.push_hints("elvira", ctx.elvira_choice())
            o.push_hints("curvature", _directions(ctx))
            c0 = time.perf_counter()
            _, _, ocyc, ores = o.ns_step(t, param.tol)
            osec = time.perf_counter() - c0
            m = ctx.mesh
            n = m.nghost
            cin = (slice(n, n + m.Nx), slice(n, n + m.Ny), slice(n, n + m.Nz))
            Pg, Po = ctx.get("P")[cin], o.get("P")[cin]
            cmp = {"oracle_ms": 1e3 * osec, "oracle_cores": cpu_threads(), "oracle_fas_cycles": ocyc,
                   "max_abs_dVF": float(np.abs(ctx.get("VF") - o.get("VF")).max()),
                   "band_equal": bool(np.array_equal(ctx.get("band"), o.get("band"))),
                   "max_abs_dP": float(np.abs((Pg - Pg.mean()) - (Po - Po.mean())).max()),
                   "max_abs_P": float(np.abs(Po - Po.mean()).max()),
                   "max_rel_du": max(float(np.abs(ctx.get(k) - o.get(k)).max() / max(np.abs(o.get(k)).max(), 1e-300))
                                     for k in ("u", "v", "w")),
                   "elvira_hints_refused": o.tie_stats("elvira")["refused"]}
            log(f"[bench] SL bubble step 0 vs oracle ({cmp['oracle_ms']:.0f} ms on {cmp['oracle_cores']} cores): "
                f"max|dVF| {cmp['max_abs_dVF']:.2e}, max|dP| {cmp['max_abs_dP']:.2e} of {cmp['max_abs_P']:.2e}")
        t = t1
    ctx.close()
    out = {"workload": f"rising bubble {N}^3 as above, semi-Lagrangian scheme, SL pressure by FAS "
                       f"(2 levels, res_iteration smoother) to 1e-10", "steps": rows,
           "ms_per_step": sum(r["ms"] for r in rows) / len(rows) if rows else None,
           "not_converged": failed}
    if o is not None and rows:
        out["step0_vs_oracle"] = cmp
        out["speedup_vs_cpu_step0"] = cmp["oracle_ms"] / rows[0]["ms"]
        o.close()
    return out


def cpu_bubble(args, param, ic, nsamples=2):
    """Oracle rising-bubble step time on the host cores, from device states of the run, and the
    device's step from the same state beside it: max |dVF|, max |dP| after mean alignment (both
    solves stop below the same 1e-10 residual), max |d divg|, max |du| relative to max |u|. The
    bubble sits on the grid's x = y symmetry: ELVIRA and the height directions take the device's
    choices where the oracle's scores tie with them to rounding."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from ns_multiphase_amd import solver
    from ns_multiphase_amd.params import FIELD
    threads = cpu_threads()
    pyoracle.set_threads(threads)
    g = solver.Context(param)
    g.init_vof(*ic)
    g.ns_init()
    o = pyoracle.OracleContext(param)
    m = g.mesh
    n = m.nghost
    cin = (slice(n, n + m.Nx), slice(n, n + m.Ny), slice(n, n + m.Nz))
    t = 0.0
    at = [0] + [args.bubble_steps // 2] * (nsamples - 1)
    done = 0
    cpu_time = 0.0
    err = {"steps_compared": 0, "max_abs_dVF": 0.0, "band_equal": True, "max_abs_dP": 0.0, "max_abs_P": 0.0,
           "max_abs_ddivg": 0.0, "max_rel_du": 0.0}
    for s in at:
        while done < s:
            t, _, _, _ = g.ns_step(t, param.tol, "fp64")
            done += 1
        for name in FIELD:
            if not name.endswith("_bar"):  # the adjoint's fields have no oracle counterpart
                o.set(name, g.get(name))
        h0 = g.elvira_choice()
        t1, _, _, _ = g.ns_step(t, param.tol, "fp64")
        done += 1
        o.push_hints("elvira", h0)
        o.push_hints("elvira", g.elvira_choice())
        o.push_hints("curvature", _directions(g))
        t0 = time.perf_counter()
        o.ns_step(t, param.tol, "fp64")
        cpu_time += time.perf_counter() - t0
        err["steps_compared"] += 1
        err["max_abs_dVF"] = max(err["max_abs_dVF"], float(np.abs(g.get("VF") - o.get("VF")).max()))
        err["band_equal"] = err["band_equal"] and bool(np.array_equal(g.get("band"), o.get("band")))
        Pg, Po = g.get("P")[cin], o.get("P")[cin]
        err["max_abs_dP"] = max(err["max_abs_dP"], float(np.abs((Pg - Pg.mean()) - (Po - Po.mean())).max()))
        err["max_abs_P"] = max(err["max_abs_P"], float(np.abs(Po - Po.mean()).max()))
        err["max_abs_ddivg"] = max(err["max_abs_ddivg"], float(np.abs(g.get("divg") - o.get("divg")).max()))
        for name in ("u", "v", "w"):
            a, b = g.get(name), o.get(name)
            err["max_rel_du"] = max(err["max_rel_du"], float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))
        t = t1
    for kind in ("elvira", "curvature"):
        st = o.tie_stats(kind)
        err[f"{kind}_ties_adopted"] = st["adopted"]
        err[f"{kind}_hints_refused"] = st["refused"]
        err[f"{kind}_max_tie_gap"] = st["gap"]
    g.close()
    log(f"  cpu bubble: {len(at)} oracle steps in {cpu_time:.1f} s; per step vs device: max|dVF| "
        f"{err['max_abs_dVF']:.2e}, max|dP| {err['max_abs_dP']:.2e} of {err['max_abs_P']:.2e}, "
        f"max|d divg| {err['max_abs_ddivg']:.2e}")
    return {"ms_per_step": 1e3 * cpu_time / len(at), "cores": threads, "kind": "port",
            "sample": f"{len(at)} oracle steps of the {param.Nx}^3 rising bubble (C++ FP64 restatement, "
                      f"FP64 multigrid to the same 1e-10 residual), from the device state at steps {at}"}, err


def run_config1(args, torch):
    """The 2D single-vortex deformation of the CPU reference's own runs (128^2, circle r=0.15 at
    (0.5, 0.75), T = 4, dt = 1/256, 1024 steps): the oracle's whole period on the host cores and
    the device's beside it, with each side's shape error at t = T."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from ns_multiphase_amd import solver
    from ns_multiphase_amd import workloads as W
    threads = cpu_threads()
    pyoracle.set_threads(threads)
    param, ic = W.deformation2d(128)
    out = {"workload": "2D deformation 128^2, circle r=0.15 at (0.5,0.75), T=4, CFL 0.5, dt=1/256, RK4, 5 tets"}
    cells = param.Nx * param.Ny
    for side in ("oracle", "device"):
        c = pyoracle.OracleContext(param) if side == "oracle" else solver.Context(param)
        c.init_vof(*ic)
        c.vof_init()
        m = c.mesh
        o = m.nghost + 3
        VF0 = c.get("VF")[o:o + m.Nx, o:o + m.Ny, o].copy()
        if side == "device":
            torch.cuda.synchronize()
        t = 0.0
        steps = 0
        t0 = time.perf_counter()
        while t < param.tFinal:
            t, _ = c.vof_step(t, param.tFinal)
            steps += 1
        if side == "device":
            torch.cuda.synchronize()
        el = time.perf_counter() - t0
        VF = c.get("VF")[o:o + m.Nx, o:o + m.Ny, o]
        out[side] = {"steps": steps, "seconds": el, "cell_updates_per_s": cells * steps / el,
                     "shape_error_at_T": float(np.abs(VF - VF0).sum()) / cells}
        if side == "oracle":
            out[side]["cores"] = threads
        out[f"{side}_VF_T"] = VF
    out["max_abs_dVF_at_T"] = float(np.abs(out.pop("oracle_VF_T") - out.pop("device_VF_T")).max())
    out["note"] = ("trajectories over 1024 steps: ulp-level differences (libm, sum order) reach ELVIRA "
                   "near-ties on the field's x = 1/2 symmetry, so the two VF at T differ at the 1e-5 to 1e-4 level "
                   "while each shape error is that of the scheme; per-step parity is in per_step_vs_oracle")
    log(f"[bench] config 1: oracle {out['oracle']['cell_updates_per_s']:.3e} cell-updates/s on {threads} threads, "
        f"device {out['device']['cell_updates_per_s']:.3e}")
    return out


def forward_mode_check(args, N=64, nsteps=64, every=8):
    """The device's gradient (checkpoints every 8) against the oracle's forward mode (dual
    numbers through the same steps, oracle/oracle_fwd.cpp) at 64^2, 64 steps, the same 16 modes
    and seeds as the adjoint workload: both are exact derivatives of the map with its branches
    held, so no step size enters."""
    import numpy as np
    from ns_multiphase_amd import adjoint as A
    from ns_multiphase_amd import solver
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.set_threads(cpu_threads())
    p = A.adjoint_params(N)
    ctx = solver.Context(p)
    ctx.init_vof("circle", 0.2, 0.5, 0.5)
    vf0 = ctx.get("VF")
    basis = A.mode_faces(ctx.mesh)
    a1 = np.random.default_rng(1).normal(0.0, 0.1, 16)
    a2 = np.random.default_rng(2).normal(0.0, 0.1, 16)
    dt = min(A.stable_dt(ctx.mesh, basis, a1), A.stable_dt(ctx.mesh, basis, a2))
    run = A.SLVOFRun(ctx, vf0, basis, dt, nsteps, every)
    target = run.forward(a2, checkpoints=False)
    L, g = run.loss_and_grad(a1, target)
    o = pyoracle.OracleContext(p)
    o.set("VF", vf0)
    Bu = np.stack([b[0] for b in basis])
    Bv = np.stack([b[1] for b in basis])
    uf = (a1 @ run._Bu).reshape(run._shape)
    vf = (a1 @ run._Bv).reshape(run._shape)
    t0 = time.perf_counter()
    VF, dVF = o.fwd_vof_run(uf, vf, np.zeros_like(uf), Bu, Bv, nsteps, dt)
    sec = time.perf_counter() - t0
    m = ctx.mesh
    oo = m.nghost + 3
    sl = (slice(oo, oo + m.Nx), slice(oo, oo + m.Ny), slice(oo, oo + m.Nz))
    diff = VF[sl] - target[sl]
    gf = np.array([float((diff * dVF[q][sl]).sum()) for q in range(16)])
    rel = np.abs(g - gf) / np.maximum(np.abs(gf), 1e-3 * np.abs(gf).max())
    ctx.close()
    o.close()
    out = {"workload": f"{N}^2 x 1, {nsteps} steps, checkpoints every {every}, same modes and seeds",
           "within_1e-6": int((rel <= 1e-6).sum()), "directions": 16, "max_rel": float(rel.max()),
           "loss_rel_diff": abs(0.5 * float((diff ** 2).sum()) - L) / L, "oracle_seconds": sec,
           "oracle_cores": cpu_threads()}
    log(f"[bench] adjoint vs forward mode at {N}^2: {out['within_1e-6']}/16 within 1e-6, max rel {rel.max():.2e}")
    return out


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
    # each time is the median of `reps` synchronised calls (host-side jitter of single calls
    # of 20-60 ms is a few per cent)
    reps = 7

    def timed(fn):
        ts, out = [], None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), out
    t_fwd, L = timed(lambda: run.loss(a1, target))
    t_grad, (L2, g) = timed(lambda: run.loss_and_grad(a1, target))
    t_grad8, (_, g8) = timed(lambda: run8.loss_and_grad(a1, target))
    # the reverse sweep alone (recomputing each segment from its checkpoint, when the interval
    # is above 1, then walking it back), after a forward that stored the checkpoints
    m = ctx.mesh
    o = m.nghost + 3
    sl = (slice(o, o + m.Nx), slice(o, o + m.Ny), slice(o, o + m.Nz))
    tb = []
    for _ in range(reps):
        VF = run.forward(a1)
        diff = np.zeros_like(VF)
        diff[sl] = VF[sl] - target[sl]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.backward(diff)
        torch.cuda.synchronize()
        tb.append(time.perf_counter() - t0)
    t_bwd = float(np.median(tb))
    fwd = None if args.no_oracle else forward_mode_check(args)
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
           "timing": f"medians of {reps} synchronised calls each",
           "checkpoint_MB": (run.nseg + every - 1) * A.checkpoint_bytes(m) / 1e6,
           "loss": L, "forward_ms": 1e3 * t_fwd, "loss_and_gradient_ms": 1e3 * t_grad,
           "gradient_cost_vs_forward": t_grad / t_fwd, "reverse_sweep_ms": 1e3 * t_bwd,
           "reverse_sweep_vs_forward": t_bwd / t_fwd,
           "loss_and_gradient_ms_every8": 1e3 * t_grad8, "gradient_cost_vs_forward_every8": t_grad8 / t_fwd,
           "every8_vs_every_max_rel": float(np.abs(g8 - g).max() / np.abs(g).max()),
           "fd_check": checks[0], "fd_check_h1e-7": checks[1], "forward_mode_check": fwd}
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
    dev_P = {}
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
        dev_P[prec] = Pn - Pn.mean()
        ctx.close()
        log(f"[bench] {N}^3 MMS {prec}: {ms:.1f} ms, {cyc} cycles, L-inf {err:.2e}")
    out["mg_speedup_mixed_vs_fp64"] = out["mg_fp64"]["solve_ms"] / out["mg_mixed"]["solve_ms"]
    if not args.no_oracle:
        P_ora, oinfo = oracle_mms(N, dt)
        for prec in ("fp64", "mixed"):
            out[f"mg_{prec}"]["vs_oracle"] = mms_vs_oracle(dev_P[prec], P_ora, oinfo)
        del P_ora
        log(f"[bench] {N}^3 MMS vs oracle: max|dP| fp64 {out['mg_fp64']['vs_oracle']['max_abs_dP']:.2e}, "
            f"mixed {out['mg_mixed']['vs_oracle']['max_abs_dP']:.2e} (oracle {oinfo['seconds']:.1f} s)")
    dev_P.clear()
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


def guarded(name, fn):
    """A section of the line: its result, or its error (the other sections and the headline still
    print)."""
    try:
        return fn()
    except Exception as e:  # noqa: BLE001 - reported in the JSON line
        log(f"[bench] {name} failed: {e!r}")
        return {"error": repr(e)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: one period, 1536 at 128^3)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--mg-N", type=int, default=128)
    ap.add_argument("--mg-reps", type=int, default=3)
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="no oracle timing (the operation counts stay)")
    ap.add_argument("--no-oracle", action="store_true", help="no oracle pass at all (no rooflines)")
    ap.add_argument("--no-vof", action="store_true", help="skip the headline deformation")
    ap.add_argument("--bubble-N", type=int, default=128)
    ap.add_argument("--bubble-steps", type=int, default=50)
    ap.add_argument("--no-bubble", action="store_true")
    ap.add_argument("--sl-N", type=int, default=32)
    ap.add_argument("--sl-steps", type=int, default=3)
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

    value = steps = warm = elapsed = None
    kernels, info = {}, None
    param, ic = W.deformation3d(args.N)
    if not args.no_vof:
        log(f"[bench] SL-VOF deformation3d N={args.N}")
        ctx, param, ic, steps, warm, elapsed, prof, info = run_vof(args, torch)
        ctx.close()
        cells = param.Nx * param.Ny * param.Nz
        value = cells * steps / elapsed
        log(f"[bench] {steps} steps in {elapsed:.3f} s -> {value:.3e} cell-updates/s")
        kernels = {k: {"ms": v[0], "launches": v[1], "avg_us": (1e3 * v[0] / v[1]) if v[1] else None,
                       "units_per_launch": v[2] / v[1] if v[1] else None}
                   for k, v in prof.items() if v[1]}

    mg = None
    if not args.no_mg:
        def _mg_section():
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
            if not args.no_oracle:
                P_ora, oinfo = oracle_mms(args.mg_N, 1e-3)
                for prec in ("fp64", "mixed"):
                    mg[prec]["vs_oracle"] = mms_vs_oracle(mg_raw[prec]["P"], P_ora, oinfo)
                log(f"[bench] mg vs oracle: max|dP| fp64 {mg['fp64']['vs_oracle']['max_abs_dP']:.2e}, "
                    f"mixed {mg['mixed']['vs_oracle']['max_abs_dP']:.2e} (oracle {oinfo['seconds']:.1f} s)")
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
            return mg

        mg = guarded("mg", _mg_section)

    bubble = None
    if not args.no_bubble:
        def _bubble_section():
            log(f"[bench] rising bubble N={args.bubble_N}, {args.bubble_steps} steps")
            bparam, bic, bubble = run_bubble(args, torch)
            if not args.no_cpu_baseline:
                bubble["cpu_baseline"], bubble["per_step_vs_oracle"] = cpu_bubble(args, bparam, bic)
                bubble["speedup_vs_cpu"] = {prec: bubble["cpu_baseline"]["ms_per_step"] / bubble[prec]["ms_per_step"]
                                            for prec in ("mixed", "fp64")}
            bubble["workload"] = (f"rising bubble {args.bubble_N}^3, r=0.125 at (0.5,0.5,0.3), rho 1000/1, "
                                  f"mu 1e-3/1e-5, sigma 0.07, g=-9.81 z, no-slip walls, dt=1e-3, CFL 0.5, "
                                  f"finite-difference scheme with correction tets, MG tol 1e-10")
            return bubble

        bubble = guarded("rising_bubble", _bubble_section)

    bubble_sl = None
    if not args.no_bubble and args.sl_steps > 0:
        bubble_sl = guarded("rising_bubble_sl", lambda: run_bubble_sl(args, torch))

    config1 = None
    if not args.no_cpu_baseline and not args.no_oracle:
        log("[bench] config 1: 2D deformation 128^2 (oracle and device)")
        config1 = guarded("deformation2d_128", lambda: run_config1(args, torch))

    adjoint = None
    if not args.no_adjoint:
        log(f"[bench] adjoint {args.adj_N}^2")
        adjoint = guarded("adjoint", lambda: run_adjoint(args, torch))
    large = None
    if not args.no_large:
        log(f"[bench] large {args.large_N}^3")
        large = guarded("large", lambda: run_large(args, torch))

    cpu = accuracy = None
    roofs = {}
    if info is not None and not args.no_oracle:
        log("[bench] oracle on the sampled states (cpu baseline, per-step error, operation counts)")
        try:
            cpu, accuracy, fc = cpu_baseline(args, param, ic, info, time_cpu=not args.no_cpu_baseline)
        except Exception as e:  # noqa: BLE001 - reported in the JSON line; no rooflines then
            log(f"[bench] cpu_baseline failed: {e!r}")
            cpu, accuracy = {"error": repr(e)}, None
    # rooflines of the two FP64 kernels of the step; the dominant one by time is `roofline`
    for key, kern, per, unit_name in (("elvira", "k_elvira_dist + k_elvira", fc["elvira_per_interface_cell"],
                                       "interface_cells"),
                                      ("sl_cut", "k_sl_tree", fc["cut_per_band_cell"], "band_cells")) \
            if accuracy is not None else ():
        ms, n, units = prof[key]
        if n <= 0 or ms <= 0:
            continue
        avg_s = ms / n / 1e3
        achieved = per * units / n / avg_s / 1e12
        roofs[key] = {"bound": "valu_fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                      "frac": achieved / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_PER_UNIT[kern] * units / n,
                      "kernel": kern, "avg_launch_us": 1e6 * avg_s, "launches": n, "total_ms": ms,
                      f"{unit_name}_per_launch": units / n, f"flops_per_{unit_name[:-1]}": per,
                      "note": "FP64 vector (VALU) peak of MI355X, 78.6 TFLOP/s with FMA = 2 flops; the kernels "
                              "form no FMA (-ffp-contract=off), so half of it is their ceiling; flops per unit "
                              f"counted by the oracle at {fc['N']}^3 on the states at steps {fc['states']}"}
    roof = max(roofs.values(), key=lambda r: r["total_ms"]) if roofs else None

    line = {
        "metric": "SL-VOF cell-updates/s at 128^3 + MG solve time, mixed vs FP64, at 1e-10 resid",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": 1e3 * elapsed / steps if steps else None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D deformation {args.N}^3, sphere r=0.15 at (0.35,0.35,0.35), T=3, CFL 0.5, "
                               f"dt=1/512, RK4, 5 tets",
                   "timed_steps": None if info is None else
                                  (f"one step from each of {steps} device states at steps {info['samples'][:3]} ... "
                                   f"{info['samples'][-1]} of the {int(round(param.tFinal / param.max_dt))}-step period "
                                   "(checkpoint restore + step, inside the timed region)") if info["sampled"] else
                                  f"the whole period from t = 0, {steps} steps",
                   "period_volume_rel_change": None if info is None else (info["V1"] - info["V0"]) / info["V0"],
                   "shape_error_at_T": None if info is None else info["shape_err"],
                   "t_end": None if info is None else info["t_end"],
                   "mg_workload": f"variable-density MMS {args.mg_N}^3, rho 1000/1, tol 1e-10"},
        "roofline": roof, "cpu_baseline": cpu, "per_step_vs_oracle": accuracy,
        "mg": mg, "rising_bubble": bubble, "rising_bubble_sl": bubble_sl, "deformation2d_128": config1,
        "adjoint": adjoint,
        "large": large, "kernels": kernels,
        "rooflines": roofs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
