"""Scenarios the sanitizer build of the oracle runs (tests/test_sanitizers.py starts this script
in a child process with GCC's AddressSanitizer runtime preloaded and MIST_ORACLE_LIB pointing
at tests/_build/liboracle_san.so). Every entry point of the oracle's C interface that the
tests use is exercised once at a small size: ICs, band, ELVIRA (with tie hints), PLIC,
transport (VOF and the momentum predictor, both schemes), the multigrid in both precisions,
the SL-SL pressure (operator, Jacobian, FAS solve), the restart branch, the per-cell probes
and the batch helpers. Any out-of-bounds access, use after free or undefined behaviour aborts
the process with the sanitizer's report."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from ns_multiphase_amd import workloads as W  # noqa: E402


def step(msg):
    print(msg, flush=True)


def vof_cases():
    for name, (param, ic) in (("deformation3d 16^3", W.deformation3d(16)),
                              ("deformation2d 32^2", W.deformation2d(32))):
        step(name)
        o = pyoracle.OracleContext(param)
        o.init_vof(*ic)
        o.vof_init()
        t = 0.0
        for _ in range(3):
            t, _ = o.vof_step(t)
        i, j, k = [n // 2 for n in (param.Nx, param.Ny, param.Nz)]
        k = k if param.Nz > 1 else 1
        err, nrm, best = o.elvira_scores(i + 1, j + 1, k)
        o.push_hints("elvira", np.full((param.Nx, param.Ny, param.Nz), -1, np.int32))
        o.compute_normal()
        o.clear_hints()
        o.sl_cell(i, j, k, 1e-3)
        o.elvira_cell(i, j, k)
        o.cell2verts(i, j, k)
        o.reduce("sum", "VF")
        o.close()


def ns_cases():
    for scheme, solver in (("finite-difference", "multigrid"), ("semi-lagrangian", "res_iteration")):
        for prec in (("fp64", "mixed") if scheme == "finite-difference" else ("fp64",)):
            step(f"rising bubble 12^3, {scheme}, {prec}")
            param, ic = W.rising_bubble(12, pressure_scheme=scheme, pressureSolver=solver, mg_lvl=2,
                                        tol=1e-8)
            o = pyoracle.OracleContext(param)
            o.init_vof(*ic)
            o.ns_init()
            t, dt, cyc, res = o.ns_step(0.0, param.tol, prec)
            o.compute_curvature()
            o.surface_force()
            i = j = k = 6
            o.sl_flux_vol(0, i, j, k, dt)
            o.correction_tets(i, j, k, dt)
            if scheme == "semi-lagrangian":
                o.sl_operator(dt)
                o.sl_jacobian(dt)
            step("  restart branch")
            b = pyoracle.OracleContext(param)
            for n in ("uf", "vf", "wf", "P", "VF"):
                b.set(n, o.get(n))
            b.restart_init()
            b.ns_step(t, param.tol, prec)
            b.close()
            o.close()


def mg_cases():
    for prec in ("fp64", "mixed"):
        step(f"MMS 16^3 {prec}")
        p = W.mms_params(16)
        o = pyoracle.OracleContext(p)
        W.mms_setup(o, 1e-3)
        o.pressure_solve(1e-3, 1e-10, prec)
        o.mg_level_coef(1e-3, 0, 0)
        o.close()


def helpers():
    step("computeAlpha, d2case, tet_vol, Anderson")
    rng = np.random.default_rng(0)
    for _ in range(50):
        m = np.abs(rng.standard_normal(3))
        m /= m.sum()
        pyoracle.compute_alpha(m, float(rng.random()))
        pyoracle.d2case(rng.standard_normal(4))
        pyoracle.tet_vol(rng.random((4, 3)))
    g = rng.standard_normal((4, 4))
    pyoracle.anderson_coefficients(g @ g.T + np.eye(4))


def forward_mode():
    step("forward mode 16^2, 2 steps")
    from ns_multiphase_amd import adjoint as A
    p = A.adjoint_params(16)
    o = pyoracle.OracleContext(p)
    o.init_vof("circle", 0.2, 0.47, 0.52)
    basis = A.mode_faces(o.mesh)
    Bu = np.stack([b[0] for b in basis])
    Bv = np.stack([b[1] for b in basis])
    a = np.random.default_rng(1).normal(0.0, 0.1, 16)
    uf = (a @ Bu.reshape(16, -1)).reshape(Bu.shape[1:])
    vf = (a @ Bv.reshape(16, -1)).reshape(Bv.shape[1:])
    o.fwd_vof_run(uf, vf, np.zeros_like(uf), Bu, Bv, 2, A.stable_dt(o.mesh, basis, a))
    o.close()


if __name__ == "__main__":
    step(f"oracle library: {pyoracle.LIB_PATH}")
    pyoracle.set_threads(2)
    helpers()
    vof_cases()
    mg_cases()
    ns_cases()
    forward_mode()
    step("all scenarios done")
