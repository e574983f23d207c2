"""Workload builders: the parameter sets and inputs of the measured configurations.

* ``deformation2d``: single-vortex deformation on a 128x128x1 slab (test/test_deformation.jl:
  circle r=0.15 at (0.5, 0.75), walls in x and y, periodic z), period T = 4, CFL 0.5, RK4,
  5 tets, plain SL transport, dt = max_dt = 0.5/N.
* ``deformation3d``: 3D deformation in the unit cube (examples/example5_deformation3D.jl:
  sphere r=0.15 at (0.35, 0.35, 0.35), walls), period T = 3, CFL 0.5, dt = max_dt = 0.25/N.
* ``rising_bubble``: gas bubble r = 0.125 at (0.5, 0.5, 0.3) in liquid filling the unit cube
  (VFbubble3d), no-slip walls (the reflect BC of examples/example_viscous_bubble3D.jl:96-146
  with zero wall velocity), grav_z = +9.81 (downward), rho 1000/1, mu 1e-3/1e-5,
  sigma 0.07, max_dt = 1e-3, CFL 0.5, finite-difference pressure scheme (transport with
  correction tets), multigrid down to 4 cells per side.
* ``mms_setup``: variable-density pressure problem after test/pressure_var_den_MMS.jl in 3D:
  exact p = cos(2 pi x) cos(2 pi y) cos(2 pi z), interface z = 0.5 + 0.05 sin(2 pi x) sin(2 pi y)
  with the heavy phase below, face velocities uf = dt/rho_f dp/dx at the faces
  (pressure_var_den_MMS.jl:92-93) and b = div(u) on the device.

The builders only prepare inputs; they work on any object with the stage methods of
``solver.Context`` (``get``/``set``/``compute_props``/``pressure_rhs``).
"""
from __future__ import annotations

import math

import numpy as np

from .params import parameters


def deformation2d(N: int = 128, T: float = 4.0, **kw) -> tuple[parameters, tuple]:
    base = dict(mu_liq=1.0, mu_gas=0.01, rho_liq=1.0, rho_gas=0.01, sigma=0.0, Lx=1.0, Ly=1.0,
                Lz=1.0 / N, tFinal=T, Nx=N, Ny=N, Nz=1, stepMax=100000, CFL=0.5, max_dt=0.5 / N,
                out_period=1, tol=1e-3, nprocx=1, nprocy=1, nprocz=1, xper=False, yper=False,
                zper=True, solveNS=False, VFVelocity="Deformation", vel_period=T,
                pressure_scheme="semi-lagrangian", projection_method="RK4", iter_type="standard")
    base.update(kw)
    return parameters(**base), ("circle", 0.15, 0.5, 0.75)


def deformation3d(N: int = 128, T: float = 3.0, **kw) -> tuple[parameters, tuple]:
    base = dict(mu_liq=1.0, mu_gas=0.01, rho_liq=1.0, rho_gas=0.01, sigma=0.0, Lx=1.0, Ly=1.0,
                Lz=1.0, tFinal=T, Nx=N, Ny=N, Nz=N, stepMax=100000, CFL=0.5, max_dt=0.25 / N,
                out_period=1, tol=1e-3, nprocx=1, nprocy=1, nprocz=1, xper=False, yper=False,
                zper=False, solveNS=False, VFVelocity="Deformation3D", vel_period=T,
                pressure_scheme="semi-lagrangian", projection_method="RK4", iter_type="standard")
    base.update(kw)
    return parameters(**base), ("sphere", 0.15, 0.35, 0.35, 0.35)


def rising_bubble(N: int = 128, **kw) -> tuple[parameters, tuple]:
    lvl = max(1, int(math.log2(N)) - 1)  # coarsest level 4 cells per side
    base = dict(mu_liq=1e-3, mu_gas=1e-5, rho_liq=1000.0, rho_gas=1.0, sigma=0.07, Lx=1.0, Ly=1.0,
                Lz=1.0, tFinal=1.0, Nx=N, Ny=N, Nz=N, stepMax=50, CFL=0.5, max_dt=1e-3,
                out_period=1, tol=1e-10, nprocx=1, nprocy=1, nprocz=1, xper=False, yper=False,
                zper=False, grav_z=9.81, solveNS=True, pressure_scheme="finite-difference",
                pressureSolver="multigrid", mg_lvl=lvl, projection_method="RK4", iter_type="standard")
    base.update(kw)
    return parameters(**base), ("bubble3d", 0.125, 0.5, 0.5, 0.3)


def mms_params(N: int = 128, rho_heavy: float = 1000.0, rho_light: float = 1.0, **kw) -> parameters:
    base = dict(mu_liq=0.0, mu_gas=0.0, rho_liq=rho_heavy, rho_gas=rho_light, sigma=0.0, Lx=1.0,
                Ly=1.0, Lz=1.0, tFinal=1.0, Nx=N, Ny=N, Nz=N, stepMax=1, CFL=0.5, out_period=1,
                tol=1e-10, nprocx=1, nprocy=1, nprocz=1, xper=False, yper=False, zper=False,
                pressure_scheme="finite-difference", pressureSolver="gauss-seidel", mg_lvl=1)
    base.update(kw)
    return parameters(**base)


def interface_vf(mesh, amp: float = 0.05, z0: float = 0.5, npts: int = 6) -> np.ndarray:
    """Liquid fraction below z = z0 + amp sin(2 pi x) sin(2 pi y) on every cell of the cell
    extents (zero in the 3 extra VF ghost layers), by npts x npts Gauss-Legendre points per
    cell column."""
    g = mesh.nghost
    nx, ny, nz = mesh.cell_shape()
    gx, gw = np.polynomial.legendre.leggauss(npts)
    xs = (mesh.x[:-1, None] + 0.5 * mesh.dx * (gx[None, :] + 1.0))   # (nx, npts)
    ys = (mesh.y[:-1, None] + 0.5 * mesh.dy * (gx[None, :] + 1.0))
    w = 0.25 * gw[:, None] * gw[None, :]                                # weights sum to 1
    h = z0 + amp * np.sin(2 * np.pi * xs)[:, None, :, None] * np.sin(2 * np.pi * ys)[None, :, None, :]
    # (nx, ny, p, q): fraction of cell k below h is clamp((h - z_k)/dz, 0, 1)
    vf = np.zeros((nx, ny, nz))
    hmin, hmax = h.min(), h.max()
    for k, zlo in enumerate(mesh.z[:-1]):
        if zlo + mesh.dz <= hmin:
            vf[:, :, k] = 1.0
        elif zlo < hmax:
            frac = np.clip((h - zlo) / mesh.dz, 0.0, 1.0)
            vf[:, :, k] = np.einsum("ijpq,pq->ij", frac, w)
    out = np.zeros(mesh.vof_shape())
    out[3:3 + nx, 3:3 + ny, 3:3 + nz] = vf
    return out


def mms_setup(ctx, dt: float, amp: float = 0.05):
    """Upload VF, densities and face velocities of the MMS problem; set RHS = div(u)."""
    mesh = ctx.mesh
    ctx.set("VF", interface_vf(mesh, amp))
    ctx.compute_props()
    denx, deny, denz = ctx.get("denx"), ctx.get("deny"), ctx.get("denz")
    nx, ny, nz = mesh.cell_shape()
    tp = 2.0 * math.pi
    X, Ym, Zm = np.meshgrid(mesh.x[:-1], mesh.ym, mesh.zm, indexing="ij")
    uf = -tp * np.sin(tp * X) * np.cos(tp * Ym) * np.cos(tp * Zm) * dt / denx[:nx]
    Xm, Y, Zm = np.meshgrid(mesh.xm, mesh.y[:-1], mesh.zm, indexing="ij")
    vf = -tp * np.cos(tp * Xm) * np.sin(tp * Y) * np.cos(tp * Zm) * dt / deny[:, :ny]
    Xm, Ym, Z = np.meshgrid(mesh.xm, mesh.ym, mesh.z[:-1], indexing="ij")
    wf = -tp * np.cos(tp * Xm) * np.cos(tp * Ym) * np.sin(tp * Z) * dt / denz[:, :, :nz]
    ctx.set("uf", uf)
    ctx.set("vf", vf)
    ctx.set("wf", wf)
    ctx.pressure_rhs()


def mms_exact(mesh) -> np.ndarray:
    """Exact pressure at the interior cell centres, shifted to zero mean."""
    g = mesh.nghost
    tp = 2.0 * math.pi
    xm = mesh.xm[g:g + mesh.Nx]
    ym = mesh.ym[g:g + mesh.Ny]
    zm = mesh.zm[g:g + mesh.Nz]
    p = (np.cos(tp * xm)[:, None, None] * np.cos(tp * ym)[None, :, None] * np.cos(tp * zm)[None, None, :])
    return p - p.mean()


def interior(mesh, a, vof=False):
    """Interior block [i, j, k] of a field array in the reference layout."""
    o = mesh.nghost + (3 if vof else 0)
    return a[o:o + mesh.Nx, o:o + mesh.Ny, o:o + mesh.Nz]
