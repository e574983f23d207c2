"""Restart round trip (run_solver's restart branch, src/NavierStokes_Parallel.jl:59-67, with
the restart arrays of src/WriteData.jl:163-184 / src/ReadData.jl:3-53): the fields a restart
file holds (VF, uf, vf, wf, P on the cell extents) are downloaded, uploaded into a fresh
context, and mist_restart_init rebuilds the rest.

* VOF transport (solveNS = false): the restarted run continues the uninterrupted one bit for
  bit (the velocity is prescribed, the step is max_dt, and nothing else carries over).
* Navier-Stokes: the reference's restart rebuilds u, v, w from the faces with ghosts from a
  fresh us, vs, ws, and starts a fresh mask, so it is not the uninterrupted run; the device's
  restart and the following step are checked against the oracle's.
"""
import numpy as np
import pytest

from ns_multiphase_amd import workloads as W
from test_gpu_ns import _close, directions

pytestmark = pytest.mark.gpu

RESTART_FIELDS = ("uf", "vf", "wf", "P")


@pytest.fixture(scope="module")
def mist():
    from ns_multiphase_amd import solver
    solver.lib()
    return solver


def restart_arrays(ctx):
    """What a restart file holds: VF, uf, vf, wf, P over the cell extents (VF's 3 extra ghost
    layers are not written)."""
    out = {n: ctx.get(n) for n in RESTART_FIELDS}
    vf = ctx.get("VF").copy()
    inner = np.zeros_like(vf, bool)
    inner[3:-3, 3:-3, 3:-3] = True
    vf[~inner] = 0.0
    out["VF"] = vf
    return out


def test_restart_vof_continues_bit_exact(mist):
    param, ic = W.deformation3d(24)
    a = mist.Context(param)
    a.init_vof(*ic)
    a.vof_init()
    t = 0.0
    for _ in range(8):
        t, _ = a.vof_step(t)
    b = mist.Context(param)
    for n, arr in restart_arrays(a).items():
        b.set(n, arr)
    b.restart_init()
    ta = tb = t
    for s in range(4):
        ta, dta = a.vof_step(ta)
        tb, dtb = b.vof_step(tb)
        assert ta == tb and dta == dtb
        assert np.array_equal(a.get("VF"), b.get("VF")), s
        assert np.array_equal(a.get("band"), b.get("band")), s
        for n in ("nx", "ny", "nz", "D"):
            assert np.array_equal(a.get(n), b.get(n)), (s, n)


@pytest.mark.parametrize("scheme", ["finite-difference", "semi-lagrangian"])
def test_restart_ns_matches_oracle(mist, oracle, scheme):
    kw = {"pressure_scheme": scheme}
    if scheme == "semi-lagrangian":
        kw.update(pressureSolver="res_iteration", mg_lvl=2)
    param, ic = W.rising_bubble(16, **kw)
    a = mist.Context(param)
    a.init_vof(*ic)
    a.ns_init()
    t = 0.0
    for _ in range(2):
        t, _, _, _ = a.ns_step(t, param.tol)
    arrays = restart_arrays(a)
    b = mist.Context(param)
    o = oracle.OracleContext(param)
    for c in (b, o):
        for n, arr in arrays.items():
            c.set(n, arr)
    dtb = b.restart_init()
    # the oracle's ELVIRA takes the device's choices where they tie (the stated bubble sits on
    # the grid's x = y symmetry)
    o.push_hints("elvira", b.elvira_choice())
    dto = o.restart_init()
    assert dtb == pytest.approx(dto, rel=1e-14)
    for n in ("u", "v", "w", "denx", "deny", "denz", "viscx", "viscy", "viscz", "divg"):
        _close(b.get(n), o.get(n), 1e-12, n)
    assert np.array_equal(b.get("band"), o.get("band"))
    for n in ("nx", "ny", "nz", "D"):
        _close(b.get(n), o.get(n), 1e-12, n)
    h0 = b.elvira_choice()
    tb, _, _, rb = b.ns_step(t, param.tol)
    o.push_hints("elvira", h0)
    o.push_hints("elvira", b.elvira_choice())
    o.push_hints("curvature", directions(b))
    to, _, _, ro = o.ns_step(t, param.tol)
    assert tb == pytest.approx(to, rel=1e-14) and rb < param.tol and ro < param.tol
    _close(b.get("VF"), o.get("VF"), 1e-12, "VF")
    for n in ("us", "vs", "ws"):
        _close(b.get(n), o.get(n), 1e-11, n)
    # the finite-difference pressure solves agree to rounding; the SL-SL FAS solves stop at
    # different iterates below the same residual (Anderson's small solves are ordered differently)
    vel_bar = 1e-12 if scheme == "finite-difference" else 1e-7
    for n in ("u", "v", "w"):
        _close(b.get(n), o.get(n), vel_bar, n)
    assert o.tie_stats("elvira")["refused"] == 0
