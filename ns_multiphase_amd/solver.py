"""Host side of the drop-in: ctypes binding of libmist_hip.so and a device-resident
``Context`` whose methods mirror the reference's stage functions.

The reference mutates caller-owned ghosted OffsetArrays; here the arrays live in device
memory inside a context and cross the boundary as flat copies of the same ghosted,
column-major layout (``get``/``set``). Errors of the C ABI become ``MistError`` with the
library's message, as the reference raises ``error(...)``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import FIELD, IC_SHAPES, REDUCE, SIDES, MistParams, create_mesh, parameters

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmist_hip.so")

_lib = None


class MistError(RuntimeError):
    """Raised for a non-zero status of the C ABI."""

    def __init__(self, status, msg):
        super().__init__(f"mist status {status}: {msg}")
        self.status = status


def lib():
    """Load libmist_hip.so (built in-tree by ``ns_multiphase_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ns_multiphase_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, d, i64, i32 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "mist_params_default": (i32, [ctypes.POINTER(MistParams)]),
        "mist_ctx_create": (i32, [ctypes.POINTER(MistParams), i32, ctypes.POINTER(vp)]),
        "mist_ctx_destroy": (i32, [vp]),
        "mist_last_error": (ctypes.c_char_p, [vp]),
        "mist_field_extents": (i32, [vp, i32, ip, ip, ip]),
        "mist_field_device_ptr": (i32, [vp, i32, ctypes.POINTER(vp)]),
        "mist_upload": (i32, [vp, i32, vp, ctypes.c_size_t]),
        "mist_download": (i32, [vp, i32, vp, ctypes.c_size_t]),
        "mist_init_vof": (i32, [vp, i32, dp, i32]),
        "mist_neumann": (i32, [vp, i32]),
        "mist_update_borders": (i32, [vp, i32]),
        "mist_define_velocity": (i32, [vp, d]),
        "mist_compute_dt": (i32, [vp, dp]),
        "mist_compute_band": (i32, [vp]),
        "mist_compute_normal": (i32, [vp]),
        "mist_compute_plic": (i32, [vp]),
        "mist_transport": (i32, [vp, d, d]),
        "mist_compute_props": (i32, [vp]),
        "mist_pressure_rhs": (i32, [vp]),
        "mist_pressure_solve": (i32, [vp, d, d, i64, ip, dp]),
        "mist_corrector": (i32, [vp, d]),
        "mist_reduce": (i32, [vp, i32, i32, dp]),
        "mist_vof_step": (i32, [vp, dp, d, dp]),
        "mist_vof_init": (i32, [vp]),
        "mist_plic_dist_batch": (i32, [vp, ip, dp, dp, dp, i64]),
        "mist_plic_vf_batch": (i32, [vp, ip, dp, dp, dp, i64]),
        "mist_plic_dist_batch_noalloc": (i32, [vp, ip, dp, dp, dp, i64]),
        "mist_stage_times": (i32, [vp, dp, i32]),
        "mist_profile_enable": (i32, [vp, i32]),
        "mist_profile_read": (i32, [vp, i32, dp, ip, ip]),
        "mist_set_wall_velocity": (i32, [vp, i32, dp]),
        "mist_apply_bc": (i32, [vp, i32, i32, i32]),
        "mist_interpolate_face": (i32, [vp, i32, i32, i32]),
        "mist_interpolate_center": (i32, [vp]),
        "mist_divergence": (i32, [vp, d]),
        "mist_compute_curvature": (i32, [vp]),
        "mist_surface_force": (i32, [vp]),
        "mist_ns_init": (i32, [vp]),
        "mist_ns_step": (i32, [vp, dp, d, dp, d, i64, ip, dp]),
        "mist_transport_vjp": (i32, [vp, d]),
        "mist_checkpoint_save": (i32, [vp, i32]),
        "mist_set_deferred_checks": (i32, [vp, i32]),
        "mist_checkpoint_restore": (i32, [vp, i32]),
        "mist_elvira_choice": (i32, [vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t]),
        "mist_pressure_operator": (i32, [vp, d]),
        "mist_restart_init": (i32, [vp, dp]),
        "mist_pressure_jacobian": (i32, [vp, d, dp, ctypes.c_size_t]),
        "mist_field_modified": (i32, [vp, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class Context:
    """Device-resident state of one mesh on one GPU (a ``mist_ctx``)."""

    def __init__(self, param: parameters, device: int = 0):
        self.param = param
        self.mesh = create_mesh(param)
        cp = param.to_c()
        h = ctypes.c_void_p()
        st = lib().mist_ctx_create(ctypes.byref(cp), device, ctypes.byref(h))
        if st:
            raise MistError(st, lib().mist_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mist_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st:
            raise MistError(st, lib().mist_last_error(self.h).decode())

    # ---- fields (reference layout, indexed [i, j, k]) ---------------------------------
    def extents(self, name):
        lo = np.zeros(3, np.int64)
        n = np.zeros(3, np.int64)
        el = ctypes.c_int64()
        self._check(lib().mist_field_extents(self.h, FIELD[name], _ip(lo), _ip(n), ctypes.byref(el)))
        return lo, n, el.value

    def get(self, name):
        lo, n, el = self.extents(name)
        a = np.zeros(tuple(n[::-1]), dtype=np.int16 if el == 2 else np.float64)
        self._check(lib().mist_download(self.h, FIELD[name], a.ctypes.data, a.nbytes))
        return a.transpose(2, 1, 0)

    def set(self, name, arr):
        lo, n, el = self.extents(name)
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.int16 if el == 2 else np.float64).transpose(2, 1, 0))
        if a.shape != tuple(n[::-1]):
            raise ValueError(f"{name}: shape {np.shape(arr)} != {tuple(n)}")
        self._check(lib().mist_upload(self.h, FIELD[name], a.ctypes.data, a.nbytes))

    def device_ptr(self, name):
        p = ctypes.c_void_p()
        self._check(lib().mist_field_device_ptr(self.h, FIELD[name], ctypes.byref(p)))
        return p.value

    def field_modified(self, name):
        """Declare a write through ``device_ptr(name)``: derived state is recomputed."""
        self._check(lib().mist_field_modified(self.h, FIELD[name]))

    def elvira_choice(self):
        """ELVIRA candidate kept per interior cell for the current VF, indexed [i, j, k]
        (-1 outside interface cells)."""
        m = self.mesh
        out = np.zeros(m.Nx * m.Ny * m.Nz, np.int32)
        self._check(lib().mist_elvira_choice(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), out.size))
        return out.reshape(m.Nz, m.Ny, m.Nx).transpose(2, 1, 0)

    # ---- stages -----------------------------------------------------------------------
    def init_vof(self, shape, *args):
        a = np.asarray(args, np.float64)
        self._check(lib().mist_init_vof(self.h, IC_SHAPES[shape], _dp(a), len(a)))

    def neumann(self, name):
        self._check(lib().mist_neumann(self.h, FIELD[name]))

    def update_borders(self, name):
        self._check(lib().mist_update_borders(self.h, FIELD[name]))

    def define_velocity(self, t):
        self._check(lib().mist_define_velocity(self.h, t))

    def compute_dt(self):
        dt = ctypes.c_double()
        self._check(lib().mist_compute_dt(self.h, ctypes.byref(dt)))
        return dt.value

    def compute_band(self):
        self._check(lib().mist_compute_band(self.h))

    def compute_normal(self):
        self._check(lib().mist_compute_normal(self.h))

    def compute_plic(self):
        self._check(lib().mist_compute_plic(self.h))

    def transport(self, dt, t=0.0):
        self._check(lib().mist_transport(self.h, dt, t))

    def compute_props(self):
        self._check(lib().mist_compute_props(self.h))

    def pressure_rhs(self):
        self._check(lib().mist_pressure_rhs(self.h))

    def corrector(self, dt):
        self._check(lib().mist_corrector(self.h, dt))

    def pressure_operator(self, dt):
        """A!(P) of the semi-Lagrangian scheme into the "divg" field."""
        self._check(lib().mist_pressure_operator(self.h, dt))

    def pressure_jacobian(self, dt):
        """jacob_single: the diagonal of dA/dP (cell extents, indexed [i, j, k])."""
        shape = self.extents("P")[1]
        out = np.zeros(tuple(shape[::-1]))
        self._check(lib().mist_pressure_jacobian(self.h, dt, _dp(out), out.size))
        return out.transpose(2, 1, 0)

    def pressure_solve(self, dt, tol, precision="fp64", check=True):
        cyc = ctypes.c_int64()
        res = ctypes.c_double()
        st = lib().mist_pressure_solve(self.h, dt, tol, {"fp64": 0, "mixed": 1}[precision],
                                       ctypes.byref(cyc), ctypes.byref(res))
        if check:
            self._check(st)
        return cyc.value, res.value

    def vof_init(self):
        self._check(lib().mist_vof_init(self.h))

    def vof_step(self, t, clamp_to=0.0):
        tt = ctypes.c_double(t)
        dt = ctypes.c_double()
        self._check(lib().mist_vof_step(self.h, ctypes.byref(tt), clamp_to, ctypes.byref(dt)))
        return tt.value, dt.value

    def reduce(self, kind, name):
        out = ctypes.c_double()
        self._check(lib().mist_reduce(self.h, REDUCE[kind], FIELD[name], ctypes.byref(out)))
        return out.value

    def plic_dist_batch(self, ijk, nrm, vf):
        ijk = np.ascontiguousarray(ijk, np.int64).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1, 3)
        vf = np.ascontiguousarray(vf, np.float64).reshape(-1)
        out = np.zeros(len(vf))
        self._check(lib().mist_plic_dist_batch(self.h, _ip(ijk), _dp(nrm), _dp(vf), _dp(out), len(vf)))
        return out

    def plic_dist_batch_noalloc(self, ijk, nrm, vf):
        """``plic_dist_batch`` without device allocation (at most 65536 entries)."""
        ijk = np.ascontiguousarray(ijk, np.int64).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1, 3)
        vf = np.ascontiguousarray(vf, np.float64).reshape(-1)
        out = np.zeros(len(vf))
        self._check(lib().mist_plic_dist_batch_noalloc(self.h, _ip(ijk), _dp(nrm), _dp(vf), _dp(out), len(vf)))
        return out

    def plic_vf_batch(self, ijk, nrm, dist):
        ijk = np.ascontiguousarray(ijk, np.int64).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1, 3)
        dist = np.ascontiguousarray(dist, np.float64).reshape(-1)
        out = np.zeros(len(dist))
        self._check(lib().mist_plic_vf_batch(self.h, _ip(ijk), _dp(nrm), _dp(dist), _dp(out), len(dist)))
        return out

    # ---- Navier-Stokes stages (same names and arguments as the oracle's context) -----
    def set_wall_velocity(self, side, vel):
        v = np.ascontiguousarray(vel, np.float64).reshape(3)
        self._check(lib().mist_set_wall_velocity(self.h, SIDES[side], _dp(v)))

    def apply_bc(self, which="u"):
        names = ("u", "v", "w") if which == "u" else ("us", "vs", "ws")
        self._check(lib().mist_apply_bc(self.h, *(FIELD[n] for n in names)))

    def interpolate_face(self, which="us"):
        names = ("u", "v", "w") if which == "u" else ("us", "vs", "ws")
        self._check(lib().mist_interpolate_face(self.h, *(FIELD[n] for n in names)))

    def interpolate_center(self):
        self._check(lib().mist_interpolate_center(self.h))

    def divergence(self, dt):
        self._check(lib().mist_divergence(self.h, dt))

    def compute_curvature(self):
        self._check(lib().mist_compute_curvature(self.h))

    def surface_force(self):
        self._check(lib().mist_surface_force(self.h))

    def ns_init(self):
        self._check(lib().mist_ns_init(self.h))

    def restart_init(self):
        """run_solver's restart initialisation after VF, uf, vf, wf, P were uploaded; returns
        compute_dt's step."""
        dt = ctypes.c_double()
        self._check(lib().mist_restart_init(self.h, ctypes.byref(dt)))
        return dt.value

    def ns_step(self, t, tol, precision="fp64", clamp_to=0.0, check=True):
        tt = ctypes.c_double(t)
        dt = ctypes.c_double()
        cyc = ctypes.c_int64()
        res = ctypes.c_double()
        st = lib().mist_ns_step(self.h, ctypes.byref(tt), clamp_to, ctypes.byref(dt), tol,
                                {"fp64": 0, "mixed": 1}[precision], ctypes.byref(cyc), ctypes.byref(res))
        if check:
            self._check(st)
        return tt.value, dt.value, cyc.value, res.value

    # ---- adjoint of the SL-VOF step (no reference counterpart) -------------------------
    def transport_vjp(self, dt):
        self._check(lib().mist_transport_vjp(self.h, dt))

    def checkpoint_save(self, slot):
        self._check(lib().mist_checkpoint_save(self.h, slot))

    def set_deferred_checks(self, on=True):
        """Transport and VJP calls stop waiting for the device to report a cut-walk overflow;
        the next call that waits (a download, a reduction, dt) reports it, and so does turning
        the deferral off."""
        self._check(lib().mist_set_deferred_checks(self.h, 1 if on else 0))

    def checkpoint_restore(self, slot):
        self._check(lib().mist_checkpoint_restore(self.h, slot))

    PROFILE_KERNELS = ("elvira", "sl_cut", "plic", "project", "mg_rbgs", "mg_resid", "sl_general",
                       "ns_cut", "ns_fields")

    def profile(self, on=True):
        self._check(lib().mist_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        """{kernel: (total_ms, launches, units)} of the kernel profile."""
        out = {}
        for q, name in enumerate(self.PROFILE_KERNELS):
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            u = ctypes.c_int64()
            self._check(lib().mist_profile_read(self.h, q, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(u)))
            out[name] = (ms.value, n.value, u.value)
        return out

    def stage_times(self):
        out = np.zeros(5)
        self._check(lib().mist_stage_times(self.h, _dp(out), 5))
        return dict(zip(("normal", "transport", "plic", "band", "mg"), out))


def run_vof(param: parameters, ic, nsteps=None, ctx_cls=Context, clamp=True, **ctx_kw):
    """``run_solver`` for ``solveNS = false`` (src/NavierStokes_Parallel.jl:36-178) without
    I/O: initial condition, ``Neumann!``/borders, band, normals, PLIC, then the time loop of
    compute_dt, defineVelocity!(t), transport!, computeBand!. ``ic`` is
    ``(shape, *args)`` of ``Context.init_vof``. The last step is clamped to ``tFinal``.
    Returns the context and the list of (t, dt)."""
    ctx = ctx_cls(param, **ctx_kw)
    ctx.init_vof(*ic)
    ctx.vof_init()
    t = 0.0
    hist = []
    nstep = 0
    limit = param.stepMax if nsteps is None else nsteps
    while nstep < limit and t < param.tFinal:
        nstep += 1
        t, dt = ctx.vof_step(t, param.tFinal if clamp else 0.0)
        hist.append((t, dt))
    return ctx, hist
