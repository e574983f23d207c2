"""Gradients through the SL-VOF transport: a checkpointed reverse sweep over the library's
`mist_transport_vjp` (the adjoint of one transport! + computeBand! step), and the velocity
parametrisation of the adjoint workload.

No reference counterpart: Mist.jl has no derivative of its transport. The workload follows
the north-star description: a 2D box, face velocities from a stream function of 16 Fourier
modes psi = sum_{kx,ky=1..4} a_k sin(kx pi x) sin(ky pi y) taken as corner differences (so the
discrete flux divergence is exactly zero and the wall fluxes vanish), a circle transported
for a number of steps, the loss 1/2 sum (VF - VF_target)^2 over interior cells, and its
gradient with respect to the 16 amplitudes.
"""
from __future__ import annotations

import math

import numpy as np

from .params import parameters

MODES = [(kx, ky) for kx in range(1, 5) for ky in range(1, 5)]


def adjoint_params(N: int = 256, **kw) -> parameters:
    """Unit square, N x N x 1 (periodic z, cubic cells), walls in x and y, SL transport."""
    base = dict(mu_liq=1.0, mu_gas=1.0, rho_liq=1.0, rho_gas=1.0, sigma=0.0, Lx=1.0, Ly=1.0,
                Lz=1.0 / N, tFinal=1.0, Nx=N, Ny=N, Nz=1, stepMax=100000, CFL=0.5, out_period=1,
                tol=1e-3, nprocx=1, nprocy=1, nprocz=1, xper=False, yper=False, zper=True,
                solveNS=False, pressure_scheme="semi-lagrangian", projection_method="RK4",
                iter_type="standard")
    base.update(kw)
    return parameters(**base)


def mode_faces(mesh):
    """Face velocities of each unit-amplitude mode on the cell-extent arrays [i, j, k]:
    uf(x_i, ym_j) = (psi(x_i, y_j+1) - psi(x_i, y_j)) / dy, vf(xm_i, y_j) = -(psi(x_i+1, y_j)
    - psi(x_i, y_j)) / dx, wf = 0."""
    x, y = mesh.x, mesh.y  # faces over the cell extents (+1)
    nz = len(mesh.zm)
    out = []
    for kx, ky in MODES:
        psi = np.sin(kx * math.pi * x)[:, None] * np.sin(ky * math.pi * y)[None, :]  # corners
        uf = (psi[:-1, 1:] - psi[:-1, :-1]) / mesh.dy
        vf = -(psi[1:, :-1] - psi[:-1, :-1]) / mesh.dx
        out.append((np.repeat(uf[:, :, None], nz, 2), np.repeat(vf[:, :, None], nz, 2)))
    return out


def faces_of(basis, amps):
    uf = sum(a * b[0] for a, b in zip(amps, basis))
    vf = sum(a * b[1] for a, b in zip(amps, basis))
    return uf, vf


def stable_dt(mesh, basis, amps, cfl=0.5):
    uf, vf = faces_of(basis, amps)
    umax = max(np.abs(uf).max(), np.abs(vf).max(), 1e-12)
    return cfl * min(mesh.dx, mesh.dy) / umax


def fold_vof_ghosts(mesh, param, lam):
    """Reverse of the ghost fills of VF (Neumann!, then update_borders! on periodic axes, as
    mist_vof_init and the end of each step apply them): the adjoint of a whole VF array
    mapped to the adjoint of the values the fills copied from. Returns a new array."""
    g = np.array(lam, dtype=np.float64, copy=True)
    ng = mesh.nghost
    N = (mesh.Nx, mesh.Ny, mesh.Nz)
    per = (param.xper, param.yper, param.zper)

    def sl(d, i):  # slab of reference index i along axis d
        idx = [slice(None)] * 3
        idx[d] = i + ng + 2
        return tuple(idx)

    for d in (2, 1, 0):
        if not per[d]:
            continue
        n = N[d]
        for s in range(ng - 1, -1, -1):
            g[sl(d, n - ng + 1 + s)] += g[sl(d, 1 - ng + s)]
            g[sl(d, 1 - ng + s)] = 0.0
        for s in range(ng):
            g[sl(d, 1 + s)] += g[sl(d, n + 1 + s)]
            g[sl(d, n + 1 + s)] = 0.0
    for d in (2, 1, 0):
        n = N[d]
        a, b = g[sl(d, n + 1)].copy(), g[sl(d, 0)].copy()
        g[sl(d, n + 1)] = 0.0
        g[sl(d, 0)] = 0.0
        g[sl(d, n)] += a
        g[sl(d, 1)] += b
    return g


MAX_SLOTS = 4096  # checkpoint slots one library context holds (mist_checkpoint_save)


def checkpoint_bytes(mesh) -> int:
    """Device bytes of one checkpoint slot (mist_checkpoint_save): VF with its ghosts (FP64),
    the band flags (int16), the normals and PLIC constants (4 x FP64), ELVIRA's choice and the
    band-cell list (2 x int32 per interior cell), each part 256-byte aligned."""
    def al(n):
        return (n + 255) // 256 * 256
    ncell_ext = int(np.prod(mesh.cell_shape()))
    ncell = mesh.Nx * mesh.Ny * mesh.Nz
    return (al(8 * int(np.prod(mesh.vof_shape()))) + al(2 * ncell_ext) + 4 * al(8 * ncell_ext)
            + 2 * al(4 * ncell) + al(4))


def checkpoint_interval(mesh, nsteps: int, mem_budget: int) -> int:
    """The smallest interval k whose slots fit in `mem_budget` bytes of device memory: ceil(nsteps
    / k) checkpoints plus k - 1 states of the segment being walked back. k = 1 keeps every state,
    and the reverse sweep then runs no transport step again."""
    per = checkpoint_bytes(mesh)
    for k in range(1, max(1, nsteps) + 1):
        slots = -(-nsteps // k) + k - 1
        if slots <= MAX_SLOTS and slots * per <= mem_budget:
            return k
    raise ValueError(f"{nsteps} steps do not fit a checkpoint budget of {mem_budget} bytes "
                     f"({per} bytes per slot)")


class SLVOFRun:
    """Forward runs and checkpointed reverse sweeps of `nsteps` transport steps of fixed dt
    under steady face velocities, on one library context. The state every `every` steps is kept
    on the device; a reverse sweep recomputes each segment forward from its checkpoint (storing
    its other states) and walks it back with `mist_transport_vjp`. With `every=None` the
    interval is the smallest whose slots fit `mem_budget` bytes (checkpoint_interval)."""

    def __init__(self, ctx, vf0, basis, dt, nsteps, every=None, mem_budget=8 << 30):
        self.ctx, self.vf0, self.basis = ctx, vf0, basis
        self.dt, self.nsteps = float(dt), int(nsteps)
        if every is None:
            every = checkpoint_interval(ctx.mesh, self.nsteps, int(mem_budget))
        self.every = int(every)
        if self.every < 1:
            raise ValueError("every must be >= 1")
        self.nseg = (self.nsteps + self.every - 1) // self.every
        if self.nseg + self.every - 1 > MAX_SLOTS:
            raise ValueError(f"{self.nseg + self.every - 1} checkpoint slots needed, at most {MAX_SLOTS}")
        # the modes as rows of two matrices, so that the faces of a set of amplitudes and the
        # projection of the face adjoints onto the modes are one matrix-vector product each
        self._shape = basis[0][0].shape
        self._Bu = np.stack([b[0].ravel() for b in basis])
        self._Bv = np.stack([b[1].ravel() for b in basis])
        self._zero = {name: np.zeros(tuple(int(v) for v in ctx.extents(name)[1]))
                      for name in ("wf", "uf_bar", "vf_bar", "wf_bar")}

    def set_velocity(self, amps):
        a = np.asarray(amps, np.float64)
        self.ctx.set("uf", (a @ self._Bu).reshape(self._shape))
        self.ctx.set("vf", (a @ self._Bv).reshape(self._shape))
        self.ctx.set("wf", self._zero["wf"])

    def _step(self):
        self.ctx.transport(self.dt)
        self.ctx.compute_band()

    def forward(self, amps, checkpoints=True):
        """VF after the run. With checkpoints, the state every `every` steps goes to its slot,
        and the states inside the last segment go to the segment slots the reverse sweep walks
        first, so that segment is not run forward again."""
        c = self.ctx
        self.set_velocity(amps)
        c.set("VF", self.vf0)
        c.vof_init()
        self._tail_kept = False
        last0 = (self.nseg - 1) * self.every
        c.set_deferred_checks(True)  # one wait for the whole run (the download below)
        try:
            for s in range(self.nsteps):
                if checkpoints:
                    if s % self.every == 0:
                        c.checkpoint_save(s // self.every)
                    elif s > last0:
                        c.checkpoint_save(self.nseg + (s - last0 - 1))
                self._step()
        finally:
            c.set_deferred_checks(False)
        self._tail_kept = bool(checkpoints)
        return c.get("VF")

    def backward(self, lam_T):
        """dL/d(amps) given dL/dVF at the end of the last forward (VF-extent array)."""
        c = self.ctx
        c.set("VF_bar", lam_T)
        for name in ("uf_bar", "vf_bar", "wf_bar"):
            c.set(name, self._zero[name])
        base = self.nseg  # slots after the checkpoints hold a segment's later states
        tail = getattr(self, "_tail_kept", False)
        self._tail_kept = False  # the sweep reuses the segment slots
        c.set_deferred_checks(True)  # one wait for the whole sweep (the downloads below)
        try:
            self._sweep(base, tail)
        finally:
            c.set_deferred_checks(False)
        guf, gvf = c.get("uf_bar"), c.get("vf_bar")
        return self._Bu @ guf.ravel() + self._Bv @ gvf.ravel()

    def _sweep(self, base, tail_kept=False):
        """Segments from the last to the first: each is run forward from its checkpoint into
        the segment slots (unless the forward pass left the last one there), then walked back."""
        c = self.ctx
        for seg in range(self.nseg - 1, -1, -1):
            s0 = seg * self.every
            s1 = min(self.nsteps, s0 + self.every)

            def slot(s):
                return seg if s == s0 else base + (s - s0 - 1)
            if s1 - s0 > 1 and not (tail_kept and seg == self.nseg - 1):
                c.checkpoint_restore(seg)  # VF, band, normals and PLIC constants of state s0
                for s in range(s0 + 1, s1):
                    self._step()
                    c.checkpoint_save(slot(s))
            for s in range(s1 - 1, s0 - 1, -1):
                c.checkpoint_restore(slot(s))
                c.transport_vjp(self.dt)

    def loss_and_grad(self, amps, target):
        """L = 1/2 sum over interior cells (VF - target)^2 and dL/d(amps)."""
        m = self.ctx.mesh
        o = m.nghost + 3
        VF = self.forward(amps)
        diff = np.zeros_like(VF)
        sl = (slice(o, o + m.Nx), slice(o, o + m.Ny), slice(o, o + m.Nz))
        diff[sl] = VF[sl] - target[sl]
        L = 0.5 * float((diff[sl] ** 2).sum())
        return L, self.backward(diff)

    def loss(self, amps, target):
        m = self.ctx.mesh
        o = m.nghost + 3
        VF = self.forward(amps, checkpoints=False)
        sl = (slice(o, o + m.Nx), slice(o, o + m.Ny), slice(o, o + m.Nz))
        return 0.5 * float(((VF[sl] - target[sl]) ** 2).sum())


def torch_function():
    """A torch.autograd.Function over a run: forward(amps) -> VF at the end, backward through
    the library's adjoint. Built on first use so that the module imports without torch."""
    import torch

    class SLVOFTransport(torch.autograd.Function):
        @staticmethod
        def forward(ctx_, amps, run):
            ctx_.run = run
            vf = run.forward(amps.detach().cpu().numpy())
            return torch.from_numpy(vf)

        @staticmethod
        def backward(ctx_, grad_vf):
            g = ctx_.run.backward(grad_vf.detach().cpu().numpy())
            return torch.from_numpy(g), None

    return SLVOFTransport
