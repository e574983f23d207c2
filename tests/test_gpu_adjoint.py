"""The adjoint of the SL-VOF transport on the GPU (mist_transport_vjp, checkpointed sweeps):
directional derivatives of one step against central differences of the library's forward
step, and the gradient of the adjoint workload's loss against central differences of the
loss computed by the library and by the oracle (the CPU restatement of the reference's
transport; no Julia / ForwardDiff here, so central differences of the reference algorithm
stand in for a forward-mode derivative of it).

The map is piecewise smooth: cut cases, clamps, the ELVIRA argmin, the band and the cell
searches switch along a path. Differences with h = 1e-6 straddle such a switch in a few
directions; the bars count the directions that agree to 1e-6 relative and bound the rest.
"""
import numpy as np
import pytest

from ns_multiphase_amd import adjoint as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mist():
    from ns_multiphase_amd import solver
    solver.lib()
    return solver


def _setup(mist, N, nsteps, every, seed=1):
    p = A.adjoint_params(N)
    ctx = mist.Context(p)
    # off the grid's symmetry: a circle centred on it has diagonal cells where two ELVIRA
    # candidates give the same normal with different derivatives (a kink of the map)
    ctx.init_vof("circle", 0.2, 0.47, 0.52)
    vf0 = ctx.get("VF")
    basis = A.mode_faces(ctx.mesh)
    amps = np.random.default_rng(seed).normal(0.0, 0.1, 16)
    dt = A.stable_dt(ctx.mesh, basis, np.random.default_rng(2).normal(0.0, 0.1, 16))
    dt = min(dt, A.stable_dt(ctx.mesh, basis, amps))
    run = A.SLVOFRun(ctx, vf0, basis, dt, nsteps, every)
    return p, ctx, run, amps


def _interior(mesh, a):
    o = mesh.nghost + 3
    return a[o:o + mesh.Nx, o:o + mesh.Ny, o:o + mesh.Nz]


def test_one_step_vjp_directional(mist):
    """<lam, dVF'> for a velocity direction and for a VF direction, one step."""
    p, ctx, run, amps = _setup(mist, 64, 1, 1)
    m = ctx.mesh
    rng = np.random.default_rng(4)
    VF1 = run.forward(amps)
    lam = np.zeros_like(VF1)
    _interior(m, lam)[...] = rng.standard_normal((m.Nx, m.Ny, m.Nz))
    g_amp = run.backward(lam)
    lam0 = A.fold_vof_ghosts(m, p, ctx.get("VF_bar"))  # the initial ghost fill, reversed
    # velocity direction
    da = rng.standard_normal(16)
    h = 1e-6
    fp = float((lam * run.forward(amps + h * da, False)).sum())
    fm = float((lam * run.forward(amps - h * da, False)).sum())
    assert float(g_amp @ da) == pytest.approx((fp - fm) / (2 * h), rel=1e-6)
    # VF direction: interface cells away from the walls (their ghosts are not filled here)
    vf0 = run.vf0
    d = np.zeros_like(vf0)
    mid = _interior(m, vf0)
    sel = np.zeros(mid.shape, bool)
    sel[4:-4, 4:-4, :] = (mid[4:-4, 4:-4, :] > 0.02) & (mid[4:-4, 4:-4, :] < 0.98)
    _interior(m, d)[sel] = rng.standard_normal(int(sel.sum()))
    saved = run.vf0
    run.vf0 = saved + h * d
    fp = float((lam * run.forward(amps, False)).sum())
    run.vf0 = saved - h * d
    fm = float((lam * run.forward(amps, False)).sum())
    run.vf0 = saved
    assert float((lam0 * d).sum()) == pytest.approx((fp - fm) / (2 * h), rel=1e-5)


def test_gradient_matches_central_differences(mist):
    """Loss gradient over 24 steps with checkpoints every 8 (64 x 64) against central
    differences of the library's forward loss, per amplitude."""
    p, ctx, run, amps = _setup(mist, 64, 24, 8)
    tgt_amps = np.random.default_rng(2).normal(0.0, 0.1, 16)
    target = run.forward(tgt_amps, False)
    L, g = run.loss_and_grad(amps, target)
    h = 1e-6
    fd = np.array([(run.loss(amps + h * e, target) - run.loss(amps - h * e, target)) / (2 * h)
                   for e in np.eye(16)])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    within = int((rel <= 1e-6).sum())
    print(f"adjoint vs FD: {within}/16 within 1e-6, max rel {rel.max():.2e}")
    assert within >= 12
    assert np.linalg.norm(g - fd) <= 1e-3 * np.linalg.norm(fd)


def test_checkpoint_intervals_give_one_gradient(mist):
    """Every state kept (interval 1), segments of 8, ragged segments of 6 (the last one shorter)
    and the interval a memory budget of 40 slots picks: the same gradient to rounding (the
    reverse kernels add with atomics, so the order of the additions may differ)."""
    p, ctx, run, amps = _setup(mist, 64, 20, 1)
    tgt_amps = np.random.default_rng(2).normal(0.0, 0.1, 16)
    target = run.forward(tgt_amps, False)
    L1, g1 = run.loss_and_grad(amps, target)
    budget = 40 * A.checkpoint_bytes(ctx.mesh)
    k_budget = A.checkpoint_interval(ctx.mesh, 20, budget)
    assert k_budget == 1
    tight = A.checkpoint_interval(ctx.mesh, 20, 8 * A.checkpoint_bytes(ctx.mesh))
    assert tight == 4  # ceil(20 / 4) + 3 = 8 slots; 3 needs 9
    for every in (8, 6, tight):
        r = A.SLVOFRun(ctx, run.vf0, run.basis, run.dt, run.nsteps, every)
        L, g = r.loss_and_grad(amps, target)
        assert L == pytest.approx(L1, rel=1e-13)
        assert np.abs(g - g1).max() <= 1e-10 * np.abs(g1).max(), (every, g, g1)
    # the forward pass leaves the last segment's states in the segment slots and the first
    # reverse sweep skips running that segment again; a second sweep after the same forward
    # finds the first segment's states there and must run it again
    r = A.SLVOFRun(ctx, run.vf0, run.basis, run.dt, run.nsteps, 8)
    VF = r.forward(amps)
    lam = np.zeros_like(VF)
    _interior(ctx.mesh, lam)[...] = _interior(ctx.mesh, VF) - _interior(ctx.mesh, target)
    for _ in range(2):
        g = r.backward(lam)
        assert np.abs(g - g1).max() <= 1e-10 * np.abs(g1).max(), (g, g1)


def test_gradient_matches_oracle_differences(mist, oracle):
    """The same gradient against central differences of the oracle's loss (the reference's
    algorithm on the CPU) at 64 x 64, 16 steps."""
    p, ctx, run, amps = _setup(mist, 64, 16, 8)
    tgt_amps = np.random.default_rng(2).normal(0.0, 0.1, 16)
    target = run.forward(tgt_amps, False)
    L, g = run.loss_and_grad(amps, target)
    o = oracle.OracleContext(p)
    m = ctx.mesh
    oo = m.nghost + 3
    sl = (slice(oo, oo + m.Nx), slice(oo, oo + m.Ny), slice(oo, oo + m.Nz))

    def oracle_loss(a):
        uf, vf = A.faces_of(run.basis, a)
        o.set("uf", uf)
        o.set("vf", vf)
        o.set("wf", np.zeros_like(uf))
        o.set("VF", run.vf0)
        o.vof_init()
        for _ in range(run.nsteps):
            o.transport(run.dt)
            o.compute_band()
        return 0.5 * float(((o.get("VF")[sl] - target[sl]) ** 2).sum())

    assert oracle_loss(amps) == pytest.approx(L, rel=1e-10)
    h = 1e-6
    fd = np.array([(oracle_loss(amps + h * e) - oracle_loss(amps - h * e)) / (2 * h) for e in np.eye(16)])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    print(f"adjoint vs oracle FD: {int((rel <= 1e-6).sum())}/16 within 1e-6, max rel {rel.max():.2e}")
    assert int((rel <= 1e-6).sum()) >= 12
    assert np.linalg.norm(g - fd) <= 1e-3 * np.linalg.norm(fd)


def test_one_step_vjp_stretched_preimages(mist):
    """The reverse step where pre-images stretch over more than four cells (3D, CFL 4): the
    reverse tree passes and the depth-first reverse walk run; a velocity direction against
    central differences of the library's step."""
    from test_oracle_pins import cube_param
    from test_gpu_parity import _sheared_faces
    ctx = mist.Context(cube_param(16, CFL=4.0, solveNS=False))
    m = ctx.mesh
    ctx.init_vof("sphere", 0.3, 0.47, 0.52, 0.5)
    ctx.vof_init()
    vf0 = ctx.get("VF")
    base = _sheared_faces(m)
    pi = np.pi
    X, Y, Z = np.meshgrid(m.x[:-1], m.ym, m.zm, indexing="ij")
    d_u = np.sin(pi * X) ** 2 * np.cos(2 * pi * Y) * np.sin(2 * pi * Z)
    X, Y, Z = np.meshgrid(m.xm, m.y[:-1], m.zm, indexing="ij")
    d_v = np.sin(pi * Y) ** 2 * np.cos(2 * pi * Z) * np.sin(3 * pi * X)
    X, Y, Z = np.meshgrid(m.xm, m.ym, m.z[:-1], indexing="ij")
    d_w = np.sin(pi * Z) ** 2 * np.sin(2 * pi * X) * np.cos(3 * pi * Y)
    dirs = (d_u, d_v, d_w)
    dt = 0.15

    def start(h):
        ctx.set("VF", vf0)
        ctx.vof_init()
        for name, b, d in zip(("uf", "vf", "wf"), base, dirs):
            ctx.set(name, b + h * d)

    def step(h):
        start(h)
        ctx.transport(dt)
        ctx.compute_band()
        return ctx.get("VF")

    VF1 = step(0.0)
    lam = np.zeros_like(VF1)
    lam[m.nghost + 3:m.nghost + 3 + m.Nx, m.nghost + 3:m.nghost + 3 + m.Ny, m.nghost + 3:m.nghost + 3 + m.Nz] = \
        np.random.default_rng(11).standard_normal((m.Nx, m.Ny, m.Nz))
    start(0.0)
    ctx.set("VF_bar", lam)
    for name in ("uf_bar", "vf_bar", "wf_bar"):
        ctx.set(name, np.zeros(ctx.get(name).shape))
    ctx.transport_vjp(dt)
    g = sum(float((ctx.get(n) * d).sum()) for n, d in zip(("uf_bar", "vf_bar", "wf_bar"), dirs))
    h = 1e-7
    fd = (float((lam * step(h)).sum()) - float((lam * step(-h)).sum())) / (2 * h)
    print(f"stretched one-step VJP {g:.10e} vs central difference {fd:.10e}")
    assert g == pytest.approx(fd, rel=1e-5)


def test_vjp_after_upload_uses_the_new_band(mist):
    """A VF uploaded after a step's band was built makes that band stale: the reverse step then
    recomputes the band (and normals, PLIC constants) of the new VF, and equals the reverse
    step of a context initialised on that VF."""
    p = A.adjoint_params(48)
    a, b = mist.Context(p), mist.Context(p)
    a.init_vof("circle", 0.2, 0.47, 0.52)
    a.vof_init()
    b.init_vof("circle", 0.15, 0.55, 0.4)
    b.vof_init()
    vf2 = b.get("VF")  # with its ghost fills
    rng = np.random.default_rng(7)
    faces = {n: 0.3 * rng.standard_normal(a.get(n).shape) for n in ("uf", "vf", "wf")}
    lam = np.zeros_like(vf2)
    _interior(a.mesh, lam)[...] = rng.standard_normal((a.mesh.Nx, a.mesh.Ny, a.mesh.Nz))
    a.set("VF", vf2)  # the band list of the first circle is stale from here
    for c in (a, b):
        for n, f in faces.items():
            c.set(n, f)
        c.set("VF_bar", lam)
        for n in ("uf_bar", "vf_bar", "wf_bar"):
            c.set(n, np.zeros(c.get(n).shape))
        c.transport_vjp(1e-3)
    # adjoints shared by several cells are added with atomics: equal up to their order
    for n in ("VF_bar", "uf_bar", "vf_bar", "wf_bar"):
        x, y = a.get(n), b.get(n)
        assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max()), n


def test_deferred_checks_change_no_result(mist):
    """mist_set_deferred_checks: transport and VJP calls that skip their per-call wait give the
    same fields as calls that wait (the forward bit for bit; the adjoints to the rounding of their
    atomic sums, whose order of arrival varies from run to run), and turning the deferral off
    reports no overflow when there was none."""
    p, ctx, run, amps = _setup(mist, 64, 6, 3)
    run.set_velocity(amps)
    out = {}
    for defer in (False, True):
        ctx.set("VF", run.vf0)
        ctx.vof_init()
        ctx.set_deferred_checks(defer)
        ctx.checkpoint_save(0)
        for _ in range(6):
            run._step()
        vf = ctx.get("VF")
        lam = np.zeros_like(vf)
        _interior(ctx.mesh, lam)[...] = 1.0
        ctx.set("VF_bar", lam)
        for name in ("uf_bar", "vf_bar", "wf_bar"):
            ctx.set(name, np.zeros(tuple(int(v) for v in ctx.extents(name)[1])))
        ctx.checkpoint_restore(0)
        ctx.transport_vjp(run.dt)
        ctx.set_deferred_checks(False)
        out[defer] = (vf, ctx.get("VF_bar"), ctx.get("uf_bar"), ctx.get("vf_bar"))
    assert np.array_equal(out[False][0], out[True][0])
    for a, b in zip(out[False][1:], out[True][1:]):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_gradient_matches_forward_mode(mist, oracle):
    """The device's reverse-mode gradient (64 x 64, 24 steps, checkpoints every 8) against the
    oracle's forward mode (dual numbers through the same 24 steps, oracle/oracle_fwd.cpp): both
    are exact derivatives of the map with its branches held, so no step size enters. All 16
    directions agree to 1e-6."""
    p, ctx, run, amps = _setup(mist, 64, 24, 8)
    tgt_amps = np.random.default_rng(2).normal(0.0, 0.1, 16)
    target = run.forward(tgt_amps, False)
    L, g = run.loss_and_grad(amps, target)
    o = oracle.OracleContext(p)
    o.set("VF", run.vf0)
    uf = (amps @ run._Bu).reshape(run._shape)  # the faces the device run set
    vf = (amps @ run._Bv).reshape(run._shape)
    Bu = np.stack([b[0] for b in run.basis])
    Bv = np.stack([b[1] for b in run.basis])
    VF, dVF = o.fwd_vof_run(uf, vf, np.zeros_like(uf), Bu, Bv, run.nsteps, run.dt)
    m = ctx.mesh
    oo = m.nghost + 3
    sl = (slice(oo, oo + m.Nx), slice(oo, oo + m.Ny), slice(oo, oo + m.Nz))
    diff = VF[sl] - target[sl]
    assert 0.5 * float((diff ** 2).sum()) == pytest.approx(L, rel=1e-10)
    gf = np.array([float((diff * dVF[q][sl]).sum()) for q in range(16)])
    rel = np.abs(g - gf) / np.maximum(np.abs(gf), 1e-3 * np.abs(gf).max())
    print(f"adjoint vs forward mode: {int((rel <= 1e-6).sum())}/16 within 1e-6, max rel {rel.max():.2e}")
    assert int((rel <= 1e-6).sum()) == 16
