"""Pins of the oracle's field builders and classifiers (CPU): the prescribed velocities of
defineVelocity!, the initial-condition builders VFcircle / VFsphere / VFbubble3d, compute_dt,
the axis normals ELVIRA! gives empty and full cells, and computeBand!.

Each is checked against something other than its own formula: closed forms evaluated
independently (numpy, mpmath quadrature), exact properties (divergence, total area),
brute-force definitions (sample counts, Chebyshev distance by sliding windows) and hand-built
cases. A plausible mistake (a sample lattice off by one, a face/centre swap, a 6- instead of
27-neighbourhood, a sign or priority swap) fails one of them.
"""
import math

import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view

from test_oracle_pins import cube_param


# ---- defineVelocity! (src/Tools.jl:648-693) --------------------------------------------------
def _closed_forms(kind, T):
    s, c, pi = np.sin, np.cos, np.pi
    if kind == "Deformation":
        return (lambda x, y, z, t: -2 * s(pi * x) ** 2 * s(pi * y) * c(pi * y) * c(pi * t / T),
                lambda x, y, z, t: 2 * s(pi * y) ** 2 * s(pi * x) * c(pi * x) * c(pi * t / T),
                lambda x, y, z, t: 0 * x)
    if kind == "Deformation3D":
        return (lambda x, y, z, t: 2 * s(pi * x) ** 2 * s(2 * pi * y) * s(2 * pi * z) * c(pi * t / T),
                lambda x, y, z, t: -s(pi * y) ** 2 * s(2 * pi * x) * s(2 * pi * z) * c(pi * t / T),
                lambda x, y, z, t: -s(pi * z) ** 2 * s(2 * pi * x) * s(2 * pi * y) * c(pi * t / T))
    if kind == "rotation":
        return (lambda x, y, z, t: 2 * pi * (0.5 - y), lambda x, y, z, t: 2 * pi * (x - 0.5), lambda x, y, z, t: 0 * x)
    if kind == "divFlow":
        return (lambda x, y, z, t: 2 * x, lambda x, y, z, t: -(2.5 - y), lambda x, y, z, t: 0 * x)
    if kind == "divFlow3D":
        return (lambda x, y, z, t: 200 * x, lambda x, y, z, t: -100 * (5 - y), lambda x, y, z, t: -50 * (5 - z))
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["Deformation", "Deformation3D", "rotation", "divFlow", "divFlow3D"])
def test_define_velocity_closed_forms_centres_and_faces(oracle, kind):
    """Cell velocities at the centres (xm, ym, zm) and face velocities at their own face
    (uf at (x, ym, zm), vf at (xm, y, zm), wf at (xm, ym, z)) over every cell incl. ghosts,
    with the period as a parameter (the reference hard-codes T = 2)."""
    T, t = 3.0, 0.7
    p = cube_param(10, Nz=6, Lz=0.6, VFVelocity=kind, vel_period=T)
    o = oracle.OracleContext(p)
    o.define_velocity(t)
    m = o.mesh
    fu, fv, fw = _closed_forms(kind, T)
    xf, yf, zf = m.x[:-1], m.y[:-1], m.z[:-1]
    C = np.meshgrid(m.xm, m.ym, m.zm, indexing="ij")
    want = {
        "u": fu(*C, t), "v": fv(*C, t), "w": fw(*C, t),
        "uf": fu(*np.meshgrid(xf, m.ym, m.zm, indexing="ij"), t),
        "vf": fv(*np.meshgrid(m.xm, yf, m.zm, indexing="ij"), t),
        "wf": fw(*np.meshgrid(m.xm, m.ym, zf, indexing="ij"), t),
    }
    for name, w in want.items():
        got = o.get(name)
        assert got.shape == w.shape
        assert np.abs(got - w).max() <= 1e-14 * max(1.0, np.abs(w).max()), name


@pytest.mark.parametrize("kind,div", [("rotation", 0.0), ("divFlow", 3.0), ("divFlow3D", 350.0)])
def test_define_velocity_discrete_divergence_exact(oracle, kind, div):
    """Fields linear in their own coordinate have an exact finite-volume divergence:
    rotation 0, divFlow 2 + 1, divFlow3D 200 + 100 + 50."""
    o = oracle.OracleContext(cube_param(8, VFVelocity=kind))
    o.define_velocity(0.2)
    m = o.mesh
    g = m.nghost
    uf, vf, wf = o.get("uf"), o.get("vf"), o.get("wf")
    sl = slice(g, g + 8)
    d = ((uf[g + 1:g + 9, sl, sl] - uf[sl, sl, sl]) / m.dx + (vf[sl, g + 1:g + 9, sl] - vf[sl, sl, sl]) / m.dy
         + (wf[sl, sl, g + 1:g + 9] - wf[sl, sl, sl]) / m.dz)
    assert np.abs(d - div).max() <= 1e-11 * max(1.0, div)


@pytest.mark.parametrize("kind,N3", [("Deformation3D", True), ("Deformation", False)])
def test_define_velocity_deformation_discretely_solenoidal(oracle, kind, N3):
    """The deformation fields have a zero finite-volume divergence on a uniform grid, to
    rounding: with faces at x_i and centres at xm_i, sin^2(pi x_{i+1}) - sin^2(pi x_i) =
    sin(2 pi xm_i) sin(pi h), so the face differences of u, v (and w) cancel exactly in
    exact arithmetic. Face values taken at centres, a wrong factor or a swapped component
    leave an O(1) divergence."""
    N = 16
    p = cube_param(N, VFVelocity=kind, vel_period=3.0) if N3 else \
        cube_param(N, Nz=1, Lz=1.0 / N, zper=True, VFVelocity=kind, vel_period=3.0)
    o = oracle.OracleContext(p)
    o.define_velocity(0.4)
    m = o.mesh
    g = m.nghost
    uf, vf, wf = o.get("uf"), o.get("vf"), o.get("wf")
    sl = slice(g, g + N)
    up = slice(g + 1, g + N + 1)
    kz = sl if N3 else slice(g, g + 1)
    kzu = up if N3 else slice(g + 1, g + 2)
    d = ((uf[up, sl, kz] - uf[sl, sl, kz]) / m.dx + (vf[sl, up, kz] - vf[sl, sl, kz]) / m.dy
         + (wf[sl, sl, kzu] - wf[sl, sl, kz]) / m.dz)
    scale = np.abs(uf).max() / m.dx
    assert np.abs(d).max() <= 1e-13 * scale
    # the same differences of the cell-centre values (a face/centre mix-up) do not cancel
    u = o.get("u")
    assert np.abs((u[up, sl, kz] - u[sl, sl, kz]) / m.dx).max() > 1e-2 * scale


def test_define_velocity_airblast_shear_profile(oracle):
    o = oracle.OracleContext(cube_param(8, VFVelocity="airblast"))
    o.define_velocity(0.0)
    m = o.mesh
    want = 50 * (1 - np.exp(-100 * np.abs(m.ym - 0.5)))
    u = o.get("u")
    assert np.abs(u - want[None, :, None]).max() <= 1e-13
    assert np.abs(o.get("v")).max() == 0.0 and np.abs(o.get("wf")).max() == 0.0


# ---- VFcircle (src/Tools.jl:790-902) ---------------------------------------------------------
def _disk_rect_area(x0, x1, y0, y1, r, xo, yo):
    """Area of the disk (xo, yo, r) inside [x0, x1] x [y0, y1], by mpmath quadrature of the
    chord length, split where the chord meets the rectangle's horizontal sides."""
    import mpmath as mp
    mp.mp.dps = 30
    a0, a1 = max(x0 - xo, -r), min(x1 - xo, r)
    if a1 <= a0:
        return 0.0
    b0, b1 = y0 - yo, y1 - yo

    def chord(x):
        s = mp.sqrt(max(r * r - x * x, 0))
        return max(0, min(b1, s) - max(b0, -s))

    pts = [a0, a1]
    for b in (b0, b1):
        if abs(b) < r:
            xb = math.sqrt(r * r - b * b)
            pts += [q for q in (-xb, xb) if a0 < q < a1]
    pts = sorted(set(pts))
    return float(mp.fsum(mp.quad(chord, [pts[q], pts[q + 1]]) for q in range(len(pts) - 1)))


def test_vf_circle_exact_areas(oracle):
    """Every cell of a 16 x 16 mesh: VFcircle against the quadrature area of the disk inside
    the cell, and the total liquid area against pi r^2."""
    N = 16
    p = cube_param(N, Nz=1, Lz=1.0 / N, zper=True)
    o = oracle.OracleContext(p)
    r, xo, yo = 0.23, 0.437, 0.551
    o.init_vof("circle", r, xo, yo)
    m = o.mesh
    g = m.nghost
    VF = o.get("VF")[3 + g:3 + g + N, 3 + g:3 + g + N, 3 + g]
    worst = 0.0
    for i in range(N):
        for j in range(N):
            x0, x1 = m.x[g + i], m.x[g + i + 1]
            y0, y1 = m.y[g + j], m.y[g + j + 1]
            exact = _disk_rect_area(x0, x1, y0, y1, r, xo, yo) / ((x1 - x0) * (y1 - y0))
            worst = max(worst, abs(VF[i, j] - exact))
    assert worst <= 1e-12, worst
    assert abs(VF.sum() * m.dx * m.dy - math.pi * r * r) <= 1e-12


# ---- VFsphere / VFbubble3d (src/Tools.jl:907-922, 1061-1076) ---------------------------------
def _sample_counts(m, cells, r, c):
    """Per cell, the number of the 20^3 samples x_min + i/21 (x_max - x_min), i = 1..20, that
    lie strictly inside the sphere, evaluated in extended precision."""
    ld = np.longdouble
    frac = np.arange(1, 21, dtype=ld) / ld(21)
    out = {}
    for (i, j, k) in cells:
        xs = ld(m.x[i]) + frac * (ld(m.x[i + 1]) - ld(m.x[i]))
        ys = ld(m.y[j]) + frac * (ld(m.y[j + 1]) - ld(m.y[j]))
        zs = ld(m.z[k]) + frac * (ld(m.z[k + 1]) - ld(m.z[k]))
        d2 = ((xs - ld(c[0])) ** 2)[:, None, None] + ((ys - ld(c[1])) ** 2)[None, :, None] + ((zs - ld(c[2])) ** 2)[None, None, :]
        out[(i, j, k)] = int((ld(r) * ld(r) - d2 > 0).sum())
    return out


def _running_sums(start, step):
    vals = [start]
    s = start
    for _ in range(8000):
        s = s + step
        vals.append(s)
    return vals


@pytest.mark.parametrize("shape", ["sphere", "bubble3d"])
def test_vf_sphere_and_bubble_sample_counts(oracle, shape):
    """VF of every cell (ghosts included) equals the reference's running sum of 1/20^3 over
    the samples inside the sphere (VFsphere adds up from 0, VFbubble3d subtracts from 1),
    with the samples counted by brute force."""
    N = 10
    p = cube_param(N)
    o = oracle.OracleContext(p)
    r, c = 0.27, (0.463, 0.518, 0.442)
    o.init_vof(shape, r, *c)
    m = o.mesh
    g = m.nghost
    VF = o.get("VF")
    cells = [(i, j, k) for i in range(N + 2 * g) for j in range(N + 2 * g) for k in range(N + 2 * g)]
    counts = _sample_counts(m, cells, r, c)
    table = _running_sums(0.0, 1.0 / 8000) if shape == "sphere" else _running_sums(1.0, -1.0 / 8000)
    cut = 0
    for (i, j, k), n in counts.items():
        assert VF[i + 3, j + 3, k + 3] == table[n], ((i, j, k), n)
        cut += 0 < n < 8000
    assert cut > 50  # the interface crosses many cells


# ---- compute_dt (src/Tools.jl:317-337) ---------------------------------------------------------
@pytest.mark.parametrize("U,mu,max_dt,expect", [
    ((2.0, 0.5, 0.25), (1e-3, 1e-3), 1.0, "x"),     # convective x binds
    ((0.1, 4.0, 0.2), (1e-3, 1e-3), 1.0, "y"),      # convective y binds
    ((0.0, 0.0, -8.0), (1e-3, 1e-3), 1.0, "z"),     # zero components give dx/0 = Inf, |w| used
    ((0.1, 0.1, 0.1), (3.0, 7.0), 1.0, "visc"),     # the larger viscosity binds
    ((0.1, 0.1, 0.1), (1e-3, 1e-3), 1e-4, "max"),   # max_dt binds (not scaled by CFL)
])
def test_compute_dt_uniform_fields(oracle, U, mu, max_dt, expect):
    p = cube_param(8, Ny=4, Ly=0.5, Nz=16, Lz=0.5, CFL=0.7, mu_liq=mu[0], mu_gas=mu[1], max_dt=max_dt)
    o = oracle.OracleContext(p)
    m = o.mesh
    for name, val in zip(("u", "v", "w"), U):
        o.set(name, np.full(m.cell_shape(), val))
    dx, dy, dz = m.dx, m.dy, m.dz
    cands = {"x": dx / abs(U[0]) if U[0] else math.inf, "y": dy / abs(U[1]) if U[1] else math.inf,
             "z": dz / abs(U[2]) if U[2] else math.inf, "visc": min(dx, dy, dz) / max(mu)}
    want = max_dt if expect == "max" else 0.7 * cands[expect]
    assert expect == "max" or cands[expect] == min(cands.values())
    assert o.compute_dt() == pytest.approx(want, rel=1e-15)


def test_compute_dt_reads_ghost_cells(oracle):
    """The maxima run over the whole arrays, ghosts included (maximum(abs.(u)))."""
    p = cube_param(8, CFL=0.5, mu_liq=1e-3, mu_gas=1e-3)
    o = oracle.OracleContext(p)
    u = np.full(o.mesh.cell_shape(), 0.5)
    o.set("u", u)
    o.set("v", u * 0)
    o.set("w", u * 0)
    dt0 = o.compute_dt()
    u[0, 0, 0] = -5.0
    o.set("u", u)
    assert dt0 == pytest.approx(0.5 * o.mesh.dx / 0.5) and o.compute_dt() == pytest.approx(0.5 * o.mesh.dx / 5.0)


# ---- ELVIRA! axis normals of empty and full cells (src/ELVIRA.jl:19-46) ----------------------
@pytest.mark.parametrize("centre,full,neigh,want", [
    (0.0, 1.0, [(-1, 0, 0)], (1, 0, 0)),                # empty cell, liquid at i-1
    (0.0, 1.0, [(1, 0, 0)], (-1, 0, 0)),
    (0.0, 1.0, [(0, -1, 0), (0, 0, 1)], (0, 1, 0)),     # y before z
    (0.0, 1.0, [(1, 0, 0), (0, -1, 0)], (-1, 0, 0)),    # x before y
    (0.0, 1.0, [(0, 0, 1)], (0, 0, -1)),
    (0.0, 1.0, [(1, 1, 0), (-1, 0, 1)], (0, 0, 0)),     # only diagonal neighbours: none
    (1.0, 0.0, [(-1, 0, 0)], (-1, 0, 0)),               # full cell, gas at i-1
    (1.0, 0.0, [(0, 1, 0), (0, 0, -1)], (0, 1, 0)),
    (1.0, 0.0, [(0, 0, 1)], (0, 0, 1)),
    (1.0, 0.0, [(0, 0, -1)], (0, 0, -1)),
])
def test_elvira_axis_normals(oracle, centre, full, neigh, want):
    p = cube_param(8)
    o = oracle.OracleContext(p)
    VF = np.full(o.mesh.vof_shape(), centre)
    g = o.mesh.nghost + 3
    i = j = k = 4
    for (a, b, c) in neigh:
        VF[g + i - 1 + a, g + j - 1 + b, g + k - 1 + c] = full
    o.set("VF", VF)
    o.compute_normal()
    gg = o.mesh.nghost
    got = tuple(float(o.get(n)[gg + i - 1, gg + j - 1, gg + k - 1]) for n in ("nx", "ny", "nz"))
    assert got == tuple(float(w) for w in want)


# ---- computeBand! (src/VF.jl:16-74) -------------------------------------------------------------
def _band_bruteforce(VF, N, ng, nband, lo, hi):
    """band by its definition: 0 on interface cells and on empty (full) cells with a full
    (empty) cell anywhere in their 3x3x3 block, +-(nband+1) elsewhere by phase; then level n
    on cells whose 3x3x3 block holds level n-1, the first ghost layer holding the Neumann copy
    of the first classification (non-periodic borders leave it)."""
    o = ng + 3
    V = VF[o - 1:o + N + 1, o - 1:o + N + 1, o - 1:o + N + 1]
    W = sliding_window_view(V, (3, 3, 3))
    c = V[1:-1, 1:-1, 1:-1]
    has_full = (W >= hi).any(axis=(3, 4, 5))
    has_empty = (W <= lo).any(axis=(3, 4, 5))
    b = np.zeros(c.shape, np.int64)
    gas, liq = c < lo, c > hi
    b[gas] = np.where(has_full[gas], 0, -(nband + 1))
    b[liq] = np.where(has_empty[liq], 0, nband + 1)
    cur = np.pad(b, 1, mode="edge")
    for n in range(1, nband + 1):
        hit = (sliding_window_view(np.abs(cur), (3, 3, 3)) == n - 1).any(axis=(3, 4, 5))
        inner = cur[1:-1, 1:-1, 1:-1]
        upd = (np.abs(inner) > n) & hit
        inner[upd] = n * np.sign(inner[upd])
    return cur[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("seed,nband", [(1, 1), (2, 2), (3, 3)])
def test_compute_band_bruteforce(oracle, seed, nband):
    """Random blobby fields (isolated diagonal contacts included, walls reached): the oracle's
    band equals the brute-force definition bit for bit. A 6-neighbour rule, a >= / > swap at
    VFhi or swept ghosts all change it."""
    N = 14
    p = cube_param(N, CFL=float(nband))
    o = oracle.OracleContext(p)
    rng = np.random.default_rng(seed)
    shape = o.mesh.vof_shape()
    # liquid boxes in gas, isolated specks of each phase (diagonal-only contacts), a few
    # fractional cells and cells exactly at the thresholds
    VF = np.zeros(shape)
    for _ in range(5):
        lo = rng.integers(0, shape[0] - 6, 3)
        hi = lo + rng.integers(3, 8, 3)
        VF[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1.0
    VF[rng.random(shape) < 0.003] = 1.0
    VF[rng.random(shape) < 0.003] = 0.0
    frac = rng.random(shape) < 0.004
    VF[frac] = rng.random(int(frac.sum()))
    VF[rng.random(shape) < 0.002] = p.VFhi
    VF[rng.random(shape) < 0.002] = p.VFlo
    o.set("VF", VF)
    o.compute_band()
    ng = o.mesh.nghost
    got = o.get("band")[ng:ng + N, ng:ng + N, ng:ng + N]
    want = _band_bruteforce(VF, N, ng, nband, p.VFlo, p.VFhi)
    assert (want == 0).sum() > 50 and (np.abs(want) == nband + 1).sum() > 5
    assert all((np.abs(want) == n).sum() > 0 for n in range(1, nband + 1))
    assert np.array_equal(got, want)
