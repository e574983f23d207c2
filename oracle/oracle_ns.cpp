// oracle_ns.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): the Navier-Stokes parts of the
// reference's time step, restated from the Julia source (src/Transport.jl, src/Velocity.jl,
// src/ELVIRA.jl, src/Tools.jl, src/VFgeom.jl, src/NavierStokes_Parallel.jl). Operation order
// follows the reference expression by expression (Julia's left-associative * and +, x^2 =
// x*x, x^(3/2) = pow(x, 1.5), /̂ = safe divide).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "oracle_core.h"

namespace ora {

// ------------------------------------------------------------------------------------
// compute_props! (src/Tools.jl:1186-1210): face density and viscosity over the ghosted
// extents, from the clamped face average of VF.
// ------------------------------------------------------------------------------------
void compute_props(Ctx& c) {
    const Mesh& m = c.m;
    const double rl = c.p.rho_liq, rg = c.p.rho_gas, ml = c.p.mu_liq, mg = c.p.mu_gas;
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i) {
                const double vx = jmin(1.0, jmax(0.0, (c.VF(i, j, k) + c.VF(i - 1, j, k)) / 2.0));
                c.denx(i, j, k) = rl * vx + rg * (1.0 - vx);
                c.viscx(i, j, k) = vx * ml + (1.0 - vx) * mg;
            }
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i) {
                const double vy = jmin(1.0, jmax(0.0, (c.VF(i, j, k) + c.VF(i, j - 1, k)) / 2.0));
                c.deny(i, j, k) = rl * vy + rg * (1.0 - vy);
                c.viscy(i, j, k) = vy * ml + (1.0 - vy) * mg;
            }
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i) {
                const double vz = jmin(1.0, jmax(0.0, (c.VF(i, j, k) + c.VF(i, j, k - 1)) / 2.0));
                c.denz(i, j, k) = rl * vz + rg * (1.0 - vz);
                c.viscz(i, j, k) = vz * ml + (1.0 - vz) * mg;
            }
}

// ------------------------------------------------------------------------------------
// SLfluxVol (src/VFgeom.jl:428-509): volume swept through face `dir` of cell (i,j,k) in dt,
// the 5-tet volume of its 4 back-traced and 4 fixed face corners. `verts` receives the
// 8 vertices (add_correction_tets builds its tets from them).
// ------------------------------------------------------------------------------------
double sl_flux_vol(const Ctx& c, int dir, int64_t i, int64_t j, int64_t k, double dt, double verts[8][3]) {
    const Mesh& m = c.m;
    const bool even = (i + j + k) % 2 == 0;
    const double tetsign = even ? 1.0 : -1.0;
    auto put = [&](int n, double x, double y, double z, bool proj) {
        verts[n][0] = x; verts[n][1] = y; verts[n][2] = z;
        if (proj) project(c, verts[n], i, j, k, dt);
    };
    const double x0 = m.x(i), x1 = m.x(i + 1), y0 = m.y(j), y1 = m.y(j + 1), z0 = m.z(k), z1 = m.z(k + 1);
    if (dir == 1) {
        if (even) {
            put(0, x0, y0, z0, true);  put(1, x0, y0, z0, false);
            put(2, x0, y1, z0, true);  put(3, x0, y1, z0, false);
            put(4, x0, y0, z1, true);  put(5, x0, y0, z1, false);
            put(6, x0, y1, z1, true);  put(7, x0, y1, z1, false);
        } else {
            put(0, x0, y0, z0, false); put(1, x0, y0, z0, true);
            put(2, x0, y1, z0, false); put(3, x0, y1, z0, true);
            put(4, x0, y0, z1, false); put(5, x0, y0, z1, true);
            put(6, x0, y1, z1, false); put(7, x0, y1, z1, true);
        }
    } else if (dir == 2) {
        if (even) {
            put(0, x0, y0, z0, true);  put(1, x1, y0, z0, true);
            put(2, x0, y0, z0, false); put(3, x1, y0, z0, false);
            put(4, x0, y0, z1, true);  put(5, x1, y0, z1, true);
            put(6, x0, y0, z1, false); put(7, x1, y0, z1, false);
        } else {
            put(0, x1, y0, z0, true);  put(1, x0, y0, z0, true);
            put(2, x1, y0, z0, false); put(3, x0, y0, z0, false);
            put(4, x1, y0, z1, true);  put(5, x0, y0, z1, true);
            put(6, x1, y0, z1, false); put(7, x0, y0, z1, false);
        }
    } else {
        if (even) {
            put(0, x0, y0, z0, true);  put(1, x1, y0, z0, true);
            put(2, x0, y1, z0, true);  put(3, x1, y1, z0, true);
            put(4, x0, y0, z0, false); put(5, x1, y0, z0, false);
            put(6, x0, y1, z0, false); put(7, x1, y1, z0, false);
        } else {
            put(0, x1, y0, z0, true);  put(1, x0, y0, z0, true);
            put(2, x1, y1, z0, true);  put(3, x0, y1, z0, true);
            put(4, x1, y0, z0, false); put(5, x0, y0, z0, false);
            put(6, x1, y1, z0, false); put(7, x0, y1, z0, false);
        }
    }
    double tets[5][4][3];
    verts2tets5(verts, tets);
    double v = 0.0;  // tets_vol (src/VFgeom.jl:683)
    for (int t = 0; t < 5; ++t) v += tet_vol(tets[t]);
    return v * tetsign;
}

// correction_tets_{x,y,z} (src/VFgeom.jl:581-640): apex e of the two tets (a,b,d,e), (b,c,d,e)
// over the quad a,b,c,d whose volume is `vol`; e starts at the quad's centroid and the
// coordinate along the face normal is solved in closed form.
static void correction_tets(int dir, const double a[3], const double b[3], const double cc[3], const double d[3],
                            double vol, double out[2][4][3]) {
    double e[3];
    for (int q = 0; q < 3; ++q) e[q] = 0.25 * (((a[q] + b[q]) + cc[q]) + d[q]);
    // 1-based names of the reference: a[1] = a[0] here
    const double a1 = a[0], a2 = a[1], a3 = a[2], b1 = b[0], b2 = b[1], b3 = b[2];
    const double c1 = cc[0], c2 = cc[1], c3 = cc[2], d1 = d[0], d2 = d[1], d3 = d[2];
    if (dir == 1) {
        const double e2 = e[1], e3 = e[2];
        double num = 6.0 * vol;
        num = num + a1 * b2 * d3;
        num = num - a1 * b3 * d2;
        num = num - a2 * b1 * d3;
        num = num + a2 * b3 * d1;
        num = num + a3 * b1 * d2;
        num = num - a3 * b2 * d1;
        num = num - a1 * b2 * e3;
        num = num + a1 * b3 * e2;
        num = num + a2 * b1 * e3;
        num = num - a3 * b1 * e2;
        num = num + b1 * c2 * d3;
        num = num - b1 * c3 * d2;
        num = num - b2 * c1 * d3;
        num = num + b2 * c3 * d1;
        num = num + b3 * c1 * d2;
        num = num - b3 * c2 * d1;
        num = num + a1 * d2 * e3;
        num = num - a1 * d3 * e2;
        num = num - a2 * d1 * e3;
        num = num + a3 * d1 * e2;
        num = num - b1 * c2 * e3;
        num = num + b1 * c3 * e2;
        num = num + b2 * c1 * e3;
        num = num - b3 * c1 * e2;
        num = num - c1 * d2 * e3;
        num = num + c1 * d3 * e2;
        num = num + c2 * d1 * e3;
        num = num - c3 * d1 * e2;
        double den = a2 * b3;
        den = den - a3 * b2;
        den = den - a2 * d3;
        den = den + a3 * d2;
        den = den + b2 * c3;
        den = den - b3 * c2;
        den = den + c2 * d3;
        den = den - c3 * d2;
        e[0] = num / den;
    } else if (dir == 2) {
        const double e1 = e[0], e3 = e[2];
        double num = 6.0 * vol;
        num = num + a1 * b2 * d3;
        num = num - a1 * b3 * d2;
        num = num - a2 * b1 * d3;
        num = num + a2 * b3 * d1;
        num = num + a3 * b1 * d2;
        num = num - a3 * b2 * d1;
        num = num - a1 * b2 * e3;
        num = num + a2 * b1 * e3;
        num = num - a2 * b3 * e1;
        num = num + a3 * b2 * e1;
        num = num + b1 * c2 * d3;
        num = num - b1 * c3 * d2;
        num = num - b2 * c1 * d3;
        num = num + b2 * c3 * d1;
        num = num + b3 * c1 * d2;
        num = num - b3 * c2 * d1;
        num = num + a1 * d2 * e3;
        num = num - a2 * d1 * e3;
        num = num + a2 * d3 * e1;
        num = num - a3 * d2 * e1;
        num = num - b1 * c2 * e3;
        num = num + b2 * c1 * e3;
        num = num - b2 * c3 * e1;
        num = num + b3 * c2 * e1;
        num = num - c1 * d2 * e3;
        num = num + c2 * d1 * e3;
        num = num - c2 * d3 * e1;
        num = num + c3 * d2 * e1;
        double den = a1 * b3;
        den = den - a3 * b1;
        den = den - a1 * d3;
        den = den + a3 * d1;
        den = den + b1 * c3;
        den = den - b3 * c1;
        den = den + c1 * d3;
        den = den - c3 * d1;
        e[1] = (-num) / den;
    } else {
        const double e1 = e[0], e2 = e[1];
        double num = 6.0 * vol;
        num = num + a1 * b2 * d3;
        num = num - a1 * b3 * d2;
        num = num - a2 * b1 * d3;
        num = num + a2 * b3 * d1;
        num = num + a3 * b1 * d2;
        num = num - a3 * b2 * d1;
        num = num + a1 * b3 * e2;
        num = num - a2 * b3 * e1;
        num = num - a3 * b1 * e2;
        num = num + a3 * b2 * e1;
        num = num + b1 * c2 * d3;
        num = num - b1 * c3 * d2;
        num = num - b2 * c1 * d3;
        num = num + b2 * c3 * d1;
        num = num + b3 * c1 * d2;
        num = num - b3 * c2 * d1;
        num = num - a1 * d3 * e2;
        num = num + a2 * d3 * e1;
        num = num + a3 * d1 * e2;
        num = num - a3 * d2 * e1;
        num = num + b1 * c3 * e2;
        num = num - b2 * c3 * e1;
        num = num - b3 * c1 * e2;
        num = num + b3 * c2 * e1;
        num = num + c1 * d3 * e2;
        num = num - c2 * d3 * e1;
        num = num - c3 * d1 * e2;
        num = num + c3 * d2 * e1;
        double den = a1 * b2;
        den = den - a2 * b1;
        den = den - a1 * d2;
        den = den + a2 * d1;
        den = den + b1 * c2;
        den = den - b2 * c1;
        den = den + c1 * d2;
        den = den - c2 * d1;
        e[2] = num / den;
    }
    for (int q = 0; q < 3; ++q) {
        out[0][0][q] = a[q]; out[0][1][q] = b[q]; out[0][2][q] = d[q]; out[0][3][q] = e[q];
        out[1][0][q] = b[q]; out[1][1][q] = cc[q]; out[1][2][q] = d[q]; out[1][3][q] = e[q];
    }
}

// add_correction_tets (src/VFgeom.jl:513-664): two tets per face of cell (i,j,k), faces x-,
// x+, y-, y+, z-, z+, making each face's SL flux volume equal to dt * A * u_face. Writes
// tets 5..16 and their cell indices (pt2index! from the cell itself).
bool add_correction_tets(const Ctx& c, int64_t i, int64_t j, int64_t k, double dt, double tets[17][4][3],
                         int64_t inds[17][4][3]) {
    const Mesh& m = c.m;
    struct Face { int dir; int64_t fi, fj, fk; };
    const Face faces[6] = {{1, i, j, k}, {1, i + 1, j, k}, {2, i, j, k}, {2, i, j + 1, k}, {3, i, j, k}, {3, i, j, k + 1}};
    int n = 5;
    for (const Face& f : faces) {
        double verts[8][3];
        const double flux_vol = sl_flux_vol(c, f.dir, f.fi, f.fj, f.fk, dt, verts);
        double divg_vol;
        if (f.dir == 1) divg_vol = ((dt * m.dy) * m.dz) * c.uf(f.fi, f.fj, f.fk);
        else if (f.dir == 2) divg_vol = ((dt * m.dx) * m.dz) * c.vf(f.fi, f.fj, f.fk);
        else divg_vol = ((dt * m.dx) * m.dy) * c.wf(f.fi, f.fj, f.fk);
        const bool even = (f.fi + f.fj + f.fk) % 2 == 0;
        const double vol = even ? (divg_vol - flux_vol) : -(divg_vol - flux_vol);
        double out[2][4][3];
        if (f.dir == 1) {
            if (even) correction_tets(1, verts[4], verts[6], verts[2], verts[0], vol, out);
            else correction_tets(1, verts[7], verts[5], verts[1], verts[3], vol, out);
        } else if (f.dir == 2) {
            correction_tets(2, verts[1], verts[5], verts[4], verts[0], vol, out);
        } else {
            correction_tets(3, verts[1], verts[0], verts[2], verts[3], vol, out);
        }
        std::memcpy(tets[n], out, sizeof(out));
        n += 2;
    }
    bool ok = true;
    for (int t = 5; t < 17; ++t)
        for (int v = 0; v < 4; ++v) ok = pt2index(c, inds[t][v], tets[t][v], i, j, k) && ok;
    return ok;
}

// ------------------------------------------------------------------------------------
// compute_curvature! (src/ELVIRA.jl:143-192): height function over the 3x3x7 column along
// the dominant normal direction; 0 where no component strictly dominates.
// ------------------------------------------------------------------------------------
// The direction the reference's strict comparisons pick: 0 x, 1 y, 2 z, -1 none dominates.
static int dominant_direction(double anx, double any, double anz) {
    if (anx > any && anx > anz) return 0;
    if (any > anx && any > anz) return 1;
    if (anz > any && anz > anx) return 2;
    return -1;
}

// A hinted height-function direction ties with the oracle's own when the normal components
// that decide between them are equal to rounding (1e-12 relative): on a symmetric interface
// |nx| = |ny| in exact arithmetic, and the last bits of the two implementations' normals
// then pick the direction. Direction d is admissible when |n_d| is the largest component to
// that tolerance; "none" when the two largest are equal to it.
static bool direction_tie(int hint, const double a[3]) {
    const double mx = std::max(a[0], std::max(a[1], a[2]));
    const double tol = 1e-12 * mx;
    if (hint >= 0 && hint < 3) return mx - a[hint] <= tol;
    if (hint != -1) return false;
    int top = 0;
    for (int d = 1; d < 3; ++d)
        if (a[d] > a[top]) top = d;
    for (int d = 0; d < 3; ++d)
        if (d != top && a[top] - a[d] <= tol) return true;
    return false;
}

static void curvature_cell(Ctx& c, int64_t i, int64_t j, int64_t k, const std::vector<int32_t>* hint,
                           int64_t stats[3]) {
    const Mesh& m = c.m;
    const double anx = std::fabs(c.nx(i, j, k)), any = std::fabs(c.ny(i, j, k)), anz = std::fabs(c.nz(i, j, k));
    double hf[3][3] = {};  // hf[a][b], a and b the two tangential offsets + 1
    double h1, h2;         // spacings of the first / second tangential axis
    int dom = dominant_direction(anx, any, anz);
    if (hint) {
        const int h = (*hint)[(size_t)((i - 1) + m.Nx * ((j - 1) + m.Ny * (k - 1)))];
        if (h != dom) {
            const double a[3] = {anx, any, anz};
            ++stats[0];
            if (direction_tie(h, a)) {
                ++stats[1];
                dom = h;
            } else {
                ++stats[2];
                if (c.refusals.size() < 64)
                    c.refusals.push_back({1, i, j, k, h, dom, h >= 0 && h < 3 ? a[h] : NAN, dom >= 0 ? a[dom] : NAN});
            }
        }
    }
    if (dom < 0) return;
    if (dom == 0) {  // hf[jj,kk], loop kk, jj, ii
        for (int kk = -1; kk <= 1; ++kk)
            for (int jj = -1; jj <= 1; ++jj)
                for (int ii = -3; ii <= 3; ++ii) hf[jj + 1][kk + 1] += c.VF(i + ii, j + jj, k + kk);
        for (auto& r : hf) for (double& v : r) v *= m.dx;
        h1 = m.dy; h2 = m.dz;
    } else if (dom == 1) {  // hf[kk,ii], loop ii, kk, jj
        for (int ii = -1; ii <= 1; ++ii)
            for (int kk = -1; kk <= 1; ++kk)
                for (int jj = -3; jj <= 3; ++jj) hf[kk + 1][ii + 1] += c.VF(i + ii, j + jj, k + kk);
        for (auto& r : hf) for (double& v : r) v *= m.dy;
        h1 = m.dz; h2 = m.dx;
    } else {  // hf[ii,jj], loop jj, ii, kk
        for (int jj = -1; jj <= 1; ++jj)
            for (int ii = -1; ii <= 1; ++ii)
                for (int kk = -3; kk <= 3; ++kk) hf[ii + 1][jj + 1] += c.VF(i + ii, j + jj, k + kk);
        for (auto& r : hf) for (double& v : r) v *= m.dz;
        h1 = m.dx; h2 = m.dy;
    }
    // first index = first tangential axis (y, z, x), second = second (z, x, y)
    const double f1 = (hf[2][1] - hf[0][1]) / (2.0 * h1);
    const double f2 = (hf[1][2] - hf[1][0]) / (2.0 * h2);
    const double f11 = ((hf[2][1] - 2.0 * hf[1][1]) + hf[0][1]) / (h1 * h1);
    const double f22 = ((hf[1][2] - 2.0 * hf[1][1]) + hf[1][0]) / (h2 * h2);
    // the reference writes the mixed difference and its denominator per direction
    const double f12 = dom == 0 ? (((hf[2][2] - hf[0][2]) - hf[2][0]) + hf[0][0]) / ((4.0 * h1) * h2)
                     : dom == 1 ? (((hf[2][2] - hf[2][0]) - hf[0][2]) + hf[0][0]) / ((4.0 * h1) * h2)
                                : (((hf[2][2] - hf[2][0]) - hf[0][2]) + hf[0][0]) / ((4.0 * h2) * h1);
    c.Curve(i, j, k) = ((((f11 + f22) + f11 * (f2 * f2)) + f22 * (f1 * f1)) - ((2.0 * f12) * f1) * f2) /
                       std::pow((1.0 + f1 * f1) + f2 * f2, 1.5);
}

void compute_curvature(Ctx& c) {
    const Mesh& m = c.m;
    std::fill(c.Curve.a.begin(), c.Curve.a.end(), 0.0);
    std::vector<int32_t> hint;
    if (!c.curv_hints.empty()) {
        hint = std::move(c.curv_hints.front());
        c.curv_hints.pop_front();
    }
    int64_t stats[3] = {};
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) curvature_cell(c, i, j, k, hint.empty() ? nullptr : &hint, stats);
    for (int q = 0; q < 3; ++q) c.tie_stats[1][q] += stats[q];
}

// mask_maker! (src/Tools.jl:1582-1598): faces next to positive curvature; never reset.
void mask_maker(Ctx& c) {
    const Mesh& m = c.m;
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                int16_t& mk = c.mask(i, j, k);
                if (c.Curve(i, j, k) > 0 || c.Curve(i - 1, j, k) > 0) mk |= 1;
                if (c.Curve(i, j, k) > 0 || c.Curve(i, j - 1, k) > 0) mk |= 2;
                if (c.Curve(i, j, k) > 0 || c.Curve(i, j, k - 1) > 0) mk |= 4;
            }
}

// compute_sf! (src/ELVIRA.jl:256-299): continuum surface force on the masked faces.
void compute_sf(Ctx& c) {
    const Mesh& m = c.m;
    const double sigma = c.p.sigma;
    std::fill(c.sfx.a.begin(), c.sfx.a.end(), 0.0);
    std::fill(c.sfy.a.begin(), c.sfy.a.end(), 0.0);
    std::fill(c.sfz.a.begin(), c.sfz.a.end(), 0.0);
    auto face = [&](double C0, double C1, double dVF, double h) {
        // C1 = curvature of the cell, C0 = of its lower neighbour
        if (C1 > 0 && C0 > 0) return ((-sigma) / (2.0 * h)) * dVF * (C1 + C0);
        if (C1 == 0) return ((-sigma) / h) * dVF * C0;
        if (C0 == 0) return ((-sigma) / h) * dVF * C1;
        return 0.0;
    };
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_ + 1; ++i)
                if (c.mask(i, j, k) & 1)
                    c.sfx(i, j, k) = face(c.Curve(i - 1, j, k), c.Curve(i, j, k), c.VF(i, j, k) - c.VF(i - 1, j, k), m.dx);
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i)
                if (c.mask(i, j, k) & 2)
                    c.sfy(i, j, k) = face(c.Curve(i, j - 1, k), c.Curve(i, j, k), c.VF(i, j, k) - c.VF(i, j - 1, k), m.dy);
    for (int64_t k = m.kmin_; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i)
                if (c.mask(i, j, k) & 4)
                    c.sfz(i, j, k) = face(c.Curve(i, j, k - 1), c.Curve(i, j, k), c.VF(i, j, k) - c.VF(i, j, k - 1), m.dz);
}

// BC! of the examples (e.g. examples/example1.jl:89-136): on every non-periodic side the first
// ghost layer is -interior + 2 * wall velocity, per component.
void apply_bc(Ctx& c, Arr3<double>& u, Arr3<double>& v, Arr3<double>& w) {
    const Mesh& m = c.m;
    Arr3<double>* F[3] = {&u, &v, &w};
    const int per[3] = {(int)c.p.xper, (int)c.p.yper, (int)c.p.zper};
    const int64_t lo[3] = {m.imin_, m.jmin_, m.kmin_}, hi[3] = {m.imax_, m.jmax_, m.kmax_};
    for (int d = 0; d < 3; ++d) {
        if (per[d]) continue;
        for (int side = 0; side < 2; ++side) {
            const int64_t g = side == 0 ? lo[d] - 1 : hi[d] + 1, in = side == 0 ? lo[d] : hi[d];
            const double* uw = c.bc_wall[2 * d + side];
            for (int q = 0; q < 3; ++q) {
                Arr3<double>& A = *F[q];
                for (int64_t b = A.e.lo[(d + 2) % 3]; b <= A.hi((d + 2) % 3); ++b)
                    for (int64_t a = A.e.lo[(d + 1) % 3]; a <= A.hi((d + 1) % 3); ++a) {
                        int64_t ig[3], ii[3];
                        ig[d] = g; ii[d] = in;
                        ig[(d + 1) % 3] = ii[(d + 1) % 3] = a;
                        ig[(d + 2) % 3] = ii[(d + 2) % 3] = b;
                        A(ig[0], ig[1], ig[2]) = -A(ii[0], ii[1], ii[2]) + 2.0 * uw[q];
                    }
            }
        }
    }
}

// interpolateFace! (src/Velocity.jl:23-37) of the cell velocities u, v, w
void interpolate_face(Ctx& c, const Arr3<double>& u, const Arr3<double>& v, const Arr3<double>& w) {
    const Mesh& m = c.m;
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 2; ++i) c.uf(i, j, k) = 0.5 * (u(i - 1, j, k) + u(i, j, k));
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 2; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 1; ++i) c.vf(i, j, k) = 0.5 * (v(i, j - 1, k) + v(i, j, k));
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 2; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 1; ++i) c.wf(i, j, k) = 0.5 * (w(i, j, k - 1) + w(i, j, k));
}

// interpolateCenter! (src/Velocity.jl:39-57): u = us everywhere (keeps the BCs), then the
// interior from the face velocities
void interpolate_center(Ctx& c) {
    const Mesh& m = c.m;
    c.u.a = c.us.a;
    c.v.a = c.vs.a;
    c.w.a = c.ws.a;
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                c.u(i, j, k) = 0.5 * (c.uf(i, j, k) + c.uf(i + 1, j, k));
                c.v(i, j, k) = 0.5 * (c.vf(i, j, k) + c.vf(i, j + 1, k));
                c.w(i, j, k) = 0.5 * (c.wf(i, j, k) + c.wf(i, j, k + 1));
            }
}

// divg_cell (src/Velocity.jl:72-122): the SL pre-image volume in band cells under the
// semi-Lagrangian scheme, finite-volume fluxes elsewhere (SL flux volumes at band-edge faces).
double divg_cell(const Ctx& c, int64_t i, int64_t j, int64_t k, double dt) {
    const Mesh& m = c.m;
    const bool sl = c.p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN;
    const int64_t nb = c.p.nband;
    const double vol = (m.dx * m.dy) * m.dz;
    double verts[8][3];
    if (sl && std::abs((int)c.band(i, j, k)) <= nb) {
        const double tetsign = cell2verts(c, i, j, k, verts);
        for (int n = 0; n < 8; ++n) project(c, verts[n], i, j, k, dt);
        double tets[5][4][3];
        verts2tets5(verts, tets);
        double tv = 0.0;
        for (int t = 0; t < 5; ++t) tv += tet_vol(tets[t]);
        const double v1 = tetsign * tv;
        return (vol - v1) / vol / dt;
    }
    auto bandat = [&](int64_t a, int64_t b, int64_t cc) { return sl && std::abs((int)c.band(a, b, cc)) <= nb; };
    const double Fxm = bandat(i - 1, j, k) ? sl_flux_vol(c, 1, i, j, k, dt, verts) / dt : (m.dy * m.dz) * c.uf(i, j, k);
    const double Fxp = bandat(i + 1, j, k) ? sl_flux_vol(c, 1, i + 1, j, k, dt, verts) / dt : (m.dy * m.dz) * c.uf(i + 1, j, k);
    const double Fym = bandat(i, j - 1, k) ? sl_flux_vol(c, 2, i, j, k, dt, verts) / dt : (m.dx * m.dz) * c.vf(i, j, k);
    const double Fyp = bandat(i, j + 1, k) ? sl_flux_vol(c, 2, i, j + 1, k, dt, verts) / dt : (m.dx * m.dz) * c.vf(i, j + 1, k);
    const double Fzm = bandat(i, j, k - 1) ? sl_flux_vol(c, 3, i, j, k, dt, verts) / dt : (m.dx * m.dy) * c.wf(i, j, k);
    const double Fzp = bandat(i, j, k + 1) ? sl_flux_vol(c, 3, i, j, k + 1, dt, verts) / dt : (m.dx * m.dy) * c.wf(i, j, k + 1);
    return (((((Fxp - Fxm) + Fyp) - Fym) + Fzp) - Fzm) / vol;
}

// divergence! (src/Velocity.jl:59-70): divg_cell on the interior, 0 elsewhere.
void divergence(Ctx& c, double dt) {
    const Mesh& m = c.m;
    std::fill(c.divg.a.begin(), c.divg.a.end(), 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) c.divg(i, j, k) = divg_cell(c, i, j, k, dt);
}

// ------------------------------------------------------------------------------------
// transport! with the momentum predictor (src/Transport.jl:2-268)
// ------------------------------------------------------------------------------------
int transport_ns(Ctx& c, double dt) {
    const Mesh& m = c.m;
    const bool sl = c.p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN;
    const int64_t nb = c.p.nband;
    const double dx = m.dx, dy = m.dy, dz = m.dz;
    compute_normal(c);
    c.status = 0;
    const Ext ce = m.cell_ext();
    Arr3<double> F[9];  // Fux Fvx Fwx Fuy Fvy Fwy Fuz Fvz Fwz
    for (auto& a : F) a.init(ce, 0.0);
    // convective fluxes (src/Transport.jl:35-105)
    for (int d = 0; d < 3; ++d) {
        const int64_t e0 = d == 0, e1 = d == 1, e2 = d == 2;
        for (int64_t k = m.kmin_; k <= m.kmax_ + e2; ++k)
            for (int64_t j = m.jmin_; j <= m.jmax_ + e1; ++j)
                for (int64_t i = m.imin_; i <= m.imax_ + e0; ++i) {
                    const int64_t a = i - e0, b = j - e1, cc = k - e2;
                    const double uface = 0.5 * (c.u(a, b, cc) + c.u(i, j, k));
                    const double vface = 0.5 * (c.v(a, b, cc) + c.v(i, j, k));
                    const double wface = 0.5 * (c.w(a, b, cc) + c.w(i, j, k));
                    const bool bm = std::abs((int)c.band(a, b, cc)) <= nb, bp = std::abs((int)c.band(i, j, k)) <= nb;
                    double fu, fv, fw;
                    if (bm && bp) {
                        fu = fv = fw = 0.0;
                    } else if ((bm || bp) && sl) {
                        double verts[8][3];
                        const double vol = sl_flux_vol(c, d + 1, i, j, k, dt, verts) / dt;
                        fu = -uface * vol; fv = -vface * vol; fw = -wface * vol;
                    } else {
                        // -uface * dy*dz*uf[i,j,k] (x faces), dx*dz (y), dx*dy (z): left to right
                        const double h1 = d == 0 ? dy : dx, h2 = d == 2 ? dy : dz;
                        const double vel = d == 0 ? c.uf(i, j, k) : (d == 1 ? c.vf(i, j, k) : c.wf(i, j, k));
                        fu = ((-uface * h1) * h2) * vel; fv = ((-vface * h1) * h2) * vel; fw = ((-wface * h1) * h2) * vel;
                    }
                    F[3 * d + 0](i, j, k) = fu; F[3 * d + 1](i, j, k) = fv; F[3 * d + 2](i, j, k) = fw;
                }
    }
    // SL cells and flux-updated cells (src/Transport.jl:107-170)
    Arr3<double> VFnew;
    VFnew.init(ce, 0.0);
    const double cv = (dx * dy) * dz;
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        std::unique_ptr<CutScratch, void (*)(CutScratch*)> s(new_cut_scratch(), delete_cut_scratch);
#pragma omp for collapse(2) schedule(dynamic, 1)
        for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
            for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
                for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                    if (std::abs((int)c.band(i, j, k)) <= nb) {
                        double verts[8][3], tets[17][4][3];
                        int64_t vInds[8][3], inds[17][4][3];
                        const double tetsign = cell2verts(c, i, j, k, verts);
                        for (int n = 0; n < 8; ++n) project(c, verts[n], i, j, k, dt);
                        double t5[5][4][3];
                        verts2tets5(verts, t5);
                        std::memcpy(tets, t5, sizeof(t5));
                        bool ok = true;
                        for (int n = 0; n < 8; ++n) ok = pt2index(c, vInds[n], verts[n], i, j, k) && ok;
                        int64_t i5[5][4][3];
                        inds2tets5(vInds, i5);
                        std::memcpy(inds, i5, sizeof(i5));
                        int ntets = 5;
                        if (!sl) {
                            ok = add_correction_tets(c, i, j, k, dt, tets, inds) && ok;
                            ntets = 17;
                        }
                        if (!ok) bad = 1;
                        double vol = 0, vLiq = 0, vU = 0, vV = 0, vW = 0, vrho = 0;
                        for (int t = 0; t < ntets; ++t) {
                            CutSums r = cut_tet(c, tets[t], inds[t], false, false, false, 1, *s);
                            vol += tetsign * r.vol; vLiq += tetsign * r.vLiq;
                            vU += tetsign * r.vU; vV += tetsign * r.vV; vW += tetsign * r.vW;
                            vrho += tetsign * r.vrho;
                        }
                        VFnew(i, j, k) = vLiq / vol;
                        c.us(i, j, k) = vU / vrho;
                        c.vs(i, j, k) = vV / vrho;
                        c.ws(i, j, k) = vW / vrho;
                    } else {
                        VFnew(i, j, k) = c.VF(i, j, k);
                        const double f = dt / cv;
                        c.us(i, j, k) = c.u(i, j, k) + f * (((((F[0](i + 1, j, k) - F[0](i, j, k)) + F[3](i, j + 1, k)) -
                                                               F[3](i, j, k)) + F[6](i, j, k + 1)) - F[6](i, j, k));
                        c.vs(i, j, k) = c.v(i, j, k) + f * (((((F[1](i + 1, j, k) - F[1](i, j, k)) + F[4](i, j + 1, k)) -
                                                               F[4](i, j, k)) + F[7](i, j, k + 1)) - F[7](i, j, k));
                        c.ws(i, j, k) = c.w(i, j, k) + f * (((((F[2](i + 1, j, k) - F[2](i, j, k)) + F[5](i, j + 1, k)) -
                                                               F[5](i, j, k)) + F[8](i, j, k + 1)) - F[8](i, j, k));
                    }
                }
    }
    if (bad) c.status = MIST_ERR_OVERFLOW;
    // new VF, normals, curvature, mask, surface force (src/Transport.jl:176-197)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) c.VF(i, j, k) = VFnew(i, j, k);
    neumann(c, c.VF);
    update_borders(c, c.VF);
    compute_normal(c);
    compute_curvature(c);
    mask_maker(c);
    compute_sf(c);
    // viscous fluxes (src/Transport.jl:205-234), from the old cell velocities
    const Arr3<double>* visc[3] = {&c.viscx, &c.viscy, &c.viscz};
    const Arr3<double>* den[3] = {&c.denx, &c.deny, &c.denz};
    for (int d = 0; d < 3; ++d) {
        const int64_t e0 = d == 0, e1 = d == 1, e2 = d == 2;
        const double h = d == 0 ? dx : (d == 1 ? dy : dz);
        const double area = d == 0 ? dy * dz : (d == 1 ? dx * dz : dx * dy);
        for (int64_t k = m.kmin_; k <= m.kmax_ + e2; ++k)
            for (int64_t j = m.jmin_; j <= m.jmax_ + e1; ++j)
                for (int64_t i = m.imin_; i <= m.imax_ + e0; ++i) {
                    const int64_t a = i - e0, b = j - e1, cc = k - e2;
                    const double mr = sdiv((*visc[d])(i, j, k), (*den[d])(i, j, k));
                    F[3 * d + 0](i, j, k) = area * (mr * ((c.u(i, j, k) - c.u(a, b, cc)) / h));
                    F[3 * d + 1](i, j, k) = area * (mr * ((c.v(i, j, k) - c.v(a, b, cc)) / h));
                    F[3 * d + 2](i, j, k) = area * (mr * ((c.w(i, j, k) - c.w(a, b, cc)) / h));
                }
    }
    // viscous, surface tension and gravity terms (src/Transport.jl:237-253)
    const double f = dt / cv;
    const double gx = c.p.grav_x, gy = c.p.grav_y, gz = c.p.grav_z;
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                c.us(i, j, k) = (c.us(i, j, k) + f * (((((F[0](i + 1, j, k) - F[0](i, j, k)) + F[3](i, j + 1, k)) -
                                                       F[3](i, j, k)) + F[6](i, j, k + 1)) - F[6](i, j, k))) +
                                dt * ((0.5 * (c.sfx(i, j, k) + c.sfx(i + 1, j, k))) /
                                          (0.5 * (c.denx(i + 1, j, k) + c.denx(i, j, k))) - gx);
                c.vs(i, j, k) = (c.vs(i, j, k) + f * (((((F[1](i + 1, j, k) - F[1](i, j, k)) + F[4](i, j + 1, k)) -
                                                       F[4](i, j, k)) + F[7](i, j, k + 1)) - F[7](i, j, k))) +
                                dt * ((0.5 * (c.sfy(i, j, k) + c.sfy(i, j + 1, k))) /
                                          (0.5 * (c.deny(i, j + 1, k) + c.deny(i, j, k))) - gy);
                c.ws(i, j, k) = (c.ws(i, j, k) + f * (((((F[2](i + 1, j, k) - F[2](i, j, k)) + F[5](i, j + 1, k)) -
                                                       F[5](i, j, k)) + F[8](i, j, k + 1)) - F[8](i, j, k))) +
                                dt * ((0.5 * (c.sfz(i, j, k) + c.sfz(i, j, k + 1))) /
                                          (0.5 * (c.denz(i, j, k + 1) + c.denz(i, j, k))) - gz);
            }
    apply_bc(c, c.us, c.vs, c.ws);
    compute_plic(c);
    update_borders(c, c.us);
    update_borders(c, c.vs);
    update_borders(c, c.ws);
    return c.status;
}

// One iteration of run_solver with solveNS = true (src/NavierStokes_Parallel.jl:125-168):
// dt, transport!, computeBand!, compute_props!, interpolateFace!, pressure_solver!,
// corrector!, interpolateCenter!, update_borders! of u, v, w, and the divergence! check.
int ns_step(Ctx& c, double* t, double clamp_to, double* dt_used, double tol, int precision, int64_t* cycles,
            double* resid) {
    double dt = compute_dt(c);
    if (clamp_to > 0.0 && *t + dt > clamp_to) dt = clamp_to - *t;
    *t += dt;
    *dt_used = dt;
    int st = transport_ns(c, dt);
    if (st) return st;
    compute_band(c);
    compute_props(c);
    interpolate_face(c, c.us, c.vs, c.ws);
    if (c.p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN) {
        // pressure_solver! with RHS = nothing: the SL operator's FAS cycles (src/Pressure.jl:5-30)
        st = pressure_solve_sl(c, dt, tol, c.p.mg_max_cycles, c.sl_exact_restore, cycles, resid, nullptr);
    } else {
        pressure_rhs(c);
        st = pressure_solve(c, dt, tol, precision, cycles, resid);
    }
    if (st && st != MIST_ERR_NOCONV) return st;
    corrector(c, dt);
    interpolate_center(c);
    update_borders(c, c.u);
    update_borders(c, c.v);
    update_borders(c, c.w);
    divergence(c, dt);
    return st;
}

}  // namespace ora
