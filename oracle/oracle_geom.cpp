// oracle_geom.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): mesh, cut tables, PLIC,
// ELVIRA, back-trace and tet cutting, restated from the reference Julia source.
#include <algorithm>
#include <cstring>

#include "oracle_core.h"
#include "oracle_tables.h"

namespace ora {

thread_local FlopCount t_flops;
bool g_count_flops = false;
#define FL(n) do { if (g_count_flops) t_flops.flops += (uint64_t)(n); } while (0)
#define FLS(n) do { if (g_count_flops) t_flops.special += (uint64_t)(n); } while (0)

// ------------------------------------------------------------------------------------
// Mesh (src/Mesh.jl:55-196 and :204-352; one rank, imin_ = 1)
// ------------------------------------------------------------------------------------
void Mesh::build(const mist_params& p, int level) {
    int64_t f = (int64_t)1 << (level - 1);  // mg_factor = 2^(lvl-1)
    Nx = std::max<int64_t>(p.Nx / f, 1);
    Ny = std::max<int64_t>(p.Ny / f, 1);
    Nz = std::max<int64_t>(p.Nz / f, 1);
    Lx = p.Lx; Ly = p.Ly; Lz = p.Lz;
    ng = 3 + std::max<int64_t>(0, (int64_t)std::ceil(p.CFL) - 2);
    imin_ = 1; imax_ = Nx; jmin_ = 1; jmax_ = Ny; kmin_ = 1; kmax_ = Nz;
    imino_ = imin_ - ng; imaxo_ = imax_ + ng;
    jmino_ = jmin_ - ng; jmaxo_ = jmax_ + ng;
    kmino_ = kmin_ - ng; kmaxo_ = kmax_ + ng;
    auto axis = [&](int64_t N, double L, std::vector<double>& fc, std::vector<double>& cc, double& h) {
        fc.assign((size_t)(N + 2 * ng + 1), 0.0);
        // range(0, stop=L, length=N+1): element q is the rounded value of q*L/N.
        for (int64_t q = 0; q <= N; ++q) fc[(size_t)(ng + q)] = ((double)q * L) / (double)N;
        h = fc[(size_t)(ng + 1)] - fc[(size_t)ng];
        for (int64_t n = 1; n <= ng; ++n) {
            fc[(size_t)(ng - n)] = fc[(size_t)ng] - (double)n * h;
            fc[(size_t)(ng + N + n)] = fc[(size_t)(ng + N)] + (double)n * h;
        }
        cc.assign((size_t)(N + 2 * ng), 0.0);
        for (size_t q = 0; q < cc.size(); ++q) cc[q] = 0.5 * (fc[q] + fc[q + 1]);
    };
    axis(Nx, Lx, xs, xms, dx);
    axis(Ny, Ly, ys, yms, dy);
    axis(Nz, Lz, zs, zms, dz);
}

Ext Mesh::cell_ext() const {
    return Ext{{imino_, jmino_, kmino_}, {imaxo_ - imino_ + 1, jmaxo_ - jmino_ + 1, kmaxo_ - kmino_ + 1}};
}
Ext Mesh::vof_ext() const {
    return Ext{{imino_ - 3, jmino_ - 3, kmino_ - 3},
               {imaxo_ - imino_ + 7, jmaxo_ - jmino_ + 7, kmaxo_ - kmino_ + 7}};
}
Ext Mesh::face_ext(int dir) const {
    Ext e = cell_ext();
    e.n[dir] += 1;
    return e;
}


// tet_vol (src/VFgeom.jl:669-681)
double tet_vol(const double v[4][3]) {
    const double a1 = v[0][0] - v[3][0], a2 = v[0][1] - v[3][1], a3 = v[0][2] - v[3][2];
    const double b1 = v[1][0] - v[3][0], b2 = v[1][1] - v[3][1], b3 = v[1][2] - v[3][2];
    const double c1 = v[2][0] - v[3][0], c2 = v[2][1] - v[3][1], c3 = v[2][2] - v[3][2];
    FL(24);
    return (-1.0 / 6.0) * ((a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3)) + a3 * (b1 * c2 - c1 * b2));
}

// d2case (src/VFgeom.jl:693-708): snap |d| < eps by majority sign, then case bits.
int d2case_inplace(double d[4]) {
    int npos = 0, nneg = 0;
    for (int n = 0; n < 4; ++n) { npos += d[n] > 0.0; nneg += d[n] < 0.0; }
    for (int n = 0; n < 4; ++n)
        if (std::fabs(d[n]) < EPS) d[n] = npos > nneg ? EPS : -EPS;
    int c = 0;
    for (int n = 0; n < 4; ++n) c += (d[n] > 0.0) << n;
    return c;  // reference case - 1
}
int d2case_copy(const double d[4]) {
    double t[4] = {d[0], d[1], d[2], d[3]};
    return d2case_inplace(t);
}

// computeAlpha (src/VF.jl:164-214), Scardovelli & Zaleski (2000) closed form.
double compute_alpha(double a, double b, double cc, double VF) {
    double m[3] = {a, b, cc};
    std::sort(m, m + 3);
    const double m1 = m[0], m2 = m[1], m3 = m[2];
    const double m12 = m1 + m2;
    const double V1 = (m1 * m1) / jmax((6.0 * m2) * m3, EPS);
    const double V2 = V1 + sdiv(0.5 * (m2 - m1), m3);
    double V3;
    FL(9);
    if (m12 <= m3) {
        V3 = sdiv(0.5 * m12, m3);
        FL(2);
    } else {
        V3 = sdiv(((m3 * m3) * (3.0 * m12 - m3) + (m1 * m1) * (m1 - 3.0 * m3)) + (m2 * m2) * (m2 - 3.0 * m3),
                  ((6.0 * m1) * m2) * m3);
        FL(19);
    }
    double alpha;
    if (0.0 <= VF && VF < V1) {
        alpha = std::cbrt((((6.0 * m1) * m2) * m3) * VF);
        FL(4); FLS(1);
    } else if (V1 <= VF && VF < V2) {
        alpha = 0.5 * (m1 + std::sqrt((m1 * m1) + ((8.0 * m2) * m3) * (VF - V1)));
        FL(8); FLS(1);
    } else if (V2 <= VF && VF < V3) {
        double a0 = ((m1 * m1) * m1 + (m2 * m2) * m2) - (((6.0 * m1) * m2) * m3) * VF;
        double a1 = -3.0 * ((m1 * m1) + (m2 * m2));
        double a2 = 3.0 * m12;
        double a3 = -1.0;
        a0 = a0 / a3; a1 = a1 / a3; a2 = a2 / a3; a3 = a3 / a3;
        const double p0 = a1 / 3.0 - (a2 * a2) / 9.0;
        const double q0 = (a1 * a2 - 3.0 * a0) / 6.0 - ((a2 * a2) * a2) / 27.0;
        const double theta = std::acos(jmin(1.0, jmax(-1.0, q0 / std::sqrt(-1.0 * ((p0 * p0) * p0))))) / 3.0;
        alpha = std::sqrt(-p0) * (std::sqrt(3.0) * std::sin(theta) - std::cos(theta)) - a2 / 3.0;
        FL(44); FLS(6);
    } else {
        if (m12 <= m3) {
            alpha = m3 * VF + 0.5 * m12;
            FL(3);
        } else {
            double a0 = (((m1 * m1) * m1 + (m2 * m2) * m2) + (m3 * m3) * m3) - (((6.0 * m1) * m2) * m3) * VF;
            double a1 = -3.0 * (((m1 * m1) + (m2 * m2)) + (m3 * m3));
            double a2 = 3.0;
            double a3 = -2.0;
            a0 = a0 / a3; a1 = a1 / a3; a2 = a2 / a3; a3 = a3 / a3;
            const double p0 = a1 / 3.0 - (a2 * a2) / 9.0;
            const double q0 = (a1 * a2 - 3.0 * a0) / 6.0 - ((a2 * a2) * a2) / 27.0;
            const double theta = std::acos(jmin(1.0, jmax(-1.0, q0 / std::sqrt(-((p0 * p0) * p0))))) / 3.0;
            alpha = std::sqrt(-p0) * (std::sqrt(3.0) * std::sin(theta) - std::cos(theta)) - a2 / 3.0;
            FL(48); FLS(6);
        }
    }
    return alpha;
}

// computeDist (src/VF.jl:109-156)
double compute_dist(const Ctx& c, int64_t i, int64_t j, int64_t k, double nx, double ny, double nz,
                    double VF) {
    const Mesh& m = c.m;
    if (VF < c.p.VFlo) return -((m.Lx + m.Ly) + m.Lz);
    if (VF > c.p.VFhi) return +((m.Lx + m.Ly) + m.Lz);
    double mm0 = nx * m.dx, mm1 = ny * m.dy, mm2 = nz * m.dz;
    const double norm = ((std::fabs(mm0) + std::fabs(mm1)) + std::fabs(mm2)) + EPS;
    mm0 = mm0 / norm; mm1 = mm1 / norm; mm2 = mm2 / norm;
    double factor = 0.0;
    if (mm0 < 0.0) factor += mm0;
    if (mm1 < 0.0) factor += mm1;
    if (mm2 < 0.0) factor += mm2;
    const double VFo = VF > 0.5 ? 1.0 - VF : VF;
    double alpha = compute_alpha(std::fabs(mm0), std::fabs(mm1), std::fabs(mm2), VFo);
    alpha = VF > 0.5 ? 1.0 - alpha : alpha;
    alpha += factor;
    alpha -= 0.5 * ((mm0 + mm1) + mm2);
    double dist = alpha * norm;
    dist += (nx * m.xm(i) + ny * m.ym(j)) + nz * m.zm(k);
    FL(24);
    return dist;
}

// oddEvenCell + cell2verts! (src/VFgeom.jl:171-216), 5-tet tesselation.
double cell2verts(const Ctx& c, int64_t i, int64_t j, int64_t k, double v[8][3]) {
    const Mesh& m = c.m;
    const double x0 = m.x(i), x1 = m.x(i + 1), y0 = m.y(j), y1 = m.y(j + 1), z0 = m.z(k), z1 = m.z(k + 1);
    if ((i + j + k) % 2 == 0) {
        const double vv[8][3] = {{x0, y0, z0}, {x1, y0, z0}, {x0, y1, z0}, {x1, y1, z0},
                                 {x0, y0, z1}, {x1, y0, z1}, {x0, y1, z1}, {x1, y1, z1}};
        std::memcpy(v, vv, sizeof(vv));
        return +1.0;
    } else {
        const double vv[8][3] = {{x1, y0, z0}, {x0, y0, z0}, {x1, y1, z0}, {x0, y1, z0},
                                 {x1, y0, z1}, {x0, y0, z1}, {x1, y1, z1}, {x0, y1, z1}};
        std::memcpy(v, vv, sizeof(vv));
        return -1.0;
    }
}

void verts2tets5(const double p[8][3], double tets[5][4][3]) {
    for (int t = 0; t < 5; ++t)
        for (int n = 0; n < 4; ++n)
            for (int q = 0; q < 3; ++q) tets[t][n][q] = p[TET5[t][n] - 1][q];
}
void inds2tets5(const int64_t p[8][3], int64_t tets[5][4][3]) {
    for (int t = 0; t < 5; ++t)
        for (int n = 0; n < 4; ++n)
            for (int q = 0; q < 3; ++q) tets[t][n][q] = p[TET5[t][n] - 1][q];
}

// computePLIC2VF (src/VF.jl:219-278). d2case snaps d in place here, so the cut-vertex
// interpolation uses the snapped distances.
double plic2vf(const Ctx& c, int64_t i, int64_t j, int64_t k, double nx, double ny, double nz,
               double dist) {
    const Mesh& m = c.m;
    double verts[8][3], tets[5][4][3];
    const double tetsign = cell2verts(c, i, j, k, verts);
    verts2tets5(verts, tets);
    double VF = 0.0;
    const double vol = (m.dx * m.dy) * m.dz;
    for (int t = 0; t < 5; ++t) {
        double vert[8][3];
        for (int n = 0; n < 4; ++n)
            for (int q = 0; q < 3; ++q) vert[n][q] = tets[t][n][q];
        double d[4];
        for (int n = 0; n < 4; ++n) d[n] = ((nx * vert[n][0] + ny * vert[n][1]) + nz * vert[n][2]) - dist;
        FL(24);
        const int cs = d2case_inplace(d);
        for (int n = 0; n < CUT_NVERT[cs]; ++n) {
            const int v1 = CUT_V1[cs][n] - 1, v2 = CUT_V2[cs][n] - 1;
            const double mu = jmin(1.0, jmax(0.0, sdiv(-d[v1], d[v2] - d[v1])));
            for (int q = 0; q < 3; ++q) vert[4 + n][q] = (1.0 - mu) * vert[v1][q] + mu * vert[v2][q];
            FL(12);
        }
        for (int n = CUT_NTETS[cs]; n >= CUT_NNTET[cs]; --n) {
            double tet[4][3];
            for (int tt = 0; tt < 4; ++tt)
                for (int q = 0; q < 3; ++q) tet[tt][q] = vert[CUT_VTET[cs][n - 1][tt] - 1][q];
            VF += (tetsign * tet_vol(tet)) / vol;
            FL(5);
        }
    }
    return VF;
}

// ------------------------------------------------------------------------------------
// ELVIRA (src/ELVIRA.jl:62-137)
// ------------------------------------------------------------------------------------
static inline double VFat(const Ctx& c, int64_t i, int64_t j, int64_t k) { return c.VF(i, j, k); }

static void check_normal(const Ctx& c, double& err, double nx_, double ny_, double nz_, int64_t i,
                         int64_t j, int64_t k, double* onx, double* ony, double* onz, int cand,
                         int* best, double* cand_err, double (*cand_n)[3]) {
    double err_ = 0.0;
    const double norm = std::sqrt((nx_ * nx_ + ny_ * ny_) + nz_ * nz_);
    nx_ = nx_ / norm; ny_ = ny_ / norm; nz_ = nz_ / norm;
    FL(8); FLS(1);
    const double dist = compute_dist(c, i, j, k, nx_, ny_, nz_, VFat(c, i, j, k));
    for (int64_t kk = k - 1; kk <= k + 1; ++kk)
        for (int64_t jj = j - 1; jj <= j + 1; ++jj)
            for (int64_t ii = i - 1; ii <= i + 1; ++ii) {
                const double VF_ = plic2vf(c, ii, jj, kk, nx_, ny_, nz_, dist);
                const double e = VFat(c, ii, jj, kk) - VF_;
                err_ += e * e;
                FL(3);
            }
    if (cand_err) cand_err[cand] = err_;
    if (cand_n) { cand_n[cand][0] = nx_; cand_n[cand][1] = ny_; cand_n[cand][2] = nz_; }
    if (err_ < err) {
        err = err_;
        *onx = nx_; *ony = ny_; *onz = nz_;
        if (best) *best = cand;
    }
}

void elvira_calc(const Ctx& c, int64_t i, int64_t j, int64_t k, double* onx, double* ony, double* onz,
                 int* best, double cand_err[27], double cand_n[27][3]) {
    const Mesh& m = c.m;
    double err = INFINITY;
    if (cand_err)
        for (int q = 0; q < 27; ++q) cand_err[q] = INFINITY;
    static const int PAIRS[3][2] = {{-1, 0}, {-1, 1}, {0, 1}};
    const double dx = m.dx, dy = m.dy, dz = m.dz;
    int cand = 0;
    if (m.Nz > 1) {  // X-Y stencils: outer jj, inner ii
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b, ++cand) {
                const int64_t i1 = i + PAIRS[b][0], i2 = i + PAIRS[b][1];
                const int64_t j1 = j + PAIRS[a][0], j2 = j + PAIRS[a][1];
                const double sx2 = (VFat(c, i2, j, k - 1) * dz + VFat(c, i2, j, k) * dz) + VFat(c, i2, j, k + 1) * dz;
                const double sx1 = (VFat(c, i1, j, k - 1) * dz + VFat(c, i1, j, k) * dz) + VFat(c, i1, j, k + 1) * dz;
                const double nx_ = -(sx2 - sx1) / (m.xm(i2) - m.xm(i1));
                const double sy2 = (VFat(c, i, j2, k - 1) * dz + VFat(c, i, j2, k) * dz) + VFat(c, i, j2, k + 1) * dz;
                const double sy1 = (VFat(c, i, j1, k - 1) * dz + VFat(c, i, j1, k) * dz) + VFat(c, i, j1, k + 1) * dz;
                const double ny_ = -(sy2 - sy1) / (m.ym(j2) - m.ym(j1));
                const double nz_ = VFat(c, i, j, k + 1) > VFat(c, i, j, k - 1) ? -1.0 : 1.0;
                FL(26);
                check_normal(c, err, nx_, ny_, nz_, i, j, k, onx, ony, onz, cand, best, cand_err, cand_n);
            }
    } else {
        cand += 9;
    }
    if (m.Ny > 1) {  // X-Z stencils: outer kk, inner ii
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b, ++cand) {
                const int64_t i1 = i + PAIRS[b][0], i2 = i + PAIRS[b][1];
                const int64_t k1 = k + PAIRS[a][0], k2 = k + PAIRS[a][1];
                const double sx2 = (VFat(c, i2, j - 1, k) * dy + VFat(c, i2, j, k) * dy) + VFat(c, i2, j + 1, k) * dy;
                const double sx1 = (VFat(c, i1, j - 1, k) * dy + VFat(c, i1, j, k) * dy) + VFat(c, i1, j + 1, k) * dy;
                const double nx_ = -(sx2 - sx1) / (m.xm(i2) - m.xm(i1));
                const double ny_ = VFat(c, i, j + 1, k) > VFat(c, i, j - 1, k) ? -1.0 : 1.0;
                const double sz2 = (VFat(c, i, j - 1, k2) * dy + VFat(c, i, j, k2) * dy) + VFat(c, i, j + 1, k2) * dy;
                const double sz1 = (VFat(c, i, j - 1, k1) * dy + VFat(c, i, j, k1) * dy) + VFat(c, i, j + 1, k1) * dy;
                const double nz_ = -(sz2 - sz1) / (m.zm(k2) - m.zm(k1));
                FL(26);
                check_normal(c, err, nx_, ny_, nz_, i, j, k, onx, ony, onz, cand, best, cand_err, cand_n);
            }
    } else {
        cand += 9;
    }
    if (m.Nx > 1) {  // Y-Z stencils: outer kk, inner jj
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b, ++cand) {
                const int64_t j1 = j + PAIRS[b][0], j2 = j + PAIRS[b][1];
                const int64_t k1 = k + PAIRS[a][0], k2 = k + PAIRS[a][1];
                const double nx_ = VFat(c, i + 1, j, k) > VFat(c, i - 1, j, k) ? -1.0 : 1.0;
                const double sy2 = (VFat(c, i - 1, j2, k) * dx + VFat(c, i, j2, k) * dx) + VFat(c, i + 1, j2, k) * dx;
                const double sy1 = (VFat(c, i - 1, j1, k) * dx + VFat(c, i, j1, k) * dx) + VFat(c, i + 1, j1, k) * dx;
                const double ny_ = -(sy2 - sy1) / (m.ym(j2) - m.ym(j1));
                const double sz2 = (VFat(c, i - 1, j, k2) * dx + VFat(c, i, j, k2) * dx) + VFat(c, i + 1, j, k2) * dx;
                const double sz1 = (VFat(c, i - 1, j, k1) * dx + VFat(c, i, j, k1) * dx) + VFat(c, i + 1, j, k1) * dx;
                const double nz_ = -(sz2 - sz1) / (m.zm(k2) - m.zm(k1));
                FL(26);
                check_normal(c, err, nx_, ny_, nz_, i, j, k, onx, ony, onz, cand, best, cand_err, cand_n);
            }
    }
}

// ------------------------------------------------------------------------------------
// Back-trace (src/Tools.jl:392-434, :516-641)
// ------------------------------------------------------------------------------------
static inline void find_cell(double p, int64_t& i, int64_t lo, int64_t hi,
                             double (Mesh::*f)(int64_t) const, const Mesh& m) {
    while (p - (m.*f)(i) < 0.0 && i > lo) i = i - 1;
    while (p - (m.*f)(i + 1) >= 0.0 && i + 1 < hi) i = i + 1;
}

static inline double trilin(const Arr3<double>& A, int64_t i, int64_t j, int64_t k, double wx1,
                            double wx2, double wy1, double wy2, double wz1, double wz2) {
    return wz1 * (wy1 * (wx1 * A(i + 1, j + 1, k + 1) + wx2 * A(i, j + 1, k + 1)) +
                  wy2 * (wx1 * A(i + 1, j, k + 1) + wx2 * A(i, j, k + 1))) +
           wz2 * (wy1 * (wx1 * A(i + 1, j + 1, k) + wx2 * A(i, j + 1, k)) +
                  wy2 * (wx1 * A(i + 1, j, k) + wx2 * A(i, j, k)));
}

double vel_uface(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k) {
    const Mesh& m = c.m;
    find_cell(pt[0], i, m.imino_, m.imaxo_, &Mesh::x, m);
    find_cell(pt[1], j, m.jmino_, m.jmaxo_, &Mesh::ym, m);
    find_cell(pt[2], k, m.kmino_, m.kmaxo_, &Mesh::zm, m);
    const double wx1 = sdiv(pt[0] - m.x(i), m.x(i + 1) - m.x(i)), wx2 = 1.0 - wx1;
    const double wy1 = sdiv(pt[1] - m.ym(j), m.ym(j + 1) - m.ym(j)), wy2 = 1.0 - wy1;
    const double wz1 = sdiv(pt[2] - m.zm(k), m.zm(k + 1) - m.zm(k)), wz2 = 1.0 - wz1;
    return trilin(c.uf, i, j, k, wx1, wx2, wy1, wy2, wz1, wz2);
}
double vel_vface(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k) {
    const Mesh& m = c.m;
    find_cell(pt[0], i, m.imino_, m.imaxo_, &Mesh::xm, m);
    find_cell(pt[1], j, m.jmino_, m.jmaxo_, &Mesh::y, m);
    find_cell(pt[2], k, m.kmino_, m.kmaxo_, &Mesh::zm, m);
    const double wx1 = sdiv(pt[0] - m.xm(i), m.xm(i + 1) - m.xm(i)), wx2 = 1.0 - wx1;
    const double wy1 = sdiv(pt[1] - m.y(j), m.y(j + 1) - m.y(j)), wy2 = 1.0 - wy1;
    const double wz1 = sdiv(pt[2] - m.zm(k), m.zm(k + 1) - m.zm(k)), wz2 = 1.0 - wz1;
    return trilin(c.vf, i, j, k, wx1, wx2, wy1, wy2, wz1, wz2);
}
double vel_wface(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k) {
    const Mesh& m = c.m;
    find_cell(pt[0], i, m.imino_, m.imaxo_, &Mesh::xm, m);
    find_cell(pt[1], j, m.jmino_, m.jmaxo_, &Mesh::ym, m);
    find_cell(pt[2], k, m.kmino_, m.kmaxo_, &Mesh::z, m);
    const double wx1 = sdiv(pt[0] - m.xm(i), m.xm(i + 1) - m.xm(i)), wx2 = 1.0 - wx1;
    const double wy1 = sdiv(pt[1] - m.ym(j), m.ym(j + 1) - m.ym(j)), wy2 = 1.0 - wy1;
    const double wz1 = sdiv(pt[2] - m.z(k), m.z(k + 1) - m.z(k)), wz2 = 1.0 - wz1;
    return trilin(c.wf, i, j, k, wx1, wx2, wy1, wy2, wz1, wz2);
}

static inline void vel_face(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k, double v[3]) {
    v[0] = vel_uface(c, pt, i, j, k);
    v[1] = vel_vface(c, pt, i, j, k);
    v[2] = vel_wface(c, pt, i, j, k);
}

// project! (src/Tools.jl:392-418)
void project(const Ctx& c, double pt[3], int64_t i, int64_t j, int64_t k, double dt) {
    if (c.p.test_case == MIST_CASE_ZALESAK) {
        const double pi2 = 2.0 * M_PI;
        const double x0 = pt[0], y0 = pt[1];
        pt[0] = (0.5 + std::cos(pi2 * dt) * (x0 - 0.5)) + std::sin(pi2 * dt) * (y0 - 0.5);
        pt[1] = (0.5 - std::sin(pi2 * dt) * (x0 - 0.5)) + std::cos(pi2 * dt) * (y0 - 0.5);
        return;
    }
    double v1[3], v2[3], v3[3], v4[3], q[3];
    if (c.p.projection_method == MIST_PROJ_RK4) {
        vel_face(c, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v1[n];
        vel_face(c, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v2[n];
        vel_face(c, q, i, j, k, v3);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - dt * v3[n];
        vel_face(c, q, i, j, k, v4);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + ((-dt) / 6.0) * (((v1[n] + 2.0 * v2[n]) + 2.0 * v3[n]) + v4[n]);
    } else if (c.p.projection_method == MIST_PROJ_EULER) {
        vel_face(c, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + (-dt) * v1[n];
    } else {
        vel_face(c, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v1[n];
        vel_face(c, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + (-dt) * v2[n];
    }
}

// pt2index! (src/Tools.jl:424-434); returns false if the walk leaves the mesh arrays.
bool pt2index(const Ctx& c, int64_t I[3], const double pt[3], int64_t i, int64_t j, int64_t k) {
    const Mesh& m = c.m;
    I[0] = i; I[1] = j; I[2] = k;
    while (pt[0] > m.x(I[0] + 1) + EPS) { I[0] += 1; if (I[0] + 1 > m.imaxo_ + 1) return false; }
    while (pt[0] < m.x(I[0]) - EPS) { I[0] -= 1; if (I[0] < m.imino_) return false; }
    while (pt[1] > m.y(I[1] + 1) + EPS) { I[1] += 1; if (I[1] + 1 > m.jmaxo_ + 1) return false; }
    while (pt[1] < m.y(I[1]) - EPS) { I[1] -= 1; if (I[1] < m.jmino_) return false; }
    while (pt[2] > m.z(I[2] + 1) + EPS) { I[2] += 1; if (I[2] + 1 > m.kmaxo_ + 1) return false; }
    while (pt[2] < m.z(I[2]) - EPS) { I[2] -= 1; if (I[2] < m.kmino_) return false; }
    return true;
}

// ------------------------------------------------------------------------------------
// cutTet (src/VFgeom.jl:713-900), recursive as in the reference. `tet` and `ind` are the
// caller's arrays and are overwritten with each child, as in the reference.
// ------------------------------------------------------------------------------------
struct CutScratch {
    static constexpr int NLEVEL = 100;  // nLevel (src/Transport.jl:23)
    double vert[NLEVEL][8][3];
    int64_t vind[NLEVEL][2][8][3];
};

CutScratch* new_cut_scratch() { return new CutScratch(); }
void delete_cut_scratch(CutScratch* s) { delete s; }

CutSums cut_tet(Ctx& c, double tet[4][3], int64_t ind[4][3], bool xdone, bool ydone, bool zdone,
                int lvl, CutScratch& s) {
    const Mesh& m = c.m;
    if (lvl >= CutScratch::NLEVEL) {
        c.status = MIST_ERR_OVERFLOW;
        return CutSums{};
    }
    int dir = -1;
    int64_t cut_ind = 0;
    double d[4];
    for (;;) {
        if (!xdone) {
            int64_t mx = std::max(std::max(ind[0][0], ind[1][0]), std::max(ind[2][0], ind[3][0]));
            int64_t mn = std::min(std::min(ind[0][0], ind[1][0]), std::min(ind[2][0], ind[3][0]));
            if (mx > mn) {
                dir = 0; cut_ind = mx;
                for (int n = 0; n < 4; ++n) d[n] = tet[n][0] - m.x(cut_ind);
                break;
            }
            xdone = true;
            continue;
        } else if (!ydone) {
            int64_t mx = std::max(std::max(ind[0][1], ind[1][1]), std::max(ind[2][1], ind[3][1]));
            int64_t mn = std::min(std::min(ind[0][1], ind[1][1]), std::min(ind[2][1], ind[3][1]));
            if (mx > mn) {
                dir = 1; cut_ind = mx;
                for (int n = 0; n < 4; ++n) d[n] = tet[n][1] - m.y(cut_ind);
                break;
            }
            ydone = true;
            continue;
        } else if (!zdone) {
            int64_t mx = std::max(std::max(ind[0][2], ind[1][2]), std::max(ind[2][2], ind[3][2]));
            int64_t mn = std::min(std::min(ind[0][2], ind[1][2]), std::min(ind[2][2], ind[3][2]));
            if (mx > mn) {
                dir = 2; cut_ind = mx;
                for (int n = 0; n < 4; ++n) d[n] = tet[n][2] - m.z(cut_ind);
                break;
            }
            zdone = true;
            continue;
        }
        break;
    }
    double(*vert)[3] = s.vert[lvl];
    if (dir >= 0) FL(4);  // plane distances of the chosen mesh cut
    if (dir < 0) {
        // Leaf: cut by the PLIC plane of cell ind[:,1] (src/VFgeom.jl:757-830).
        CutSums r;
        r.lvl = lvl;
        for (int n = 0; n < 4; ++n)
            for (int q = 0; q < 3; ++q) vert[n][q] = tet[n][q];
        const int64_t i = ind[0][0], j = ind[0][1], k = ind[0][2];
        const double nx = c.nx(i, j, k), ny = c.ny(i, j, k), nz = c.nz(i, j, k), D = c.D(i, j, k);
        for (int n = 0; n < 4; ++n) d[n] = ((nx * vert[n][0] + ny * vert[n][1]) + nz * vert[n][2]) - D;
        FL(20);
        const int cs = d2case_copy(d);
        for (int n = 0; n < CUT_NVERT[cs]; ++n) {
            const int v1 = CUT_V1[cs][n] - 1, v2 = CUT_V2[cs][n] - 1;
            const double mu = jmin(1.0, jmax(0.0, sdiv(-d[v1], d[v2] - d[v1])));
            for (int q = 0; q < 3; ++q) vert[4 + n][q] = (1.0 - mu) * vert[v1][q] + mu * vert[v2][q];
            FL(13);
        }
        // counted: vol and vLiq; the momentum sums (not needed by the VOF path) are not
        FL(2 * (CUT_NTETS[cs] - CUT_NNTET[cs] + 1) + (CUT_NNTET[cs] - 1));
        const double uu = c.u(i, j, k), vv = c.v(i, j, k), ww = c.w(i, j, k);
        const double rl = c.p.rho_liq, rg = c.p.rho_gas;
        for (int n = CUT_NTETS[cs]; n >= CUT_NNTET[cs]; --n) {
            double nt[4][3];
            for (int t = 0; t < 4; ++t)
                for (int q = 0; q < 3; ++q) nt[t][q] = vert[CUT_VTET[cs][n - 1][t] - 1][q];
            const double tv = tet_vol(nt);
            r.vol += tv; r.vLiq += tv;
            r.vU += (tv * uu) * rl; r.vV += (tv * vv) * rl; r.vW += (tv * ww) * rl; r.vrho += tv * rl;
        }
        for (int n = 1; n <= CUT_NNTET[cs] - 1; ++n) {
            double nt[4][3];
            for (int t = 0; t < 4; ++t)
                for (int q = 0; q < 3; ++q) nt[t][q] = vert[CUT_VTET[cs][n - 1][t] - 1][q];
            const double tv = tet_vol(nt);
            r.vol += tv;
            r.vU += (tv * uu) * rg; r.vV += (tv * vv) * rg; r.vW += (tv * ww) * rg; r.vrho += tv * rg;
        }
        return r;
    }
    // Cut by the mesh plane (src/VFgeom.jl:832-899).
    const int cs = d2case_copy(d);
    int64_t(*vi1)[3] = s.vind[lvl][0];
    int64_t(*vi2)[3] = s.vind[lvl][1];
    for (int n = 0; n < 4; ++n) {
        for (int q = 0; q < 3; ++q) {
            vert[n][q] = tet[n][q];
            vi1[n][q] = ind[n][q];
            vi2[n][q] = ind[n][q];
        }
        vi1[n][dir] = std::min(vi1[n][dir], cut_ind - 1);
        vi2[n][dir] = std::max(vi1[n][dir], cut_ind);
    }
    for (int n = 0; n < CUT_NVERT[cs]; ++n) {
        const int v1 = CUT_V1[cs][n] - 1, v2 = CUT_V2[cs][n] - 1;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-d[v1], d[v2] - d[v1])));
        for (int q = 0; q < 3; ++q) vert[4 + n][q] = (1.0 - mu) * vert[v1][q] + mu * vert[v2][q];
        FL(13);
        // the walk starts at the side-1 index of the vertex numbered like this cut vertex
        if (!pt2index(c, vi1[4 + n], vert[4 + n], vi1[n][0], vi1[n][1], vi1[n][2])) {
            c.status = MIST_ERR_OVERFLOW;
        }
        for (int q = 0; q < 3; ++q) {
            vi1[4 + n][q] = std::max(vi1[4 + n][q], std::min(vi1[v1][q], vi1[v2][q]));
            vi1[4 + n][q] = std::min(vi1[4 + n][q], std::max(vi1[v1][q], vi1[v2][q]));
        }
        for (int q = 0; q < 3; ++q) vi2[4 + n][q] = vi1[4 + n][q];
        vi1[4 + n][dir] = cut_ind - 1;
        vi2[4 + n][dir] = cut_ind;
    }
    CutSums r;
    int maxlvl = lvl;
    for (int n = 0; n < CUT_NTETS[cs]; ++n) {
        const int side = CUT_SIDE[cs][n];
        for (int nn = 0; nn < 4; ++nn) {
            const int vt = CUT_VTET[cs][n][nn] - 1;
            for (int q = 0; q < 3; ++q) {
                tet[nn][q] = vert[vt][q];
                ind[nn][q] = side == 1 ? vi1[vt][q] : vi2[vt][q];
            }
        }
        CutSums cr = cut_tet(c, tet, ind, xdone, ydone, zdone, lvl + 1, s);
        r.vol += cr.vol; r.vLiq += cr.vLiq; r.vU += cr.vU; r.vV += cr.vV; r.vW += cr.vW; r.vrho += cr.vrho;
        maxlvl = cr.lvl;
    }
    r.lvl = maxlvl;
    return r;
}

}  // namespace ora
