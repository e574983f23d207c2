// oracle_fwd.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): forward-mode derivatives of the
// SL-VOF step by dual numbers.
//
// The VOF part of transport! (src/Transport.jl:16-20, 108-153, 178-184, 259-260) and the
// computeBand! that follows it, restated on a dual-number scalar that carries, beside each
// value, its derivatives with respect to up to ND_FWD parameters (the amplitudes of the face
// velocity modes of the adjoint workload). Every value is computed with the same operations in
// the same order as the oracle's double-precision code (oracle_geom.cpp, oracle_stages.cpp), so
// the value parts equal the oracle's step bit for bit (checked by the tests); the derivative
// parts follow each operation's derivative. The branch decisions (the cut cases, the cell
// searches, ELVIRA's argmin, the band, the clamps) are taken on the values and held fixed,
// which is what a forward-mode AD of the reference's code does (ForwardDiff's dual numbers
// compare by value).
//
// ELVIRA's argmin comes from the oracle's own elvira_calc on the values; the kept candidate's
// normal is then evaluated on dual liquid fractions. Where another candidate gives a bitwise
// identical normal, the map has a kink there (the sign of a perturbation decides which one
// is kept), and the derivative is the mean of the two candidates' derivatives, the slope a
// central difference sees.
#include <algorithm>
#include <cstring>
#include <vector>

#include "oracle_core.h"
#include "oracle_fill.h"
#include "oracle_tables.h"

namespace ora {

constexpr int ND_FWD = 16;

struct Dual {
    double v = 0.0;
    double d[ND_FWD] = {};
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants have zero derivative
};

inline Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] + b.d[q];
    return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] - b.d[q];
    return r;
}
inline Dual operator-(const Dual& a) {
    Dual r(-a.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = -a.d[q];
    return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
    return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
    Dual r(a.v / b.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = (a.d[q] - r.v * b.d[q]) / b.v;
    return r;
}
inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
inline Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }

inline Dual dsqrt(const Dual& a) {
    Dual r(std::sqrt(a.v));
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] * (0.5 / r.v);
    return r;
}
inline Dual dcbrt(const Dual& a) {
    Dual r(std::cbrt(a.v));
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] / (3.0 * (r.v * r.v));
    return r;
}
inline Dual dacos(const Dual& a) {
    Dual r(std::acos(a.v));
    const double s = -1.0 / std::sqrt(1.0 - a.v * a.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] * s;
    return r;
}
inline Dual dsin(const Dual& a) {
    Dual r(std::sin(a.v));
    const double c = std::cos(a.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] * c;
    return r;
}
inline Dual dcos(const Dual& a) {
    Dual r(std::cos(a.v));
    const double s = -std::sin(a.v);
    for (int q = 0; q < ND_FWD; ++q) r.d[q] = a.d[q] * s;
    return r;
}
inline Dual dfabs(const Dual& a) { return std::signbit(a.v) ? -a : a; }
// Julia's max/min on the values; the operand kept carries its derivative
inline Dual djmax(const Dual& a, const Dual& b) {
    if (a.v != a.v) return a;
    if (b.v != b.v) return b;
    if (a.v == b.v) return std::signbit(a.v) ? b : a;
    return a.v > b.v ? a : b;
}
inline Dual djmin(const Dual& a, const Dual& b) {
    if (a.v != a.v) return a;
    if (b.v != b.v) return b;
    if (a.v == b.v) return std::signbit(a.v) ? a : b;
    return a.v < b.v ? a : b;
}
inline Dual dsdiv(const Dual& a, const Dual& b) { return std::fabs(b.v) <= EPS ? Dual(INFINITY) : a / b; }

namespace {

// tet_vol (src/VFgeom.jl:669-681)
Dual tet_vol_d(const Dual v[4][3]) {
    const Dual a1 = v[0][0] - v[3][0], a2 = v[0][1] - v[3][1], a3 = v[0][2] - v[3][2];
    const Dual b1 = v[1][0] - v[3][0], b2 = v[1][1] - v[3][1], b3 = v[1][2] - v[3][2];
    const Dual c1 = v[2][0] - v[3][0], c2 = v[2][1] - v[3][1], c3 = v[2][2] - v[3][2];
    return Dual(-1.0 / 6.0) * ((a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3)) + a3 * (b1 * c2 - c1 * b2));
}

int d2case_of(const Dual d[4]) {
    const double t[4] = {d[0].v, d[1].v, d[2].v, d[3].v};
    return d2case_copy(t);
}

// computeAlpha (src/VF.jl:164-214)
Dual compute_alpha_d(const Dual& a, const Dual& b, const Dual& cc, const Dual& VF) {
    Dual m[3] = {a, b, cc};
    for (int p = 1; p < 3; ++p)  // ascending by value (ties keep their order)
        for (int q = p; q > 0 && m[q].v < m[q - 1].v; --q) std::swap(m[q], m[q - 1]);
    const Dual m1 = m[0], m2 = m[1], m3 = m[2];
    const Dual m12 = m1 + m2;
    const Dual V1 = (m1 * m1) / djmax((Dual(6.0) * m2) * m3, Dual(EPS));
    const Dual V2 = V1 + dsdiv(Dual(0.5) * (m2 - m1), m3);
    Dual V3;
    if (m12.v <= m3.v) {
        V3 = dsdiv(Dual(0.5) * m12, m3);
    } else {
        V3 = dsdiv(((m3 * m3) * (Dual(3.0) * m12 - m3) + (m1 * m1) * (m1 - Dual(3.0) * m3)) +
                       (m2 * m2) * (m2 - Dual(3.0) * m3),
                   ((Dual(6.0) * m1) * m2) * m3);
    }
    Dual alpha;
    if (0.0 <= VF.v && VF.v < V1.v) {
        alpha = dcbrt((((Dual(6.0) * m1) * m2) * m3) * VF);
    } else if (V1.v <= VF.v && VF.v < V2.v) {
        alpha = Dual(0.5) * (m1 + dsqrt((m1 * m1) + ((Dual(8.0) * m2) * m3) * (VF - V1)));
    } else if (V2.v <= VF.v && VF.v < V3.v) {
        Dual a0 = ((m1 * m1) * m1 + (m2 * m2) * m2) - (((Dual(6.0) * m1) * m2) * m3) * VF;
        Dual a1 = Dual(-3.0) * ((m1 * m1) + (m2 * m2));
        Dual a2 = Dual(3.0) * m12;
        Dual a3 = Dual(-1.0);
        a0 = a0 / a3; a1 = a1 / a3; a2 = a2 / a3; a3 = a3 / a3;
        const Dual p0 = a1 / Dual(3.0) - (a2 * a2) / Dual(9.0);
        const Dual q0 = (a1 * a2 - Dual(3.0) * a0) / Dual(6.0) - ((a2 * a2) * a2) / Dual(27.0);
        const Dual theta =
            dacos(djmin(Dual(1.0), djmax(Dual(-1.0), q0 / dsqrt(Dual(-1.0) * ((p0 * p0) * p0))))) / Dual(3.0);
        alpha = dsqrt(-p0) * (Dual(std::sqrt(3.0)) * dsin(theta) - dcos(theta)) - a2 / Dual(3.0);
    } else {
        if (m12.v <= m3.v) {
            alpha = m3 * VF + Dual(0.5) * m12;
        } else {
            Dual a0 = (((m1 * m1) * m1 + (m2 * m2) * m2) + (m3 * m3) * m3) - (((Dual(6.0) * m1) * m2) * m3) * VF;
            Dual a1 = Dual(-3.0) * (((m1 * m1) + (m2 * m2)) + (m3 * m3));
            Dual a2 = Dual(3.0);
            Dual a3 = Dual(-2.0);
            a0 = a0 / a3; a1 = a1 / a3; a2 = a2 / a3; a3 = a3 / a3;
            const Dual p0 = a1 / Dual(3.0) - (a2 * a2) / Dual(9.0);
            const Dual q0 = (a1 * a2 - Dual(3.0) * a0) / Dual(6.0) - ((a2 * a2) * a2) / Dual(27.0);
            const Dual theta = dacos(djmin(Dual(1.0), djmax(Dual(-1.0), q0 / dsqrt(-((p0 * p0) * p0))))) / Dual(3.0);
            alpha = dsqrt(-p0) * (Dual(std::sqrt(3.0)) * dsin(theta) - dcos(theta)) - a2 / Dual(3.0);
        }
    }
    return alpha;
}

// The dual arrays of the step beside the oracle context that holds the values for the
// branch decisions (ELVIRA's argmin, computeBand!).
struct Fwd {
    Ctx& c;
    Arr3<Dual> VF, nx, ny, nz, D, uf, vf, wf;
    int status = 0;
    explicit Fwd(Ctx& cc) : c(cc) {
        VF.init(c.VF.e); nx.init(c.nx.e); ny.init(c.ny.e); nz.init(c.nz.e); D.init(c.D.e);
        uf.init(c.uf.e); vf.init(c.vf.e); wf.init(c.wf.e);
    }
    void fail(int st) {
#pragma omp atomic write
        status = st;
    }
    void sync_values() {
        for (size_t q = 0; q < VF.a.size(); ++q) c.VF.a[q] = VF.a[q].v;
    }
};

// computeDist (src/VF.jl:109-156)
Dual compute_dist_d(const Fwd& f, int64_t i, int64_t j, int64_t k, const Dual& nx, const Dual& ny, const Dual& nz,
                    const Dual& VF) {
    const Mesh& m = f.c.m;
    if (VF.v < f.c.p.VFlo) return Dual(-((m.Lx + m.Ly) + m.Lz));
    if (VF.v > f.c.p.VFhi) return Dual(+((m.Lx + m.Ly) + m.Lz));
    Dual mm0 = nx * Dual(m.dx), mm1 = ny * Dual(m.dy), mm2 = nz * Dual(m.dz);
    const Dual norm = ((dfabs(mm0) + dfabs(mm1)) + dfabs(mm2)) + Dual(EPS);
    mm0 = mm0 / norm; mm1 = mm1 / norm; mm2 = mm2 / norm;
    Dual factor(0.0);
    if (mm0.v < 0.0) factor += mm0;
    if (mm1.v < 0.0) factor += mm1;
    if (mm2.v < 0.0) factor += mm2;
    const Dual VFo = VF.v > 0.5 ? Dual(1.0) - VF : VF;
    Dual alpha = compute_alpha_d(dfabs(mm0), dfabs(mm1), dfabs(mm2), VFo);
    alpha = VF.v > 0.5 ? Dual(1.0) - alpha : alpha;
    alpha += factor;
    alpha -= Dual(0.5) * ((mm0 + mm1) + mm2);
    Dual dist = alpha * norm;
    dist += (nx * Dual(m.xm(i)) + ny * Dual(m.ym(j))) + nz * Dual(m.zm(k));
    return dist;
}

// The normal of ELVIRA candidate `cand` (src/ELVIRA.jl:62-100, the candidate numbering of
// elvira_calc), normalised as check_normal! does.
void candidate_normal(const Fwd& f, int64_t i, int64_t j, int64_t k, int cand, Dual n[3]) {
    static const int PAIRS[3][2] = {{-1, 0}, {-1, 1}, {0, 1}};
    const Mesh& m = f.c.m;
    const int fam = cand / 9, a = (cand % 9) / 3, b = cand % 3;
    const Dual dx(m.dx), dy(m.dy), dz(m.dz);
    auto V = [&](int64_t ii, int64_t jj, int64_t kk) -> const Dual& { return f.VF(ii, jj, kk); };
    Dual nx_, ny_, nz_;
    if (fam == 0) {
        const int64_t i1 = i + PAIRS[b][0], i2 = i + PAIRS[b][1], j1 = j + PAIRS[a][0], j2 = j + PAIRS[a][1];
        const Dual sx2 = (V(i2, j, k - 1) * dz + V(i2, j, k) * dz) + V(i2, j, k + 1) * dz;
        const Dual sx1 = (V(i1, j, k - 1) * dz + V(i1, j, k) * dz) + V(i1, j, k + 1) * dz;
        nx_ = -(sx2 - sx1) / Dual(m.xm(i2) - m.xm(i1));
        const Dual sy2 = (V(i, j2, k - 1) * dz + V(i, j2, k) * dz) + V(i, j2, k + 1) * dz;
        const Dual sy1 = (V(i, j1, k - 1) * dz + V(i, j1, k) * dz) + V(i, j1, k + 1) * dz;
        ny_ = -(sy2 - sy1) / Dual(m.ym(j2) - m.ym(j1));
        nz_ = Dual(V(i, j, k + 1).v > V(i, j, k - 1).v ? -1.0 : 1.0);
    } else if (fam == 1) {
        const int64_t i1 = i + PAIRS[b][0], i2 = i + PAIRS[b][1], k1 = k + PAIRS[a][0], k2 = k + PAIRS[a][1];
        const Dual sx2 = (V(i2, j - 1, k) * dy + V(i2, j, k) * dy) + V(i2, j + 1, k) * dy;
        const Dual sx1 = (V(i1, j - 1, k) * dy + V(i1, j, k) * dy) + V(i1, j + 1, k) * dy;
        nx_ = -(sx2 - sx1) / Dual(m.xm(i2) - m.xm(i1));
        ny_ = Dual(V(i, j + 1, k).v > V(i, j - 1, k).v ? -1.0 : 1.0);
        const Dual sz2 = (V(i, j - 1, k2) * dy + V(i, j, k2) * dy) + V(i, j + 1, k2) * dy;
        const Dual sz1 = (V(i, j - 1, k1) * dy + V(i, j, k1) * dy) + V(i, j + 1, k1) * dy;
        nz_ = -(sz2 - sz1) / Dual(m.zm(k2) - m.zm(k1));
    } else {
        const int64_t j1 = j + PAIRS[b][0], j2 = j + PAIRS[b][1], k1 = k + PAIRS[a][0], k2 = k + PAIRS[a][1];
        nx_ = Dual(V(i + 1, j, k).v > V(i - 1, j, k).v ? -1.0 : 1.0);
        const Dual sy2 = (V(i - 1, j2, k) * dx + V(i, j2, k) * dx) + V(i + 1, j2, k) * dx;
        const Dual sy1 = (V(i - 1, j1, k) * dx + V(i, j1, k) * dx) + V(i + 1, j1, k) * dx;
        ny_ = -(sy2 - sy1) / Dual(m.ym(j2) - m.ym(j1));
        const Dual sz2 = (V(i - 1, j, k2) * dx + V(i, j, k2) * dx) + V(i + 1, j, k2) * dx;
        const Dual sz1 = (V(i - 1, j, k1) * dx + V(i, j, k1) * dx) + V(i + 1, j, k1) * dx;
        nz_ = -(sz2 - sz1) / Dual(m.zm(k2) - m.zm(k1));
    }
    const Dual norm = dsqrt((nx_ * nx_ + ny_ * ny_) + nz_ * nz_);
    n[0] = nx_ / norm;
    n[1] = ny_ / norm;
    n[2] = nz_ / norm;
}

// computeNormal! with ELVIRA (oracle_stages.cpp compute_normal on dual liquid fractions)
void compute_normal_d(Fwd& f) {
    Ctx& c = f.c;
    const Mesh& m = c.m;
    f.sync_values();
    std::fill(f.nx.a.begin(), f.nx.a.end(), Dual(0.0));
    std::fill(f.ny.a.begin(), f.ny.a.end(), Dual(0.0));
    std::fill(f.nz.a.begin(), f.nz.a.end(), Dual(0.0));
    const double VFlo = c.p.VFlo, VFhi = c.p.VFhi;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                const double vf = c.VF(i, j, k);
                if (vf >= VFlo && vf <= VFhi) {
                    double a = 0, b = 0, d = 0, err[27], cn[27][3];
                    int best = -1;
                    elvira_calc(c, i, j, k, &a, &b, &d, &best, err, cn);
                    if (best < 0) continue;
                    Dual n[3];
                    candidate_normal(f, i, j, k, best, n);
                    for (int q = 0; q < 27; ++q) {  // a candidate with the same normal: a kink
                        if (q == best || !(err[q] < INFINITY)) continue;
                        if (cn[q][0] == cn[best][0] && cn[q][1] == cn[best][1] && cn[q][2] == cn[best][2]) {
                            Dual t[3];
                            candidate_normal(f, i, j, k, q, t);
                            for (int r = 0; r < 3; ++r)
                                for (int p = 0; p < ND_FWD; ++p) n[r].d[p] = 0.5 * (n[r].d[p] + t[r].d[p]);
                            break;
                        }
                    }
                    f.nx(i, j, k) = n[0]; f.ny(i, j, k) = n[1]; f.nz(i, j, k) = n[2];
                } else if (vf < VFlo) {
                    if (c.VF(i - 1, j, k) > VFhi) f.nx(i, j, k) = Dual(1);
                    else if (c.VF(i + 1, j, k) > VFhi) f.nx(i, j, k) = Dual(-1);
                    else if (c.VF(i, j - 1, k) > VFhi) f.ny(i, j, k) = Dual(1);
                    else if (c.VF(i, j + 1, k) > VFhi) f.ny(i, j, k) = Dual(-1);
                    else if (c.VF(i, j, k - 1) > VFhi) f.nz(i, j, k) = Dual(1);
                    else if (c.VF(i, j, k + 1) > VFhi) f.nz(i, j, k) = Dual(-1);
                } else if (vf > VFhi) {
                    if (c.VF(i - 1, j, k) < VFlo) f.nx(i, j, k) = Dual(-1);
                    else if (c.VF(i + 1, j, k) < VFlo) f.nx(i, j, k) = Dual(1);
                    else if (c.VF(i, j - 1, k) < VFlo) f.ny(i, j, k) = Dual(-1);
                    else if (c.VF(i, j + 1, k) < VFlo) f.ny(i, j, k) = Dual(1);
                    else if (c.VF(i, j, k - 1) < VFlo) f.nz(i, j, k) = Dual(-1);
                    else if (c.VF(i, j, k + 1) < VFlo) f.nz(i, j, k) = Dual(1);
                }
            }
    borders_t(c, f.nx);
    borders_t(c, f.ny);
    borders_t(c, f.nz);
}

// computePLIC! (src/VF.jl:94-103): every cell including ghosts, then borders
void compute_plic_d(Fwd& f) {
    const Mesh& m = f.c.m;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i)
                f.D(i, j, k) = compute_dist_d(f, i, j, k, f.nx(i, j, k), f.ny(i, j, k), f.nz(i, j, k), f.VF(i, j, k));
    borders_t(f.c, f.D);
}

// get_velocity_{u,v,w}face (src/Tools.jl:516-641) on dual faces and positions
void find_cell_v(double p, int64_t& i, int64_t lo, int64_t hi, double (Mesh::*fc)(int64_t) const, const Mesh& m) {
    while (p - (m.*fc)(i) < 0.0 && i > lo) i = i - 1;
    while (p - (m.*fc)(i + 1) >= 0.0 && i + 1 < hi) i = i + 1;
}

Dual trilin_d(const Arr3<Dual>& A, int64_t i, int64_t j, int64_t k, const Dual& wx1, const Dual& wx2, const Dual& wy1,
              const Dual& wy2, const Dual& wz1, const Dual& wz2) {
    return wz1 * (wy1 * (wx1 * A(i + 1, j + 1, k + 1) + wx2 * A(i, j + 1, k + 1)) +
                  wy2 * (wx1 * A(i + 1, j, k + 1) + wx2 * A(i, j, k + 1))) +
           wz2 * (wy1 * (wx1 * A(i + 1, j + 1, k) + wx2 * A(i, j + 1, k)) +
                  wy2 * (wx1 * A(i + 1, j, k) + wx2 * A(i, j, k)));
}

Dual vel_comp(const Fwd& f, const Arr3<Dual>& A, int axis, const Dual pt[3], int64_t i, int64_t j, int64_t k) {
    const Mesh& m = f.c.m;
    using Fn = double (Mesh::*)(int64_t) const;
    const Fn fx = axis == 0 ? &Mesh::x : &Mesh::xm, fy = axis == 1 ? &Mesh::y : &Mesh::ym,
             fz = axis == 2 ? &Mesh::z : &Mesh::zm;
    find_cell_v(pt[0].v, i, m.imino_, m.imaxo_, fx, m);
    find_cell_v(pt[1].v, j, m.jmino_, m.jmaxo_, fy, m);
    find_cell_v(pt[2].v, k, m.kmino_, m.kmaxo_, fz, m);
    const Dual wx1 = dsdiv(pt[0] - Dual((m.*fx)(i)), Dual((m.*fx)(i + 1) - (m.*fx)(i))), wx2 = Dual(1.0) - wx1;
    const Dual wy1 = dsdiv(pt[1] - Dual((m.*fy)(j)), Dual((m.*fy)(j + 1) - (m.*fy)(j))), wy2 = Dual(1.0) - wy1;
    const Dual wz1 = dsdiv(pt[2] - Dual((m.*fz)(k)), Dual((m.*fz)(k + 1) - (m.*fz)(k))), wz2 = Dual(1.0) - wz1;
    return trilin_d(A, i, j, k, wx1, wx2, wy1, wy2, wz1, wz2);
}

void vel_face_d(const Fwd& f, const Dual pt[3], int64_t i, int64_t j, int64_t k, Dual v[3]) {
    v[0] = vel_comp(f, f.uf, 0, pt, i, j, k);
    v[1] = vel_comp(f, f.vf, 1, pt, i, j, k);
    v[2] = vel_comp(f, f.wf, 2, pt, i, j, k);
}

// project! (src/Tools.jl:392-418), RK4, Euler or midpoint
void project_d(const Fwd& f, Dual pt[3], int64_t i, int64_t j, int64_t k, double dt) {
    Dual v1[3], v2[3], v3[3], v4[3], q[3];
    const Dual hdt(0.5 * dt), fdt(dt);
    if (f.c.p.projection_method == MIST_PROJ_RK4) {
        vel_face_d(f, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - hdt * v1[n];
        vel_face_d(f, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - hdt * v2[n];
        vel_face_d(f, q, i, j, k, v3);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - fdt * v3[n];
        vel_face_d(f, q, i, j, k, v4);
        for (int n = 0; n < 3; ++n)
            pt[n] = pt[n] + Dual((-dt) / 6.0) * (((v1[n] + Dual(2.0) * v2[n]) + Dual(2.0) * v3[n]) + v4[n]);
    } else if (f.c.p.projection_method == MIST_PROJ_EULER) {
        vel_face_d(f, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + Dual(-dt) * v1[n];
    } else {
        vel_face_d(f, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - hdt * v1[n];
        vel_face_d(f, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + Dual(-dt) * v2[n];
    }
}

bool pt2index_d(const Fwd& f, int64_t I[3], const Dual pt[3], int64_t i, int64_t j, int64_t k) {
    const double p[3] = {pt[0].v, pt[1].v, pt[2].v};
    return pt2index(f.c, I, p, i, j, k);
}

struct SumsD {
    Dual vol, vLiq;
};

// cutTet's per-level vertices (NLEVEL x 8 x 3) and indices (NLEVEL x 2 x 8 x 3), per thread
struct ScratchD {
    std::vector<Dual> vert = std::vector<Dual>(100 * 8 * 3);
    std::vector<int64_t> vind = std::vector<int64_t>(100 * 2 * 8 * 3);
};

// cutTet (src/VFgeom.jl:713-900) on dual vertices; the structure of oracle_geom.cpp's cut_tet
SumsD cut_tet_d(Fwd& f, ScratchD& s, Dual tet[4][3], int64_t ind[4][3], bool xdone, bool ydone, bool zdone,
                int lvl) {
    const Mesh& m = f.c.m;
    if (lvl >= 100) {
        f.fail(MIST_ERR_OVERFLOW);
        return SumsD{};
    }
    int dir = -1;
    int64_t cut_ind = 0;
    Dual d[4];
    for (;;) {
        if (!xdone) {
            const int64_t mx = std::max(std::max(ind[0][0], ind[1][0]), std::max(ind[2][0], ind[3][0]));
            const int64_t mn = std::min(std::min(ind[0][0], ind[1][0]), std::min(ind[2][0], ind[3][0]));
            if (mx > mn) {
                dir = 0; cut_ind = mx;
                for (int n = 0; n < 4; ++n) d[n] = tet[n][0] - Dual(m.x(cut_ind));
                break;
            }
            xdone = true;
            continue;
        } else if (!ydone) {
            const int64_t mx = std::max(std::max(ind[0][1], ind[1][1]), std::max(ind[2][1], ind[3][1]));
            const int64_t mn = std::min(std::min(ind[0][1], ind[1][1]), std::min(ind[2][1], ind[3][1]));
            if (mx > mn) {
                dir = 1; cut_ind = mx;
                for (int n = 0; n < 4; ++n) d[n] = tet[n][1] - Dual(m.y(cut_ind));
                break;
            }
            ydone = true;
            continue;
        } else if (!zdone) {
            const int64_t mx = std::max(std::max(ind[0][2], ind[1][2]), std::max(ind[2][2], ind[3][2]));
            const int64_t mn = std::min(std::min(ind[0][2], ind[1][2]), std::min(ind[2][2], ind[3][2]));
            if (mx > mn) {
                dir = 2; cut_ind = mx;
                for (int n = 0; n < 4; ++n) d[n] = tet[n][2] - Dual(m.z(cut_ind));
                break;
            }
            zdone = true;
            continue;
        }
        break;
    }
    Dual(*vert)[3] = reinterpret_cast<Dual(*)[3]>(&s.vert[(size_t)lvl * 24]);
    if (dir < 0) {  // leaf: the PLIC plane of cell ind[:,1] (src/VFgeom.jl:757-830)
        SumsD r;
        for (int n = 0; n < 4; ++n)
            for (int q = 0; q < 3; ++q) vert[n][q] = tet[n][q];
        const int64_t i = ind[0][0], j = ind[0][1], k = ind[0][2];
        const Dual nx = f.nx(i, j, k), ny = f.ny(i, j, k), nz = f.nz(i, j, k), D = f.D(i, j, k);
        for (int n = 0; n < 4; ++n) d[n] = ((nx * vert[n][0] + ny * vert[n][1]) + nz * vert[n][2]) - D;
        const int cs = d2case_of(d);
        for (int n = 0; n < CUT_NVERT[cs]; ++n) {
            const int v1 = CUT_V1[cs][n] - 1, v2 = CUT_V2[cs][n] - 1;
            const Dual mu = djmin(Dual(1.0), djmax(Dual(0.0), dsdiv(-d[v1], d[v2] - d[v1])));
            for (int q = 0; q < 3; ++q) vert[4 + n][q] = (Dual(1.0) - mu) * vert[v1][q] + mu * vert[v2][q];
        }
        for (int n = CUT_NTETS[cs]; n >= CUT_NNTET[cs]; --n) {
            Dual nt[4][3];
            for (int t = 0; t < 4; ++t)
                for (int q = 0; q < 3; ++q) nt[t][q] = vert[CUT_VTET[cs][n - 1][t] - 1][q];
            const Dual tv = tet_vol_d(nt);
            r.vol += tv;
            r.vLiq += tv;
        }
        for (int n = 1; n <= CUT_NNTET[cs] - 1; ++n) {
            Dual nt[4][3];
            for (int t = 0; t < 4; ++t)
                for (int q = 0; q < 3; ++q) nt[t][q] = vert[CUT_VTET[cs][n - 1][t] - 1][q];
            r.vol += tet_vol_d(nt);
        }
        return r;
    }
    // cut by the mesh plane (src/VFgeom.jl:832-899)
    const int cs = d2case_of(d);
    int64_t(*vi1)[3] = reinterpret_cast<int64_t(*)[3]>(&s.vind[(size_t)lvl * 48]);
    int64_t(*vi2)[3] = reinterpret_cast<int64_t(*)[3]>(&s.vind[(size_t)lvl * 48 + 24]);
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
        const Dual mu = djmin(Dual(1.0), djmax(Dual(0.0), dsdiv(-d[v1], d[v2] - d[v1])));
        for (int q = 0; q < 3; ++q) vert[4 + n][q] = (Dual(1.0) - mu) * vert[v1][q] + mu * vert[v2][q];
        if (!pt2index_d(f, vi1[4 + n], vert[4 + n], vi1[n][0], vi1[n][1], vi1[n][2])) f.fail(MIST_ERR_OVERFLOW);
        for (int q = 0; q < 3; ++q) {
            vi1[4 + n][q] = std::max(vi1[4 + n][q], std::min(vi1[v1][q], vi1[v2][q]));
            vi1[4 + n][q] = std::min(vi1[4 + n][q], std::max(vi1[v1][q], vi1[v2][q]));
        }
        for (int q = 0; q < 3; ++q) vi2[4 + n][q] = vi1[4 + n][q];
        vi1[4 + n][dir] = cut_ind - 1;
        vi2[4 + n][dir] = cut_ind;
    }
    SumsD r;
    for (int n = 0; n < CUT_NTETS[cs]; ++n) {
        const int side = CUT_SIDE[cs][n];
        for (int nn = 0; nn < 4; ++nn) {
            const int vt = CUT_VTET[cs][n][nn] - 1;
            for (int q = 0; q < 3; ++q) {
                tet[nn][q] = vert[vt][q];
                ind[nn][q] = side == 1 ? vi1[vt][q] : vi2[vt][q];
            }
        }
        const SumsD cr = cut_tet_d(f, s, tet, ind, xdone, ydone, zdone, lvl + 1);
        r.vol += cr.vol;
        r.vLiq += cr.vLiq;
    }
    return r;
}

// transport!, VOF part (oracle_stages.cpp transport_vof), then computeBand! on the values
int vof_step_d(Fwd& f, double dt) {
    Ctx& c = f.c;
    const Mesh& m = c.m;
    compute_normal_d(f);
    f.status = 0;
    Arr3<Dual> VFnew;
    VFnew.init(m.cell_ext());
#pragma omp parallel
    {
    ScratchD s;
#pragma omp for collapse(2) schedule(dynamic, 1)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                if (std::abs((int)c.band(i, j, k)) <= c.p.nband) {
                    double vd[8][3];
                    const double tetsign = cell2verts(c, i, j, k, vd);
                    Dual verts[8][3], tets[5][4][3];
                    int64_t vInds[8][3], inds[5][4][3];
                    for (int n = 0; n < 8; ++n) {
                        for (int q = 0; q < 3; ++q) verts[n][q] = Dual(vd[n][q]);
                        project_d(f, verts[n], i, j, k, dt);
                    }
                    for (int t = 0; t < 5; ++t)
                        for (int n = 0; n < 4; ++n)
                            for (int q = 0; q < 3; ++q) tets[t][n][q] = verts[TET5[t][n] - 1][q];
                    for (int n = 0; n < 8; ++n)
                        if (!pt2index_d(f, vInds[n], verts[n], i, j, k)) f.fail(MIST_ERR_OVERFLOW);
                    inds2tets5(vInds, inds);
                    Dual vol(0.0), vLiq(0.0);
                    for (int t = 0; t < 5; ++t) {
                        const SumsD r = cut_tet_d(f, s, tets[t], inds[t], false, false, false, 1);
                        vol += Dual(tetsign) * r.vol;
                        vLiq += Dual(tetsign) * r.vLiq;
                    }
                    VFnew(i, j, k) = vLiq / vol;
                } else {
                    VFnew(i, j, k) = f.VF(i, j, k);
                }
            }
    }
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) f.VF(i, j, k) = VFnew(i, j, k);
    neumann_t(m, f.VF);
    borders_t(c, f.VF);
    compute_normal_d(f);
    compute_plic_d(f);
    f.sync_values();
    compute_band(c);
    return f.status;
}

}  // namespace

}  // namespace ora

using namespace ora;

extern "C" {

// Forward-mode run of `nsteps` VOF steps (transport! + computeBand!) of fixed dt from the VF in
// the oracle context h (VF extents, ghosts as given): the faces are uf, vf, wf (cell extents)
// with derivatives Bu[q], Bv[q] (q < nmodes <= 16; wf has none). Runs the initialisation of
// run_solver (Neumann!, borders, computeBand!, ELVIRA, computePLIC!) first. Writes VF (VF
// extents) into vf_out and its nmodes derivatives into dvf_out[q * n + ...]. The oracle
// context's VF, band and PLIC arrays hold the values afterwards.
int ora_fwd_vof_run(void* h, const double* uf, const double* vf, const double* wf, const double* Bu, const double* Bv,
                    int nmodes, int nsteps, double dt, double* vf_out, double* dvf_out) {
    Ctx& c = *(Ctx*)h;
    if (nmodes < 0 || nmodes > ND_FWD || c.p.solveNS || c.p.test_case == MIST_CASE_ZALESAK) return MIST_ERR_PARAM;
    Fwd f(c);
    const size_t nface = c.uf.a.size();
    for (size_t q = 0; q < nface; ++q) {
        f.uf.a[q] = Dual(uf[q]);
        f.vf.a[q] = Dual(vf[q]);
        f.wf.a[q] = Dual(wf[q]);
        for (int p = 0; p < nmodes; ++p) {
            f.uf.a[q].d[p] = Bu[(size_t)p * nface + q];
            f.vf.a[q].d[p] = Bv[(size_t)p * nface + q];
        }
    }
    for (size_t q = 0; q < c.VF.a.size(); ++q) f.VF.a[q] = Dual(c.VF.a[q]);
    neumann_t(c.m, f.VF);
    borders_t(c, f.VF);
    f.sync_values();
    compute_band(c);
    compute_normal_d(f);
    compute_plic_d(f);
    int st = 0;
    for (int s = 0; s < nsteps && !st; ++s) st = vof_step_d(f, dt);
    for (size_t q = 0; q < c.D.a.size(); ++q) c.D.a[q] = f.D.a[q].v;
    const size_t nvf = f.VF.a.size();
    for (size_t q = 0; q < nvf; ++q) {
        vf_out[q] = f.VF.a[q].v;
        for (int p = 0; p < nmodes; ++p) dvf_out[(size_t)p * nvf + q] = f.VF.a[q].d[p];
    }
    return st;
}

}  // extern "C"
