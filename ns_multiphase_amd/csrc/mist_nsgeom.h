// mist_nsgeom.h -- geometry of the momentum predictor shared by mist_ns.hip and the cut trees
// of mist_vof.hip: SLfluxVol, the apex of the flux-correction pyramids and the 17 tets of a
// band cell under the finite-difference scheme.
#pragma once

#include "mist_kernels.h"

namespace mist {

// ------------------------------------------------------------------------------------
// SLfluxVol (src/VFgeom.jl:428-509): the volume swept through face `dir` (1, 2, 3) of cell
// (i,j,k) in one step, 5 tets over its 4 back-traced and 4 fixed corners. The back-traced
// corners are the projected vertices pv of the band cells. `verts` receives the 8 vertices.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ bool in_band(const int16_t* band, int64_t c, int nband) { return abs((int)band[c]) <= nband; }

// Back-traced vertices read from the array of all grid vertices (3 doubles per vertex).
struct VertPtr {
    const DevMesh* m;
    const double* pv;
    __device__ __forceinline__ double at(int I, int J, int K, int c) const { return pv[3 * vtx_id(*m, I, J, K) + c]; }
};

template <class PV>
__device__ inline double sl_flux_vol_acc(const DevMesh& m, const PV& pv, int dir, int i, int j, int k,
                                         double verts[8][3]) {
    const bool even = ((i + j + k) % 2) == 0;
    const int odd = even ? 0 : 1;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        int I, J, K;
        bool proj;
        if (dir == 1) {  // pairs over (y, z) corners: projected first for even faces
            const int n = s >> 1;
            I = i; J = j + (n & 1); K = k + (n >> 1);
            proj = ((s & 1) == 0) == even;
        } else if (dir == 2) {
            I = i + ((s & 1) ^ odd); J = j; K = k + (s >> 2);
            proj = (s & 2) == 0;
        } else {
            I = i + ((s & 1) ^ odd); J = j + ((s >> 1) & 1); K = k;
            proj = s < 4;
        }
        if (proj) {
            verts[s][0] = pv.at(I, J, K, 0); verts[s][1] = pv.at(I, J, K, 1); verts[s][2] = pv.at(I, J, K, 2);
        } else {
            verts[s][0] = m.x[I]; verts[s][1] = m.y[J]; verts[s][2] = m.z[K];
        }
    }
    double v = 0.0;  // tets_vol (src/VFgeom.jl:683)
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const unsigned q = TET5[t];
        v += tet_vol(verts[nib(q, 0)], verts[nib(q, 1)], verts[nib(q, 2)], verts[nib(q, 3)]);
    }
    return v * (even ? 1.0 : -1.0);
}

__device__ inline double sl_flux_vol(const DevMesh& m, const double* pv, int dir, int i, int j, int k,
                                     double verts[8][3]) {
    return sl_flux_vol_acc(m, VertPtr{&m, pv}, dir, i, j, k, verts);
}

// divg_cell (src/Velocity.jl:79-122) of interior cell (i, j, k), linear index q. Under the
// semi-Lagrangian scheme a band cell's value is 1 - (pre-image volume) / vol over dt and a
// face beside a band cell carries its SLfluxVol; the other faces carry A u_face. The
// back-traced vertices come from the accessor pv (at(I, J, K, c)), the face velocities from fa.
template <class PV, class FA>
__device__ inline double divg_cell_eval(const DevMesh& m, int nband, int sl, const int16_t* band, const PV& pv,
                                        const FA& fa, double dt, int64_t q, int i, int j, int k) {
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    const double vol = (m.dx * m.dy) * m.dz;
    if (sl && in_band(band, q, nband)) {
        const bool even = ((i + j + k) % 2) == 0;
        double tv = 0.0;
        for (int t = 0; t < 5; ++t) {
            double tet[4][3];
            const unsigned tq = TET5[t];
            for (int v4 = 0; v4 < 4; ++v4) {
                const int c = vert_corner(nib(tq, v4), even);
                const int I = i + (c & 1), J = j + ((c >> 1) & 1), K = k + (c >> 2);
                tet[v4][0] = pv.at(I, J, K, 0); tet[v4][1] = pv.at(I, J, K, 1); tet[v4][2] = pv.at(I, J, K, 2);
            }
            tv += tet_vol(tet[0], tet[1], tet[2], tet[3]);
        }
        const double v1 = (even ? 1.0 : -1.0) * tv;
        return (vol - v1) / vol / dt;
    }
    double verts[8][3];
    auto flux = [&](int d, int up, int64_t nb) {
        const int fi = i + (d == 1 ? up : 0), fj = j + (d == 2 ? up : 0), fk = k + (d == 3 ? up : 0);
        if (sl && in_band(band, nb, nband)) return sl_flux_vol_acc(m, pv, d, fi, fj, fk, verts) / dt;
        const int64_t fc = m.cell(fi, fj, fk);
        if (d == 1) return (m.dy * m.dz) * fa.fu(fc);
        if (d == 2) return (m.dx * m.dz) * fa.fv(fc);
        return (m.dx * m.dy) * fa.fw(fc);
    };
    const double Fxm = flux(1, 0, q - 1), Fxp = flux(1, 1, q + 1);
    const double Fym = flux(2, 0, q - sy), Fyp = flux(2, 1, q + sy);
    const double Fzm = flux(3, 0, q - sz), Fzp = flux(3, 1, q + sz);
    return (((((Fxp - Fxm) + Fyp) - Fym) + Fzp) - Fzm) / vol;
}

// correction_tets_{x,y,z} (src/VFgeom.jl:581-640): the coordinate of the apex e along the face
// normal that gives the pyramid (a,b,d,e) + (b,c,d,e) the volume `vol`; e starts at the
// centroid of a, b, c, d. The terms are added left to right as the reference writes them.
__device__ inline void correction_apex(int dir, const double* a, const double* b, const double* cc, const double* d,
                                       double vol, double e[3]) {
    for (int q = 0; q < 3; ++q) e[q] = 0.25 * (((a[q] + b[q]) + cc[q]) + d[q]);
    const double a1 = a[0], a2 = a[1], a3 = a[2], b1 = b[0], b2 = b[1], b3 = b[2];
    const double c1 = cc[0], c2 = cc[1], c3 = cc[2], d1 = d[0], d2 = d[1], d3 = d[2];
    const double e1 = e[0], e2 = e[1], e3 = e[2];
    double num = 6.0 * vol;
    num = num + a1 * b2 * d3; num = num - a1 * b3 * d2; num = num - a2 * b1 * d3; num = num + a2 * b3 * d1;
    num = num + a3 * b1 * d2; num = num - a3 * b2 * d1;
    double den;
    if (dir == 1) {
        num = num - a1 * b2 * e3; num = num + a1 * b3 * e2; num = num + a2 * b1 * e3; num = num - a3 * b1 * e2;
        num = num + b1 * c2 * d3; num = num - b1 * c3 * d2; num = num - b2 * c1 * d3; num = num + b2 * c3 * d1;
        num = num + b3 * c1 * d2; num = num - b3 * c2 * d1; num = num + a1 * d2 * e3; num = num - a1 * d3 * e2;
        num = num - a2 * d1 * e3; num = num + a3 * d1 * e2; num = num - b1 * c2 * e3; num = num + b1 * c3 * e2;
        num = num + b2 * c1 * e3; num = num - b3 * c1 * e2; num = num - c1 * d2 * e3; num = num + c1 * d3 * e2;
        num = num + c2 * d1 * e3; num = num - c3 * d1 * e2;
        den = a2 * b3; den = den - a3 * b2; den = den - a2 * d3; den = den + a3 * d2; den = den + b2 * c3;
        den = den - b3 * c2; den = den + c2 * d3; den = den - c3 * d2;
        e[0] = num / den;
    } else if (dir == 2) {
        num = num - a1 * b2 * e3; num = num + a2 * b1 * e3; num = num - a2 * b3 * e1; num = num + a3 * b2 * e1;
        num = num + b1 * c2 * d3; num = num - b1 * c3 * d2; num = num - b2 * c1 * d3; num = num + b2 * c3 * d1;
        num = num + b3 * c1 * d2; num = num - b3 * c2 * d1; num = num + a1 * d2 * e3; num = num - a2 * d1 * e3;
        num = num + a2 * d3 * e1; num = num - a3 * d2 * e1; num = num - b1 * c2 * e3; num = num + b2 * c1 * e3;
        num = num - b2 * c3 * e1; num = num + b3 * c2 * e1; num = num - c1 * d2 * e3; num = num + c2 * d1 * e3;
        num = num - c2 * d3 * e1; num = num + c3 * d2 * e1;
        den = a1 * b3; den = den - a3 * b1; den = den - a1 * d3; den = den + a3 * d1; den = den + b1 * c3;
        den = den - b3 * c1; den = den + c1 * d3; den = den - c3 * d1;
        e[1] = (-num) / den;
    } else {
        num = num + a1 * b3 * e2; num = num - a2 * b3 * e1; num = num - a3 * b1 * e2; num = num + a3 * b2 * e1;
        num = num + b1 * c2 * d3; num = num - b1 * c3 * d2; num = num - b2 * c1 * d3; num = num + b2 * c3 * d1;
        num = num + b3 * c1 * d2; num = num - b3 * c2 * d1; num = num - a1 * d3 * e2; num = num + a2 * d3 * e1;
        num = num + a3 * d1 * e2; num = num - a3 * d2 * e1; num = num + b1 * c3 * e2; num = num - b2 * c3 * e1;
        num = num - b3 * c1 * e2; num = num + b3 * c2 * e1; num = num + c1 * d3 * e2; num = num - c2 * d3 * e1;
        num = num - c3 * d1 * e2; num = num + c3 * d2 * e1;
        den = a1 * b2; den = den - a2 * b1; den = den - a1 * d2; den = den + a2 * d1; den = den + b1 * c2;
        den = den - b2 * c1; den = den + c1 * d2; den = den - c2 * d1;
        e[2] = num / den;
    }
}

// Tet `t` (0-based) of band cell (i,j,k) in transport!: 0..4 the back-traced 5-tet
// decomposition (cell2tets!, src/VFgeom.jl:345-420), 5..16 the correction tets of
// add_correction_tets (src/VFgeom.jl:513-664): faces x-, x+, y-, y+, z-, z+, two tets each.
__device__ inline void ns_tet(const DevMesh& m, const double* pv, const double* uf, const double* vf,
                              const double* wf, double dt, int i, int j, int k, int t, double tet[4][3]) {
    const bool even = ((i + j + k) % 2) == 0;
    if (t < 5) {
        const unsigned tv = TET5[t];
        for (int n = 0; n < 4; ++n) {
            const int c = vert_corner(nib(tv, n), even);
            const int64_t vq = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
            tet[n][0] = pv[3 * vq + 0]; tet[n][1] = pv[3 * vq + 1]; tet[n][2] = pv[3 * vq + 2];
        }
        return;
    }
    const int f = (t - 5) >> 1, half = (t - 5) & 1;
    const int dir = 1 + (f >> 1), up = f & 1;
    const int fi = i + (dir == 1 ? up : 0), fj = j + (dir == 2 ? up : 0), fk = k + (dir == 3 ? up : 0);
    double verts[8][3];
    const double flux_vol = sl_flux_vol(m, pv, dir, fi, fj, fk, verts);
    double divg_vol;
    const int64_t fc = m.cell(fi, fj, fk);
    if (dir == 1) divg_vol = ((dt * m.dy) * m.dz) * uf[fc];
    else if (dir == 2) divg_vol = ((dt * m.dx) * m.dz) * vf[fc];
    else divg_vol = ((dt * m.dx) * m.dy) * wf[fc];
    const bool feven = ((fi + fj + fk) % 2) == 0;
    const double vol = feven ? (divg_vol - flux_vol) : -(divg_vol - flux_vol);
    int ia, ib, ic, id;
    if (dir == 1) {
        if (feven) { ia = 4; ib = 6; ic = 2; id = 0; }
        else { ia = 7; ib = 5; ic = 1; id = 3; }
    } else if (dir == 2) { ia = 1; ib = 5; ic = 4; id = 0; }
    else { ia = 1; ib = 0; ic = 2; id = 3; }
    double e[3];
    correction_apex(dir, verts[ia], verts[ib], verts[ic], verts[id], vol, e);
    const int p0 = half ? ib : ia, p1 = half ? ic : ib;
    for (int q = 0; q < 3; ++q) {
        tet[0][q] = verts[p0][q];
        tet[1][q] = verts[p1][q];
        tet[2][q] = verts[id][q];
        tet[3][q] = e[q];
    }
}

}  // namespace mist
