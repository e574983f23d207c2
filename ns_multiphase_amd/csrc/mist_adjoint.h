// mist_adjoint.h -- reverse-mode (adjoint) pieces of the SL-VOF step, callable from host and
// device code: the derivative of a tet volume, of a PLIC leaf cut and of an edge cut, of the
// staggered trilinear velocity and the RK4 back-trace, of computeAlpha/computeDist, and of
// an ELVIRA candidate normal. Each function re-evaluates the forward quantities it needs in
// the same arithmetic as the forward kernels, then adds the adjoints of its inputs. Branch
// choices (cut cases, clamps, the alpha branch, the ELVIRA argmin, cell searches) are held
// fixed: the adjoint is that of the piecewise-smooth map on the current piece.
#pragma once

#include "../../include/mist_hip.h"
#include "mist_device.h"

namespace mist {

// g_k += w * d tet_vol(p0, p1, p2, p3) / d p_k, tet_vol = -det(p0-p3, p1-p3, p2-p3) / 6
__host__ __device__ inline void tet_vol_grad(const double* p0, const double* p1, const double* p2, const double* p3,
                                             double w, double* g0, double* g1, double* g2, double* g3) {
    const double a[3] = {p0[0] - p3[0], p0[1] - p3[1], p0[2] - p3[2]};
    const double b[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
    const double c[3] = {p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
    const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double s = -w / 6.0;
    for (int q = 0; q < 3; ++q) {
        g0[q] += s * bc[q];
        g1[q] += s * ca[q];
        g2[q] += s * ab[q];
        g3[q] -= s * ((bc[q] + ca[q]) + ab[q]);
    }
}

// Cut point x = (1 - mu) xa + mu xb, mu = clamp(-da / (db - da), 0, 1) (the /^ of the
// reference gives mu = 1 when |db - da| <= eps). Adds the adjoint gx of x to ga, gb and
// returns the adjoints of da, db (zero where mu is clamped).
__host__ __device__ inline void lerp_adj(const double* xa, const double* xb, double da, double db, const double* gx,
                                         double* ga, double* gb, double& gda, double& gdb) {
    const double den = db - da;
    double mu;
    bool live = false;
    if (fabs(den) <= EPS) {
        mu = 1.0;
    } else {
        const double raw = -da / den;
        mu = jmin(1.0, jmax(0.0, raw));
        live = raw > 0.0 && raw < 1.0;
    }
    double s = 0.0;
    for (int q = 0; q < 3; ++q) {
        ga[q] += (1.0 - mu) * gx[q];
        gb[q] += mu * gx[q];
        s += gx[q] * (xb[q] - xa[q]);
    }
    if (live) {
        const double den2 = den * den;
        gda = s * (-db / den2);
        gdb = s * (da / den2);
    } else {
        gda = 0.0;
        gdb = 0.0;
    }
}

// The leaf of cutTet (src/VFgeom.jl:757-830): tet x cut by the plane n.x = D; the liquid
// volume L (n.x < D side) and the total V. Adds gL dL + gV dV to gx (vertices), gn, gD.
__host__ __device__ inline void leaf_adj(const double x[4][3], double n0, double n1, double n2, double Dc, double gL,
                                         double gV, double gx[4][3], double gn[3], double& gD) {
    double d[4], dd[4];
    for (int v = 0; v < 4; ++v) {
        d[v] = ((n0 * x[v][0] + n1 * x[v][1]) + n2 * x[v][2]) - Dc;
        dd[v] = d[v];
    }
    const int cs = d2case_snap(dd);
    double Q[8][3];
    for (int v = 0; v < 4; ++v)
        for (int q = 0; q < 3; ++q) Q[v][q] = x[v][q];
    const int nv = CutTab::nvert[cs];
    for (int e = 0; e < nv; ++e) {
        const int a = CutTab::edge[cs][e] & 15, b = CutTab::edge[cs][e] >> 4;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-d[a], d[b] - d[a])));
        for (int q = 0; q < 3; ++q) Q[4 + e][q] = (1.0 - mu) * Q[a][q] + mu * Q[b][q];
    }
    double g8[8][3] = {};
    const int NT = CutTab::ntets[cs], NN = CutTab::nntet[cs];
    for (int n = NT; n >= NN; --n) {
        const int v0 = vtet_of(cs, n - 1, 0), v1 = vtet_of(cs, n - 1, 1), v2 = vtet_of(cs, n - 1, 2),
                  v3 = vtet_of(cs, n - 1, 3);
        tet_vol_grad(Q[v0], Q[v1], Q[v2], Q[v3], gL + gV, g8[v0], g8[v1], g8[v2], g8[v3]);
    }
    for (int n = 1; n <= NN - 1; ++n) {
        const int v0 = vtet_of(cs, n - 1, 0), v1 = vtet_of(cs, n - 1, 1), v2 = vtet_of(cs, n - 1, 2),
                  v3 = vtet_of(cs, n - 1, 3);
        tet_vol_grad(Q[v0], Q[v1], Q[v2], Q[v3], gV, g8[v0], g8[v1], g8[v2], g8[v3]);
    }
    double gd[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = nv - 1; e >= 0; --e) {
        const int a = CutTab::edge[cs][e] & 15, b = CutTab::edge[cs][e] >> 4;
        double gda, gdb;
        lerp_adj(Q[a], Q[b], d[a], d[b], g8[4 + e], g8[a], g8[b], gda, gdb);
        gd[a] += gda;
        gd[b] += gdb;
    }
    const double n[3] = {n0, n1, n2};
    for (int v = 0; v < 4; ++v)
        for (int q = 0; q < 3; ++q) {
            gx[v][q] += g8[v][q] + gd[v] * n[q];
            gn[q] += gd[v] * x[v][q];
        }
    gD -= ((gd[0] + gd[1]) + gd[2]) + gd[3];
}

// ---- velocity interpolation and back-trace ---------------------------------------------
// One component of get_velocity_{u,v,w}face: trilinear weights (upper weights wx1, wy1, wz1
// with spacings hx, hy, hz) of the 8 values around base cell (i, j, k) of array A.
struct TriStencil {
    int64_t base;
    double wx1, wy1, wz1, hx, hy, hz;
};

__host__ __device__ inline void vel_stencils(const DevMesh& m, const double pt[3], int i0, int j0, int k0,
                                             TriStencil S[3]) {
    const int lo = 1 - m.ng;
    int i, j, k;
    i = find_cell(pt[0], i0, lo, m.Nx + m.ng, m.x);
    j = find_cell(pt[1], j0, lo, m.Ny + m.ng, m.ym);
    k = find_cell(pt[2], k0, lo, m.Nz + m.ng, m.zm);
    double hx = m.x[i + 1] - m.x[i], hy = m.ym[j + 1] - m.ym[j], hz = m.zm[k + 1] - m.zm[k];
    double wx1 = sdiv(pt[0] - m.x[i], hx), wy1 = sdiv(pt[1] - m.ym[j], hy), wz1 = sdiv(pt[2] - m.zm[k], hz);
    S[0] = TriStencil{m.cell(i, j, k), wx1, wy1, wz1, hx, hy, hz};
    i = find_cell(pt[0], i0, lo, m.Nx + m.ng, m.xm);
    j = find_cell(pt[1], j0, lo, m.Ny + m.ng, m.y);
    hx = m.xm[i + 1] - m.xm[i];
    hy = m.y[j + 1] - m.y[j];
    wx1 = sdiv(pt[0] - m.xm[i], hx);
    wy1 = sdiv(pt[1] - m.y[j], hy);
    S[1] = TriStencil{m.cell(i, j, k), wx1, wy1, wz1, hx, hy, hz};
    j = find_cell(pt[1], j0, lo, m.Ny + m.ng, m.ym);
    k = find_cell(pt[2], k0, lo, m.Nz + m.ng, m.z);
    hy = m.ym[j + 1] - m.ym[j];
    hz = m.z[k + 1] - m.z[k];
    wy1 = sdiv(pt[1] - m.ym[j], hy);
    wz1 = sdiv(pt[2] - m.z[k], hz);
    S[2] = TriStencil{m.cell(i, j, k), wx1, wy1, wz1, hx, hy, hz};
}

// Adjoint of v = vel_face(pt): gpt += (dv/dpt)^T gv, and acc(c, idx, w) adds w to the
// adjoint of component c's face array at idx.
template <class Acc>
__host__ __device__ inline void vel_face_adj(const DevMesh& m, const double* const F[3], const double pt[3], int i0,
                                             int j0, int k0, const double gv[3], double gpt[3], Acc&& acc) {
    TriStencil S[3];
    vel_stencils(m, pt, i0, j0, k0, S);
    const int64_t sx = 1, sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int c = 0; c < 3; ++c) {
        if (gv[c] == 0.0) continue;
        const TriStencil& T = S[c];
        const double* A = F[c];
        const double W[2][3] = {{1.0 - T.wx1, 1.0 - T.wy1, 1.0 - T.wz1}, {T.wx1, T.wy1, T.wz1}};
        double dwx = 0.0, dwy = 0.0, dwz = 0.0;  // dv / d(wx1, wy1, wz1)
        for (int cz = 0; cz < 2; ++cz)
            for (int cy = 0; cy < 2; ++cy)
                for (int cx = 0; cx < 2; ++cx) {
                    const int64_t idx = T.base + cx * sx + cy * sy + cz * sz;
                    const double a = A[idx];
                    acc(c, idx, gv[c] * ((W[cx][0] * W[cy][1]) * W[cz][2]));
                    const double sxg = cx ? 1.0 : -1.0, syg = cy ? 1.0 : -1.0, szg = cz ? 1.0 : -1.0;
                    dwx += sxg * (W[cy][1] * W[cz][2]) * a;
                    dwy += syg * (W[cx][0] * W[cz][2]) * a;
                    dwz += szg * (W[cx][0] * W[cy][1]) * a;
                }
        if (fabs(T.hx) > EPS) gpt[0] += gv[c] * dwx / T.hx;
        if (fabs(T.hy) > EPS) gpt[1] += gv[c] * dwy / T.hy;
        if (fabs(T.hz) > EPS) gpt[2] += gv[c] * dwz / T.hz;
    }
}

// Adjoint of project! (src/Tools.jl:392-418) with respect to the face velocities: the
// back-traced point out(x) receives the adjoint g; the start point x is a fixed mesh vertex.
template <class Acc>
__host__ __device__ inline void project_adj(const DevMesh& m, const DevParams& p, const double* const F[3],
                                            const double x[3], int i, int j, int k, double dt, const double g[3],
                                            Acc&& acc) {
    if (p.zalesak) return;  // closed-form rotation: no velocity dependence
    double v1[3], v2[3], v3[3], q2[3], q3[3], q4[3];
    if (p.projection == MIST_PROJ_RK4) {
        vel_face(m, F[0], F[1], F[2], x, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q2[n] = x[n] - (0.5 * dt) * v1[n];
        vel_face(m, F[0], F[1], F[2], q2, i, j, k, v2);
        for (int n = 0; n < 3; ++n) q3[n] = x[n] - (0.5 * dt) * v2[n];
        vel_face(m, F[0], F[1], F[2], q3, i, j, k, v3);
        for (int n = 0; n < 3; ++n) q4[n] = x[n] - dt * v3[n];
        // out = x + (-dt/6) (((v1 + 2 v2) + 2 v3) + v4)
        const double c = (-dt) / 6.0;
        double g1[3], g2[3], g3[3], g4[3], gq[3];
        for (int n = 0; n < 3; ++n) {
            g1[n] = c * g[n];
            g2[n] = 2.0 * c * g[n];
            g3[n] = 2.0 * c * g[n];
            g4[n] = c * g[n];
        }
        gq[0] = gq[1] = gq[2] = 0.0;
        vel_face_adj(m, F, q4, i, j, k, g4, gq, acc);
        for (int n = 0; n < 3; ++n) g3[n] += -dt * gq[n];
        gq[0] = gq[1] = gq[2] = 0.0;
        vel_face_adj(m, F, q3, i, j, k, g3, gq, acc);
        for (int n = 0; n < 3; ++n) g2[n] += -(0.5 * dt) * gq[n];
        gq[0] = gq[1] = gq[2] = 0.0;
        vel_face_adj(m, F, q2, i, j, k, g2, gq, acc);
        for (int n = 0; n < 3; ++n) g1[n] += -(0.5 * dt) * gq[n];
        gq[0] = gq[1] = gq[2] = 0.0;
        vel_face_adj(m, F, x, i, j, k, g1, gq, acc);
    } else if (p.projection == MIST_PROJ_EULER) {
        double g1[3], gq[3] = {0.0, 0.0, 0.0};
        for (int n = 0; n < 3; ++n) g1[n] = (-dt) * g[n];
        vel_face_adj(m, F, x, i, j, k, g1, gq, acc);
    } else {  // midpoint: out = x - dt v2, v2 = U(x - dt/2 v1)
        vel_face(m, F[0], F[1], F[2], x, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q2[n] = x[n] - (0.5 * dt) * v1[n];
        double g2[3], g1[3], gq[3] = {0.0, 0.0, 0.0};
        for (int n = 0; n < 3; ++n) g2[n] = (-dt) * g[n];
        vel_face_adj(m, F, q2, i, j, k, g2, gq, acc);
        for (int n = 0; n < 3; ++n) g1[n] = -(0.5 * dt) * gq[n];
        gq[0] = gq[1] = gq[2] = 0.0;
        vel_face_adj(m, F, x, i, j, k, g1, gq, acc);
    }
}

// ---- computeAlpha / computeDist ---------------------------------------------------------
// Adjoint of alpha = compute_alpha(a, b, c, VF): adds galpha * d alpha / d(a, b, c, VF).
// The branch is that of the forward evaluation; the cubic branches are differentiated
// implicitly through the cubic whose root the trigonometric formula returns.
__host__ __device__ inline void compute_alpha_adj(double a, double b, double c, double VF, double galpha,
                                                  double& ga, double& gb, double& gc, double& gVF) {
    double m[3] = {a, b, c};
    int id[3] = {0, 1, 2};
    double t;
    int ti;
    if (m[0] > m[1]) { t = m[0]; m[0] = m[1]; m[1] = t; ti = id[0]; id[0] = id[1]; id[1] = ti; }
    if (m[1] > m[2]) { t = m[1]; m[1] = m[2]; m[2] = t; ti = id[1]; id[1] = id[2]; id[2] = ti; }
    if (m[0] > m[1]) { t = m[0]; m[0] = m[1]; m[1] = t; ti = id[0]; id[0] = id[1]; id[1] = ti; }
    const double m1 = m[0], m2 = m[1], m3 = m[2];
    const double m12 = m1 + m2;
    const double den1 = (6.0 * m2) * m3;
    const double V1 = (m1 * m1) / jmax(den1, EPS);
    const double V2 = V1 + sdiv(0.5 * (m2 - m1), m3);
    double V3;
    if (m12 <= m3) V3 = sdiv(0.5 * m12, m3);
    else V3 = sdiv(((m3 * m3) * (3.0 * m12 - m3) + (m1 * m1) * (m1 - 3.0 * m3)) + (m2 * m2) * (m2 - 3.0 * m3),
                   ((6.0 * m1) * m2) * m3);
    const double alpha = compute_alpha(a, b, c, VF);
    double d1 = 0.0, d2 = 0.0, d3 = 0.0, dV = 0.0;  // d alpha / d(m1, m2, m3, VF)
    if (0.0 <= VF && VF < V1) {
        const double P = (((6.0 * m1) * m2) * m3) * VF;
        if (P > 0.0) {
            const double dP = alpha / (3.0 * P);
            d1 = dP * (6.0 * m2 * m3 * VF);
            d2 = dP * (6.0 * m1 * m3 * VF);
            d3 = dP * (6.0 * m1 * m2 * VF);
            dV = dP * (6.0 * m1 * m2 * m3);
        }
    } else if (V1 <= VF && VF < V2) {
        const double S = (m1 * m1) + ((8.0 * m2) * m3) * (VF - V1);
        if (S > 0.0) {
            const double dS = 0.25 / sqrt(S);
            double dV1_1, dV1_2, dV1_3;
            if (den1 > EPS) {
                dV1_1 = (2.0 * m1) / den1;
                dV1_2 = -V1 / m2;
                dV1_3 = -V1 / m3;
            } else {
                dV1_1 = (2.0 * m1) / EPS;
                dV1_2 = dV1_3 = 0.0;
            }
            const double k8 = 8.0 * m2 * m3;
            d1 = 0.5 + dS * (2.0 * m1 - k8 * dV1_1);
            d2 = dS * (8.0 * m3 * (VF - V1) - k8 * dV1_2);
            d3 = dS * (8.0 * m2 * (VF - V1) - k8 * dV1_3);
            dV = dS * k8;
        } else {
            d1 = 0.5;
        }
    } else if (V2 <= VF && VF < V3) {
        // P = -A^3 + 3 m12 A^2 - 3 (m1^2 + m2^2) A + m1^3 + m2^3 - 6 m1 m2 m3 VF
        const double A = alpha;
        const double dPdA = -3.0 * A * A + 6.0 * m12 * A - 3.0 * (m1 * m1 + m2 * m2);
        if (fabs(dPdA) > 0.0) {
            const double p1 = 3.0 * A * A - 6.0 * m1 * A + 3.0 * m1 * m1 - 6.0 * m2 * m3 * VF;
            const double p2 = 3.0 * A * A - 6.0 * m2 * A + 3.0 * m2 * m2 - 6.0 * m1 * m3 * VF;
            const double p3 = -6.0 * m1 * m2 * VF;
            const double pV = -6.0 * m1 * m2 * m3;
            d1 = -p1 / dPdA; d2 = -p2 / dPdA; d3 = -p3 / dPdA; dV = -pV / dPdA;
        }
    } else if (m12 <= m3) {
        d1 = 0.5; d2 = 0.5; d3 = VF; dV = m3;
    } else {
        // P = -2 A^3 + 3 A^2 - 3 (m1^2 + m2^2 + m3^2) A + m1^3 + m2^3 + m3^3 - 6 m1 m2 m3 VF
        const double A = alpha;
        const double dPdA = -6.0 * A * A + 6.0 * A - 3.0 * (m1 * m1 + m2 * m2 + m3 * m3);
        if (fabs(dPdA) > 0.0) {
            const double p1 = -6.0 * m1 * A + 3.0 * m1 * m1 - 6.0 * m2 * m3 * VF;
            const double p2 = -6.0 * m2 * A + 3.0 * m2 * m2 - 6.0 * m1 * m3 * VF;
            const double p3 = -6.0 * m3 * A + 3.0 * m3 * m3 - 6.0 * m1 * m2 * VF;
            const double pV = -6.0 * m1 * m2 * m3;
            d1 = -p1 / dPdA; d2 = -p2 / dPdA; d3 = -p3 / dPdA; dV = -pV / dPdA;
        }
    }
    double g[3] = {0.0, 0.0, 0.0};
    g[id[0]] = galpha * d1;
    g[id[1]] = galpha * d2;
    g[id[2]] = galpha * d3;
    ga += g[0];
    gb += g[1];
    gc += g[2];
    gVF += galpha * dV;
}

// Adjoint of D = compute_dist(n, VF) of cell (i, j, k): adds gD dD/dn to gn and gD dD/dVF
// to gVF. Empty and full cells have a constant D.
__host__ __device__ inline void compute_dist_adj(const DevMesh& m, const DevParams& p, int i, int j, int k, double nx,
                                                 double ny, double nz, double VF, double gD, double gn[3],
                                                 double& gVF) {
    if (VF < p.VFlo || VF > p.VFhi) return;
    const double h[3] = {m.dx, m.dy, m.dz};
    const double n[3] = {nx, ny, nz};
    const double tt[3] = {nx * m.dx, ny * m.dy, nz * m.dz};
    const double norm = ((fabs(tt[0]) + fabs(tt[1])) + fabs(tt[2])) + EPS;
    double mm[3];
    for (int q = 0; q < 3; ++q) mm[q] = tt[q] / norm;
    double factor = 0.0;
    for (int q = 0; q < 3; ++q)
        if (mm[q] < 0.0) factor += mm[q];
    const bool hi = VF > 0.5;
    const double VFo = hi ? 1.0 - VF : VF;
    double alpha0 = compute_alpha(fabs(mm[0]), fabs(mm[1]), fabs(mm[2]), VFo);
    double alpha = hi ? 1.0 - alpha0 : alpha0;
    alpha += factor;
    alpha -= 0.5 * ((mm[0] + mm[1]) + mm[2]);
    // dist = alpha norm + n . centre
    const double xc[3] = {m.xm[i], m.ym[j], m.zm[k]};
    for (int q = 0; q < 3; ++q) gn[q] += gD * xc[q];
    const double g_alpha = gD * norm;
    double g_norm = gD * alpha;
    double g_mm[3];
    for (int q = 0; q < 3; ++q) g_mm[q] = g_alpha * ((mm[q] < 0.0 ? 1.0 : 0.0) - 0.5);
    const double g_alpha0 = hi ? -g_alpha : g_alpha;
    double ga[3] = {0.0, 0.0, 0.0}, gVFo = 0.0;
    compute_alpha_adj(fabs(mm[0]), fabs(mm[1]), fabs(mm[2]), VFo, g_alpha0, ga[0], ga[1], ga[2], gVFo);
    gVF += hi ? -gVFo : gVFo;
    for (int q = 0; q < 3; ++q) g_mm[q] += ga[q] * (mm[q] > 0.0 ? 1.0 : (mm[q] < 0.0 ? -1.0 : 0.0));
    // mm = tt / norm, norm = sum |tt| + eps
    double g_tt[3];
    double s = 0.0;
    for (int q = 0; q < 3; ++q) {
        g_tt[q] = g_mm[q] / norm;
        s += g_mm[q] * tt[q];
    }
    g_norm -= s / (norm * norm);
    for (int q = 0; q < 3; ++q) {
        g_tt[q] += g_norm * (tt[q] > 0.0 ? 1.0 : (tt[q] < 0.0 ? -1.0 : 0.0));
        gn[q] += g_tt[q] * h[q];
    }
    (void)n;
}

// ---- ELVIRA candidate normal -----------------------------------------------------------
// Candidate `cand` of ELVIRA_calc! (src/ELVIRA.jl:62-100: family cand / 9, outer stencil pair
// (cand % 9) / 3, inner cand % 3) of cell (i, j, k): the raw vector (two column-sum slopes and
// a +-1 component) and, for its normalisation n = r / |r|, the adjoint gn of n mapped to the
// liquid fractions: acc(di, dj, dk, w) adds w to the adjoint of VF(i+di, j+dj, k+dk).
template <class VFAt, class Acc>
__host__ __device__ inline void elvira_cand_adj(const DevMesh& m, int i, int j, int k, int cand, VFAt&& at,
                                                const double gn[3], Acc&& acc) {
    const int fam = cand / 9, a = (cand % 9) / 3, b = cand % 3;
    auto P0 = [](int q) { return q == 2 ? 0 : -1; };
    auto P1 = [](int q) { return q == 0 ? 0 : 1; };
    double r[3];
    int s1a = 0, s1b = 0, s2a = 0, s2b = 0;  // stencil pairs of the two slopes
    double w1 = 0.0, w2 = 0.0, h1 = 0.0, h2 = 0.0;
    int ax1 = 0, ax2 = 0, axs = 0;  // slope axes and the +-1 axis
    if (fam == 0) {
        ax1 = 0; ax2 = 1; axs = 2;
        s1a = P0(b); s1b = P1(b); s2a = P0(a); s2b = P1(a);
        w1 = w2 = m.dz;
        h1 = m.xm[i + s1b] - m.xm[i + s1a];
        h2 = m.ym[j + s2b] - m.ym[j + s2a];
    } else if (fam == 1) {
        ax1 = 0; ax2 = 2; axs = 1;
        s1a = P0(b); s1b = P1(b); s2a = P0(a); s2b = P1(a);
        w1 = w2 = m.dy;
        h1 = m.xm[i + s1b] - m.xm[i + s1a];
        h2 = m.zm[k + s2b] - m.zm[k + s2a];
    } else {
        ax1 = 1; ax2 = 2; axs = 0;
        s1a = P0(b); s1b = P1(b); s2a = P0(a); s2b = P1(a);
        w1 = w2 = m.dx;
        h1 = m.ym[j + s1b] - m.ym[j + s1a];
        h2 = m.zm[k + s2b] - m.zm[k + s2a];
    }
    // the column sums run along the +-1 axis: offset o along the slope axis `ax`, t along axs
    auto off = [&](int ax, int o, int t, int& di, int& dj, int& dk) {
        int d[3] = {0, 0, 0};
        d[ax] = o;
        d[axs] = t;
        di = d[0]; dj = d[1]; dk = d[2];
    };
    auto colsum = [&](int ax, int o, double w) {
        double s = 0.0;
        for (int t = -1; t <= 1; ++t) {
            int di, dj, dk;
            off(ax, o, t, di, dj, dk);
            s = s + at(di, dj, dk) * w;
        }
        return s;
    };
    r[ax1] = -(colsum(ax1, s1b, w1) - colsum(ax1, s1a, w1)) / h1;
    r[ax2] = -(colsum(ax2, s2b, w2) - colsum(ax2, s2a, w2)) / h2;
    {
        int di, dj, dk, ei, ej, ek;
        int d[3] = {0, 0, 0};
        d[axs] = 1;
        di = d[0]; dj = d[1]; dk = d[2];
        d[axs] = -1;
        ei = d[0]; ej = d[1]; ek = d[2];
        r[axs] = at(di, dj, dk) > at(ei, ej, ek) ? -1.0 : 1.0;
    }
    const double len = sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    const double n[3] = {r[0] / len, r[1] / len, r[2] / len};
    const double ng = (n[0] * gn[0] + n[1] * gn[1]) + n[2] * gn[2];
    double gr[3];
    for (int q = 0; q < 3; ++q) gr[q] = (gn[q] - n[q] * ng) / len;
    for (int t = -1; t <= 1; ++t) {
        int di, dj, dk;
        off(ax1, s1b, t, di, dj, dk);
        acc(di, dj, dk, -gr[ax1] * w1 / h1);
        off(ax1, s1a, t, di, dj, dk);
        acc(di, dj, dk, gr[ax1] * w1 / h1);
        off(ax2, s2b, t, di, dj, dk);
        acc(di, dj, dk, -gr[ax2] * w2 / h2);
        off(ax2, s2a, t, di, dj, dk);
        acc(di, dj, dk, gr[ax2] * w2 / h2);
    }
}

// ---- the leaf cut on V3 vertices (the reverse of cut_case and cut_canonical) ----------
// Gradient of tet_vol3 (the volume -det(p0-p3, p1-p3, p2-p3) / 6) with weight w, into g0..g3.
__host__ __device__ __forceinline__ void tet_vol3_grad(const V3& p0, const V3& p1, const V3& p2, const V3& p3,
                                                       double w, V3& g0, V3& g1, V3& g2, V3& g3) {
    const double a1 = p0.x - p3.x, a2 = p0.y - p3.y, a3 = p0.z - p3.z;
    const double b1 = p1.x - p3.x, b2 = p1.y - p3.y, b3 = p1.z - p3.z;
    const double c1 = p2.x - p3.x, c2 = p2.y - p3.y, c3 = p2.z - p3.z;
    const double s = -w / 6.0;
    const V3 ga{s * (b2 * c3 - c2 * b3), s * (c1 * b3 - b1 * c3), s * (b1 * c2 - c1 * b2)};
    const V3 gb{s * (c2 * a3 - a2 * c3), s * (a1 * c3 - c1 * a3), s * (c1 * a2 - a1 * c2)};
    const V3 gc{s * (a2 * b3 - b2 * a3), s * (b1 * a3 - a1 * b3), s * (a1 * b2 - b1 * a2)};
    g0.x += ga.x; g0.y += ga.y; g0.z += ga.z;
    g1.x += gb.x; g1.y += gb.y; g1.z += gb.z;
    g2.x += gc.x; g2.y += gc.y; g2.z += gc.z;
    g3.x -= (ga.x + gb.x) + gc.x; g3.y -= (ga.y + gb.y) + gc.y; g3.z -= (ga.z + gb.z) + gc.z;
}

// lerp_adj on V3: x = lerp3(a, b, mu(da, db)); adds gx's share to ga, gb, returns gda, gdb.
__host__ __device__ __forceinline__ void lerp3_adj(const V3& a, const V3& b, double da, double db, const V3& gx,
                                                   V3& ga, V3& gb, double& gda, double& gdb) {
    const double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, G[3] = {gx.x, gx.y, gx.z};
    double sa[3] = {0.0, 0.0, 0.0}, sb[3] = {0.0, 0.0, 0.0};
    lerp_adj(A, B, da, db, G, sa, sb, gda, gdb);
    ga.x += sa[0]; ga.y += sa[1]; ga.z += sa[2];
    gb.x += sb[0]; gb.y += sb[1]; gb.z += sb[2];
}

// Reverse of cut_case<CS, true> (the PLIC cut of a leaf): the liquid sub-tets carry gL + gV,
// the gas ones gV; the cut vertices pass their adjoints to their edge ends and to the vertex
// distances dm. Adds to gP and gd.
template <int CS>
__host__ __device__ __forceinline__ void cut_case_adj(const V3 (&P)[4], const double (&dm)[4], double gL,
                                                      double gV, V3 (&gP)[4], double (&gd)[4]) {
    constexpr int NV = CutTab::nvert[CS], NT = CutTab::ntets[CS], NN = CutTab::nntet[CS];
    V3 Q[8];
    Q[0] = P[0]; Q[1] = P[1]; Q[2] = P[2]; Q[3] = P[3];
    sfor<0, NV>([&](auto I) {
        constexpr int n = decltype(I)::value;
        constexpr int v1 = CutTab::edge[CS][n] & 15, v2 = CutTab::edge[CS][n] >> 4;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-dm[v1], dm[v2] - dm[v1])));
        Q[4 + n] = lerp3(Q[v1], Q[v2], mu);
    });
    V3 g[8];
    sfor<0, 8>([&](auto I) { g[decltype(I)::value] = V3{0.0, 0.0, 0.0}; });
    sfor<0, NT - NN + 1>([&](auto I) {
        constexpr int n = NT - decltype(I)::value;
        constexpr int a = vtet_of(CS, n - 1, 0), b = vtet_of(CS, n - 1, 1), c = vtet_of(CS, n - 1, 2),
                      d = vtet_of(CS, n - 1, 3);
        tet_vol3_grad(Q[a], Q[b], Q[c], Q[d], gL + gV, g[a], g[b], g[c], g[d]);
    });
    sfor<1, NN>([&](auto I) {
        constexpr int n = decltype(I)::value;
        constexpr int a = vtet_of(CS, n - 1, 0), b = vtet_of(CS, n - 1, 1), c = vtet_of(CS, n - 1, 2),
                      d = vtet_of(CS, n - 1, 3);
        tet_vol3_grad(Q[a], Q[b], Q[c], Q[d], gV, g[a], g[b], g[c], g[d]);
    });
    sfor<0, NV>([&](auto I) {
        constexpr int n = NV - 1 - decltype(I)::value;
        constexpr int v1 = CutTab::edge[CS][n] & 15, v2 = CutTab::edge[CS][n] >> 4;
        double ga, gb;
        lerp3_adj(Q[v1], Q[v2], dm[v1], dm[v2], g[4 + n], g[v1], g[v2], ga, gb);
        gd[v1] += ga;
        gd[v2] += gb;
    });
    sfor<0, 4>([&](auto I) {
        constexpr int v = decltype(I)::value;
        gP[v].x += g[v].x; gP[v].y += g[v].y; gP[v].z += g[v].z;
    });
}

__host__ __device__ __forceinline__ void cut_dispatch_adj(int cs, const V3 (&P)[4], const double (&dm)[4],
                                                          double gL, double gV, V3 (&gP)[4], double (&gd)[4]) {
    switch (cs) {
#define MIST_CASE(k) \
    case k: cut_case_adj<k>(P, dm, gL, gV, gP, gd); break;
        MIST_CASE(0) MIST_CASE(1) MIST_CASE(2) MIST_CASE(3) MIST_CASE(4) MIST_CASE(5) MIST_CASE(6) MIST_CASE(7)
        MIST_CASE(8) MIST_CASE(9) MIST_CASE(10) MIST_CASE(11) MIST_CASE(12) MIST_CASE(13) MIST_CASE(14) MIST_CASE(15)
#undef MIST_CASE
        default: break;
    }
}

// Reverse of cut_canonical: the same renumbering onto case 1 or case 3, whose liquid and gas
// pieces carry the weights of the leaf's own liquid (gL + gV) and gas (gV) volumes after the
// swap of a lone vertex below the plane and the sign of an odd renumbering; the adjoints of the
// renumbered vertices and distances go back to gP and gd. Cases 0 and 15 (the leaf whole) give
// the tet's volume gradient.
__host__ __device__ __forceinline__ void cut_canonical_adj(int cs, const V3 (&P)[4], const double (&dm)[4],
                                                           double gL, double gV, V3 (&gP)[4], double (&gd)[4]) {
    if (cs == 0 || cs == 15) {
        tet_vol3_grad(P[0], P[1], P[2], P[3], cs == 0 ? gL + gV : gV, gP[0], gP[1], gP[2], gP[3]);
        return;
    }
    const unsigned c = canon_case(cs);
    const double s = (c & (1u << 16)) ? -1.0 : 1.0;
    // weights of the canonical case's liquid and gas pieces
    double wl = s * (gL + gV), wg = s * gV;
    if (c & (1u << 17)) {
        const double t = wl;
        wl = wg;
        wg = t;
    }
    auto run = [&](auto A, auto B, auto C, auto D) {
        constexpr int a = decltype(A)::value, b = decltype(B)::value, cc = decltype(C)::value,
                      d = decltype(D)::value;
        const V3 Q[4] = {P[a], P[b], P[cc], P[d]};
        const double e[4] = {dm[a], dm[b], dm[cc], dm[d]};
        V3 gQ[4] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double ge[4] = {0.0, 0.0, 0.0, 0.0};
        // cut_case_adj weighs liquid pieces by its gL + gV and gas pieces by its gV
        if (c & (1u << 18)) cut_case_adj<3>(Q, e, wl - wg, wg, gQ, ge);
        else cut_case_adj<1>(Q, e, wl - wg, wg, gQ, ge);
        gP[a].x += gQ[0].x; gP[a].y += gQ[0].y; gP[a].z += gQ[0].z;
        gP[b].x += gQ[1].x; gP[b].y += gQ[1].y; gP[b].z += gQ[1].z;
        gP[cc].x += gQ[2].x; gP[cc].y += gQ[2].y; gP[cc].z += gQ[2].z;
        gP[d].x += gQ[3].x; gP[d].y += gQ[3].y; gP[d].z += gQ[3].z;
        gd[a] += ge[0];
        gd[b] += ge[1];
        gd[cc] += ge[2];
        gd[d] += ge[3];
    };
    switch (cs) {
        case 2: case 13: run(IC<1>{}, IC<0>{}, IC<2>{}, IC<3>{}); break;
        case 4: case 11: run(IC<2>{}, IC<0>{}, IC<1>{}, IC<3>{}); break;
        case 8: case 7: run(IC<3>{}, IC<0>{}, IC<1>{}, IC<2>{}); break;
        case 5: run(IC<0>{}, IC<2>{}, IC<1>{}, IC<3>{}); break;
        case 6: run(IC<1>{}, IC<2>{}, IC<0>{}, IC<3>{}); break;
        case 9: run(IC<0>{}, IC<3>{}, IC<1>{}, IC<2>{}); break;
        case 10: run(IC<1>{}, IC<3>{}, IC<0>{}, IC<2>{}); break;
        case 12: run(IC<2>{}, IC<3>{}, IC<0>{}, IC<1>{}); break;
        default: run(IC<0>{}, IC<1>{}, IC<2>{}, IC<3>{}); break;
    }
}

}  // namespace mist
