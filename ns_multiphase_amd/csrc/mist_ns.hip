// mist_ns.hip -- gfx950 kernels of the Navier-Stokes parts of the time step
// (run_solver with solveNS, src/NavierStokes_Parallel.jl:125-168): the momentum predictor of
// transport! (src/Transport.jl:2-268) with the flux-correction tets of the finite-difference
// pressure scheme, the height-function curvature, the CSF surface force, viscous fluxes and
// gravity, the reflect boundary condition, and the face/centre velocity interpolations and
// divergence of src/Velocity.jl. Launchers at the bottom are called by the C ABI.
//
// Layout: every field is the reference's ghosted column-major array (i fastest); face arrays
// of density, viscosity and surface force carry one extra face along their own axis.
#include <algorithm>
#include <cmath>

#include "mist_internal.h"
#include "mist_kernels.h"
#include "mist_nsgeom.h"

namespace mist {

// index of face (i,j,k) in an x / y face array (one extra face along x / y); z face arrays
// index like cell fields
__device__ __forceinline__ int64_t fxi(const DevMesh& m, int i, int j, int k) {
    return (int64_t)(i - 1 + m.ng) + (int64_t)(m.cnx + 1) * ((int64_t)(j - 1 + m.ng) + (int64_t)m.cny * (k - 1 + m.ng));
}
__device__ __forceinline__ int64_t fyi(const DevMesh& m, int i, int j, int k) {
    return (int64_t)(i - 1 + m.ng) + (int64_t)m.cnx * ((int64_t)(j - 1 + m.ng) + (int64_t)(m.cny + 1) * (k - 1 + m.ng));
}
__device__ __forceinline__ int64_t fzi(const DevMesh& m, int i, int j, int k) { return m.cell(i, j, k); }


// Tets FIRST .. FIRST+T-1 of the cells of `list` through the general cut walk with the six
// momentum sums, one lane per (cell, tet), 64/T cells per block. Without FINAL the cell's
// sums go to sums[6 q ..]; with FINAL they are added to the sums already there (those of tets
// 0 .. FIRST-1) and the cell's VF (in place: the cuts read only normals and D) and
// us, vs, ws = (rho u vol) / (rho vol) are written (src/Transport.jl:107-150). The tetsign of
// the reference multiplies every tet of a cell alike and cancels in these ratios.
template <int T, int FIRST, bool FINAL>
__global__ __launch_bounds__(64) void k_ns_walk(DevMesh m, const double* pv, const int* list, const int* count,
                                                const double* nx, const double* ny, const double* nz,
                                                const double* D, LeafMomentum lm, const double* uf,
                                                const double* vf, const double* wf, double dt, double* sums,
                                                double* VF, double* us, double* vs, double* ws, int* ovf) {
    constexpr int CPB = 64 / T;
    __shared__ double s_acc[6][64];
    const int lane = threadIdx.x;
    const int cl = lane / T, t = lane - T * cl;
    const int nb = *count;
    for (int base = blockIdx.x * CPB; base < nb; base += gridDim.x * CPB) {
        const int cidx = base + cl;
        const bool active = cl < CPB && cidx < nb;
        double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int q = 0, i = 1, j = 1, k = 1;
        if (active) {
            q = list[cidx];
            interior_of(m, q, i, j, k);
            double tet[4][3];
            ns_tet(m, pv, uf, vf, wf, dt, i, j, k, FIRST + t, tet);
            int ind[4][3];
            int o = 0;
            for (int n = 0; n < 4; ++n) pt2index(m, ind[n], tet[n], i, j, k, &o);
            CutFrame<6> st[CUT_MAXLVL];
            int depth = 0;
            if (!cut_tet<6>(m, nx, ny, nz, D, lm, tet, ind, st, r, depth)) o = 1;
            if (o) atomicOr(ovf, 1);
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) s_acc[a][lane] = r[a];
        __syncthreads();
        if (active && t == 0) {
            double S[6];
            double* o = sums + 6 * (int64_t)q;
#pragma unroll
            for (int a = 0; a < 6; ++a) S[a] = (FINAL && FIRST > 0) ? o[a] : 0.0;
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int a = 0; a < 6; ++a) S[a] += s_acc[a][lane + n];
            if (FINAL) {
                VF[m.vof(i, j, k)] = S[1] / S[0];
                const int64_t c = m.cell(i, j, k);
                us[c] = S[2] / S[5];
                vs[c] = S[3] / S[5];
                ws[c] = S[4] / S[5];
            } else {
#pragma unroll
                for (int a = 0; a < 6; ++a) o[a] = S[a];
            }
        }
        __syncthreads();
    }
}

// Convective fluxes and the flux update of the cells outside the band (src/Transport.jl:35-105,
// :151-170). Each cell evaluates its six face fluxes: zero between two band cells, the SL flux
// volume at a band edge under the semi-Lagrangian scheme, the FD flux otherwise.
__global__ void k_ns_convect(DevMesh m, int nband, int sl, const int16_t* band, const double* pv, const double* u,
                             const double* v, const double* w, const double* uf, const double* vf,
                             const double* wf, double dt, double* us, double* vs, double* ws) {
    const int n = m.Nx * m.Ny * m.Nz;
    const double f = dt / ((m.dx * m.dy) * m.dz);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        if (in_band(band, c, nband)) continue;
        double F[3][2][3];  // axis, lower/upper face, component
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int64_t st = d == 0 ? 1 : (d == 1 ? (int64_t)m.cnx : (int64_t)m.cnx * m.cny);
            const double h1 = d == 0 ? m.dy : m.dx, h2 = d == 2 ? m.dy : m.dz;
            const double* vel = d == 0 ? uf : (d == 1 ? vf : wf);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int64_t b = c + s * st, a = b - st;  // cells below / above the face
                const double uface = 0.5 * (u[a] + u[b]);
                const double vface = 0.5 * (v[a] + v[b]);
                const double wface = 0.5 * (w[a] + w[b]);
                const bool bm = in_band(band, a, nband), bp = in_band(band, b, nband);
                if (bm && bp) {
                    F[d][s][0] = F[d][s][1] = F[d][s][2] = 0.0;
                } else if ((bm || bp) && sl) {
                    double verts[8][3];
                    const int fi = i + (d == 0 ? s : 0), fj = j + (d == 1 ? s : 0), fk = k + (d == 2 ? s : 0);
                    const double vol = sl_flux_vol(m, pv, d + 1, fi, fj, fk, verts) / dt;
                    F[d][s][0] = -uface * vol; F[d][s][1] = -vface * vol; F[d][s][2] = -wface * vol;
                } else {
                    const double vn = vel[b];
                    F[d][s][0] = ((-uface * h1) * h2) * vn;
                    F[d][s][1] = ((-vface * h1) * h2) * vn;
                    F[d][s][2] = ((-wface * h1) * h2) * vn;
                }
            }
        }
        const double* old[3] = {u, v, w};
        double* out[3] = {us, vs, ws};
#pragma unroll
        for (int q = 0; q < 3; ++q)
            out[q][c] = old[q][c] + f * (((((F[0][1][q] - F[0][0][q]) + F[1][1][q]) - F[1][0][q]) + F[2][1][q]) -
                                         F[2][0][q]);
    }
}

// compute_curvature! (src/ELVIRA.jl:143-192), with fill!(Curve, 0) for the cells it skips:
// the height function of the 3x3x7 column along the strictly dominant normal component.
__global__ __launch_bounds__(256) void k_curvature(DevMesh m, const double* VF, const double* nx, const double* ny, const double* nz,
                            double* Curve) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        double K = 0.0;
        if (i >= 1 && i <= m.Nx && j >= 1 && j <= m.Ny && k >= 1 && k <= m.Nz) {
            const double anx = fabs(nx[q]), any = fabs(ny[q]), anz = fabs(nz[q]);
            int dom = -1;
            if (anx > any && anx > anz) dom = 0;
            else if (any > anx && any > anz) dom = 1;
            else if (anz > any && anz > anx) dom = 2;
            if (dom >= 0) {
                // hf[a][b]: a, b = offsets + 1 along the first / second tangential axis
                // (x: y, z; y: z, x; z: x, y), each a sum of 7 cells along `dom`
                double hf[3][3];
                const int ea[3] = {dom == 2 ? 1 : 0, dom == 0 ? 1 : 0, dom == 1 ? 1 : 0};  // first axis
                const int eb[3] = {dom == 1 ? 1 : 0, dom == 2 ? 1 : 0, dom == 0 ? 1 : 0};  // second axis
                const int ed[3] = {dom == 0 ? 1 : 0, dom == 1 ? 1 : 0, dom == 2 ? 1 : 0};
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const int oi = (a - 1) * ea[0] + (b - 1) * eb[0];
                        const int oj = (a - 1) * ea[1] + (b - 1) * eb[1];
                        const int ok = (a - 1) * ea[2] + (b - 1) * eb[2];
                        double s = 0.0;
#pragma unroll
                        for (int r = -3; r <= 3; ++r)
                            s += VF[m.vof(i + oi + r * ed[0], j + oj + r * ed[1], k + ok + r * ed[2])];
                        hf[a][b] = s;
                    }
                const double hd = dom == 0 ? m.dx : (dom == 1 ? m.dy : m.dz);
                const double h1 = dom == 0 ? m.dy : (dom == 1 ? m.dz : m.dx);
                const double h2 = dom == 0 ? m.dz : (dom == 1 ? m.dx : m.dy);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) hf[a][b] *= hd;
                const double f1 = (hf[2][1] - hf[0][1]) / (2.0 * h1);
                const double f2 = (hf[1][2] - hf[1][0]) / (2.0 * h2);
                const double f11 = ((hf[2][1] - 2.0 * hf[1][1]) + hf[0][1]) / (h1 * h1);
                const double f22 = ((hf[1][2] - 2.0 * hf[1][1]) + hf[1][0]) / (h2 * h2);
                // the mixed difference and its denominator are written per direction
                const double f12 = dom == 0 ? (((hf[2][2] - hf[0][2]) - hf[2][0]) + hf[0][0]) / ((4.0 * h1) * h2)
                                 : dom == 1 ? (((hf[2][2] - hf[2][0]) - hf[0][2]) + hf[0][0]) / ((4.0 * h1) * h2)
                                            : (((hf[2][2] - hf[2][0]) - hf[0][2]) + hf[0][0]) / ((4.0 * h2) * h1);
                K = ((((f11 + f22) + f11 * (f2 * f2)) + f22 * (f1 * f1)) - ((2.0 * f12) * f1) * f2) /
                    pow((1.0 + f1 * f1) + f2 * f2, 1.5);
            }
        }
        Curve[q] = K;
    }
}

// mask_maker! (src/Tools.jl:1582-1598): faces next to positive curvature; bits are only set.
__global__ void k_mask(DevMesh m, const double* Curve, int16_t* mask) {
    const int n = m.Nx * m.Ny * m.Nz;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        const bool here = Curve[c] > 0;
        int bits = mask[c];
        if (here || Curve[c - 1] > 0) bits |= 1;
        if (here || Curve[c - sy] > 0) bits |= 2;
        if (here || Curve[c - sz] > 0) bits |= 4;
        mask[c] = (int16_t)bits;
    }
}

__device__ __forceinline__ double csf_face(double sigma, double h, double C0, double C1, double dVF) {
    // C1 = curvature of the cell, C0 = of its lower neighbour
    if (C1 > 0 && C0 > 0) return (((-sigma) / (2.0 * h)) * dVF) * (C1 + C0);
    if (C1 == 0) return (((-sigma) / h) * dVF) * C0;
    if (C0 == 0) return (((-sigma) / h) * dVF) * C1;
    return 0.0;
}

// compute_sf! (src/ELVIRA.jl:256-299) over the whole face arrays (fill!(sf, 0) elsewhere).
__global__ void k_surface_force(DevMesh m, double sigma, const double* VF, const double* Curve,
                                const int16_t* mask, double* sfx, double* sfy, double* sfz) {
    const int ax = m.cnx + 1, ay = m.cny + 1, az = m.cnz + 1;
    const int64_t n = (int64_t)ax * ay * az;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(q % ax);
        const int b = (int)((q / ax) % ay);
        const int cz = (int)(q / ((int64_t)ax * ay));
        const int i = a + 1 - m.ng, j = b + 1 - m.ng, k = cz + 1 - m.ng;
        const bool jin = j >= 1 && j <= m.Ny, kin = k >= 1 && k <= m.Nz, iin = i >= 1 && i <= m.Nx;
        if (b < m.cny && cz < m.cnz) {  // x faces
            double s = 0.0;
            if (i >= 1 && i <= m.Nx + 1 && jin && kin) {
                const int64_t c = m.cell(i, j, k);
                if (mask[c] & 1)
                    s = csf_face(sigma, m.dx, Curve[c - 1], Curve[c], VF[m.vof(i, j, k)] - VF[m.vof(i - 1, j, k)]);
            }
            sfx[a + (int64_t)ax * (b + (int64_t)m.cny * cz)] = s;
        }
        if (a < m.cnx && cz < m.cnz) {  // y faces
            double s = 0.0;
            if (j >= 1 && j <= m.Ny + 1 && iin && kin) {
                const int64_t c = m.cell(i, j, k);
                if (mask[c] & 2)
                    s = csf_face(sigma, m.dy, Curve[c - sy], Curve[c], VF[m.vof(i, j, k)] - VF[m.vof(i, j - 1, k)]);
            }
            sfy[a + (int64_t)m.cnx * (b + (int64_t)ay * cz)] = s;
        }
        if (a < m.cnx && b < m.cny) {  // z faces
            double s = 0.0;
            if (k >= 1 && k <= m.Nz + 1 && iin && jin) {
                const int64_t c = m.cell(i, j, k);
                if (mask[c] & 4)
                    s = csf_face(sigma, m.dz, Curve[c - sz], Curve[c], VF[m.vof(i, j, k)] - VF[m.vof(i, j, k - 1)]);
            }
            sfz[a + (int64_t)m.cnx * (b + (int64_t)m.cny * cz)] = s;
        }
    }
}

// Viscous fluxes from the old cell velocities, surface tension and gravity
// (src/Transport.jl:205-253): us += dt/V (flux differences) + dt (sf / rho_face - g).
__global__ void k_ns_viscous(DevMesh m, double gx, double gy, double gz, double dt, const double* u,
                             const double* v, const double* w, const double* viscx, const double* viscy,
                             const double* viscz, const double* denx, const double* deny, const double* denz,
                             const double* sfx, const double* sfy, const double* sfz, double* us, double* vs,
                             double* ws) {
    const int n = m.Nx * m.Ny * m.Nz;
    const double f = dt / ((m.dx * m.dy) * m.dz);
    const double ax = m.dy * m.dz, ay = m.dx * m.dz, az = m.dx * m.dy;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        // face (i,j,k) and its upper neighbour along each axis
        const int64_t x0 = fxi(m, i, j, k), x1 = x0 + 1;
        const int64_t y0 = fyi(m, i, j, k), y1 = y0 + m.cnx;
        const int64_t z0 = fzi(m, i, j, k), z1 = z0 + sz;
        const double mx0 = sdiv(viscx[x0], denx[x0]), mx1 = sdiv(viscx[x1], denx[x1]);
        const double my0 = sdiv(viscy[y0], deny[y0]), my1 = sdiv(viscy[y1], deny[y1]);
        const double mz0 = sdiv(viscz[z0], denz[z0]), mz1 = sdiv(viscz[z1], denz[z1]);
        const double* old[3] = {u, v, w};
        double* out[3] = {us, vs, ws};
        double src[3];
        src[0] = dt * ((0.5 * (sfx[x0] + sfx[x1])) / (0.5 * (denx[x1] + denx[x0])) - gx);
        src[1] = dt * ((0.5 * (sfy[y0] + sfy[y1])) / (0.5 * (deny[y1] + deny[y0])) - gy);
        src[2] = dt * ((0.5 * (sfz[z0] + sfz[z1])) / (0.5 * (denz[z1] + denz[z0])) - gz);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double* A = old[p];
            const double Fx0 = ax * (mx0 * ((A[c] - A[c - 1]) / m.dx)), Fx1 = ax * (mx1 * ((A[c + 1] - A[c]) / m.dx));
            const double Fy0 = ay * (my0 * ((A[c] - A[c - sy]) / m.dy)), Fy1 = ay * (my1 * ((A[c + sy] - A[c]) / m.dy));
            const double Fz0 = az * (mz0 * ((A[c] - A[c - sz]) / m.dz)), Fz1 = az * (mz1 * ((A[c + sz] - A[c]) / m.dz));
            out[p][c] = (out[p][c] + f * (((((Fx1 - Fx0) + Fy1) - Fy0) + Fz1) - Fz0)) + src[p];
        }
    }
}

// Reflect boundary condition on the two sides of axis d: first ghost layer
// = -interior + 2 U_wall, over the full extents of the other two axes.
__global__ void k_reflect_axis(DevMesh m, int d, double lo0, double lo1, double lo2, double hi0, double hi1,
                               double hi2, double* u, double* v, double* w) {
    const int n0 = d == 0 ? m.cny : m.cnx;
    const int n1 = d == 2 ? m.cny : m.cnz;
    const int64_t tot = (int64_t)n0 * n1;
    const int64_t st = d == 0 ? 1 : (d == 1 ? (int64_t)m.cnx : (int64_t)m.cnx * m.cny);
    const int N = d == 0 ? m.Nx : (d == 1 ? m.Ny : m.Nz);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(q % n0), b = (int)(q / n0);
        int I[3];
        I[d] = m.ng;  // 0-based position of reference index 1 along d
        if (d == 0) { I[1] = a; I[2] = b; }
        else if (d == 1) { I[0] = a; I[2] = b; }
        else { I[0] = a; I[1] = b; }
        const int64_t first = (int64_t)I[0] + (int64_t)m.cnx * ((int64_t)I[1] + (int64_t)m.cny * I[2]);
        const int64_t last = first + (int64_t)(N - 1) * st;
        double* F[3] = {u, v, w};
        const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            F[p][first - st] = -F[p][first] + 2.0 * lo[p];
            F[p][last + st] = -F[p][last] + 2.0 * hi[p];
        }
    }
}

// interpolateFace! (src/Velocity.jl:23-37)
__global__ void k_interp_face(DevMesh m, const double* u, const double* v, const double* w, double* uf,
                              double* vf, double* wf) {
    const int nxr = m.Nx + 3, nyr = m.Ny + 3, nzr = m.Nz + 3;  // reference indices 0..N+2
    const int64_t n = (int64_t)nxr * nyr * nzr;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nxr), j = (int)((q / nxr) % nyr), k = (int)(q / ((int64_t)nxr * nyr));
        const int64_t c = m.cell(i, j, k);
        if (i <= m.Nx + 2 && j <= m.Ny + 1 && k <= m.Nz + 1) uf[c] = 0.5 * (u[c - 1] + u[c]);
        if (i <= m.Nx + 1 && j <= m.Ny + 2 && k <= m.Nz + 1) vf[c] = 0.5 * (v[c - sy] + v[c]);
        if (i <= m.Nx + 1 && j <= m.Ny + 1 && k <= m.Nz + 2) wf[c] = 0.5 * (w[c - sz] + w[c]);
    }
}

// interpolateCenter! (src/Velocity.jl:39-57): u = us outside the interior (the predictor's
// boundary values), the face average inside.
__global__ void k_interp_center(DevMesh m, const double* us, const double* vs, const double* ws, const double* uf,
                                const double* vf, const double* wf, double* u, double* v, double* w) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        if (i >= 1 && i <= m.Nx && j >= 1 && j <= m.Ny && k >= 1 && k <= m.Nz) {
            u[q] = 0.5 * (uf[q] + uf[q + 1]);
            v[q] = 0.5 * (vf[q] + vf[q + sy]);
            w[q] = 0.5 * (wf[q] + wf[q + sz]);
        } else {
            u[q] = us[q];
            v[q] = vs[q];
            w[q] = ws[q];
        }
    }
}

// divergence! / divg_cell (src/Velocity.jl:59-122); ghosts 0 (fill!(divg, 0)). Under the
// semi-Lagrangian scheme pv holds the vertices back-traced with the current face velocities.
__global__ void k_divergence(DevMesh m, int nband, int sl, const int16_t* band, const double* pv,
                             const double* uf, const double* vf, const double* wf, double dt, double* divg) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    const VertPtr pva{&m, pv};
    const FacePtrs fa{uf, vf, wf};
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        if (!(i >= 1 && i <= m.Nx && j >= 1 && j <= m.Ny && k >= 1 && k <= m.Nz)) {
            divg[q] = 0.0;
            continue;
        }
        divg[q] = divg_cell_eval(m, nband, sl, band, pva, fa, dt, q, i, j, k);
    }
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
static inline int ns_grid(int64_t n, int block, int cap = 65536) {
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

#define NSL(kernel, grid, block, ...)                                                    \
    do {                                                                                 \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ctx->stream, __VA_ARGS__); \
        MIST_HIP_TRY(hipGetLastError());                                                 \
    } while (0)

#define FLD(id) ((double*)ctx->f[id])

// First half of transport! with the momentum predictor: the band cells' new VF and us, vs, ws
// from the cut pre-images, and the flux update of the other cells.
int launch_ns_advect(mist_ctx* ctx, double dt) {
    int st;
    if ((st = launch_project_vertices(ctx, dt))) return st;
    const bool fd = ctx->p.pressure_scheme == MIST_PS_FINITE_DIFFERENCE;
    const LeafMomentum lm{FLD(MIST_FIELD_U), FLD(MIST_FIELD_V), FLD(MIST_FIELD_W), ctx->p.rho_liq, ctx->p.rho_gas};
    const int64_t ncap = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    double* sums = ctx->ns_sums;
    const double* nx = (const double*)FLD(MIST_FIELD_NX);
    const double* ny = (const double*)FLD(MIST_FIELD_NY);
    const double* nz = (const double*)FLD(MIST_FIELD_NZ);
    const double* D = (const double*)FLD(MIST_FIELD_PLIC_D);
    const double* uf = (const double*)FLD(MIST_FIELD_UF);
    const double* vf = (const double*)FLD(MIST_FIELD_VF);
    const double* wf = (const double*)FLD(MIST_FIELD_WF);
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_NS_CUT);
    // the five back-traced tets and, under the FD scheme, the twelve correction tets of every
    // band cell: cut trees per work-group writing VF and us, vs, ws; the general walk for the
    // cells they leave
    if ((st = launch_tree_momentum(ctx, fd, dt))) return st;
    if (fd) {
        NSL((k_ns_walk<5, 0, false>), 32, 64, ctx->m, (const double*)ctx->pv, (const int*)ctx->defer_list,
            (const int*)(ctx->counters + 6), nx, ny, nz, D, lm, uf, vf, wf, dt, sums, FLD(MIST_FIELD_VOF),
            FLD(MIST_FIELD_US), FLD(MIST_FIELD_VS), FLD(MIST_FIELD_WS), ctx->counters + 2);
        NSL((k_ns_walk<12, 5, true>), 32, 64, ctx->m, (const double*)ctx->pv, (const int*)ctx->defer_list,
            (const int*)(ctx->counters + 6), nx, ny, nz, D, lm, uf, vf, wf, dt, sums, FLD(MIST_FIELD_VOF),
            FLD(MIST_FIELD_US), FLD(MIST_FIELD_VS), FLD(MIST_FIELD_WS), ctx->counters + 2);
    } else {
        NSL((k_ns_walk<5, 0, true>), 32, 64, ctx->m, (const double*)ctx->pv, (const int*)ctx->defer_list,
            (const int*)(ctx->counters + 6), nx, ny, nz, D, lm, uf, vf, wf, dt, sums, FLD(MIST_FIELD_VOF),
            FLD(MIST_FIELD_US), FLD(MIST_FIELD_VS), FLD(MIST_FIELD_WS), ctx->counters + 2);
    }
    prof_end(ctx, MIST_PROF_NS_CUT, ev, ctx->counters + 0, 0);
    ev = prof_begin(ctx, MIST_PROF_NS_FIELDS);
    NSL(k_ns_convect, ns_grid(ncap, 256), 256, ctx->m, (int)ctx->p.nband, fd ? 0 : 1,
        (const int16_t*)ctx->f[MIST_FIELD_BAND], (const double*)ctx->pv, (const double*)FLD(MIST_FIELD_U),
        (const double*)FLD(MIST_FIELD_V), (const double*)FLD(MIST_FIELD_W), (const double*)FLD(MIST_FIELD_UF),
        (const double*)FLD(MIST_FIELD_VF), (const double*)FLD(MIST_FIELD_WF), dt, FLD(MIST_FIELD_US),
        FLD(MIST_FIELD_VS), FLD(MIST_FIELD_WS));
    prof_end(ctx, MIST_PROF_NS_FIELDS, ev, nullptr, ncap);
    if ((st = launch_neumann(ctx, MIST_FIELD_VOF))) return st;
    return launch_borders(ctx, MIST_FIELD_VOF);
}

int launch_curvature(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    NSL(k_curvature, ns_grid(n, 256), 256, ctx->m, (const double*)FLD(MIST_FIELD_VOF), (const double*)FLD(MIST_FIELD_NX),
        (const double*)FLD(MIST_FIELD_NY), (const double*)FLD(MIST_FIELD_NZ), FLD(MIST_FIELD_CURV));
    return MIST_OK;
}

int launch_surface_force(mist_ctx* ctx) {
    const int ni = ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    NSL(k_mask, ns_grid(ni, 256), 256, ctx->m, (const double*)FLD(MIST_FIELD_CURV), (int16_t*)ctx->f[MIST_FIELD_MASK]);
    const int64_t n = (int64_t)(ctx->m.cnx + 1) * (ctx->m.cny + 1) * (ctx->m.cnz + 1);
    NSL(k_surface_force, ns_grid(n, 256), 256, ctx->m, ctx->p.sigma, (const double*)FLD(MIST_FIELD_VOF),
        (const double*)FLD(MIST_FIELD_CURV), (const int16_t*)ctx->f[MIST_FIELD_MASK], FLD(MIST_FIELD_SFX),
        FLD(MIST_FIELD_SFY), FLD(MIST_FIELD_SFZ));
    return MIST_OK;
}

// Second half of transport!: curvature, mask, surface force, then viscous, surface tension
// and gravity terms (normals of the new VF computed by the caller).
int launch_ns_forces(mist_ctx* ctx, double dt) {
    int st;
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_NS_FIELDS);
    if ((st = launch_curvature(ctx))) return st;
    if ((st = launch_surface_force(ctx))) return st;
    const int ni = ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    NSL(k_ns_viscous, ns_grid(ni, 256), 256, ctx->m, ctx->p.grav_x, ctx->p.grav_y, ctx->p.grav_z, dt,
        (const double*)FLD(MIST_FIELD_U), (const double*)FLD(MIST_FIELD_V), (const double*)FLD(MIST_FIELD_W),
        (const double*)FLD(MIST_FIELD_VISCX), (const double*)FLD(MIST_FIELD_VISCY),
        (const double*)FLD(MIST_FIELD_VISCZ), (const double*)FLD(MIST_FIELD_DENX),
        (const double*)FLD(MIST_FIELD_DENY), (const double*)FLD(MIST_FIELD_DENZ), (const double*)FLD(MIST_FIELD_SFX),
        (const double*)FLD(MIST_FIELD_SFY), (const double*)FLD(MIST_FIELD_SFZ), FLD(MIST_FIELD_US),
        FLD(MIST_FIELD_VS), FLD(MIST_FIELD_WS));
    prof_end(ctx, MIST_PROF_NS_FIELDS, ev, nullptr, ni);
    return MIST_OK;
}

int launch_reflect_bc(mist_ctx* ctx, int fu, int fv, int fw) {
    const int64_t per[3] = {ctx->p.xper, ctx->p.yper, ctx->p.zper};
    const mist::DevMesh& m = ctx->m;
    for (int d = 0; d < 3; ++d) {
        if (per[d]) continue;
        const int64_t tot = (int64_t)(d == 0 ? m.cny : m.cnx) * (d == 2 ? m.cny : m.cnz);
        const double* lo = ctx->bc_wall[2 * d];
        const double* hi = ctx->bc_wall[2 * d + 1];
        NSL(k_reflect_axis, ns_grid(tot, 256), 256, ctx->m, d, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], FLD(fu),
            FLD(fv), FLD(fw));
    }
    return MIST_OK;
}

int launch_interp_face(mist_ctx* ctx, int fu, int fv, int fw) {
    const int64_t n = (int64_t)(ctx->m.Nx + 3) * (ctx->m.Ny + 3) * (ctx->m.Nz + 3);
    NSL(k_interp_face, ns_grid(n, 256), 256, ctx->m, (const double*)FLD(fu), (const double*)FLD(fv),
        (const double*)FLD(fw), FLD(MIST_FIELD_UF), FLD(MIST_FIELD_VF), FLD(MIST_FIELD_WF));
    return MIST_OK;
}

int launch_interp_center(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    NSL(k_interp_center, ns_grid(n, 256), 256, ctx->m, (const double*)FLD(MIST_FIELD_US),
        (const double*)FLD(MIST_FIELD_VS), (const double*)FLD(MIST_FIELD_WS), (const double*)FLD(MIST_FIELD_UF),
        (const double*)FLD(MIST_FIELD_VF), (const double*)FLD(MIST_FIELD_WF), FLD(MIST_FIELD_U), FLD(MIST_FIELD_V),
        FLD(MIST_FIELD_W));
    return MIST_OK;
}

int launch_divergence(mist_ctx* ctx, double dt, bool reuse_vertex_list) {
    const bool sl = ctx->p.pressure_scheme == MIST_PS_SEMI_LAGRANGIAN;
    int st;
    if (sl && (st = launch_project_vertices(ctx, dt, reuse_vertex_list))) return st;
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    NSL(k_divergence, ns_grid(n, 256), 256, ctx->m, (int)ctx->p.nband, sl ? 1 : 0,
        (const int16_t*)ctx->f[MIST_FIELD_BAND], (const double*)ctx->pv, (const double*)FLD(MIST_FIELD_UF),
        (const double*)FLD(MIST_FIELD_VF), (const double*)FLD(MIST_FIELD_WF), dt, FLD(MIST_FIELD_DIVG));
    return MIST_OK;
}

}  // namespace mist
