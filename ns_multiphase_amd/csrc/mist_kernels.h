// mist_kernels.h -- device helpers shared by the kernels of mist_vof.hip and mist_ns.hip:
// cell indexing, work-group-aggregated appends, the back-trace of a point, and cutTet as an
// explicit depth-first walk carrying either the two volume sums of the VOF transport or the
// six sums of the momentum predictor.
#pragma once

#include "mist_internal.h"

namespace mist {

constexpr double PI = 3.141592653589793;
constexpr int CUT_MAXLVL = 16;

__device__ __forceinline__ void cell_of(const DevMesh& m, int64_t q, int& i, int& j, int& k) {
    const int a = (int)(q % m.cnx);
    const int64_t r = q / m.cnx;
    const int b = (int)(r % m.cny);
    const int c = (int)(r / m.cny);
    i = a + 1 - m.ng; j = b + 1 - m.ng; k = c + 1 - m.ng;
}
__device__ __forceinline__ void interior_of(const DevMesh& m, int q, int& i, int& j, int& k) {
    i = q % m.Nx + 1;
    const int r = q / m.Nx;
    j = r % m.Ny + 1;
    k = r / m.Ny + 1;
}
__device__ __forceinline__ int interior_id(const DevMesh& m, int i, int j, int k) {
    return (i - 1) + m.Nx * ((j - 1) + m.Ny * (k - 1));
}

// Work-group-aggregated append: the waves take their slices of the group's range in LDS and
// one thread adds the group's total to *count, so a sparse list over a large grid costs one
// global atomic per group instead of one per wave. Every thread of the group must call it the
// same number of times (it synchronises the group); s is 2 ints of LDS.
__device__ __forceinline__ void append_group(bool flag, int item, int* list, int* count, int* s) {
    const unsigned long long mask = __ballot(flag);
    const int lane = __lane_id();
    if (threadIdx.x == 0) s[0] = 0;
    __syncthreads();
    int woff = 0;
    if (lane == 0 && mask) woff = atomicAdd(&s[0], __popcll(mask));
    woff = __shfl(woff, 0);
    __syncthreads();
    if (threadIdx.x == 0) s[1] = s[0] ? atomicAdd(count, s[0]) : 0;
    __syncthreads();
    if (flag) list[s[1] + woff + __popcll(mask & ((1ull << lane) - 1ull))] = item;
    __syncthreads();  // s is reused by the next call
}

// ------------------------------------------------------------------------------------
// Back-trace of cell vertices (project!, src/Tools.jl:392-418). The back-traced position of
// a vertex does not depend on which cell asks for it (the interpolation index search
// converges to the same cell from any start), so each vertex of a band cell is traced once.
// ------------------------------------------------------------------------------------
template <class FA>
__host__ __device__ inline void project_pt_acc(const DevMesh& m, const DevParams& p, const FA& fa, double pt[3], int i,
                                               int j, int k, double dt) {
    if (p.zalesak) {
        const double pi2 = 2.0 * PI;
        const double x0 = pt[0], y0 = pt[1];
        pt[0] = (0.5 + cos(pi2 * dt) * (x0 - 0.5)) + sin(pi2 * dt) * (y0 - 0.5);
        pt[1] = (0.5 - sin(pi2 * dt) * (x0 - 0.5)) + cos(pi2 * dt) * (y0 - 0.5);
        return;
    }
    double v1[3], v2[3], v3[3], v4[3], q[3];
    if (p.projection == MIST_PROJ_RK4) {
        vel_face_acc(m, fa, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v1[n];
        vel_face_acc(m, fa, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v2[n];
        vel_face_acc(m, fa, q, i, j, k, v3);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - dt * v3[n];
        vel_face_acc(m, fa, q, i, j, k, v4);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + ((-dt) / 6.0) * (((v1[n] + 2.0 * v2[n]) + 2.0 * v3[n]) + v4[n]);
    } else if (p.projection == MIST_PROJ_EULER) {
        vel_face_acc(m, fa, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + (-dt) * v1[n];
    } else {
        vel_face_acc(m, fa, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v1[n];
        vel_face_acc(m, fa, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + (-dt) * v2[n];
    }
}

__host__ __device__ inline void project_pt(const DevMesh& m, const DevParams& p, const double* uf, const double* vf,
                           const double* wf, double pt[3], int i, int j, int k, double dt) {
    project_pt_acc(m, p, FacePtrs{uf, vf, wf}, pt, i, j, k, dt);
}

__device__ __forceinline__ int64_t vtx_id(const DevMesh& m, int I, int J, int K) {
    return (int64_t)(I - 1) + (int64_t)(m.Nx + 1) * ((int64_t)(J - 1) + (int64_t)(m.Ny + 1) * (K - 1));
}

// ------------------------------------------------------------------------------------
// cutTet (src/VFgeom.jl:713-900) as an explicit depth-first walk. Each frame keeps what a
// recursion level of the reference keeps in vert[:,:,lvl] and vert_ind[:,:,:,lvl], plus the
// running sums of its children, so the sums nest exactly as the recursion adds them.
// NS = 2 sums (vol, vLiq) or 6 (vol, vLiq, vrhoU, vrhoV, vrhoW, vrho: src/VFgeom.jl:791-827,
// with the cell-centre velocity and the phase density of the leaf cell).
// ------------------------------------------------------------------------------------
template <int NS>
struct CutFrame {
    double v[8][3];
    short vi[2][8][3];
    double acc[NS];
    unsigned char cs, child, nchild, flags;
};

// Velocity and densities the momentum sums read at a leaf (unused for NS = 2).
struct LeafMomentum {
    const double* u;
    const double* v;
    const double* w;
    double rho_liq, rho_gas;
};

template <int NS>
__device__ __forceinline__ void cut_leaf(const DevMesh& m, const double* nx, const double* ny, const double* nz,
                                         const double* D, const LeafMomentum& lm, const double tet[4][3],
                                         const int ind0[3], double out[NS]) {
    const int64_t c = m.cell(ind0[0], ind0[1], ind0[2]);
    const double n0 = nx[c], n1 = ny[c], n2 = nz[c], Dc = D[c];
    double vert[8][3], d[4], dd[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        vert[n][0] = tet[n][0]; vert[n][1] = tet[n][1]; vert[n][2] = tet[n][2];
        d[n] = ((n0 * vert[n][0] + n1 * vert[n][1]) + n2 * vert[n][2]) - Dc;
        dd[n] = d[n];
    }
    const int cs = d2case_snap(dd);  // the reference snaps a copy here
    const int nv = CUT_NVERT[cs];
    for (int n = 0; n < nv; ++n) {
        const unsigned e = CUT_EDGE[cs][n];
        const int v1 = e & 15, v2 = e >> 4;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-d[v1], d[v2] - d[v1])));
        vert[4 + n][0] = (1.0 - mu) * vert[v1][0] + mu * vert[v2][0];
        vert[4 + n][1] = (1.0 - mu) * vert[v1][1] + mu * vert[v2][1];
        vert[4 + n][2] = (1.0 - mu) * vert[v1][2] + mu * vert[v2][2];
    }
    double S[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) S[q] = 0.0;
    double uc = 0.0, vc = 0.0, wc = 0.0;
    if (NS == 6) { uc = lm.u[c]; vc = lm.v[c]; wc = lm.w[c]; }
    const int hi = CUT_NTETS[cs], lo = CUT_NNTET[cs];
    for (int n = hi; n >= lo; --n) {  // liquid side
        const unsigned q = CUT_VTET[cs][n - 1];
        const double tv = tet_vol(vert[nib(q, 0)], vert[nib(q, 1)], vert[nib(q, 2)], vert[nib(q, 3)]);
        S[0] += tv; S[1] += tv;
        if (NS == 6) {
            S[2] += (tv * uc) * lm.rho_liq; S[3] += (tv * vc) * lm.rho_liq; S[4] += (tv * wc) * lm.rho_liq;
            S[5] += tv * lm.rho_liq;
        }
    }
    for (int n = 1; n <= lo - 1; ++n) {  // gas side
        const unsigned q = CUT_VTET[cs][n - 1];
        const double tv = tet_vol(vert[nib(q, 0)], vert[nib(q, 1)], vert[nib(q, 2)], vert[nib(q, 3)]);
        S[0] += tv;
        if (NS == 6) {
            S[2] += (tv * uc) * lm.rho_gas; S[3] += (tv * vc) * lm.rho_gas; S[4] += (tv * wc) * lm.rho_gas;
            S[5] += tv * lm.rho_gas;
        }
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) out[q] = S[q];
}

// Returns false when the walk needs more than CUT_MAXLVL levels or a walk leaves the mesh.
template <int NS>
__device__ inline bool cut_tet(const DevMesh& m, const double* nx, const double* ny, const double* nz,
                               const double* D, const LeafMomentum& lm, double tet[4][3], int ind[4][3],
                               CutFrame<NS>* st, double out[NS], int& depth) {
    int top = -1;
    unsigned flags = 0;  // bit 0/1/2: x/y/z cuts done
    bool have_node = true;
    int ovf = 0;
    depth = 0;
    for (;;) {
        if (have_node) {
            int dir = -1, cut_ind = 0;
            for (int a = 0; a < 3; ++a) {
                if (flags & (1u << a)) continue;
                const int mx = max(max(ind[0][a], ind[1][a]), max(ind[2][a], ind[3][a]));
                const int mn = min(min(ind[0][a], ind[1][a]), min(ind[2][a], ind[3][a]));
                if (mx > mn) { dir = a; cut_ind = mx; break; }
                flags |= 1u << a;
            }
            if (dir < 0) {
                double lf[NS];
                cut_leaf<NS>(m, nx, ny, nz, D, lm, tet, ind[0], lf);
                if (top < 0) {
                    for (int q = 0; q < NS; ++q) out[q] = lf[q];
                    return ovf == 0;
                }
                for (int q = 0; q < NS; ++q) st[top].acc[q] += lf[q];
            } else {
                if (top + 1 >= CUT_MAXLVL) return false;
                CutFrame<NS>& F = st[++top];
                depth = max(depth, top + 1);
                const double* cc = dir == 0 ? m.x : (dir == 1 ? m.y : m.z);
                const double plane = cc[cut_ind];
                double d[4], dd[4];
                for (int n = 0; n < 4; ++n) { d[n] = tet[n][dir] - plane; dd[n] = d[n]; }
                const int cs = d2case_snap(dd);
                for (int n = 0; n < 4; ++n) {
                    for (int q = 0; q < 3; ++q) {
                        F.v[n][q] = tet[n][q];
                        F.vi[0][n][q] = (short)ind[n][q];
                        F.vi[1][n][q] = (short)ind[n][q];
                    }
                    const int lo1 = min(ind[n][dir], cut_ind - 1);
                    F.vi[0][n][dir] = (short)lo1;
                    F.vi[1][n][dir] = (short)max(lo1, cut_ind);
                }
                const int nv = CUT_NVERT[cs];
                for (int n = 0; n < nv; ++n) {
                    const unsigned e = CUT_EDGE[cs][n];
                    const int v1 = e & 15, v2 = e >> 4;
                    const double mu = jmin(1.0, jmax(0.0, sdiv(-d[v1], d[v2] - d[v1])));
                    double pt[3];
                    for (int q = 0; q < 3; ++q) {
                        pt[q] = (1.0 - mu) * F.v[v1][q] + mu * F.v[v2][q];
                        F.v[4 + n][q] = pt[q];
                    }
                    int I[3];
                    pt2index(m, I, pt, F.vi[0][n][0], F.vi[0][n][1], F.vi[0][n][2], &ovf);
                    for (int q = 0; q < 3; ++q) {
                        int w = max(I[q], (int)min(F.vi[0][v1][q], F.vi[0][v2][q]));
                        w = min(w, (int)max(F.vi[0][v1][q], F.vi[0][v2][q]));
                        F.vi[0][4 + n][q] = (short)w;
                        F.vi[1][4 + n][q] = (short)w;
                    }
                    F.vi[0][4 + n][dir] = (short)(cut_ind - 1);
                    F.vi[1][4 + n][dir] = (short)cut_ind;
                }
                F.cs = (unsigned char)cs;
                F.child = 0;
                F.nchild = CUT_NTETS[cs];
                F.flags = (unsigned char)flags;
                for (int q = 0; q < NS; ++q) F.acc[q] = 0.0;
            }
            have_node = false;
        }
        CutFrame<NS>& F = st[top];
        if (F.child < F.nchild) {
            const int n = F.child++;
            const unsigned q = CUT_VTET[F.cs][n];
            const int side = (CUT_SIDE2[F.cs] >> n) & 1;
            for (int nn = 0; nn < 4; ++nn) {
                const int v = nib(q, nn);
                for (int a = 0; a < 3; ++a) {
                    tet[nn][a] = F.v[v][a];
                    ind[nn][a] = F.vi[side][v][a];
                }
            }
            flags = F.flags;
            have_node = true;
        } else {
            double acc[NS];
            for (int q = 0; q < NS; ++q) acc[q] = F.acc[q];
            --top;
            if (top < 0) {
                for (int q = 0; q < NS; ++q) out[q] = acc[q];
                return ovf == 0;
            }
            for (int q = 0; q < NS; ++q) st[top].acc[q] += acc[q];
        }
    }
}

}  // namespace mist
