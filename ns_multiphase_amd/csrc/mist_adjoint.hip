// mist_adjoint.hip -- reverse mode of one SL-VOF transport step (the VOF part of transport!,
// src/Transport.jl:2-268, with computeBand! held fixed): given the adjoint of the new liquid
// fractions, the adjoint of the old ones and of the face velocities. No reference
// counterpart: the reference has no derivative of its transport.
//
// The forward step is VF' = T(VF, uf): normals n(VF) by ELVIRA, plane constants D(n, VF) by
// computeDist, back-traced corners x(uf) by RK4, and in each band cell VF' = vLiq / vol of
// the cut pre-image (elsewhere VF' = VF); then the ghost fills of VF'. The reverse pass runs
// these in the opposite order on the state of the step (its VF, normals, D, band and the
// ELVIRA candidate each interface cell kept):
//   1. ghost fills of VF' reversed (periodic copies, then the one-layer Neumann copies);
//   2. outside the band the adjoint passes through; in each band cell the five tets are cut
//      again, then walked in reverse with seeds dVF'/dvLiq = 1/vol, dVF'/dvol = -vLiq/vol^2:
//      leaf cuts give plane adjoints (n, D) of their cells, mesh-plane cuts pass vertex
//      adjoints up the tree to the cell's back-traced corners;
//   3. the corners' adjoints through the RK4 back-trace to uf, vf, wf;
//   4. the plane-constant adjoints through computeDist to n and VF (after D's periodic copies
//      are reversed), then the normal adjoints through the chosen ELVIRA candidate to VF.
// Adjoints shared by several threads (corners, planes, faces, stencil neighbours) are added
// with atomics, so their rounding depends on the order of arrival.
#include <algorithm>
#include <cmath>

#include "mist_adjoint.h"
#include "mist_internal.h"
#include "mist_kernels.h"
#include "mist_tree.h"

namespace mist {

// ---- reverse of the ghost fills ------------------------------------------------------
struct AdjExt {
    int lo[3];
    int n[3];
    __device__ __forceinline__ int64_t at(int i, int j, int k) const {
        return (int64_t)(i - lo[0]) + (int64_t)n[0] * ((int64_t)(j - lo[1]) + (int64_t)n[1] * (k - lo[2]));
    }
};

// reverse of the one-layer Neumann copy along one axis: ghost[mn-1] = A[mn], ghost[mx+1] = A[mx]
__global__ void k_adj_neumann_axis(double* g, AdjExt e, int d, int mn, int mx) {
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    const int64_t tot = (int64_t)e.n[d1] * e.n[d2];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int id[3];
        id[d1] = e.lo[d1] + (int)(q % e.n[d1]);
        id[d2] = e.lo[d2] + (int)(q / e.n[d1]);
        id[d] = mx + 1;
        const int64_t ghi = e.at(id[0], id[1], id[2]);
        id[d] = mn - 1;
        const int64_t glo = e.at(id[0], id[1], id[2]);
        const double a = g[ghi], b = g[glo];
        g[ghi] = 0.0;
        g[glo] = 0.0;
        id[d] = mx;
        g[e.at(id[0], id[1], id[2])] += a;
        id[d] = mn;
        g[e.at(id[0], id[1], id[2])] += b;
    }
}

// Up to four arrays of the same extents, treated alike by one launch.
struct Arrays4 {
    double* a[4];
    int n;
};

// reverse of k_borders_axis (the statements of each column undone in reverse order)
__global__ void k_adj_borders_axis(Arrays4 G, AdjExt e, int d, int mn, int mx, int mno, int mxo, int ng) {
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    const int64_t tot = (int64_t)e.n[d1] * e.n[d2];
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < tot * G.n;
         r += (int64_t)gridDim.x * blockDim.x) {
        double* const g = G.a[r / tot];
        const int64_t q = r % tot;
        int id[3], is[3];
        id[d1] = is[d1] = e.lo[d1] + (int)(q % e.n[d1]);
        id[d2] = is[d2] = e.lo[d2] + (int)(q / e.n[d1]);
        const int len2 = mn - mno;
        for (int s = len2 - 1; s >= 0; --s) {
            id[d] = mno + s; is[d] = mx - ng + 1 + s;
            const int64_t a = e.at(id[0], id[1], id[2]), b = e.at(is[0], is[1], is[2]);
            const double v = g[a];
            g[a] = 0.0;
            g[b] += v;
        }
        const int len1 = mxo - mx;
        for (int s = 0; s < len1; ++s) {
            id[d] = mx + 1 + s; is[d] = mn + s;
            const int64_t a = e.at(id[0], id[1], id[2]), b = e.at(is[0], is[1], is[2]);
            const double v = g[a];
            g[a] = 0.0;
            g[b] += v;
        }
    }
}

// zero three arrays of lengths n0, n1, n2
__global__ void k_adj_zero3(double* a0, int64_t n0, double* a1, int64_t n1, double* a2, int64_t n2) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n0 + n1 + n2; q += st) {
        if (q < n0) a0[q] = 0.0;
        else if (q < n0 + n1) a1[q - n0] = 0.0;
        else a2[q - n0 - n1] = 0.0;
    }
}

// interior cells outside the band: VF' = VF
__global__ void k_adj_identity(DevMesh m, int nband, const int16_t* band, const double* lam, double* lam_new) {
    const int n = m.Nx * m.Ny * m.Nz;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        if (abs((int)band[m.cell(i, j, k)]) <= nband) continue;
        const int64_t v = m.vof(i, j, k);
        lam_new[v] += lam[v];
    }
}

// ---- reverse cut walk -----------------------------------------------------------------
struct AdjFrame {
    double v[8][3];
    double g[8][3];
    double d[4];
    short vi[2][8][3];
    unsigned char cs, child, nchild, flags, dir;
};

// cutTet of one tet in reverse: the walk of cut_tet, each frame keeping the adjoints of its 8
// vertices; a leaf adds its vertex adjoints to its parent's slots and its plane adjoints to
// the leaf cell (gpl: 4 arrays nx, ny, nz, D); a finished frame passes its cut vertices'
// adjoints to their edge ends, then its 4 vertices' adjoints to the parent. groot receives
// the adjoints of the tet's vertices. Returns false when the walk overflows.
__device__ bool cut_tet_adj(const DevMesh& m, const double* nx, const double* ny, const double* nz, const double* D,
                            double tet[4][3], int ind[4][3], unsigned flags0, AdjFrame* st, double gL, double gV,
                            double groot[4][3], double* const gpl[4]) {
    int top = -1;
    unsigned flags = flags0;
    bool have_node = true;
    int ovf = 0;
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
                const int64_t c = m.cell(ind[0][0], ind[0][1], ind[0][2]);
                double gx[4][3] = {}, gn[3] = {0.0, 0.0, 0.0}, gD = 0.0;
                leaf_adj(tet, nx[c], ny[c], nz[c], D[c], gL, gV, gx, gn, gD);
                if (gn[0] != 0.0) atomicAdd(gpl[0] + c, gn[0]);
                if (gn[1] != 0.0) atomicAdd(gpl[1] + c, gn[1]);
                if (gn[2] != 0.0) atomicAdd(gpl[2] + c, gn[2]);
                if (gD != 0.0) atomicAdd(gpl[3] + c, gD);
                if (top < 0) {
                    for (int v = 0; v < 4; ++v)
                        for (int q = 0; q < 3; ++q) groot[v][q] = gx[v][q];
                    return ovf == 0;
                }
                AdjFrame& P = st[top];
                const unsigned qv = CUT_VTET[P.cs][P.child - 1];
                for (int nn = 0; nn < 4; ++nn)
                    for (int q = 0; q < 3; ++q) P.g[nib(qv, nn)][q] += gx[nn][q];
            } else {
                if (top + 1 >= CUT_MAXLVL) return false;
                AdjFrame& F = st[++top];
                const double* cc = dir == 0 ? m.x : (dir == 1 ? m.y : m.z);
                const double plane = cc[cut_ind];
                double dd[4];
                for (int n = 0; n < 4; ++n) {
                    F.d[n] = tet[n][dir] - plane;
                    dd[n] = F.d[n];
                }
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
                    const double mu = jmin(1.0, jmax(0.0, sdiv(-F.d[v1], F.d[v2] - F.d[v1])));
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
                for (int n = 0; n < 8; ++n)
                    for (int q = 0; q < 3; ++q) F.g[n][q] = 0.0;
                F.cs = (unsigned char)cs;
                F.child = 0;
                F.nchild = CUT_NTETS[cs];
                F.flags = (unsigned char)flags;
                F.dir = (unsigned char)dir;
            }
            have_node = false;
        }
        AdjFrame& F = st[top];
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
            // the mesh-plane cut of this node in reverse: d = x[dir] - plane
            const int nv = CUT_NVERT[F.cs];
            for (int e = nv - 1; e >= 0; --e) {
                const unsigned ed = CUT_EDGE[F.cs][e];
                const int a = ed & 15, b = ed >> 4;
                double gda, gdb;
                lerp_adj(F.v[a], F.v[b], F.d[a], F.d[b], F.g[4 + e], F.g[a], F.g[b], gda, gdb);
                F.g[a][F.dir] += gda;
                F.g[b][F.dir] += gdb;
            }
            double g4[4][3];
            for (int v = 0; v < 4; ++v)
                for (int q = 0; q < 3; ++q) g4[v][q] = F.g[v][q];
            --top;
            if (top < 0) {
                for (int v = 0; v < 4; ++v)
                    for (int q = 0; q < 3; ++q) groot[v][q] = g4[v][q];
                return ovf == 0;
            }
            AdjFrame& P = st[top];
            const unsigned qv = CUT_VTET[P.cs][P.child - 1];
            for (int nn = 0; nn < 4; ++nn)
                for (int q = 0; q < 3; ++q) P.g[nib(qv, nn)][q] += g4[nn][q];
        }
    }
}

// Cells left to the general walk by the tree pass: their five tets' vol and vLiq into sums.
__global__ __launch_bounds__(64) void k_adj_gen_sums(DevMesh m, const double* pv, const int* list, const int* count,
                                                     const double* nx, const double* ny, const double* nz,
                                                     const double* D, double* sums, int* ovf) {
    __shared__ double s_vol[64], s_liq[64];
    const int lane = threadIdx.x;
    const int cl = lane / 5, t = lane - 5 * cl;
    const int nb = *count;
    const LeafMomentum none{nullptr, nullptr, nullptr, 0.0, 0.0};
    for (int base = blockIdx.x * 12; base < nb; base += gridDim.x * 12) {
        const int cidx = base + cl;
        const bool active = lane < 60 && cidx < nb;
        double r[2] = {0.0, 0.0};
        int q = 0;
        if (active) {
            q = list[cidx];
            int i, j, k;
            interior_of(m, q, i, j, k);
            const bool even = ((i + j + k) % 2) == 0;
            const unsigned tv = TET5[t];
            double tet[4][3];
            int ind[4][3], o = 0;
            for (int n = 0; n < 4; ++n) {
                const int c = vert_corner(nib(tv, n), even);
                const int64_t vq = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
                tet[n][0] = pv[3 * vq + 0]; tet[n][1] = pv[3 * vq + 1]; tet[n][2] = pv[3 * vq + 2];
                pt2index(m, ind[n], tet[n], i, j, k, &o);
            }
            CutFrame<2> st[CUT_MAXLVL];
            int depth = 0;
            if (!cut_tet<2>(m, nx, ny, nz, D, none, tet, ind, st, r, depth)) o = 1;
            if (o) atomicOr(ovf, 1);
        }
        s_vol[lane] = r[0];
        s_liq[lane] = r[1];
        __syncthreads();
        if (active && t == 0) {
            double V = 0.0, L = 0.0;
            for (int n = 0; n < 5; ++n) { V += s_vol[lane + n]; L += s_liq[lane + n]; }
            sums[6 * (int64_t)q + 0] = V;
            sums[6 * (int64_t)q + 1] = L;
        }
        __syncthreads();
    }
}

// Band cells in reverse: one lane per (cell, tet, child of the tet's first cut), 30 lanes per
// cell, two cells per block. Each lane replays the first cut of its tet, walks its child's
// subtree in reverse with the seeds of VF' = vLiq / vol (vol and vLiq of the cell from the
// forward sums), takes the child's vertex adjoints back through the first cut to the tet's
// vertices and adds them to the cell's back-traced corners. A tet without cuts is a leaf,
// reversed by its first lane alone.
__global__ __launch_bounds__(64) void k_adj_cut(DevMesh m, const double* pv, const int* list, const int* count,
                                                const double* nx, const double* ny, const double* nz,
                                                const double* D, const double* lam, const double* sums,
                                                double* gpv, double* gnx, double* gny, double* gnz, double* gD,
                                                int* ovf) {
    const int lane = threadIdx.x;
    const int cl = lane / 30, r = lane - 30 * cl, t = r / 6, ch = r - 6 * t;
    const int nb = *count;
    double* const gpl[4] = {gnx, gny, gnz, gD};
    for (int base = blockIdx.x * 2; base < nb; base += gridDim.x * 2) {
        const int cidx = base + cl;
        if (cl >= 2 || cidx >= nb) continue;
        const int q = list[cidx];
        int i, j, k;
        interior_of(m, q, i, j, k);
        const double lv = lam[m.vof(i, j, k)];
        if (lv == 0.0) continue;
        const double V = sums[6 * (int64_t)q + 0], L = sums[6 * (int64_t)q + 1];
        const double gL = lv / V, gV = -lv * L / (V * V);  // VF' = L / V
        const bool even = ((i + j + k) % 2) == 0;
        const unsigned tv = TET5[t];
        double tet[4][3];
        int ind[4][3], o = 0;
        int64_t vq[4];
        for (int n = 0; n < 4; ++n) {
            const int c = vert_corner(nib(tv, n), even);
            vq[n] = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
            tet[n][0] = pv[3 * vq[n] + 0]; tet[n][1] = pv[3 * vq[n] + 1]; tet[n][2] = pv[3 * vq[n] + 2];
            pt2index(m, ind[n], tet[n], i, j, k, &o);
        }
        // the first node of the walk: which plane cuts the tet (src/VFgeom.jl:738-756)
        unsigned flags = 0;
        int dir = -1, cut_ind = 0;
        for (int a = 0; a < 3; ++a) {
            const int mx = max(max(ind[0][a], ind[1][a]), max(ind[2][a], ind[3][a]));
            const int mn = min(min(ind[0][a], ind[1][a]), min(ind[2][a], ind[3][a]));
            if (mx > mn) { dir = a; cut_ind = mx; break; }
            flags |= 1u << a;
        }
        double g4[4][3] = {};
        if (dir < 0) {
            if (ch != 0) continue;
            const int64_t c = m.cell(ind[0][0], ind[0][1], ind[0][2]);
            double gn[3] = {0.0, 0.0, 0.0}, gd = 0.0;
            leaf_adj(tet, nx[c], ny[c], nz[c], D[c], gL, gV, g4, gn, gd);
            atomicAdd(gpl[0] + c, gn[0]);
            atomicAdd(gpl[1] + c, gn[1]);
            atomicAdd(gpl[2] + c, gn[2]);
            atomicAdd(gpl[3] + c, gd);
        } else {
            const double* cc = dir == 0 ? m.x : (dir == 1 ? m.y : m.z);
            const double plane = cc[cut_ind];
            double v[8][3], d[4], dd[4];
            short vi[2][8][3];
            for (int n = 0; n < 4; ++n) {
                d[n] = tet[n][dir] - plane;
                dd[n] = d[n];
            }
            const int cs = d2case_snap(dd);
            if (ch >= CUT_NTETS[cs]) continue;
            for (int n = 0; n < 4; ++n) {
                for (int a = 0; a < 3; ++a) {
                    v[n][a] = tet[n][a];
                    vi[0][n][a] = (short)ind[n][a];
                    vi[1][n][a] = (short)ind[n][a];
                }
                const int lo1 = min(ind[n][dir], cut_ind - 1);
                vi[0][n][dir] = (short)lo1;
                vi[1][n][dir] = (short)max(lo1, cut_ind);
            }
            const int nv = CUT_NVERT[cs];
            for (int n = 0; n < nv; ++n) {
                const unsigned e = CUT_EDGE[cs][n];
                const int v1 = e & 15, v2 = e >> 4;
                const double mu = jmin(1.0, jmax(0.0, sdiv(-d[v1], d[v2] - d[v1])));
                double pt[3];
                for (int a = 0; a < 3; ++a) {
                    pt[a] = (1.0 - mu) * v[v1][a] + mu * v[v2][a];
                    v[4 + n][a] = pt[a];
                }
                int I[3];
                pt2index(m, I, pt, vi[0][n][0], vi[0][n][1], vi[0][n][2], &o);
                for (int a = 0; a < 3; ++a) {
                    int w = max(I[a], (int)min(vi[0][v1][a], vi[0][v2][a]));
                    w = min(w, (int)max(vi[0][v1][a], vi[0][v2][a]));
                    vi[0][4 + n][a] = (short)w;
                    vi[1][4 + n][a] = (short)w;
                }
                vi[0][4 + n][dir] = (short)(cut_ind - 1);
                vi[1][4 + n][dir] = (short)cut_ind;
            }
            // this lane's child: its subtree in reverse
            const unsigned qv = CUT_VTET[cs][ch];
            const int side = (CUT_SIDE2[cs] >> ch) & 1;
            double ct[4][3];
            int ci[4][3];
            for (int nn = 0; nn < 4; ++nn) {
                const int vv = nib(qv, nn);
                for (int a = 0; a < 3; ++a) {
                    ct[nn][a] = v[vv][a];
                    ci[nn][a] = vi[side][vv][a];
                }
            }
            AdjFrame st[CUT_MAXLVL];
            double gc[4][3];
            if (!cut_tet_adj(m, nx, ny, nz, D, ct, ci, flags, st, gL, gV, gc, gpl)) o = 1;
            double g8[8][3] = {};
            for (int nn = 0; nn < 4; ++nn)
                for (int a = 0; a < 3; ++a) g8[nib(qv, nn)][a] += gc[nn][a];
            for (int e = nv - 1; e >= 0; --e) {
                const unsigned ed = CUT_EDGE[cs][e];
                const int a = ed & 15, b = ed >> 4;
                double gda, gdb;
                lerp_adj(v[a], v[b], d[a], d[b], g8[4 + e], g8[a], g8[b], gda, gdb);
                g8[a][dir] += gda;
                g8[b][dir] += gdb;
            }
            for (int n = 0; n < 4; ++n)
                for (int a = 0; a < 3; ++a) g4[n][a] = g8[n][a];
        }
        for (int n = 0; n < 4; ++n)
            for (int a = 0; a < 3; ++a)
                if (g4[n][a] != 0.0) atomicAdd(gpv + 3 * vq[n] + a, g4[n][a]);
        if (o) atomicOr(ovf, 1);
    }
}

// ---- reverse of the cut tree ----------------------------------------------------------
template <class C>
struct TreeAdjShared {
    unsigned par[C::PV];   // how each pool vertex was made (par_pack)
    double g[3][C::PV];    // adjoints of the pool vertices
    double gnb[64][4];     // adjoints of the box cells' nx, ny, nz, D
    int lvl[12];           // pool size at the start of each mesh-plane level, then at the end
    double VL[2];          // the cell's vol and vLiq
};

template <class C>
__device__ __forceinline__ void pool_add(TreeAdjShared<C>& A, unsigned id, const V3& g) {
    if (g.x != 0.0) atomicAdd(&A.g[0][id], g.x);
    if (g.y != 0.0) atomicAdd(&A.g[1][id], g.y);
    if (g.z != 0.0) atomicAdd(&A.g[2][id], g.z);
}

// Band cells in reverse, one work-group per cell as in k_sl_tree: the tree of the cell is built
// again (recording how each pool vertex was made), the leaves give vol and vLiq in the forward
// order, then the leaves are reversed with the seeds of VF' = vLiq / vol (their PLIC cuts give
// the box cells' plane adjoints, their last z cuts pass vertex adjoints to the node's
// vertices), the mesh-plane levels are reversed from the last one down, and the adjoints of
// the 8 back-traced corners and of the planes are added to gpv and gpl. Cells with a zero
// adjoint are skipped; cells that outgrow the pools go to `defer` like the forward pass.
template <class C>
__global__ __launch_bounds__(C::NT) void k_adj_tree(DevMesh m, const double* pv, const CutCell* cut, const int* idx,
                                                  const int* count, const double* nx, const double* ny,
                                                  const double* nz, const double* D, const double* lam,
                                                  double* gpv, double* gnx, double* gny, double* gnz, double* gD,
                                                  int* defer, int* ndefer, int defer_cell, int* queue) {
    constexpr int NT = C::NT;
    __shared__ TreeShared<C> S;
    __shared__ TreeAdjShared<C> A;
    const int tid = threadIdx.x;
    const int nb = *count;
    double* const gpl[4] = {gnx, gny, gnz, gD};
    tab_load<C>(S);
    if (tid == 0) S.next = atomicAdd(queue, 1);
    __syncthreads();
    int cidx = S.next;
    while (cidx < nb) {
        __syncthreads();
        int claim = 0;
        if (tid == 0) claim = atomicAdd(queue, 1);
        const int entry = idx ? idx[cidx] : cidx;
        const CutCell cc = cut[entry];
        const int q = cc.q;
        int i, j, k;
        interior_of(m, q, i, j, k);
        const double lv = lam[m.vof(i, j, k)];
        if (lv != 0.0) {
            const bool even = ((i + j + k) % 2) == 0;
            const int lo[3] = {cc.lo0, cc.lo1, cc.lo2};
            int span[3] = {0, 0, 0};
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int a = 0; a < 3; ++a) span[a] = max(span[a], (int)((cc.offs >> (6 * c + 2 * a)) & 3ull));
            const int bx = span[0] + 1, by = span[1] + 1, bz = span[2] + 1;
            const int ncell = bx * by * bz;
            if (tid < 8) {
                const int64_t vq = vtx_id(m, i + (tid & 1), j + ((tid >> 1) & 1), k + (tid >> 2));
                S.px[tid] = pv[3 * vq + 0];
                S.py[tid] = pv[3 * vq + 1];
                S.pz[tid] = pv[3 * vq + 2];
            }
            if (tid < 15) {
                const int a = tid / 5, off = tid - 5 * a;
                const int la = a == 0 ? lo[0] : (a == 1 ? lo[1] : lo[2]);
                const int sa = a == 0 ? span[0] : (a == 1 ? span[1] : span[2]);
                if (off <= sa + 1) S.plane[a][off] = (a == 0 ? m.x : (a == 1 ? m.y : m.z))[la + off];
            }
            for (int l = tid; l < ncell; l += NT) {
                const int ox = l % bx, oy = (l / bx) % by, oz = l / (bx * by);
                const int64_t c = m.cell(lo[0] + ox, lo[1] + oy, lo[2] + oz);
                const double n0 = nx[c], n1 = ny[c], n2 = nz[c], Dc = D[c];
                S.nb[l][0] = n0;
                S.nb[l][1] = n1;
                S.nb[l][2] = n2;
                S.nb[l][3] = Dc;
                S.cls[l] = (signed char)box_class(n0, n1, n2, Dc, m.x[lo[0] + ox], m.x[lo[0] + ox + 1],
                                                  m.y[lo[1] + oy], m.y[lo[1] + oy + 1], m.z[lo[2] + oz],
                                                  m.z[lo[2] + oz + 1]);
                A.gnb[l][0] = 0.0;
                A.gnb[l][1] = 0.0;
                A.gnb[l][2] = 0.0;
                A.gnb[l][3] = 0.0;
            }
            if (tid < 5) {
                const unsigned tv = TET5[tid];
                unsigned ids[4], bits = 0;
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    const int c = vert_corner(nib(tv, n), even);
                    ids[n] = (unsigned)c;
                    bits |= (unsigned)((cc.offs >> (6 * c)) & 63ull) << (6 * n);
                }
                node_set(S, 0, tid, TreeNode{ids[0] | (ids[1] << 16), ids[2] | (ids[3] << 16), bits});
            }
            __syncthreads();
            // ---- the tree again, recording each pool vertex's edge and plane ---------------
            int n = 5, npool = 8, cur = 0, nl = 0;
            bool ok = true;
            for (int axis = 0; axis < 3 && ok; ++axis) {
                const int minspan = axis == 2 ? 2 : 1;
                for (int rep = 0; rep < 3 && ok; ++rep) {
                    int any = 0;
                    for (int e = tid; e < n; e += NT) {
                        const unsigned bits = node_bits(S, cur, e);
                        const int o0 = boff(bits, 0, axis), o1 = boff(bits, 1, axis), o2 = boff(bits, 2, axis),
                                  o3 = boff(bits, 3, axis);
                        any |= max(max(o0, o1), max(o2, o3)) - min(min(o0, o1), min(o2, o3)) >= minspan;
                    }
                    if (!__syncthreads_or(any)) break;
                    if (tid == 0) A.lvl[nl] = npool;
                    ++nl;
                    ok = tree_level<C>(S, axis, minspan, cur, n, npool, Span3{{span[0], span[1], span[2]}}, A.par);
                    cur ^= 1;
                }
            }
            if (tid == 0) A.lvl[nl] = npool;
            int nleaf = 0;
            for (int b = 0; b < n && ok; b += NT) {
                const int e = b + tid;
                int cnt = 0;
                if (e < n) {
                    const TreeNode N = node_get(S, cur, e);
                    const int o0 = boff(N.bits, 0, 2), o1 = boff(N.bits, 1, 2), o2 = boff(N.bits, 2, 2),
                              o3 = boff(N.bits, 3, 2);
                    const int mx = max(max(o0, o1), max(o2, o3));
                    if (mx > min(min(o0, o1), min(o2, o3))) {
                        const double P = S.plane[2][mx];
                        double dd[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) dd[v] = S.pz[tn_id(N, v)] - P;
                        cnt = S.tab.ntets[d2case_snap(dd)];
                    } else {
                        cnt = 1;
                    }
                }
                int tot;
                const int o = block_scan<NT>(cnt, tot, S.scan);
                if (nleaf + tot > C::NLEAF) {
                    ok = false;
                    break;
                }
                if (e < n)
                    for (int c = 0; c < cnt; ++c) S.lnode[nleaf + o + c] = leaf_entry(e, c);
                nleaf += tot;
            }
            if (!ok) {
                if (tid == 0) defer[atomicAdd(ndefer, 1)] = defer_cell ? q : entry;
            } else {
                for (int v = tid; v < npool; v += NT) {
                    A.g[0][v] = 0.0;
                    A.g[1][v] = 0.0;
                    A.g[2][v] = 0.0;
                }
                __syncthreads();
                // ---- vol and vLiq of the cell, as the forward pass adds them ------------------
                double Vs = 0.0, Ls = 0.0;
                for (int L = tid; L < nleaf; L += NT) {
                    TreeLeaf T;
                    tree_leaf<C>(S, cur, L, span, bx, by, T);
                    const int lcls = S.cls[T.l];
                    auto on_liq = [&](double tv) { Vs += tv; Ls += tv; };
                    auto on_gas = [&](double tv) { Vs += tv; };
                    if (lcls != 0) {
                        const double tv = tet_vol3(T.Q[0], T.Q[1], T.Q[2], T.Q[3]);
                        if (lcls > 0) on_liq(tv);
                        else on_gas(tv);
                        continue;
                    }
                    const double n0 = S.nb[T.l][0], n1 = S.nb[T.l][1], n2 = S.nb[T.l][2], Dc = S.nb[T.l][3];
                    double d[4], dd[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        d[v] = ((n0 * T.Q[v].x + n1 * T.Q[v].y) + n2 * T.Q[v].z) - Dc;
                        dd[v] = d[v];
                    }
                    const int cs = d2case_snap(dd);
                    const double tv = tet_vol3(T.Q[0], T.Q[1], T.Q[2], T.Q[3]);
                    Vs += tv;
                    Ls += cs == 0 ? tv : (cs == 15 ? 0.0 : tv * tet_liquid_fraction(cs, dd));
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    Vs += __shfl_xor(Vs, off);
                    Ls += __shfl_xor(Ls, off);
                }
                if constexpr (NT > 64) {
                    if ((tid & 63) == 0) {
                        S.red[0][tid >> 6] = Vs;
                        S.red[1][tid >> 6] = Ls;
                    }
                    __syncthreads();
                    if (tid == 0) {
                        Vs = S.red[0][0];
                        Ls = S.red[1][0];
                        for (int w = 1; w < NT / 64; ++w) {
                            Vs += S.red[0][w];
                            Ls += S.red[1][w];
                        }
                    }
                }
                if (tid == 0) {
                    A.VL[0] = Vs;
                    A.VL[1] = Ls;
                }
                __syncthreads();
                const double V = A.VL[0], Lq = A.VL[1];
                const double gL = lv / V, gV = -lv * Lq / (V * V);  // VF' = vLiq / vol
                // ---- leaves in reverse ---------------------------------------------------------
                for (int L = tid; L < nleaf; L += NT) {
                    TreeLeaf T;
                    tree_leaf<C>(S, cur, L, span, bx, by, T);
                    const int lcls = S.cls[T.l];
                    V3 gQ[4] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
                    if (lcls != 0) {
                        tet_vol3_grad(T.Q[0], T.Q[1], T.Q[2], T.Q[3], lcls > 0 ? gL + gV : gV, gQ[0], gQ[1], gQ[2],
                                      gQ[3]);
                    } else {
                        const double n0 = S.nb[T.l][0], n1 = S.nb[T.l][1], n2 = S.nb[T.l][2], Dc = S.nb[T.l][3];
                        double d[4], dd[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            d[v] = ((n0 * T.Q[v].x + n1 * T.Q[v].y) + n2 * T.Q[v].z) - Dc;
                            dd[v] = d[v];
                        }
                        const int cs = d2case_snap(dd);
                        double gd[4] = {0.0, 0.0, 0.0, 0.0};
                        cut_canonical_adj(cs, T.Q, d, gL, gV, gQ, gd);
                        // d_v = n . Q_v - D
                        double gn0 = 0.0, gn1 = 0.0, gn2 = 0.0;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            gQ[v].x += gd[v] * n0;
                            gQ[v].y += gd[v] * n1;
                            gQ[v].z += gd[v] * n2;
                            gn0 += gd[v] * T.Q[v].x;
                            gn1 += gd[v] * T.Q[v].y;
                            gn2 += gd[v] * T.Q[v].z;
                        }
                        const double gDc = -(((gd[0] + gd[1]) + gd[2]) + gd[3]);
                        if (gn0 != 0.0) atomicAdd(&A.gnb[T.l][0], gn0);
                        if (gn1 != 0.0) atomicAdd(&A.gnb[T.l][1], gn1);
                        if (gn2 != 0.0) atomicAdd(&A.gnb[T.l][2], gn2);
                        if (gDc != 0.0) atomicAdd(&A.gnb[T.l][3], gDc);
                    }
                    // back through the node's last z cut, or straight to the node's vertices
                    if (T.zcut) {
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) {
                            const int lvx = nib(T.qv, kk);
                            if (lvx < 4) {
                                pool_add(A, tn_id(T.N, lvx), gQ[kk]);
                            } else {
                                const unsigned eg = S.tab.edge[T.cs][lvx - 4];
                                const unsigned ia = tn_id(T.N, eg & 15), ib = tn_id(T.N, eg >> 4);
                                const V3 pa = pool_v(S, ia), pb = pool_v(S, ib);
                                V3 ga{0.0, 0.0, 0.0}, gb{0.0, 0.0, 0.0};
                                double gda, gdb;
                                lerp3_adj(pa, pb, pa.z - T.P, pb.z - T.P, gQ[kk], ga, gb, gda, gdb);
                                ga.z += gda;
                                gb.z += gdb;
                                pool_add(A, ia, ga);
                                pool_add(A, ib, gb);
                            }
                        }
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v) pool_add(A, tn_id(T.N, v), gQ[v]);
                    }
                }
                __syncthreads();
                // ---- mesh-plane levels in reverse ----------------------------------------------
                for (int lev = nl - 1; lev >= 0; --lev) {
                    const int v0 = A.lvl[lev], v1 = A.lvl[lev + 1];
                    for (int v = v0 + tid; v < v1; v += NT) {
                        const V3 g{A.g[0][v], A.g[1][v], A.g[2][v]};
                        if (g.x == 0.0 && g.y == 0.0 && g.z == 0.0) continue;
                        const unsigned pk = A.par[v];
                        const unsigned ia = pk & 2047u, ib = (pk >> 11) & 2047u;
                        const int axis = (int)((pk >> 22) & 3u), mx = (int)(pk >> 24);
                        const double P = S.plane[axis][mx];
                        const V3 pa = pool_v(S, ia), pb = pool_v(S, ib);
                        const double da = (axis == 0 ? pa.x : (axis == 1 ? pa.y : pa.z)) - P;
                        const double db = (axis == 0 ? pb.x : (axis == 1 ? pb.y : pb.z)) - P;
                        V3 ga{0.0, 0.0, 0.0}, gb{0.0, 0.0, 0.0};
                        double gda, gdb;
                        lerp3_adj(pa, pb, da, db, g, ga, gb, gda, gdb);
                        if (axis == 0) { ga.x += gda; gb.x += gdb; }
                        else if (axis == 1) { ga.y += gda; gb.y += gdb; }
                        else { ga.z += gda; gb.z += gdb; }
                        pool_add(A, ia, ga);
                        pool_add(A, ib, gb);
                    }
                    __syncthreads();
                }
                // ---- the back-traced corners and the box cells' planes --------------------------
                if (tid < 24) {
                    const int c = tid / 3, a = tid - 3 * c;
                    const double gg = A.g[a][c];
                    if (gg != 0.0) {
                        const int64_t vq = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
                        atomicAdd(gpv + 3 * vq + a, gg);
                    }
                }
                for (int l = tid; l < ncell; l += NT) {
                    const int ox = l % bx, oy = (l / bx) % by, oz = l / (bx * by);
                    const int64_t c = m.cell(lo[0] + ox, lo[1] + oy, lo[2] + oz);
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const double gg = A.gnb[l][a];
                        if (gg != 0.0) atomicAdd(gpl[a] + c, gg);
                    }
                }
            }
        }
        if (tid == 0) S.next = claim;
        __syncthreads();
        cidx = S.next;
    }
}

// Corners of band cells: the adjoint of the back-traced position through project!
__global__ __launch_bounds__(128) void k_adj_backtrace(DevMesh m, DevParams p, const int16_t* band, const double* uf, const double* vf,
                                const double* wf, double dt, const double* gpv, double* guf, double* gvf,
                                double* gwf) {
    const int64_t n = (int64_t)(m.Nx + 1) * (m.Ny + 1) * (m.Nz + 1);
    const double* const F[3] = {uf, vf, wf};
    double* const G[3] = {guf, gvf, gwf};
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const double g[3] = {gpv[3 * q + 0], gpv[3 * q + 1], gpv[3 * q + 2]};
        if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0) continue;
        const int I = (int)(q % (m.Nx + 1)) + 1;
        const int64_t r = q / (m.Nx + 1);
        const int J = (int)(r % (m.Ny + 1)) + 1;
        const int K = (int)(r / (m.Ny + 1)) + 1;
        const double x[3] = {m.x[I], m.y[J], m.z[K]};
        project_adj(m, p, F, x, min(I, m.Nx), min(J, m.Ny), min(K, m.Nz), dt, g,
                    [&](int c, int64_t idx, double w) {
                        if (w != 0.0) atomicAdd(G[c] + idx, w);
                    });
    }
}

// computeDist of every cell (computePLIC!, src/VF.jl:94-103) in reverse
__global__ void k_adj_plic(DevMesh m, DevParams p, const double* VF, const double* nx, const double* ny,
                           const double* nz, const double* gD, double* gnx, double* gny, double* gnz,
                           double* lam_new) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const double g = gD[q];
        if (g == 0.0) continue;
        int i, j, k;
        cell_of(m, q, i, j, k);
        double gn[3] = {0.0, 0.0, 0.0}, gv = 0.0;
        compute_dist_adj(m, p, i, j, k, nx[q], ny[q], nz[q], VF[m.vof(i, j, k)], g, gn, gv);
        gnx[q] += gn[0];
        gny[q] += gn[1];
        gnz[q] += gn[2];
        if (gv != 0.0) lam_new[m.vof(i, j, k)] += gv;
    }
}

// ELVIRA normals of the interface cells in reverse, through the candidate the forward kept
__global__ void k_adj_elvira(DevMesh m, DevParams p, const double* VF, const int* best, const double* gnx,
                             const double* gny, const double* gnz, double* lam_new) {
    const int n = m.Nx * m.Ny * m.Nz;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        const double g[3] = {gnx[c], gny[c], gnz[c]};
        if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0) continue;
        const double vf = VF[m.vof(i, j, k)];
        if (!(vf >= p.VFlo && vf <= p.VFhi)) continue;  // axis normals: constant
        const int code = best[q];
        if (code < 0) continue;
        // the chosen candidate; where a second one gave the very same normal (a kink, whose
        // side a perturbation's sign picks) the mean of the two, the slope a central
        // difference sees
        const int cand = code & 255, tie = (code >> 8) - 1;
        const double h = tie >= 0 ? 0.5 : 1.0;
        const double gh[3] = {h * g[0], h * g[1], h * g[2]};
        auto at = [&](int di, int dj, int dk) { return VF[m.vof(i + di, j + dj, k + dk)]; };
        auto acc = [&](int di, int dj, int dk, double w) {
            if (w != 0.0) atomicAdd(lam_new + m.vof(i + di, j + dj, k + dk), w);
        };
        elvira_cand_adj(m, i, j, k, cand, at, gh, acc);
        if (tie >= 0) elvira_cand_adj(m, i, j, k, tie, at, gh, acc);
    }
}

// ------------------------------------------------------------------------------------
// launcher
// ------------------------------------------------------------------------------------
static inline int adj_grid(int64_t n, int block, int cap = 65536) {
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

#define ADJL(kernel, grid, block, ...)                                                   \
    do {                                                                                 \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ctx->stream, __VA_ARGS__); \
        MIST_HIP_TRY(hipGetLastError());                                                 \
    } while (0)

static AdjExt adj_ext(mist_ctx* ctx, int field) {
    AdjExt e;
    for (int d = 0; d < 3; ++d) { e.lo[d] = (int)ctx->flo[field][d]; e.n[d] = (int)ctx->fn[field][d]; }
    return e;
}

// reverse of launch_borders on up to four double arrays with the extents of `field`
static int adj_borders(mist_ctx* ctx, Arrays4 G, int field) {
    const AdjExt e = adj_ext(ctx, field);
    const int64_t per[3] = {ctx->p.xper, ctx->p.yper, ctx->p.zper};
    const int mx[3] = {ctx->m.Nx, ctx->m.Ny, ctx->m.Nz};
    const int ng = ctx->m.ng;
    for (int d = 2; d >= 0; --d) {
        if (!per[d]) continue;
        const int64_t tot = (int64_t)e.n[(d + 1) % 3] * e.n[(d + 2) % 3];
        ADJL(k_adj_borders_axis, adj_grid(tot * G.n, 256), 256, G, e, d, 1, mx[d], 1 - ng, mx[d] + ng, ng);
    }
    return MIST_OK;
}

static int adj_neumann(mist_ctx* ctx, double* g, int field) {
    const AdjExt e = adj_ext(ctx, field);
    const int mx[3] = {ctx->m.Nx, ctx->m.Ny, ctx->m.Nz};
    for (int d = 2; d >= 0; --d) {
        const int64_t tot = (int64_t)e.n[(d + 1) % 3] * e.n[(d + 2) % 3];
        ADJL(k_adj_neumann_axis, adj_grid(tot, 256), 256, g, e, d, 1, mx[d]);
    }
    return MIST_OK;
}

// Resident work-groups of the reverse cut pass (queried once; outside any graph capture).
int vjp_prepare(mist_ctx* ctx) {
    if (!ctx->tree_blocks_adj) {
        int per = 0, dev = 0, cus = 0;
        MIST_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_adj_tree<TreeMain>, TreeMain::NT, 0));
        MIST_HIP_TRY(hipGetDevice(&dev));
        MIST_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ctx->tree_blocks_adj = std::max(1, per) * std::max(1, cus);
        if (!ctx->tree_cus) ctx->tree_cus = std::max(1, cus);
    }
    return MIST_OK;
}

int launch_transport_vjp(mist_ctx* ctx, double dt) {
    int st;
    const mist::DevMesh& m = ctx->m;
    const int64_t ncell_ext = (int64_t)m.cnx * m.cny * m.cnz;
    const int64_t nvof = (int64_t)m.vnx * m.vny * m.vnz;
    const int64_t nv = (int64_t)(m.Nx + 1) * (m.Ny + 1) * (m.Nz + 1);
    const int64_t ncap = (int64_t)m.Nx * m.Ny * m.Nz;
    double* lam = (double*)ctx->f[MIST_FIELD_VOF_BAR];
    double* lam_new = ctx->adj_lam;
    double* gpl[4] = {ctx->adj_plane, ctx->adj_plane + ncell_ext, ctx->adj_plane + 2 * ncell_ext,
                      ctx->adj_plane + 3 * ncell_ext};
    // pre-images of this state's band cells under the current face velocities
    if ((st = launch_project_vertices(ctx, dt))) return st;
    // 1. ghost fills of VF' in reverse: periodic copies, then Neumann
    if ((st = adj_borders(ctx, Arrays4{{lam, nullptr, nullptr, nullptr}, 1}, MIST_FIELD_VOF))) return st;
    if ((st = adj_neumann(ctx, lam, MIST_FIELD_VOF))) return st;
    ADJL(k_adj_zero3, adj_grid(nvof + 4 * ncell_ext + 3 * nv, 256), 256, lam_new, nvof, ctx->adj_plane,
         4 * ncell_ext, ctx->adj_pv, 3 * nv);
    // 2. identity outside the band; the band cells' cut in reverse
    ADJL(k_adj_identity, adj_grid(ncap, 256), 256, m, (int)ctx->p.nband, (const int16_t*)ctx->f[MIST_FIELD_BAND],
         (const double*)lam, lam_new);
    const double* nx = (const double*)ctx->f[MIST_FIELD_NX];
    const double* ny = (const double*)ctx->f[MIST_FIELD_NY];
    const double* nz = (const double*)ctx->f[MIST_FIELD_NZ];
    const double* D = (const double*)ctx->f[MIST_FIELD_PLIC_D];
    // the band cells' cut trees in reverse: bulk cells (every leaf on one side of its plane)
    // have a zero adjoint and stay out of the cut list; one work-group per cell, the wider pass
    // for trees that outgrow the first, the general walk for the rest
    if ((st = launch_cut_list(ctx))) return st;
    if ((st = vjp_prepare(ctx))) return st;
    const CutCell* cut = (const CutCell*)ctx->cut_list;
    ADJL(k_adj_tree<TreeMain>, ctx->tree_blocks_adj, TreeMain::NT, m, (const double*)ctx->pv, cut,
         (const int*)nullptr, (const int*)(ctx->counters + 8), nx, ny, nz, D, (const double*)lam, ctx->adj_pv, gpl[0],
         gpl[1], gpl[2], gpl[3], ctx->iface_list, ctx->counters + 4, 0, ctx->counters + 5);
    ADJL(k_adj_tree<TreeWide>, ctx->tree_cus, TreeWide::NT, m, (const double*)ctx->pv, cut,
         (const int*)ctx->iface_list, (const int*)(ctx->counters + 4), nx, ny, nz, D, (const double*)lam, ctx->adj_pv,
         gpl[0], gpl[1], gpl[2], gpl[3], ctx->defer_list, ctx->counters + 6, 1, ctx->counters + 7);
    ADJL(k_adj_gen_sums, 32, 64, m, (const double*)ctx->pv, (const int*)ctx->defer_list,
         (const int*)(ctx->counters + 6), nx, ny, nz, D, ctx->adj_sums, ctx->counters + 2);
    ADJL(k_adj_cut, 32, 64, m, (const double*)ctx->pv, (const int*)ctx->defer_list, (const int*)(ctx->counters + 6),
         nx, ny, nz, D, (const double*)lam, (const double*)ctx->adj_sums, ctx->adj_pv, gpl[0], gpl[1], gpl[2], gpl[3],
         ctx->counters + 2);
    // 3. the back-trace in reverse, into the face-velocity adjoints
    ADJL(k_adj_backtrace, adj_grid(nv, 128), 128, m, ctx->dp, (const int16_t*)ctx->f[MIST_FIELD_BAND],
         (const double*)ctx->f[MIST_FIELD_UF], (const double*)ctx->f[MIST_FIELD_VF],
         (const double*)ctx->f[MIST_FIELD_WF], dt, (const double*)ctx->adj_pv, (double*)ctx->f[MIST_FIELD_UF_BAR],
         (double*)ctx->f[MIST_FIELD_VFACE_BAR], (double*)ctx->f[MIST_FIELD_WF_BAR]);
    // 4. computePLIC! (after D's periodic copies) and ELVIRA (after the normals' copies)
    if ((st = adj_borders(ctx, Arrays4{{gpl[3], nullptr, nullptr, nullptr}, 1}, MIST_FIELD_PLIC_D))) return st;
    ADJL(k_adj_plic, adj_grid(ncell_ext, 256), 256, m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF], nx, ny, nz,
         (const double*)gpl[3], gpl[0], gpl[1], gpl[2], lam_new);
    if ((st = adj_borders(ctx, Arrays4{{gpl[0], gpl[1], gpl[2], nullptr}, 3}, MIST_FIELD_NX))) return st;
    ADJL(k_adj_elvira, adj_grid(ncap, 256), 256, m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF],
         (const int*)ctx->elv_best, (const double*)gpl[0], (const double*)gpl[1], (const double*)gpl[2], lam_new);
    MIST_HIP_TRY(hipMemcpyAsync(lam, lam_new, nvof * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return MIST_OK;
}

}  // namespace mist
