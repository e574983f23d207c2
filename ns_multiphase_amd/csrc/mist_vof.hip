// mist_vof.hip -- gfx950 kernels of the SL-VOF transport stage and its helpers:
// prescribed velocity, ghost fills, band flags, ELVIRA normals, PLIC constants,
// back-trace of cell vertices, iterative tet cutting, initial conditions, densities,
// reductions. Launchers at the bottom are called by the C ABI (mist_capi.hip).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mist_internal.h"
#include "mist_kernels.h"
#include "mist_tree.h"
#include "mist_nsgeom.h"

namespace mist {

// ------------------------------------------------------------------------------------
// defineVelocity! (src/Tools.jl:648-693), period T of cos(pi t / T) as a parameter
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void vel_fun(int kind, double T, double x, double y, double z, double t, double& u,
                                        double& v, double& w) {
    const double twopi = 2.0 * PI;
    switch (kind) {
        case MIST_VEL_DEFORMATION: {
            const double sx = sin(PI * x), sy = sin(PI * y);
            u = ((-2.0 * (sx * sx) * sin(PI * y)) * cos(PI * y)) * cos((PI * t) / T);
            v = ((2.0 * (sy * sy) * sin(PI * x)) * cos(PI * x)) * cos((PI * t) / T);
            w = 0.0;
            break;
        }
        case MIST_VEL_DEFORMATION3D: {
            const double sx = sin(PI * x), sy = sin(PI * y), sz = sin(PI * z);
            u = ((2.0 * (sx * sx) * sin(twopi * y)) * sin(twopi * z)) * cos((PI * t) / T);
            v = ((-(sy * sy) * sin(twopi * x)) * sin(twopi * z)) * cos((PI * t) / T);
            w = ((-(sz * sz) * sin(twopi * x)) * sin(twopi * y)) * cos((PI * t) / T);
            break;
        }
        case MIST_VEL_ROTATION: u = twopi * (0.5 - y); v = twopi * (x - 0.5); w = 0.0; break;
        case MIST_VEL_DIVFLOW: u = 2.0 * x; v = -1.0 * (2.5 - y); w = 0.0; break;
        case MIST_VEL_DIVFLOW3D: u = 200.0 * x; v = -100.0 * (5.0 - y); w = -50.0 * (5.0 - z); break;
        case MIST_VEL_AIRBLAST: u = 50.0 * (1.0 - exp(-100.0 * fabs(y - 0.5))); v = 0.0; w = 0.0; break;
        default: u = v = w = 0.0;
    }
}

// sin(pi c), sin(2 pi c) and cos(pi c) of every coordinate entry, computed once per context:
// the prescribed fields are products of these per-axis factors, so a step multiplies table
// entries in the reference's order instead of evaluating ~24 FP64 trig calls per cell.
__global__ void k_trig_tables(const double* g, int64_t n, double* t) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const double c = g[q];
        t[q] = sin(PI * c);
        t[n + q] = sin((2.0 * PI) * c);
        t[2 * n + q] = cos(PI * c);
    }
}

// vel_fun at (x, y, z) = (faces or centres per axis: CX, CY, CZ = 1 for centres) of cell (i, j, k)
template <int CX, int CY, int CZ>
__device__ __forceinline__ void vel_at(int kind, double T, double t, double ct, const DevMesh& m, const VelTab& V,
                                       int i, int j, int k, double& u, double& v, double& w) {
    constexpr int X = CX ? 3 : 0, Y = CY ? 4 : 1, Z = CZ ? 5 : 2;
    if (kind == MIST_VEL_DEFORMATION) {
        const double sx = V.s1[X][i], sy = V.s1[Y][j];
        u = ((-2.0 * (sx * sx) * V.s1[Y][j]) * V.c1[Y][j]) * ct;
        v = ((2.0 * (sy * sy) * V.s1[X][i]) * V.c1[X][i]) * ct;
        w = 0.0;
    } else if (kind == MIST_VEL_DEFORMATION3D) {
        const double sx = V.s1[X][i], sy = V.s1[Y][j], sz = V.s1[Z][k];
        u = ((2.0 * (sx * sx) * V.s2[Y][j]) * V.s2[Z][k]) * ct;
        v = ((-(sy * sy) * V.s2[X][i]) * V.s2[Z][k]) * ct;
        w = ((-(sz * sz) * V.s2[X][i]) * V.s2[Y][j]) * ct;
    } else {
        vel_fun(kind, T, CX ? m.xm[i] : m.x[i], CY ? m.ym[j] : m.y[j], CZ ? m.zm[k] : m.z[k], t, u, v, w);
    }
}

// With `part` (blocks of 256 threads), each block also writes the maxima of |u|, |v|, |w| over
// its cells to part[3 blockIdx + c] (what k_maxabs3 gives: a maximum does not depend on order).
__global__ __launch_bounds__(256) void k_define_velocity(DevMesh m, VelTab V, int kind, double T, double t, double* u,
                                                         double* v, double* w, double* uf, double* vf, double* wf,
                                                         double* part) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    const double ct = cos((PI * t) / T);
    double r[3] = {0.0, 0.0, 0.0};
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        double a, b, c;
        vel_at<1, 1, 1>(kind, T, t, ct, m, V, i, j, k, a, b, c);
        u[q] = a; v[q] = b; w[q] = c;
        r[0] = jmax(r[0], fabs(a)); r[1] = jmax(r[1], fabs(b)); r[2] = jmax(r[2], fabs(c));
        vel_at<0, 1, 1>(kind, T, t, ct, m, V, i, j, k, a, b, c);
        uf[q] = a;
        vel_at<1, 0, 1>(kind, T, t, ct, m, V, i, j, k, a, b, c);
        vf[q] = b;
        vel_at<1, 1, 0>(kind, T, t, ct, m, V, i, j, k, a, b, c);
        wf[q] = c;
    }
    if (part) {  // wave maxima by shuffles, then the block's four waves
        __shared__ double s_w[4][3];
        for (int c = 0; c < 3; ++c)
            for (int off = 32; off >= 1; off >>= 1) r[c] = jmax(r[c], __shfl_xor(r[c], off));
        if ((threadIdx.x & 63) == 0)
            for (int c = 0; c < 3; ++c) s_w[threadIdx.x >> 6][c] = r[c];
        __syncthreads();
        if (threadIdx.x < 3) {
            const int c = threadIdx.x;
            part[3 * blockIdx.x + c] = jmax(jmax(s_w[0][c], s_w[1][c]), jmax(s_w[2][c], s_w[3][c]));
        }
    }
}

// ------------------------------------------------------------------------------------
// Reductions (block tree in LDS, fixed order; then one block over the partials)
// ------------------------------------------------------------------------------------
constexpr int RED_THREADS = 256;

__device__ __forceinline__ double red_op(int kind, double a, double b) {
    switch (kind) {
        case MIST_REDUCE_SUM: return a + b;
        case MIST_REDUCE_MIN: return jmin(a, b);
        default: return jmax(a, b);
    }
}
__device__ __forceinline__ double red_init(int kind) {
    return kind == MIST_REDUCE_SUM ? 0.0 : (kind == MIST_REDUCE_MIN ? __builtin_huge_val() : -__builtin_huge_val());
}

__device__ double block_reduce(int kind, double v) {
    __shared__ double s[RED_THREADS];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = red_op(kind, s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// max |a| over whole arrays (compute_dt takes maxima including ghosts)
__global__ __launch_bounds__(RED_THREADS) void k_maxabs3(const double* a0, const double* a1, const double* a2,
                                                         int64_t n, double* part) {
    double r[3] = {0.0, 0.0, 0.0};
    const double* a[3] = {a0, a1, a2};
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) r[c] = jmax(r[c], fabs(a[c][q]));
    for (int c = 0; c < 3; ++c) {
        const double b = block_reduce(MIST_REDUCE_MAX, r[c]);
        if (threadIdx.x == 0) part[3 * blockIdx.x + c] = b;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_interior(DevMesh m, const double* A, int vof, int kind,
                                                                 double* part) {
    const int64_t n = (int64_t)m.Nx * m.Ny * m.Nz;
    double r = red_init(kind);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, (int)q, i, j, k);
        const double v = A[vof ? m.vof(i, j, k) : m.cell(i, j, k)];
        r = red_op(kind, r, kind == MIST_REDUCE_MAXABS ? fabs(v) : v);
    }
    const double b = block_reduce(kind == MIST_REDUCE_MAXABS ? MIST_REDUCE_MAX : kind, r);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// block b reduces the n partials part[b + q * stride] into out[b]
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partials(const double* part, int n, int stride, int kind,
                                                                 double* out) {
    double r = red_init(kind);
    for (int q = threadIdx.x; q < n; q += blockDim.x) r = red_op(kind, r, part[blockIdx.x + q * stride]);
    const double b = block_reduce(kind, r);
    if (threadIdx.x == 0) out[blockIdx.x] = b;
}

// ------------------------------------------------------------------------------------
// Neumann! (src/Tools.jl:4-18) and single-rank update_borders! (src/Parallel.jl:63-143)
// ------------------------------------------------------------------------------------
struct Ext3 {
    int lo[3];
    int n[3];
    __device__ __forceinline__ int64_t at(int i, int j, int k) const {
        return (int64_t)(i - lo[0]) + (int64_t)n[0] * ((int64_t)(j - lo[1]) + (int64_t)n[1] * (k - lo[2]));
    }
};

// Neumann! on the three axes in one launch. Applied axis after axis over the full extents of
// the other axes, the one-layer copies leave in every first-layer ghost the value of the cell
// with each first-layer ghost coordinate moved to the adjacent interior index (0 -> 1,
// mx+1 -> mx) and deeper coordinates kept; those sources are never written, so one thread per
// ghost position of the six faces copies its value directly (edge and corner positions get the
// same value from each face that contains them).
template <class T>
__global__ void k_neumann3(T* A, Ext3 e, int mx0, int mx1, int mx2) {
    const int mx[3] = {mx0, mx1, mx2};
    const int64_t f0 = (int64_t)e.n[1] * e.n[2], f1 = (int64_t)e.n[2] * e.n[0], f2 = (int64_t)e.n[0] * e.n[1];
    const int64_t tot = f0 + f1 + f2;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int d = q < f0 ? 0 : (q < f0 + f1 ? 1 : 2);
        const int64_t r = q - (d == 0 ? 0 : (d == 1 ? f0 : f0 + f1));
        const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        int id[3];
        id[d1] = e.lo[d1] + (int)(r % e.n[d1]);
        id[d2] = e.lo[d2] + (int)(r / e.n[d1]);
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            id[d] = side ? mx[d] + 1 : 0;
            int s[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) s[a] = id[a] == 0 ? 1 : (id[a] == mx[a] + 1 ? mx[a] : id[a]);
            A[e.at(id[0], id[1], id[2])] = A[e.at(s[0], s[1], s[2])];
        }
    }
}

// Periodic self exchange along axis d: A[mx+1..mxo] = A[mn..], then A[mno..mn-1] =
// A[mx-ng+1..mx]; each column is copied in the order that gives the snapshot semantics of
// the reference's Sendrecv! buffers even when the extent is smaller than nghost.
template <class T>
__global__ void k_borders_axis(T* A, Ext3 e, int d, int mn, int mx, int mno, int mxo, int ng) {
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    const int64_t tot = (int64_t)e.n[d1] * e.n[d2];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int id[3], is[3];
        id[d1] = is[d1] = e.lo[d1] + (int)(q % e.n[d1]);
        id[d2] = is[d2] = e.lo[d2] + (int)(q / e.n[d1]);
        const int len1 = mxo - mx;
        for (int s = len1 - 1; s >= 0; --s) {
            id[d] = mx + 1 + s; is[d] = mn + s;
            A[e.at(id[0], id[1], id[2])] = A[e.at(is[0], is[1], is[2])];
        }
        const int len2 = mn - mno;
        for (int s = 0; s < len2; ++s) {
            id[d] = mno + s; is[d] = mx - ng + 1 + s;
            A[e.at(id[0], id[1], id[2])] = A[e.at(is[0], is[1], is[2])];
        }
    }
}

// ------------------------------------------------------------------------------------
// computeBand! (src/VF.jl:16-74)
// ------------------------------------------------------------------------------------
// One sweep of the band growth. In place: a sweep only lowers |band| from >n to n and only
// compares neighbours with n-1, so concurrent updates do not change any outcome. With `list`
// (the last sweep, n = nband) each cell's value is final once its own update is done, and the
// cells with |band| <= nband are appended to the band-cell list (what k_band_list gives).
__global__ void k_band_sweep(DevMesh m, int nstep, int16_t* band, int* list, int* count) {
    __shared__ int s_app[2];
    const int n = m.Nx * m.Ny * m.Nz;
    const int nb = (n + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole work-groups take part
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += gridDim.x * blockDim.x) {
        bool f = false;
        if (q < n) {
            int i, j, k;
            interior_of(m, q, i, j, k);
            int b = band[m.cell(i, j, k)];
            if (abs(b) > nstep) {
                bool hit = false;
                for (int kk = k - 1; kk <= k + 1; ++kk)
                    for (int jj = j - 1; jj <= j + 1; ++jj)
                        for (int ii = i - 1; ii <= i + 1; ++ii)
                            hit |= abs((int)band[m.cell(ii, jj, kk)]) == nstep - 1;
                if (hit) {
                    b = nstep * (b > 0 ? 1 : -1);
                    band[m.cell(i, j, k)] = (int16_t)b;
                }
            }
            f = abs(b) <= nstep;
        }
        if (list) append_group(f, q, list, count, s_app);
    }
}

// Cells with |band| <= nband (interior linear index) appended to `list`.
__global__ void k_band_list(DevMesh m, int nband, const int16_t* band, int* list, int* count) {
    __shared__ int s_app[2];
    const int n = m.Nx * m.Ny * m.Nz;
    const int nb = (n + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole work-groups take part
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += gridDim.x * blockDim.x) {
        bool f = false;
        if (q < n) {
            int i, j, k;
            interior_of(m, q, i, j, k);
            f = abs((int)band[m.cell(i, j, k)]) <= nband;
        }
        append_group(f, q, list, count, s_app);
    }
}

// Classification: a cell takes band 0 when it is an interface cell or one of its 27 neighbours
// lies in the other phase, +-(nband+1) otherwise. One work-group of 256 threads per tile of
// interior cells (8x8x4, 16x8x2 or 16x16x1 by the mesh's depth), the tile's VF and a one-cell
// halo staged in LDS, so a cell's 27 neighbours come from LDS. (The same staging of the int16
// band values made the sweep slower than its 27 coalesced loads per wave: 32 against 15 us.)
struct BandTile {
    int tx, ty, tz;  // tile extents
    int nbx, nby;    // tiles along x and y
};
constexpr int BAND_TILE_LDS = 1024;  // >= (tx + 2)(ty + 2)(tz + 2) of every tile shape

__device__ __forceinline__ void band_tile_origin(const BandTile& T, int& i0, int& j0, int& k0) {
    const int b = blockIdx.x;
    i0 = 1 + (b % T.nbx) * T.tx;
    j0 = 1 + ((b / T.nbx) % T.nby) * T.ty;
    k0 = 1 + (b / (T.nbx * T.nby)) * T.tz;
}

__global__ __launch_bounds__(256) void k_band_classify_t(DevMesh m, DevParams p, BandTile T, const double* VF,
                                                         int16_t* band, int* count) {
    __shared__ double s[BAND_TILE_LDS];
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;  // the band-cell list, built after the sweeps
    int i0, j0, k0;
    band_tile_origin(T, i0, j0, k0);
    const int hx = T.tx + 2, hy = T.ty + 2, nh = hx * hy * (T.tz + 2);
    for (int e = threadIdx.x; e < nh; e += blockDim.x) {
        const int i = i0 - 1 + e % hx, j = j0 - 1 + (e / hx) % hy, k = k0 - 1 + e / (hx * hy);
        if (i <= m.Nx + 1 && j <= m.Ny + 1 && k <= m.Nz + 1) s[e] = VF[m.vof(i, j, k)];
    }
    __syncthreads();
    const int a = threadIdx.x % T.tx, b = (threadIdx.x / T.tx) % T.ty, c = threadIdx.x / (T.tx * T.ty);
    const int i = i0 + a, j = j0 + b, k = k0 + c;
    if (c >= T.tz || i > m.Nx || j > m.Ny || k > m.Nz) return;
    const double* o = s + (a + 1) + hx * ((b + 1) + hy * (c + 1));
    const double v = o[0];
    int bb;
    if (v < p.VFlo) {
        bb = -p.nband - 1;
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di)
                    if (o[di + hx * (dj + hy * dk)] >= p.VFhi) bb = 0;
    } else if (v > p.VFhi) {
        bb = p.nband + 1;
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di)
                    if (o[di + hx * (dj + hy * dk)] <= p.VFlo) bb = 0;
    } else {
        bb = 0;
    }
    band[m.cell(i, j, k)] = (int16_t)bb;
}

__global__ void k_fill_int(int* a, int n, int v) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = v;
}

// ------------------------------------------------------------------------------------
// ELVIRA! (src/ELVIRA.jl:6-55): normals of non-interface cells, interface list
// ------------------------------------------------------------------------------------
__global__ void k_normals_prep(DevMesh m, DevParams p, const double* VF, double* nx, double* ny, double* nz,
                               int* list, int* count) {
    __shared__ int s_app[2];
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    const int64_t nb = (n + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole work-groups take part
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nb; q += (int64_t)gridDim.x * blockDim.x) {
        bool iface = false;
        int item = 0;
        if (q < n) {
            int i, j, k;
            cell_of(m, q, i, j, k);
            double a = 0.0, b = 0.0, c = 0.0;
            if (i >= 1 && i <= m.Nx && j >= 1 && j <= m.Ny && k >= 1 && k <= m.Nz) {
                const double v = VF[m.vof(i, j, k)];
                if (v >= p.VFlo && v <= p.VFhi) {
                    iface = true;
                    item = interior_id(m, i, j, k);
                } else if (v < p.VFlo) {
                    if (VF[m.vof(i - 1, j, k)] > p.VFhi) a = 1;
                    else if (VF[m.vof(i + 1, j, k)] > p.VFhi) a = -1;
                    else if (VF[m.vof(i, j - 1, k)] > p.VFhi) b = 1;
                    else if (VF[m.vof(i, j + 1, k)] > p.VFhi) b = -1;
                    else if (VF[m.vof(i, j, k - 1)] > p.VFhi) c = 1;
                    else if (VF[m.vof(i, j, k + 1)] > p.VFhi) c = -1;
                } else if (v > p.VFhi) {
                    if (VF[m.vof(i - 1, j, k)] < p.VFlo) a = -1;
                    else if (VF[m.vof(i + 1, j, k)] < p.VFlo) a = 1;
                    else if (VF[m.vof(i, j - 1, k)] < p.VFlo) b = -1;
                    else if (VF[m.vof(i, j + 1, k)] < p.VFlo) b = 1;
                    else if (VF[m.vof(i, j, k - 1)] < p.VFlo) c = -1;
                    else if (VF[m.vof(i, j, k + 1)] < p.VFlo) c = 1;
                }
            }
            nx[q] = a; ny[q] = b; nz[q] = c;
        }
        append_group(iface, item, list, count, s_app);
    }
}

// The normal of ELVIRA candidate c of cell (i, j, k) (ELVIRA_calc!, src/ELVIRA.jl:62-100:
// family c / 9 = X-Y, X-Z or Y-Z stencils, outer stencil pair (c % 9) / 3, inner c % 3) from the
// 3x3x3 liquid fractions at(di, dj, dk), normalised.
template <class At>
__device__ __forceinline__ void elvira_normal(const DevMesh& m, int i, int j, int k, int c, At&& at, double& cnx,
                                              double& cny, double& cnz) {
    const int fam = c / 9, a = (c % 9) / 3, b = c % 3;
    // stencil pairs [[-1,0],[-1,1],[0,1]] of src/ELVIRA.jl:72
    auto P0 = [](int q) { return q == 2 ? 0 : -1; };
    auto P1 = [](int q) { return q == 0 ? 0 : 1; };
    if (fam == 0) {  // X-Y stencils (src/ELVIRA.jl:71-78)
        const int i1 = P0(b), i2 = P1(b), j1 = P0(a), j2 = P1(a);
        const double dz = m.dz;
        const double sx2 = (at(i2, 0, -1) * dz + at(i2, 0, 0) * dz) + at(i2, 0, 1) * dz;
        const double sx1 = (at(i1, 0, -1) * dz + at(i1, 0, 0) * dz) + at(i1, 0, 1) * dz;
        cnx = -(sx2 - sx1) / (m.xm[i + i2] - m.xm[i + i1]);
        const double sy2 = (at(0, j2, -1) * dz + at(0, j2, 0) * dz) + at(0, j2, 1) * dz;
        const double sy1 = (at(0, j1, -1) * dz + at(0, j1, 0) * dz) + at(0, j1, 1) * dz;
        cny = -(sy2 - sy1) / (m.ym[j + j2] - m.ym[j + j1]);
        cnz = at(0, 0, 1) > at(0, 0, -1) ? -1.0 : 1.0;
    } else if (fam == 1) {  // X-Z stencils (src/ELVIRA.jl:81-88)
        const int i1 = P0(b), i2 = P1(b), k1 = P0(a), k2 = P1(a);
        const double dy = m.dy;
        const double sx2 = (at(i2, -1, 0) * dy + at(i2, 0, 0) * dy) + at(i2, 1, 0) * dy;
        const double sx1 = (at(i1, -1, 0) * dy + at(i1, 0, 0) * dy) + at(i1, 1, 0) * dy;
        cnx = -(sx2 - sx1) / (m.xm[i + i2] - m.xm[i + i1]);
        cny = at(0, 1, 0) > at(0, -1, 0) ? -1.0 : 1.0;
        const double sz2 = (at(0, -1, k2) * dy + at(0, 0, k2) * dy) + at(0, 1, k2) * dy;
        const double sz1 = (at(0, -1, k1) * dy + at(0, 0, k1) * dy) + at(0, 1, k1) * dy;
        cnz = -(sz2 - sz1) / (m.zm[k + k2] - m.zm[k + k1]);
    } else {  // Y-Z stencils (src/ELVIRA.jl:91-98)
        const int j1 = P0(b), j2 = P1(b), k1 = P0(a), k2 = P1(a);
        const double dx = m.dx;
        cnx = at(1, 0, 0) > at(-1, 0, 0) ? -1.0 : 1.0;
        const double sy2 = (at(-1, j2, 0) * dx + at(0, j2, 0) * dx) + at(1, j2, 0) * dx;
        const double sy1 = (at(-1, j1, 0) * dx + at(0, j1, 0) * dx) + at(1, j1, 0) * dx;
        cny = -(sy2 - sy1) / (m.ym[j + j2] - m.ym[j + j1]);
        const double sz2 = (at(-1, 0, k2) * dx + at(0, 0, k2) * dx) + at(1, 0, k2) * dx;
        const double sz1 = (at(-1, 0, k1) * dx + at(0, 0, k1) * dx) + at(1, 0, k1) * dx;
        cnz = -(sz2 - sz1) / (m.zm[k + k2] - m.zm[k + k1]);
    }
    const double norm = sqrt((cnx * cnx + cny * cny) + cnz * cnz);
    cnx = cnx / norm; cny = cny / norm; cnz = cnz / norm;
}

// computeDist (src/VF.jl:109-156) of every ELVIRA candidate of every interface cell, into
// dist[27 cidx + c]. Its own kernel: the closed-form root's registers would otherwise spill
// in the scoring kernel's loops.
__global__ __launch_bounds__(256) void k_elvira_dist(DevMesh m, DevParams p, const double* VF, const int* list,
                                                     const int* count, int fam_mask, double* dist, int* ntie,
                                                     int* queue) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *ntie = 0;   // k_elvira, which follows, appends the near ties
        *queue = 0;  // and takes its cells from this counter
    }
    const int64_t n = (int64_t)(*count) * 27;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int cidx = (int)(q / 27), c = (int)(q - 27 * (int64_t)cidx);
        if (!((fam_mask >> (c / 9)) & 1)) continue;
        int i, j, k;
        interior_of(m, list[cidx], i, j, k);
        auto at = [&](int di, int dj, int dk) { return VF[m.vof(i + di, j + dj, k + dk)]; };
        double cnx, cny, cnz;
        elvira_normal(m, i, j, k, c, at, cnx, cny, cnz);
        dist[q] = compute_dist(m, p, i, j, k, cnx, cny, cnz, at(0, 0, 0));
    }
}

// ELVIRA_calc! + check_normal! (src/ELVIRA.jl:62-137) with one interface cell at a time per
// work-group of 256 threads, the cells taken from a queue. Threads 0..26 load the 3x3x3 liquid
// fractions and build the 27 candidate normals once; their PLIC constants (computeDist) come
// from k_elvira_dist, which evaluates the same normals bit for bit. Each candidate's error is
// the sum over the 27 neighbours, in the reference's order, of (VF - computePLIC2VF)^2, where
// computePLIC2VF adds, tet by tet, the liquid sub-tet volumes of the cut (src/VF.jl:219-278).
//
// The threads score the 729 candidate-neighbour pairs in turn. A plane that misses the
// neighbour by a wide margin gives the exact all-gas / all-liquid value at once. Otherwise the
// plane's distances at the neighbour's 8 corners (the products and sums the reference forms at
// each tet vertex, so each tet's four distances are its values bit for bit) give each of the
// five tets its d2case case: an all-liquid tet adds its whole term (ts vol / cellvol, computed
// once per neighbour) and a cut tet that term times its liquid share in closed form
// (tet_liquid_fraction). That equals the reference's sub-tet sum in exact arithmetic; in
// floating point the two differ by the rounding of the reference's sub-tet determinants, about
// 1e-16 |x| / h per neighbour. One thread per candidate adds the squared errors in neighbour
// order. The best candidate is the lowest-numbered one with the smallest finite error (the
// reference's strict `<` scan); a candidate with another normal whose error lies within what
// that deviation can change of the best error is a near tie, and the reference's own
// arithmetic decides between them (k_elvira_ties).
constexpr int ELV_THREADS = 256;
constexpr int ELV_PAIRS = 27 * 27;  // candidate-neighbour pairs
// Six work-groups per CU (80 VGPRs): the kernel arguments spill to scratch in the prologue,
// once per resident work-group; four (123 VGPRs, no spills) were 2% slower.

template <bool POW2>
__global__ __launch_bounds__(ELV_THREADS, 6) void k_elvira(DevMesh m, DevParams p, const double* VF, const int* list,
                                                        const int* count, int fam_mask, const double* dist,
                                                        double* nxo, double* nyo, double* nzo, int* best, int* ties,
                                                        int* ntie, int* queue) {
    __shared__ double sVF[27];
    __shared__ double sC[27][4];
    __shared__ int sOk[27];
    __shared__ double sT[27][5];      // all-liquid term (ts * vol(tet)) / cellvol per neighbour and tet
    __shared__ double sFull[27];      // computePLIC2VF of an all-liquid neighbour
    __shared__ double sE[ELV_PAIRS];  // squared error of each candidate-neighbour pair
    __shared__ double sErr[27];
    // faces x[i-1..i+2], y[j-1..j+2], z[k-1..k+2] and centres xm[i-1..i+1], ym, zm of the cell's
    // neighbourhood: the pair step reads them from LDS
    __shared__ double sF[3][4], sM[3][3];
    __shared__ int sCell;
    const int tid = threadIdx.x;
    const int ncell = *count;
    const VolScale<POW2> vscale((m.dx * m.dy) * m.dz);
    const double hmin = jmin(jmin(m.dx, m.dy), m.dz);
    for (;;) {
        // the next cell from the queue: a work-group per resident slot, each taking cells until
        // the list is done (the barrier at the loop's end orders this write after every read)
        if (tid == 0) sCell = atomicAdd(queue, 1);
        __syncthreads();
        const int cidx = sCell;
        if (cidx >= ncell) break;
        int i, j, k;
        interior_of(m, list[cidx], i, j, k);
        if (tid < 27) {
            const int ii = i - 1 + tid % 3, jj = j - 1 + (tid / 3) % 3, kk = k - 1 + tid / 9;
            sVF[tid] = VF[m.vof(ii, jj, kk)];
        }
        if (tid >= 160 && tid < 181) {
            const int e = tid - 160;
            if (e < 12) {
                const int a = e / 4, o = e % 4;
                sF[a][o] = (a == 0 ? m.x[i - 1 + o] : (a == 1 ? m.y[j - 1 + o] : m.z[k - 1 + o]));
            } else {
                const int a = (e - 12) / 3, o = (e - 12) % 3;
                sM[a][o] = (a == 0 ? m.xm[i - 1 + o] : (a == 1 ? m.ym[j - 1 + o] : m.zm[k - 1 + o]));
            }
        }
        if (tid < 27 * 5) {
            const int nb = tid / 5, t = tid % 5;
            const int ii = i - 1 + nb % 3, jj = j - 1 + (nb / 3) % 3, kk = k - 1 + nb / 9;
            const bool even = ((ii + jj + kk) % 2) == 0;
            V3 P[4];
            tet5_vertices(t, even, m.x[ii], m.x[ii + 1], m.y[jj], m.y[jj + 1], m.z[kk], m.z[kk + 1], P);
            sT[nb][t] = vscale(even ? 1.0 : -1.0, tet_vol3(P[0], P[1], P[2], P[3]));
        }
        __syncthreads();
        if (tid < 27) {
            double f = 0.0;
            for (int t = 0; t < 5; ++t) f += sT[tid][t];
            sFull[tid] = f;
        }
        if (tid < 27) {
            const int ok = (fam_mask >> (tid / 9)) & 1;
            if (ok) {
                auto at = [&](int di, int dj, int dk) { return sVF[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)]; };
                double cnx, cny, cnz;
                elvira_normal(m, i, j, k, tid, at, cnx, cny, cnz);
                sC[tid][0] = cnx; sC[tid][1] = cny; sC[tid][2] = cnz;
                sC[tid][3] = dist[27 * (int64_t)cidx + tid];
            }
            sOk[tid] = ok;
        }
        __syncthreads();
        for (int pr = tid; pr < ELV_PAIRS; pr += ELV_THREADS) {
            const int c = pr % 27, nb = pr / 27;
            double vf_ = 0.0;
            if (sOk[c]) {
                const double cnx = sC[c][0], cny = sC[c][1], cnz = sC[c][2], dist = sC[c][3];
                const int oi = nb % 3, oj = (nb / 3) % 3, ol = nb / 9;
                const double hx = fabs(cnx) * m.dx, hy = fabs(cny) * m.dy, hz = fabs(cnz) * m.dz;
                const double xc = sM[0][oi], yc = sM[1][oj], zc = sM[2][ol];
                const double dcen = ((cnx * xc + cny * yc) + cnz * zc) - dist;
                const double hext = 0.5 * ((hx + hy) + hz);
                const double margin = 1e-9 * (((1.0 + fabs(dist)) + fabs(xc)) + (fabs(yc) + fabs(zc)));
                if (dcen + hext < -margin) {
                    vf_ = sFull[nb];
                } else if (!(dcen - hext > margin)) {
                    const bool even = (((i - 1 + oi) + (j - 1 + oj) + (k - 1 + ol)) % 2) == 0;
                    // d of the cell's 8 reference vertices, each the expression the tets'
                    // vertices evaluate ((nx x + ny y) + nz z) - dist with the same products
                    // (odd cells mirror x), then each tet's four taken from them
                    const double px0 = cnx * sF[0][oi], px1 = cnx * sF[0][oi + 1];
                    const double py0 = cny * sF[1][oj], py1 = cny * sF[1][oj + 1];
                    const double pz0 = cnz * sF[2][ol], pz1 = cnz * sF[2][ol + 1];
                    const double qx0 = even ? px0 : px1, qx1 = even ? px1 : px0;
                    double e8[8];
                    unsigned pos8 = 0;  // corners on the gas side
                    bool tiny = false;  // a corner d2case would snap
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        e8[v] = (((v & 1) ? qx1 : qx0) + ((v & 2) ? py1 : py0) + ((v & 4) ? pz1 : pz0)) - dist;
                        pos8 |= (e8[v] > 0.0 ? 1u : 0u) << v;
                        tiny = tiny || fabs(e8[v]) < EPS;
                    }
                    sfor<0, 5>([&](auto T) {
                        constexpr int t = decltype(T)::value;
                        constexpr unsigned tv = CutTab::tet5[t];
                        double d[4] = {e8[tv & 15], e8[(tv >> 4) & 15], e8[(tv >> 8) & 15], e8[(tv >> 12) & 15]};
                        // with no corner below d2case's snapping threshold its case is the tet's
                        // corners' sides and the distances stay as they are
                        int cs;
                        if (!tiny)
                            cs = (int)(((pos8 >> (tv & 15)) & 1u) | (((pos8 >> ((tv >> 4) & 15)) & 1u) << 1) |
                                       (((pos8 >> ((tv >> 8) & 15)) & 1u) << 2) |
                                       (((pos8 >> ((tv >> 12) & 15)) & 1u) << 3));
                        else
                            cs = d2case_snap(d);
                        if (cs == 0) vf_ += sT[nb][t];
                        else if (cs != 15) vf_ += sT[nb][t] * tet_liquid_fraction(cs, d);
                    });
                }
            }
            const double df = sVF[nb] - vf_;
            sE[pr] = df * df;
        }
        __syncthreads();
        if (tid < 27) {
            double e = 0.0;
            for (int nb = 0; nb < 27; ++nb) e += sE[27 * nb + tid];
            sErr[tid] = e;
        }
        __syncthreads();
        if (tid < 32) {
            const int lane = tid;
            double err = __builtin_huge_val();
            if (lane < 27 && sOk[lane]) err = sErr[lane];
            const bool ok = err < __builtin_huge_val();
            double be = ok ? err : __builtin_huge_val();
            int bl = ok ? lane : 64;
            for (int off = 16; off > 0; off >>= 1) {
                const double oe = __shfl_xor(be, off, 32);
                const int ol = __shfl_xor(bl, off, 32);
                if (ol < 64 && (bl == 64 || oe < be || (oe == be && ol < bl))) { be = oe; bl = ol; }
            }
            // another candidate with the very same normal: the choice between the two is a kink
            // of the map (a perturbation's sign decides it), recorded for the adjoint
            const bool same = bl < 64 && lane < 27 && lane != bl && sOk[lane] && sC[lane][0] == sC[bl][0] &&
                              sC[lane][1] == sC[bl][1] && sC[lane][2] == sC[bl][2];
            const unsigned long long tmask = __ballot(same);
            const int tie = tmask ? __ffsll((long long)tmask) - 1 : -1;
            // candidates with another normal whose errors lie within what the volume fractions'
            // deviation from the reference's arithmetic can change of the minimum: the reference's
            // own arithmetic decides between them and the best (k_elvira_ties). A deviation dv
            // per neighbour moves an error E = sum of 27 squares by at most 2 dv sqrt(27 E) +
            // 27 dv^2, and two candidates' errors twice that; dv is ten times the measured
            // 1e-16 |x| / h. A candidate with the very same normal has the very same error in any
            // arithmetic, and the lowest-numbered one is kept either way.
            double R = 0.0;
            for (int a = 0; a < 3; ++a) R = jmax(R, jmax(fabs(sF[a][0]), fabs(sF[a][3])));
            const double dv = 1e-15 * (R / hmin + 1.0);
            const double slack = (1e-11 * be + (21.0 * dv) * sqrt(be)) + (54.0 * dv) * dv;
            const bool near = bl < 64 && lane < 27 && lane != bl && sOk[lane] && err <= be + slack &&
                              !(sC[lane][0] == sC[bl][0] && sC[lane][1] == sC[bl][1] && sC[lane][2] == sC[bl][2]);
            const unsigned long long fmask = __ballot(near) | (bl < 64 ? (1ull << bl) : 0ull);
            if (lane == 0) {
                if (bl < 64) {
                    const int64_t c = m.cell(i, j, k);
                    nxo[c] = sC[bl][0]; nyo[c] = sC[bl][1]; nzo[c] = sC[bl][2];
                }
                best[list[cidx]] = bl < 64 ? (bl | ((tie + 1) << 8)) : -1;
                if (__popcll(fmask) > 1) {
                    const int slot = atomicAdd(ntie, 1);
                    ties[2 * slot] = cidx;
                    ties[2 * slot + 1] = (int)fmask;
                }
            }
        }
        __syncthreads();
    }
}

// The cells k_elvira flagged, their near-tied candidates scored again in the reference's own
// arithmetic: computePLIC2VF of each neighbour with the fourteen case bodies and the sub-tets
// in the reference's order (plic2vf_reg), the squared errors added in neighbour order. Between
// candidates whose errors are equal in exact arithmetic (mirrored candidates of a symmetric
// interface) this keeps the one the reference keeps; the rest of the cell's result is
// k_elvira's. One work-group of 64 threads per flagged cell, a thread per candidate.
__global__ __launch_bounds__(64) void k_elvira_ties(DevMesh m, const double* VF, const int* list, int fam_mask,
                                                    const double* dist, const int* ties, const int* ntie,
                                                    double* nxo, double* nyo, double* nzo, int* best) {
    __shared__ double sErr[27];
    __shared__ double sN[27][3];
    const double cellvol = (m.dx * m.dy) * m.dz;
    const int n = *ntie;
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const int cidx = ties[2 * e];
        const unsigned fmask = (unsigned)ties[2 * e + 1];
        int i, j, k;
        interior_of(m, list[cidx], i, j, k);
        const int c = threadIdx.x;
        auto at = [&](int di, int dj, int dk) { return VF[m.vof(i + di, j + dj, k + dk)]; };
        if (c < 27) {
            double cnx = 0.0, cny = 0.0, cnz = 0.0;
            if ((fam_mask >> (c / 9)) & 1) elvira_normal(m, i, j, k, c, at, cnx, cny, cnz);
            sN[c][0] = cnx; sN[c][1] = cny; sN[c][2] = cnz;
            double err = __builtin_huge_val();
            if ((fmask >> c) & 1) {
                const double dc = dist[27 * (int64_t)cidx + c];
                double acc = 0.0;
                for (int nb = 0; nb < 27; ++nb) {
                    const int oi = nb % 3, oj = (nb / 3) % 3, ol = nb / 9;
                    const int ii = i - 1 + oi, jj = j - 1 + oj, kk = k - 1 + ol;
                    const bool even = ((ii + jj + kk) % 2) == 0;
                    const double vf_ = plic2vf_reg(m.x[ii], m.x[ii + 1], m.y[jj], m.y[jj + 1], m.z[kk], m.z[kk + 1],
                                                   cnx, cny, cnz, dc, even, cellvol);
                    const double df = at(oi - 1, oj - 1, ol - 1) - vf_;
                    acc += df * df;
                }
                err = acc;
            }
            sErr[c] = err;
        }
        __syncthreads();
        if (c == 0) {
            int bl = -1;
            double be = __builtin_huge_val();
            for (int q = 0; q < 27; ++q)  // the reference's strict < scan over the flagged candidates
                if (((fmask >> q) & 1) && sErr[q] < be) { be = sErr[q]; bl = q; }
            if (bl >= 0) {
                int tie = -1;  // another candidate with the very same normal (k_elvira's rule)
                for (int q = 0; q < 27 && tie < 0; ++q)
                    if (q != bl && ((fam_mask >> (q / 9)) & 1) && sN[q][0] == sN[bl][0] && sN[q][1] == sN[bl][1] &&
                        sN[q][2] == sN[bl][2])
                        tie = q;
                const int64_t cc = m.cell(i, j, k);
                nxo[cc] = sN[bl][0]; nyo[cc] = sN[bl][1]; nzo[cc] = sN[bl][2];
                best[list[cidx]] = bl | ((tie + 1) << 8);
            }
        }
        __syncthreads();
    }
}

// computePLIC! (src/VF.jl:94-103) over every cell including ghosts
__global__ void k_plic(DevMesh m, DevParams p, const double* VF, const double* nx, const double* ny,
                       const double* nz, double* D) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        D[q] = compute_dist(m, p, i, j, k, nx[q], ny[q], nz[q], VF[m.vof(i, j, k)]);
    }
}

// Every vertex of a band cell, ghost cells included (a periodic image of a band cell puts an
// SL flux on the boundary face of its interior neighbour), appended to `list` (linear vertex
// index over (Nx+1)(Ny+1)(Nz+1)), so that the back-trace runs on full waves.
__global__ void k_vertex_list(DevMesh m, DevParams p, const int16_t* band, int* list, int* count) {
    __shared__ int s_app[2];
    const int64_t n = (int64_t)(m.Nx + 1) * (m.Ny + 1) * (m.Nz + 1);
    const int64_t nb = (n + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole work-groups take part
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nb; q += (int64_t)gridDim.x * blockDim.x) {
        bool need = false;
        if (q < n) {
            const int I = (int)(q % (m.Nx + 1)) + 1;
            const int64_t r = q / (m.Nx + 1);
            const int J = (int)(r % (m.Ny + 1)) + 1;
            const int K = (int)(r / (m.Ny + 1)) + 1;
            for (int c = 0; c < 8; ++c) {
                const int i = I - 1 + (c & 1), j = J - 1 + ((c >> 1) & 1), k = K - 1 + (c >> 2);
                need |= abs((int)band[m.cell(i, j, k)]) <= p.nband;
            }
        }
        append_group(need, (int)q, list, count, s_app);
    }
}

// project! (src/Tools.jl:392-418) of the listed vertices into pv[3 q ..].
__global__ void k_project_list(DevMesh m, DevParams p, const int* list, const int* count, const double* uf,
                               const double* vf, const double* wf, double dt, double* pv) {
    const int n = *count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int64_t q = list[e];
        const int I = (int)(q % (m.Nx + 1)) + 1;
        const int64_t r = q / (m.Nx + 1);
        const int J = (int)(r % (m.Ny + 1)) + 1;
        const int K = (int)(r / (m.Ny + 1)) + 1;
        double pt[3] = {m.x[I], m.y[J], m.z[K]};
        project_pt(m, p, uf, vf, wf, pt, min(I, m.Nx), min(J, m.Ny), min(K, m.Nz), dt);
        pv[3 * q + 0] = pt[0];
        pv[3 * q + 1] = pt[1];
        pv[3 * q + 2] = pt[2];
    }
}

// Semi-Lagrangian VF of band cells (src/Transport.jl:111-145). 12 cells per wave, one lane
// per (cell, tet); the five tet results of a cell are summed in tet order by its first lane.
// VF is written in place: the cuts read only normals and D, never VF.
__global__ __launch_bounds__(64) void k_sl_cut(DevMesh m, const double* pv, const int* list, const int* count,
                                               const double* nx, const double* ny, const double* nz,
                                               const double* D, double* VF, int* ovf) {
    __shared__ double s_vol[64], s_liq[64];
    const int lane = threadIdx.x;
    const int cl = lane / 5, t = lane - 5 * cl;
    const int nb = *count;
    for (int base = blockIdx.x * 12; base < nb; base += gridDim.x * 12) {
        const int cidx = base + cl;
        const bool active = lane < 60 && cidx < nb;
        double vol = 0.0, liq = 0.0;
        int i = 1, j = 1, k = 1;
        if (active) {
            interior_of(m, list[cidx], i, j, k);
            const bool even = ((i + j + k) % 2) == 0;
            const unsigned tv = TET5[t];
            double tet[4][3];
            int ind[4][3];
            int o = 0;
            for (int n = 0; n < 4; ++n) {
                const int c = vert_corner(nib(tv, n), even);
                const int64_t vq = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
                tet[n][0] = pv[3 * vq + 0]; tet[n][1] = pv[3 * vq + 1]; tet[n][2] = pv[3 * vq + 2];
                pt2index(m, ind[n], tet[n], i, j, k, &o);
            }
            CutFrame<2> st[CUT_MAXLVL];
            int depth = 0;
            double r[2];
            const LeafMomentum none{nullptr, nullptr, nullptr, 0.0, 0.0};
            if (!cut_tet<2>(m, nx, ny, nz, D, none, tet, ind, st, r, depth)) o = 1;
            vol = r[0];
            liq = r[1];
            if (o) atomicOr(ovf, 1);
        }
        s_vol[lane] = vol;
        s_liq[lane] = liq;
        __syncthreads();
        if (active && t == 0) {
            const double ts = ((i + j + k) % 2) == 0 ? 1.0 : -1.0;
            double V = 0.0, L = 0.0;
            for (int q = 0; q < 5; ++q) { V += ts * s_vol[lane + q]; L += ts * s_liq[lane + q]; }
            VF[m.vof(i, j, k)] = L / V;
        }
        __syncthreads();
    }
}

// Corners, index box and bulk test of every band cell, one thread per cell (the set-up of the
// cut walk, kept out of the work-group kernel where its chain of dependent loads would stall a
// whole block per cell). Corner cells come from pt2index! started at the cell itself
// (src/Transport.jl:128-131). When every cell of the box lies on one side of its PLIC plane,
// each leaf of the reference's cutTet is all-liquid (or all-gas), its vol and vLiq sums receive
// the same additions (or vLiq none) and the new VF is exactly 1 (or 0): it is written here. The
// other cells go to the cut list; those whose corners span more than 4 cells along an axis to
// the general walk.
// With the momentum sums (`mom`) every cell is cut: the leaves' cell velocities weight the
// momentum even where every leaf is on one side of its plane.
// Without `VF` (the reverse pass) the bulk cells are only left out of the lists.
__global__ __launch_bounds__(256) void k_sl_classify(DevMesh m, const double* pv, const int* list, const int* count,
                                                     const double* nx, const double* ny, const double* nz,
                                                     const double* D, double* VF, int* ovf, CutCell* cut,
                                                     int* ncut, int* general, int* ngeneral, int mom) {
    const int nb = *count;
    const int nloop = (nb + 63) & ~63;  // whole waves in the loop (the appends use ballots)
    for (int cidx = blockIdx.x * blockDim.x + threadIdx.x; cidx < nloop; cidx += gridDim.x * blockDim.x) {
        const bool valid = cidx < nb;
        int q = 0, i = 1, j = 1, k = 1;
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        unsigned long long offs = 0;
        int kind = 0;  // 0: none, 1: cut list, 2: general walk
        if (valid) {
            q = list[cidx];
            interior_of(m, q, i, j, k);
            int I[8][3];
            int o = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int64_t vq = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
                const double pt[3] = {pv[3 * vq + 0], pv[3 * vq + 1], pv[3 * vq + 2]};
                pt2index(m, I[c], pt, i, j, k, &o);
            }
            if (o) atomicOr(ovf, 1);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = I[0][a];
                hi[a] = I[0][a];
#pragma unroll
                for (int c = 1; c < 8; ++c) {
                    lo[a] = min(lo[a], I[c][a]);
                    hi[a] = max(hi[a], I[c][a]);
                }
            }
            if (hi[0] - lo[0] > 3 || hi[1] - lo[1] > 3 || hi[2] - lo[2] > 3) {
                kind = 2;
            } else {
#pragma unroll
                for (int c = 0; c < 8; ++c)
#pragma unroll
                    for (int a = 0; a < 3; ++a) offs |= (unsigned long long)(I[c][a] - lo[a]) << (6 * c + 2 * a);
                bool liq = !mom, gas = !mom;
                for (int cz = lo[2]; cz <= hi[2] && (liq || gas); ++cz)
                    for (int cy = lo[1]; cy <= hi[1] && (liq || gas); ++cy)
                        for (int cx = lo[0]; cx <= hi[0] && (liq || gas); ++cx) {
                            const int64_t c = m.cell(cx, cy, cz);
                            const int cl = box_class(nx[c], ny[c], nz[c], D[c], m.x[cx], m.x[cx + 1], m.y[cy],
                                                     m.y[cy + 1], m.z[cz], m.z[cz + 1]);
                            liq = liq && cl == 1;
                            gas = gas && cl == -1;
                        }
                if (liq || gas) {
                    if (VF) VF[m.vof(i, j, k)] = liq ? 1.0 : 0.0;
                } else {
                    kind = 1;
                }
            }
        }
        const int sc = wave_slot(kind == 1, ncut);
        if (sc >= 0) cut[sc] = CutCell{q, lo[0], lo[1], lo[2], offs};
        const int sg = wave_slot(kind == 2, ngeneral);
        if (sg >= 0) general[sg] = q;
    }
}

// MOM: the six sums of the momentum predictor (vol, vLiq, rho u vol, rho v vol, rho w vol,
// rho vol with the leaf cell's centre velocity, src/VFgeom.jl:791-827) give VF and the predicted
// velocities us, vs, ws of the cell. CORR: the twelve flux-correction tets of the
// finite-difference scheme are cut in the same tree as the five back-traced ones; a cell whose
// tets do not fit one index box of 4 cells per axis goes to `defer`.
template <class C, bool MOM, bool CORR = false>
__global__ __launch_bounds__(C::NT) void k_sl_tree(DevMesh m, const double* pv, const CutCell* cut, const int* idx,
                                                const int* count, const double* nx, const double* ny,
                                                const double* nz, const double* D, double* VF, int* defer,
                                                int* ndefer, int defer_cell, int* queue, TreeNS ns) {
    static_assert(MOM || !CORR, "the correction tets belong to the momentum predictor");
    constexpr int NT = C::NT;
    const LeafMomentum& lm = ns.lm;
    __shared__ TreeShared<C> S;
    __shared__ double s_vel[MOM ? 64 : 1][3];
    __shared__ double s_red[MOM ? 4 : 1][NT / 64];
    __shared__ int s_cind[CORR ? 48 : 1][3];  // cells of the correction tets' vertices
    __shared__ int s_box[2][3];               // index box over corners and correction vertices
    const int tid = threadIdx.x;
    const int nb = *count;
    tab_load<C>(S);
    // the next cell is claimed while the current one is cut (the atomic's latency is hidden)
    if (tid == 0) S.next = atomicAdd(queue, 1);
    __syncthreads();
    int cidx = S.next;
    while (cidx < nb) {
        __syncthreads();  // every thread has read S.next
        int claim = 0;
        if (tid == 0) claim = atomicAdd(queue, 1);
        const int entry = idx ? idx[cidx] : cidx;
        const CutCell cc = cut[entry];
        const int q = cc.q;
        int i, j, k;
        interior_of(m, q, i, j, k);
        const bool even = ((i + j + k) % 2) == 0;
        int lo[3] = {cc.lo0, cc.lo1, cc.lo2};
        int span[3] = {0, 0, 0};
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) span[a] = max(span[a], (int)((cc.offs >> (6 * c + 2 * a)) & 3ull));
        int shift[3] = {0, 0, 0};  // of the corners' offsets when the box widens
        bool box_ok = true;
        if constexpr (CORR) {
            // the twelve correction pyramids (add_correction_tets, src/VFgeom.jl:513-664): their
            // vertices go to pool ids 8 + 4t + n, their cells (pt2index! from the cell, as
            // transport! walks them) widen the index box
            if (tid == 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    s_box[0][a] = lo[a];
                    s_box[1][a] = lo[a] + span[a];
                }
            }
            __syncthreads();
            if (tid < 12) {
                double tet[4][3];
                ns_tet(m, pv, ns.uf, ns.vf, ns.wf, ns.dt, i, j, k, 5 + tid, tet);
                int o = 0;
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    S.px[8 + 4 * tid + n] = tet[n][0];
                    S.py[8 + 4 * tid + n] = tet[n][1];
                    S.pz[8 + 4 * tid + n] = tet[n][2];
                    int I[3];
                    pt2index(m, I, tet[n], i, j, k, &o);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        s_cind[4 * tid + n][a] = I[a];
                        atomicMin(&s_box[0][a], I[a]);
                        atomicMax(&s_box[1][a], I[a]);
                    }
                }
                if (o) atomicOr(ns.ovf, 1);
            }
            __syncthreads();
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                shift[a] = lo[a] - s_box[0][a];
                lo[a] = s_box[0][a];
                span[a] = s_box[1][a] - s_box[0][a];
                box_ok = box_ok && span[a] <= 3;
            }
        }
        const int bx = span[0] + 1, by = span[1] + 1, bz = span[2] + 1;
        const int ncell = box_ok ? bx * by * bz : 0;
        // ---- projected corners, mesh planes and the box cells' PLIC planes, all at once ------
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
            if (box_ok && off <= sa + 1) S.plane[a][off] = (a == 0 ? m.x : (a == 1 ? m.y : m.z))[la + off];
        }
        for (int l = tid; l < ncell; l += NT) {
            const int ox = l % bx, oy = (l / bx) % by, oz = l / (bx * by);
            const int64_t c = m.cell(lo[0] + ox, lo[1] + oy, lo[2] + oz);
            const double n0 = nx[c], n1 = ny[c], n2 = nz[c], Dc = D[c];
            S.nb[l][0] = n0;
            S.nb[l][1] = n1;
            S.nb[l][2] = n2;
            S.nb[l][3] = Dc;
            S.cls[l] = (signed char)box_class(n0, n1, n2, Dc, m.x[lo[0] + ox], m.x[lo[0] + ox + 1], m.y[lo[1] + oy],
                                              m.y[lo[1] + oy + 1], m.z[lo[2] + oz], m.z[lo[2] + oz + 1]);
            if constexpr (MOM) {
                s_vel[l][0] = lm.u[c];
                s_vel[l][1] = lm.v[c];
                s_vel[l][2] = lm.w[c];
            }
        }
        // ---- the five tets (src/VFgeom.jl:235-256), then the correction tets -------------------
        if (tid < 5) {
            const unsigned tv = TET5[tid];
            unsigned ids[4], bits = 0;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const int c = vert_corner(nib(tv, n), even);
                ids[n] = (unsigned)c;
                unsigned ob = 0;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    ob |= (unsigned)(((cc.offs >> (6 * c + 2 * a)) & 3ull) + shift[a]) << (2 * a);
                bits |= ob << (6 * n);
            }
            node_set(S, 0, tid, TreeNode{ids[0] | (ids[1] << 16), ids[2] | (ids[3] << 16), bits});
        }
        if constexpr (CORR) {
            if (tid >= 5 && tid < 17 && box_ok) {
                const int t = tid - 5;
                unsigned bits = 0;
#pragma unroll
                for (int n = 0; n < 4; ++n)
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        bits |= (unsigned)(s_cind[4 * t + n][a] - lo[a]) << (6 * n + 2 * a);
                const unsigned id0 = 8 + 4 * t;
                node_set(S, 0, tid, TreeNode{id0 | ((id0 + 1) << 16), (id0 + 2) | ((id0 + 3) << 16), bits});
            }
        }
        __syncthreads();
        // ---- mesh-plane levels -------------------------------------------------------------
        int n = CORR ? 17 : 5, npool = CORR ? 56 : 8, cur = 0;
        bool ok = box_ok;
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
                ok = tree_level<C>(S, axis, minspan, cur, n, npool, Span3{{span[0], span[1], span[2]}});
                cur ^= 1;
            }
        }
        // ---- leaves per node (the last z-plane is cut again by each leaf) ------------------
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
        if (!ok) {  // pool or lists too small for this cell: the next pass
            if (tid == 0) defer[atomicAdd(ndefer, 1)] = defer_cell ? q : entry;
        } else {
            __syncthreads();
            // ---- leaves: last z cut and the PLIC cut (src/VFgeom.jl:757-830) ------------------
            double Vs = 0.0, Ls = 0.0;
            double Ms[4] = {0.0, 0.0, 0.0, 0.0};  // rho u vol, rho v vol, rho w vol, rho vol
            for (int L = tid; L < nleaf; L += NT) {
                const unsigned le = S.lnode[L];
                const int e = (int)(le & 0x7FFu), c = (int)(le >> 11);
                const TreeNode N = node_get(S, cur, e);
                const int o0 = boff(N.bits, 0, 2), o1 = boff(N.bits, 1, 2), o2 = boff(N.bits, 2, 2),
                          o3 = boff(N.bits, 3, 2);
                const int mx = max(max(o0, o1), max(o2, o3));
                V3 Q[4];
                unsigned ob;  // cell offsets (x, y, z: 2 bits each) of the leaf's first vertex
                if (mx > min(min(o0, o1), min(o2, o3))) {
                    // child c of the node's last z cut, its vertices straight from the pool: an
                    // original vertex is the edge (v, v) at mu = 0, which lerp3 returns exactly
                    const double P = S.plane[2][mx];
                    double dd[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) dd[v] = S.pz[tn_id(N, v)] - P;
                    const int cs = d2case_snap(dd);
                    const unsigned qv = S.tab.vtet[cs][c];
                    const bool s2 = (S.tab.side2[cs] >> c) & 1;
                    int e1v = 0, e2v = 0;  // edge of the first vertex
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        const int lv = nib(qv, kk);
                        const bool cutv = lv >= 4;
                        const unsigned eg = S.tab.edge[cs][cutv ? lv - 4 : 0];
                        const int v1 = cutv ? (int)(eg & 15) : lv, v2 = cutv ? (int)(eg >> 4) : lv;
                        const V3 A = pool_v(S, tn_id(N, v1)), B = pool_v(S, tn_id(N, v2));
                        const double d1 = A.z - P, d2 = B.z - P;
                        const double mu = cutv ? jmin(1.0, jmax(0.0, sdiv(-d1, d2 - d1))) : 0.0;
                        Q[kk] = lerp3(A, B, mu);
                        if (kk == 0) { e1v = v1; e2v = v2; }
                    }
                    // cell of the leaf: that of its first vertex (side-1/2 index along z, the walk
                    // and clamp of src/VFgeom.jl:853-862 along x and y for a cut vertex)
                    const int lv0 = nib(qv, 0);
                    if (lv0 < 4) {
                        ob = (N.bits >> (6 * lv0)) & 63u;
                        const unsigned oa = ob >> 4;
                        ob = (ob & 15u) | ((s2 ? (unsigned)mx : min(oa, (unsigned)(mx - 1))) << 4);
                    } else {
                        ob = (unsigned)(s2 ? mx : mx - 1) << 4;
#pragma unroll
                        for (int ax = 0; ax < 2; ++ax) {
                            const int ea = boff(N.bits, e1v, ax), eb = boff(N.bits, e2v, ax);
                            const unsigned oo = ea == eb ? (unsigned)ea
                                                         : walk_clamp(ax == 0 ? Q[0].x : Q[0].y, boff(N.bits, lv0 - 4, ax),
                                                                      ea, eb, S.plane[ax], span[ax]);
                            ob |= oo << (2 * ax);
                        }
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Q[v] = pool_v(S, tn_id(N, v));
                    ob = N.bits & 63u;
                }
                const unsigned ox = ob & 3u, oy = (ob >> 2) & 3u, oz = (ob >> 4) & 3u;
                const int l = (int)ox + bx * ((int)oy + by * (int)oz);
                const int lcls = S.cls[l];
                double uc = 0.0, vc = 0.0, wc = 0.0;
                if constexpr (MOM) {
                    uc = s_vel[l][0];
                    vc = s_vel[l][1];
                    wc = s_vel[l][2];
                }
                auto on_liq = [&](double tv) {
                    Vs += tv;
                    Ls += tv;
                    if constexpr (MOM) {
                        Ms[0] += (tv * uc) * lm.rho_liq; Ms[1] += (tv * vc) * lm.rho_liq;
                        Ms[2] += (tv * wc) * lm.rho_liq; Ms[3] += tv * lm.rho_liq;
                    }
                };
                auto on_gas = [&](double tv) {
                    Vs += tv;
                    if constexpr (MOM) {
                        Ms[0] += (tv * uc) * lm.rho_gas; Ms[1] += (tv * vc) * lm.rho_gas;
                        Ms[2] += (tv * wc) * lm.rho_gas; Ms[3] += tv * lm.rho_gas;
                    }
                };
                if (lcls != 0) {  // every vertex strictly on one side: the d2case case is 1 or 16
                    const double tv = tet_vol3(Q[0], Q[1], Q[2], Q[3]);
                    if (lcls > 0) on_liq(tv);
                    else on_gas(tv);
                    continue;
                }
                const double n0 = S.nb[l][0], n1 = S.nb[l][1], n2 = S.nb[l][2], Dc = S.nb[l][3];
                double d[4], dd[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    d[v] = ((n0 * Q[v].x + n1 * Q[v].y) + n2 * Q[v].z) - Dc;
                    dd[v] = d[v];
                }
                const int cs = d2case_snap(dd);
                if constexpr (!MOM) {
                    // the leaf's volume is its own tet's (the sum of its pieces to rounding) and its
                    // liquid volume that times its liquid share in closed form (the sum of the
                    // reference's liquid sub-tets in exact arithmetic)
                    const double tv = tet_vol3(Q[0], Q[1], Q[2], Q[3]);
                    Vs += tv;
                    Ls += cs == 0 ? tv : (cs == 15 ? 0.0 : tv * tet_liquid_fraction(cs, dd));
                } else if (cs == 0 || cs == 15) {  // one sub-tet, the leaf itself (vertex order 1,2,3,4)
                    const double tv = tet_vol3(Q[0], Q[1], Q[2], Q[3]);
                    if (cs == 0) on_liq(tv);
                    else on_gas(tv);
                } else {
                    cut_canonical(cs, Q, d, on_liq, on_gas);
                }
            }
            // ---- block sums in a fixed order ----------------------------------------------------
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                Vs += __shfl_xor(Vs, off);
                Ls += __shfl_xor(Ls, off);
                if constexpr (MOM) {
#pragma unroll
                    for (int a = 0; a < 4; ++a) Ms[a] += __shfl_xor(Ms[a], off);
                }
            }
            if constexpr (NT > 64) {
                if ((tid & 63) == 0) {
                    S.red[0][tid >> 6] = Vs;
                    S.red[1][tid >> 6] = Ls;
                    if constexpr (MOM) {
#pragma unroll
                        for (int a = 0; a < 4; ++a) s_red[a][tid >> 6] = Ms[a];
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    Vs = S.red[0][0];
                    Ls = S.red[1][0];
                    if constexpr (MOM) {
#pragma unroll
                        for (int a = 0; a < 4; ++a) Ms[a] = s_red[a][0];
                    }
                    for (int w = 1; w < NT / 64; ++w) {
                        Vs += S.red[0][w];
                        Ls += S.red[1][w];
                        if constexpr (MOM) {
#pragma unroll
                            for (int a = 0; a < 4; ++a) Ms[a] += s_red[a][w];
                        }
                    }
                }
            }
            if (tid == 0) {
                VF[m.vof(i, j, k)] = Ls / Vs;
                if constexpr (MOM) {  // the tetsign of the reference cancels in these ratios
                    const int64_t c = m.cell(i, j, k);
                    ns.us[c] = Ms[0] / Ms[3];
                    ns.vs[c] = Ms[1] / Ms[3];
                    ns.ws[c] = Ms[2] / Ms[3];
                }
            }
        }
        if (tid == 0) S.next = claim;
        __syncthreads();
        cidx = S.next;
    }
}

// ------------------------------------------------------------------------------------
// Initial conditions: VFcircle (src/Tools.jl:790-902), VFsphere / VFbubble3d (:907-922,
// :1061-1076). For the sampled shapes the result depends only on the number of samples
// inside, through a table of the reference's sequential sums; boxes far inside or outside
// the sphere (margin far above rounding) skip the 20^3 sample loop.
// ------------------------------------------------------------------------------------
__device__ double vf_circle(double xmin, double xmax, double ymin, double ymax, double rad, double xo, double yo) {
    const double dx = jmax(xmax - xmin, ymax - ymin);
    const double xm = 0.5 * (xmin + xmax), ym = 0.5 * (ymin + ymax);
    const double G = rad - sqrt((xm - xo) * (xm - xo) + (ym - yo) * (ym - yo));
    if (G > 2.0 * dx) return 1.0;
    if (G < -2.0 * dx) return 0.0;
    double VF = 0.0;
    bool xdone = false;
    int xcycl = 1;
    const double r2 = rad * rad;
    while (!xdone) {
        double x_min, x_max;
        if ((xmax - xo) * (xmin - xo) < 0.0) {
            if (xcycl == 1) { x_min = 0.0; x_max = fabs(xmin - xo); xcycl = 2; }
            else { x_min = 0.0; x_max = fabs(xmax - xo); xdone = true; }
        } else {
            x_min = jmin(fabs(xmin - xo), fabs(xmax - xo));
            x_max = jmax(fabs(xmin - xo), fabs(xmax - xo));
            xdone = true;
        }
        bool ydone = false;
        int ycycl = 1;
        while (!ydone) {
            double y_min, y_max;
            if ((ymin - yo) * (ymax - yo) < 0.0) {
                if (ycycl == 1) { y_min = 0.0; y_max = fabs(ymin - yo); ycycl = 2; }
                else { y_min = 0.0; y_max = fabs(ymax - yo); ydone = true; }
            } else {
                y_min = jmin(fabs(ymin - yo), fabs(ymax - yo));
                y_max = jmax(fabs(ymin - yo), fabs(ymax - yo));
                ydone = true;
            }
            double x_1, x_2;
            if (r2 - x_min * x_min > 0.0) {
                const double yint = sqrt(r2 - x_min * x_min);
                if (yint < y_min) continue;
                else if (yint < y_max) x_1 = x_min;
                else x_1 = sqrt(r2 - y_max * y_max);
            } else {
                continue;
            }
            if (r2 - x_max * x_max > 0.0) {
                const double yint = sqrt(r2 - x_max * x_max);
                if (yint > y_max) {
                    VF += ((x_max - x_min) * (y_max - y_min)) / ((xmax - xmin) * (ymax - ymin));
                    continue;
                } else if (yint > y_min) x_2 = x_max;
                else x_2 = sqrt(r2 - y_min * y_min);
            } else {
                x_2 = sqrt(r2 - y_min * y_min);
            }
            VF += (((((x_1 - x_min) * (y_max - y_min) + (r2 * asin(x_2 / rad)) / 2.0) - (r2 * asin(x_1 / rad)) / 2.0) -
                    (x_1 * pow(r2 - x_1 * x_1, 0.5)) / 2.0) +
                   (x_2 * pow(r2 - x_2 * x_2, 0.5)) / 2.0 - (y_min * (x_2 - x_1))) /
                  ((xmax - xmin) * (ymax - ymin));
        }
    }
    return VF;
}

__device__ double vf_sphere(const double* tab, double xmin, double xmax, double ymin, double ymax, double zmin,
                            double zmax, double rad, double xo, double yo, double zo) {
    auto clampd = [](double p, double a, double b) { return p < a ? a - p : (p > b ? p - b : 0.0); };
    const double ex = clampd(xo, xmin, xmax), ey = clampd(yo, ymin, ymax), ez = clampd(zo, zmin, zmax);
    const double dmin2 = ex * ex + ey * ey + ez * ez;
    const double fx = jmax(fabs(xmin - xo), fabs(xmax - xo));
    const double fy = jmax(fabs(ymin - yo), fabs(ymax - yo));
    const double fz = jmax(fabs(zmin - zo), fabs(zmax - zo));
    const double dmax2 = fx * fx + fy * fy + fz * fz;
    const double r2 = rad * rad;
    if (dmin2 > r2 * (1.0 + 1e-9)) return tab[0];
    if (dmax2 < r2 * (1.0 - 1e-9)) return tab[8000];
    int cnt = 0;
    for (int k = 1; k <= 20; ++k)
        for (int j = 1; j <= 20; ++j)
            for (int i = 1; i <= 20; ++i) {
                const double xh = xmin + ((double)i / 21.0) * (xmax - xmin);
                const double yh = ymin + ((double)j / 21.0) * (ymax - ymin);
                const double zh = zmin + ((double)k / 21.0) * (zmax - zmin);
                const double G = r2 - (((xh - xo) * (xh - xo) + (yh - yo) * (yh - yo)) + (zh - zo) * (zh - zo));
                cnt += G > 0.0;
            }
    return tab[cnt];
}

__global__ void k_init_vof(DevMesh m, int shape, double a0, double a1, double a2, double a3, const double* tab,
                           double* VF) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        double v;
        if (shape == MIST_IC_CIRCLE) v = vf_circle(m.x[i], m.x[i + 1], m.y[j], m.y[j + 1], a0, a1, a2);
        else v = vf_sphere(tab, m.x[i], m.x[i + 1], m.y[j], m.y[j + 1], m.z[k], m.z[k + 1], a0, a1, a2, a3);
        VF[m.vof(i, j, k)] = v;
    }
}

// compute_props! (src/Tools.jl:1186-1210): face density and viscosity from the clamped face
// average of VF
__global__ void k_props(DevMesh m, double rl, double rg, double ml, double mg, const double* VF, double* denx,
                        double* deny, double* denz, double* viscx, double* viscy, double* viscz) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        const double vc = VF[m.vof(i, j, k)];
        const double vx = jmin(1.0, jmax(0.0, (vc + VF[m.vof(i - 1, j, k)]) / 2.0));
        const double vy = jmin(1.0, jmax(0.0, (vc + VF[m.vof(i, j - 1, k)]) / 2.0));
        const double vz = jmin(1.0, jmax(0.0, (vc + VF[m.vof(i, j, k - 1)]) / 2.0));
        // face arrays carry one extra face along their axis
        const int a = i - 1 + m.ng, b = j - 1 + m.ng, c = k - 1 + m.ng;
        const int64_t fx = a + (int64_t)(m.cnx + 1) * (b + (int64_t)m.cny * c);
        const int64_t fy = a + (int64_t)m.cnx * (b + (int64_t)(m.cny + 1) * c);
        const int64_t fz = a + (int64_t)m.cnx * (b + (int64_t)m.cny * c);
        denx[fx] = rl * vx + rg * (1.0 - vx);
        deny[fy] = rl * vy + rg * (1.0 - vy);
        denz[fz] = rl * vz + rg * (1.0 - vz);
        viscx[fx] = vx * ml + (1.0 - vx) * mg;
        viscy[fy] = vy * ml + (1.0 - vy) * mg;
        viscz[fz] = vz * ml + (1.0 - vz) * mg;
    }
}

// FD right-hand side div(u*) (src/Pressure.jl:16-22)
__global__ void k_pressure_rhs(DevMesh m, const double* uf, const double* vf, const double* wf, double* rhs) {
    const int n = m.Nx * m.Ny * m.Nz;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        rhs[c] = ((uf[c + 1] - uf[c]) / m.dx + (vf[c + sy] - vf[c]) / m.dy) + (wf[c + sz] - wf[c]) / m.dz;
    }
}

// corrector! (src/Velocity.jl:3-20)
__global__ void k_corrector(DevMesh m, double dt, const double* P, const double* denx, const double* deny,
                            const double* denz, double* uf, double* vf, double* wf) {
    // ranges: x faces i in 0..N+2, j,k in 0..N+1 (reference indices imin_-1..imax_+2 etc.)
    const int nxr = m.Nx + 3, nyr = m.Ny + 3, nzr = m.Nz + 3;
    const int64_t n = (int64_t)nxr * nyr * nzr;
    const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nxr);
        const int j = (int)((q / nxr) % nyr);
        const int k = (int)(q / ((int64_t)nxr * nyr));
        const int64_t c = m.cell(i, j, k);
        const int a = i - 1 + m.ng, b = j - 1 + m.ng, cc = k - 1 + m.ng;
        if (i <= m.Nx + 2 && j <= m.Ny + 1 && k <= m.Nz + 1) {
            const double den = denx[a + (int64_t)(m.cnx + 1) * (b + (int64_t)m.cny * cc)];
            uf[c] = uf[c] - dt / den * ((P[c] - P[c - 1]) / m.dx);
        }
        if (i <= m.Nx + 1 && j <= m.Ny + 2 && k <= m.Nz + 1) {
            const double den = deny[a + (int64_t)m.cnx * (b + (int64_t)(m.cny + 1) * cc)];
            vf[c] = vf[c] - dt / den * ((P[c] - P[c - sy]) / m.dy);
        }
        if (i <= m.Nx + 1 && j <= m.Ny + 1 && k <= m.Nz + 2) {
            const double den = denz[a + (int64_t)m.cnx * (b + (int64_t)m.cny * cc)];
            wf[c] = wf[c] - dt / den * ((P[c] - P[c - sz]) / m.dz);
        }
    }
}

// Batched PLIC primitives
__global__ void k_plic_dist_batch(DevMesh m, DevParams p, const int64_t* ijk, const double* nrm, const double* vf,
                                  double* out, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = compute_dist(m, p, (int)ijk[3 * q], (int)ijk[3 * q + 1], (int)ijk[3 * q + 2], nrm[3 * q],
                              nrm[3 * q + 1], nrm[3 * q + 2], vf[q]);
}

__global__ void k_plic_vf_batch(DevMesh m, const int64_t* ijk, const double* nrm, const double* dist, double* out,
                                int64_t n) {
    const double vol = (m.dx * m.dy) * m.dz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)ijk[3 * q], j = (int)ijk[3 * q + 1], k = (int)ijk[3 * q + 2];
        const double cnx = nrm[3 * q], cny = nrm[3 * q + 1], cnz = nrm[3 * q + 2];
        out[q] = plic2vf_reg(m.x[i], m.x[i + 1], m.y[j], m.y[j + 1], m.z[k], m.z[k + 1], cnx, cny, cnz, dist[q],
                             ((i + j + k) % 2) == 0, vol);
    }
}

// ====================================================================================
// Launchers
// ====================================================================================
static inline int grid_for(int64_t n, int block, int cap = 65536) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

static Ext3 ext_of(mist_ctx* ctx, int field) {
    Ext3 e;
    for (int d = 0; d < 3; ++d) { e.lo[d] = (int)ctx->flo[field][d]; e.n[d] = (int)ctx->fn[field][d]; }
    return e;
}

#define LAUNCH(kernel, grid, block, ...)                                                 \
    do {                                                                                 \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ctx->stream, __VA_ARGS__); \
        MIST_HIP_TRY(hipGetLastError());                                                 \
    } while (0)

int launch_trig_tables(mist_ctx* ctx, int64_t n) {
    LAUNCH(k_trig_tables, grid_for(n, 256), 256, (const double*)ctx->geom, n, ctx->trig);
    return MIST_OK;
}

int launch_define_velocity(mist_ctx* ctx, double t) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    LAUNCH(k_define_velocity, grid_for(n, 256), 256, ctx->m, ctx->vt, (int)ctx->p.vf_velocity, ctx->p.vel_period, t,
           (double*)ctx->f[MIST_FIELD_U], (double*)ctx->f[MIST_FIELD_V], (double*)ctx->f[MIST_FIELD_W],
           (double*)ctx->f[MIST_FIELD_UF], (double*)ctx->f[MIST_FIELD_VF], (double*)ctx->f[MIST_FIELD_WF],
           (double*)nullptr);
    return MIST_OK;
}

// defineVelocity! with the maxima of |u|, |v|, |w| reduced behind it into ctx->h_vmax (pinned);
// ctx->vmax_ev marks their arrival.
int launch_define_velocity_vmax(mist_ctx* ctx, double t) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    const int g = std::min(ctx->red_blocks, grid_for(n, RED_THREADS));
    LAUNCH(k_define_velocity, g, RED_THREADS, ctx->m, ctx->vt, (int)ctx->p.vf_velocity, ctx->p.vel_period, t,
           (double*)ctx->f[MIST_FIELD_U], (double*)ctx->f[MIST_FIELD_V], (double*)ctx->f[MIST_FIELD_W],
           (double*)ctx->f[MIST_FIELD_UF], (double*)ctx->f[MIST_FIELD_VF], (double*)ctx->f[MIST_FIELD_WF], ctx->red);
    LAUNCH(k_reduce_partials, 3, RED_THREADS, ctx->red, g, 3, (int)MIST_REDUCE_MAX, ctx->red + 3 * ctx->red_blocks);
    MIST_HIP_TRY(hipMemcpyAsync(ctx->h_vmax, ctx->red + 3 * ctx->red_blocks, 3 * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipEventRecord(ctx->vmax_ev, ctx->stream));
    return MIST_OK;
}

int launch_maxabs_uvw(mist_ctx* ctx, double out[3]) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    const int g = std::min(ctx->red_blocks, grid_for(n, RED_THREADS));
    LAUNCH(k_maxabs3, g, RED_THREADS, (const double*)ctx->f[MIST_FIELD_U], (const double*)ctx->f[MIST_FIELD_V],
           (const double*)ctx->f[MIST_FIELD_W], n, ctx->red);
    LAUNCH(k_reduce_partials, 3, RED_THREADS, ctx->red, g, 3, (int)MIST_REDUCE_MAX, ctx->red + 3 * ctx->red_blocks);
    MIST_HIP_TRY(hipMemcpyAsync(ctx->h_red, ctx->red + 3 * ctx->red_blocks, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    out[0] = ctx->h_red[0]; out[1] = ctx->h_red[1]; out[2] = ctx->h_red[2];
    return MIST_OK;
}

int launch_reduce(mist_ctx* ctx, int kind, int field, double* out) {
    const int64_t n = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    const int g = std::min(ctx->red_blocks, grid_for(n, RED_THREADS));
    LAUNCH(k_reduce_interior, g, RED_THREADS, ctx->m, (const double*)ctx->f[field],
           (field == MIST_FIELD_VOF || field == MIST_FIELD_VOF_BAR) ? 1 : 0,
           kind, ctx->red);
    LAUNCH(k_reduce_partials, 1, RED_THREADS, ctx->red, g, 1, kind == MIST_REDUCE_MAXABS ? (int)MIST_REDUCE_MAX : kind,
           ctx->red + 3 * ctx->red_blocks);
    MIST_HIP_TRY(hipMemcpyAsync(ctx->h_red, ctx->red + 3 * ctx->red_blocks, sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_red[0];
    return MIST_OK;
}

int launch_neumann(mist_ctx* ctx, int field) {
    const Ext3 e = ext_of(ctx, field);
    const int64_t tot = (int64_t)e.n[1] * e.n[2] + (int64_t)e.n[2] * e.n[0] + (int64_t)e.n[0] * e.n[1];
    if (field == MIST_FIELD_BAND)
        LAUNCH(k_neumann3<int16_t>, grid_for(tot, 256), 256, (int16_t*)ctx->f[field], e, ctx->m.Nx, ctx->m.Ny, ctx->m.Nz);
    else
        LAUNCH(k_neumann3<double>, grid_for(tot, 256), 256, (double*)ctx->f[field], e, ctx->m.Nx, ctx->m.Ny, ctx->m.Nz);
    return MIST_OK;
}

// The 3 extra ghost layers of VF beyond the cell extents set to 0.
__global__ void k_clear_vof_outer(DevMesh m, double* VF) {
    const int64_t n = (int64_t)m.vnx * m.vny * m.vnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(q % m.vnx), b = (int)((q / m.vnx) % m.vny), c = (int)(q / ((int64_t)m.vnx * m.vny));
        if (a < 3 || b < 3 || c < 3 || a >= m.vnx - 3 || b >= m.vny - 3 || c >= m.vnz - 3) VF[q] = 0.0;
    }
}

int launch_clear_vof_outer(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.vnx * ctx->m.vny * ctx->m.vnz;
    LAUNCH(k_clear_vof_outer, grid_for(n, 256), 256, ctx->m, (double*)ctx->f[MIST_FIELD_VOF]);
    return MIST_OK;
}

int launch_neumann_cells(mist_ctx* ctx, double* A) {
    const Ext3 e = ext_of(ctx, MIST_FIELD_P);
    const int64_t tot = (int64_t)e.n[1] * e.n[2] + (int64_t)e.n[2] * e.n[0] + (int64_t)e.n[0] * e.n[1];
    LAUNCH(k_neumann3<double>, grid_for(tot, 256), 256, A, e, ctx->m.Nx, ctx->m.Ny, ctx->m.Nz);
    return MIST_OK;
}

int launch_borders(mist_ctx* ctx, int field) {
    const Ext3 e = ext_of(ctx, field);
    const int64_t per[3] = {ctx->p.xper, ctx->p.yper, ctx->p.zper};
    const int mx[3] = {ctx->m.Nx, ctx->m.Ny, ctx->m.Nz};
    const int ng = ctx->m.ng;
    for (int d = 0; d < 3; ++d) {
        if (!per[d]) continue;
        const int64_t tot = (int64_t)e.n[(d + 1) % 3] * e.n[(d + 2) % 3];
        if (field == MIST_FIELD_BAND)
            LAUNCH(k_borders_axis<int16_t>, grid_for(tot, 256), 256, (int16_t*)ctx->f[field], e, d, 1, mx[d], 1 - ng,
                   mx[d] + ng, ng);
        else
            LAUNCH(k_borders_axis<double>, grid_for(tot, 256), 256, (double*)ctx->f[field], e, d, 1, mx[d], 1 - ng,
                   mx[d] + ng, ng);
    }
    return MIST_OK;
}

int launch_band(mist_ctx* ctx) {
    ctx->vtx_list_ok = false;
    const DevMesh& m = ctx->m;
    const int n = (int)(m.Nx * m.Ny * m.Nz);
    int16_t* band = (int16_t*)ctx->f[MIST_FIELD_BAND];
    BandTile T{8, 8, 4, 0, 0};
    if (m.Nz < 2) T = BandTile{16, 16, 1, 0, 0};
    else if (m.Nz < 4) T = BandTile{16, 8, 2, 0, 0};
    T.nbx = (m.Nx + T.tx - 1) / T.tx;
    T.nby = (m.Ny + T.ty - 1) / T.ty;
    const int64_t tiles = (int64_t)T.nbx * T.nby * ((m.Nz + T.tz - 1) / T.tz);
    LAUNCH(k_band_classify_t, (int)tiles, 256, m, ctx->dp, T, (const double*)ctx->f[MIST_FIELD_VOF], band,
           ctx->counters + 0);
    int st = launch_neumann(ctx, MIST_FIELD_BAND);
    if (st) return st;
    if ((st = launch_borders(ctx, MIST_FIELD_BAND))) return st;
    // the last sweep also lists the band cells
    for (int s = 1; s <= ctx->dp.nband; ++s) {
        const bool last = s == ctx->dp.nband;
        LAUNCH(k_band_sweep, grid_for(n, 256), 256, m, s, band, last ? ctx->band_list : nullptr,
               last ? ctx->counters + 0 : nullptr);
        if ((st = launch_borders(ctx, MIST_FIELD_BAND))) return st;
    }
    if (ctx->dp.nband < 1)
        LAUNCH(k_band_list, grid_for(n, 256), 256, m, ctx->dp.nband, (const int16_t*)band, ctx->band_list,
               ctx->counters + 0);
    return MIST_OK;
}

int launch_normals(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    double* nx = (double*)ctx->f[MIST_FIELD_NX];
    double* ny = (double*)ctx->f[MIST_FIELD_NY];
    double* nz = (double*)ctx->f[MIST_FIELD_NZ];
    LAUNCH(k_fill_int, 1, 64, ctx->counters + 1, 1, 0);
    LAUNCH(k_normals_prep, grid_for(n, 256), 256, ctx->m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF], nx, ny, nz,
           ctx->iface_list, ctx->counters + 1);
    const int fam = (ctx->m.Nz > 1 ? 1 : 0) | (ctx->m.Ny > 1 ? 2 : 0) | (ctx->m.Nx > 1 ? 4 : 0);
    const int64_t ncap = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    if (!ctx->elv_ties && hipMalloc(&ctx->elv_ties, (size_t)ncap * 2 * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        ctx->elv_ties = nullptr;
        ctx->err = "ELVIRA tie list: device allocation failed";
        return MIST_ERR_OOM;
    }
    if (!ctx->elv_dist && hipMalloc(&ctx->elv_dist, (size_t)ncap * 27 * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        ctx->elv_dist = nullptr;
        ctx->err = "ELVIRA plane constants: device allocation failed";
        return MIST_ERR_OOM;
    }
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_ELVIRA);  // both kernels of ELVIRA
    LAUNCH(k_elvira_dist, grid_for(ncap * 27, 256, 8192), 256, ctx->m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF],
           (const int*)ctx->iface_list, (const int*)(ctx->counters + 1), fam, ctx->elv_dist, ctx->counters + 10,
           ctx->counters + 11);
    // the scoring kernel's work-groups take cells from a queue: as many as stay resident
    if (!ctx->elv_blocks) {
        int per = 0, dev = 0, cus = 0;
        MIST_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_elvira<true>, ELV_THREADS, 0));
        MIST_HIP_TRY(hipGetDevice(&dev));
        MIST_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ctx->elv_blocks = std::max(1, per) * std::max(1, cus);
    }
    const int elv_grid = grid_for(ncap, 1, ctx->elv_blocks);
    if (vol_is_pow2((ctx->m.dx * ctx->m.dy) * ctx->m.dz))
        LAUNCH(k_elvira<true>, elv_grid, ELV_THREADS, ctx->m, ctx->dp,
               (const double*)ctx->f[MIST_FIELD_VOF], (const int*)ctx->iface_list, (const int*)(ctx->counters + 1), fam,
               (const double*)ctx->elv_dist, nx, ny, nz, ctx->elv_best, ctx->elv_ties, ctx->counters + 10,
               ctx->counters + 11);
    else
        LAUNCH(k_elvira<false>, elv_grid, ELV_THREADS, ctx->m, ctx->dp,
               (const double*)ctx->f[MIST_FIELD_VOF], (const int*)ctx->iface_list, (const int*)(ctx->counters + 1), fam,
               (const double*)ctx->elv_dist, nx, ny, nz, ctx->elv_best, ctx->elv_ties, ctx->counters + 10,
               ctx->counters + 11);
    LAUNCH(k_elvira_ties, grid_for(ncap, 1, 2048), 64, ctx->m, (const double*)ctx->f[MIST_FIELD_VOF],
           (const int*)ctx->iface_list, fam, (const double*)ctx->elv_dist, (const int*)ctx->elv_ties,
           (const int*)(ctx->counters + 10), nx, ny, nz, ctx->elv_best);
    prof_end(ctx, MIST_PROF_ELVIRA, ev, ctx->counters + 1, 0);
    int st;
    if ((st = launch_borders(ctx, MIST_FIELD_NX))) return st;
    if ((st = launch_borders(ctx, MIST_FIELD_NY))) return st;
    return launch_borders(ctx, MIST_FIELD_NZ);
}

int launch_plic(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_PLIC);
    LAUNCH(k_plic, grid_for(n, 256), 256, ctx->m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF],
           (const double*)ctx->f[MIST_FIELD_NX], (const double*)ctx->f[MIST_FIELD_NY],
           (const double*)ctx->f[MIST_FIELD_NZ], (double*)ctx->f[MIST_FIELD_PLIC_D]);
    prof_end(ctx, MIST_PROF_PLIC, ev, nullptr, n);
    return launch_borders(ctx, MIST_FIELD_PLIC_D);
}

int launch_project_vertices(mist_ctx* ctx, double dt, bool reuse_list, bool count_cleared) {
    const int64_t nv = (int64_t)(ctx->m.Nx + 1) * (ctx->m.Ny + 1) * (ctx->m.Nz + 1);
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_PROJECT);
    if (!(reuse_list && ctx->vtx_list_ok)) {
        if (!count_cleared) LAUNCH(k_fill_int, 1, 64, ctx->counters + 9, 1, 0);
        LAUNCH(k_vertex_list, grid_for(nv, 256, 8192), 256, ctx->m, ctx->dp, (const int16_t*)ctx->f[MIST_FIELD_BAND],
               ctx->vtx_list, ctx->counters + 9);
    }
    ctx->vtx_list_ok = reuse_list;
    LAUNCH(k_project_list, grid_for(nv, 256, 2048), 256, ctx->m, ctx->dp, (const int*)ctx->vtx_list,
           (const int*)(ctx->counters + 9), (const double*)ctx->f[MIST_FIELD_UF], (const double*)ctx->f[MIST_FIELD_VF],
           (const double*)ctx->f[MIST_FIELD_WF], dt, ctx->pv);
    prof_end(ctx, MIST_PROF_PROJECT, ev, nullptr, nv);
    return MIST_OK;
}

int launch_transport(mist_ctx* ctx, double dt) {
    // deferred cells, work queues and cut-list size [4..8], and the back-trace's vertex count [9]
    LAUNCH(k_fill_int, 1, 64, ctx->counters + 4, 6, 0);
    if (int st = launch_project_vertices(ctx, dt, false, true)) return st;
    hipEvent_t ev;
    const int64_t ncap = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    ev = prof_begin(ctx, MIST_PROF_SL_CUT);
    if (!ctx->tree_blocks) {
        int per = 0, dev = 0, cus = 0;
        MIST_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sl_tree<TreeMain, false>, TreeMain::NT, 0));
        MIST_HIP_TRY(hipGetDevice(&dev));
        MIST_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ctx->tree_cus = std::max(1, cus);
        ctx->tree_blocks = std::max(1, per) * ctx->tree_cus;
    }
    const double* nx = (const double*)ctx->f[MIST_FIELD_NX];
    const double* ny = (const double*)ctx->f[MIST_FIELD_NY];
    const double* nz = (const double*)ctx->f[MIST_FIELD_NZ];
    const double* D = (const double*)ctx->f[MIST_FIELD_PLIC_D];
    double* VF = (double*)ctx->f[MIST_FIELD_VOF];
    CutCell* cut = (CutCell*)ctx->cut_list;
    // bulk cells, index boxes; cells to cut go to the cut list
    LAUNCH(k_sl_classify, grid_for(ncap, 256, 8192), 256, ctx->m, (const double*)ctx->pv, (const int*)ctx->band_list,
           (const int*)(ctx->counters + 0), nx, ny, nz, D, VF, ctx->counters + 2, cut, ctx->counters + 8,
           ctx->defer_list, ctx->counters + 6, 0);
    const TreeNS none{{nullptr, nullptr, nullptr, 0.0, 0.0}, nullptr, nullptr, nullptr, 0.0,
                      nullptr, nullptr, nullptr, nullptr};
    LAUNCH((k_sl_tree<TreeMain, false>), ctx->tree_blocks, TreeMain::NT, ctx->m, (const double*)ctx->pv,
           (const CutCell*)cut, (const int*)nullptr, (const int*)(ctx->counters + 8), nx, ny, nz, D, VF,
           ctx->iface_list, ctx->counters + 4, 0, ctx->counters + 5, none);
    prof_end(ctx, MIST_PROF_SL_CUT, ev, ctx->counters + 0, 0);
    // cells whose cut tree outgrows the main pass: the same walk with larger capacities, one
    // block per compute unit; then the general walk for those and for cells whose corners span
    // more than four cells
    ev = prof_begin(ctx, MIST_PROF_SL_GENERAL);
    LAUNCH((k_sl_tree<TreeWide, false>), ctx->tree_cus, TreeWide::NT, ctx->m, (const double*)ctx->pv, (const CutCell*)cut,
           (const int*)ctx->iface_list, (const int*)(ctx->counters + 4), nx, ny, nz, D, VF, ctx->defer_list,
           ctx->counters + 6, 1, ctx->counters + 7, none);
    LAUNCH(k_sl_cut, 32, 64, ctx->m, (const double*)ctx->pv, (const int*)ctx->defer_list,
           (const int*)(ctx->counters + 6), nx, ny, nz, D, VF, ctx->counters + 2);
    prof_end(ctx, MIST_PROF_SL_GENERAL, ev, ctx->counters + 6, 0);
    int st;
    if ((st = launch_neumann(ctx, MIST_FIELD_VOF))) return st;
    return launch_borders(ctx, MIST_FIELD_VOF);
}

// The momentum predictor's cut pass (src/Transport.jl:107-150): every band cell's five
// back-traced tets, with `corr` also its twelve correction tets, through the cut trees with the
// six momentum sums, writing VF and us, vs, ws (the classify / main tree / wide tree pipeline of
// launch_transport without the bulk shortcut). Cells left for the general walk are listed in
// defer_list (count counters[6]) for the caller.
int launch_tree_momentum(mist_ctx* ctx, bool corr, double dt) {
    const int64_t ncap = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    LAUNCH(k_fill_int, 1, 64, ctx->counters + 4, 5, 0);
    if (!ctx->tree_blocks_mom) {
        int per = 0, dev = 0, cus = 0;
        MIST_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sl_tree<TreeMain, true, true>, TreeMain::NT, 0));
        MIST_HIP_TRY(hipGetDevice(&dev));
        MIST_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ctx->tree_cus = std::max(1, cus);
        ctx->tree_blocks_mom = std::max(1, per) * ctx->tree_cus;
    }
    const double* nx = (const double*)ctx->f[MIST_FIELD_NX];
    const double* ny = (const double*)ctx->f[MIST_FIELD_NY];
    const double* nz = (const double*)ctx->f[MIST_FIELD_NZ];
    const double* D = (const double*)ctx->f[MIST_FIELD_PLIC_D];
    double* VF = (double*)ctx->f[MIST_FIELD_VOF];
    CutCell* cut = (CutCell*)ctx->cut_list;
    const TreeNS ns{{(const double*)ctx->f[MIST_FIELD_U], (const double*)ctx->f[MIST_FIELD_V],
                     (const double*)ctx->f[MIST_FIELD_W], ctx->p.rho_liq, ctx->p.rho_gas},
                    (const double*)ctx->f[MIST_FIELD_UF], (const double*)ctx->f[MIST_FIELD_VF],
                    (const double*)ctx->f[MIST_FIELD_WF], dt, (double*)ctx->f[MIST_FIELD_US],
                    (double*)ctx->f[MIST_FIELD_VS], (double*)ctx->f[MIST_FIELD_WS], ctx->counters + 2};
    LAUNCH(k_sl_classify, grid_for(ncap, 256, 8192), 256, ctx->m, (const double*)ctx->pv, (const int*)ctx->band_list,
           (const int*)(ctx->counters + 0), nx, ny, nz, D, VF, ctx->counters + 2, cut, ctx->counters + 8,
           ctx->defer_list, ctx->counters + 6, 1);
    // the wide pass gets one block per compute unit: without the bulk shortcut, and with the
    // correction tets, more cells outgrow the main pass
    if (corr) {
        LAUNCH((k_sl_tree<TreeMain, true, true>), ctx->tree_blocks_mom, TreeMain::NT, ctx->m, (const double*)ctx->pv,
               (const CutCell*)cut, (const int*)nullptr, (const int*)(ctx->counters + 8), nx, ny, nz, D, VF,
               ctx->iface_list, ctx->counters + 4, 0, ctx->counters + 5, ns);
        LAUNCH((k_sl_tree<TreeWide, true, true>), ctx->tree_cus, TreeWide::NT, ctx->m, (const double*)ctx->pv,
               (const CutCell*)cut, (const int*)ctx->iface_list, (const int*)(ctx->counters + 4), nx, ny, nz, D, VF,
               ctx->defer_list, ctx->counters + 6, 1, ctx->counters + 7, ns);
    } else {
        LAUNCH((k_sl_tree<TreeMain, true, false>), ctx->tree_blocks_mom, TreeMain::NT, ctx->m, (const double*)ctx->pv,
               (const CutCell*)cut, (const int*)nullptr, (const int*)(ctx->counters + 8), nx, ny, nz, D, VF,
               ctx->iface_list, ctx->counters + 4, 0, ctx->counters + 5, ns);
        LAUNCH((k_sl_tree<TreeWide, true, false>), ctx->tree_cus, TreeWide::NT, ctx->m, (const double*)ctx->pv,
               (const CutCell*)cut, (const int*)ctx->iface_list, (const int*)(ctx->counters + 4), nx, ny, nz, D, VF,
               ctx->defer_list, ctx->counters + 6, 1, ctx->counters + 7, ns);
    }
    return MIST_OK;
}

// The cut list of the band cells for the reverse pass: work counters cleared, bulk cells
// (zero adjoint) left out without writing VF, cells for the general walk in defer_list.
int launch_cut_list(mist_ctx* ctx) {
    const int64_t ncap = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    LAUNCH(k_fill_int, 1, 64, ctx->counters + 4, 5, 0);
    LAUNCH(k_sl_classify, grid_for(ncap, 256, 8192), 256, ctx->m, (const double*)ctx->pv,
           (const int*)ctx->band_list, (const int*)(ctx->counters + 0), (const double*)ctx->f[MIST_FIELD_NX],
           (const double*)ctx->f[MIST_FIELD_NY], (const double*)ctx->f[MIST_FIELD_NZ],
           (const double*)ctx->f[MIST_FIELD_PLIC_D], (double*)nullptr, ctx->counters + 2, (CutCell*)ctx->cut_list,
           ctx->counters + 8, ctx->defer_list, ctx->counters + 6, 0);
    return MIST_OK;
}

int launch_init_vof(mist_ctx* ctx, int shape, const double* args) {
    static double* d_tab = nullptr;  // per process: sequential sums of 1/8000 (sphere) and 1 - them (bubble)
    static int tab_dev = -1;
    if (!d_tab || tab_dev != ctx->device) {
        std::vector<double> tab(2 * 8001);
        const double sub = 1.0 / 8000.0;
        tab[0] = 0.0; tab[8001] = 1.0;
        for (int n = 1; n <= 8000; ++n) { tab[n] = tab[n - 1] + sub; tab[8001 + n] = tab[8001 + n - 1] - sub; }
        double* p = nullptr;
        MIST_HIP_TRY(hipMalloc(&p, tab.size() * sizeof(double)));
        MIST_HIP_TRY(hipMemcpy(p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        d_tab = p;
        tab_dev = ctx->device;
    }
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    const double* tab = shape == MIST_IC_BUBBLE3D ? d_tab + 8001 : d_tab;
    LAUNCH(k_init_vof, grid_for(n, 128), 128, ctx->m, shape, args[0], args[1], args[2],
           shape == MIST_IC_CIRCLE ? 0.0 : args[3], tab, (double*)ctx->f[MIST_FIELD_VOF]);
    return MIST_OK;
}

int launch_props(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    LAUNCH(k_props, grid_for(n, 256), 256, ctx->m, ctx->p.rho_liq, ctx->p.rho_gas, ctx->p.mu_liq, ctx->p.mu_gas,
           (const double*)ctx->f[MIST_FIELD_VOF], (double*)ctx->f[MIST_FIELD_DENX], (double*)ctx->f[MIST_FIELD_DENY],
           (double*)ctx->f[MIST_FIELD_DENZ], (double*)ctx->f[MIST_FIELD_VISCX], (double*)ctx->f[MIST_FIELD_VISCY],
           (double*)ctx->f[MIST_FIELD_VISCZ]);
    return MIST_OK;
}

int launch_pressure_rhs(mist_ctx* ctx) {
    const int n = ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    LAUNCH(k_pressure_rhs, grid_for(n, 256), 256, ctx->m, (const double*)ctx->f[MIST_FIELD_UF],
           (const double*)ctx->f[MIST_FIELD_VF], (const double*)ctx->f[MIST_FIELD_WF], (double*)ctx->f[MIST_FIELD_RHS]);
    return MIST_OK;
}

int launch_corrector(mist_ctx* ctx, double dt) {
    const int64_t n = (int64_t)(ctx->m.Nx + 3) * (ctx->m.Ny + 3) * (ctx->m.Nz + 3);
    LAUNCH(k_corrector, grid_for(n, 256), 256, ctx->m, dt, (const double*)ctx->f[MIST_FIELD_P],
           (const double*)ctx->f[MIST_FIELD_DENX], (const double*)ctx->f[MIST_FIELD_DENY],
           (const double*)ctx->f[MIST_FIELD_DENZ], (double*)ctx->f[MIST_FIELD_UF], (double*)ctx->f[MIST_FIELD_VF],
           (double*)ctx->f[MIST_FIELD_WF]);
    return MIST_OK;
}

int launch_plic_dist_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* vf, double* out,
                           int64_t n) {
    LAUNCH(k_plic_dist_batch, grid_for(n, 128), 128, ctx->m, ctx->dp, ijk, nrm, vf, out, n);
    return MIST_OK;
}

int launch_plic_vf_batch(mist_ctx* ctx, const int64_t* ijk, const double* nrm, const double* dist, double* out,
                         int64_t n) {
    LAUNCH(k_plic_vf_batch, grid_for(n, 128), 128, ctx->m, ijk, nrm, dist, out, n);
    return MIST_OK;
}

}  // namespace mist
