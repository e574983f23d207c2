// mist_vof.hip -- gfx950 kernels of the SL-VOF transport stage and its helpers:
// prescribed velocity, ghost fills, band flags, ELVIRA normals, PLIC constants,
// back-trace of cell vertices, iterative tet cutting, initial conditions, densities,
// reductions. Launchers at the bottom are called by the C ABI (mist_capi.hip).
#include <algorithm>
#include <cmath>
#include <vector>

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

// Wave-aggregated append of `flag` items to list[*count].
__device__ __forceinline__ void append(bool flag, int item, int* list, int* count) {
    const unsigned long long mask = __ballot(flag);
    if (mask == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, leader);
    if (flag) {
        const unsigned long long lower = mask & ((1ull << lane) - 1ull);
        list[base + __popcll(lower)] = item;
    }
}

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
        default: u = v = w = 0.0;
    }
}

__global__ void k_define_velocity(DevMesh m, int kind, double T, double t, double* u, double* v, double* w,
                                  double* uf, double* vf, double* wf) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        double a, b, c;
        vel_fun(kind, T, m.xm[i], m.ym[j], m.zm[k], t, a, b, c);
        u[q] = a; v[q] = b; w[q] = c;
        vel_fun(kind, T, m.x[i], m.ym[j], m.zm[k], t, a, b, c);
        uf[q] = a;
        vel_fun(kind, T, m.xm[i], m.y[j], m.zm[k], t, a, b, c);
        vf[q] = b;
        vel_fun(kind, T, m.xm[i], m.ym[j], m.z[k], t, a, b, c);
        wf[q] = c;
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

__global__ __launch_bounds__(RED_THREADS) void k_reduce_partials(const double* part, int n, int stride, int kind,
                                                                 double* out) {
    double r = red_init(kind);
    for (int q = threadIdx.x; q < n; q += blockDim.x) r = red_op(kind, r, part[q * stride]);
    const double b = block_reduce(kind, r);
    if (threadIdx.x == 0) *out = b;
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

template <class T>
__global__ void k_neumann_axis(T* A, Ext3 e, int d, int mn, int mx) {
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    const int64_t tot = (int64_t)e.n[d1] * e.n[d2];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int id[3];
        id[d1] = e.lo[d1] + (int)(q % e.n[d1]);
        id[d2] = e.lo[d2] + (int)(q / e.n[d1]);
        id[d] = mn;
        const T lo_v = A[e.at(id[0], id[1], id[2])];
        id[d] = mx;
        const T hi_v = A[e.at(id[0], id[1], id[2])];
        id[d] = mn - 1;
        A[e.at(id[0], id[1], id[2])] = lo_v;
        id[d] = mx + 1;
        A[e.at(id[0], id[1], id[2])] = hi_v;
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
__global__ void k_band_classify(DevMesh m, DevParams p, const double* VF, int16_t* band) {
    const int n = m.Nx * m.Ny * m.Nz;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const double v = VF[m.vof(i, j, k)];
        int b;
        if (v < p.VFlo) {
            b = -p.nband - 1;
            for (int kk = k - 1; kk <= k + 1; ++kk)
                for (int jj = j - 1; jj <= j + 1; ++jj)
                    for (int ii = i - 1; ii <= i + 1; ++ii)
                        if (VF[m.vof(ii, jj, kk)] >= p.VFhi) b = 0;
        } else if (v > p.VFhi) {
            b = p.nband + 1;
            for (int kk = k - 1; kk <= k + 1; ++kk)
                for (int jj = j - 1; jj <= j + 1; ++jj)
                    for (int ii = i - 1; ii <= i + 1; ++ii)
                        if (VF[m.vof(ii, jj, kk)] <= p.VFlo) b = 0;
        } else {
            b = 0;
        }
        band[m.cell(i, j, k)] = (int16_t)b;
    }
}

// One sweep of the band growth. In place: a sweep only lowers |band| from >n to n and only
// compares neighbours with n-1, so concurrent updates do not change any outcome.
__global__ void k_band_sweep(DevMesh m, int nstep, int16_t* band) {
    const int n = m.Nx * m.Ny * m.Nz;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, q, i, j, k);
        const int b = band[m.cell(i, j, k)];
        if (abs(b) > nstep) {
            bool hit = false;
            for (int kk = k - 1; kk <= k + 1; ++kk)
                for (int jj = j - 1; jj <= j + 1; ++jj)
                    for (int ii = i - 1; ii <= i + 1; ++ii)
                        hit |= abs((int)band[m.cell(ii, jj, kk)]) == nstep - 1;
            if (hit) band[m.cell(i, j, k)] = (int16_t)(nstep * (b > 0 ? 1 : -1));
        }
    }
}

__global__ void k_band_list(DevMesh m, int nband, const int16_t* band, int* list, int* count) {
    const int n = m.Nx * m.Ny * m.Nz;
    const int nb = (n + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole waves take part
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += gridDim.x * blockDim.x) {
        bool f = false;
        if (q < n) {
            int i, j, k;
            interior_of(m, q, i, j, k);
            f = abs((int)band[m.cell(i, j, k)]) <= nband;
        }
        append(f, q, list, count);
    }
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
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    const int64_t nb = (n + blockDim.x - 1) / blockDim.x * blockDim.x;
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
        append(iface, item, list, count);
    }
}

// ELVIRA_calc! + check_normal! (src/ELVIRA.jl:62-137). Two interface cells per wave: lane
// c < 27 of each half scores candidate c (family c/9: X-Y, X-Z, Y-Z; outer pair (c%9)/3,
// inner pair c%3, the reference's loop order). The 27 liquid fractions of the 3x3x3 block and
// the full-cell fractions of its cells sit in LDS. The best candidate is the lowest-numbered
// one with the smallest finite error, which is what the reference's strict `<` scan keeps.
__global__ __launch_bounds__(64) void k_elvira(DevMesh m, DevParams p, const double* VF, const int* list,
                                               const int* count, int fam_mask, double* nxo, double* nyo,
                                               double* nzo) {
    __shared__ double sVF[2][27];
    __shared__ double sFull[2][27];
    const int half = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const int ncell = *count;
    const double vol = (m.dx * m.dy) * m.dz;
    for (int base = blockIdx.x * 2; base < ncell; base += gridDim.x * 2) {
        const int cidx = base + half;
        const bool valid = cidx < ncell;
        int i = 1, j = 1, k = 1;
        if (valid) interior_of(m, list[cidx], i, j, k);
        if (valid && lane < 27) {
            const int ii = i - 1 + lane % 3, jj = j - 1 + (lane / 3) % 3, kk = k - 1 + lane / 9;
            sVF[half][lane] = VF[m.vof(ii, jj, kk)];
            // fraction of a cell entirely below the plane: case 1 for all five tets
            const double cx[2] = {m.x[ii], m.x[ii + 1]}, cy[2] = {m.y[jj], m.y[jj + 1]}, cz[2] = {m.z[kk], m.z[kk + 1]};
            double dc[8];
            for (int c = 0; c < 8; ++c) dc[c] = -1.0;
            sFull[half][lane] = plic2vf_corners(cx, cy, cz, dc, ((ii + jj + kk) % 2) == 0, 0.0, vol);
        }
        __syncthreads();
        double err = __builtin_huge_val();
        double bx = 0.0, by = 0.0, bz = 0.0;
        const int fam = lane / 9, a = (lane % 9) / 3, b = lane % 3;
        if (valid && lane < 27 && ((fam_mask >> fam) & 1)) {
            const double* V = sVF[half];
            auto at = [&](int di, int dj, int dk) { return V[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)]; };
            const int P0[3] = {-1, -1, 0}, P1[3] = {0, 1, 1};
            double cnx, cny, cnz;
            if (fam == 0) {  // X-Y stencils (src/ELVIRA.jl:71-78)
                const int i1 = P0[b], i2 = P1[b], j1 = P0[a], j2 = P1[a];
                const double dz = m.dz;
                const double sx2 = (at(i2, 0, -1) * dz + at(i2, 0, 0) * dz) + at(i2, 0, 1) * dz;
                const double sx1 = (at(i1, 0, -1) * dz + at(i1, 0, 0) * dz) + at(i1, 0, 1) * dz;
                cnx = -(sx2 - sx1) / (m.xm[i + i2] - m.xm[i + i1]);
                const double sy2 = (at(0, j2, -1) * dz + at(0, j2, 0) * dz) + at(0, j2, 1) * dz;
                const double sy1 = (at(0, j1, -1) * dz + at(0, j1, 0) * dz) + at(0, j1, 1) * dz;
                cny = -(sy2 - sy1) / (m.ym[j + j2] - m.ym[j + j1]);
                cnz = at(0, 0, 1) > at(0, 0, -1) ? -1.0 : 1.0;
            } else if (fam == 1) {  // X-Z stencils (src/ELVIRA.jl:81-88)
                const int i1 = P0[b], i2 = P1[b], k1 = P0[a], k2 = P1[a];
                const double dy = m.dy;
                const double sx2 = (at(i2, -1, 0) * dy + at(i2, 0, 0) * dy) + at(i2, 1, 0) * dy;
                const double sx1 = (at(i1, -1, 0) * dy + at(i1, 0, 0) * dy) + at(i1, 1, 0) * dy;
                cnx = -(sx2 - sx1) / (m.xm[i + i2] - m.xm[i + i1]);
                cny = at(0, 1, 0) > at(0, -1, 0) ? -1.0 : 1.0;
                const double sz2 = (at(0, -1, k2) * dy + at(0, 0, k2) * dy) + at(0, 1, k2) * dy;
                const double sz1 = (at(0, -1, k1) * dy + at(0, 0, k1) * dy) + at(0, 1, k1) * dy;
                cnz = -(sz2 - sz1) / (m.zm[k + k2] - m.zm[k + k1]);
            } else {  // Y-Z stencils (src/ELVIRA.jl:91-98)
                const int j1 = P0[b], j2 = P1[b], k1 = P0[a], k2 = P1[a];
                const double dx = m.dx;
                cnx = at(1, 0, 0) > at(-1, 0, 0) ? -1.0 : 1.0;
                const double sy2 = (at(-1, j2, 0) * dx + at(0, j2, 0) * dx) + at(1, j2, 0) * dx;
                const double sy1 = (at(-1, j1, 0) * dx + at(0, j1, 0) * dx) + at(1, j1, 0) * dx;
                cny = -(sy2 - sy1) / (m.ym[j + j2] - m.ym[j + j1]);
                const double sz2 = (at(-1, 0, k2) * dx + at(0, 0, k2) * dx) + at(1, 0, k2) * dx;
                const double sz1 = (at(-1, 0, k1) * dx + at(0, 0, k1) * dx) + at(1, 0, k1) * dx;
                cnz = -(sz2 - sz1) / (m.zm[k + k2] - m.zm[k + k1]);
            }
            // check_normal!
            const double norm = sqrt((cnx * cnx + cny * cny) + cnz * cnz);
            cnx = cnx / norm; cny = cny / norm; cnz = cnz / norm;
            const double dist = compute_dist(m, p, i, j, k, cnx, cny, cnz, at(0, 0, 0));
            double e = 0.0;
            for (int nb = 0; nb < 27; ++nb) {
                const int ii = i - 1 + nb % 3, jj = j - 1 + (nb / 3) % 3, kk = k - 1 + nb / 9;
                const double cx[2] = {m.x[ii], m.x[ii + 1]}, cy[2] = {m.y[jj], m.y[jj + 1]},
                             cz[2] = {m.z[kk], m.z[kk + 1]};
                double dc[8];
                bool allneg = true, allpos = true;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    dc[c] = ((cnx * cx[c & 1] + cny * cy[(c >> 1) & 1]) + cnz * cz[c >> 2]) - dist;
                    allneg &= dc[c] <= -EPS;
                    allpos &= dc[c] >= EPS;
                }
                double vf_;
                if (allneg) vf_ = sFull[half][nb];
                else if (allpos) vf_ = 0.0;
                else vf_ = plic2vf_corners(cx, cy, cz, dc, ((ii + jj + kk) % 2) == 0, 0.0, vol);
                const double df = V[nb] - vf_;
                e += df * df;
            }
            err = e;
            bx = cnx; by = cny; bz = cnz;
        }
        // argmin over the 32 lanes of this half: smallest finite error, lowest lane on ties
        const bool ok = err < __builtin_huge_val();
        double be = ok ? err : __builtin_huge_val();
        int bl = ok ? lane : 64;
        for (int off = 16; off > 0; off >>= 1) {
            const double oe = __shfl_xor(be, off, 32);
            const int ol = __shfl_xor(bl, off, 32);
            if (ol < 64 && (bl == 64 || oe < be || (oe == be && ol < bl))) { be = oe; bl = ol; }
        }
        const double wx = __shfl(bx, bl & 31, 32), wy = __shfl(by, bl & 31, 32), wz = __shfl(bz, bl & 31, 32);
        if (valid && lane == 0 && bl < 64) {
            const int64_t c = m.cell(i, j, k);
            nxo[c] = wx; nyo[c] = wy; nzo[c] = wz;
        }
        __syncthreads();
    }
}

// ELVIRA with one interface cell per work-group of 256 threads. Threads 0..26 load the
// 3x3x3 liquid fractions and build the 27 candidate normals and their PLIC constants once
// (computeDist); then half-wave h scores every candidate on neighbours h, h+8, h+16, h+24,
// writing (VF - VF_)^2 to LDS; the first half-wave sums each candidate's 27 terms in the
// reference's neighbour order and takes the lowest-numbered minimum. Neighbours that every
// candidate plane misses by a wide margin take the exact all-liquid / all-gas value.
constexpr int ELV_THREADS = 256;

__global__ __launch_bounds__(ELV_THREADS) void k_elvira2(DevMesh m, DevParams p, const double* VF, const int* list,
                                                         const int* count, int fam_mask, double* nxo, double* nyo,
                                                         double* nzo) {
    __shared__ double sVF[27];
    __shared__ double sFull[27];
    __shared__ double sC[27][4];
    __shared__ int sOk[27];
    __shared__ double sE[27][28];
    const int tid = threadIdx.x;
    const int half = tid >> 5, lane = tid & 31;
    const int ncell = *count;
    const double cellvol = (m.dx * m.dy) * m.dz;
    for (int cidx = blockIdx.x; cidx < ncell; cidx += gridDim.x) {
        int i, j, k;
        interior_of(m, list[cidx], i, j, k);
        if (tid < 27) {
            const int ii = i - 1 + tid % 3, jj = j - 1 + (tid / 3) % 3, kk = k - 1 + tid / 9;
            sVF[tid] = VF[m.vof(ii, jj, kk)];
            sFull[tid] = plic2vf_full(m.x[ii], m.x[ii + 1], m.y[jj], m.y[jj + 1], m.z[kk], m.z[kk + 1],
                                      ((ii + jj + kk) % 2) == 0, cellvol);
        }
        __syncthreads();
        if (tid < 27) {
            const int fam = tid / 9, a = (tid % 9) / 3, b = tid % 3;
            int ok = (fam_mask >> fam) & 1;
            if (ok) {
                auto at = [&](int di, int dj, int dk) { return sVF[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)]; };
                // stencil pairs [[-1,0],[-1,1],[0,1]] of src/ELVIRA.jl:72
                auto P0 = [](int q) { return q == 2 ? 0 : -1; };
                auto P1 = [](int q) { return q == 0 ? 0 : 1; };
                double cnx, cny, cnz;
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
                sC[tid][0] = cnx; sC[tid][1] = cny; sC[tid][2] = cnz;
                sC[tid][3] = compute_dist(m, p, i, j, k, cnx, cny, cnz, at(0, 0, 0));
            }
            sOk[tid] = ok;
        }
        __syncthreads();
        if (lane < 27 && sOk[lane]) {
            const double cnx = sC[lane][0], cny = sC[lane][1], cnz = sC[lane][2], dist = sC[lane][3];
            const double hx = fabs(cnx) * m.dx, hy = fabs(cny) * m.dy, hz = fabs(cnz) * m.dz;
            for (int nb = half; nb < 27; nb += 8) {
                const int ii = i - 1 + nb % 3, jj = j - 1 + (nb / 3) % 3, kk = k - 1 + nb / 9;
                const double dcen = ((cnx * m.xm[ii] + cny * m.ym[jj]) + cnz * m.zm[kk]) - dist;
                const double hext = 0.5 * ((hx + hy) + hz);
                const double margin =
                    1e-9 * (((1.0 + fabs(dist)) + fabs(m.xm[ii])) + (fabs(m.ym[jj]) + fabs(m.zm[kk])));
                double vf_;
                if (dcen - hext > margin) {
                    vf_ = 0.0;
                } else if (dcen + hext < -margin) {
                    vf_ = sFull[nb];
                } else {
                    vf_ = plic2vf_reg(m.x[ii], m.x[ii + 1], m.y[jj], m.y[jj + 1], m.z[kk], m.z[kk + 1], cnx, cny,
                                      cnz, dist, ((ii + jj + kk) % 2) == 0, cellvol);
                }
                const double df = sVF[nb] - vf_;
                sE[lane][nb] = df * df;
            }
        }
        __syncthreads();
        if (tid < 32) {
            double err = __builtin_huge_val();
            if (lane < 27 && sOk[lane]) {
                double e = 0.0;
                for (int nb = 0; nb < 27; ++nb) e += sE[lane][nb];
                err = e;
            }
            const bool ok = err < __builtin_huge_val();
            double be = ok ? err : __builtin_huge_val();
            int bl = ok ? lane : 64;
            for (int off = 16; off > 0; off >>= 1) {
                const double oe = __shfl_xor(be, off, 32);
                const int ol = __shfl_xor(bl, off, 32);
                if (ol < 64 && (bl == 64 || oe < be || (oe == be && ol < bl))) { be = oe; bl = ol; }
            }
            if (lane == 0 && bl < 64) {
                const int64_t c = m.cell(i, j, k);
                nxo[c] = sC[bl][0]; nyo[c] = sC[bl][1]; nzo[c] = sC[bl][2];
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

// ------------------------------------------------------------------------------------
// Back-trace of cell vertices (project!, src/Tools.jl:392-418). The back-traced position of
// a vertex does not depend on which cell asks for it (the interpolation index search
// converges to the same cell from any start), so each vertex of a band cell is traced once.
// ------------------------------------------------------------------------------------
__device__ void project_pt(const DevMesh& m, const DevParams& p, const double* uf, const double* vf,
                           const double* wf, double pt[3], int i, int j, int k, double dt) {
    if (p.zalesak) {
        const double pi2 = 2.0 * PI;
        const double x0 = pt[0], y0 = pt[1];
        pt[0] = (0.5 + cos(pi2 * dt) * (x0 - 0.5)) + sin(pi2 * dt) * (y0 - 0.5);
        pt[1] = (0.5 - sin(pi2 * dt) * (x0 - 0.5)) + cos(pi2 * dt) * (y0 - 0.5);
        return;
    }
    double v1[3], v2[3], v3[3], v4[3], q[3];
    if (p.projection == MIST_PROJ_RK4) {
        vel_face(m, uf, vf, wf, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v1[n];
        vel_face(m, uf, vf, wf, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v2[n];
        vel_face(m, uf, vf, wf, q, i, j, k, v3);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - dt * v3[n];
        vel_face(m, uf, vf, wf, q, i, j, k, v4);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + ((-dt) / 6.0) * (((v1[n] + 2.0 * v2[n]) + 2.0 * v3[n]) + v4[n]);
    } else if (p.projection == MIST_PROJ_EULER) {
        vel_face(m, uf, vf, wf, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + (-dt) * v1[n];
    } else {
        vel_face(m, uf, vf, wf, pt, i, j, k, v1);
        for (int n = 0; n < 3; ++n) q[n] = pt[n] - (0.5 * dt) * v1[n];
        vel_face(m, uf, vf, wf, q, i, j, k, v2);
        for (int n = 0; n < 3; ++n) pt[n] = pt[n] + (-dt) * v2[n];
    }
}

__device__ __forceinline__ int64_t vtx_id(const DevMesh& m, int I, int J, int K) {
    return (int64_t)(I - 1) + (int64_t)(m.Nx + 1) * ((int64_t)(J - 1) + (int64_t)(m.Ny + 1) * (K - 1));
}

__global__ void k_project_vertices(DevMesh m, DevParams p, const int16_t* band, const double* uf, const double* vf,
                                   const double* wf, double dt, double* pv) {
    const int64_t n = (int64_t)(m.Nx + 1) * (m.Ny + 1) * (m.Nz + 1);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(q % (m.Nx + 1)) + 1;
        const int64_t r = q / (m.Nx + 1);
        const int J = (int)(r % (m.Ny + 1)) + 1;
        const int K = (int)(r / (m.Ny + 1)) + 1;
        bool need = false;
        for (int c = 0; c < 8; ++c) {
            const int i = I - 1 + (c & 1), j = J - 1 + ((c >> 1) & 1), k = K - 1 + (c >> 2);
            if (i >= 1 && i <= m.Nx && j >= 1 && j <= m.Ny && k >= 1 && k <= m.Nz)
                need |= abs((int)band[m.cell(i, j, k)]) <= p.nband;
        }
        if (!need) continue;
        double pt[3] = {m.x[I], m.y[J], m.z[K]};
        project_pt(m, p, uf, vf, wf, pt, min(I, m.Nx), min(J, m.Ny), min(K, m.Nz), dt);
        pv[3 * q + 0] = pt[0];
        pv[3 * q + 1] = pt[1];
        pv[3 * q + 2] = pt[2];
    }
}

// ------------------------------------------------------------------------------------
// cutTet (src/VFgeom.jl:713-900) as an explicit depth-first walk. Each frame keeps what a
// recursion level of the reference keeps in vert[:,:,lvl] and vert_ind[:,:,:,lvl], plus the
// running sums of its children, so the sums nest exactly as the recursion adds them.
// ------------------------------------------------------------------------------------
struct CutFrame {
    double v[8][3];
    short vi[2][8][3];
    double acc_vol, acc_liq;
    unsigned char cs, child, nchild, flags;
};

__device__ __forceinline__ void cut_leaf(const DevMesh& m, const double* nx, const double* ny, const double* nz,
                                         const double* D, const double tet[4][3], const int ind0[3], double& vol,
                                         double& liq) {
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
    double V = 0.0, L = 0.0;
    const int hi = CUT_NTETS[cs], lo = CUT_NNTET[cs];
    for (int n = hi; n >= lo; --n) {
        const unsigned q = CUT_VTET[cs][n - 1];
        const double tv = tet_vol(vert[nib(q, 0)], vert[nib(q, 1)], vert[nib(q, 2)], vert[nib(q, 3)]);
        V += tv; L += tv;
    }
    for (int n = 1; n <= lo - 1; ++n) {
        const unsigned q = CUT_VTET[cs][n - 1];
        V += tet_vol(vert[nib(q, 0)], vert[nib(q, 1)], vert[nib(q, 2)], vert[nib(q, 3)]);
    }
    vol = V;
    liq = L;
}

// Returns false when the walk needs more than CUT_MAXLVL levels or a walk leaves the mesh.
__device__ bool cut_tet(const DevMesh& m, const double* nx, const double* ny, const double* nz, const double* D,
                        double tet[4][3], int ind[4][3], CutFrame* st, double& out_vol, double& out_liq,
                        int& depth) {
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
                double lv, ll;
                cut_leaf(m, nx, ny, nz, D, tet, ind[0], lv, ll);
                if (top < 0) { out_vol = lv; out_liq = ll; return ovf == 0; }
                st[top].acc_vol += lv;
                st[top].acc_liq += ll;
            } else {
                if (top + 1 >= CUT_MAXLVL) return false;
                CutFrame& F = st[++top];
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
                F.acc_vol = 0.0;
                F.acc_liq = 0.0;
            }
            have_node = false;
        }
        CutFrame& F = st[top];
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
            const double av = F.acc_vol, al = F.acc_liq;
            --top;
            if (top < 0) { out_vol = av; out_liq = al; return ovf == 0; }
            st[top].acc_vol += av;
            st[top].acc_liq += al;
        }
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
            CutFrame st[CUT_MAXLVL];
            int depth = 0;
            if (!cut_tet(m, nx, ny, nz, D, tet, ind, st, vol, liq, depth)) o = 1;
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

// ------------------------------------------------------------------------------------
// cutTet for tets whose vertices lie in at most two cells along each axis (all but a few
// pre-image tets at CFL <= 1): then the reference recursion is at most one x cut, one y cut
// and one z cut before the PLIC leaf, and every vertex index is lo or lo+1 per axis. The
// three levels are nested loops over sub-tets with their vertices in registers; a vertex's
// cell is three bits relative to lo. A cut vertex's index is the pt2index! walk clamped
// between the indices of its edge ends (src/VFgeom.jl:853-862), which for two candidate
// cells reduces to one comparison with the plane between them.
// ------------------------------------------------------------------------------------
struct FastTet {
    V3 p[4];
    unsigned bits;  // 3 bits per vertex (x, y, z offset from lo)
};

struct FastLevel {
    V3 q[8];
    unsigned bits1;  // side-1 bits of the 8 level vertices
    int cs, nchild, axis;
};

__device__ __forceinline__ V3 sel8(const V3 (&q)[8], int v) {
    V3 r = q[0];
#pragma unroll
    for (int k = 1; k < 8; ++k)
        if (v == k) r = q[k];
    return r;
}
__device__ __forceinline__ double comp(const V3& a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }

// True when the four vertices do not share their cell index along `axis`.
__device__ __forceinline__ bool spans(unsigned bits, int axis) {
    const unsigned b0 = (bits >> axis) & 1u, b1 = (bits >> (3 + axis)) & 1u, b2 = (bits >> (6 + axis)) & 1u,
                   b3 = (bits >> (9 + axis)) & 1u;
    return (b0 | b1 | b2 | b3) != (b0 & b1 & b2 & b3);
}

// Mesh-plane cut of node T along `axis` at the plane between lo and lo+1 (cut_ind = lo+1).
__device__ __forceinline__ void fast_cut(const FastTet& T, int axis, const double planes[3], FastLevel& F) {
    const double P = planes[axis];
    double d[4], dd[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) { d[n] = comp(T.p[n], axis) - P; dd[n] = d[n]; }
    const int cs = d2case_snap(dd);  // the reference snaps a copy
    F.cs = cs;
    F.axis = axis;
    F.nchild = CUT_NTETS[cs];
    // original vertices: side-1 index along the axis is min(ind, cut_ind - 1) = lo
    unsigned b1 = T.bits & ~((1u << axis) | (1u << (3 + axis)) | (1u << (6 + axis)) | (1u << (9 + axis)));
#pragma unroll
    for (int n = 0; n < 4; ++n) F.q[n] = T.p[n];
    const int nv = CUT_NVERT[cs];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        if (n < nv) {
            const unsigned e = CUT_EDGE[cs][n];
            const int v1 = e & 15, v2 = e >> 4;
            double d1 = d[0], d2 = d[0];
#pragma unroll
            for (int k = 1; k < 4; ++k) { if (v1 == k) d1 = d[k]; if (v2 == k) d2 = d[k]; }
            V3 a = T.p[0], b = T.p[0];
#pragma unroll
            for (int k = 1; k < 4; ++k) { if (v1 == k) a = T.p[k]; if (v2 == k) b = T.p[k]; }
            const double mu = jmin(1.0, jmax(0.0, sdiv(-d1, d2 - d1)));
            const V3 c = lerp3(a, b, mu);
            F.q[4 + n] = c;
            // index of the cut vertex: walk from vertex n's side-1 index, clamp to [v1, v2]
            const unsigned s = (b1 >> (3 * n)) & 7u;
            const unsigned e1 = (b1 >> (3 * v1)) & 7u, e2 = (b1 >> (3 * v2)) & 7u;
            unsigned nb = 0;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const unsigned u1 = (e1 >> ax) & 1u, u2 = (e2 >> ax) & 1u;
                unsigned bit;
                if (ax == axis) bit = 0u;  // side 1: cut_ind - 1
                else if (u1 == u2) bit = u1;
                else {
                    const double pc = comp(c, ax), pl = planes[ax];
                    bit = ((s >> ax) & 1u) ? (pc < pl - EPS ? 0u : 1u) : (pc > pl + EPS ? 1u : 0u);
                }
                nb |= bit << ax;
            }
            b1 |= nb << (3 * (4 + n));
        }
    }
    F.bits1 = b1;
}

__device__ __forceinline__ FastTet fast_child(const FastLevel& F, int n) {
    FastTet T;
    const unsigned q = CUT_VTET[F.cs][n];
    const bool side2 = (CUT_SIDE2[F.cs] >> n) & 1;
    unsigned bits = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int v = nib(q, k);
        T.p[k] = sel8(F.q, v);
        unsigned b = (F.bits1 >> (3 * v)) & 7u;
        if (side2) b |= 1u << F.axis;
        bits |= b << (3 * k);
    }
    T.bits = bits;
    return T;
}

__device__ __forceinline__ void fast_leaf(const DevMesh& m, const double* nx, const double* ny, const double* nz,
                                          const double* D, const FastTet& T, int lo0, int lo1, int lo2, double& vol,
                                          double& liq) {
    const unsigned b = T.bits & 7u;
    const int64_t c = m.cell(lo0 + (b & 1u), lo1 + ((b >> 1) & 1u), lo2 + ((b >> 2) & 1u));
    const double n0 = nx[c], n1 = ny[c], n2 = nz[c], Dc = D[c];
    double d[4], dd[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        d[n] = ((n0 * T.p[n].x + n1 * T.p[n].y) + n2 * T.p[n].z) - Dc;
        dd[n] = d[n];
    }
    const int cs = d2case_snap(dd);
    double V = 0.0, L = 0.0;
    cut_dispatch<true>(cs, T.p, d, [&](double tv) { V += tv; L += tv; }, [&](double tv) { V += tv; });
    vol = V;
    liq = L;
}

// Volumes of the tet after the x, y, z and PLIC cuts, summed as the recursion nests them.
__device__ void cut_tet_fast(const DevMesh& m, const double* nx, const double* ny, const double* nz, const double* D,
                             const FastTet& root, int lo0, int lo1, int lo2, double& vol, double& liq) {
    const double planes[3] = {m.x[lo0 + 1], m.y[lo1 + 1], m.z[lo2 + 1]};
    FastLevel FX, FY, FZ;
    const bool cx = spans(root.bits, 0);
    if (cx) fast_cut(root, 0, planes, FX);
    const int nxc = cx ? FX.nchild : 1;
    double vX = 0.0, lX = 0.0;
    for (int a = 0; a < nxc; ++a) {
        const FastTet TY = cx ? fast_child(FX, a) : root;
        const bool cy = spans(TY.bits, 1);
        if (cy) fast_cut(TY, 1, planes, FY);
        const int nyc = cy ? FY.nchild : 1;
        double vY = 0.0, lY = 0.0;
        for (int b = 0; b < nyc; ++b) {
            const FastTet TZ = cy ? fast_child(FY, b) : TY;
            const bool cz = spans(TZ.bits, 2);
            if (cz) fast_cut(TZ, 2, planes, FZ);
            const int nzc = cz ? FZ.nchild : 1;
            double vZ = 0.0, lZ = 0.0;
            for (int c = 0; c < nzc; ++c) {
                const FastTet TL = cz ? fast_child(FZ, c) : TZ;
                double v, l;
                fast_leaf(m, nx, ny, nz, D, TL, lo0, lo1, lo2, v, l);
                if (cz) { vZ += v; lZ += l; } else { vZ = v; lZ = l; }
            }
            if (cy) { vY += vZ; lY += lZ; } else { vY = vZ; lY = lZ; }
        }
        if (cx) { vX += vY; lX += lY; } else { vX = vY; lX = lY; }
    }
    vol = vX;
    liq = lX;
}

// SL volume fraction of band cells: 12 cells per wave, one lane per (cell, tet). Cells with a
// tet spanning more than two cells along an axis go to the general kernel (k_sl_cut) through
// `defer`.
__global__ __launch_bounds__(64) void k_sl_fast(DevMesh m, const double* pv, const int* list, const int* count,
                                                const double* nx, const double* ny, const double* nz,
                                                const double* D, double* VF, int* ovf, int* defer,
                                                int* ndefer) {
    __shared__ double s_vol[64], s_liq[64];
    __shared__ int s_gen[64];
    const int lane = threadIdx.x;
    const int cl = lane / 5, t = lane - 5 * cl;
    const int nb = *count;
    for (int base = blockIdx.x * 12; base < nb; base += gridDim.x * 12) {
        const int cidx = base + cl;
        const bool active = lane < 60 && cidx < nb;
        double vol = 0.0, liq = 0.0;
        int general = 0;
        int i = 1, j = 1, k = 1, q = 0;
        if (active) {
            q = list[cidx];
            interior_of(m, q, i, j, k);
            const bool even = ((i + j + k) % 2) == 0;
            const unsigned tv = TET5[t];
            FastTet T;
            int ind[4][3];
            int o = 0;
            for (int n = 0; n < 4; ++n) {
                const int c = vert_corner(nib(tv, n), even);
                const int64_t vq = vtx_id(m, i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
                T.p[n].x = pv[3 * vq + 0];
                T.p[n].y = pv[3 * vq + 1];
                T.p[n].z = pv[3 * vq + 2];
                const double pt[3] = {T.p[n].x, T.p[n].y, T.p[n].z};
                pt2index(m, ind[n], pt, i, j, k, &o);
            }
            if (o) atomicOr(ovf, 1);
            int lo[3];
            unsigned bits = 0;
            for (int a = 0; a < 3; ++a) {
                const int mn = min(min(ind[0][a], ind[1][a]), min(ind[2][a], ind[3][a]));
                const int mx = max(max(ind[0][a], ind[1][a]), max(ind[2][a], ind[3][a]));
                lo[a] = mn;
                if (mx - mn > 1) general = 1;
            }
            if (!general) {
                for (int n = 0; n < 4; ++n)
                    for (int a = 0; a < 3; ++a) bits |= (unsigned)(ind[n][a] - lo[a]) << (3 * n + a);
                T.bits = bits;
                cut_tet_fast(m, nx, ny, nz, D, T, lo[0], lo[1], lo[2], vol, liq);
            }
        }
        s_vol[lane] = vol;
        s_liq[lane] = liq;
        s_gen[lane] = general;
        __syncthreads();
        if (active && t == 0) {
            const int gen = s_gen[lane] | s_gen[lane + 1] | s_gen[lane + 2] | s_gen[lane + 3] | s_gen[lane + 4];
            if (gen) {
                defer[atomicAdd(ndefer, 1)] = q;
            } else {
                const double ts = ((i + j + k) % 2) == 0 ? 1.0 : -1.0;
                double V = 0.0, L = 0.0;
                for (int r = 0; r < 5; ++r) { V += ts * s_vol[lane + r]; L += ts * s_liq[lane + r]; }
                VF[m.vof(i, j, k)] = L / V;
            }
        }
        __syncthreads();
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

// compute_dens! (src/Tools.jl:1212-1232)
__global__ void k_dens(DevMesh m, double rl, double rg, const double* VF, double* denx, double* deny,
                       double* denz) {
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
        denx[a + (int64_t)(m.cnx + 1) * (b + (int64_t)m.cny * c)] = rl * vx + rg * (1.0 - vx);
        deny[a + (int64_t)m.cnx * (b + (int64_t)(m.cny + 1) * c)] = rl * vy + rg * (1.0 - vy);
        denz[a + (int64_t)m.cnx * (b + (int64_t)m.cny * c)] = rl * vz + rg * (1.0 - vz);
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

int launch_define_velocity(mist_ctx* ctx, double t) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    LAUNCH(k_define_velocity, grid_for(n, 256), 256, ctx->m, (int)ctx->p.vf_velocity, ctx->p.vel_period, t,
           (double*)ctx->f[MIST_FIELD_U], (double*)ctx->f[MIST_FIELD_V], (double*)ctx->f[MIST_FIELD_W],
           (double*)ctx->f[MIST_FIELD_UF], (double*)ctx->f[MIST_FIELD_VF], (double*)ctx->f[MIST_FIELD_WF]);
    return MIST_OK;
}

int launch_maxabs_uvw(mist_ctx* ctx, double out[3]) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    const int g = std::min(ctx->red_blocks, grid_for(n, RED_THREADS));
    LAUNCH(k_maxabs3, g, RED_THREADS, (const double*)ctx->f[MIST_FIELD_U], (const double*)ctx->f[MIST_FIELD_V],
           (const double*)ctx->f[MIST_FIELD_W], n, ctx->red);
    for (int c = 0; c < 3; ++c)
        LAUNCH(k_reduce_partials, 1, RED_THREADS, ctx->red + c, g, 3, (int)MIST_REDUCE_MAX, ctx->red + 3 * ctx->red_blocks + c);
    MIST_HIP_TRY(hipMemcpyAsync(ctx->h_red, ctx->red + 3 * ctx->red_blocks, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    out[0] = ctx->h_red[0]; out[1] = ctx->h_red[1]; out[2] = ctx->h_red[2];
    return MIST_OK;
}

int launch_reduce(mist_ctx* ctx, int kind, int field, double* out) {
    const int64_t n = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    const int g = std::min(ctx->red_blocks, grid_for(n, RED_THREADS));
    LAUNCH(k_reduce_interior, g, RED_THREADS, ctx->m, (const double*)ctx->f[field], field == MIST_FIELD_VOF ? 1 : 0,
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
    const int mn[3] = {1, 1, 1}, mx[3] = {ctx->m.Nx, ctx->m.Ny, ctx->m.Nz};
    for (int d = 0; d < 3; ++d) {
        const int64_t tot = (int64_t)e.n[(d + 1) % 3] * e.n[(d + 2) % 3];
        if (field == MIST_FIELD_BAND)
            LAUNCH(k_neumann_axis<int16_t>, grid_for(tot, 256), 256, (int16_t*)ctx->f[field], e, d, mn[d], mx[d]);
        else
            LAUNCH(k_neumann_axis<double>, grid_for(tot, 256), 256, (double*)ctx->f[field], e, d, mn[d], mx[d]);
    }
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
    const int n = (int)(ctx->m.Nx * ctx->m.Ny * ctx->m.Nz);
    int16_t* band = (int16_t*)ctx->f[MIST_FIELD_BAND];
    LAUNCH(k_band_classify, grid_for(n, 256), 256, ctx->m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF], band);
    int st = launch_neumann(ctx, MIST_FIELD_BAND);
    if (st) return st;
    if ((st = launch_borders(ctx, MIST_FIELD_BAND))) return st;
    for (int s = 1; s <= ctx->dp.nband; ++s) {
        LAUNCH(k_band_sweep, grid_for(n, 256), 256, ctx->m, s, band);
        if ((st = launch_borders(ctx, MIST_FIELD_BAND))) return st;
    }
    LAUNCH(k_fill_int, 1, 64, ctx->counters, 1, 0);
    LAUNCH(k_band_list, grid_for(n, 256), 256, ctx->m, ctx->dp.nband, (const int16_t*)band, ctx->band_list,
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
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_ELVIRA);
    LAUNCH(k_elvira2, grid_for(ncap, 1, 16384), ELV_THREADS, ctx->m, ctx->dp, (const double*)ctx->f[MIST_FIELD_VOF],
           (const int*)ctx->iface_list, (const int*)(ctx->counters + 1), fam, nx, ny, nz);
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

int launch_transport(mist_ctx* ctx, double dt) {
    const int64_t nv = (int64_t)(ctx->m.Nx + 1) * (ctx->m.Ny + 1) * (ctx->m.Nz + 1);
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_PROJECT);
    LAUNCH(k_project_vertices, grid_for(nv, 256), 256, ctx->m, ctx->dp, (const int16_t*)ctx->f[MIST_FIELD_BAND],
           (const double*)ctx->f[MIST_FIELD_UF], (const double*)ctx->f[MIST_FIELD_VF],
           (const double*)ctx->f[MIST_FIELD_WF], dt, ctx->pv);
    prof_end(ctx, MIST_PROF_PROJECT, ev, nullptr, nv);
    const int64_t ncap = (int64_t)ctx->m.Nx * ctx->m.Ny * ctx->m.Nz;
    ev = prof_begin(ctx, MIST_PROF_SL_CUT);
    LAUNCH(k_fill_int, 1, 64, ctx->counters + 4, 1, 0);
    LAUNCH(k_sl_fast, grid_for((ncap + 11) / 12, 64, 16384), 64, ctx->m, (const double*)ctx->pv,
           (const int*)ctx->band_list, (const int*)(ctx->counters + 0), (const double*)ctx->f[MIST_FIELD_NX],
           (const double*)ctx->f[MIST_FIELD_NY], (const double*)ctx->f[MIST_FIELD_NZ],
           (const double*)ctx->f[MIST_FIELD_PLIC_D], (double*)ctx->f[MIST_FIELD_VOF], ctx->counters + 2,
           ctx->iface_list, ctx->counters + 4);
    // cells with a tet spanning three or more cells along an axis: general walk
    LAUNCH(k_sl_cut, 256, 64, ctx->m, (const double*)ctx->pv, (const int*)ctx->iface_list,
           (const int*)(ctx->counters + 4), (const double*)ctx->f[MIST_FIELD_NX], (const double*)ctx->f[MIST_FIELD_NY],
           (const double*)ctx->f[MIST_FIELD_NZ], (const double*)ctx->f[MIST_FIELD_PLIC_D],
           (double*)ctx->f[MIST_FIELD_VOF], ctx->counters + 2);
    prof_end(ctx, MIST_PROF_SL_CUT, ev, ctx->counters + 0, 0);
    int st;
    if ((st = launch_neumann(ctx, MIST_FIELD_VOF))) return st;
    return launch_borders(ctx, MIST_FIELD_VOF);
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

int launch_dens(mist_ctx* ctx) {
    const int64_t n = (int64_t)ctx->m.cnx * ctx->m.cny * ctx->m.cnz;
    LAUNCH(k_dens, grid_for(n, 256), 256, ctx->m, ctx->p.rho_liq, ctx->p.rho_gas, (const double*)ctx->f[MIST_FIELD_VOF],
           (double*)ctx->f[MIST_FIELD_DENX], (double*)ctx->f[MIST_FIELD_DENY], (double*)ctx->f[MIST_FIELD_DENZ]);
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
