// mist_mg.hip -- variable-density pressure Poisson multigrid on gfx950.
//
// Operator (src/Pressure.jl:909-1033 scaling): (A p)_c = sum_f c_f (p_nb - p_c) with
// c_f = dt / (den_f h^2), wall faces dropped, periodic faces wrapped, b = div(u*)
// (src/Pressure.jl:16-22). V-cycle of mg_vc_lin! (src/Multigrid.jl:678-764): red-black
// Gauss-Seidel (src/Pressure.jl:1096-1133, colour (i+j+k)%2, red first, omega 1), 8-cell
// average restriction (src/Multigrid.jl:183-197), trilinear prolongation from coarse centres
// (src/Multigrid.jl:5-54). Coarse face coefficients are the mean of the fine faces covering
// them (MIST_COARSE_FACE_AVERAGE), or the reference's face densities of the liquid fraction
// restricted to the level (MIST_COARSE_RESTRICTED_VF, src/Multigrid.jl:725-737). The outer
// loop evaluates the FP64 residual in flux form and applies one V-cycle in
// precision T (double, or float for the mixed mode) to it per iteration, until
// max|b - A p| <= tol; the solution is shifted to zero mean.
//
// Storage: interior cells only (n_x n_y n_z per level, x fastest); neighbours across a wall
// have a zero coefficient and read the cell itself, so no ghost layers are needed. The levels
// relaxed by per-level launches store each x-row as its red cells then its black cells, so a
// one-colour sweep reads and writes unit-stride vectors; the coarse end (one work-group, in
// LDS) and the FP64 outer arrays use the plain order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mist_internal.h"

namespace mist {

struct LevelGeom {
    int n[3];
    int per[3];
    int coarsened[3];  // coarser than the level above along axis d
    int split;         // rows stored red half then black half (n[0] even)
    __host__ __device__ int64_t size() const { return (int64_t)n[0] * n[1] * n[2]; }
};

// storage index of cell (i, j, k); red = (i+1)+(j+1)+(k+1) even = i+j+k odd (second half)
__host__ __device__ __forceinline__ int64_t lidx(const LevelGeom& g, int i, int j, int k) {
    const int64_t row = (int64_t)g.n[0] * (j + (int64_t)g.n[1] * k);
    if (!g.split) return row + i;
    return row + (((i + j + k) & 1) ? (g.n[0] >> 1) : 0) + (i >> 1);
}

// cell (i, j, k) stored at index q
__device__ __forceinline__ void lcell(const LevelGeom& g, int64_t q, int& i, int& j, int& k) {
    const int64_t row = q / g.n[0];
    const int r = (int)(q - row * g.n[0]);
    j = (int)(row % g.n[1]);
    k = (int)(row / g.n[1]);
    if (!g.split) {
        i = r;
        return;
    }
    const int hn = g.n[0] >> 1;
    const int half = r >= hn ? 1 : 0;
    i = 2 * (r - half * hn) + ((half + j + k) & 1);
}

template <class T>
struct LevelBuf {
    T* c[3] = {nullptr, nullptr, nullptr};
    T* e = nullptr;
    T* b = nullptr;
    T* r = nullptr;
};

struct MGState {
    std::vector<LevelGeom> g;         // plain order (the FP64 outer arrays use g[0])
    std::vector<LevelGeom> gq[2];     // storage of the FP64 and FP32 hierarchies
    std::vector<LevelBuf<double>> d;  // FP64 hierarchy
    std::vector<LevelBuf<float>> s;   // FP32 hierarchy
    double* c64[3] = {nullptr, nullptr, nullptr};  // fine FP64 operator
    double* p64 = nullptr;
    double* b64 = nullptr;
    std::vector<double*> vfl;  // restricted liquid fraction per level (plain order; 1..)
    double dt = 0.0;           // of the solve being set up
    double* part = nullptr;
    double* h_part = nullptr;  // pinned
    int nblocks = 0;
    int64_t nfine = 0;
    int first_small[2] = {0, 0};              // first level solved in one block (FP64, FP32)
    size_t small_lds[2] = {0, 0};             // its LDS bytes
    void* small_desc[2] = {nullptr, nullptr};  // SmallLevels<T> in device memory
    hipGraph_t graph[2] = {nullptr, nullptr};  // one defect-correction iteration
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<void*> allocs;
};

void mg_free(MGState* s) {
    if (!s) return;
    for (int q = 0; q < 2; ++q) {
        if (s->exec[q]) (void)hipGraphExecDestroy(s->exec[q]);
        if (s->graph[q]) (void)hipGraphDestroy(s->graph[q]);
    }
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->h_part) (void)hipHostFree(s->h_part);
    delete s;
}

// ---- kernels ---------------------------------------------------------------------------
// Neighbours (storage indices) and face coefficients of cell (i, j, k): faces x-, x+, y-, y+,
// z-, z+; C[d][c] is the face between c and its lower neighbour along d. A neighbour across a
// wall is the cell itself with coefficient 0; across a periodic side, the wrapped cell.
template <class T, bool MAY_SPLIT = true>
__device__ __forceinline__ void stencil(const LevelGeom& g, T* const* C, int i, int j, int k, int64_t cell, int64_t nb[6],
                                       T cf[6]) {
    auto lidx = [&](int a, int b, int c) -> int64_t {
        if constexpr (MAY_SPLIT) return ::mist::lidx(g, a, b, c);
        return a + (int64_t)g.n[0] * (b + (int64_t)g.n[1] * c);
    };
    const int q[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int nd = g.n[d];
        int up[3] = {i, j, k}, dn[3] = {i, j, k};
        cf[2 * d] = C[d][cell];
        if (q[d] + 1 < nd || g.per[d]) {
            up[d] = q[d] + 1 < nd ? q[d] + 1 : 0;
            nb[2 * d + 1] = lidx(up[0], up[1], up[2]);
            cf[2 * d + 1] = C[d][nb[2 * d + 1]];
        } else {
            nb[2 * d + 1] = cell;
            cf[2 * d + 1] = T(0);
        }
        if (q[d] > 0 || g.per[d]) {
            dn[d] = q[d] > 0 ? q[d] - 1 : nd - 1;
            nb[2 * d] = lidx(dn[0], dn[1], dn[2]);
        } else {
            nb[2 * d] = cell;
        }
    }
}

// One colour of a red-black Gauss-Seidel sweep (rbgs_update!, src/Pressure.jl:1096-1133)
template <class T>
__global__ void k_rbgs(LevelGeom g, T* cx, T* cy, T* cz, T* e, const T* b, int color) {
    T* C[3] = {cx, cy, cz};
    const int hx = (g.n[0] + 1) / 2;
    const int64_t tot = (int64_t)hx * g.n[1] * g.n[2];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int h = (int)(q % hx);
        const int64_t row = q / hx;
        const int j = (int)(row % g.n[1]);
        const int k = (int)(row / g.n[1]);
        const int i = 2 * h + ((color + 1 + j + k) & 1);  // ((i+1)+(j+1)+(k+1)) % 2 == color
        if (i >= g.n[0]) continue;
        const int64_t cell = lidx(g, i, j, k);
        int64_t nb[6];
        T cf[6];
        stencil(g, C, i, j, k, cell, nb, cf);
        const T S = ((((cf[0] + cf[1]) + cf[2]) + cf[3]) + cf[4]) + cf[5];
        if (!(S > T(0))) continue;
        const T sum = ((((cf[0] * e[nb[0]] + cf[1] * e[nb[1]]) + cf[2] * e[nb[2]]) + cf[3] * e[nb[3]]) +
                       cf[4] * e[nb[4]]) + cf[5] * e[nb[5]];
        e[cell] = (sum - b[cell]) / S;
    }
}

// k_rbgs on a split level: thread (h, j, k) relaxes the cell of `color` at position h of row
// (j, k), i = 2h + p. Its x-neighbours sit at h-1+p and h+p of the other half of the same row,
// its y- and z-neighbours at h of the other half of the adjacent rows, so every load and the
// store are unit-stride across the wave; no index divisions.
template <class T>
__global__ __launch_bounds__(256) void k_rbgs_split(LevelGeom g, const T* __restrict__ cx, const T* __restrict__ cy,
                                                    const T* __restrict__ cz, T* e, const T* __restrict__ b,
                                                    int color) {
    const int hn = g.n[0] >> 1;
    const int h = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (h >= hn || j >= g.n[1]) return;
    const int p = (color + 1 + j + k) & 1;
    const int i = 2 * h + p;
    const int64_t sy = g.n[0], sz = (int64_t)g.n[0] * g.n[1];
    const int64_t row = sy * j + sz * k;
    const int own_off = color == 0 ? hn : 0, oth_off = hn - own_off;
    const int64_t cell = row + own_off + h;
    const int64_t oth = row + oth_off;
    // neighbours (storage index) and whether the face is open (not a wall)
    int64_t nxm = cell, nxp = cell, nym = cell, nyp = cell, nzm = cell, nzp = cell;
    bool oxm = true, oxp = true, oym = true, oyp = true, ozm = true, ozp = true;
    if (i > 0) nxm = oth + h - 1 + p; else if (g.per[0]) nxm = oth + hn - 1; else oxm = false;
    if (i < g.n[0] - 1) nxp = oth + h + p; else if (g.per[0]) nxp = oth; else oxp = false;
    if (j > 0) nym = oth - sy + h; else if (g.per[1]) nym = oth + sy * (g.n[1] - 1) + h; else oym = false;
    if (j < g.n[1] - 1) nyp = oth + sy + h; else if (g.per[1]) nyp = oth - sy * (g.n[1] - 1) + h; else oyp = false;
    if (k > 0) nzm = oth - sz + h; else if (g.per[2]) nzm = oth + sz * (g.n[2] - 1) + h; else ozm = false;
    if (k < g.n[2] - 1) nzp = oth + sz + h; else if (g.per[2]) nzp = oth - sz * (g.n[2] - 1) + h; else ozp = false;
    (void)oxm; (void)oym; (void)ozm;
    const T c0 = cx[cell], c1 = oxp ? cx[nxp] : T(0);
    const T c2 = cy[cell], c3 = oyp ? cy[nyp] : T(0);
    const T c4 = cz[cell], c5 = ozp ? cz[nzp] : T(0);
    const T S = ((((c0 + c1) + c2) + c3) + c4) + c5;
    if (!(S > T(0))) return;
    const T sum = ((((c0 * e[nxm] + c1 * e[nxp]) + c2 * e[nym]) + c3 * e[nyp]) + c4 * e[nzm]) + c5 * e[nzp];
    e[cell] = (sum - b[cell]) / S;
}

template <class T>
__global__ void k_resid(LevelGeom g, T* cx, T* cy, T* cz, const T* e, const T* b, T* r) {
    T* C[3] = {cx, cy, cz};
    const int64_t tot = g.size();
    for (int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < tot;
         cell += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        lcell(g, cell, i, j, k);
        int64_t nb[6];
        T cf[6];
        stencil(g, C, i, j, k, cell, nb, cf);
        const T pc = e[cell];
        const T Ap = ((((cf[0] * (e[nb[0]] - pc) + cf[1] * (e[nb[1]] - pc)) + cf[2] * (e[nb[2]] - pc)) +
                       cf[3] * (e[nb[3]] - pc)) + cf[4] * (e[nb[4]] - pc)) + cf[5] * (e[nb[5]] - pc);
        r[cell] = b[cell] - Ap;
    }
}

template <class T>
__global__ void k_restrict(LevelGeom F, LevelGeom C, const T* rf, T* bc) {
    const int f0 = C.coarsened[0] ? 2 : 1, f1 = C.coarsened[1] ? 2 : 1, f2 = C.coarsened[2] ? 2 : 1;
    const T inv = T(1) / T(f0 * f1 * f2);
    const int64_t tot = C.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int I, J, K;
        lcell(C, q, I, J, K);
        T s = T(0);
        for (int c = 0; c < f2; ++c)
            for (int b = 0; b < f1; ++b)
                for (int a = 0; a < f0; ++a) s += rf[lidx(F, I * f0 + a, J * f1 + b, K * f2 + c)];
        bc[q] = s * inv;
    }
}

template <class T>
__global__ void k_prolong_add(LevelGeom C, LevelGeom F, const T* ec, T* ef) {
    const int64_t tot = F.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int qq[3];
        lcell(F, q, qq[0], qq[1], qq[2]);
        int lo[3], hi[3];
        T w1[3], w2[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (!C.coarsened[d]) { lo[d] = hi[d] = qq[d]; w1[d] = T(1); w2[d] = T(0); continue; }
            const int Q = qq[d] / 2;
            if ((qq[d] & 1) == 0) { lo[d] = Q - 1; hi[d] = Q; w1[d] = T(0.75); w2[d] = T(0.25); }
            else { lo[d] = Q; hi[d] = Q + 1; w1[d] = T(0.25); w2[d] = T(0.75); }
            if (lo[d] < 0) lo[d] = C.per[d] ? C.n[d] - 1 : 0;
            if (hi[d] >= C.n[d]) hi[d] = C.per[d] ? 0 : C.n[d] - 1;
        }
        auto E = [&](int a, int b, int c) { return ec[lidx(C, a, b, c)]; };
        const T v = w1[2] * (w1[1] * (w1[0] * E(hi[0], hi[1], hi[2]) + w2[0] * E(lo[0], hi[1], hi[2])) +
                             w2[1] * (w1[0] * E(hi[0], lo[1], hi[2]) + w2[0] * E(lo[0], lo[1], hi[2]))) +
                    w2[2] * (w1[1] * (w1[0] * E(hi[0], hi[1], lo[2]) + w2[0] * E(lo[0], hi[1], lo[2])) +
                             w2[1] * (w1[0] * E(hi[0], lo[1], lo[2]) + w2[0] * E(lo[0], lo[1], lo[2])));
        ef[q] += v;
    }
}

template <class T>
__global__ void k_coarsen_coef(LevelGeom F, LevelGeom C, int d, const T* cf, T* cc) {
    const int f[3] = {C.coarsened[0] ? 2 : 1, C.coarsened[1] ? 2 : 1, C.coarsened[2] ? 2 : 1};
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    const T scale = T(1) / T(f[d] * f[d]) / T(f[d1] * f[d2]);
    const int64_t tot = C.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int Q[3];
        lcell(C, q, Q[0], Q[1], Q[2]);
        T s = T(0);
        for (int b = 0; b < f[d2]; ++b)
            for (int a = 0; a < f[d1]; ++a) {
                int x[3];
                x[d] = Q[d] * f[d];
                x[d1] = Q[d1] * f[d1] + a;
                x[d2] = Q[d2] * f[d2] + b;
                s += cf[lidx(F, x[0], x[1], x[2])];
            }
        cc[q] = s * scale;
    }
}

// The liquid fraction of level C (plain order): the mean of the finer level's over each coarse
// cell, summed in i, then j, then k order and scaled by 1/8 (restrict!,
// src/Multigrid.jl:183-197). From the context's VF (interior cells) when `src` is null.
__global__ void k_restrict_vf(DevMesh m, LevelGeom F, LevelGeom C, const double* vf_field, const double* src,
                              double* dst) {
    const int f0 = C.coarsened[0] ? 2 : 1, f1 = C.coarsened[1] ? 2 : 1, f2 = C.coarsened[2] ? 2 : 1;
    const double inv = 1.0 / (double)(f0 * f1 * f2);
    const int64_t tot = C.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(q % C.n[0]);
        const int J = (int)((q / C.n[0]) % C.n[1]);
        const int K = (int)(q / ((int64_t)C.n[0] * C.n[1]));
        double s = 0.0;
        for (int c = 0; c < f2; ++c)
            for (int b = 0; b < f1; ++b)
                for (int a = 0; a < f0; ++a) {
                    const int i = I * f0 + a, j = J * f1 + b, k = K * f2 + c;
                    s += src ? src[i + (int64_t)F.n[0] * (j + (int64_t)F.n[1] * k)]
                             : vf_field[m.vof(i + 1, j + 1, k + 1)];
                }
        dst[q] = s * inv;
    }
}

// Face coefficients of level C along d from its liquid fraction vf (plain order), in the
// level's storage: compute_dens! (src/Tools.jl:1212-1232) gives the face density
// rho_liq v + rho_gas (1 - v), v = clamp((vf_c + vf_{c-1}) / 2, 0, 1), and the coefficient is
// dt / (den h^2) as on the fine level; wall faces 0, periodic faces wrapped.
template <class T>
__global__ void k_coef_from_vf(LevelGeom C, int d, double dt, double rl, double rg, double h2, const double* vf,
                               T* c) {
    const int64_t tot = C.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int Q[3];
        lcell(C, q, Q[0], Q[1], Q[2]);
        int lo[3] = {Q[0], Q[1], Q[2]};
        bool wall = false;
        if (Q[d] > 0) lo[d] = Q[d] - 1;
        else if (C.per[d]) lo[d] = C.n[d] - 1;
        else wall = true;
        if (wall) {
            c[q] = T(0);
            continue;
        }
        const double a = vf[Q[0] + (int64_t)C.n[0] * (Q[1] + (int64_t)C.n[1] * Q[2])];
        const double b = vf[lo[0] + (int64_t)C.n[0] * (lo[1] + (int64_t)C.n[1] * lo[2])];
        const double v = fmin(1.0, fmax(0.0, (a + b) / 2));
        const double den = rl * v + rg * (1 - v);
        c[q] = (T)(dt / (den * h2));
    }
}

// fine coefficients from the face densities (FP64), plus the rhs and initial guess
__global__ void k_fine_setup(DevMesh m, LevelGeom g, double dt, const double* denx, const double* deny,
                             const double* denz, const double* rhs, const double* P, double* cx, double* cy,
                             double* cz, double* b, double* p) {
    const int64_t tot = g.size();
    const double h2[3] = {m.dx * m.dx, m.dy * m.dy, m.dz * m.dz};
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.n[0]) + 1;
        const int j = (int)((q / g.n[0]) % g.n[1]) + 1;
        const int k = (int)(q / ((int64_t)g.n[0] * g.n[1])) + 1;
        const int a = i - 1 + m.ng, bb = j - 1 + m.ng, c = k - 1 + m.ng;
        const double dx = denx[a + (int64_t)(m.cnx + 1) * (bb + (int64_t)m.cny * c)];
        const double dy = deny[a + (int64_t)m.cnx * (bb + (int64_t)(m.cny + 1) * c)];
        const double dz = denz[a + (int64_t)m.cnx * (bb + (int64_t)m.cny * c)];
        cx[q] = (i > 1 || g.per[0]) ? dt / (dx * h2[0]) : 0.0;
        cy[q] = (j > 1 || g.per[1]) ? dt / (dy * h2[1]) : 0.0;
        cz[q] = (k > 1 || g.per[2]) ? dt / (dz * h2[2]) : 0.0;
        b[q] = rhs[m.cell(i, j, k)];
        p[q] = P[m.cell(i, j, k)];
    }
}

// level storage of g from a plain-order FP64 array
template <class T>
__global__ void k_to_level(LevelGeom g, const double* a, T* b) {
    const int64_t n = g.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        lcell(g, q, i, j, k);
        b[q] = (T)a[i + (int64_t)g.n[0] * (j + (int64_t)g.n[1] * k)];
    }
}

// p64 += e (p64 in plain order, e in the storage of g)
template <class T>
__global__ void k_add_correction(LevelGeom g, const T* e, double* p) {
    const int64_t n = g.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        lcell(g, q, i, j, k);
        p[i + (int64_t)g.n[0] * (j + (int64_t)g.n[1] * k)] += (double)e[q];
    }
}

template <class T>
__global__ void k_zero(T* a, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        a[q] = T(0);
}

__global__ void k_shift(double* a, int64_t n, const double* sum) {
    const double mean = *sum / (double)n;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        a[q] -= mean;
}

constexpr int MG_RED = 256;
__device__ double mg_block_sum(double v, bool maxabs) {
    __shared__ double s[MG_RED];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = MG_RED / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] = maxabs ? fmax(s[threadIdx.x], s[threadIdx.x + w]) : s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// The outer loop's FP64 residual r = b - A p (plain order, flux form) handed straight to the
// V-cycle of precision T: bT = (T) r and eT = 0 in the storage of gT, max |r| of the block into
// part[blockIdx.x]; r itself is not stored.
template <class T>
__global__ __launch_bounds__(MG_RED) void k_outer_resid(LevelGeom g, LevelGeom gT, const double* cx, const double* cy,
                                                       const double* cz, const double* p, const double* b, T* bT,
                                                       T* eT, double* part) {
    const double* C[3] = {cx, cy, cz};
    const int64_t tot = g.size();
    double vmax = 0.0;
    for (int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < tot;
         cell += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(cell % g.n[0]);
        const int64_t row = cell / g.n[0];
        const int j = (int)(row % g.n[1]), k = (int)(row / g.n[1]);
        int64_t nb[6];
        double cf[6];
        stencil<double, false>(g, const_cast<double* const*>(C), i, j, k, cell, nb, cf);
        const double pc = p[cell];
        const double Ap = ((((cf[0] * (p[nb[0]] - pc) + cf[1] * (p[nb[1]] - pc)) + cf[2] * (p[nb[2]] - pc)) +
                            cf[3] * (p[nb[3]] - pc)) + cf[4] * (p[nb[4]] - pc)) + cf[5] * (p[nb[5]] - pc);
        const double r = b[cell] - Ap;
        const int64_t qs = lidx(gT, i, j, k);
        bT[qs] = (T)r;
        eT[qs] = T(0);
        vmax = fmax(vmax, fabs(r));
    }
    const double m = mg_block_sum(vmax, true);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

template <class T>
__global__ __launch_bounds__(MG_RED) void k_sum_part(const T* a, int64_t n, int maxabs, double* part) {
    double v = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        v = maxabs ? fmax(v, fabs((double)a[q])) : v + (double)a[q];
    const double r = mg_block_sum(v, maxabs != 0);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(MG_RED) void k_sum_final(const double* part, int n, int maxabs, double* out) {
    double v = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = maxabs ? fmax(v, part[q]) : v + part[q];
    const double r = mg_block_sum(v, maxabs != 0);
    if (threadIdx.x == 0) *out = r;
}

// ---- the coarse end of the V-cycle in one work-group --------------------------------------
// Levels from `first` down to the coarsest run in one block of MG_SMALL_THREADS threads with
// e, b and r of every such level in LDS (coefficients stay in global memory): the same
// smoothing, residual, restriction and prolongation as the per-level kernels, separated by
// block barriers instead of launches. The coarsest level removes the mean of b, zeroes e and
// runs `coarse_sweeps` red-black sweeps.
constexpr int MG_SMALL_THREADS = 1024;
constexpr int MG_SMALL_MAXLEV = 12;
constexpr size_t MG_SMALL_LDS = 120 * 1024;

template <class T>
struct SmallLevels {
    int count;  // number of levels handled in the block
    int nu1, nu2, coarse_sweeps;
    LevelGeom g[MG_SMALL_MAXLEV];
    T* c[MG_SMALL_MAXLEV][3];
    int off[MG_SMALL_MAXLEV];  // LDS element offset of level l: e, then b, then r
    const T* b_top;            // b of the first level (global, from the restriction)
    T* e_top;                  // e of the first level (global, for the prolongation)
};

// The cells of one colour that a thread relaxes on a level of at most 2 * MG_SMALL_THREADS
// cells per colour, with their stencils: loaded once per level visit, then every sweep reads
// only the neighbours' e from LDS.
template <class T>
struct CellSet {
    int cell[2];
    int nb[2][6];
    T cf[2][6];
    T S[2], b[2];
    bool on[2];
};

template <class T>
__device__ __forceinline__ void load_cells(const LevelGeom& g, T* const* C, const T* b, int color, int first,
                                           int stride, CellSet<T>& cs) {
    const int hx = (g.n[0] + 1) / 2;
    const int tot = hx * g.n[1] * g.n[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int q = first + s * stride;
        cs.on[s] = false;
        cs.cell[s] = 0;
        cs.S[s] = T(1);
        cs.b[s] = T(0);
#pragma unroll
        for (int f = 0; f < 6; ++f) { cs.nb[s][f] = 0; cs.cf[s][f] = T(0); }
        if (q >= tot) continue;
        const int h = q % hx, row = q / hx;
        const int j = row % g.n[1], k = row / g.n[1];
        const int i = 2 * h + ((color + 1 + j + k) & 1);
        if (i >= g.n[0]) continue;
        const int cell = i + g.n[0] * (j + g.n[1] * k);
        int64_t nb[6];
        T cf[6];
        stencil<T, false>(g, C, i, j, k, cell, nb, cf);
        const T S = ((((cf[0] + cf[1]) + cf[2]) + cf[3]) + cf[4]) + cf[5];
        if (!(S > T(0))) continue;
        cs.on[s] = true;
        cs.cell[s] = cell;
        cs.S[s] = S;
        cs.b[s] = b[cell];
#pragma unroll
        for (int f = 0; f < 6; ++f) { cs.nb[s][f] = (int)nb[f]; cs.cf[s][f] = cf[f]; }
    }
}

template <class T>
__device__ __forceinline__ void relax_cells(T* e, const CellSet<T>& cs) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (!cs.on[s]) continue;
        const T* cf = cs.cf[s];
        const int* nb = cs.nb[s];
        const T sum = ((((cf[0] * e[nb[0]] + cf[1] * e[nb[1]]) + cf[2] * e[nb[2]]) + cf[3] * e[nb[3]]) +
                       cf[4] * e[nb[4]]) + cf[5] * e[nb[5]];
        e[cs.cell[s]] = (sum - cs.b[s]) / cs.S[s];
    }
}

// Orders the LDS accesses of one wave (its lanes run in lockstep; the fences keep the
// compiler from moving loads above the previous phase's stores).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// `sweeps` red-black sweeps of a level in LDS. Levels of at most 64 cells run on wave 0 alone
// with wave-level ordering; larger ones (at most 2 * MG_SMALL_THREADS cells per colour) on the
// whole block with a barrier per colour.
template <class T>
__device__ __forceinline__ void blk_smooth(const LevelGeom& g, T* const* C, T* e, const T* b, int sweeps) {
    if (sweeps <= 0) return;
    const int tid = threadIdx.x;
    if (g.size() <= 64) {
        if (tid < 64) {
            CellSet<T> red, black;
            load_cells(g, C, b, 0, tid, 64, red);
            load_cells(g, C, b, 1, tid, 64, black);
            for (int sw = 0; sw < sweeps; ++sw) {
                relax_cells(e, red);
                wave_sync();
                relax_cells(e, black);
                wave_sync();
            }
        }
        __syncthreads();
        return;
    }
    CellSet<T> red, black;
    load_cells(g, C, b, 0, tid, MG_SMALL_THREADS, red);
    load_cells(g, C, b, 1, tid, MG_SMALL_THREADS, black);
    for (int sw = 0; sw < sweeps; ++sw) {
        relax_cells(e, red);
        __syncthreads();
        relax_cells(e, black);
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(MG_SMALL_THREADS) void k_small_vcycle(const SmallLevels<T>* P) {
    extern __shared__ double lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const int tid = threadIdx.x;
    const int nl = P->count;
    {
        const int n = (int)P->g[0].size();
        T* e = lds + P->off[0];
        T* b = e + n;
        for (int q = tid; q < n; q += MG_SMALL_THREADS) {
            b[q] = P->b_top[q];
            e[q] = T(0);
        }
        __syncthreads();
    }
    // down: smooth, residual, restrict
    for (int l = 0; l + 1 < nl; ++l) {
        const LevelGeom g = P->g[l];
        const int n = (int)g.size();
        T* C[3] = {P->c[l][0], P->c[l][1], P->c[l][2]};
        T* e = lds + P->off[l];
        T* b = e + n;
        T* r = b + n;
        blk_smooth(g, C, e, b, P->nu1);
        for (int cell = tid; cell < n; cell += MG_SMALL_THREADS) {
            const int i = cell % g.n[0], j = (cell / g.n[0]) % g.n[1], k = cell / (g.n[0] * g.n[1]);
            int64_t nb[6];
            T cf[6];
            stencil<T, false>(g, C, i, j, k, cell, nb, cf);
            const T pc = e[cell];
            const T Ap = ((((cf[0] * (e[nb[0]] - pc) + cf[1] * (e[nb[1]] - pc)) + cf[2] * (e[nb[2]] - pc)) +
                           cf[3] * (e[nb[3]] - pc)) + cf[4] * (e[nb[4]] - pc)) + cf[5] * (e[nb[5]] - pc);
            r[cell] = b[cell] - Ap;
        }
        __syncthreads();
        const LevelGeom gc = P->g[l + 1];
        const int nc = (int)gc.size();
        T* ec = lds + P->off[l + 1];
        T* bc = ec + nc;
        const int f0 = gc.coarsened[0] ? 2 : 1, f1 = gc.coarsened[1] ? 2 : 1, f2 = gc.coarsened[2] ? 2 : 1;
        const T inv = T(1) / T(f0 * f1 * f2);
        for (int q = tid; q < nc; q += MG_SMALL_THREADS) {
            const int I = q % gc.n[0], J = (q / gc.n[0]) % gc.n[1], K = q / (gc.n[0] * gc.n[1]);
            T sum = T(0);
            for (int c = 0; c < f2; ++c)
                for (int bb = 0; bb < f1; ++bb)
                    for (int a = 0; a < f0; ++a)
                        sum += r[(I * f0 + a) + g.n[0] * ((J * f1 + bb) + g.n[1] * (K * f2 + c))];
            bc[q] = sum * inv;
            ec[q] = T(0);
        }
        __syncthreads();
    }
    // coarsest: consistent right-hand side, then sweeps from zero
    {
        const int l = nl - 1;
        const LevelGeom g = P->g[l];
        const int n = (int)g.size();
        T* C[3] = {P->c[l][0], P->c[l][1], P->c[l][2]};
        T* e = lds + P->off[l];
        T* b = e + n;
        __shared__ double red[MG_SMALL_THREADS / 64];
        double v = 0.0;
        for (int q = tid; q < n; q += MG_SMALL_THREADS) v += (double)b[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
        if ((tid & 63) == 0) red[tid >> 6] = v;
        __syncthreads();
        double tot = 0.0;
        for (int w = 0; w < MG_SMALL_THREADS / 64; ++w) tot += red[w];
        const T mean = (T)(tot / (double)n);
        for (int q = tid; q < n; q += MG_SMALL_THREADS) {
            b[q] -= mean;
            e[q] = T(0);
        }
        __syncthreads();
        blk_smooth(g, C, e, (const T*)b, P->coarse_sweeps);
    }
    // up: prolongate and smooth
    for (int l = nl - 2; l >= 0; --l) {
        const LevelGeom g = P->g[l], gc = P->g[l + 1];
        const int n = (int)g.size();
        T* C[3] = {P->c[l][0], P->c[l][1], P->c[l][2]};
        T* e = lds + P->off[l];
        T* b = e + n;
        const T* ec = lds + P->off[l + 1];
        for (int q = tid; q < n; q += MG_SMALL_THREADS) {
            const int qq[3] = {q % g.n[0], (q / g.n[0]) % g.n[1], q / (g.n[0] * g.n[1])};
            int lo[3], hi[3];
            T w1[3], w2[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (!gc.coarsened[d]) { lo[d] = hi[d] = qq[d]; w1[d] = T(1); w2[d] = T(0); continue; }
                const int Q = qq[d] / 2;
                if ((qq[d] & 1) == 0) { lo[d] = Q - 1; hi[d] = Q; w1[d] = T(0.75); w2[d] = T(0.25); }
                else { lo[d] = Q; hi[d] = Q + 1; w1[d] = T(0.25); w2[d] = T(0.75); }
                if (lo[d] < 0) lo[d] = gc.per[d] ? gc.n[d] - 1 : 0;
                if (hi[d] >= gc.n[d]) hi[d] = gc.per[d] ? 0 : gc.n[d] - 1;
            }
            auto E = [&](int a, int bb, int c) { return ec[a + gc.n[0] * (bb + gc.n[1] * c)]; };
            const T v = w1[2] * (w1[1] * (w1[0] * E(hi[0], hi[1], hi[2]) + w2[0] * E(lo[0], hi[1], hi[2])) +
                                 w2[1] * (w1[0] * E(hi[0], lo[1], hi[2]) + w2[0] * E(lo[0], lo[1], hi[2]))) +
                        w2[2] * (w1[1] * (w1[0] * E(hi[0], hi[1], lo[2]) + w2[0] * E(lo[0], hi[1], lo[2])) +
                                 w2[1] * (w1[0] * E(hi[0], lo[1], lo[2]) + w2[0] * E(lo[0], lo[1], lo[2])));
            e[q] += v;
        }
        __syncthreads();
        blk_smooth(g, C, e, (const T*)b, P->nu2);
    }
    {
        const int n = (int)P->g[0].size();
        const T* e = lds + P->off[0];
        for (int q = tid; q < n; q += MG_SMALL_THREADS) P->e_top[q] = e[q];
    }
}

__global__ void k_store_p(DevMesh m, LevelGeom g, const double* p, double* P) {
    const int64_t tot = g.size();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.n[0]) + 1;
        const int j = (int)((q / g.n[0]) % g.n[1]) + 1;
        const int k = (int)(q / ((int64_t)g.n[0] * g.n[1])) + 1;
        P[m.cell(i, j, k)] = p[q];
    }
}

// ---- host driver -----------------------------------------------------------------------
static inline int gridn(int64_t n, int b = 256) {
    int64_t g = (n + b - 1) / b;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

#define MGL(kernel, grid, block, ...)                                                    \
    do {                                                                                 \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ctx->stream, __VA_ARGS__); \
        MIST_HIP_TRY(hipGetLastError());                                                 \
    } while (0)

template <class T>
static int alloc_level(mist_ctx* ctx, MGState* s, const LevelGeom& g, LevelBuf<T>& L) {
    const size_t bytes = (size_t)g.size() * sizeof(T);
    void** ptrs[6] = {(void**)&L.c[0], (void**)&L.c[1], (void**)&L.c[2], (void**)&L.e, (void**)&L.b, (void**)&L.r};
    for (void** p : ptrs) {
        if (hipMalloc(p, bytes) != hipSuccess) { ctx->err = "multigrid: device allocation failed"; return MIST_ERR_OOM; }
        s->allocs.push_back(*p);
    }
    return MIST_OK;
}

// The coarse end of the hierarchy that fits in MG_SMALL_LDS bytes of LDS (e, b, r per level)
// goes to k_small_vcycle; its descriptor is copied to device memory once.
template <class T>
static int build_small(mist_ctx* ctx, MGState* s, std::vector<LevelBuf<T>>& L, int q) {
    const int nlev = (int)s->g.size();
    int first = nlev - 1;
    size_t bytes = 3 * sizeof(T) * (size_t)s->g[first].size();
    const int64_t maxcells = 4 * MG_SMALL_THREADS;  // two cells per thread and colour
    while (first > 0 && nlev - first < MG_SMALL_MAXLEV && s->g[first - 1].size() <= maxcells &&
           bytes + 3 * sizeof(T) * (size_t)s->g[first - 1].size() <= MG_SMALL_LDS) {
        --first;
        bytes += 3 * sizeof(T) * (size_t)s->g[first].size();
    }
    if (bytes > MG_SMALL_LDS || s->g[first].size() > maxcells) { ctx->err = "multigrid: coarsest level too large for one work-group (raise mg_lvl)"; return MIST_ERR_PARAM; }
    SmallLevels<T> h{};
    h.count = nlev - first;
    h.nu1 = (int)ctx->p.mg_nu1;
    h.nu2 = (int)ctx->p.mg_nu2;
    h.coarse_sweeps = (int)ctx->p.mg_coarse_sweeps;
    int off = 0;
    for (int l = first; l < nlev; ++l) {
        h.g[l - first] = s->g[l];
        for (int d = 0; d < 3; ++d) h.c[l - first][d] = L[l].c[d];
        h.off[l - first] = off;
        off += 3 * (int)s->g[l].size();
    }
    h.b_top = L[first].b;
    h.e_top = L[first].e;
    void* dsc = nullptr;
    if (hipMalloc(&dsc, sizeof(h)) != hipSuccess) { ctx->err = "multigrid: device allocation failed"; return MIST_ERR_OOM; }
    s->allocs.push_back(dsc);
    MIST_HIP_TRY(hipMemcpy(dsc, &h, sizeof(h), hipMemcpyHostToDevice));
    MIST_HIP_TRY(hipFuncSetAttribute((const void*)k_small_vcycle<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    s->first_small[q] = first;
    s->small_lds[q] = bytes;
    s->small_desc[q] = dsc;
    return MIST_OK;
}

static int build_state(mist_ctx* ctx) {
    if (ctx->mg) return MIST_OK;
    const mist_params& p = ctx->p;
    MGState* s = new MGState();
    LevelGeom g0{};
    g0.n[0] = (int)p.Nx; g0.n[1] = (int)p.Ny; g0.n[2] = (int)p.Nz;
    g0.per[0] = p.xper && p.Nx > 1; g0.per[1] = p.yper && p.Ny > 1; g0.per[2] = p.zper && p.Nz > 1;
    int64_t nlev = p.mg_lvl;
    if (nlev <= 1) {
        nlev = 1;
        int n[3] = {g0.n[0], g0.n[1], g0.n[2]};
        for (;;) {
            const int mx = std::max(n[0], std::max(n[1], n[2]));
            bool ok = mx > 4;
            for (int d = 0; d < 3; ++d) if (n[d] > 1 && (n[d] % 2)) ok = false;
            if (!ok) break;
            for (int d = 0; d < 3; ++d) if (n[d] > 1) n[d] /= 2;
            ++nlev;
        }
    }
    s->g.push_back(g0);
    for (int64_t l = 1; l < nlev; ++l) {
        LevelGeom c{};
        const LevelGeom& f = s->g.back();
        for (int d = 0; d < 3; ++d) {
            if (f.n[d] > 1 && (f.n[d] % 2)) { mg_free(s); ctx->err = "multigrid: odd extent cannot be coarsened"; return MIST_ERR_PARAM; }
            c.coarsened[d] = f.n[d] > 1;
            c.n[d] = f.n[d] > 1 ? f.n[d] / 2 : 1;
            c.per[d] = f.per[d] && c.n[d] > 1;
        }
        s->g.push_back(c);
    }
    s->d.resize(s->g.size());
    s->s.resize(s->g.size());
    s->vfl.assign(s->g.size(), nullptr);
    if (p.mg_coarse_op == MIST_COARSE_RESTRICTED_VF)
        for (size_t l = 1; l < s->g.size(); ++l) {
            if (hipMalloc((void**)&s->vfl[l], (size_t)s->g[l].size() * sizeof(double)) != hipSuccess) { mg_free(s); ctx->err = "multigrid: device allocation failed"; return MIST_ERR_OOM; }
            s->allocs.push_back(s->vfl[l]);
        }
    int st;
    for (size_t l = 0; l < s->g.size(); ++l) {
        if ((st = alloc_level(ctx, s, s->g[l], s->d[l]))) { mg_free(s); return st; }
        if ((st = alloc_level(ctx, s, s->g[l], s->s[l]))) { mg_free(s); return st; }
    }
    if ((st = build_small<double>(ctx, s, s->d, 0)) || (st = build_small<float>(ctx, s, s->s, 1))) {
        mg_free(s);
        return st;
    }
    for (int q = 0; q < 2; ++q) {
        s->gq[q] = s->g;
        for (int l = 0; l < s->first_small[q]; ++l) s->gq[q][l].split = s->g[l].n[0] % 2 == 0;
    }
    s->nfine = g0.size();
    const size_t fb = (size_t)s->nfine * sizeof(double);
    double** f64[5] = {&s->c64[0], &s->c64[1], &s->c64[2], &s->p64, &s->b64};
    for (double** q : f64) {
        if (hipMalloc((void**)q, fb) != hipSuccess) { mg_free(s); ctx->err = "multigrid: device allocation failed"; return MIST_ERR_OOM; }
        s->allocs.push_back(*q);
    }
    s->nblocks = 4096;
    if (hipMalloc((void**)&s->part, (s->nblocks + 8) * sizeof(double)) != hipSuccess) { mg_free(s); ctx->err = "multigrid: device allocation failed"; return MIST_ERR_OOM; }
    s->allocs.push_back(s->part);
    if (hipHostMalloc((void**)&s->h_part, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) { mg_free(s); ctx->err = "multigrid: pinned allocation failed"; return MIST_ERR_OOM; }
    ctx->mg = s;
    return MIST_OK;
}

// sum (maxabs = 0) or max |.| (maxabs = 1) of a device vector, result in s->part[nblocks]
template <class T>
static int dev_reduce(mist_ctx* ctx, const T* a, int64_t n, int maxabs) {
    MGState* s = ctx->mg;
    const int g = std::min<int>(s->nblocks, gridn(n, MG_RED));
    MGL(k_sum_part<T>, g, MG_RED, a, n, maxabs, s->part);
    MGL(k_sum_final, 1, MG_RED, (const double*)s->part, g, maxabs, s->part + s->nblocks);
    return MIST_OK;
}

// `sweeps` red-black sweeps on level g (one launch per colour, red first)
template <class T>
static int smooth(mist_ctx* ctx, const LevelGeom& g, LevelBuf<T>& B, int sweeps, bool prof) {
    const int64_t nh = (int64_t)((g.n[0] + 1) / 2) * g.n[1] * g.n[2];
    const dim3 grid((g.n[0] / 2 + 63) / 64, (g.n[1] + 3) / 4, g.n[2]);
    for (int sw = 0; sw < sweeps; ++sw) {
        hipEvent_t ev = prof ? prof_begin(ctx, MIST_PROF_MG_RBGS) : nullptr;
        for (int c = 0; c < 2; ++c) {
            if (g.split) {
                hipLaunchKernelGGL(k_rbgs_split<T>, grid, dim3(64, 4), 0, ctx->stream, g, (const T*)B.c[0],
                                   (const T*)B.c[1], (const T*)B.c[2], B.e, (const T*)B.b, c);
                MIST_HIP_TRY(hipGetLastError());
            } else {
                MGL(k_rbgs<T>, gridn(nh), 256, g, B.c[0], B.c[1], B.c[2], B.e, (const T*)B.b, c);
            }
        }
        prof_end(ctx, MIST_PROF_MG_RBGS, ev, nullptr, g.size());
    }
    return MIST_OK;
}

template <class T>
static int vcycle(mist_ctx* ctx, std::vector<LevelBuf<T>>& L, size_t l) {
    MGState* s = ctx->mg;
    const int q = sizeof(T) == sizeof(double) ? 0 : 1;
    const LevelGeom& g = s->gq[q][l];
    LevelBuf<T>& B = L[l];
    const int nu1 = (int)ctx->p.mg_nu1, nu2 = (int)ctx->p.mg_nu2;
    const int64_t n = g.size();
    if ((int)l == s->first_small[q]) {
        hipLaunchKernelGGL(k_small_vcycle<T>, dim3(1), dim3(MG_SMALL_THREADS), s->small_lds[q], ctx->stream,
                           (const SmallLevels<T>*)s->small_desc[q]);
        MIST_HIP_TRY(hipGetLastError());
        return MIST_OK;
    }
    if (l > 0) MGL(k_zero<T>, gridn(n), 256, B.e, n);
    int st;
    if ((st = smooth(ctx, g, B, nu1, l == 0))) return st;
    MGL(k_resid<T>, gridn(n), 256, g, B.c[0], B.c[1], B.c[2], (const T*)B.e, (const T*)B.b, B.r);
    const LevelGeom& gc = s->gq[q][l + 1];
    MGL(k_restrict<T>, gridn(gc.size()), 256, g, gc, (const T*)B.r, L[l + 1].b);
    if ((st = vcycle(ctx, L, l + 1))) return st;
    MGL(k_prolong_add<T>, gridn(n), 256, gc, g, (const T*)L[l + 1].e, B.e);
    return smooth(ctx, g, B, nu2, l == 0);
}

template <class T>
static int setup_hierarchy(mist_ctx* ctx, std::vector<LevelBuf<T>>& L) {
    MGState* s = ctx->mg;
    const std::vector<LevelGeom>& G = s->gq[sizeof(T) == sizeof(double) ? 0 : 1];
    for (int d = 0; d < 3; ++d) MGL(k_to_level<T>, gridn(s->nfine), 256, G[0], (const double*)s->c64[d], L[0].c[d]);
    if (ctx->p.mg_coarse_op == MIST_COARSE_RESTRICTED_VF) {
        double h[3] = {ctx->m.dx, ctx->m.dy, ctx->m.dz};
        for (size_t l = 1; l < L.size(); ++l) {
            MGL(k_restrict_vf, gridn(G[l].size()), 256, ctx->m, s->g[l - 1], s->g[l],
                (const double*)ctx->f[MIST_FIELD_VOF], l == 1 ? (const double*)nullptr : (const double*)s->vfl[l - 1],
                s->vfl[l]);
            for (int d = 0; d < 3; ++d) {
                if (G[l].coarsened[d]) h[d] *= 2.0;
                MGL(k_coef_from_vf<T>, gridn(G[l].size()), 256, G[l], d, s->dt, ctx->p.rho_liq, ctx->p.rho_gas,
                    h[d] * h[d], (const double*)s->vfl[l], L[l].c[d]);
            }
        }
        return MIST_OK;
    }
    for (size_t l = 1; l < L.size(); ++l)
        for (int d = 0; d < 3; ++d)
            MGL(k_coarsen_coef<T>, gridn(G[l].size()), 256, G[l - 1], G[l], d, (const T*)L[l - 1].c[d], L[l].c[d]);
    return MIST_OK;
}

// FP64 residual of the outer loop handed to the V-cycle of precision T (b of level 0, e = 0)
// and max|r| into s->part[nblocks]
template <class T>
static int outer_residual(mist_ctx* ctx, std::vector<LevelBuf<T>>& L) {
    MGState* s = ctx->mg;
    const int64_t n = s->nfine;
    const int g = std::min<int>(s->nblocks, gridn(n, MG_RED));
    hipEvent_t ev = prof_begin(ctx, MIST_PROF_MG_RESID);
    MGL(k_outer_resid<T>, g, MG_RED, s->g[0], s->gq[sizeof(T) == sizeof(double) ? 0 : 1][0], (const double*)s->c64[0],
        (const double*)s->c64[1], (const double*)s->c64[2], (const double*)s->p64, (const double*)s->b64, L[0].b,
        L[0].e, s->part);
    prof_end(ctx, MIST_PROF_MG_RESID, ev, nullptr, n);
    MGL(k_sum_final, 1, MG_RED, (const double*)s->part, g, 1, s->part + s->nblocks);
    return MIST_OK;
}

// One defect-correction iteration: e = V-cycle(r) in precision T, p += e, the new residual
// handed to the next V-cycle.
template <class T>
static int outer_iteration(mist_ctx* ctx, std::vector<LevelBuf<T>>& L) {
    MGState* s = ctx->mg;
    const int64_t n = s->nfine;
    const LevelGeom& g0 = s->gq[sizeof(T) == sizeof(double) ? 0 : 1][0];
    int st = vcycle(ctx, L, 0);
    if (st) return st;
    MGL(k_add_correction<T>, gridn(n), 256, g0, (const T*)L[0].e, s->p64);
    return outer_residual(ctx, L);
}

template <class T>
static int solve_t(mist_ctx* ctx, std::vector<LevelBuf<T>>& L, double tol, int64_t* cycles, double* resid) {
    MGState* s = ctx->mg;
    int st = setup_hierarchy(ctx, L);
    if (st) return st;
    const int q = sizeof(T) == sizeof(double) ? 0 : 1;
    const int64_t maxc = ctx->p.mg_max_cycles > 0 ? ctx->p.mg_max_cycles : 10000;
    int64_t cyc = 0;
    double res = 0.0;
    int status = MIST_OK;
    if ((st = outer_residual(ctx, L))) return st;
    for (;;) {
        MIST_HIP_TRY(hipMemcpyAsync(s->h_part, s->part + s->nblocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
        res = s->h_part[0];
        if (!(res > tol)) break;  // also stops on NaN, reported through the residual
        if (cyc >= maxc) { status = MIST_ERR_NOCONV; break; }
        if (ctx->prof.on) {
            // per-kernel timing needs the kernels launched one by one
            if ((st = outer_iteration(ctx, L))) return st;
        } else {
            // the ~200 launches of an iteration replayed as one graph (captured once)
            if (!s->exec[q]) {
                MIST_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                st = outer_iteration(ctx, L);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
                if (st) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                MIST_HIP_TRY(ec);
                s->graph[q] = g;
                MIST_HIP_TRY(hipGraphInstantiate(&s->exec[q], g, nullptr, nullptr, 0));
            }
            MIST_HIP_TRY(hipGraphLaunch(s->exec[q], ctx->stream));
        }
        ++cyc;
    }
    if (res != res) status = MIST_ERR_NOCONV;
    *cycles = cyc;
    *resid = res;
    return status;
}

int pressure_solve(mist_ctx* ctx, double dt, double tol, int precision, int64_t* cycles, double* resid) {
    int st = build_state(ctx);
    if (st) return st;
    MGState* s = ctx->mg;
    const int64_t n = s->nfine;
    s->dt = dt;
    MGL(k_fine_setup, gridn(n), 256, ctx->m, s->g[0], dt, (const double*)ctx->f[MIST_FIELD_DENX],
        (const double*)ctx->f[MIST_FIELD_DENY], (const double*)ctx->f[MIST_FIELD_DENZ],
        (const double*)ctx->f[MIST_FIELD_RHS], (const double*)ctx->f[MIST_FIELD_P], s->c64[0], s->c64[1], s->c64[2],
        s->b64, s->p64);
    // compatibility: remove the mean of b (constants are the null space of A)
    if ((st = dev_reduce<double>(ctx, s->b64, n, 0))) return st;
    MGL(k_shift, gridn(n), 256, s->b64, n, (const double*)(s->part + s->nblocks));
    int status = precision == MIST_MG_MIXED ? solve_t<float>(ctx, s->s, tol, cycles, resid)
                                            : solve_t<double>(ctx, s->d, tol, cycles, resid);
    if (status != MIST_OK && status != MIST_ERR_NOCONV) return status;
    // zero-mean gauge, back into P, then ghosts
    if ((st = dev_reduce<double>(ctx, s->p64, n, 0))) return st;
    MGL(k_shift, gridn(n), 256, s->p64, n, (const double*)(s->part + s->nblocks));
    MGL(k_store_p, gridn(n), 256, ctx->m, s->g[0], (const double*)s->p64, (double*)ctx->f[MIST_FIELD_P]);
    if ((st = launch_neumann(ctx, MIST_FIELD_P))) return st;
    if ((st = launch_borders(ctx, MIST_FIELD_P))) return st;
    if (status == MIST_ERR_NOCONV) ctx->err = "pressure solve did not converge";
    return status;
}

}  // namespace mist
