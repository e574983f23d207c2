// mist_fas.hip -- the semi-Lagrangian pressure equation of the SL-SL scheme and its FAS
// multigrid on the device:
//
//   A!             src/Pressure.jl:83-114   corrected faces u* - dt/rho grad P, then divg_cell
//                                           (the SL pre-image volume in the band)
//   jacob_single   src/Pressure.jl:699-717  diagonal Jacobian by central differences, delta 5
//   res_iteration  src/Pressure.jl:777-884  diagonal-Newton steps with Anderson mixing (m 10)
//   par_anderson_accel, weighted_sum        src/Pressure.jl:720-775
//   restrict!, restrict_{x,y,z}_face!, prolong!   src/Multigrid.jl:5-54, 183-197, 242-279
//   mg_fas!, mg_cycler                      src/Multigrid.jl:333-500
//
// Each multigrid level is a child context (the level's mesh, VF, band and band list, face
// densities) on the parent's stream, so the operator reuses the vertex back-trace and the
// divergence kernels of the fine mesh: the child's UF, VF, WF fields hold the corrected face
// velocities when the divergence runs. Each level keeps its predicted faces, its P arrays, the
// Anderson history (10 slots of the interior) and its Gram matrix on the device; the Newton
// step, the history update, the Gram row, the 11 x 11 KKT solve and the mixing are kernels, so
// an iteration returns to the host only for its convergence test.
//
// Decisions (the reference's own form is kept by the oracle for comparison):
//   * res_iteration's first iteration takes the diagonal-Newton step Gn_loc; the reference
//     copies its zero-filled Gn argument into P there (src/Pressure.jl:843), which restarts
//     every smoother call from P = 0 and stalls the cycle;
//   * the Jacobian's columns are independent: each cell's two perturbed evaluations start from
//     the unperturbed P (the reference adds and subtracts delta in place); one thread per cell
//     evaluates both at that cell.
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mist_kernels.h"
#include "mist_nsgeom.h"

namespace mist {

#define RET_IF_ST(x)            \
    do {                        \
        const int st_ = (x);    \
        if (st_) return st_;    \
    } while (0)

constexpr int FAS_M = 10;          // Anderson history length
constexpr int FAS_T = 256;         // threads per block of the reductions
constexpr int FAS_BLOCKS = 1024;   // partial sums per reduction

struct FasLevel {
    mist_ctx* c = nullptr;  // child context of the level
    int64_t ncell = 0;      // cell-extent count
    int64_t nint = 0;       // interior count
    double *uf0 = nullptr, *vf0 = nullptr, *wf0 = nullptr;  // predicted face velocities
    double *P_h = nullptr, *AP_f = nullptr, *AP_c = nullptr, *P_bar_H = nullptr, *P_H = nullptr;
    double *tau = nullptr, *jac = nullptr;
    double* hist = nullptr;  // F slots then G slots, FAS_M each, interior size
    double* small = nullptr; // gram (FAS_M^2), alpha (FAS_M), scalars
};

// Convergence state of one res_iteration call on the device: once `done` is set (max|AP| <
// tol), the iterations already queued after it change nothing, so the host queues iterations
// in batches and reads the state once per batch.
struct FasCtrl {
    int done;         // converged
    int iters;        // iterations of this call (residual evaluations)
    double res;       // the last max|AP|
    long long total;  // iterations since the solve began
};

struct FasState {
    std::vector<FasLevel> lv;
    double* part = nullptr;   // FAS_BLOCKS x (FAS_M + 1) partial sums
    double* h_red = nullptr;  // pinned: max |AP|
    FasCtrl* ctrl = nullptr;  // device
    FasCtrl* h_ctrl = nullptr;  // pinned
    int nlev = 0;
};

void fas_free(FasState* s) {
    if (!s) return;
    for (FasLevel& L : s->lv) {
        double* bufs[] = {L.uf0, L.vf0, L.wf0, L.P_h, L.AP_f, L.AP_c, L.P_bar_H, L.P_H, L.tau, L.jac,
                          L.hist, L.small};
        for (double* b : bufs)
            if (b) (void)hipFree(b);
        if (L.c) {
            L.c->stream = nullptr;  // the parent's
            mist_ctx_destroy(L.c);
        }
    }
    if (s->part) (void)hipFree(s->part);
    if (s->h_red) (void)hipHostFree(s->h_red);
    if (s->ctrl) (void)hipFree(s->ctrl);
    if (s->h_ctrl) (void)hipHostFree(s->h_ctrl);
    delete s;
}

// ------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int64_t face_x(const DevMesh& m, int i, int j, int k) {
    return (int64_t)(i - 1 + m.ng) + (int64_t)(m.cnx + 1) * ((int64_t)(j - 1 + m.ng) + (int64_t)m.cny * (k - 1 + m.ng));
}
__device__ __forceinline__ int64_t face_y(const DevMesh& m, int i, int j, int k) {
    return (int64_t)(i - 1 + m.ng) + (int64_t)m.cnx * ((int64_t)(j - 1 + m.ng) + (int64_t)(m.cny + 1) * (k - 1 + m.ng));
}

// The corrected faces of A! (src/Pressure.jl:92-108) over the reference's ranges, 0 elsewhere
// (fill!). x: plain division; y and z: the safe division /^ of the reference.
__global__ void k_fas_grad(DevMesh m, double dt, const double* uf0, const double* vf0, const double* wf0,
                           const double* denx, const double* deny, const double* denz, const double* P, double* gx,
                           double* gy, double* gz) {
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        const bool kin = k >= 0 && k <= m.Nz + 1, jin = j >= 0 && j <= m.Ny + 1, iin = i >= 0 && i <= m.Nx + 1;
        double a = 0.0, b = 0.0, c = 0.0;
        if (kin && jin && i >= 0 && i <= m.Nx + 2)
            a = uf0[q] - ((dt / denx[face_x(m, i, j, k)]) * (P[q] - P[q - 1])) / m.dx;
        if (kin && iin && j >= 0 && j <= m.Ny + 2)
            b = vf0[q] - sdiv(sdiv(dt, deny[face_y(m, i, j, k)]) * (P[q] - P[q - m.cnx]), m.dy);
        if (jin && iin && k >= 0 && k <= m.Nz + 2)
            c = wf0[q] - sdiv(sdiv(dt, denz[q]) * (P[q] - P[q - (int64_t)m.cnx * m.cny]), m.dz);
        gx[q] = a; gy[q] = b; gz[q] = c;
    }
}

// The corrected faces of A! with P + s at one interior cell c and at the ghost copies of it
// that the one-layer Neumann fill makes (the cells whose coordinates clamp to c's: 0 -> 1,
// N + 1 -> N). A face neither of whose cells is perturbed is read from the faces k_fas_grad
// computed for the unperturbed P; the others are recomputed with k_fas_grad's arithmetic.
struct PertFaces {
    const DevMesh* m;
    const double *gx, *gy, *gz;
    const double *uf0, *vf0, *wf0, *denx, *deny, *denz, *P;
    double dt, s;
    int64_t img[8];
    int nimg;
    __device__ __forceinline__ bool pert(int64_t q) const {
        for (int n = 0; n < nimg; ++n)
            if (img[n] == q) return true;
        return false;
    }
    __device__ __forceinline__ double Pp(int64_t q) const { return pert(q) ? P[q] + s : P[q]; }
    __device__ double fu(int64_t q) const {
        if (!pert(q) && !pert(q - 1)) return gx[q];
        int i, j, k;
        cell_of(*m, q, i, j, k);
        if (!(k >= 0 && k <= m->Nz + 1 && j >= 0 && j <= m->Ny + 1 && i >= 0 && i <= m->Nx + 2)) return 0.0;
        return uf0[q] - ((dt / denx[face_x(*m, i, j, k)]) * (Pp(q) - Pp(q - 1))) / m->dx;
    }
    __device__ double fv(int64_t q) const {
        const int64_t sy = m->cnx;
        if (!pert(q) && !pert(q - sy)) return gy[q];
        int i, j, k;
        cell_of(*m, q, i, j, k);
        if (!(k >= 0 && k <= m->Nz + 1 && i >= 0 && i <= m->Nx + 1 && j >= 0 && j <= m->Ny + 2)) return 0.0;
        return vf0[q] - sdiv(sdiv(dt, deny[face_y(*m, i, j, k)]) * (Pp(q) - Pp(q - sy)), m->dy);
    }
    __device__ double fw(int64_t q) const {
        const int64_t sz = (int64_t)m->cnx * m->cny;
        if (!pert(q) && !pert(q - sz)) return gz[q];
        int i, j, k;
        cell_of(*m, q, i, j, k);
        if (!(j >= 0 && j <= m->Ny + 1 && i >= 0 && i <= m->Nx + 1 && k >= 0 && k <= m->Nz + 2)) return 0.0;
        return wf0[q] - sdiv(sdiv(dt, denz[q]) * (Pp(q) - Pp(q - sz)), m->dz);
    }
};

// The 8 corners of one cell back-traced with perturbed faces (the divergence of a cell reads
// no other vertex).
struct CornerVerts {
    double v[8][3];
    int i, j, k;
    __device__ __forceinline__ double at(int I, int J, int K, int c) const {
        return v[(I - i) + 2 * (J - j) + 4 * (K - k)][c];
    }
};

// jacob_single (src/Pressure.jl:699-717), one thread per interior cell c: A!(P + delta e_c)
// and A!(P - delta e_c) at c, each from the unperturbed state (the columns are independent),
// J = (A+ - A-) /^ 2 delta. The divergence at c reads the faces of c and, through the
// back-trace of c's corners, faces near them: PertFaces gives every face as A! would compute
// it over the whole mesh with P perturbed at c, and the corners are traced as
// k_vertex_list selects them (only vertices of band cells are read), so each column equals
// a whole-mesh evaluation of A! with only c perturbed.
__global__ void k_fas_jac_local(DevMesh m, DevParams p, int nband, const int16_t* band, double dt, const double* uf0,
                                const double* vf0, const double* wf0, const double* denx, const double* deny,
                                const double* denz, const double* P, const double* gx, const double* gy,
                                const double* gz, double delta, double* J) {
    const int64_t n = (int64_t)m.Nx * m.Ny * m.Nz;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(r % m.Nx), j = 1 + (int)((r / m.Nx) % m.Ny), k = 1 + (int)(r / ((int64_t)m.Nx * m.Ny));
        const int64_t q = m.cell(i, j, k);
        PertFaces fa{&m, gx, gy, gz, uf0, vf0, wf0, denx, deny, denz, P, dt, 0.0, {}, 0};
        {
            int xs[2], ys[2], zs[2], nx = 0, ny = 0, nz = 0;
            xs[nx++] = i;
            if (i == 1) xs[nx++] = 0;
            if (i == m.Nx) xs[nx++] = m.Nx + 1;
            ys[ny++] = j;
            if (j == 1) ys[ny++] = 0;
            if (j == m.Ny) ys[ny++] = m.Ny + 1;
            zs[nz++] = k;
            if (k == 1) zs[nz++] = 0;
            if (k == m.Nz) zs[nz++] = m.Nz + 1;
            for (int c = 0; c < nz; ++c)
                for (int b = 0; b < ny; ++b)
                    for (int a = 0; a < nx; ++a) fa.img[fa.nimg++] = m.cell(xs[a], ys[b], zs[c]);
        }
        // the divergence reads back-traced vertices only in a band cell or beside one
        const int64_t sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
        const bool trace = in_band(band, q, nband) || in_band(band, q - 1, nband) || in_band(band, q + 1, nband) ||
                           in_band(band, q - sy, nband) || in_band(band, q + sy, nband) ||
                           in_band(band, q - sz, nband) || in_band(band, q + sz, nband);
        double A[2];
        for (int sgn = 0; sgn < 2; ++sgn) {
            fa.s = sgn == 0 ? delta : -delta;
            CornerVerts cv;
            cv.i = i; cv.j = j; cv.k = k;
            for (int c = 0; c < 8; ++c) {
                const int I = i + (c & 1), Jv = j + ((c >> 1) & 1), K = k + (c >> 2);
                cv.v[c][0] = cv.v[c][1] = cv.v[c][2] = 0.0;
                if (!trace) continue;
                bool need = false;  // the vertex selection of k_vertex_list
                for (int e = 0; e < 8; ++e)
                    need |= abs((int)band[m.cell(I - 1 + (e & 1), Jv - 1 + ((e >> 1) & 1), K - 1 + (e >> 2))]) <= p.nband;
                if (!need) continue;
                double pt[3] = {m.x[I], m.y[Jv], m.z[K]};
                project_pt_acc(m, p, fa, pt, min(I, m.Nx), min(Jv, m.Ny), min(K, m.Nz), dt);
                cv.v[c][0] = pt[0]; cv.v[c][1] = pt[1]; cv.v[c][2] = pt[2];
            }
            A[sgn] = divg_cell_eval(m, nband, 1, band, cv, fa, dt, q, i, j, k);
        }
        J[q] = sdiv(A[0] - A[1], 2.0 * delta);
    }
}

// AP = divg (+ tau) on the interior, 0 on the ghosts; block maxima of |AP|.
__global__ __launch_bounds__(FAS_T) void k_fas_residual(DevMesh m, const double* divg, const double* tau, double* AP,
                                                        double* part, const FasCtrl* ctl) {
    if (ctl && ctl->done) return;
    __shared__ double s[FAS_T];
    const int64_t n = (int64_t)m.cnx * m.cny * m.cnz;
    double mx = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of(m, q, i, j, k);
        double v = 0.0;
        if (i >= 1 && i <= m.Nx && j >= 1 && j <= m.Ny && k >= 1 && k <= m.Nz) {
            v = divg[q];
            if (tau) v += tau[q];
        }
        AP[q] = v;
        mx = jmax(mx, fabs(v));
    }
    s[threadIdx.x] = mx;
    __syncthreads();
    for (int w = FAS_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = jmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// max of the block maxima into *out; with ctl, the iteration's convergence test
// (res_iteration's `res < tol`, src/Pressure.jl:819-823) unless the call has already converged.
__global__ __launch_bounds__(FAS_T) void k_fas_max_final(const double* part, int n, double* out, FasCtrl* ctl,
                                                         double tol) {
    if (ctl && ctl->done) return;
    __shared__ double s[FAS_T];
    double mx = 0.0;
    for (int q = threadIdx.x; q < n; q += FAS_T) mx = jmax(mx, part[q]);
    s[threadIdx.x] = mx;
    __syncthreads();
    for (int w = FAS_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = jmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (out) *out = s[0];
        if (ctl) {
            ctl->res = s[0];
            ctl->iters += 1;
            ctl->total += 1;
            if (s[0] < tol) ctl->done = 1;
        }
    }
}

__global__ void k_fas_ctrl_reset(FasCtrl* ctl, int total_too) {
    ctl->done = 0;
    ctl->iters = 0;
    ctl->res = INFINITY;
    if (total_too) ctl->total = 0;
}

// Newton step and history slot: Gn = P - omega AP / max(|J|, 1e-12), Fn = Gn - P into slot
// `slot`; block partial sums of Fn . F_s for the n filled slots (the new Gram row).
__global__ __launch_bounds__(FAS_T) void k_fas_step(DevMesh m, const double* P, const double* AP, const double* J,
                                                    double omega, double* hist, int64_t nint, int slot, int n,
                                                    double* part, const FasCtrl* ctl) {
    if (ctl->done) return;
    __shared__ double s[FAS_M][FAS_T];
    double acc[FAS_M];
#pragma unroll
    for (int t = 0; t < FAS_M; ++t) acc[t] = 0.0;
    double* F = hist;
    double* G = hist + (int64_t)FAS_M * nint;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nint; r += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, (int)r, i, j, k);
        const int64_t q = m.cell(i, j, k);
        const double Js = jmax(fabs(J[q]), 1e-12);
        const double gn = P[q] - (omega * AP[q]) / Js;
        const double fn = gn - P[q];
        F[(int64_t)slot * nint + r] = fn;
        G[(int64_t)slot * nint + r] = gn;
#pragma unroll
        for (int t = 0; t < FAS_M; ++t)
            if (t < n) acc[t] += fn * (t == slot ? fn : F[(int64_t)t * nint + r]);
    }
#pragma unroll
    for (int t = 0; t < FAS_M; ++t) s[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int w = FAS_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int t = 0; t < n; ++t) s[t][threadIdx.x] += s[t][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < n) part[(int64_t)blockIdx.x * FAS_M + threadIdx.x] = s[threadIdx.x][0];
}

// The Gram row of the new slot from the partials (fixed order), then, when two or more slots
// are filled, the KKT system [G + eps I, 1; 1', 0] a = [0; 1] with eps = 1e-10 max diag G
// (par_anderson_accel, src/Pressure.jl:747-775) by LU with partial pivoting (the first largest
// pivot): the system lives in LDS, each elimination step updates its rows in parallel (every
// entry is computed by one thread, as a one-thread elimination computes it), and one thread
// substitutes back.
__global__ __launch_bounds__(FAS_T) void k_fas_gram_solve(const double* part, int nblocks, int slot, int n,
                                                          double* small, const FasCtrl* ctl) {
    if (ctl->done) return;
    constexpr int NM = FAS_M + 1;
    __shared__ double row[FAS_M];
    __shared__ double K[NM][NM], bv[NM], fr[NM];
    __shared__ int piv;
    double* gram = small;
    double* alpha = small + FAS_M * FAS_M;
    const int tid = threadIdx.x;
    if (tid < n) {
        double sum = 0.0;
        for (int b = 0; b < nblocks; ++b) sum += part[(int64_t)b * FAS_M + tid];
        row[tid] = sum;
    }
    __syncthreads();
    if (tid < n) {
        gram[slot * FAS_M + tid] = row[tid];
        gram[tid * FAS_M + slot] = row[tid];
    }
    if (n < 2) return;
    __syncthreads();  // the new row and column are in `gram` for the whole block
    const int N = n + 1;
    if (tid == 0) {
        double mx = gram[0];
        for (int i = 1; i < n; ++i) mx = jmax(mx, gram[i * FAS_M + i]);
        fr[0] = 1e-10 * mx;  // eps
    }
    __syncthreads();
    const double eps = fr[0];
    __syncthreads();
    if (tid < N * N) {
        const int i = tid / N, j = tid - N * (tid / N);
        double v;
        if (i < n && j < n) v = gram[i * FAS_M + j] + (i == j ? eps : 0.0);
        else if (i == n && j == n) v = 0.0;
        else v = 1.0;
        K[i][j] = v;
        if (j == 0) bv[i] = i == n ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int col = 0; col < N; ++col) {
        if (tid == 0) {
            int p = col;
            for (int r = col + 1; r < N; ++r)
                if (fabs(K[r][col]) > fabs(K[p][col])) p = r;
            piv = p;
        }
        __syncthreads();
        const int p = piv;
        if (p != col) {
            if (tid < N) {
                const double t = K[col][tid]; K[col][tid] = K[p][tid]; K[p][tid] = t;
            }
            if (tid == 0) {
                const double t = bv[col]; bv[col] = bv[p]; bv[p] = t;
            }
            __syncthreads();
        }
        if (tid > col && tid < N) fr[tid] = K[tid][col] / K[col][col];
        __syncthreads();
        const int w = N - col - 1;  // rows and columns below and right of the pivot
        if (tid < w * w) {
            const int r = col + 1 + tid / w, q = col + 1 + tid % w;
            K[r][q] -= fr[r] * K[col][q];
        }
        if (tid < w) {
            const int r = col + 1 + tid;
            K[r][col] = fr[r];
            bv[r] -= fr[r] * bv[col];
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int r = N - 1; r >= 0; --r) {
            double s = bv[r];
            for (int q = r + 1; q < N; ++q) s -= K[r][q] * bv[q];
            bv[r] = s / K[r][r];
        }
        for (int i = 0; i < n; ++i) alpha[i] = bv[i];
    }
}

// P = beta (sum_s alpha_s G_s) + (1 - beta) P with two or more slots (weighted_sum starts from
// zeros and adds in slot order), the Newton step Gn of the new slot with one; block partial
// sums of the new interior P for the mean.
__global__ __launch_bounds__(FAS_T) void k_fas_mix(DevMesh m, double* P, const double* hist, int64_t nint, int slot,
                                                   int n, const double* small, double beta, double* part,
                                                   const FasCtrl* ctl) {
    if (ctl->done) return;
    __shared__ double s[FAS_T];
    const double* G = hist + (int64_t)FAS_M * nint;
    const double* alpha = small + FAS_M * FAS_M;
    double acc = 0.0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nint; r += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, (int)r, i, j, k);
        const int64_t q = m.cell(i, j, k);
        double v;
        if (n > 1) {
            double pn = 0.0;
            for (int t = 0; t < n; ++t) pn += alpha[t] * G[(int64_t)t * nint + r];
            v = beta * pn + (1.0 - beta) * P[q];
        } else {
            v = G[(int64_t)slot * nint + r];
        }
        P[q] = v;
        acc += v;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = FAS_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// P .-= mean(P[interior]) over the whole array, ghosts included (src/Pressure.jl:881).
__global__ __launch_bounds__(FAS_T) void k_fas_sub_mean(double* P, int64_t n, const double* part, int nblocks,
                                                        int64_t count, const FasCtrl* ctl) {
    if (ctl->done) return;
    __shared__ double mean;
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int b = 0; b < nblocks; ++b) sum += part[b];
        mean = sum / (double)count;
    }
    __syncthreads();
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        P[q] -= mean;
}

struct Ext {
    int lo[3], n[3];
    __device__ __forceinline__ int64_t at(int i, int j, int k) const {
        return (int64_t)(i - lo[0]) + (int64_t)n[0] * ((int64_t)(j - lo[1]) + (int64_t)n[1] * (k - lo[2]));
    }
};

// restrict! (src/Multigrid.jl:183-197): 8-cell mean on the coarse interior.
__global__ void k_fas_restrict(int Nx, int Ny, int Nz, double* C, Ext ce, const double* F, Ext fe) {
    const int64_t n = (int64_t)Nx * Ny * Nz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % Nx) + 1, j = (int)((q / Nx) % Ny) + 1, k = (int)(q / ((int64_t)Nx * Ny)) + 1;
        const int ii = 2 * i - 1, jj = 2 * j - 1, kk = 2 * k - 1;
        C[ce.at(i, j, k)] = (((((((F[fe.at(ii, jj, kk)] + F[fe.at(ii + 1, jj, kk)]) + F[fe.at(ii, jj + 1, kk)]) +
                                 F[fe.at(ii + 1, jj + 1, kk)]) + F[fe.at(ii, jj, kk + 1)]) + F[fe.at(ii + 1, jj, kk + 1)]) +
                               F[fe.at(ii, jj + 1, kk + 1)]) + F[fe.at(ii + 1, jj + 1, kk + 1)]) / 8.0;
    }
}

// restrict_{x,y,z}_face! (src/Multigrid.jl:242-279): the 4 fine faces of each coarse face.
__global__ void k_fas_restrict_faces(int Nx, int Ny, int Nz, Ext e, Ext fe, double* cu, double* cv, double* cw,
                                     const double* fu, const double* fv, const double* fw) {
    const int64_t n = (int64_t)(Nx + 1) * (Ny + 1) * (Nz + 1);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % (Nx + 1)) + 1, j = (int)((q / (Nx + 1)) % (Ny + 1)) + 1, k = (int)(q / ((int64_t)(Nx + 1) * (Ny + 1))) + 1;
        const int ii = 2 * i - 1, jj = 2 * j - 1, kk = 2 * k - 1;
        if (j <= Ny && k <= Nz)
            cu[e.at(i, j, k)] = (((fu[fe.at(ii, jj, kk)] + fu[fe.at(ii, jj, kk + 1)]) + fu[fe.at(ii, jj + 1, kk)]) +
                                 fu[fe.at(ii, jj + 1, kk + 1)]) / 4.0;
        if (i <= Nx && k <= Nz)
            cv[e.at(i, j, k)] = (((fv[fe.at(ii, jj, kk)] + fv[fe.at(ii + 1, jj, kk)]) + fv[fe.at(ii, jj, kk + 1)]) +
                                 fv[fe.at(ii + 1, jj, kk + 1)]) / 4.0;
        if (i <= Nx && j <= Ny)
            cw[e.at(i, j, k)] = (((fw[fe.at(ii, jj, kk)] + fw[fe.at(ii + 1, jj, kk)]) + fw[fe.at(ii, jj + 1, kk)]) +
                                 fw[fe.at(ii + 1, jj + 1, kk)]) / 4.0;
    }
}

// prolong! (src/Multigrid.jl:5-54): the fine interior from the coarse centres (ghosts
// included), the coarse cell found by the reference's walk from the clamped fine index.
__global__ void k_fas_prolong(DevMesh fm, DevMesh cm, double* Fo, const double* C) {
    const int64_t n = (int64_t)fm.Nx * fm.Ny * fm.Nz;
    const int lo = 1 - cm.ng;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i_f, jf, kf;
        interior_of(fm, (int)q, i_f, jf, kf);
        const double p0 = fm.xm[i_f], p1 = fm.ym[jf], p2 = fm.zm[kf];
        int i = min(max(i_f, lo), cm.Nx + cm.ng), j = min(max(jf, lo), cm.Ny + cm.ng), k = min(max(kf, lo), cm.Nz + cm.ng);
        while (p0 - cm.xm[i] < 0.0 && i > lo) --i;
        while (p0 - cm.xm[i + 1] >= 0.0 && i + 1 < cm.Nx + cm.ng) ++i;
        while (p1 - cm.ym[j] < 0.0 && j > lo) --j;
        while (p1 - cm.ym[j + 1] >= 0.0 && j + 1 < cm.Ny + cm.ng) ++j;
        while (p2 - cm.zm[k] < 0.0 && k > lo) --k;
        while (p2 - cm.zm[k + 1] >= 0.0 && k + 1 < cm.Nz + cm.ng) ++k;
        const double wx1 = (p0 - cm.xm[i]) / (cm.xm[i + 1] - cm.xm[i]), wx2 = 1.0 - wx1;
        const double wy1 = (p1 - cm.ym[j]) / (cm.ym[j + 1] - cm.ym[j]), wy2 = 1.0 - wy1;
        const double wz1 = (p2 - cm.zm[k]) / (cm.zm[k + 1] - cm.zm[k]), wz2 = 1.0 - wz1;
        auto c = [&](int a, int b, int d) { return C[cm.cell(a, b, d)]; };
        Fo[fm.cell(i_f, jf, kf)] = wz1 * (wy1 * (wx1 * c(i + 1, j + 1, k + 1) + wx2 * c(i, j + 1, k + 1)) +
                                          wy2 * (wx1 * c(i + 1, j, k + 1) + wx2 * c(i, j, k + 1))) +
                                   wz2 * (wy1 * (wx1 * c(i + 1, j + 1, k) + wx2 * c(i, j + 1, k)) +
                                          wy2 * (wx1 * c(i + 1, j, k) + wx2 * c(i, j, k)));
    }
}

__global__ void k_fas_axpy(double* y, const double* x, double a, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        y[q] += a * x[q];
}
__global__ void k_fas_diff(double* out, const double* a, const double* b, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = a[q] - b[q];
}
__global__ void k_fas_interior_add(DevMesh m, double* y, const double* x) {
    const int64_t n = (int64_t)m.Nx * m.Ny * m.Nz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, (int)q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        y[c] += x[c];
    }
}
__global__ void k_fas_interior_copy(DevMesh m, double* y, const double* x) {
    const int64_t n = (int64_t)m.Nx * m.Ny * m.Nz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_of(m, (int)q, i, j, k);
        const int64_t c = m.cell(i, j, k);
        y[c] = x[c];
    }
}
__global__ void k_fas_zero(double* y, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        y[q] = 0.0;
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
static inline int fgrid(int64_t n, int block, int cap = FAS_BLOCKS) {
    const int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

#define FL(kernel, grid, block, ...)                                                          \
    do {                                                                                      \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ctx->stream, __VA_ARGS__);     \
        MIST_HIP_TRY(hipGetLastError());                                                      \
    } while (0)

static Ext cell_ext(const mist_ctx* c) {
    Ext e;
    for (int d = 0; d < 3; ++d) { e.lo[d] = (int)c->flo[MIST_FIELD_P][d]; e.n[d] = (int)c->fn[MIST_FIELD_P][d]; }
    return e;
}
static Ext vof_ext(const mist_ctx* c) {
    Ext e;
    for (int d = 0; d < 3; ++d) { e.lo[d] = (int)c->flo[MIST_FIELD_VOF][d]; e.n[d] = (int)c->fn[MIST_FIELD_VOF][d]; }
    return e;
}

static int fas_setup(mist_ctx* ctx) {
    if (ctx->fas) return MIST_OK;
    FasState* S = new FasState();
    ctx->fas = S;
    S->nlev = (int)std::max<int64_t>(1, ctx->p.mg_lvl);
    S->lv.resize((size_t)S->nlev);
    for (int l = 1; l <= S->nlev; ++l) {
        FasLevel& L = S->lv[(size_t)l - 1];
        mist_params lp = ctx->p;
        const int64_t f = (int64_t)1 << (l - 1);
        lp.Nx = std::max<int64_t>(ctx->p.Nx / f, 1);
        lp.Ny = std::max<int64_t>(ctx->p.Ny / f, 1);
        lp.Nz = std::max<int64_t>(ctx->p.Nz / f, 1);
        lp.solveNS = 0;
        mist_ctx* c = nullptr;
        int st = mist_ctx_create(&lp, ctx->device, &c);
        if (st) { ctx->err = std::string("multigrid level context: ") + mist_last_error(nullptr); return st; }
        // the level runs on the parent's stream (its set-up work is finished first)
        MIST_HIP_TRY(hipStreamSynchronize(c->stream));
        (void)hipStreamDestroy(c->stream);
        c->stream = ctx->stream;
        L.c = c;
        L.ncell = (int64_t)c->m.cnx * c->m.cny * c->m.cnz;
        L.nint = (int64_t)c->m.Nx * c->m.Ny * c->m.Nz;
        double** cells[] = {&L.uf0, &L.vf0, &L.wf0, &L.P_h, &L.AP_f, &L.AP_c, &L.P_bar_H, &L.P_H, &L.tau, &L.jac};
        for (double** b : cells) {
            if (hipMalloc(b, (size_t)L.ncell * sizeof(double)) != hipSuccess) { ctx->err = "FAS level allocation"; return MIST_ERR_OOM; }
            MIST_HIP_TRY(hipMemsetAsync(*b, 0, (size_t)L.ncell * sizeof(double), ctx->stream));
        }
        if (hipMalloc(&L.hist, (size_t)2 * FAS_M * L.nint * sizeof(double)) != hipSuccess ||
            hipMalloc(&L.small, (size_t)(FAS_M * FAS_M + FAS_M + 8) * sizeof(double)) != hipSuccess) {
            ctx->err = "FAS history allocation";
            return MIST_ERR_OOM;
        }
    }
    if (hipMalloc(&S->part, (size_t)FAS_BLOCKS * (FAS_M + 1) * sizeof(double)) != hipSuccess) {
        ctx->err = "FAS reduction allocation";
        return MIST_ERR_OOM;
    }
    if (hipHostMalloc(&S->h_red, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&S->h_ctrl, sizeof(FasCtrl), hipHostMallocDefault) != hipSuccess) {
        ctx->err = "FAS pinned allocation";
        return MIST_ERR_OOM;
    }
    if (hipMalloc(&S->ctrl, sizeof(FasCtrl)) != hipSuccess) {
        ctx->err = "FAS control allocation";
        return MIST_ERR_OOM;
    }
    return MIST_OK;
}

// A! (src/Pressure.jl:83-114) on a level: P's ghosts, the corrected faces into the child's
// UF, VF, WF, the back-traced band vertices, divg_cell into the child's DIVG.
#define CHILD(x)                                  \
    do {                                          \
        const int st_ = (x);                      \
        if (st_) { ctx->err = c->err; return st_; } \
    } while (0)

// The Neumann fill of P and the corrected faces of A! into the level's UF, VF, WF.
static int fas_faces(mist_ctx* ctx, FasLevel& L, double* P, double dt) {
    mist_ctx* c = L.c;
    CHILD(launch_neumann_cells(c, P));
    FL(k_fas_grad, fgrid(L.ncell, 256, 65536), 256, c->m, dt, (const double*)L.uf0, (const double*)L.vf0,
       (const double*)L.wf0, (const double*)c->f[MIST_FIELD_DENX], (const double*)c->f[MIST_FIELD_DENY],
       (const double*)c->f[MIST_FIELD_DENZ], (const double*)P, (double*)c->f[MIST_FIELD_UF],
       (double*)c->f[MIST_FIELD_VF], (double*)c->f[MIST_FIELD_WF]);
    return MIST_OK;
}

static int fas_A(mist_ctx* ctx, FasLevel& L, double* P, double dt) {
    mist_ctx* c = L.c;
    RET_IF_ST(fas_faces(ctx, L, P, dt));
    CHILD(launch_divergence(c, dt, true));  // the level's band is fixed during a solve
    return MIST_OK;
}

// AP = A!(P) (+ tau). With ctl: one iteration's residual and convergence test on the device
// (nothing is written once ctl is done). Without: max |AP| into *res on the host when res is
// given.
static int fas_residual(mist_ctx* ctx, FasLevel& L, double* P, double* AP, bool use_tau, double dt, double* res,
                        FasCtrl* ctl = nullptr, double tol = 0.0) {
    FasState* S = ctx->fas;
    RET_IF_ST(fas_A(ctx, L, P, dt));
    const int nb = fgrid(L.ncell, FAS_T);
    FL(k_fas_residual, nb, FAS_T, L.c->m, (const double*)L.c->f[MIST_FIELD_DIVG], use_tau ? (const double*)L.tau : nullptr,
       AP, S->part, (const FasCtrl*)ctl);
    FL(k_fas_max_final, 1, FAS_T, (const double*)S->part, nb, ctl ? nullptr : S->h_red, ctl, tol);
    if (!ctl && res) {
        MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
        *res = S->h_red[0];
    }
    return MIST_OK;
}

// jacob_single (src/Pressure.jl:699-717) with independent columns: the faces of the
// unperturbed P, then every column evaluated locally by its own thread.
static int fas_jacobian(mist_ctx* ctx, FasLevel& L, double* P, double dt) {
    mist_ctx* c = L.c;
    RET_IF_ST(fas_faces(ctx, L, P, dt));
    FL(k_fas_jac_local, fgrid(L.nint, 64, 65536), 64, c->m, c->dp, (int)c->p.nband,
       (const int16_t*)c->f[MIST_FIELD_BAND], dt, (const double*)L.uf0, (const double*)L.vf0, (const double*)L.wf0,
       (const double*)c->f[MIST_FIELD_DENX], (const double*)c->f[MIST_FIELD_DENY],
       (const double*)c->f[MIST_FIELD_DENZ], (const double*)P, (const double*)c->f[MIST_FIELD_UF],
       (const double*)c->f[MIST_FIELD_VF], (const double*)c->f[MIST_FIELD_WF], 5.0, L.jac);
    return MIST_OK;
}

// res_iteration (src/Pressure.jl:777-884). The iterations are queued in batches without a host
// round trip: the convergence test runs on the device (FasCtrl), and the kernels of the
// iterations queued after convergence change nothing, so the iterates are those of the
// reference's loop. With `result` the call waits for the device at the end and returns
// *converged and *res (the last max|AP|); otherwise it returns without waiting.
static int fas_smooth(mist_ctx* ctx, FasLevel& L, double* P, double* AP, double dt, double tol, int64_t max_iter,
                      bool use_tau, bool result, bool* converged, double* res) {
    FasState* S = ctx->fas;
    mist_ctx* c = L.c;
    const double omega = 0.8, beta = 0.5;
    // the reference evaluates A(P) before its Jacobian; of it only the ghost fill of P persists
    {
        mist_ctx* c = L.c;
        CHILD(launch_neumann_cells(c, P));
    }
    RET_IF_ST(fas_jacobian(ctx, L, P, dt));
    FL(k_fas_ctrl_reset, 1, 1, S->ctrl, 0);
    int hist_idx = 0, n_hist = 0;
    int64_t queued = 0;
    int batch = 8;
    const int nbi = fgrid(L.nint, FAS_T);
    while (queued < max_iter) {
        const int64_t nb = std::min<int64_t>(max_iter <= 16 ? max_iter : batch, max_iter - queued);
        for (int64_t q = 0; q < nb; ++q) {
            ++queued;
            RET_IF_ST(fas_residual(ctx, L, P, AP, use_tau, dt, nullptr, S->ctrl, tol));
            // the slot and history length of this iteration if it does not converge (the
            // kernels of a converged call skip, so later values are never used)
            hist_idx = hist_idx % FAS_M + 1;
            if (n_hist < FAS_M) ++n_hist;
            const int slot = hist_idx - 1;
            FL(k_fas_step, nbi, FAS_T, c->m, (const double*)P, (const double*)AP, (const double*)L.jac, omega, L.hist,
               L.nint, slot, n_hist, S->part, (const FasCtrl*)S->ctrl);
            FL(k_fas_gram_solve, 1, FAS_T, (const double*)S->part, nbi, slot, n_hist, L.small, (const FasCtrl*)S->ctrl);
            FL(k_fas_mix, nbi, FAS_T, c->m, P, (const double*)L.hist, L.nint, slot, n_hist, (const double*)L.small, beta,
               S->part, (const FasCtrl*)S->ctrl);
            FL(k_fas_sub_mean, fgrid(L.ncell, FAS_T, 65536), FAS_T, P, L.ncell, (const double*)S->part, nbi, L.nint,
               (const FasCtrl*)S->ctrl);
        }
        if (queued >= max_iter) break;
        // a long call: wait for the batch and stop once converged
        MIST_HIP_TRY(hipMemcpyAsync(S->h_ctrl, S->ctrl, sizeof(FasCtrl), hipMemcpyDeviceToHost, ctx->stream));
        MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (S->h_ctrl->done) break;
        batch = std::min(batch * 2, 32);
    }
    if (result) {
        MIST_HIP_TRY(hipMemcpyAsync(S->h_ctrl, S->ctrl, sizeof(FasCtrl), hipMemcpyDeviceToHost, ctx->stream));
        MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (converged) *converged = S->h_ctrl->done != 0;
        if (res) *res = S->h_ctrl->res;
    }
    return MIST_OK;
}

// mg_fas! (src/Multigrid.jl:375-500), lvl 1-based.
static int fas_cycle(mist_ctx* ctx, int lvl, double dt, double tol, bool* conv, double* res_out) {
    FasState* S = ctx->fas;
    FasLevel& L = S->lv[(size_t)lvl - 1];
    const int v1 = 10;
    if (lvl == S->nlev) return fas_smooth(ctx, L, L.P_h, L.AP_f, dt, 1e-10, 50000, true, false, nullptr, nullptr);
    RET_IF_ST(fas_smooth(ctx, L, L.P_h, L.AP_f, dt, tol, v1, lvl != 1, false, nullptr, nullptr));
    FasLevel& C = S->lv[(size_t)lvl];
    mist_ctx* cc = C.c;
    const DevMesh& cm = cc->m;
    const int64_t nci = C.nint;
    // the coarse VF, its band and densities
    FL(k_fas_restrict, fgrid(nci, 256, 65536), 256, cm.Nx, cm.Ny, cm.Nz, (double*)cc->f[MIST_FIELD_VOF], vof_ext(cc),
       (const double*)L.c->f[MIST_FIELD_VOF], vof_ext(L.c));
    {
        mist_ctx* c = cc;
        CHILD(launch_neumann(cc, MIST_FIELD_VOF));
        CHILD(launch_band(cc));
    }
    // the restricted approximation and A(P) + tau of this level
    FL(k_fas_restrict, fgrid(nci, 256, 65536), 256, cm.Nx, cm.Ny, cm.Nz, C.P_h, cell_ext(cc), (const double*)L.P_h,
       cell_ext(L.c));
    {
        mist_ctx* c = cc;
        CHILD(launch_neumann_cells(cc, C.P_h));
    }
    FL(k_fas_restrict, fgrid(nci, 256, 65536), 256, cm.Nx, cm.Ny, cm.Nz, C.AP_f, cell_ext(cc), (const double*)L.AP_f,
       cell_ext(L.c));
    {
        mist_ctx* c = cc;
        CHILD(launch_props(cc));
    }
    FL(k_fas_restrict_faces, fgrid((int64_t)(cm.Nx + 1) * (cm.Ny + 1) * (cm.Nz + 1), 256, 65536), 256, cm.Nx, cm.Ny,
       cm.Nz, cell_ext(cc), cell_ext(L.c), C.uf0, C.vf0, C.wf0, (const double*)L.uf0, (const double*)L.vf0,
       (const double*)L.wf0);
    // tau = R(A_h(P_h)) - A_H(R P_h)
    RET_IF_ST(fas_residual(ctx, C, C.P_h, C.AP_c, false, dt, nullptr));
    FL(k_fas_diff, fgrid(C.ncell, 256, 65536), 256, C.tau, (const double*)C.AP_f, (const double*)C.AP_c, C.ncell);
    MIST_HIP_TRY(hipMemcpyAsync(C.P_bar_H, C.P_h, (size_t)C.ncell * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(L.P_H, L.P_h, (size_t)L.ncell * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    RET_IF_ST(fas_cycle(ctx, lvl + 1, dt, tol, conv, res_out));
    FL(k_fas_axpy, fgrid(C.ncell, 256, 65536), 256, C.P_h, (const double*)C.P_bar_H, -1.0, C.ncell);
    FL(k_fas_zero, fgrid(L.ncell, 256, 65536), 256, L.AP_f, L.ncell);
    FL(k_fas_prolong, fgrid(L.nint, 256, 65536), 256, L.c->m, cm, L.AP_f, (const double*)C.P_h);
    FL(k_fas_interior_add, fgrid(L.nint, 256, 65536), 256, L.c->m, L.P_H, (const double*)L.AP_f);
    FL(k_fas_zero, fgrid(L.ncell, 256, 65536), 256, L.AP_f, L.ncell);
    bool c1 = false;
    double res = 0.0;
    RET_IF_ST(fas_smooth(ctx, L, L.P_H, L.AP_f, dt, tol, v1, lvl != 1, lvl == 1, &c1, &res));
    if (lvl == 1) {
        *conv = c1;
        *res_out = res;
    }
    MIST_HIP_TRY(hipMemcpyAsync(L.P_h, L.P_H, (size_t)L.ncell * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return MIST_OK;
}

// copy_to_mg! (src/Tools.jl:1414-1428) on level 1, and the fine VF.
static int fas_load(mist_ctx* ctx) {
    RET_IF_ST(fas_setup(ctx));
    FasLevel& F = ctx->fas->lv[0];
    mist_ctx* c = F.c;
    const size_t cb = (size_t)F.ncell * sizeof(double);
    MIST_HIP_TRY(hipMemcpyAsync(F.P_h, ctx->f[MIST_FIELD_P], cb, hipMemcpyDeviceToDevice, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(F.uf0, ctx->f[MIST_FIELD_UF], cb, hipMemcpyDeviceToDevice, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(F.vf0, ctx->f[MIST_FIELD_VF], cb, hipMemcpyDeviceToDevice, ctx->stream));
    MIST_HIP_TRY(hipMemcpyAsync(F.wf0, ctx->f[MIST_FIELD_WF], cb, hipMemcpyDeviceToDevice, ctx->stream));
    const int copy[] = {MIST_FIELD_VOF, MIST_FIELD_BAND, MIST_FIELD_DENX, MIST_FIELD_DENY, MIST_FIELD_DENZ};
    for (int f : copy) {
        const size_t bytes = (size_t)(ctx->fn[f][0] * ctx->fn[f][1] * ctx->fn[f][2] * ctx->felem[f]);
        MIST_HIP_TRY(hipMemcpyAsync(c->f[f], ctx->f[f], bytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    c->vtx_list_ok = false;  // a new band: the next A! lists its vertices again
    return MIST_OK;
}

// A!(P) of the current fields (src/Pressure.jl:83-114) into MIST_FIELD_DIVG.
int pressure_operator_sl(mist_ctx* ctx, double dt) {
    RET_IF_ST(fas_load(ctx));
    FasLevel& F = ctx->fas->lv[0];
    RET_IF_ST(fas_residual(ctx, F, F.P_h, F.AP_f, false, dt, nullptr));
    MIST_HIP_TRY(hipMemcpyAsync(ctx->f[MIST_FIELD_DIVG], F.AP_f, (size_t)F.ncell * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MIST_OK;
}

// jacob_single (src/Pressure.jl:699-717) of the current fields, columns independent, into the
// host array out (cell extents; the interior is set, the rest 0).
int pressure_jacobian_sl(mist_ctx* ctx, double dt, double* out) {
    RET_IF_ST(fas_load(ctx));
    FasLevel& F = ctx->fas->lv[0];
    {
        mist_ctx* c = F.c;
        CHILD(launch_neumann_cells(c, F.P_h));
    }
    FL(k_fas_zero, fgrid(F.ncell, 256, 65536), 256, F.jac, F.ncell);
    RET_IF_ST(fas_jacobian(ctx, F, F.P_h, dt));
    MIST_HIP_TRY(hipMemcpyAsync(out, F.jac, (size_t)F.ncell * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MIST_OK;
}

// The SL-SL pressure solve (pressure_solver! with the semi-Lagrangian scheme,
// src/Pressure.jl:5-30): FAS cycles to max|A(P)| < tol for mg_lvl > 1, res_iteration on the
// fine level for mg_lvl = 1. P (MIST_FIELD_P) is the initial guess and the result.
int pressure_solve_sl(mist_ctx* ctx, double dt, double tol, int64_t* cycles, double* resid, int64_t* smooth_iters) {
    if (ctx->p.xper || ctx->p.yper || ctx->p.zper) {
        ctx->err = "semi-Lagrangian pressure: periodic axes are not supported";
        return MIST_ERR_PARAM;
    }
    RET_IF_ST(fas_load(ctx));
    FasState* S = ctx->fas;
    FasLevel& F = S->lv[0];
    bool conv = false;
    int64_t it = 0;
    double res = INFINITY;
    FL(k_fas_ctrl_reset, 1, 1, S->ctrl, 1);
    if (S->nlev == 1) {
        RET_IF_ST(fas_smooth(ctx, F, F.P_h, F.AP_f, dt, tol, 50000, false, true, &conv, &res));
        it = S->h_ctrl->iters;
    } else {
        while (it < ctx->p.mg_max_cycles) {
            ++it;
            RET_IF_ST(fas_cycle(ctx, 1, dt, tol, &conv, &res));
            if (conv) break;
        }
    }
    // P[interior] = P_h[1][interior]; Neumann!; update_borders!
    FL(k_fas_interior_copy, fgrid(F.nint, 256, 65536), 256, ctx->m, (double*)ctx->f[MIST_FIELD_P], (const double*)F.P_h);
    RET_IF_ST(launch_neumann(ctx, MIST_FIELD_P));
    MIST_HIP_TRY(hipMemcpyAsync(S->h_ctrl, S->ctrl, sizeof(FasCtrl), hipMemcpyDeviceToHost, ctx->stream));
    MIST_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (cycles) *cycles = it;
    if (resid) *resid = res;
    if (smooth_iters) *smooth_iters = S->h_ctrl->total;
    if (!conv) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "P-solve did not converge: max|A(P)| = %.3e after %lld cycles (tol %.1e)", res,
                      (long long)it, tol);
        ctx->err = buf;
        return MIST_ERR_NOCONV;
    }
    return MIST_OK;
}

}  // namespace mist
