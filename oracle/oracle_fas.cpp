// oracle_fas.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): the semi-Lagrangian pressure
// equation of the SL-SL scheme and its FAS multigrid, restated from the reference:
//
//   A!              src/Pressure.jl:83-114    corrected faces u* - dt/rho grad P, then divg_cell
//   A! (one cell)   src/Pressure.jl:117-136   the same around one cell, for the Jacobian
//   jacob_single    src/Pressure.jl:699-717   diagonal Jacobian by central differences, delta 5
//   add/remove_perturb!  src/Tools.jl:1348-1412
//   weighted_sum, par_anderson_accel          src/Pressure.jl:720-775
//   res_iteration   src/Pressure.jl:777-884   diagonal-Newton steps with Anderson mixing
//   restrict_{x,y,z}_face!                    src/Multigrid.jl:242-279
//   mg_fas!         src/Multigrid.jl:375-500
//   mg_cycler       src/Multigrid.jl:333-371  (semi-Lagrangian branch)
//
// The multigrid arrays of the reference (mg_initArrays, src/Tools.jl:194-311) persist for a
// run; every read of them here happens after a write in the same solve, or reads a ghost
// layer the reference never writes (zero), so a solve allocates them afresh and zeroes them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "oracle_core.h"

namespace ora {

namespace {

// *^ (src/Tools.jl:379-385): 0 when either factor is within eps of zero.
inline double smul(double a, double b) { return (std::fabs(a) <= EPS || std::fabs(b) <= EPS) ? 0.0 : a * b; }

struct FasLevel {
    std::unique_ptr<Ctx> c;          // the level's mesh, VF (tmplrg), band and densities; its
                                     // uf, vf, wf are gradx, grady, gradz of the reference
    Arr3<double> uf, vf, wf;         // mg_arrays.uf/vf/wf: the predicted face velocities
    Arr3<double> P_h, AP_f, AP_c, P_bar_H, P_H, tau, Gn, jac, L1, L2;
};

std::unique_ptr<Ctx> level_ctx(const Ctx& fine, int level) {
    auto c = std::make_unique<Ctx>();
    c->p = fine.p;
    c->sl_first_step_newton = fine.sl_first_step_newton;
    c->m.build(c->p, level);
    const Ext ce = c->m.cell_ext();
    c->VF.init(c->m.vof_ext());
    for (Arr3<double>* a : {&c->uf, &c->vf, &c->wf, &c->divg, &c->P}) a->init(ce);
    c->denx.init(c->m.face_ext(0));
    c->deny.init(c->m.face_ext(1));
    c->denz.init(c->m.face_ext(2));
    c->band.init(ce);
    return c;
}

void level_arrays(FasLevel& L) {
    const Ext ce = L.c->m.cell_ext();
    for (Arr3<double>* a : {&L.uf, &L.vf, &L.wf, &L.P_h, &L.AP_f, &L.AP_c, &L.P_bar_H, &L.P_H, &L.tau, &L.Gn,
                            &L.jac, &L.L1, &L.L2})
        a->init(ce, 0.0);
}

template <class F>
void interior_for(const Mesh& m, F&& f) {
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) f(i, j, k);
}

// A! (src/Pressure.jl:83-114) on a level: LHS = divg_cell of the corrected faces.
void fas_A(FasLevel& L, Arr3<double>& P, Arr3<double>& LHS, double dt) {
    Ctx& c = *L.c;
    const Mesh& m = c.m;
    std::fill(c.uf.a.begin(), c.uf.a.end(), 0.0);
    std::fill(c.vf.a.begin(), c.vf.a.end(), 0.0);
    std::fill(c.wf.a.begin(), c.wf.a.end(), 0.0);
    neumann(c, P);
    update_borders(c, P);
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 2; ++i)
                c.uf(i, j, k) = L.uf(i, j, k) - ((dt / c.denx(i, j, k)) * (P(i, j, k) - P(i - 1, j, k))) / m.dx;
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 2; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 1; ++i)
                c.vf(i, j, k) = L.vf(i, j, k) - sdiv(sdiv(dt, c.deny(i, j, k)) * (P(i, j, k) - P(i, j - 1, k)), m.dy);
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 2; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 1; ++i)
                c.wf(i, j, k) = L.wf(i, j, k) - sdiv(sdiv(dt, c.denz(i, j, k)) * (P(i, j, k) - P(i, j, k - 1)), m.dz);
    std::fill(LHS.a.begin(), LHS.a.end(), 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) LHS(i, j, k) = divg_cell(c, i, j, k, dt);
}

// A! around one cell (src/Pressure.jl:117-136): the corrected faces of the box around
// (i, j, k), then divg_cell there. No ghost fill of P (add_perturb! keeps the wall copy).
void fas_A_local(FasLevel& L, Arr3<double>& P, int64_t i, int64_t j, int64_t k, Arr3<double>& LHS, double dt) {
    Ctx& c = *L.c;
    const Mesh& m = c.m;
    const int64_t Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
    for (int64_t kk = std::max<int64_t>(k - 1, 0); kk <= std::min(k + 1, Nz + 1); ++kk)
        for (int64_t jj = std::max<int64_t>(j - 1, 0); jj <= std::min(j + 1, Ny + 1); ++jj)
            for (int64_t ii = std::max<int64_t>(i - 1, 0); ii <= std::min(i + 2, Nx + 2); ++ii)
                c.uf(ii, jj, kk) = L.uf(ii, jj, kk) - sdiv(smul(sdiv(dt, c.denx(ii, jj, kk)), P(ii, jj, kk) - P(ii - 1, jj, kk)), m.dx);
    for (int64_t kk = std::max<int64_t>(k - 1, 0); kk <= std::min(k + 1, Nz + 1); ++kk)
        for (int64_t jj = std::max<int64_t>(j - 1, 0); jj <= std::min(j + 2, Ny + 2); ++jj)
            for (int64_t ii = std::max<int64_t>(i - 1, 0); ii <= std::min(i + 1, Nx + 1); ++ii)
                c.vf(ii, jj, kk) = L.vf(ii, jj, kk) - sdiv(smul(sdiv(dt, c.deny(ii, jj, kk)), P(ii, jj, kk) - P(ii, jj - 1, kk)), m.dy);
    for (int64_t kk = std::max<int64_t>(k - 1, 0); kk <= std::min(k + 2, Nz + 2); ++kk)
        for (int64_t jj = std::max<int64_t>(j - 1, 0); jj <= std::min(j + 1, Ny + 1); ++jj)
            for (int64_t ii = std::max<int64_t>(i - 1, 0); ii <= std::min(i + 1, Nx + 1); ++ii)
                c.wf(ii, jj, kk) = L.wf(ii, jj, kk) - sdiv(smul(sdiv(dt, c.denz(ii, jj, kk)), P(ii, jj, kk) - P(ii, jj, kk - 1)), m.dz);
    LHS(i, j, k) = divg_cell(c, i, j, k, dt);
}

// add_perturb! / remove_perturb! (src/Tools.jl:1348-1412): P += s at one cell, and the
// cell's wall ghosts follow it.
void perturb(const Mesh& m, Arr3<double>& P, double s, int64_t i, int64_t j, int64_t k) {
    P(i, j, k) += s;
    if (i == m.imin_) P(i - 1, j, k) = P(i, j, k);
    if (i == m.imax_) P(i + 1, j, k) = P(i, j, k);
    if (j == m.jmin_) P(i, j - 1, k) = P(i, j, k);
    if (j == m.jmax_) P(i, j + 1, k) = P(i, j, k);
    if (k == m.kmin_) P(i, j, k - 1) = P(i, j, k);
    if (k == m.kmax_) P(i, j, k + 1) = P(i, j, k);
}

// jacob_single (src/Pressure.jl:699-717). With exact_restore (not the reference) each cell's
// columns start from the unperturbed P and corrected faces: P is put back bit for bit and the
// faces a cell's box rewrote are restored, so no cell sees another's perturbation; the
// reference adds and subtracts delta (P drifts by its rounding) and leaves the box's faces
// to the next cell.
void jacob_single(FasLevel& L, Arr3<double>& P, double dt, bool exact_restore) {
    Ctx& c = *L.c;
    const Mesh& m = c.m;
    std::fill(L.L1.a.begin(), L.L1.a.end(), 0.0);
    std::fill(L.L2.a.begin(), L.L2.a.end(), 0.0);
    const double delta = 5.0;
    std::vector<double> P0, g0[3];
    if (exact_restore) {
        P0 = P.a;
        g0[0] = c.uf.a; g0[1] = c.vf.a; g0[2] = c.wf.a;
    }
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                perturb(m, P, delta, i, j, k);
                fas_A_local(L, P, i, j, k, L.L1, dt);
                perturb(m, P, -delta, i, j, k);
                if (exact_restore) P.a = P0;
                perturb(m, P, -delta, i, j, k);
                fas_A_local(L, P, i, j, k, L.L2, dt);
                perturb(m, P, delta, i, j, k);
                if (exact_restore) {
                    P.a = P0;
                    c.uf.a = g0[0]; c.vf.a = g0[1]; c.wf.a = g0[2];
                }
                L.jac(i, j, k) = sdiv(L.L1(i, j, k) - L.L2(i, j, k), 2.0 * delta);
            }
}

// LU with partial pivoting (first largest pivot) and substitution: KKT \ rhs.
std::vector<double> solve_dense(std::vector<double> A, std::vector<double> b, int n) {
    std::vector<int> piv(n);
    for (int col = 0; col < n; ++col) {
        int p = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[r * n + col]) > std::fabs(A[p * n + col])) p = r;
        if (p != col) {
            for (int q = 0; q < n; ++q) std::swap(A[col * n + q], A[p * n + q]);
            std::swap(b[col], b[p]);
        }
        for (int r = col + 1; r < n; ++r) {
            const double f = A[r * n + col] / A[col * n + col];
            A[r * n + col] = f;
            for (int q = col + 1; q < n; ++q) A[r * n + q] -= f * A[col * n + q];
            b[r] -= f * b[col];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = b[r];
        for (int q = r + 1; q < n; ++q) s -= A[r * n + q] * b[q];
        b[r] = s / A[r * n + r];
    }
    return b;
}

struct Anderson {
    int m = 10;
    int64_t len = 0;
    std::vector<std::vector<double>> F, G;  // Fhist, Phist: interior arrays, x fastest
    int hist_idx = 0, n_hist = 0;
    std::vector<double> gram;               // the Gram matrix entries of the slots
};

// par_anderson_accel (src/Pressure.jl:747-775): min ||sum a_i F_i|| with sum a_i = 1,
// G + 1e-10 max(diag G) I regularised, by the KKT system [G 1; 1' 0].
std::vector<double> anderson_alpha(const Anderson& A, int n) {
    std::vector<double> G((size_t)n * n);
    double mx = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) G[(size_t)(i * n + j)] = A.gram[(size_t)(i * A.m + j)];
    for (int i = 0; i < n; ++i) mx = i == 0 ? G[0] : jmax(mx, G[(size_t)(i * n + i)]);
    const double eps = 1e-10 * mx;
    for (int i = 0; i < n; ++i) G[(size_t)(i * n + i)] += eps;
    const int N = n + 1;
    std::vector<double> K((size_t)N * N, 0.0), rhs((size_t)N, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) K[(size_t)(i * N + j)] = G[(size_t)(i * n + j)];
        K[(size_t)(i * N + n)] = 1.0;
        K[(size_t)(n * N + i)] = 1.0;
    }
    rhs[(size_t)n] = 1.0;
    std::vector<double> a = solve_dense(K, rhs, N);
    a.resize((size_t)n);
    return a;
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t q = 0; q < a.size(); ++q) s += a[q] * b[q];
    return s;
}

}  // namespace

// res_iteration (src/Pressure.jl:777-884) on level L with P = L.P_h (or the level's working
// copy), tau added when use_tau. Returns the iterations run; *res the last max|A(P)+tau|.
static int64_t res_iteration(FasLevel& L, Arr3<double>& P, double dt, double tol, int64_t max_iter, bool use_tau,
                             bool* converged, Arr3<double>& AP, bool exact_restore, double* res) {
    Ctx& c = *L.c;
    const Mesh& m = c.m;
    std::fill(L.Gn.a.begin(), L.Gn.a.end(), 0.0);
    std::fill(L.jac.a.begin(), L.jac.a.end(), 0.0);
    const double omega = 0.8, beta = 0.5;
    Anderson H;
    H.len = m.Nx * m.Ny * m.Nz;
    H.F.assign((size_t)H.m, std::vector<double>((size_t)H.len, 0.0));
    H.G.assign((size_t)H.m, std::vector<double>((size_t)H.len, 0.0));
    H.gram.assign((size_t)(H.m * H.m), 0.0);
    fas_A(L, P, AP, dt);
    jacob_single(L, P, dt, exact_restore);
    int64_t it = 0;
    double rn = INFINITY;
    std::vector<double> Gn_loc((size_t)H.len), Fn_loc((size_t)H.len);
    while (it < max_iter) {
        ++it;
        fas_A(L, P, AP, dt);
        if (use_tau) interior_for(m, [&](int64_t i, int64_t j, int64_t k) { AP(i, j, k) += L.tau(i, j, k); });
        // parallel_max_all(abs.(AP)): the ghosts of AP are 0
        rn = 0.0;
        for (double v : AP.a) rn = jmax(rn, std::fabs(v));
        *res = rn;
        if (rn < tol) {
            if (converged) *converged = true;
            return it;
        }
        size_t q = 0;
        interior_for(m, [&](int64_t i, int64_t j, int64_t k) {
            const double Js = jmax(std::fabs(L.jac(i, j, k)), 1e-12);
            Gn_loc[q] = P(i, j, k) - (omega * AP(i, j, k)) / Js;
            Fn_loc[q] = Gn_loc[q] - P(i, j, k);
            ++q;
        });
        H.hist_idx = H.hist_idx % H.m + 1;
        if (H.n_hist < H.m) ++H.n_hist;
        const int slot = H.hist_idx - 1;
        H.F[(size_t)slot] = Fn_loc;
        H.G[(size_t)slot] = Gn_loc;
        // the Gram entries of the new slot (the others are unchanged)
        for (int s = 0; s < H.n_hist; ++s) {
            const double g = dot(H.F[(size_t)slot], H.F[(size_t)s]);
            H.gram[(size_t)(slot * H.m + s)] = g;
            H.gram[(size_t)(s * H.m + slot)] = g;
        }
        if (H.n_hist > 1) {
            const std::vector<double> a = anderson_alpha(H, H.n_hist);
            // weighted_sum (src/Pressure.jl:720-726): zeros, then += a_i * Phist_i in order
            std::vector<double> Pn((size_t)H.len, 0.0);
            for (int s = 0; s < H.n_hist; ++s)
                for (size_t r = 0; r < Pn.size(); ++r) Pn[r] += a[(size_t)s] * H.G[(size_t)s][r];
            size_t r = 0;
            interior_for(m, [&](int64_t i, int64_t j, int64_t k) {
                P(i, j, k) = beta * Pn[r] + (1.0 - beta) * P(i, j, k);
                ++r;
            });
        } else if (c.sl_first_step_newton) {
            size_t r = 0;
            interior_for(m, [&](int64_t i, int64_t j, int64_t k) { P(i, j, k) = Gn_loc[r++]; });
        } else {
            // P = Gn[interior]: the Gn argument, zero-filled above and never written
            interior_for(m, [&](int64_t i, int64_t j, int64_t k) { P(i, j, k) = L.Gn(i, j, k); });
        }
        // P .-= parallel_mean_all(P[interior]): the whole array, ghosts included
        double sum = 0.0;
        interior_for(m, [&](int64_t i, int64_t j, int64_t k) { sum += P(i, j, k); });
        const double mean = sum / (double)H.len;
        for (double& v : P.a) v -= mean;
    }
    return it;
}

namespace {

void restrict_cells(const Mesh& cm, Arr3<double>& C, const Arr3<double>& F) {
    for (int64_t k = cm.kmin_; k <= cm.kmax_; ++k)
        for (int64_t j = cm.jmin_; j <= cm.jmax_; ++j)
            for (int64_t i = cm.imin_; i <= cm.imax_; ++i) {
                const int64_t ii = 2 * i - 1, jj = 2 * j - 1, kk = 2 * k - 1;
                C(i, j, k) = (((((((F(ii, jj, kk) + F(ii + 1, jj, kk)) + F(ii, jj + 1, kk)) + F(ii + 1, jj + 1, kk)) +
                                 F(ii, jj, kk + 1)) + F(ii + 1, jj, kk + 1)) + F(ii, jj + 1, kk + 1)) +
                              F(ii + 1, jj + 1, kk + 1)) / 8.0;
            }
}

// restrict_{x,y,z}_face! (src/Multigrid.jl:242-279): the 4 fine faces covering a coarse face.
void restrict_faces(const Mesh& cm, FasLevel& C, const FasLevel& F) {
    for (int64_t k = cm.kmin_; k <= cm.kmax_; ++k)
        for (int64_t j = cm.jmin_; j <= cm.jmax_; ++j)
            for (int64_t i = cm.imin_; i <= cm.imax_ + 1; ++i) {
                const int64_t ii = 2 * i - 1, jj = 2 * j - 1, kk = 2 * k - 1;
                C.uf(i, j, k) = (((F.uf(ii, jj, kk) + F.uf(ii, jj, kk + 1)) + F.uf(ii, jj + 1, kk)) + F.uf(ii, jj + 1, kk + 1)) / 4.0;
            }
    for (int64_t k = cm.kmin_; k <= cm.kmax_; ++k)
        for (int64_t j = cm.jmin_; j <= cm.jmax_ + 1; ++j)
            for (int64_t i = cm.imin_; i <= cm.imax_; ++i) {
                const int64_t ii = 2 * i - 1, jj = 2 * j - 1, kk = 2 * k - 1;
                C.vf(i, j, k) = (((F.vf(ii, jj, kk) + F.vf(ii + 1, jj, kk)) + F.vf(ii, jj, kk + 1)) + F.vf(ii + 1, jj, kk + 1)) / 4.0;
            }
    for (int64_t k = cm.kmin_; k <= cm.kmax_ + 1; ++k)
        for (int64_t j = cm.jmin_; j <= cm.jmax_; ++j)
            for (int64_t i = cm.imin_; i <= cm.imax_; ++i) {
                const int64_t ii = 2 * i - 1, jj = 2 * j - 1, kk = 2 * k - 1;
                C.wf(i, j, k) = (((F.wf(ii, jj, kk) + F.wf(ii + 1, jj, kk)) + F.wf(ii, jj + 1, kk)) + F.wf(ii + 1, jj + 1, kk)) / 4.0;
            }
}

// prolong! (src/Multigrid.jl:5-54): trilinear from coarse centres (ghosts included).
void prolong(const Mesh& fm, const Mesh& cm, Arr3<double>& Fo, const Arr3<double>& C) {
    for (int64_t kf = fm.kmin_; kf <= fm.kmax_; ++kf)
        for (int64_t jf = fm.jmin_; jf <= fm.jmax_; ++jf)
            for (int64_t i_f = fm.imin_; i_f <= fm.imax_; ++i_f) {
                const double pt[3] = {fm.xm(i_f), fm.ym(jf), fm.zm(kf)};
                int64_t i = std::min(std::max(i_f, cm.imino_), cm.imaxo_);
                int64_t j = std::min(std::max(jf, cm.jmino_), cm.jmaxo_);
                int64_t k = std::min(std::max(kf, cm.kmino_), cm.kmaxo_);
                while (pt[0] - cm.xm(i) < 0.0 && i > cm.imino_) --i;
                while (pt[0] - cm.xm(i + 1) >= 0.0 && i + 1 < cm.imaxo_) ++i;
                while (pt[1] - cm.ym(j) < 0.0 && j > cm.jmino_) --j;
                while (pt[1] - cm.ym(j + 1) >= 0.0 && j + 1 < cm.jmaxo_) ++j;
                while (pt[2] - cm.zm(k) < 0.0 && k > cm.kmino_) --k;
                while (pt[2] - cm.zm(k + 1) >= 0.0 && k + 1 < cm.kmaxo_) ++k;
                const double wx1 = (pt[0] - cm.xm(i)) / (cm.xm(i + 1) - cm.xm(i)), wx2 = 1.0 - wx1;
                const double wy1 = (pt[1] - cm.ym(j)) / (cm.ym(j + 1) - cm.ym(j)), wy2 = 1.0 - wy1;
                const double wz1 = (pt[2] - cm.zm(k)) / (cm.zm(k + 1) - cm.zm(k)), wz2 = 1.0 - wz1;
                Fo(i_f, jf, kf) = wz1 * (wy1 * (wx1 * C(i + 1, j + 1, k + 1) + wx2 * C(i, j + 1, k + 1)) +
                                         wy2 * (wx1 * C(i + 1, j, k + 1) + wx2 * C(i, j, k + 1))) +
                                  wz2 * (wy1 * (wx1 * C(i + 1, j + 1, k) + wx2 * C(i, j + 1, k)) +
                                         wy2 * (wx1 * C(i + 1, j, k) + wx2 * C(i, j, k)));
            }
}

struct FasRun {
    std::vector<FasLevel> lv;  // index 0 = level 1 (finest)
    double dt = 0, tol = 0;
    bool exact_restore = false;
    int64_t smooth_iters = 0;  // res_iteration iterations over the whole solve
    double last_res = 0;
};

// mg_fas! (src/Multigrid.jl:375-500); lvl is 1-based as in the reference.
bool mg_fas(FasRun& R, int lvl) {
    const int nlev = (int)R.lv.size();
    FasLevel& L = R.lv[(size_t)lvl - 1];
    const int v1 = 10;
    double res = 0.0;
    if (lvl == nlev) {
        R.smooth_iters += res_iteration(L, L.P_h, R.dt, 1e-10, 50000, true, nullptr, L.AP_f, R.exact_restore, &res);
        return false;
    }
    // pre-smoothing (the tol of the level: param tol on level 1; coarse levels keep it too)
    R.smooth_iters += res_iteration(L, L.P_h, R.dt, R.tol, v1, lvl != 1, nullptr, L.AP_f, R.exact_restore, &res);
    FasLevel& C = R.lv[(size_t)lvl];
    Ctx& cc = *C.c;
    const Mesh& cm = cc.m;
    // VF of the coarse level, its band and densities
    restrict_cells(cm, cc.VF, L.c->VF);
    update_borders(cc, cc.VF);
    neumann(cc, cc.VF);
    compute_band(cc);
    // restricted approximation, with its ghosts
    restrict_cells(cm, C.P_h, L.P_h);
    neumann(cc, C.P_h);
    update_borders(cc, C.P_h);
    // restricted A(P) + tau of the fine level
    restrict_cells(cm, C.AP_f, L.AP_f);
    compute_dens(cc);
    restrict_faces(cm, C, L);
    // tau = R(A_h(P_h)) - A_H(R P_h)
    std::fill(C.AP_c.a.begin(), C.AP_c.a.end(), 0.0);
    fas_A(C, C.P_h, C.AP_c, R.dt);
    for (size_t q = 0; q < C.tau.a.size(); ++q) C.tau.a[q] = C.AP_f.a[q] - C.AP_c.a[q];
    C.P_bar_H.a = C.P_h.a;
    L.P_H.a = L.P_h.a;
    mg_fas(R, lvl + 1);
    for (size_t q = 0; q < C.P_h.a.size(); ++q) C.P_h.a[q] -= C.P_bar_H.a[q];
    std::fill(L.AP_f.a.begin(), L.AP_f.a.end(), 0.0);
    prolong(L.c->m, cm, L.AP_f, C.P_h);
    update_borders(*L.c, L.AP_f);
    interior_for(L.c->m, [&](int64_t i, int64_t j, int64_t k) { L.P_H(i, j, k) += L.AP_f(i, j, k); });
    std::fill(L.AP_f.a.begin(), L.AP_f.a.end(), 0.0);
    bool conv = false;
    R.smooth_iters += res_iteration(L, L.P_H, R.dt, R.tol, v1, lvl != 1, lvl == 1 ? &conv : nullptr, L.AP_f,
                                    R.exact_restore, &res);
    if (lvl == 1) R.last_res = res;
    L.P_h.a = L.P_H.a;
    return conv;
}

}  // namespace

// The SL-SL pressure solve, as pressure_solver! dispatches it (src/Pressure.jl:5-30):
// mg_lvl > 1: mg_cycler's FAS cycles (src/Multigrid.jl:333-371) until converged (at most
// max_cycles); mg_lvl == 1: res_iteration on the fine level (at most 50000 iterations).
int pressure_solve_sl(Ctx& c, double dt, double tol, int64_t max_cycles, bool exact_restore, int64_t* cycles,
                      double* resid, int64_t* smooth_iters) {
    const int nlev = (int)std::max<int64_t>(1, c.p.mg_lvl);
    FasRun R;
    R.dt = dt;
    R.tol = tol;
    R.exact_restore = exact_restore;
    R.lv.resize((size_t)nlev);
    for (int l = 1; l <= nlev; ++l) {
        FasLevel& L = R.lv[(size_t)l - 1];
        L.c = level_ctx(c, l);
        level_arrays(L);
    }
    // copy_to_mg! (src/Tools.jl:1414-1428) on level 1, and the fine VF
    FasLevel& F = R.lv[0];
    F.P_h.a = c.P.a;
    F.uf.a = c.uf.a; F.vf.a = c.vf.a; F.wf.a = c.wf.a;
    F.c->denx.a = c.denx.a; F.c->deny.a = c.deny.a; F.c->denz.a = c.denz.a;
    F.c->band.a = c.band.a;
    F.c->VF.a = c.VF.a;
    int64_t it = 0;
    bool conv = false;
    double res = 0.0;
    if (nlev == 1) {
        it = res_iteration(F, F.P_h, dt, tol, 50000, false, &conv, F.AP_f, exact_restore, &res);
        R.smooth_iters = it;
    } else {
        while (it < max_cycles) {
            ++it;
            conv = mg_fas(R, 1);
            res = R.last_res;
            if (conv) break;
        }
    }
    interior_for(c.m, [&](int64_t i, int64_t j, int64_t k) { c.P(i, j, k) = F.P_h(i, j, k); });
    neumann(c, c.P);
    update_borders(c, c.P);
    if (cycles) *cycles = it;
    if (resid) *resid = res;
    if (smooth_iters) *smooth_iters = R.smooth_iters;
    c.mg_cycles = it;
    c.mg_resid = res;
    return conv ? MIST_OK : MIST_ERR_NOCONV;
}

// A! on the fine level of c (the operator alone): LHS into c.divg, the corrected faces into
// gx, gy, gz. For the tests.
void sl_operator(Ctx& c, double dt, Arr3<double>& gx, Arr3<double>& gy, Arr3<double>& gz) {
    FasLevel L;
    L.c = level_ctx(c, 1);
    level_arrays(L);
    L.uf.a = c.uf.a; L.vf.a = c.vf.a; L.wf.a = c.wf.a;
    L.c->denx.a = c.denx.a; L.c->deny.a = c.deny.a; L.c->denz.a = c.denz.a;
    L.c->band.a = c.band.a;
    L.c->VF.a = c.VF.a;
    Arr3<double> P = c.P;
    fas_A(L, P, c.divg, dt);
    gx.a = L.c->uf.a; gy.a = L.c->vf.a; gz.a = L.c->wf.a;
}

// jacob_single on the fine level of c, into out (cell extents; interior set).
void sl_jacobian(Ctx& c, double dt, bool exact_restore, Arr3<double>& out) {
    FasLevel L;
    L.c = level_ctx(c, 1);
    level_arrays(L);
    L.uf.a = c.uf.a; L.vf.a = c.vf.a; L.wf.a = c.wf.a;
    L.c->denx.a = c.denx.a; L.c->deny.a = c.deny.a; L.c->denz.a = c.denz.a;
    L.c->band.a = c.band.a;
    L.c->VF.a = c.VF.a;
    Arr3<double> P = c.P;
    Arr3<double> AP;
    AP.init(L.c->m.cell_ext());
    fas_A(L, P, AP, dt);
    jacob_single(L, P, dt, exact_restore);
    out.a = L.jac.a;
}

// One KKT solve of par_anderson_accel from a given Gram matrix (for the tests).
void anderson_coefficients(const double* gram, int n, double* alpha) {
    Anderson H;
    H.gram.assign((size_t)(H.m * H.m), 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) H.gram[(size_t)(i * H.m + j)] = gram[i * n + j];
    const std::vector<double> a = anderson_alpha(H, n);
    for (int i = 0; i < n; ++i) alpha[i] = a[(size_t)i];
}

}  // namespace ora
