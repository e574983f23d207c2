// oracle_mg.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): variable-density pressure
// Poisson multigrid.
//
// Operator (src/Pressure.jl:909-1033 scaling): (A p)_c = sum_f c_f (p_nb - p_c),
// c_f = dt / (den_f * h^2), faces on walls dropped (Neumann), periodic faces wrapped;
// b = div(u*) (src/Pressure.jl:16-22). V-cycle structure of mg_vc_lin!
// (src/Multigrid.jl:678-764): red-black Gauss-Seidel smoothing (src/Pressure.jl:1096-1133,
// colour (i+j+k)%2 with red first, omega = 1), 8-cell average restriction
// (src/Multigrid.jl:183-197), trilinear prolongation from coarse centres
// (src/Multigrid.jl:5-54), correction added to the fine level. Coarse operators: the mean of
// the fine face coefficients (MIST_COARSE_FACE_AVERAGE), or the reference's face densities
// of the restricted liquid fraction (MIST_COARSE_RESTRICTED_VF). The outer loop evaluates
// the FP64 residual in flux form and runs one V-cycle (in T = double or float) on it per
// iteration until max|r| <= tol; the result is shifted to zero mean (src/Pressure.jl:1160).
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle_core.h"

namespace ora {

namespace {

template <class T>
struct Level {
    int64_t n[3] = {1, 1, 1};
    bool per[3] = {false, false, false};
    bool coarsened[3] = {false, false, false};  // this level is coarser than the previous in axis d
    std::vector<T> c[3];  // coefficient of the low face of each cell, per axis
    std::vector<T> e, b, r;
    int64_t size() const { return n[0] * n[1] * n[2]; }
    inline int64_t id(int64_t i, int64_t j, int64_t k) const { return i + n[0] * (j + n[1] * k); }
};

template <class T>
inline void neighbours(const Level<T>& L, int64_t i, int64_t j, int64_t k, int64_t nb[6], T cf[6]) {
    const int64_t q[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        int64_t lo[3] = {i, j, k}, hi[3] = {i, j, k};
        const int64_t nd = L.n[d];
        T clo = L.c[d][(size_t)L.id(i, j, k)];
        T chi;
        if (q[d] + 1 < nd) {
            hi[d] = q[d] + 1;
            chi = L.c[d][(size_t)L.id(hi[0], hi[1], hi[2])];
        } else if (L.per[d]) {
            hi[d] = 0;
            chi = L.c[d][(size_t)L.id(hi[0], hi[1], hi[2])];
        } else {
            chi = T(0);
        }
        if (q[d] > 0) lo[d] = q[d] - 1;
        else if (L.per[d]) lo[d] = nd - 1;
        nb[2 * d] = L.id(lo[0], lo[1], lo[2]);
        nb[2 * d + 1] = L.id(hi[0], hi[1], hi[2]);
        cf[2 * d] = clo;
        cf[2 * d + 1] = chi;
    }
}

template <class T>
void smooth(Level<T>& L, int sweeps) {
    for (int s = 0; s < sweeps; ++s)
        for (int color = 0; color < 2; ++color) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int64_t k = 0; k < L.n[2]; ++k)
                for (int64_t j = 0; j < L.n[1]; ++j)
                    for (int64_t i = 0; i < L.n[0]; ++i) {
                        if (((i + 1) + (j + 1) + (k + 1)) % 2 != color) continue;
                        int64_t nb[6];
                        T cf[6];
                        neighbours(L, i, j, k, nb, cf);
                        const T S = ((((cf[0] + cf[1]) + cf[2]) + cf[3]) + cf[4]) + cf[5];
                        if (!(S > T(0))) continue;
                        const T sum = ((((cf[0] * L.e[(size_t)nb[0]] + cf[1] * L.e[(size_t)nb[1]]) +
                                         cf[2] * L.e[(size_t)nb[2]]) + cf[3] * L.e[(size_t)nb[3]]) +
                                       cf[4] * L.e[(size_t)nb[4]]) + cf[5] * L.e[(size_t)nb[5]];
                        L.e[(size_t)L.id(i, j, k)] = (sum - L.b[(size_t)L.id(i, j, k)]) / S;
                    }
        }
}

template <class T>
void residual(Level<T>& L) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < L.n[2]; ++k)
        for (int64_t j = 0; j < L.n[1]; ++j)
            for (int64_t i = 0; i < L.n[0]; ++i) {
                int64_t nb[6];
                T cf[6];
                neighbours(L, i, j, k, nb, cf);
                const int64_t c = L.id(i, j, k);
                const T pc = L.e[(size_t)c];
                const T Ap = ((((cf[0] * (L.e[(size_t)nb[0]] - pc) + cf[1] * (L.e[(size_t)nb[1]] - pc)) +
                                cf[2] * (L.e[(size_t)nb[2]] - pc)) + cf[3] * (L.e[(size_t)nb[3]] - pc)) +
                              cf[4] * (L.e[(size_t)nb[4]] - pc)) + cf[5] * (L.e[(size_t)nb[5]] - pc);
                L.r[(size_t)c] = L.b[(size_t)c] - Ap;
            }
}

template <class T>
void restrict_to(const Level<T>& F, Level<T>& C) {
    const int64_t f0 = C.coarsened[0] ? 2 : 1, f1 = C.coarsened[1] ? 2 : 1, f2 = C.coarsened[2] ? 2 : 1;
    const T inv = T(1) / T(f0 * f1 * f2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t K = 0; K < C.n[2]; ++K)
        for (int64_t J = 0; J < C.n[1]; ++J)
            for (int64_t I = 0; I < C.n[0]; ++I) {
                T s = T(0);
                for (int64_t c = 0; c < f2; ++c)
                    for (int64_t b = 0; b < f1; ++b)
                        for (int64_t a = 0; a < f0; ++a) s += F.r[(size_t)F.id(I * f0 + a, J * f1 + b, K * f2 + c)];
                C.b[(size_t)C.id(I, J, K)] = s * inv;
            }
}

// Coarse face coefficient = mean of the fine face coefficients covering it, rescaled from
// fine to coarse spacing along the face normal.
template <class T>
void coarsen_coefficients(const Level<T>& F, Level<T>& C) {
    const int64_t f[3] = {C.coarsened[0] ? 2 : 1, C.coarsened[1] ? 2 : 1, C.coarsened[2] ? 2 : 1};
    for (int d = 0; d < 3; ++d) {
        C.c[d].assign((size_t)C.size(), T(0));
        const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        const T scale = T(1) / T(f[d] * f[d]) / T(f[d1] * f[d2]);
        for (int64_t K = 0; K < C.n[2]; ++K)
            for (int64_t J = 0; J < C.n[1]; ++J)
                for (int64_t I = 0; I < C.n[0]; ++I) {
                    const int64_t Q[3] = {I, J, K};
                    T s = T(0);
                    for (int64_t b = 0; b < f[d2]; ++b)
                        for (int64_t a = 0; a < f[d1]; ++a) {
                            int64_t q[3];
                            q[d] = Q[d] * f[d];
                            q[d1] = Q[d1] * f[d1] + a;
                            q[d2] = Q[d2] * f[d2] + b;
                            s += F.c[d][(size_t)F.id(q[0], q[1], q[2])];
                        }
                    C.c[d][(size_t)C.id(I, J, K)] = s * scale;
                }
    }
}

// The liquid fraction of level C: the mean of the finer level's over each coarse cell
// (restrict!, src/Multigrid.jl:183-197: the 8 fine values summed in i, then j, then k order,
// divided by 8), vf of the finer level in plain order.
template <class T>
void restrict_vf(const Level<T>& F, const Level<T>& C, const std::vector<double>& vf, std::vector<double>& out) {
    const int64_t f0 = C.coarsened[0] ? 2 : 1, f1 = C.coarsened[1] ? 2 : 1, f2 = C.coarsened[2] ? 2 : 1;
    const double inv = 1.0 / (double)(f0 * f1 * f2);
    out.assign((size_t)C.size(), 0.0);
    for (int64_t K = 0; K < C.n[2]; ++K)
        for (int64_t J = 0; J < C.n[1]; ++J)
            for (int64_t I = 0; I < C.n[0]; ++I) {
                double s = 0.0;
                for (int64_t c = 0; c < f2; ++c)
                    for (int64_t b = 0; b < f1; ++b)
                        for (int64_t a = 0; a < f0; ++a) s += vf[(size_t)F.id(I * f0 + a, J * f1 + b, K * f2 + c)];
                out[(size_t)C.id(I, J, K)] = s * inv;
            }
}

// Face coefficients of a coarse level from its liquid fraction (mg_vc_lin!,
// src/Multigrid.jl:725-737): compute_dens! (src/Tools.jl:1212-1232) gives the face density
// rho_liq vf + rho_gas (1 - vf) with vf = clamp((VF_c + VF_{c-1}) / 2, 0, 1), and the
// coefficient is dt / (den h^2) with the level's spacing h, as on the fine level. Wall faces
// are dropped and periodic faces wrap, as on the fine level.
template <class T>
void coefficients_from_vf(Level<T>& C, const std::vector<double>& vf, double dt, double rl, double rg,
                          const double h[3]) {
    for (int d = 0; d < 3; ++d) {
        C.c[d].assign((size_t)C.size(), T(0));
        const double h2 = h[d] * h[d];
        for (int64_t K = 0; K < C.n[2]; ++K)
            for (int64_t J = 0; J < C.n[1]; ++J)
                for (int64_t I = 0; I < C.n[0]; ++I) {
                    const int64_t Q[3] = {I, J, K};
                    int64_t lo[3] = {I, J, K};
                    if (Q[d] > 0) lo[d] = Q[d] - 1;
                    else if (C.per[d]) lo[d] = C.n[d] - 1;
                    else continue;
                    const double v = std::min(1.0, std::max(0.0, (vf[(size_t)C.id(I, J, K)] +
                                                                  vf[(size_t)C.id(lo[0], lo[1], lo[2])]) / 2));
                    const double den = rl * v + rg * (1 - v);
                    C.c[d][(size_t)C.id(I, J, K)] = (T)(dt / (den * h2));
                }
    }
}

template <class T>
void prolong_add(const Level<T>& C, Level<T>& F) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < F.n[2]; ++k)
        for (int64_t j = 0; j < F.n[1]; ++j)
            for (int64_t i = 0; i < F.n[0]; ++i) {
                int64_t lo[3], hi[3];
                T w1[3], w2[3];
                const int64_t q[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    if (!C.coarsened[d]) { lo[d] = hi[d] = q[d]; w1[d] = T(1); w2[d] = T(0); continue; }
                    const int64_t Q = q[d] / 2;
                    if (q[d] % 2 == 0) { lo[d] = Q - 1; hi[d] = Q; w1[d] = T(0.75); w2[d] = T(0.25); }
                    else { lo[d] = Q; hi[d] = Q + 1; w1[d] = T(0.25); w2[d] = T(0.75); }
                    if (lo[d] < 0) lo[d] = C.per[d] ? C.n[d] - 1 : 0;
                    if (hi[d] >= C.n[d]) hi[d] = C.per[d] ? 0 : C.n[d] - 1;
                }
                auto E = [&](int64_t a, int64_t b, int64_t c) { return C.e[(size_t)C.id(a, b, c)]; };
                const T v = w1[2] * (w1[1] * (w1[0] * E(hi[0], hi[1], hi[2]) + w2[0] * E(lo[0], hi[1], hi[2])) +
                                     w2[1] * (w1[0] * E(hi[0], lo[1], hi[2]) + w2[0] * E(lo[0], lo[1], hi[2]))) +
                            w2[2] * (w1[1] * (w1[0] * E(hi[0], hi[1], lo[2]) + w2[0] * E(lo[0], hi[1], lo[2])) +
                                     w2[1] * (w1[0] * E(hi[0], lo[1], lo[2]) + w2[0] * E(lo[0], lo[1], lo[2])));
                F.e[(size_t)F.id(i, j, k)] += v;
            }
}

template <class T>
void remove_mean(std::vector<T>& a) {
    double s = 0.0;
    for (T v : a) s += (double)v;
    const T mean = (T)(s / (double)a.size());
    for (T& v : a) v -= mean;
}

template <class T>
void vcycle(std::vector<Level<T>>& lv, size_t l, int nu1, int nu2, int coarse_sweeps) {
    Level<T>& L = lv[l];
    if (l + 1 == lv.size()) {
        std::fill(L.e.begin(), L.e.end(), T(0));
        remove_mean(L.b);
        smooth(L, coarse_sweeps);
        return;
    }
    if (l > 0) std::fill(L.e.begin(), L.e.end(), T(0));
    smooth(L, nu1);
    residual(L);
    restrict_to(L, lv[l + 1]);
    vcycle(lv, l + 1, nu1, nu2, coarse_sweeps);
    prolong_add(lv[l + 1], L);
    smooth(L, nu2);
}

// The fine FP64 operator and the hierarchy of precision T with its coarse operators.
template <class T>
int build_levels(Ctx& c, double dt, Level<double>& fine, std::vector<Level<T>>& lv) {
    const Mesh& m = c.m;
    const mist_params& p = c.p;
    fine.n[0] = m.Nx; fine.n[1] = m.Ny; fine.n[2] = m.Nz;
    fine.per[0] = p.xper && m.Nx > 1; fine.per[1] = p.yper && m.Ny > 1; fine.per[2] = p.zper && m.Nz > 1;
    const double h2[3] = {m.dx * m.dx, m.dy * m.dy, m.dz * m.dz};
    const Arr3<double>* den[3] = {&c.denx, &c.deny, &c.denz};
    for (int d = 0; d < 3; ++d) fine.c[d].assign((size_t)fine.size(), 0.0);
    for (int64_t k = 0; k < m.Nz; ++k)
        for (int64_t j = 0; j < m.Ny; ++j)
            for (int64_t i = 0; i < m.Nx; ++i) {
                const int64_t q[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    double cf = 0.0;
                    if (q[d] > 0 || fine.per[d]) cf = dt / ((*den[d])(i + 1, j + 1, k + 1) * h2[d]);
                    fine.c[d][(size_t)fine.id(i, j, k)] = cf;
                }
            }
    // hierarchy
    int64_t nlev = p.mg_lvl;
    if (nlev <= 1) {
        nlev = 1;
        int64_t n[3] = {m.Nx, m.Ny, m.Nz};
        for (;;) {
            const int64_t mx = std::max(n[0], std::max(n[1], n[2]));
            bool ok = mx > 4;
            for (int d = 0; d < 3; ++d) if (n[d] > 1 && n[d] % 2) ok = false;
            if (!ok) break;
            for (int d = 0; d < 3; ++d) if (n[d] > 1) n[d] /= 2;
            ++nlev;
        }
    }
    lv.resize((size_t)nlev);
    for (int d = 0; d < 3; ++d) { lv[0].n[d] = fine.n[d]; lv[0].per[d] = fine.per[d]; }
    for (int d = 0; d < 3; ++d) {
        lv[0].c[d].resize(fine.c[d].size());
        for (size_t q = 0; q < fine.c[d].size(); ++q) lv[0].c[d][q] = (T)fine.c[d][q];
    }
    for (size_t l = 1; l < lv.size(); ++l) {
        for (int d = 0; d < 3; ++d) {
            const int64_t nf = lv[l - 1].n[d];
            if (nf > 1 && nf % 2) { c.err = "multigrid: odd extent cannot be coarsened"; return MIST_ERR_PARAM; }
            lv[l].coarsened[d] = nf > 1;
            lv[l].n[d] = nf > 1 ? nf / 2 : 1;
            lv[l].per[d] = lv[l - 1].per[d] && lv[l].n[d] > 1;
        }
        if (p.mg_coarse_op != MIST_COARSE_RESTRICTED_VF) coarsen_coefficients(lv[l - 1], lv[l]);
    }
    if (p.mg_coarse_op == MIST_COARSE_RESTRICTED_VF) {
        std::vector<double> vf((size_t)fine.size()), vc;
        for (int64_t k = 0; k < m.Nz; ++k)
            for (int64_t j = 0; j < m.Ny; ++j)
                for (int64_t i = 0; i < m.Nx; ++i) vf[(size_t)fine.id(i, j, k)] = c.VF(i + 1, j + 1, k + 1);
        double h[3] = {m.dx, m.dy, m.dz};
        for (size_t l = 1; l < lv.size(); ++l) {
            restrict_vf(lv[l - 1], lv[l], vf, vc);
            for (int d = 0; d < 3; ++d)
                if (lv[l].coarsened[d]) h[d] *= 2.0;
            coefficients_from_vf(lv[l], vc, dt, p.rho_liq, p.rho_gas, h);
            vf.swap(vc);
        }
    }
    for (auto& L : lv) {
        L.e.assign((size_t)L.size(), T(0));
        L.b.assign((size_t)L.size(), T(0));
        L.r.assign((size_t)L.size(), T(0));
    }
    return MIST_OK;
}

template <class T>
int solve(Ctx& c, double dt, double tol, int64_t* cycles_out, double* resid_out) {
    const Mesh& m = c.m;
    const mist_params& p = c.p;
    std::vector<Level<T>> lv;
    Level<double> fine;  // FP64 operator for the outer residual
    const int bst = build_levels(c, dt, fine, lv);
    if (bst) return bst;
    // right-hand side (mean removed) and initial guess
    fine.b.resize((size_t)fine.size());
    fine.e.resize((size_t)fine.size());
    fine.r.resize((size_t)fine.size());
    for (int64_t k = 0; k < m.Nz; ++k)
        for (int64_t j = 0; j < m.Ny; ++j)
            for (int64_t i = 0; i < m.Nx; ++i) {
                fine.b[(size_t)fine.id(i, j, k)] = c.RHS(i + 1, j + 1, k + 1);
                fine.e[(size_t)fine.id(i, j, k)] = c.P(i + 1, j + 1, k + 1);
            }
    remove_mean(fine.b);
    int64_t cycles = 0;
    double res = 0.0;
    int st = MIST_OK;
    const int64_t maxc = p.mg_max_cycles > 0 ? p.mg_max_cycles : 10000;
    for (;;) {
        residual(fine);
        res = 0.0;
        bool nan = false;
        for (double v : fine.r) {
            nan |= v != v;
            res = std::max(res, std::fabs(v));
        }
        if (nan) {  // a diverged iteration: reported, not taken for convergence
            res = NAN;
            st = MIST_ERR_NOCONV;
            break;
        }
        if (res <= tol) break;
        if (cycles >= maxc) { st = MIST_ERR_NOCONV; break; }
        for (size_t q = 0; q < fine.r.size(); ++q) { lv[0].b[q] = (T)fine.r[q]; lv[0].e[q] = T(0); }
        vcycle(lv, 0, (int)p.mg_nu1, (int)p.mg_nu2, (int)p.mg_coarse_sweeps);
        for (size_t q = 0; q < fine.e.size(); ++q) fine.e[q] += (double)lv[0].e[q];
        ++cycles;
    }
    remove_mean(fine.e);
    for (int64_t k = 0; k < m.Nz; ++k)
        for (int64_t j = 0; j < m.Ny; ++j)
            for (int64_t i = 0; i < m.Nx; ++i) c.P(i + 1, j + 1, k + 1) = fine.e[(size_t)fine.id(i, j, k)];
    neumann(c, c.P);
    update_borders(c, c.P);
    *cycles_out = cycles;
    *resid_out = res;
    c.mg_cycles = cycles;
    c.mg_resid = res;
    return st;
}

}  // namespace

int mg_level_coefficients(Ctx& c, double dt, int level, int axis, std::vector<double>& out, int64_t n[3]) {
    Level<double> fine;
    std::vector<Level<double>> lv;
    const int st = build_levels(c, dt, fine, lv);
    if (st) return st;
    if (level < 0 || (size_t)level >= lv.size() || axis < 0 || axis > 2) return MIST_ERR_PARAM;
    const Level<double>& L = lv[(size_t)level];
    for (int d = 0; d < 3; ++d) n[d] = L.n[d];
    out = L.c[axis];
    return MIST_OK;
}

int pressure_solve(Ctx& c, double dt, double tol, int precision, int64_t* cycles, double* resid) {
    if (precision == MIST_MG_MIXED) return solve<float>(c, dt, tol, cycles, resid);
    return solve<double>(c, dt, tol, cycles, resid);
}

}  // namespace ora
