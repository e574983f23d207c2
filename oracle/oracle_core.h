/*
 * oracle_core.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * FP64 CPU restatement of the reference (julianfox8/NS_Multiphase, "Mist.jl") for the
 * SL-VOF transport and pressure multigrid path, written from reading the Julia source.
 * Each function cites the reference file:line it follows. Only tests/, the package's
 * smoke() and bench.py's cpu_baseline leg load this code, and only as the checker; the
 * product library never links or calls it.
 *
 * Arithmetic follows the reference operation by operation (same association order,
 * Julia's literal powers x^2 = x*x and x^3 = (x*x)*x, sequential left-fold sums for the
 * short `sum`s, Julia min/max semantics). It is compiled with -ffp-contract=off and no
 * fast-math so no FMA is formed.
 */
#ifndef MIST_ORACLE_CORE_H
#define MIST_ORACLE_CORE_H

#include <cmath>
#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "../include/mist_hip.h"

namespace ora {

constexpr double EPS = 2.220446049250313e-16;  // Julia eps() for Float64

// Julia's min/max on Float64 (NaN-propagating; -0.0 < 0.0).
inline double jmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}
inline double jmin(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
// Safe divide  a /^ b  (src/Tools.jl:375-377): typemax(Float64) = Inf when |b| <= eps().
inline double sdiv(double a, double b) { return std::fabs(b) <= EPS ? INFINITY : a / b; }

struct Ext {
    int64_t lo[3];
    int64_t n[3];
    int64_t size() const { return n[0] * n[1] * n[2]; }
};

template <class T>
struct Arr3 {
    Ext e{};
    std::vector<T> a;
    void init(const Ext& ext, T v = T(0)) { e = ext; a.assign((size_t)ext.size(), v); }
    inline int64_t idx(int64_t i, int64_t j, int64_t k) const {
        return (i - e.lo[0]) + e.n[0] * ((j - e.lo[1]) + e.n[1] * (k - e.lo[2]));
    }
    inline T& operator()(int64_t i, int64_t j, int64_t k) { return a[(size_t)idx(i, j, k)]; }
    inline const T& operator()(int64_t i, int64_t j, int64_t k) const { return a[(size_t)idx(i, j, k)]; }
    inline int64_t hi(int d) const { return e.lo[d] + e.n[d] - 1; }
};

// mesh_struct + create_mesh for one rank (src/Mesh.jl:1-196, :204-352 for MG levels).
struct Mesh {
    int64_t Nx = 0, Ny = 0, Nz = 0, ng = 0;
    double Lx = 0, Ly = 0, Lz = 0, dx = 0, dy = 0, dz = 0;
    int64_t imin_ = 1, imax_ = 0, jmin_ = 1, jmax_ = 0, kmin_ = 1, kmax_ = 0;
    int64_t imino_ = 0, imaxo_ = 0, jmino_ = 0, jmaxo_ = 0, kmino_ = 0, kmaxo_ = 0;
    std::vector<double> xs, ys, zs, xms, yms, zms;  // offset by imino_ etc.
    inline double x(int64_t i) const { return xs[(size_t)(i - imino_)]; }
    inline double y(int64_t j) const { return ys[(size_t)(j - jmino_)]; }
    inline double z(int64_t k) const { return zs[(size_t)(k - kmino_)]; }
    inline double xm(int64_t i) const { return xms[(size_t)(i - imino_)]; }
    inline double ym(int64_t j) const { return yms[(size_t)(j - jmino_)]; }
    inline double zm(int64_t k) const { return zms[(size_t)(k - kmino_)]; }
    void build(const mist_params& p, int level);  // level 1 = finest
    Ext cell_ext() const;
    Ext vof_ext() const;
    Ext face_ext(int dir) const;
};

struct CutScratch;  // per-thread cutTet storage

struct CutSums {
    double vol = 0, vLiq = 0, vU = 0, vV = 0, vW = 0, vrho = 0;
    int lvl = 0;
};

// The oracle's context: parameters, mesh and the reference arrays (src/Tools.jl:136-191).
struct Ctx {
    mist_params p{};
    Mesh m;
    Arr3<double> VF, nx, ny, nz, D, u, v, w, uf, vf, wf, P, divg, denx, deny, denz, RHS;
    Arr3<double> us, vs, ws, viscx, viscy, viscz, Curve, sfx, sfy, sfz;
    Arr3<int16_t> band;
    Arr3<int16_t> mask;          // bits 0..2: mask[i,j,k][1..3] of the reference
    double bc_wall[6][3] = {};  // wall velocity of sides x-, x+, y-, y+, z-, z+
    std::string err;
    int status = 0;
    int64_t max_cut_level = 0;
    int64_t mg_cycles = 0;
    double mg_resid = 0;
    // Tie hints of a parity check against another implementation. Each compute_normal call
    // takes the front array of elv_hints (if any): the ELVIRA candidate the other side kept
    // per interior cell (x fastest, -1 none); each compute_curvature call takes the front of
    // curv_hints: the height-function direction it used (0 x, 1 y, 2 z, -1 none). A hinted
    // choice replaces the oracle's own only where the oracle's scores tie with it to rounding
    // (see elvira_tie / direction_tie); tie_stats[kind] = {cells whose hint differed, adopted,
    // refused}, tie_gap[kind] = the largest relative score gap adopted.
    std::deque<std::vector<int32_t>> elv_hints, curv_hints;
    int64_t tie_stats[2][3] = {};
    double tie_gap[2] = {};
    struct Refusal {
        int kind;
        int64_t i, j, k;
        int hint, own;
        double e_hint, e_own;  // ELVIRA errors; for directions the |n| components of hint and own
    };
    std::vector<Refusal> refusals;  // the first 64 refused hints
    // SL pressure, the build's decisions on two reference quirks (false gives the reference's
    // literal form): jacob_single restores P and the corrected faces exactly per cell (see
    // jacob_single in oracle_fas.cpp), and res_iteration's first iteration takes the
    // diagonal-Newton step Gn_loc (the reference copies its zero-filled Gn argument into P)
    bool sl_exact_restore = true;
    bool sl_first_step_newton = true;
};

// ---- geometry primitives (src/VFgeom.jl, src/VF.jl, src/Tools.jl) -----------------
double tet_vol(const double v[4][3]);
int d2case_inplace(double d[4]);
int d2case_copy(const double d[4]);
double compute_alpha(double m0, double m1, double m2, double VF);
double compute_dist(const Ctx& c, int64_t i, int64_t j, int64_t k, double nx, double ny,
                    double nz, double VF);
double cell2verts(const Ctx& c, int64_t i, int64_t j, int64_t k, double verts[8][3]);
void verts2tets5(const double p[8][3], double tets[5][4][3]);
void inds2tets5(const int64_t p[8][3], int64_t tets[5][4][3]);
double plic2vf(const Ctx& c, int64_t i, int64_t j, int64_t k, double nx, double ny,
               double nz, double dist);
void project(const Ctx& c, double pt[3], int64_t i, int64_t j, int64_t k, double dt);
bool pt2index(const Ctx& c, int64_t I[3], const double pt[3], int64_t i, int64_t j, int64_t k);
double vel_uface(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k);
double vel_vface(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k);
double vel_wface(const Ctx& c, const double pt[3], int64_t i, int64_t j, int64_t k);
CutSums cut_tet(Ctx& c, double tet[4][3], int64_t ind[4][3], bool xdone, bool ydone,
                bool zdone, int lvl, CutScratch& s);
void elvira_calc(const Ctx& c, int64_t i, int64_t j, int64_t k, double* onx, double* ony,
                 double* onz, int* best_candidate, double cand_err[27] = nullptr,
                 double cand_n[27][3] = nullptr);

// ---- stages ----------------------------------------------------------------------
Arr3<double>* field_f64(Ctx& c, int field);
Arr3<int16_t>* field_i16(Ctx& c, int field);
void neumann(Ctx& c, Arr3<double>& A);
void neumann_i16(Ctx& c, Arr3<int16_t>& A);
void update_borders(Ctx& c, Arr3<double>& A);
void update_borders_i16(Ctx& c, Arr3<int16_t>& A);
void define_velocity(Ctx& c, double t);
double compute_dt(const Ctx& c);
void compute_band(Ctx& c);
void compute_normal(Ctx& c);
void compute_plic(Ctx& c);
int transport_vof(Ctx& c, double dt);
void compute_dens(Ctx& c);
void pressure_rhs(Ctx& c);
int pressure_solve(Ctx& c, double dt, double tol, int precision, int64_t* cycles, double* resid);
int mg_level_coefficients(Ctx& c, double dt, int level, int axis, std::vector<double>& out, int64_t n[3]);
int pressure_solve_sl(Ctx& c, double dt, double tol, int64_t max_cycles, bool exact_restore, int64_t* cycles,
                      double* resid, int64_t* smooth_iters);
void sl_operator(Ctx& c, double dt, Arr3<double>& gx, Arr3<double>& gy, Arr3<double>& gz);
void sl_jacobian(Ctx& c, double dt, bool exact_restore, Arr3<double>& out);
void anderson_coefficients(const double* gram, int n, double* alpha);
void corrector(Ctx& c, double dt);
void compute_props(Ctx& c);
double sl_flux_vol(const Ctx& c, int dir, int64_t i, int64_t j, int64_t k, double dt, double verts[8][3]);
bool add_correction_tets(const Ctx& c, int64_t i, int64_t j, int64_t k, double dt, double tets[17][4][3],
                         int64_t inds[17][4][3]);
void compute_curvature(Ctx& c);
void mask_maker(Ctx& c);
void compute_sf(Ctx& c);
void apply_bc(Ctx& c, Arr3<double>& u, Arr3<double>& v, Arr3<double>& w);
void interpolate_face(Ctx& c, const Arr3<double>& u, const Arr3<double>& v, const Arr3<double>& w);
void interpolate_center(Ctx& c);
void divergence(Ctx& c, double dt);
double divg_cell(const Ctx& c, int64_t i, int64_t j, int64_t k, double dt);
int transport_ns(Ctx& c, double dt);
int ns_step(Ctx& c, double* t, double clamp_to, double* dt_used, double tol, int precision, int64_t* cycles,
            double* resid);
CutScratch* new_cut_scratch();
void delete_cut_scratch(CutScratch* s);
double vf_circle(double xmin, double xmax, double ymin, double ymax, double rad, double xo, double yo);
double vf_sphere(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
                 double rad, double xo, double yo, double zo, bool bubble);

// Operation counter for the ELVIRA path (+ - * / = 1 flop each; sqrt, cbrt and trig are
// counted separately). Per-thread accumulation, summed by the caller.
struct FlopCount {
    uint64_t flops = 0;
    uint64_t special = 0;
};
extern thread_local FlopCount t_flops;
extern bool g_count_flops;

}  // namespace ora

#endif
