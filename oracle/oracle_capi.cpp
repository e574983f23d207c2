// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): extern "C" entry points used by
// tests/ (through ctypes) and by bench.py's cpu_baseline leg. They mirror the product's
// C ABI (include/mist_hip.h) on host arrays, plus single-cell primitives for the pins.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

using namespace ora;

namespace ora {
CutScratch* new_cut_scratch();
void delete_cut_scratch(CutScratch* s);
}

static std::string g_err;

static int check_params(const mist_params* p, std::string& err) {
    if (p->Nx < 1 || p->Ny < 1 || p->Nz < 1) { err = "Nx, Ny, Nz must be >= 1"; return MIST_ERR_PARAM; }
    if (!(p->Lx > 0 && p->Ly > 0 && p->Lz > 0)) { err = "Lx, Ly, Lz must be > 0"; return MIST_ERR_PARAM; }
    if (p->tesselation != MIST_TESS_5_TETS) { err = "incompatible tesselation called"; return MIST_ERR_PARAM; }
    if (p->normal_method != MIST_NORMAL_ELVIRA) { err = "Unknown method to copute interface normal"; return MIST_ERR_PARAM; }
    if (p->projection_method < 0 || p->projection_method > 2) { err = "Unknown projection_method in project!"; return MIST_ERR_PARAM; }
    if (p->mg_coarse_op != MIST_COARSE_FACE_AVERAGE && p->mg_coarse_op != MIST_COARSE_RESTRICTED_VF) { err = "unknown mg_coarse_op"; return MIST_ERR_PARAM; }
    return MIST_OK;
}

extern "C" {

void ora_params_default(mist_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->max_dt = INFINITY;
    p->solveNS = 1;
    p->pressure_scheme = MIST_PS_SEMI_LAGRANGIAN;
    p->mg_lvl = 1;
    p->VFlo = 1e-10;
    p->VFhi = 1.0 - 1e-10;
    p->nband = -1;
    p->vel_period = 2.0;
    p->mg_precision = MIST_MG_FP64;
    p->mg_nu1 = 10;
    p->mg_nu2 = 10;
    p->mg_max_cycles = 10000;
    p->mg_coarse_op = MIST_COARSE_FACE_AVERAGE;
    p->mg_coarse_sweeps = 64;
}

int ora_ctx_create(const mist_params* p, void** out) {
    *out = nullptr;
    std::string err;
    int st = check_params(p, err);
    if (st) { g_err = err; return st; }
    auto c = std::make_unique<Ctx>();
    c->p = *p;
    if (c->p.nband < 0) c->p.nband = (int64_t)std::ceil(c->p.CFL);
    c->m.build(c->p, 1);
    const Ext ce = c->m.cell_ext();
    c->VF.init(c->m.vof_ext());
    for (Arr3<double>* a : {&c->nx, &c->ny, &c->nz, &c->D, &c->u, &c->v, &c->w, &c->uf, &c->vf, &c->wf,
                            &c->P, &c->divg, &c->RHS, &c->us, &c->vs, &c->ws, &c->Curve})
        a->init(ce);
    c->denx.init(c->m.face_ext(0));
    c->deny.init(c->m.face_ext(1));
    c->denz.init(c->m.face_ext(2));
    c->viscx.init(c->m.face_ext(0));
    c->viscy.init(c->m.face_ext(1));
    c->viscz.init(c->m.face_ext(2));
    c->sfx.init(c->m.face_ext(0));
    c->sfy.init(c->m.face_ext(1));
    c->sfz.init(c->m.face_ext(2));
    c->band.init(ce);
    c->mask.init(ce);
    *out = c.release();
    return MIST_OK;
}

void ora_ctx_destroy(void* h) { delete (Ctx*)h; }

const char* ora_last_error(void* h) { return h ? ((Ctx*)h)->err.c_str() : g_err.c_str(); }

int ora_field_extents(void* h, int field, int64_t lo[3], int64_t n[3], int64_t* elem) {
    Ctx& c = *(Ctx*)h;
    const Ext* e = nullptr;
    if (Arr3<int16_t>* b = field_i16(c, field)) { e = &b->e; *elem = 2; }
    else { Arr3<double>* a = field_f64(c, field); if (!a) return MIST_ERR_PARAM; e = &a->e; *elem = 8; }
    for (int d = 0; d < 3; ++d) { lo[d] = e->lo[d]; n[d] = e->n[d]; }
    return MIST_OK;
}

int ora_set_field(void* h, int field, const void* src, size_t nbytes) {
    Ctx& c = *(Ctx*)h;
    if (Arr3<int16_t>* b = field_i16(c, field)) {
        if (nbytes != b->a.size() * 2) return MIST_ERR_PARAM;
        std::memcpy(b->a.data(), src, nbytes);
        return MIST_OK;
    }
    Arr3<double>* a = field_f64(c, field);
    if (!a || nbytes != a->a.size() * 8) return MIST_ERR_PARAM;
    std::memcpy(a->a.data(), src, nbytes);
    return MIST_OK;
}

int ora_get_field(void* h, int field, void* dst, size_t nbytes) {
    Ctx& c = *(Ctx*)h;
    if (Arr3<int16_t>* b = field_i16(c, field)) {
        if (nbytes != b->a.size() * 2) return MIST_ERR_PARAM;
        std::memcpy(dst, b->a.data(), nbytes);
        return MIST_OK;
    }
    Arr3<double>* a = field_f64(c, field);
    if (!a || nbytes != a->a.size() * 8) return MIST_ERR_PARAM;
    std::memcpy(dst, a->a.data(), nbytes);
    return MIST_OK;
}

int ora_init_vof(void* h, int shape, const double* args, int nargs) {
    Ctx& c = *(Ctx*)h;
    const Mesh& m = c.m;
    if ((shape == MIST_IC_CIRCLE && nargs < 3) || (shape != MIST_IC_CIRCLE && nargs < 4)) return MIST_ERR_PARAM;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i) {
                double v;
                if (shape == MIST_IC_CIRCLE)
                    v = vf_circle(m.x(i), m.x(i + 1), m.y(j), m.y(j + 1), args[0], args[1], args[2]);
                else
                    v = vf_sphere(m.x(i), m.x(i + 1), m.y(j), m.y(j + 1), m.z(k), m.z(k + 1), args[0], args[1],
                                  args[2], args[3], shape == MIST_IC_BUBBLE3D);
                c.VF(i, j, k) = v;
            }
    return MIST_OK;
}

int ora_neumann(void* h, int field) {
    Ctx& c = *(Ctx*)h;
    if (field == MIST_FIELD_BAND) { neumann_i16(c, c.band); return MIST_OK; }
    Arr3<double>* a = field_f64(c, field);
    if (!a) return MIST_ERR_PARAM;
    neumann(c, *a);
    return MIST_OK;
}

int ora_update_borders(void* h, int field) {
    Ctx& c = *(Ctx*)h;
    if (field == MIST_FIELD_BAND) { update_borders_i16(c, c.band); return MIST_OK; }
    Arr3<double>* a = field_f64(c, field);
    if (!a) return MIST_ERR_PARAM;
    update_borders(c, *a);
    return MIST_OK;
}

int ora_define_velocity(void* h, double t) { define_velocity(*(Ctx*)h, t); return MIST_OK; }
int ora_compute_dt(void* h, double* dt) { *dt = compute_dt(*(Ctx*)h); return MIST_OK; }
int ora_compute_band(void* h) { compute_band(*(Ctx*)h); return MIST_OK; }
int ora_compute_normal(void* h) { compute_normal(*(Ctx*)h); return MIST_OK; }
int ora_compute_plic(void* h) { compute_plic(*(Ctx*)h); return MIST_OK; }
// transport! (src/Transport.jl:2-268): with solveNS the momentum predictor too
int ora_transport(void* h, double dt, double /*t*/) {
    Ctx& c = *(Ctx*)h;
    return c.p.solveNS ? transport_ns(c, dt) : transport_vof(c, dt);
}
int ora_compute_props(void* h) { compute_props(*(Ctx*)h); return MIST_OK; }
int ora_pressure_rhs(void* h) { pressure_rhs(*(Ctx*)h); return MIST_OK; }
int ora_corrector(void* h, double dt) { corrector(*(Ctx*)h, dt); return MIST_OK; }
int ora_pressure_solve(void* h, double dt, double tol, int64_t precision, int64_t* cycles, double* resid) {
    return pressure_solve(*(Ctx*)h, dt, tol, (int)precision, cycles, resid);
}
// the FP64 face coefficients of multigrid level `level` (0 = fine) along `axis` that a
// pressure solve with this dt would build from the context's densities and VF, plain order
// (n[0] x n[1] x n[2] values, at most `cap`)
// The SL-SL pressure solve (mg_cycler / res_iteration); smooth_iters = res_iteration
// iterations over the solve.
int ora_pressure_solve_sl(void* h, double dt, double tol, int exact_restore, int64_t* cycles, double* resid,
                          int64_t* smooth_iters) {
    Ctx& c = *(Ctx*)h;
    if (c.p.xper || c.p.yper || c.p.zper) { c.err = "SL pressure: periodic axes are not supported"; return MIST_ERR_PARAM; }
    int st = pressure_solve_sl(c, dt, tol, c.p.mg_max_cycles, exact_restore != 0, cycles, resid, smooth_iters);
    if (st) c.err = "P-solve did not converge";
    return st;
}
void ora_set_sl_exact_restore(void* h, int on) { ((Ctx*)h)->sl_exact_restore = on != 0; }
void ora_set_sl_first_step_newton(void* h, int on) { ((Ctx*)h)->sl_first_step_newton = on != 0; }
// A! of the current P, uf, vf, wf, densities and band: LHS into divg, the corrected faces
// into gx, gy, gz (cell extents; any may be null).
int ora_sl_operator(void* h, double dt, double* gx, double* gy, double* gz) {
    Ctx& c = *(Ctx*)h;
    Arr3<double> a, b, d;
    sl_operator(c, dt, a, b, d);
    const size_t bytes = a.a.size() * sizeof(double);
    if (gx) std::memcpy(gx, a.a.data(), bytes);
    if (gy) std::memcpy(gy, b.a.data(), bytes);
    if (gz) std::memcpy(gz, d.a.data(), bytes);
    return MIST_OK;
}
int ora_sl_jacobian(void* h, double dt, int exact_restore, double* out) {
    Ctx& c = *(Ctx*)h;
    Arr3<double> J;
    sl_jacobian(c, dt, exact_restore != 0, J);
    std::memcpy(out, J.a.data(), J.a.size() * sizeof(double));
    return MIST_OK;
}
void ora_anderson_coefficients(const double* gram, int n, double* alpha) { anderson_coefficients(gram, n, alpha); }

int ora_mg_level_coef(void* h, double dt, int level, int axis, double* out, int64_t cap, int64_t n[3]) {
    std::vector<double> c;
    const int st = mg_level_coefficients(*(Ctx*)h, dt, level, axis, c, n);
    if (st) return st;
    if ((int64_t)c.size() > cap) return MIST_ERR_PARAM;
    std::memcpy(out, c.data(), c.size() * sizeof(double));
    return MIST_OK;
}
int64_t ora_max_cut_level(void* h) { return ((Ctx*)h)->max_cut_level; }

// run_solver initialisation of the VOF fields (src/NavierStokes_Parallel.jl:73-98)
int ora_vof_init(void* h) {
    Ctx& c = *(Ctx*)h;
    neumann(c, c.VF);
    update_borders(c, c.VF);
    compute_band(c);
    compute_normal(c);
    compute_plic(c);
    return MIST_OK;
}

// One run_solver step with solveNS = false (src/NavierStokes_Parallel.jl:125-146)
int ora_vof_step(void* h, double* t, double clamp_to, double* dt_used) {
    Ctx& c = *(Ctx*)h;
    double dt = compute_dt(c);
    if (clamp_to > 0.0 && *t + dt > clamp_to) dt = clamp_to - *t;
    *t += dt;
    define_velocity(c, *t);
    int st = transport_vof(c, dt);
    compute_band(c);
    *dt_used = dt;
    return st;
}

// ---- Navier-Stokes stages --------------------------------------------------------------
int ora_set_wall_velocity(void* h, int side, const double vel[3]) {
    Ctx& c = *(Ctx*)h;
    if (side < 0 || side > 5) return MIST_ERR_PARAM;
    for (int q = 0; q < 3; ++q) c.bc_wall[side][q] = vel[q];
    return MIST_OK;
}

static bool velocity_triple(Ctx& c, int fu, int fv, int fw, Arr3<double>** a) {
    const bool cell = fu == MIST_FIELD_U && fv == MIST_FIELD_V && fw == MIST_FIELD_W;
    const bool pred = fu == MIST_FIELD_US && fv == MIST_FIELD_VS && fw == MIST_FIELD_WS;
    if (!cell && !pred) return false;
    a[0] = field_f64(c, fu); a[1] = field_f64(c, fv); a[2] = field_f64(c, fw);
    return true;
}

int ora_apply_bc(void* h, int fu, int fv, int fw) {
    Ctx& c = *(Ctx*)h;
    Arr3<double>* a[3];
    if (!velocity_triple(c, fu, fv, fw, a)) return MIST_ERR_PARAM;
    apply_bc(c, *a[0], *a[1], *a[2]);
    return MIST_OK;
}

int ora_interpolate_face(void* h, int fu, int fv, int fw) {
    Ctx& c = *(Ctx*)h;
    Arr3<double>* a[3];
    if (!velocity_triple(c, fu, fv, fw, a)) return MIST_ERR_PARAM;
    interpolate_face(c, *a[0], *a[1], *a[2]);
    return MIST_OK;
}

int ora_interpolate_center(void* h) { interpolate_center(*(Ctx*)h); return MIST_OK; }
int ora_divergence(void* h, double dt) { divergence(*(Ctx*)h, dt); return MIST_OK; }
int ora_compute_curvature(void* h) { compute_curvature(*(Ctx*)h); return MIST_OK; }
int ora_surface_force(void* h) {
    Ctx& c = *(Ctx*)h;
    mask_maker(c);
    compute_sf(c);
    return MIST_OK;
}

// run_solver initialisation with solveNS (src/NavierStokes_Parallel.jl:73-105): Neumann!(VF),
// BC!(u,v,w), update_borders! of u, v, w and VF, interpolateFace!, compute_props!,
// computeBand!, computeNormal!, computePLIC!, and the divergence diagnostic at the first dt.
int ora_ns_init(void* h) {
    Ctx& c = *(Ctx*)h;
    neumann(c, c.VF);
    apply_bc(c, c.u, c.v, c.w);
    update_borders(c, c.u);
    update_borders(c, c.v);
    update_borders(c, c.w);
    update_borders(c, c.VF);
    interpolate_face(c, c.u, c.v, c.w);
    compute_props(c);
    compute_band(c);
    compute_normal(c);
    compute_plic(c);
    divergence(c, compute_dt(c));
    return MIST_OK;
}

// The restart branch of run_solver's initialisation (src/NavierStokes_Parallel.jl:59-67,
// :87-100) after fillArrays (src/ReadData.jl:3-53) filled VF, uf, vf, wf, P on the cell extents.
int ora_restart_init(void* h, double* dt) {
    Ctx& c = *(Ctx*)h;
    const Mesh& m = c.m;
    for (int64_t k = c.VF.e.lo[2]; k <= c.VF.hi(2); ++k)
        for (int64_t j = c.VF.e.lo[1]; j <= c.VF.hi(1); ++j)
            for (int64_t i = c.VF.e.lo[0]; i <= c.VF.hi(0); ++i)
                if (i < m.imino_ || i > m.imaxo_ || j < m.jmino_ || j > m.jmaxo_ || k < m.kmino_ || k > m.kmaxo_)
                    c.VF(i, j, k) = 0.0;
    interpolate_center(c);
    update_borders(c, c.u);
    update_borders(c, c.v);
    update_borders(c, c.w);
    compute_props(c);
    compute_band(c);
    compute_normal(c);
    compute_plic(c);
    const double d = compute_dt(c);
    divergence(c, d);
    if (dt) *dt = d;
    return MIST_OK;
}

int ora_ns_step(void* h, double* t, double clamp_to, double* dt_used, double tol, int64_t precision,
                int64_t* cycles, double* resid) {
    return ns_step(*(Ctx*)h, t, clamp_to, dt_used, tol, (int)precision, cycles, resid);
}

double ora_sl_flux_vol(void* h, int dir, int64_t i, int64_t j, int64_t k, double dt, double verts[24]) {
    double v[8][3];
    const double r = sl_flux_vol(*(Ctx*)h, dir, i, j, k, dt, v);
    std::memcpy(verts, v, sizeof(v));
    return r;
}

// the 17 tets of an FD-scheme band cell: its 5 back-traced tets, then the 12 correction tets
int ora_correction_tets(void* h, int64_t i, int64_t j, int64_t k, double dt, double out[17 * 12]) {
    const Ctx& c = *(const Ctx*)h;
    double tets[17][4][3] = {}, verts[8][3], t5[5][4][3];
    int64_t inds[17][4][3];
    cell2verts(c, i, j, k, verts);
    for (int n = 0; n < 8; ++n) project(c, verts[n], i, j, k, dt);
    verts2tets5(verts, t5);
    std::memcpy(tets, t5, sizeof(t5));
    const bool ok = add_correction_tets(c, i, j, k, dt, tets, inds);
    std::memcpy(out, tets, sizeof(tets));
    return ok ? MIST_OK : MIST_ERR_OVERFLOW;
}

// Reductions over interior cells in FP64 (sum in long double for a stable checker).
int ora_reduce(void* h, int kind, int field, double* out) {
    Ctx& c = *(Ctx*)h;
    const Mesh& m = c.m;
    Arr3<double>* a = field_f64(c, field);
    if (!a) return MIST_ERR_PARAM;
    long double s = 0.0L;
    double r = kind == MIST_REDUCE_MIN ? INFINITY : (kind == MIST_REDUCE_MAX ? -INFINITY : 0.0);
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                const double v = (*a)(i, j, k);
                if (kind == MIST_REDUCE_SUM) s += v;
                else if (kind == MIST_REDUCE_MAX) r = jmax(r, v);
                else if (kind == MIST_REDUCE_MIN) r = jmin(r, v);
                else r = jmax(r, std::fabs(v));
            }
    *out = kind == MIST_REDUCE_SUM ? (double)s : r;
    return MIST_OK;
}

// ---- single-cell primitives for the pins ---------------------------------------------
double ora_compute_alpha(double m0, double m1, double m2, double VF) { return compute_alpha(m0, m1, m2, VF); }
double ora_compute_dist(void* h, int64_t i, int64_t j, int64_t k, double nx, double ny, double nz, double VF) {
    return compute_dist(*(Ctx*)h, i, j, k, nx, ny, nz, VF);
}
double ora_plic2vf(void* h, int64_t i, int64_t j, int64_t k, double nx, double ny, double nz, double dist) {
    return plic2vf(*(Ctx*)h, i, j, k, nx, ny, nz, dist);
}
int ora_d2case(const double d[4], int inplace, double dout[4]) {
    double t[4] = {d[0], d[1], d[2], d[3]};
    int cs = d2case_inplace(t);
    if (dout) for (int n = 0; n < 4; ++n) dout[n] = inplace ? t[n] : d[n];
    return cs + 1;  // reference numbering
}
double ora_tet_vol(const double v[12]) {
    double t[4][3];
    std::memcpy(t, v, sizeof(t));
    return tet_vol(t);
}
double ora_cell2verts(void* h, int64_t i, int64_t j, int64_t k, double out[24]) {
    double v[8][3];
    const double s = cell2verts(*(Ctx*)h, i, j, k, v);
    std::memcpy(out, v, sizeof(v));
    return s;
}
void ora_project(void* h, double pt[3], int64_t i, int64_t j, int64_t k, double dt) { project(*(Ctx*)h, pt, i, j, k, dt); }
int ora_pt2index(void* h, const double pt[3], int64_t i, int64_t j, int64_t k, int64_t out[3]) {
    return pt2index(*(Ctx*)h, out, pt, i, j, k) ? 1 : 0;
}
void ora_vel_face(void* h, const double pt[3], int64_t i, int64_t j, int64_t k, double out[3]) {
    Ctx& c = *(Ctx*)h;
    out[0] = vel_uface(c, pt, i, j, k);
    out[1] = vel_vface(c, pt, i, j, k);
    out[2] = vel_wface(c, pt, i, j, k);
}
// cutTet of one tet with vertex indices: out = (vol, vLiq, vU, vV, vW, vrho); returns level.
int ora_cut_tet(void* h, const double tet_in[12], const int64_t ind_in[12], double out[6]) {
    Ctx& c = *(Ctx*)h;
    double tet[4][3];
    int64_t ind[4][3];
    std::memcpy(tet, tet_in, sizeof(tet));
    std::memcpy(ind, ind_in, sizeof(ind));
    std::unique_ptr<CutScratch, void (*)(CutScratch*)> s(new_cut_scratch(), delete_cut_scratch);
    CutSums r = cut_tet(c, tet, ind, false, false, false, 1, *s);
    out[0] = r.vol; out[1] = r.vLiq; out[2] = r.vU; out[3] = r.vV; out[4] = r.vW; out[5] = r.vrho;
    return r.lvl;
}
// Semi-Lagrangian volume of one band cell (src/Transport.jl:117-145): out = (vol, vLiq).
int ora_sl_cell(void* h, int64_t i, int64_t j, int64_t k, double dt, double out[2]) {
    Ctx& c = *(Ctx*)h;
    double verts[8][3], tets[5][4][3];
    int64_t vInds[8][3], inds[5][4][3];
    const double tetsign = cell2verts(c, i, j, k, verts);
    for (int n = 0; n < 8; ++n) project(c, verts[n], i, j, k, dt);
    verts2tets5(verts, tets);
    for (int n = 0; n < 8; ++n) pt2index(c, vInds[n], verts[n], i, j, k);
    inds2tets5(vInds, inds);
    std::unique_ptr<CutScratch, void (*)(CutScratch*)> s(new_cut_scratch(), delete_cut_scratch);
    double vol = 0.0, vLiq = 0.0;
    int lvl = 0;
    for (int t = 0; t < 5; ++t) {
        CutSums r = cut_tet(c, tets[t], inds[t], false, false, false, 1, *s);
        vol += tetsign * r.vol;
        vLiq += tetsign * r.vLiq;
        lvl = std::max(lvl, r.lvl);
    }
    out[0] = vol;
    out[1] = vLiq;
    return lvl;
}
int ora_elvira_cell(void* h, int64_t i, int64_t j, int64_t k, double out[3]) {
    int best = -1;
    out[0] = out[1] = out[2] = 0.0;
    elvira_calc(*(Ctx*)h, i, j, k, &out[0], &out[1], &out[2], &best);
    return best;
}

// Scores of the ELVIRA candidates of one cell (error, normal; INFINITY where a family is
// skipped), for the tie checks of the tests.
int ora_elvira_scores(void* h, int64_t i, int64_t j, int64_t k, double err[27], double nrm[27 * 3]) {
    int best = -1;
    double a = 0, b = 0, d = 0;
    double cn[27][3] = {};
    elvira_calc(*(Ctx*)h, i, j, k, &a, &b, &d, &best, err, cn);
    std::memcpy(nrm, cn, sizeof(cn));
    return best;
}

// Tie hints (kind 0: ELVIRA candidates, kind 1: height-function directions), one array of
// Nx*Ny*Nz entries (x fastest) per later compute_normal / compute_curvature call.
int ora_push_hints(void* h, int kind, const int32_t* hint, int64_t n) {
    Ctx& c = *(Ctx*)h;
    if (n != c.p.Nx * c.p.Ny * c.p.Nz || kind < 0 || kind > 1) { c.err = "bad hint array"; return MIST_ERR_PARAM; }
    (kind == 0 ? c.elv_hints : c.curv_hints).emplace_back(hint, hint + n);
    return MIST_OK;
}
// The refused hints recorded so far (at most 64): per entry kind, i, j, k, hint, own choice
// (ints) and the two scores.
int64_t ora_tie_refusals(void* h, int64_t* ints, double* scores, int64_t cap) {
    const Ctx& c = *(Ctx*)h;
    const int64_t n = std::min<int64_t>(cap, (int64_t)c.refusals.size());
    for (int64_t q = 0; q < n; ++q) {
        const Ctx::Refusal& r = c.refusals[(size_t)q];
        const int64_t v[6] = {r.kind, r.i, r.j, r.k, r.hint, r.own};
        for (int t = 0; t < 6; ++t) ints[6 * q + t] = v[t];
        scores[2 * q] = r.e_hint;
        scores[2 * q + 1] = r.e_own;
    }
    return n;
}
void ora_clear_hints(void* h) {
    Ctx& c = *(Ctx*)h;
    c.refusals.clear();
    c.elv_hints.clear();
    c.curv_hints.clear();
    std::memset(c.tie_stats, 0, sizeof(c.tie_stats));
    c.tie_gap[0] = c.tie_gap[1] = 0.0;
}
// {cells whose hint differed, adopted, refused, pending hint arrays} and the largest adopted
// relative score gap (ELVIRA only).
int ora_tie_stats(void* h, int kind, int64_t out[4], double* gap) {
    Ctx& c = *(Ctx*)h;
    if (kind < 0 || kind > 1) return MIST_ERR_PARAM;
    for (int q = 0; q < 3; ++q) out[q] = c.tie_stats[kind][q];
    out[3] = (int64_t)(kind == 0 ? c.elv_hints.size() : c.curv_hints.size());
    *gap = c.tie_gap[kind];
    return MIST_OK;
}

void ora_set_count_flops(int on) { g_count_flops = on != 0; }
void ora_get_flops(uint64_t out[2]) { out[0] = t_flops.flops; out[1] = t_flops.special; }
void ora_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int ora_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
