// oracle_stages.cpp -- TEST INFRASTRUCTURE ONLY (the oracle): ghost fills, band, normals,
// PLIC, SL-VOF transport, prescribed velocity, time step, initial conditions, densities.
#include <algorithm>
#include <cstring>
#include <memory>

#include "oracle_core.h"
#include "oracle_fill.h"

namespace ora {

CutScratch* new_cut_scratch();
void delete_cut_scratch(CutScratch* s);

Arr3<int16_t>* field_i16(Ctx& c, int f) {
    if (f == MIST_FIELD_BAND) return &c.band;
    if (f == MIST_FIELD_MASK) return &c.mask;
    return nullptr;
}

Arr3<double>* field_f64(Ctx& c, int f) {
    switch (f) {
        case MIST_FIELD_VOF: return &c.VF;
        case MIST_FIELD_NX: return &c.nx;
        case MIST_FIELD_NY: return &c.ny;
        case MIST_FIELD_NZ: return &c.nz;
        case MIST_FIELD_PLIC_D: return &c.D;
        case MIST_FIELD_U: return &c.u;
        case MIST_FIELD_V: return &c.v;
        case MIST_FIELD_W: return &c.w;
        case MIST_FIELD_UF: return &c.uf;
        case MIST_FIELD_VF: return &c.vf;
        case MIST_FIELD_WF: return &c.wf;
        case MIST_FIELD_P: return &c.P;
        case MIST_FIELD_DIVG: return &c.divg;
        case MIST_FIELD_DENX: return &c.denx;
        case MIST_FIELD_DENY: return &c.deny;
        case MIST_FIELD_DENZ: return &c.denz;
        case MIST_FIELD_RHS: return &c.RHS;
        case MIST_FIELD_US: return &c.us;
        case MIST_FIELD_VS: return &c.vs;
        case MIST_FIELD_WS: return &c.ws;
        case MIST_FIELD_VISCX: return &c.viscx;
        case MIST_FIELD_VISCY: return &c.viscy;
        case MIST_FIELD_VISCZ: return &c.viscz;
        case MIST_FIELD_CURV: return &c.Curve;
        case MIST_FIELD_SFX: return &c.sfx;
        case MIST_FIELD_SFY: return &c.sfy;
        case MIST_FIELD_SFZ: return &c.sfz;
        default: return nullptr;
    }
}

void neumann(Ctx& c, Arr3<double>& A) { neumann_t(c.m, A); }
void neumann_i16(Ctx& c, Arr3<int16_t>& A) { neumann_t(c.m, A); }

void update_borders(Ctx& c, Arr3<double>& A) { borders_t(c, A); }
void update_borders_i16(Ctx& c, Arr3<int16_t>& A) { borders_t(c, A); }

// ------------------------------------------------------------------------------------
// computeBand! (src/VF.jl:16-74)
// ------------------------------------------------------------------------------------
void compute_band(Ctx& c) {
    const Mesh& m = c.m;
    const double VFlo = c.p.VFlo, VFhi = c.p.VFhi;
    const int64_t nband = c.p.nband;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                const double vf = c.VF(i, j, k);
                int16_t b;
                if (vf < VFlo) {
                    b = (int16_t)(-nband - 1);
                    for (int64_t kk = k - 1; kk <= k + 1 && b != 0; ++kk)
                        for (int64_t jj = j - 1; jj <= j + 1 && b != 0; ++jj)
                            for (int64_t ii = i - 1; ii <= i + 1; ++ii)
                                if (c.VF(ii, jj, kk) >= VFhi) { b = 0; break; }
                } else if (vf > VFhi) {
                    b = (int16_t)(nband + 1);
                    for (int64_t kk = k - 1; kk <= k + 1 && b != 0; ++kk)
                        for (int64_t jj = j - 1; jj <= j + 1 && b != 0; ++jj)
                            for (int64_t ii = i - 1; ii <= i + 1; ++ii)
                                if (c.VF(ii, jj, kk) <= VFlo) { b = 0; break; }
                } else {
                    b = 0;
                }
                c.band(i, j, k) = b;
            }
    neumann_i16(c, c.band);
    update_borders_i16(c, c.band);
    // A sweep only lowers |band| from >n to n and only tests neighbours against n-1, so a
    // snapshot read gives the same result as the reference's in-place serial sweep.
    for (int64_t n = 1; n <= nband; ++n) {
        const Arr3<int16_t> snap = c.band;
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
            for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
                for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                    const int16_t b = snap(i, j, k);
                    if (std::abs((int)b) > n) {
                        bool hit = false;
                        for (int64_t kk = k - 1; kk <= k + 1; ++kk)
                            for (int64_t jj = j - 1; jj <= j + 1; ++jj)
                                for (int64_t ii = i - 1; ii <= i + 1; ++ii)
                                    if (std::abs((int)snap(ii, jj, kk)) == n - 1) hit = true;
                        if (hit) c.band(i, j, k) = (int16_t)(n * (b > 0 ? 1 : (b < 0 ? -1 : 0)));
                    }
                }
        update_borders_i16(c, c.band);
    }
}

// ------------------------------------------------------------------------------------
// ELVIRA! (src/ELVIRA.jl:6-55)
// ------------------------------------------------------------------------------------
// A hinted ELVIRA candidate is a tie of the oracle's own choice when its error exceeds the
// smallest one by no more than rounding: 1e-11 relative (the ties of the stated sphere and
// bubble differ by at most 8e-13 relative; a genuinely worse candidate differs at O(1)
// relative), plus 1e-30 absolute for planar interfaces whose errors are rounding
// noise themselves. The errors of mirrored candidates on a symmetric field are equal in exact
// arithmetic and differ in their last bits between implementations (libm cbrt/acos/cos in
// computeDist, the association of the transport's volume sums).
static bool elvira_tie(double e_hint, double e_best) {
    if (!(e_hint == e_hint) || !(e_hint < INFINITY)) return false;
    return e_hint <= e_best * (1.0 + 1e-11) + 1e-30;
}

void compute_normal(Ctx& c) {
    const Mesh& m = c.m;
    std::fill(c.nx.a.begin(), c.nx.a.end(), 0.0);
    std::fill(c.ny.a.begin(), c.ny.a.end(), 0.0);
    std::fill(c.nz.a.begin(), c.nz.a.end(), 0.0);
    const double VFlo = c.p.VFlo, VFhi = c.p.VFhi;
    std::vector<int32_t> hint;
    if (!c.elv_hints.empty()) {
        hint = std::move(c.elv_hints.front());
        c.elv_hints.pop_front();
    }
    int64_t h_diff = 0, h_adopt = 0, h_refuse = 0;
    double h_gap = 0.0;
    const FlopCount before = t_flops;  // the operation counter of the calling thread
    uint64_t fl = 0, fs = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 1) reduction(+ : fl, fs, h_diff, h_adopt, h_refuse) \
    reduction(max : h_gap)
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j) {
            t_flops = FlopCount{};
            for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                const double vf = c.VF(i, j, k);
                if (vf >= VFlo && vf <= VFhi) {
                    double a = 0, b = 0, d = 0;
                    if (hint.empty()) {
                        elvira_calc(c, i, j, k, &a, &b, &d, nullptr);
                    } else {
                        int best = -1;
                        double err[27], cn[27][3];
                        elvira_calc(c, i, j, k, &a, &b, &d, &best, err, cn);
                        const int h = hint[(size_t)((i - 1) + m.Nx * ((j - 1) + m.Ny * (k - 1)))];
                        if (h != best) {
                            ++h_diff;
                            if (h >= 0 && h < 27 && best >= 0 && elvira_tie(err[h], err[best])) {
                                ++h_adopt;
                                h_gap = std::max(h_gap, (err[h] - err[best]) / std::max(err[best], 1e-300));
                                a = cn[h][0]; b = cn[h][1]; d = cn[h][2];
                            } else {
                                ++h_refuse;
#pragma omp critical(ora_refusals)
                                if (c.refusals.size() < 64)
                                    c.refusals.push_back({0, i, j, k, h, best, (h >= 0 && h < 27) ? err[h] : NAN,
                                                          best >= 0 ? err[best] : NAN});
                            }
                        }
                    }
                    c.nx(i, j, k) = a; c.ny(i, j, k) = b; c.nz(i, j, k) = d;
                } else if (vf < VFlo) {
                    if (c.VF(i - 1, j, k) > VFhi) c.nx(i, j, k) = 1;
                    else if (c.VF(i + 1, j, k) > VFhi) c.nx(i, j, k) = -1;
                    else if (c.VF(i, j - 1, k) > VFhi) c.ny(i, j, k) = 1;
                    else if (c.VF(i, j + 1, k) > VFhi) c.ny(i, j, k) = -1;
                    else if (c.VF(i, j, k - 1) > VFhi) c.nz(i, j, k) = 1;
                    else if (c.VF(i, j, k + 1) > VFhi) c.nz(i, j, k) = -1;
                } else if (vf > VFhi) {
                    if (c.VF(i - 1, j, k) < VFlo) c.nx(i, j, k) = -1;
                    else if (c.VF(i + 1, j, k) < VFlo) c.nx(i, j, k) = 1;
                    else if (c.VF(i, j - 1, k) < VFlo) c.ny(i, j, k) = -1;
                    else if (c.VF(i, j + 1, k) < VFlo) c.ny(i, j, k) = 1;
                    else if (c.VF(i, j, k - 1) < VFlo) c.nz(i, j, k) = -1;
                    else if (c.VF(i, j, k + 1) < VFlo) c.nz(i, j, k) = 1;
                }
            }
            fl += t_flops.flops;
            fs += t_flops.special;
        }
    t_flops.flops = before.flops + fl;
    t_flops.special = before.special + fs;
    c.tie_stats[0][0] += h_diff;
    c.tie_stats[0][1] += h_adopt;
    c.tie_stats[0][2] += h_refuse;
    c.tie_gap[0] = std::max(c.tie_gap[0], h_gap);
    update_borders(c, c.nx);
    update_borders(c, c.ny);
    update_borders(c, c.nz);
}

// computePLIC! (src/VF.jl:94-103): every cell including ghosts, then borders.
void compute_plic(Ctx& c) {
    const Mesh& m = c.m;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i)
                c.D(i, j, k) = compute_dist(c, i, j, k, c.nx(i, j, k), c.ny(i, j, k), c.nz(i, j, k), c.VF(i, j, k));
    update_borders(c, c.D);
}

// ------------------------------------------------------------------------------------
// transport!, VOF part (src/Transport.jl:16-20, 108-153, 178-184, 259-260)
// ------------------------------------------------------------------------------------
int transport_vof(Ctx& c, double dt) {
    const Mesh& m = c.m;
    compute_normal(c);
    c.status = 0;
    Arr3<double> VFnew;
    VFnew.init(m.cell_ext(), 0.0);
    int64_t maxlvl = 0;
    const FlopCount before = t_flops;  // the operation counter of the calling thread
    uint64_t fl = 0, fs = 0;
#pragma omp parallel reduction(max : maxlvl) reduction(+ : fl, fs)
    {
        std::unique_ptr<CutScratch, void (*)(CutScratch*)> s(new_cut_scratch(), delete_cut_scratch);
        t_flops = FlopCount{};
#pragma omp for collapse(2) schedule(dynamic, 1)
        for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
            for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
                for (int64_t i = m.imin_; i <= m.imax_; ++i) {
                    if (std::abs((int)c.band(i, j, k)) <= c.p.nband) {
                        double verts[8][3], tets[5][4][3];
                        int64_t vInds[8][3], inds[5][4][3];
                        const double tetsign = cell2verts(c, i, j, k, verts);
                        for (int n = 0; n < 8; ++n) project(c, verts[n], i, j, k, dt);
                        verts2tets5(verts, tets);
                        for (int n = 0; n < 8; ++n)
                            if (!pt2index(c, vInds[n], verts[n], i, j, k)) c.status = MIST_ERR_OVERFLOW;
                        inds2tets5(vInds, inds);
                        double vol = 0.0, vLiq = 0.0;
                        for (int t = 0; t < 5; ++t) {
                            CutSums r = cut_tet(c, tets[t], inds[t], false, false, false, 1, *s);
                            vol += tetsign * r.vol;
                            vLiq += tetsign * r.vLiq;
                            maxlvl = std::max<int64_t>(maxlvl, r.lvl);
                        }
                        VFnew(i, j, k) = vLiq / vol;
                    } else {
                        VFnew(i, j, k) = c.VF(i, j, k);
                    }
                }
        fl += t_flops.flops;
        fs += t_flops.special;
    }
    t_flops.flops = before.flops + fl;
    t_flops.special = before.special + fs;
    c.max_cut_level = maxlvl;
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i) c.VF(i, j, k) = VFnew(i, j, k);
    neumann(c, c.VF);
    update_borders(c, c.VF);  // update_VF_borders! behaves as update_borders!
    compute_normal(c);
    compute_plic(c);
    return c.status;
}

// ------------------------------------------------------------------------------------
// defineVelocity! (src/Tools.jl:648-693) with the period of cos(pi t / T) as a parameter
// (the reference hard-codes T = 2).
// ------------------------------------------------------------------------------------
static void vel_fun(int kind, double T, double x, double y, double z, double t, double* u, double* v, double* w) {
    const double pi = M_PI, twopi = 2.0 * M_PI;
    switch (kind) {
        case MIST_VEL_DEFORMATION: {
            const double sx = std::sin(pi * x), sy = std::sin(pi * y);
            *u = ((-2.0 * (sx * sx) * std::sin(pi * y)) * std::cos(pi * y)) * std::cos((pi * t) / T);
            *v = ((2.0 * (sy * sy) * std::sin(pi * x)) * std::cos(pi * x)) * std::cos((pi * t) / T);
            *w = 0.0;
            break;
        }
        case MIST_VEL_DEFORMATION3D: {
            const double sx = std::sin(pi * x), sy = std::sin(pi * y), sz = std::sin(pi * z);
            *u = ((2.0 * (sx * sx) * std::sin(twopi * y)) * std::sin(twopi * z)) * std::cos((pi * t) / T);
            *v = ((-(sy * sy) * std::sin(twopi * x)) * std::sin(twopi * z)) * std::cos((pi * t) / T);
            *w = ((-(sz * sz) * std::sin(twopi * x)) * std::sin(twopi * y)) * std::cos((pi * t) / T);
            break;
        }
        case MIST_VEL_ROTATION:
            *u = twopi * (0.5 - y); *v = twopi * (x - 0.5); *w = 0.0;
            break;
        case MIST_VEL_DIVFLOW:
            *u = 2.0 * x; *v = -1.0 * (2.5 - y); *w = 0.0;
            break;
        case MIST_VEL_DIVFLOW3D:
            *u = 200.0 * x; *v = -100.0 * (5.0 - y); *w = -50.0 * (5.0 - z);
            break;
        case MIST_VEL_AIRBLAST:
            *u = 50.0 * (1.0 - std::exp(-100.0 * std::fabs(y - 0.5))); *v = 0.0; *w = 0.0;
            break;
        default:
            *u = *v = *w = 0.0;
    }
}

void define_velocity(Ctx& c, double t) {
    const Mesh& m = c.m;
    const int kind = (int)c.p.vf_velocity;
    const double T = c.p.vel_period;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i) {
                double a, b, d;
                vel_fun(kind, T, m.xm(i), m.ym(j), m.zm(k), t, &a, &b, &d);
                c.u(i, j, k) = a; c.v(i, j, k) = b; c.w(i, j, k) = d;
                vel_fun(kind, T, m.x(i), m.ym(j), m.zm(k), t, &a, &b, &d);
                c.uf(i, j, k) = a;
                vel_fun(kind, T, m.xm(i), m.y(j), m.zm(k), t, &a, &b, &d);
                c.vf(i, j, k) = b;
                vel_fun(kind, T, m.xm(i), m.ym(j), m.z(k), t, &a, &b, &d);
                c.wf(i, j, k) = d;
            }
}

// compute_dt (src/Tools.jl:317-337); maxima over the arrays including ghosts.
double compute_dt(const Ctx& c) {
    double mu = 0, mv = 0, mw = 0;
    for (double a : c.u.a) mu = jmax(mu, std::fabs(a));
    for (double a : c.v.a) mv = jmax(mv, std::fabs(a));
    for (double a : c.w.a) mw = jmax(mw, std::fabs(a));
    const Mesh& m = c.m;
    const double convec_dt = jmin(jmin(m.dx / mu, m.dy / mv), m.dz / mw);
    const double viscous_dt = jmin(jmin(m.dx, m.dy), m.dz) / jmax(c.p.mu_liq, c.p.mu_gas);
    return jmin(c.p.max_dt, c.p.CFL * jmin(convec_dt, viscous_dt));
}

// ------------------------------------------------------------------------------------
// Initial conditions
// ------------------------------------------------------------------------------------
// VFcircle (src/Tools.jl:790-902): exact area fraction of a circle in a rectangle.
double vf_circle(double xmin, double xmax, double ymin, double ymax, double rad, double xo, double yo) {
    const double dx = jmax(xmax - xmin, ymax - ymin);
    const double xm = 0.5 * (xmin + xmax), ym = 0.5 * (ymin + ymax);
    const double G = rad - std::sqrt((xm - xo) * (xm - xo) + (ym - yo) * (ym - yo));
    if (G > 2.0 * dx) return 1.0;
    if (G < -2.0 * dx) return 0.0;
    double VF = 0.0;
    bool xdone = false;
    int xcycl = 1;
    const double r2 = rad * rad;
    while (!xdone) {
        double x_min, x_max;
        if ((xmax - xo) * (xmin - xo) < 0.0) {
            if (xcycl == 1) { x_min = 0.0; x_max = std::fabs(xmin - xo); xcycl = 2; }
            else { x_min = 0.0; x_max = std::fabs(xmax - xo); xdone = true; }
        } else {
            x_min = jmin(std::fabs(xmin - xo), std::fabs(xmax - xo));
            x_max = jmax(std::fabs(xmin - xo), std::fabs(xmax - xo));
            xdone = true;
        }
        bool ydone = false;
        int ycycl = 1;
        while (!ydone) {
            double y_min, y_max;
            if ((ymin - yo) * (ymax - yo) < 0.0) {
                if (ycycl == 1) { y_min = 0.0; y_max = std::fabs(ymin - yo); ycycl = 2; }
                else { y_min = 0.0; y_max = std::fabs(ymax - yo); ydone = true; }
            } else {
                y_min = jmin(std::fabs(ymin - yo), std::fabs(ymax - yo));
                y_max = jmax(std::fabs(ymin - yo), std::fabs(ymax - yo));
                ydone = true;
            }
            double x_1, x_2;
            if (r2 - x_min * x_min > 0.0) {
                const double yint = std::sqrt(r2 - x_min * x_min);
                if (yint < y_min) { VF += 0.0; continue; }
                else if (yint < y_max) x_1 = x_min;
                else x_1 = std::sqrt(r2 - y_max * y_max);
            } else {
                VF += 0.0;
                continue;
            }
            if (r2 - x_max * x_max > 0.0) {
                const double yint = std::sqrt(r2 - x_max * x_max);
                if (yint > y_max) {
                    VF += ((x_max - x_min) * (y_max - y_min)) / ((xmax - xmin) * (ymax - ymin));
                    continue;
                } else if (yint > y_min) x_2 = x_max;
                else x_2 = std::sqrt(r2 - y_min * y_min);
            } else {
                x_2 = std::sqrt(r2 - y_min * y_min);
            }
            VF += (((((x_1 - x_min) * (y_max - y_min) + (r2 * std::asin(x_2 / rad)) / 2.0) -
                     (r2 * std::asin(x_1 / rad)) / 2.0) -
                    (x_1 * std::pow(r2 - x_1 * x_1, 0.5)) / 2.0) +
                   (x_2 * std::pow(r2 - x_2 * x_2, 0.5)) / 2.0 - (y_min * (x_2 - x_1))) /
                  ((xmax - xmin) * (ymax - ymin));
        }
    }
    return VF;
}

// VFsphere / VFbubble3d (src/Tools.jl:907-922, :1061-1076): 20^3 interior samples. The
// result only depends on the number n of samples inside, through the sequential sum of
// n copies of 1/8000 (or 1 minus it), which is tabulated; cells whose box is clearly
// inside or outside the sphere (margin far above rounding) skip the sample loop.
static const std::vector<double>& sample_table(bool bubble) {
    struct Tables {
        std::vector<double> S, B;
        Tables() : S(8001), B(8001) {
            const double sub = 1.0 / 8000.0;
            S[0] = 0.0; B[0] = 1.0;
            for (int n = 1; n <= 8000; ++n) { S[n] = S[n - 1] + sub; B[n] = B[n - 1] - sub; }
        }
    };
    static const Tables tabs;  // initialised once, thread-safe
    return bubble ? tabs.B : tabs.S;
}

double vf_sphere(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
                 double rad, double xo, double yo, double zo, bool bubble) {
    const std::vector<double>& tab = sample_table(bubble);
    auto clampd = [](double p, double a, double b) { return p < a ? a - p : (p > b ? p - b : 0.0); };
    const double ex = clampd(xo, xmin, xmax), ey = clampd(yo, ymin, ymax), ez = clampd(zo, zmin, zmax);
    const double dmin2 = ex * ex + ey * ey + ez * ez;
    const double fx = jmax(std::fabs(xmin - xo), std::fabs(xmax - xo));
    const double fy = jmax(std::fabs(ymin - yo), std::fabs(ymax - yo));
    const double fz = jmax(std::fabs(zmin - zo), std::fabs(zmax - zo));
    const double dmax2 = fx * fx + fy * fy + fz * fz;
    const double r2 = rad * rad;
    if (dmin2 > r2 * (1.0 + 1e-9)) return tab[0];
    if (dmax2 < r2 * (1.0 - 1e-9)) return tab[8000];
    const int nF = 20;
    int cnt = 0;
    for (int k = 1; k <= nF; ++k)
        for (int j = 1; j <= nF; ++j)
            for (int i = 1; i <= nF; ++i) {
                const double xh = xmin + ((double)i / (double)(nF + 1)) * (xmax - xmin);
                const double yh = ymin + ((double)j / (double)(nF + 1)) * (ymax - ymin);
                const double zh = zmin + ((double)k / (double)(nF + 1)) * (zmax - zmin);
                const double G = r2 - (((xh - xo) * (xh - xo) + (yh - yo) * (yh - yo)) + (zh - zo) * (zh - zo));
                cnt += G > 0.0;
            }
    return tab[cnt];
}

// compute_dens! (src/Tools.jl:1212-1232)
void compute_dens(Ctx& c) {
    const Mesh& m = c.m;
    const double rl = c.p.rho_liq, rg = c.p.rho_gas;
    for (int64_t k = m.kmino_; k <= m.kmaxo_; ++k)
        for (int64_t j = m.jmino_; j <= m.jmaxo_; ++j)
            for (int64_t i = m.imino_; i <= m.imaxo_; ++i) {
                const double vx = jmin(1.0, jmax(0.0, (c.VF(i, j, k) + c.VF(i - 1, j, k)) / 2.0));
                c.denx(i, j, k) = rl * vx + rg * (1.0 - vx);
                const double vy = jmin(1.0, jmax(0.0, (c.VF(i, j, k) + c.VF(i, j - 1, k)) / 2.0));
                c.deny(i, j, k) = rl * vy + rg * (1.0 - vy);
                const double vz = jmin(1.0, jmax(0.0, (c.VF(i, j, k) + c.VF(i, j, k - 1)) / 2.0));
                c.denz(i, j, k) = rl * vz + rg * (1.0 - vz);
            }
}

// FD right-hand side div(u*) (src/Pressure.jl:16-22)
void pressure_rhs(Ctx& c) {
    const Mesh& m = c.m;
    std::fill(c.RHS.a.begin(), c.RHS.a.end(), 0.0);
    for (int64_t k = m.kmin_; k <= m.kmax_; ++k)
        for (int64_t j = m.jmin_; j <= m.jmax_; ++j)
            for (int64_t i = m.imin_; i <= m.imax_; ++i)
                c.RHS(i, j, k) = ((c.uf(i + 1, j, k) - c.uf(i, j, k)) / m.dx + (c.vf(i, j + 1, k) - c.vf(i, j, k)) / m.dy) +
                                 (c.wf(i, j, k + 1) - c.wf(i, j, k)) / m.dz;
}

// corrector! (src/Velocity.jl:3-20)
void corrector(Ctx& c, double dt) {
    const Mesh& m = c.m;
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 2; ++i)
                c.uf(i, j, k) = c.uf(i, j, k) - dt / c.denx(i, j, k) * ((c.P(i, j, k) - c.P(i - 1, j, k)) / m.dx);
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 1; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 2; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 1; ++i)
                c.vf(i, j, k) = c.vf(i, j, k) - dt / c.deny(i, j, k) * ((c.P(i, j, k) - c.P(i, j - 1, k)) / m.dy);
    for (int64_t k = m.kmin_ - 1; k <= m.kmax_ + 2; ++k)
        for (int64_t j = m.jmin_ - 1; j <= m.jmax_ + 1; ++j)
            for (int64_t i = m.imin_ - 1; i <= m.imax_ + 1; ++i)
                c.wf(i, j, k) = c.wf(i, j, k) - dt / c.denz(i, j, k) * ((c.P(i, j, k) - c.P(i, j, k - 1)) / m.dz);
}

}  // namespace ora
