// adjoint_harness.hip -- TEST INFRASTRUCTURE: host entry points over the library's adjoint
// pieces (ns_multiphase_amd/csrc/mist_adjoint.h) and the forward functions they reverse,
// so the tests can compare each adjoint with central differences of its forward function on
// the CPU; and over the two forms of the leaf cut (the fourteen case bodies of cut_dispatch and
// the two of cut_canonical), forward and reverse, so the tests can compare them. Built by
// tests/test_adjoint_pieces.py with hipcc; never part of the library.
#include <cmath>
#include <vector>

#include "../ns_multiphase_amd/csrc/mist_adjoint.h"
#include "../ns_multiphase_amd/csrc/mist_kernels.h"

using namespace mist;

namespace {

struct HostMesh {
    std::vector<double> xs[6];  // x, y, z faces; xm, ym, zm centres
    DevMesh m{};
    DevParams p{};
};
HostMesh g_mesh;

void axis(int64_t N, double L, int ng, std::vector<double>& fc, std::vector<double>& cc, double& h) {
    fc.assign((size_t)(N + 2 * ng + 1), 0.0);
    for (int64_t q = 0; q <= N; ++q) fc[(size_t)(ng + q)] = ((double)q * L) / (double)N;
    h = fc[(size_t)(ng + 1)] - fc[(size_t)ng];
    for (int n = 1; n <= ng; ++n) {
        fc[(size_t)(ng - n)] = fc[(size_t)ng] - (double)n * h;
        fc[(size_t)(ng + N + n)] = fc[(size_t)(ng + N)] + (double)n * h;
    }
    cc.assign((size_t)(N + 2 * ng), 0.0);
    for (size_t q = 0; q < cc.size(); ++q) cc[q] = 0.5 * (fc[q] + fc[q + 1]);
}

// leaf forward: the liquid and total volumes of cut_leaf in mist_kernels.h
void leaf_fwd(const double x[4][3], double n0, double n1, double n2, double Dc, double& L, double& V) {
    double d[4], dd[4];
    for (int v = 0; v < 4; ++v) {
        d[v] = ((n0 * x[v][0] + n1 * x[v][1]) + n2 * x[v][2]) - Dc;
        dd[v] = d[v];
    }
    const int cs = d2case_snap(dd);
    double Q[8][3];
    for (int v = 0; v < 4; ++v)
        for (int q = 0; q < 3; ++q) Q[v][q] = x[v][q];
    for (int e = 0; e < CutTab::nvert[cs]; ++e) {
        const int a = CutTab::edge[cs][e] & 15, b = CutTab::edge[cs][e] >> 4;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-d[a], d[b] - d[a])));
        for (int q = 0; q < 3; ++q) Q[4 + e][q] = (1.0 - mu) * Q[a][q] + mu * Q[b][q];
    }
    L = 0.0;
    V = 0.0;
    for (int n = CutTab::ntets[cs]; n >= CutTab::nntet[cs]; --n) {
        const double tv = tet_vol(Q[vtet_of(cs, n - 1, 0)], Q[vtet_of(cs, n - 1, 1)], Q[vtet_of(cs, n - 1, 2)],
                                  Q[vtet_of(cs, n - 1, 3)]);
        L += tv;
        V += tv;
    }
    for (int n = 1; n <= CutTab::nntet[cs] - 1; ++n)
        V += tet_vol(Q[vtet_of(cs, n - 1, 0)], Q[vtet_of(cs, n - 1, 1)], Q[vtet_of(cs, n - 1, 2)],
                     Q[vtet_of(cs, n - 1, 3)]);
}

}  // namespace

extern "C" {

void h_mesh(const int64_t N[3], const double L[3], int ng, double VFlo) {
    HostMesh& H = g_mesh;
    double h[3];
    for (int a = 0; a < 3; ++a) axis(N[a], L[a], ng, H.xs[a], H.xs[3 + a], h[a]);
    DevMesh& m = H.m;
    m.Nx = (int)N[0]; m.Ny = (int)N[1]; m.Nz = (int)N[2]; m.ng = ng;
    m.cnx = m.Nx + 2 * ng; m.cny = m.Ny + 2 * ng; m.cnz = m.Nz + 2 * ng;
    m.vnx = m.cnx + 6; m.vny = m.cny + 6; m.vnz = m.cnz + 6;
    m.dx = h[0]; m.dy = h[1]; m.dz = h[2];
    m.Lx = L[0]; m.Ly = L[1]; m.Lz = L[2];
    const int64_t imino = 1 - ng;
    m.x = H.xs[0].data() - imino; m.y = H.xs[1].data() - imino; m.z = H.xs[2].data() - imino;
    m.xm = H.xs[3].data() - imino; m.ym = H.xs[4].data() - imino; m.zm = H.xs[5].data() - imino;
    H.p.VFlo = VFlo;
    H.p.VFhi = 1.0 - VFlo;
    H.p.nband = 1;
    H.p.projection = MIST_PROJ_RK4;
    H.p.zalesak = 0;
}

void h_leaf(const double* x, const double* n, double D, double* out) {
    leaf_fwd((const double(*)[3])x, n[0], n[1], n[2], D, out[0], out[1]);
}

void h_leaf_adj(const double* x, const double* n, double D, double gL, double gV, double* gx, double* gn, double* gD) {
    double g4[4][3] = {}, gnn[3] = {0.0, 0.0, 0.0}, gd = 0.0;
    leaf_adj((const double(*)[3])x, n[0], n[1], n[2], D, gL, gV, g4, gnn, gd);
    for (int v = 0; v < 4; ++v)
        for (int q = 0; q < 3; ++q) gx[3 * v + q] = g4[v][q];
    for (int q = 0; q < 3; ++q) gn[q] = gnn[q];
    *gD = gd;
}

// liquid and total volume of tet x cut by n.x = D: out[0..1] by cut_dispatch, out[2..3] by
// cut_canonical, out[4] the snapped case, out[5] the liquid volume by cut_canonical_liquid,
// out[6] tet_liquid_fraction times the tet's volume and out[7] cut_dispatch's liquid volume,
// both on the snapped distances (out holds 8 doubles)
void h_cut_compare(const double* x, const double* n, double D, double* out) {
    V3 P[4];
    double d[4], dd[4];
    for (int v = 0; v < 4; ++v) {
        P[v] = V3{x[3 * v], x[3 * v + 1], x[3 * v + 2]};
        d[v] = ((n[0] * P[v].x + n[1] * P[v].y) + n[2] * P[v].z) - D;
        dd[v] = d[v];
    }
    const int cs = d2case_snap(dd);
    double L = 0.0, G = 0.0, Lc = 0.0, Gc = 0.0;
    cut_dispatch<true>(cs, P, d, [&](double v) { L += v; }, [&](double v) { G += v; });
    cut_canonical(cs, P, d, [&](double v) { Lc += v; }, [&](double v) { Gc += v; });
    out[0] = L; out[1] = L + G; out[2] = Lc; out[3] = Lc + Gc; out[4] = cs;
    out[5] = cut_canonical_liquid(cs, P, d);
    const double tv = tet_vol3(P[0], P[1], P[2], P[3]);
    out[6] = cs == 0 ? tv : (cs == 15 ? 0.0 : tv * tet_liquid_fraction(cs, dd));
    double Ls = 0.0;
    cut_dispatch<false>(cs, P, dd, [&](double v) { Ls += v; }, [](double) {});
    out[7] = Ls;
}

// computePLIC2VF of the cell [x0,x1]x[y0,y1]x[z0,z1] (corners numbered x fastest, x mirrored
// in odd cells, the reference's 5-tet decomposition) for the plane n.x = D: out[0] with the
// reference's fourteen case bodies on the snapped distances, out[1] as the ELVIRA scoring
// kernel forms it, each cut tet's liquid share (tet_liquid_fraction) times its whole term
void h_cell_vf_compare(const double* box, const double* n, double D, int even, double* out) {
    const double cellvol = ((box[1] - box[0]) * (box[3] - box[2])) * (box[5] - box[4]);
    const double ts = even ? 1.0 : -1.0;
    double ref = 0.0, frac = 0.0;
    for (int t = 0; t < 5; ++t) {
        V3 P[4];
        double d[4];
        for (int v = 0; v < 4; ++v) {
            const int c = vert_corner((CutTab::tet5[t] >> (4 * v)) & 15, even != 0);
            P[v] = V3{(c & 1) ? box[1] : box[0], (c & 2) ? box[3] : box[2], (c & 4) ? box[5] : box[4]};
            d[v] = ((n[0] * P[v].x + n[1] * P[v].y) + n[2] * P[v].z) - D;
        }
        const int cs = d2case_snap(d);
        cut_dispatch<false>(cs, P, d, [&](double v) { ref += (ts * v) / cellvol; }, [](double) {});
        const double whole = (ts * tet_vol3(P[0], P[1], P[2], P[3])) / cellvol;
        frac += cs == 0 ? whole : (cs == 15 ? 0.0 : whole * tet_liquid_fraction(cs, d));
    }
    out[0] = ref;
    out[1] = frac;
}

// adjoint of gL * liquid + gV * total volume of the cut of h_cut_compare, with respect to
// x (12), n (3) and D (1): out[0..15] through cut_dispatch_adj, out[16..31] through
// cut_canonical_adj
void h_cut_adj_compare(const double* x, const double* n, double D, double gL, double gV, double* out) {
    V3 P[4];
    double d[4], dd[4];
    for (int v = 0; v < 4; ++v) {
        P[v] = V3{x[3 * v], x[3 * v + 1], x[3 * v + 2]};
        d[v] = ((n[0] * P[v].x + n[1] * P[v].y) + n[2] * P[v].z) - D;
        dd[v] = d[v];
    }
    const int cs = d2case_snap(dd);
    for (int form = 0; form < 2; ++form) {
        V3 gP[4] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double gd[4] = {0.0, 0.0, 0.0, 0.0};
        if (form == 0) cut_dispatch_adj(cs, P, d, gL, gV, gP, gd);
        else cut_canonical_adj(cs, P, d, gL, gV, gP, gd);
        double* o = out + 16 * form;
        for (int q = 0; q < 16; ++q) o[q] = 0.0;
        for (int v = 0; v < 4; ++v) {
            o[3 * v + 0] = gP[v].x + gd[v] * n[0];
            o[3 * v + 1] = gP[v].y + gd[v] * n[1];
            o[3 * v + 2] = gP[v].z + gd[v] * n[2];
            o[12] += gd[v] * P[v].x;
            o[13] += gd[v] * P[v].y;
            o[14] += gd[v] * P[v].z;
            o[15] -= gd[v];
        }
    }
}

double h_alpha(double a, double b, double c, double VF) { return compute_alpha(a, b, c, VF); }

void h_alpha_adj(double a, double b, double c, double VF, double* out) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    compute_alpha_adj(a, b, c, VF, 1.0, out[0], out[1], out[2], out[3]);
}

double h_dist(int i, int j, int k, const double* n, double VF) {
    return compute_dist(g_mesh.m, g_mesh.p, i, j, k, n[0], n[1], n[2], VF);
}

void h_dist_adj(int i, int j, int k, const double* n, double VF, double* out) {
    double gn[3] = {0.0, 0.0, 0.0}, gv = 0.0;
    compute_dist_adj(g_mesh.m, g_mesh.p, i, j, k, n[0], n[1], n[2], VF, 1.0, gn, gv);
    out[0] = gn[0]; out[1] = gn[1]; out[2] = gn[2]; out[3] = gv;
}

// uf, vf, wf: cell-extent arrays of the mesh set by h_mesh
void h_project(const double* uf, const double* vf, const double* wf, const double* x, int i, int j, int k, double dt,
               double* out) {
    double pt[3] = {x[0], x[1], x[2]};
    project_pt(g_mesh.m, g_mesh.p, uf, vf, wf, pt, i, j, k, dt);
    out[0] = pt[0]; out[1] = pt[1]; out[2] = pt[2];
}

void h_project_adj(const double* uf, const double* vf, const double* wf, const double* x, int i, int j, int k,
                   double dt, const double* g, double* guf, double* gvf, double* gwf) {
    const double* const F[3] = {uf, vf, wf};
    double* const G[3] = {guf, gvf, gwf};
    project_adj(g_mesh.m, g_mesh.p, F, x, i, j, k, dt, g, [&](int c, int64_t idx, double w) { G[c][idx] += w; });
}

// VF: VF-extent array of the mesh set by h_mesh; the candidate's normalised vector
void h_elvira_cand(const double* VF, int i, int j, int k, int cand, double* nout) {
    const DevMesh& m = g_mesh.m;
    auto at = [&](int di, int dj, int dk) { return VF[m.vof(i + di, j + dj, k + dk)]; };
    const double gn[3] = {0.0, 0.0, 0.0};
    // forward: the raw vector as elvira_cand_adj builds it (seed zero), normalised
    const int fam = cand / 9, a = (cand % 9) / 3, b = cand % 3;
    auto P0 = [](int q) { return q == 2 ? 0 : -1; };
    auto P1 = [](int q) { return q == 0 ? 0 : 1; };
    double r[3];
    if (fam == 0) {
        const int i1 = P0(b), i2 = P1(b), j1 = P0(a), j2 = P1(a);
        const double dz = m.dz;
        r[0] = -(((at(i2, 0, -1) * dz + at(i2, 0, 0) * dz) + at(i2, 0, 1) * dz) -
                 ((at(i1, 0, -1) * dz + at(i1, 0, 0) * dz) + at(i1, 0, 1) * dz)) / (m.xm[i + i2] - m.xm[i + i1]);
        r[1] = -(((at(0, j2, -1) * dz + at(0, j2, 0) * dz) + at(0, j2, 1) * dz) -
                 ((at(0, j1, -1) * dz + at(0, j1, 0) * dz) + at(0, j1, 1) * dz)) / (m.ym[j + j2] - m.ym[j + j1]);
        r[2] = at(0, 0, 1) > at(0, 0, -1) ? -1.0 : 1.0;
    } else if (fam == 1) {
        const int i1 = P0(b), i2 = P1(b), k1 = P0(a), k2 = P1(a);
        const double dy = m.dy;
        r[0] = -(((at(i2, -1, 0) * dy + at(i2, 0, 0) * dy) + at(i2, 1, 0) * dy) -
                 ((at(i1, -1, 0) * dy + at(i1, 0, 0) * dy) + at(i1, 1, 0) * dy)) / (m.xm[i + i2] - m.xm[i + i1]);
        r[1] = at(0, 1, 0) > at(0, -1, 0) ? -1.0 : 1.0;
        r[2] = -(((at(0, -1, k2) * dy + at(0, 0, k2) * dy) + at(0, 1, k2) * dy) -
                 ((at(0, -1, k1) * dy + at(0, 0, k1) * dy) + at(0, 1, k1) * dy)) / (m.zm[k + k2] - m.zm[k + k1]);
    } else {
        const int j1 = P0(b), j2 = P1(b), k1 = P0(a), k2 = P1(a);
        const double dx = m.dx;
        r[0] = at(1, 0, 0) > at(-1, 0, 0) ? -1.0 : 1.0;
        r[1] = -(((at(-1, j2, 0) * dx + at(0, j2, 0) * dx) + at(1, j2, 0) * dx) -
                 ((at(-1, j1, 0) * dx + at(0, j1, 0) * dx) + at(1, j1, 0) * dx)) / (m.ym[j + j2] - m.ym[j + j1]);
        r[2] = -(((at(-1, 0, k2) * dx + at(0, 0, k2) * dx) + at(1, 0, k2) * dx) -
                 ((at(-1, 0, k1) * dx + at(0, 0, k1) * dx) + at(1, 0, k1) * dx)) / (m.zm[k + k2] - m.zm[k + k1]);
    }
    const double len = sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    nout[0] = r[0] / len; nout[1] = r[1] / len; nout[2] = r[2] / len;
    (void)gn;
}

void h_elvira_cand_adj(const double* VF, int i, int j, int k, int cand, const double* gn, double* gVF) {
    const DevMesh& m = g_mesh.m;
    elvira_cand_adj(
        m, i, j, k, cand, [&](int di, int dj, int dk) { return VF[m.vof(i + di, j + dj, k + dk)]; }, gn,
        [&](int di, int dj, int dk, double w) { gVF[m.vof(i + di, j + dj, k + dk)] += w; });
}

}  // extern "C"
