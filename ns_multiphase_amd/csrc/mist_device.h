// mist_device.h -- device-side geometry shared by the SL-VOF kernels (gfx950).
//
// The arithmetic of every function follows the reference operation by operation (same
// association order, literal powers as products, no FMA: the library is compiled with
// -ffp-contract=off) so that results agree with an FP64 evaluation of the reference to the
// last few ulps; cbrt/acos/sin/cos come from the device math library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mist {

constexpr double EPS = 2.220446049250313e-16;  // Julia eps() for Float64

// Geometry of one mesh level as seen by kernels: reference indices (interior 1..N, ghosts
// 1-ng..N+ng), coordinate pointers pre-offset so x[i] addresses reference index i.
struct DevMesh {
    int Nx, Ny, Nz, ng;
    int cnx, cny, cnz;   // cell-field extents (N + 2 ng)
    int vnx, vny, vnz;   // VF extents (N + 2 ng + 6)
    double dx, dy, dz;
    double Lx, Ly, Lz;
    const double* x;  const double* y;  const double* z;    // faces, index imino..imaxo+1
    const double* xm; const double* ym; const double* zm;   // centres, index imino..imaxo
    __host__ __device__ inline int64_t cell(int i, int j, int k) const {
        return (int64_t)(i - 1 + ng) + (int64_t)cnx * ((int64_t)(j - 1 + ng) + (int64_t)cny * (k - 1 + ng));
    }
    __host__ __device__ inline int64_t vof(int i, int j, int k) const {
        return (int64_t)(i + 2 + ng) + (int64_t)vnx * ((int64_t)(j + 2 + ng) + (int64_t)vny * (k + 2 + ng));
    }
    __host__ __device__ inline int imino() const { return 1 - ng; }
    __host__ __device__ inline int imaxo() const { return Nx + ng; }
    __host__ __device__ inline int jmaxo() const { return Ny + ng; }
    __host__ __device__ inline int kmaxo() const { return Nz + ng; }
};

// Per-axis trig factors of the prescribed velocities, pre-offset like DevMesh's coordinate
// pointers; index 0..2 = faces x, y, z, 3..5 = centres xm, ym, zm.
struct VelTab {
    const double* s1[6];  // sin(pi c)
    const double* s2[6];  // sin(2 pi c)
    const double* c1[6];  // cos(pi c)
};

struct DevParams {
    double VFlo, VFhi;
    int nband;
    int projection;      // MIST_PROJ_*
    int zalesak;         // test_case == "Zalesak"
    double rho_liq, rho_gas;
};

// Julia min/max for non-NaN arguments with signed zeros ordered -0.0 < 0.0.
__host__ __device__ __forceinline__ double jmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return __builtin_signbit(a) ? b : a;
    return a > b ? a : b;
}
__host__ __device__ __forceinline__ double jmin(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return __builtin_signbit(a) ? a : b;
    return a < b ? a : b;
}
// a /^ b  (src/Tools.jl:375-377)
__host__ __device__ __forceinline__ double sdiv(double a, double b) {
    return fabs(b) <= EPS ? __builtin_huge_val() : a / b;
}

// ---- cut tables (src/VFgeom.jl:54-165), case c = reference case - 1 ----------------------
// Vertex numbers 1..8 of the reference are stored 0-based. The 4 vertices of tet n of case c
// are packed as 4 nibbles of CUT_VTET[c][n]; CUT_EDGE packs the (v1, v2) pairs of the cut
// vertices as nibbles v1 | v2 << 4 per byte.
__constant__ const unsigned char CUT_NVERT[16] = {0, 3, 3, 4, 3, 4, 4, 3, 3, 4, 4, 3, 4, 3, 3, 0};
__constant__ const unsigned char CUT_NTETS[16] = {1, 4, 4, 6, 4, 6, 6, 4, 4, 6, 6, 4, 6, 4, 4, 1};
__constant__ const unsigned char CUT_NNTET[16] = {1, 2, 2, 4, 2, 4, 4, 4, 2, 4, 4, 4, 4, 4, 4, 2};

#define E2(a, b) (unsigned char)(((a) - 1) | (((b) - 1) << 4))
__constant__ const unsigned char CUT_EDGE[16][4] = {
    {0, 0, 0, 0},
    {E2(1, 2), E2(1, 3), E2(1, 4), 0},
    {E2(2, 3), E2(2, 4), E2(2, 1), 0},
    {E2(1, 4), E2(2, 4), E2(1, 3), E2(2, 3)},
    {E2(3, 4), E2(3, 1), E2(3, 2), 0},
    {E2(1, 4), E2(3, 4), E2(1, 2), E2(3, 2)},
    {E2(2, 4), E2(3, 4), E2(2, 1), E2(3, 1)},
    {E2(4, 1), E2(4, 2), E2(4, 3), 0},
    {E2(4, 1), E2(4, 2), E2(4, 3), 0},
    {E2(1, 3), E2(4, 3), E2(1, 2), E2(4, 2)},
    {E2(2, 3), E2(4, 3), E2(2, 1), E2(4, 1)},
    {E2(3, 4), E2(3, 1), E2(3, 2), 0},
    {E2(3, 2), E2(4, 2), E2(3, 1), E2(4, 1)},
    {E2(2, 3), E2(2, 4), E2(2, 1), 0},
    {E2(1, 2), E2(1, 3), E2(1, 4), 0},
    {0, 0, 0, 0}};
#undef E2

#define T4(a, b, c, d) (unsigned short)(((a) - 1) | (((b) - 1) << 4) | (((c) - 1) << 8) | (((d) - 1) << 12))
__constant__ const unsigned short CUT_VTET[16][6] = {
    {T4(1, 2, 3, 4), 0, 0, 0, 0, 0},
    {T4(5, 7, 6, 1), T4(6, 2, 3, 4), T4(4, 2, 5, 6), T4(5, 6, 7, 4), 0, 0},
    {T4(7, 5, 6, 2), T4(1, 3, 4, 6), T4(1, 5, 3, 6), T4(5, 7, 6, 1), 0, 0},
    {T4(5, 8, 6, 2), T4(5, 7, 8, 1), T4(5, 1, 8, 2), T4(5, 6, 8, 4), T4(5, 8, 7, 3), T4(5, 8, 3, 4)},
    {T4(6, 5, 7, 3), T4(2, 1, 4, 6), T4(6, 5, 4, 2), T4(6, 7, 5, 2), 0, 0},
    {T4(5, 6, 8, 3), T4(5, 8, 7, 1), T4(5, 8, 1, 3), T4(5, 8, 6, 4), T4(5, 7, 8, 2), T4(5, 8, 4, 2)},
    {T4(8, 6, 5, 3), T4(5, 7, 8, 2), T4(8, 5, 2, 3), T4(8, 5, 6, 4), T4(5, 8, 7, 1), T4(5, 8, 1, 4)},
    {T4(1, 2, 3, 7), T4(1, 2, 7, 6), T4(5, 7, 6, 1), T4(5, 6, 7, 4), 0, 0},
    {T4(5, 6, 7, 4), T4(1, 2, 3, 6), T4(5, 1, 3, 6), T4(5, 7, 6, 3), 0, 0},
    {T4(5, 8, 6, 4), T4(5, 7, 8, 1), T4(5, 8, 4, 1), T4(5, 6, 8, 3), T4(5, 8, 7, 2), T4(5, 8, 2, 3)},
    {T4(8, 5, 6, 4), T4(5, 8, 7, 2), T4(8, 2, 5, 4), T4(8, 6, 5, 3), T4(5, 7, 8, 1), T4(5, 8, 3, 1)},
    {T4(1, 4, 2, 7), T4(4, 1, 6, 7), T4(6, 7, 5, 4), T4(6, 5, 7, 3), 0, 0},
    {T4(8, 6, 5, 4), T4(5, 7, 8, 3), T4(8, 4, 5, 3), T4(8, 5, 6, 2), T4(5, 8, 7, 1), T4(5, 8, 1, 2)},
    {T4(3, 4, 1, 7), T4(7, 6, 3, 4), T4(7, 6, 5, 3), T4(7, 5, 6, 2), 0, 0},
    {T4(7, 4, 2, 3), T4(2, 3, 6, 7), T4(5, 6, 7, 2), T4(5, 7, 6, 1), 0, 0},
    {T4(1, 2, 3, 4), 0, 0, 0, 0, 0}};
#undef T4
// side of the mesh plane of each sub-tet (1 = below, 2 = above), bit n set when side 2
__constant__ const unsigned char CUT_SIDE2[16] = {
    0x00, 0x01, 0x01, 0x07, 0x01, 0x07, 0x07, 0x07, 0x01, 0x07, 0x07, 0x07, 0x07, 0x07, 0x07, 0x01};

// 5-tet decomposition (src/VFgeom.jl:235-256), 0-based corner numbers packed per tet.
__constant__ const unsigned short TET5[5] = {
    (0 | (1 << 4) | (3 << 8) | (5 << 12)), (0 | (3 << 4) | (2 << 8) | (6 << 12)),
    (0 | (4 << 4) | (5 << 8) | (6 << 12)), (3 | (6 << 4) | (5 << 8) | (7 << 12)),
    (0 | (3 << 4) | (6 << 8) | (5 << 12))};

__host__ __device__ __forceinline__ int nib(unsigned v, int q) { return (v >> (4 * q)) & 15; }

// tet_vol (src/VFgeom.jl:669-681)
__host__ __device__ __forceinline__ double tet_vol(const double* p0, const double* p1, const double* p2, const double* p3) {
    const double a1 = p0[0] - p3[0], a2 = p0[1] - p3[1], a3 = p0[2] - p3[2];
    const double b1 = p1[0] - p3[0], b2 = p1[1] - p3[1], b3 = p1[2] - p3[2];
    const double c1 = p2[0] - p3[0], c2 = p2[1] - p3[1], c3 = p2[2] - p3[2];
    return (-1.0 / 6.0) * ((a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3)) + a3 * (b1 * c2 - c1 * b2));
}

// d2case (src/VFgeom.jl:693-708) on a copy held in registers; d is snapped in place. With no
// distance below the snapping threshold (the common case) the snap changes nothing and the
// case is the vertices' sides.
__host__ __device__ __forceinline__ int d2case_snap(double d[4]) {
    if (!(fabs(d[0]) < EPS || fabs(d[1]) < EPS || fabs(d[2]) < EPS || fabs(d[3]) < EPS))
        return (d[0] > 0.0) | ((d[1] > 0.0) << 1) | ((d[2] > 0.0) << 2) | ((d[3] > 0.0) << 3);
    int npos = 0, nneg = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) { npos += d[n] > 0.0; nneg += d[n] < 0.0; }
    const double s = npos > nneg ? EPS : -EPS;
#pragma unroll
    for (int n = 0; n < 4; ++n)
        if (fabs(d[n]) < EPS) d[n] = s;
    return (d[0] > 0.0) | ((d[1] > 0.0) << 1) | ((d[2] > 0.0) << 2) | ((d[3] > 0.0) << 3);
}

// computeAlpha (src/VF.jl:164-214), Scardovelli & Zaleski (2000).
__host__ __device__ inline double compute_alpha(double a, double b, double c, double VF) {
    // sort ascending
    double m1 = a, m2 = b, m3 = c, t;
    if (m1 > m2) { t = m1; m1 = m2; m2 = t; }
    if (m2 > m3) { t = m2; m2 = m3; m3 = t; }
    if (m1 > m2) { t = m1; m1 = m2; m2 = t; }
    const double m12 = m1 + m2;
    const double V1 = (m1 * m1) / jmax((6.0 * m2) * m3, EPS);
    const double V2 = V1 + sdiv(0.5 * (m2 - m1), m3);
    double V3;
    if (m12 <= m3) V3 = sdiv(0.5 * m12, m3);
    else V3 = sdiv(((m3 * m3) * (3.0 * m12 - m3) + (m1 * m1) * (m1 - 3.0 * m3)) + (m2 * m2) * (m2 - 3.0 * m3),
                   ((6.0 * m1) * m2) * m3);
    double alpha;
    if (0.0 <= VF && VF < V1) {
        alpha = cbrt((((6.0 * m1) * m2) * m3) * VF);
    } else if (V1 <= VF && VF < V2) {
        alpha = 0.5 * (m1 + sqrt((m1 * m1) + ((8.0 * m2) * m3) * (VF - V1)));
    } else if (V2 <= VF && VF < V3) {
        double a0 = ((m1 * m1) * m1 + (m2 * m2) * m2) - (((6.0 * m1) * m2) * m3) * VF;
        double a1 = -3.0 * ((m1 * m1) + (m2 * m2));
        double a2 = 3.0 * m12;
        const double a3 = -1.0;
        a0 = a0 / a3; a1 = a1 / a3; a2 = a2 / a3;
        const double p0 = a1 / 3.0 - (a2 * a2) / 9.0;
        const double q0 = (a1 * a2 - 3.0 * a0) / 6.0 - ((a2 * a2) * a2) / 27.0;
        const double theta = acos(jmin(1.0, jmax(-1.0, q0 / sqrt(-1.0 * ((p0 * p0) * p0))))) / 3.0;
        alpha = sqrt(-p0) * (sqrt(3.0) * sin(theta) - cos(theta)) - a2 / 3.0;
    } else if (m12 <= m3) {
        alpha = m3 * VF + 0.5 * m12;
    } else {
        double a0 = (((m1 * m1) * m1 + (m2 * m2) * m2) + (m3 * m3) * m3) - (((6.0 * m1) * m2) * m3) * VF;
        double a1 = -3.0 * (((m1 * m1) + (m2 * m2)) + (m3 * m3));
        double a2 = 3.0;
        const double a3 = -2.0;
        a0 = a0 / a3; a1 = a1 / a3; a2 = a2 / a3;
        const double p0 = a1 / 3.0 - (a2 * a2) / 9.0;
        const double q0 = (a1 * a2 - 3.0 * a0) / 6.0 - ((a2 * a2) * a2) / 27.0;
        const double theta = acos(jmin(1.0, jmax(-1.0, q0 / sqrt(-((p0 * p0) * p0))))) / 3.0;
        alpha = sqrt(-p0) * (sqrt(3.0) * sin(theta) - cos(theta)) - a2 / 3.0;
    }
    return alpha;
}

// computeDist (src/VF.jl:109-156)
__host__ __device__ inline double compute_dist(const DevMesh& m, const DevParams& p, int i, int j, int k, double nx,
                                      double ny, double nz, double VF) {
    if (VF < p.VFlo) return -((m.Lx + m.Ly) + m.Lz);
    if (VF > p.VFhi) return +((m.Lx + m.Ly) + m.Lz);
    double mm0 = nx * m.dx, mm1 = ny * m.dy, mm2 = nz * m.dz;
    const double norm = ((fabs(mm0) + fabs(mm1)) + fabs(mm2)) + EPS;
    mm0 = mm0 / norm; mm1 = mm1 / norm; mm2 = mm2 / norm;
    double factor = 0.0;
    if (mm0 < 0.0) factor += mm0;
    if (mm1 < 0.0) factor += mm1;
    if (mm2 < 0.0) factor += mm2;
    const double VFo = VF > 0.5 ? 1.0 - VF : VF;
    double alpha = compute_alpha(fabs(mm0), fabs(mm1), fabs(mm2), VFo);
    alpha = VF > 0.5 ? 1.0 - alpha : alpha;
    alpha += factor;
    alpha -= 0.5 * ((mm0 + mm1) + mm2);
    double dist = alpha * norm;
    dist += (nx * m.xm[i] + ny * m.ym[j]) + nz * m.zm[k];
    return dist;
}

// Corner c (bits: x = c&1, y = c>>1&1, z = c>>2) of cell (i,j,k) in the reference vertex
// numbering of cell2verts! (src/VFgeom.jl:190-216): even cells number corners in lexical
// order, odd cells mirror x. Returns the lexical corner index of reference vertex v.
__host__ __device__ __forceinline__ int vert_corner(int v, bool even) { return even ? v : (v ^ 1); }


// ---- register-resident tet cutting -------------------------------------------------
// The same tables as compile-time constants, so that a case-specialised cut indexes its
// vertex arrays with constants and keeps them in registers.
struct CutTab {
    static constexpr unsigned char nvert[16] = {0, 3, 3, 4, 3, 4, 4, 3, 3, 4, 4, 3, 4, 3, 3, 0};
    // the 5-tet table TET5 (reference vertex n of tet t in nibble n)
    static constexpr unsigned short tet5[5] = {0x5310, 0x6230, 0x6540, 0x7563, 0x5630};
    static constexpr unsigned char ntets[16] = {1, 4, 4, 6, 4, 6, 6, 4, 4, 6, 6, 4, 6, 4, 4, 1};
    static constexpr unsigned char nntet[16] = {1, 2, 2, 4, 2, 4, 4, 4, 2, 4, 4, 4, 4, 4, 4, 2};
    // cut vertex n of case c interpolates reference vertices (v1, v2): nibbles v1 | v2 << 4
    static constexpr unsigned char edge[16][4] = {
        {0, 0, 0, 0},
        {0x10, 0x20, 0x30, 0},
        {0x21, 0x31, 0x01, 0},
        {0x30, 0x31, 0x20, 0x21},
        {0x32, 0x02, 0x12, 0},
        {0x30, 0x32, 0x10, 0x12},
        {0x31, 0x32, 0x01, 0x02},
        {0x03, 0x13, 0x23, 0},
        {0x03, 0x13, 0x23, 0},
        {0x20, 0x23, 0x10, 0x13},
        {0x21, 0x23, 0x01, 0x03},
        {0x32, 0x02, 0x12, 0},
        {0x12, 0x13, 0x02, 0x03},
        {0x21, 0x31, 0x01, 0},
        {0x10, 0x20, 0x30, 0},
        {0, 0, 0, 0}};
};

struct V3 {
    double x, y, z;
};

__host__ __device__ __forceinline__ V3 lerp3(const V3& a, const V3& b, double mu) {
    return V3{(1.0 - mu) * a.x + mu * b.x, (1.0 - mu) * a.y + mu * b.y, (1.0 - mu) * a.z + mu * b.z};
}
__host__ __device__ __forceinline__ double tet_vol3(const V3& p0, const V3& p1, const V3& p2, const V3& p3) {
    const double a1 = p0.x - p3.x, a2 = p0.y - p3.y, a3 = p0.z - p3.z;
    const double b1 = p1.x - p3.x, b2 = p1.y - p3.y, b3 = p1.z - p3.z;
    const double c1 = p2.x - p3.x, c2 = p2.y - p3.y, c3 = p2.z - p3.z;
    return (-1.0 / 6.0) * ((a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3)) + a3 * (b1 * c2 - c1 * b2));
}

template <int V>
struct IC {
    static constexpr int value = V;
};
template <int B, int E, class F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < E) {
        f(IC<B>{});
        sfor<B + 1, E>(f);
    }
}

// Vertex numbers 5..8 of a case: (0-based) tet numbers of the sub-tets, as constants.
__host__ __device__ constexpr int vtet_of(int c, int n, int k) {
    constexpr unsigned short T[16][6] = {
        {0x3210, 0, 0, 0, 0, 0},
        {0x0564, 0x3215, 0x5413, 0x3654, 0, 0},
        {0x1546, 0x5320, 0x5240, 0x0564, 0, 0},
        {0x1574, 0x0764, 0x1704, 0x3754, 0x2674, 0x3274},
        {0x2645, 0x5301, 0x1345, 0x1465, 0, 0},
        {0x2754, 0x0674, 0x2074, 0x3574, 0x1764, 0x1374},
        {0x2457, 0x1764, 0x2147, 0x3547, 0x0674, 0x3074},
        {0x6210, 0x5610, 0x0564, 0x3654, 0, 0},
        {0x3654, 0x5210, 0x5204, 0x2564, 0, 0},
        {0x3574, 0x0764, 0x0374, 0x2754, 0x1674, 0x2174},
        {0x3547, 0x1674, 0x3417, 0x2457, 0x0764, 0x0274},
        {0x6130, 0x6503, 0x3465, 0x2645, 0, 0},
        {0x3457, 0x2764, 0x2437, 0x1547, 0x0674, 0x1074},
        {0x6032, 0x3256, 0x2456, 0x1546, 0, 0},
        {0x2136, 0x6521, 0x1654, 0x0564, 0, 0},
        {0x3210, 0, 0, 0, 0, 0}};
    return (T[c][n] >> (4 * k)) & 15;
}

// Liquid (negative side) and total volume of tet P cut by a plane with vertex distances dm,
// sub-tets in the reference's order (src/VFgeom.jl:784-827): liquid tets from ntets down
// to nntet, then gas tets 1..nntet-1. on_liq(tv) / on_gas(tv) receive each sub-tet volume.
template <int CS, bool GAS, class FL, class FG>
__host__ __device__ __forceinline__ void cut_case(const V3 (&P)[4], const double (&dm)[4], FL&& on_liq, FG&& on_gas) {
    constexpr int NV = CutTab::nvert[CS], NT = CutTab::ntets[CS], NN = CutTab::nntet[CS];
    V3 Q[8];
    Q[0] = P[0]; Q[1] = P[1]; Q[2] = P[2]; Q[3] = P[3];
    sfor<0, NV>([&](auto I) {
        constexpr int n = decltype(I)::value;
        constexpr int v1 = CutTab::edge[CS][n] & 15, v2 = CutTab::edge[CS][n] >> 4;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-dm[v1], dm[v2] - dm[v1])));
        Q[4 + n] = lerp3(Q[v1], Q[v2], mu);
    });
    sfor<0, NT - NN + 1>([&](auto I) {
        constexpr int n = NT - decltype(I)::value;  // 1-based sub-tet number
        on_liq(tet_vol3(Q[vtet_of(CS, n - 1, 0)], Q[vtet_of(CS, n - 1, 1)], Q[vtet_of(CS, n - 1, 2)],
                        Q[vtet_of(CS, n - 1, 3)]));
    });
    if constexpr (GAS) {
        sfor<1, NN>([&](auto I) {
            constexpr int n = decltype(I)::value;
            on_gas(tet_vol3(Q[vtet_of(CS, n - 1, 0)], Q[vtet_of(CS, n - 1, 1)], Q[vtet_of(CS, n - 1, 2)],
                            Q[vtet_of(CS, n - 1, 3)]));
        });
    }
}

template <bool GAS, class FL, class FG>
__host__ __device__ __forceinline__ void cut_dispatch(int cs, const V3 (&P)[4], const double (&dm)[4], FL&& on_liq,
                                                      FG&& on_gas) {
    switch (cs) {
#define MIST_CASE(k) \
    case k: cut_case<k, GAS>(P, dm, on_liq, on_gas); break;
        MIST_CASE(0) MIST_CASE(1) MIST_CASE(2) MIST_CASE(3) MIST_CASE(4) MIST_CASE(5) MIST_CASE(6) MIST_CASE(7)
        MIST_CASE(8) MIST_CASE(9) MIST_CASE(10) MIST_CASE(11) MIST_CASE(12) MIST_CASE(13) MIST_CASE(14) MIST_CASE(15)
#undef MIST_CASE
        default: break;
    }
}

// The liquid and gas volumes of a leaf cut by its PLIC plane (cut_dispatch<true> summed per
// side) through two case bodies instead of fourteen, so the lanes of a wave that cut leaves of
// different cases run at most two bodies. Every case with three cut vertices is the reference's
// case 1 (vertex 1 alone above the plane) on the vertices renumbered with the lone vertex first;
// every case with four is case 3 (vertices 1 and 2 above) with the two above first. The pieces
// are the reference's, so the volumes agree with cut_dispatch to rounding (the sub-tets' vertex
// order changes); a renumbering of odd parity flips the sign of every piece, and when the lone
// vertex is below the plane its piece is the liquid one.
// Per case: renumbering (4 nibbles, the order the switch of cut_canonical copies the vertices
// in), parity (bit 16), lone vertex below (bit 17), four cut vertices (bit 18).
__host__ __device__ __forceinline__ unsigned canon_case(int cs) {
    switch (cs) {
        case 1: return 0x3210u;
        case 2: return 0x3201u | (1u << 16);
        case 4: return 0x3102u;
        case 8: return 0x2103u | (1u << 16);
        case 14: return 0x3210u | (1u << 17);
        case 13: return 0x3201u | (1u << 16) | (1u << 17);
        case 11: return 0x3102u | (1u << 17);
        case 7: return 0x2103u | (1u << 16) | (1u << 17);
        case 3: return 0x3210u | (1u << 18);
        case 5: return 0x3120u | (1u << 16) | (1u << 18);
        case 6: return 0x3021u | (1u << 18);
        case 9: return 0x2130u | (1u << 18);
        case 10: return 0x2031u | (1u << 16) | (1u << 18);
        case 12: return 0x1032u | (1u << 18);
        default: return 0x3210u;
    }
}

// The renumbering of canon_case as register copies per case (an index taken from the table
// would make the compiler load the vertices from a private array in scratch memory).
__host__ __device__ __forceinline__ void canon_perm(int cs, const V3 (&P)[4], const double (&dm)[4], V3 (&Q)[4],
                                                    double (&e)[4]) {
    auto perm = [&](auto A, auto B, auto C, auto D) {
        constexpr int a = decltype(A)::value, b = decltype(B)::value, cc = decltype(C)::value,
                      d = decltype(D)::value;
        Q[0] = P[a]; Q[1] = P[b]; Q[2] = P[cc]; Q[3] = P[d];
        e[0] = dm[a]; e[1] = dm[b]; e[2] = dm[cc]; e[3] = dm[d];
    };
    switch (cs) {
        case 2: case 13: perm(IC<1>{}, IC<0>{}, IC<2>{}, IC<3>{}); break;
        case 4: case 11: perm(IC<2>{}, IC<0>{}, IC<1>{}, IC<3>{}); break;
        case 8: case 7: perm(IC<3>{}, IC<0>{}, IC<1>{}, IC<2>{}); break;
        case 5: perm(IC<0>{}, IC<2>{}, IC<1>{}, IC<3>{}); break;
        case 6: perm(IC<1>{}, IC<2>{}, IC<0>{}, IC<3>{}); break;
        case 9: perm(IC<0>{}, IC<3>{}, IC<1>{}, IC<2>{}); break;
        case 10: perm(IC<1>{}, IC<3>{}, IC<0>{}, IC<2>{}); break;
        case 12: perm(IC<2>{}, IC<3>{}, IC<0>{}, IC<1>{}); break;
        default: perm(IC<0>{}, IC<1>{}, IC<2>{}, IC<3>{}); break;
    }
}

template <class FL, class FG>
__host__ __device__ __forceinline__ void cut_canonical(int cs, const V3 (&P)[4], const double (&dm)[4], FL&& on_liq,
                                                       FG&& on_gas) {
    if (cs == 0 || cs == 15) {  // one piece, the leaf itself
        const double tv = tet_vol3(P[0], P[1], P[2], P[3]);
        on_liq(cs == 0 ? tv : 0.0);
        on_gas(cs == 0 ? 0.0 : tv);
        return;
    }
    const unsigned c = canon_case(cs);
    V3 Q[4];
    double e[4];
    canon_perm(cs, P, dm, Q, e);
    double L = 0.0, G = 0.0;
    auto lq = [&](double v) { L += v; };
    auto gs = [&](double v) { G += v; };
    if (c & (1u << 18)) cut_case<3, true>(Q, e, lq, gs);
    else cut_case<1, true>(Q, e, lq, gs);
    if (c & (1u << 17)) {
        const double t = L;
        L = G;
        G = t;
    }
    if (c & (1u << 16)) {
        L = -L;
        G = -G;
    }
    on_liq(L);
    on_gas(G);
}

// One side of canonical case CS (1 or 3): the liquid sub-tets (LIQ) or the gas ones, added in
// the order cut_case hands them to on_liq / on_gas.
template <int CS, bool LIQ>
__host__ __device__ __forceinline__ double cut_side(const V3 (&P)[4], const double (&dm)[4]) {
    constexpr int NV = CutTab::nvert[CS], NT = CutTab::ntets[CS], NN = CutTab::nntet[CS];
    V3 Q[8];
    Q[0] = P[0]; Q[1] = P[1]; Q[2] = P[2]; Q[3] = P[3];
    sfor<0, NV>([&](auto I) {
        constexpr int n = decltype(I)::value;
        constexpr int v1 = CutTab::edge[CS][n] & 15, v2 = CutTab::edge[CS][n] >> 4;
        const double mu = jmin(1.0, jmax(0.0, sdiv(-dm[v1], dm[v2] - dm[v1])));
        Q[4 + n] = lerp3(Q[v1], Q[v2], mu);
    });
    double S = 0.0;
    if constexpr (LIQ) {
        sfor<0, NT - NN + 1>([&](auto I) {
            constexpr int n = NT - decltype(I)::value;
            S += tet_vol3(Q[vtet_of(CS, n - 1, 0)], Q[vtet_of(CS, n - 1, 1)], Q[vtet_of(CS, n - 1, 2)],
                          Q[vtet_of(CS, n - 1, 3)]);
        });
    } else {
        sfor<1, NN>([&](auto I) {
            constexpr int n = decltype(I)::value;
            S += tet_vol3(Q[vtet_of(CS, n - 1, 0)], Q[vtet_of(CS, n - 1, 1)], Q[vtet_of(CS, n - 1, 2)],
                          Q[vtet_of(CS, n - 1, 3)]);
        });
    }
    return S;
}

// The liquid volume cut_canonical gives (its on_liq argument, bit for bit), computing only the
// side it needs: case 1 or case 3 on the renumbered vertices, the gas side of case 1 when the
// lone vertex lies below the plane.
__host__ __device__ __forceinline__ double cut_canonical_liquid(int cs, const V3 (&P)[4], const double (&dm)[4]) {
    if (cs == 0) return tet_vol3(P[0], P[1], P[2], P[3]);
    if (cs == 15) return 0.0;
    const unsigned c = canon_case(cs);
    V3 Q[4];
    double e[4];
    canon_perm(cs, P, dm, Q, e);
    double v;
    if (c & (1u << 18)) v = cut_side<3, true>(Q, e);
    else if (c & (1u << 17)) v = cut_side<1, false>(Q, e);
    else v = cut_side<1, true>(Q, e);
    return (c & (1u << 16)) ? -v : v;
}

// The share of a tet's volume on the liquid side (d < 0) of the plane, from its four vertex
// distances snapped by d2case_snap (so none is zero) for a cut case cs (1..14). The point of
// the tet with barycentric weights w lies at distance sum w_v d_v, and the liquid share is the
// volume of {sum w_v d_v < 0}: with one liquid vertex (liquid depth a, gas heights g) it is
// a^3 / prod(a + g); with one gas vertex it is 1 minus the same corner share of that vertex;
// with two liquid vertices (depths p, q) and two gas ones (heights s, t) it is
//   (p^2 q^2 + p q (p + q)(s + t) + s t (p^2 + p q + q^2)) / ((p + s)(p + t)(q + s)(q + t)),
// the sum of the two liquid corner shares over their common denominator. Every term is
// positive, so the value is accurate to a few rounding errors; times the tet's volume it
// equals, in exact arithmetic, the sum of the reference's liquid sub-tets (src/VFgeom.jl:55-165).
__host__ __device__ __forceinline__ double tet_liquid_fraction(int cs, const double (&d)[4]) {
    const int ngas = __builtin_popcount((unsigned)cs);
    if (ngas == 2) {
        // depths of the two liquid vertices and heights of the two gas ones, in vertex order
        const bool l0 = d[0] < 0.0, l1 = d[1] < 0.0, l2 = d[2] < 0.0;
        const double p = -(l0 ? d[0] : (l1 ? d[1] : d[2]));
        const double q = -(d[3] < 0.0 ? d[3] : (l2 ? d[2] : d[1]));
        const double s = (!l0 ? d[0] : (!l1 ? d[1] : d[2]));
        const double t = (d[3] > 0.0 ? d[3] : (!l2 ? d[2] : d[1]));
        const double pq = p * q;
        const double num = (pq * pq + (pq * (p + q)) * (s + t)) + (s * t) * ((p * p + pq) + q * q);
        const double den = ((p + s) * (p + t)) * ((q + s) * (q + t));
        return num / den;
    }
    // one vertex alone on its side: its depth or height a and the corner share a^3 / prod(a + |d|)
    const bool lone_liquid = ngas == 3;
    double a = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        if ((d[v] < 0.0) == lone_liquid) a = fabs(d[v]);
    double den = 1.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) den *= ((d[v] < 0.0) == lone_liquid) ? 1.0 : (a + fabs(d[v]));
    const double corner = ((a * a) * a) / den;
    return lone_liquid ? corner : 1.0 - corner;
}

// (ts * v) / cellvol, the reference's per-sub-tet term of computePLIC2VF (ts = +-1). When the
// cell volume is a power of two (vol_is_pow2) the quotient is exact scaling, so v * (ts /
// cellvol) rounds to the same double (both are the correctly rounded value of ts v 2^-e)
// without a division; kernels take the form as a template parameter.
__host__ __device__ inline bool vol_is_pow2(double cellvol) {
    int e = 0;
    const double fr = frexp(cellvol, &e);
    return fr == 0.5 && e > -1000 && e < 1000;
}
template <bool POW2>
struct VolScale {
    double cv, inv;
    __device__ explicit VolScale(double cellvol) : cv(cellvol), inv(1.0 / cellvol) {}
    __device__ __forceinline__ double operator()(double ts, double v) const {
        if constexpr (POW2) return v * (ts * inv);
        else return (ts * v) / cv;
    }
};

// computePLIC2VF (src/VF.jl:219-278) of the cell [x0,x1]x[y0,y1]x[z0,z1] for the plane
// (nx,ny,nz,dist): 5 tets (parity `even` selects the vertex numbering and tetsign), d per tet
// vertex as the reference computes it, snapped in place by d2case, then the liquid sub-tets.
__device__ __forceinline__ void tet5_vertices(int t, bool even, double x0, double x1, double y0, double y1,
                                              double z0, double z1, V3 (&P)[4]) {
    const unsigned tv = TET5[t];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const int c = vert_corner(nib(tv, n), even);
        P[n].x = (c & 1) ? x1 : x0;
        P[n].y = (c & 2) ? y1 : y0;
        P[n].z = (c & 4) ? z1 : z0;
    }
}

__device__ __forceinline__ double plic2vf_reg(double x0, double x1, double y0, double y1, double z0, double z1,
                                              double nx, double ny, double nz, double dist, bool even,
                                              double cellvol) {
    const double ts = even ? 1.0 : -1.0;
    double VF = 0.0;
    for (int t = 0; t < 5; ++t) {
        V3 P[4];
        tet5_vertices(t, even, x0, x1, y0, y1, z0, z1, P);
        double d[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) d[n] = ((nx * P[n].x + ny * P[n].y) + nz * P[n].z) - dist;
        const int cs = d2case_snap(d);
        cut_dispatch<false>(cs, P, d, [&](double v) { VF += (ts * v) / cellvol; }, [](double) {});
    }
    return VF;
}

// computePLIC2VF of a cell lying entirely on the liquid side: case 1 in every tet, so each
// whole tet is one liquid sub-tet.
__device__ __forceinline__ double plic2vf_full(double x0, double x1, double y0, double y1, double z0, double z1,
                                               bool even, double cellvol) {
    const double ts = even ? 1.0 : -1.0;
    double VF = 0.0;
    for (int t = 0; t < 5; ++t) {
        V3 P[4];
        tet5_vertices(t, even, x0, x1, y0, y1, z0, z1, P);
        VF += (ts * tet_vol3(P[0], P[1], P[2], P[3])) / cellvol;
    }
    return VF;
}

// pt2index! (src/Tools.jl:424-434). The walk is bounded by the coordinate arrays; a walk
// that would leave them sets *ovf and stops at the last valid cell.
__host__ __device__ __forceinline__ void pt2index(const DevMesh& m, int I[3], const double pt[3], int i, int j, int k,
                                         int* ovf) {
    const int lo = 1 - m.ng;
    I[0] = i; I[1] = j; I[2] = k;
    while (pt[0] > m.x[I[0] + 1] + EPS) { if (I[0] + 1 > m.Nx + m.ng) { *ovf = 1; break; } I[0] += 1; }
    while (pt[0] < m.x[I[0]] - EPS) { if (I[0] - 1 < lo) { *ovf = 1; break; } I[0] -= 1; }
    while (pt[1] > m.y[I[1] + 1] + EPS) { if (I[1] + 1 > m.Ny + m.ng) { *ovf = 1; break; } I[1] += 1; }
    while (pt[1] < m.y[I[1]] - EPS) { if (I[1] - 1 < lo) { *ovf = 1; break; } I[1] -= 1; }
    while (pt[2] > m.z[I[2] + 1] + EPS) { if (I[2] + 1 > m.Nz + m.ng) { *ovf = 1; break; } I[2] += 1; }
    while (pt[2] < m.z[I[2]] - EPS) { if (I[2] - 1 < lo) { *ovf = 1; break; } I[2] -= 1; }
}

// Index search of get_velocity_{u,v,w}face (src/Tools.jl:520-540): i with c[i] <= p < c[i+1]
// clamped to [lo, hi-1]. The result does not depend on the starting index.
__host__ __device__ __forceinline__ int find_cell(double p, int i, int lo, int hi, const double* c) {
    while (p - c[i] < 0.0 && i > lo) i = i - 1;
    while (p - c[i + 1] >= 0.0 && i + 1 < hi) i = i + 1;
    return i;
}

// Face-velocity arrays read directly (the accessor of vel_face_acc and project_pt_acc).
struct FacePtrs {
    const double *u, *v, *w;
    __host__ __device__ __forceinline__ double fu(int64_t q) const { return u[q]; }
    __host__ __device__ __forceinline__ double fv(int64_t q) const { return v[q]; }
    __host__ __device__ __forceinline__ double fw(int64_t q) const { return w[q]; }
};

template <int C, class FA>
__host__ __device__ __forceinline__ double face_at(const FA& fa, int64_t q) {
    if constexpr (C == 0) return fa.fu(q);
    else if constexpr (C == 1) return fa.fv(q);
    else return fa.fw(q);
}

template <int C, class FA>
__host__ __device__ __forceinline__ double trilin_acc(const FA& fa, const DevMesh& m, int i, int j, int k, double wx1,
                                                     double wx2, double wy1, double wy2, double wz1, double wz2) {
    const int64_t sx = 1, sy = m.cnx, sz = (int64_t)m.cnx * m.cny;
    const int64_t b = m.cell(i, j, k);
    return wz1 * (wy1 * (wx1 * face_at<C>(fa, b + sx + sy + sz) + wx2 * face_at<C>(fa, b + sy + sz)) +
                  wy2 * (wx1 * face_at<C>(fa, b + sx + sz) + wx2 * face_at<C>(fa, b + sz))) +
           wz2 * (wy1 * (wx1 * face_at<C>(fa, b + sx + sy) + wx2 * face_at<C>(fa, b + sy)) +
                  wy2 * (wx1 * face_at<C>(fa, b + sx) + wx2 * face_at<C>(fa, b)));
}

__host__ __device__ __forceinline__ double trilin(const double* A, const DevMesh& m, int i, int j, int k, double wx1,
                                         double wx2, double wy1, double wy2, double wz1, double wz2) {
    return trilin_acc<0>(FacePtrs{A, A, A}, m, i, j, k, wx1, wx2, wy1, wy2, wz1, wz2);
}

// get_velocity_{u,v,w}face (src/Tools.jl:516-641), starting index (i,j,k); the face
// velocities come from the accessor fa (fu, fv, fw at a cell-extent index).
template <class FA>
__host__ __device__ inline void vel_face_acc(const DevMesh& m, const FA& fa, const double pt[3], int i0, int j0, int k0,
                                             double v[3]) {
    const int lo = 1 - m.ng;
    int i, j, k;
    double wx1, wy1, wz1;
    // u on x faces
    i = find_cell(pt[0], i0, lo, m.Nx + m.ng, m.x);
    j = find_cell(pt[1], j0, lo, m.Ny + m.ng, m.ym);
    k = find_cell(pt[2], k0, lo, m.Nz + m.ng, m.zm);
    wx1 = sdiv(pt[0] - m.x[i], m.x[i + 1] - m.x[i]);
    wy1 = sdiv(pt[1] - m.ym[j], m.ym[j + 1] - m.ym[j]);
    wz1 = sdiv(pt[2] - m.zm[k], m.zm[k + 1] - m.zm[k]);
    v[0] = trilin_acc<0>(fa, m, i, j, k, wx1, 1.0 - wx1, wy1, 1.0 - wy1, wz1, 1.0 - wz1);
    // v on y faces
    i = find_cell(pt[0], i0, lo, m.Nx + m.ng, m.xm);
    j = find_cell(pt[1], j0, lo, m.Ny + m.ng, m.y);
    wx1 = sdiv(pt[0] - m.xm[i], m.xm[i + 1] - m.xm[i]);
    wy1 = sdiv(pt[1] - m.y[j], m.y[j + 1] - m.y[j]);
    v[1] = trilin_acc<1>(fa, m, i, j, k, wx1, 1.0 - wx1, wy1, 1.0 - wy1, wz1, 1.0 - wz1);
    // w on z faces
    j = find_cell(pt[1], j0, lo, m.Ny + m.ng, m.ym);
    k = find_cell(pt[2], k0, lo, m.Nz + m.ng, m.z);
    wy1 = sdiv(pt[1] - m.ym[j], m.ym[j + 1] - m.ym[j]);
    wz1 = sdiv(pt[2] - m.z[k], m.z[k + 1] - m.z[k]);
    v[2] = trilin_acc<2>(fa, m, i, j, k, wx1, 1.0 - wx1, wy1, 1.0 - wy1, wz1, 1.0 - wz1);
}

__host__ __device__ inline void vel_face(const DevMesh& m, const double* uf, const double* vf, const double* wf,
                                const double pt[3], int i0, int j0, int k0, double v[3]) {
    vel_face_acc(m, FacePtrs{uf, vf, wf}, pt, i0, j0, k0, v);
}

}  // namespace mist
