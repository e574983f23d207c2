// mist_tree.h -- the breadth-first cut tree of a band cell shared by the SL-VOF cut pass
// (mist_vof.hip) and its reverse (mist_adjoint.hip): capacities, the LDS layout, one level of
// mesh-plane cuts, the bulk test of a cell against its PLIC plane and the cut-list entry.
#pragma once

#include "mist_kernels.h"

namespace mist {

// ------------------------------------------------------------------------------------
// Band-cell transport, one block per cell: cutTet (src/VFgeom.jl:713-900) of the five
// pre-image tets of a cell (src/Transport.jl:108-150) as a breadth-first walk of the cut tree.
//
// All tets of a cell are cut level by level: the x-planes while any node spans two or more
// x-cells, then y, then z while a node spans three or more z-cells. Each level is one pass of
// the block over the node list: a thread cuts one node by the plane at its largest index
// (src/VFgeom.jl:738-756), appends its new vertices to the cell's vertex pool and its child
// tets to the next list, at offsets from block prefix sums. The last z-plane (a node spanning
// two z-cells) is cut again by each thread that evaluates one of its children, so the leaves
// (the PLIC cuts, src/VFgeom.jl:757-830) run one per thread with nothing stored per leaf.
//
// Cell indices are kept as 2-bit offsets from the smallest index `lo` of the projected
// corners (6 bits per vertex: x, y, z). The index of a cut vertex is the pt2index! walk from
// the cell of the vertex numbered like it, clamped between the indices of its edge ends
// (src/VFgeom.jl:853-862); the walk is bounded by the corners' index box, which the clamp
// makes exact.
//
// The liquid and total volumes of the cell are the sums of the leaf volumes, added per thread
// and then over the block in a fixed order. When every cell that the corners' index box covers
// lies entirely on one side of its PLIC plane, every leaf is all-liquid (or all-gas), the
// reference's two sums are identical (or the liquid one is zero) and the new VF is exactly
// 1 (or 0): such cells are written without cutting.
// ------------------------------------------------------------------------------------
// Threads per cell and capacities: vertex pool (8 corners + mesh-cut vertices), nodes per
// list, leaves per cell.
template <int NT_, int PV_, int NL_, int NLEAF_>
struct TreeCfg {
    static constexpr int NT = NT_, PV = PV_, NL = NL_, NLEAF = NLEAF_;
};
// Sized from the oracle's counts on the 3D deformation (99.9% of cells need at most ~320
// pool vertices and ~230 nodes after the y cuts): the main pass takes almost all cells at
// ~17 KB of LDS per block (8-byte nodes); the few that overflow it go to a second pass with 256 threads
// and ~90 KB.
using TreeMain = TreeCfg<64, 256, 384, 1152>;
using TreeWide = TreeCfg<256, 1536, 1536, 6144>;

struct TreeNode {
    unsigned w0, w1;  // pool ids of vertices 0,1 and 2,3 (16 bits each)
    unsigned bits;    // cell offset from lo of vertex n along axis q at bit 6n + 2q
};
// A node as the lists in LDS hold it: with at most 256 pool vertices the four ids take 8 bits
// each and a node 8 bytes (more blocks fit a compute unit's LDS); otherwise the TreeNode itself.
struct TreeNode8 {
    unsigned ids;  // pool id of vertex n at bits 8n
    unsigned bits;
};
template <int PV>
struct NodeSlot {
    using type = TreeNode;
};
template <>
struct NodeSlot<256> {
    using type = TreeNode8;
};
__device__ __forceinline__ TreeNode node_unpack(const TreeNode& N) { return N; }
__device__ __forceinline__ TreeNode node_unpack(const TreeNode8& P) {
    return TreeNode{(P.ids & 0xFFu) | (((P.ids >> 8) & 0xFFu) << 16), ((P.ids >> 16) & 0xFFu) | ((P.ids >> 24) << 16),
                    P.bits};
}
__device__ __forceinline__ void node_pack(TreeNode& D, const TreeNode& N) { D = N; }
__device__ __forceinline__ void node_pack(TreeNode8& D, const TreeNode& N) {
    D.ids = (N.w0 & 0xFFu) | (((N.w0 >> 16) & 0xFFu) << 8) | ((N.w1 & 0xFFu) << 16) | (((N.w1 >> 16) & 0xFFu) << 24);
    D.bits = N.bits;
}

// The cut tables in LDS: every lane indexes them with its own case, and LDS answers such a
// gather with less latency than the constant-memory path.
struct CutTabS {
    unsigned short vtet[16][6];
    unsigned char edge[16][4];
    unsigned char ntets[16], nvert[16], side2[16];
};

template <class C>
struct TreeShared {
    CutTabS tab;
    double px[C::PV], py[C::PV], pz[C::PV];
    typename NodeSlot<C::PV>::type node[2][C::NL];
    unsigned short lnode[C::NLEAF];  // node (bits 0-10) and child (bits 11-13) of each leaf
    double plane[3][5];                // faces lo+0 .. lo+span+1 per axis
    double nb[64][4];                  // nx, ny, nz, D of the cells of the index box
    signed char cls[64];               // +1 / -1: the cell lies on the liquid / gas side of its plane
    int next;                          // cell taken from the work queue
    int scan[C::NT / 64];
    double red[2][C::NT / 64];
};

template <class C>
__device__ __forceinline__ TreeNode node_get(const TreeShared<C>& S, int list, int e) {
    return node_unpack(S.node[list][e]);
}
template <class C>
__device__ __forceinline__ void node_set(TreeShared<C>& S, int list, int e, const TreeNode& N) {
    node_pack(S.node[list][e], N);
}
template <class C>
__device__ __forceinline__ unsigned node_bits(const TreeShared<C>& S, int list, int e) {
    return S.node[list][e].bits;
}
// the leaf entry: node e of the last list and child c of its last z cut
__device__ __forceinline__ unsigned short leaf_entry(int e, int c) { return (unsigned short)(e | (c << 11)); }

__device__ __forceinline__ unsigned tn_id(const TreeNode& N, int n) {
    return ((n < 2 ? N.w0 : N.w1) >> ((n & 1) * 16)) & 0xFFFFu;
}
__device__ __forceinline__ unsigned boff(unsigned bits, int n, int q) { return (bits >> (6 * n + 2 * q)) & 3u; }
// a[n] for a register array and a runtime n. The empty asm makes the four values opaque so
// the selects stay selects of registers; otherwise the optimizer turns them into an indexed
// load, which puts the array in scratch memory.
__device__ __forceinline__ void opaque(double& v) { asm("" : "+v"(v)); }
__device__ __forceinline__ void opaque(unsigned& v) { asm("" : "+v"(v)); }
__device__ __forceinline__ void opaque(V3& v) {
    opaque(v.x);
    opaque(v.y);
    opaque(v.z);
}
template <class T>
__device__ __forceinline__ T sel4(const T (&a)[4], int n) {
    T a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    opaque(a0);
    opaque(a1);
    opaque(a2);
    opaque(a3);
    T r = a0;
    if (n == 1) r = a1;
    if (n == 2) r = a2;
    if (n == 3) r = a3;
    return r;
}
template <class C>
__device__ __forceinline__ V3 pool_v(const TreeShared<C>& S, unsigned id) {
    return V3{S.px[id], S.py[id], S.pz[id]};
}
template <class C>
__device__ __forceinline__ double pool_c(const TreeShared<C>& S, unsigned id, int axis) {
    return axis == 0 ? S.px[id] : (axis == 1 ? S.py[id] : S.pz[id]);
}

// Copies the cut tables into the block's LDS; a barrier must follow before their first use.
template <class C>
__device__ __forceinline__ void tab_load(TreeShared<C>& S) {
    for (int q = threadIdx.x; q < 16 * 6; q += C::NT) S.tab.vtet[q / 6][q % 6] = CUT_VTET[q / 6][q % 6];
    for (int q = threadIdx.x; q < 16 * 4; q += C::NT) S.tab.edge[q / 4][q % 4] = CUT_EDGE[q / 4][q % 4];
    if (threadIdx.x < 16) {
        S.tab.ntets[threadIdx.x] = CUT_NTETS[threadIdx.x];
        S.tab.nvert[threadIdx.x] = CUT_NVERT[threadIdx.x];
        S.tab.side2[threadIdx.x] = CUT_SIDE2[threadIdx.x];
    }
}

// Exclusive prefix sum over the block; *total receives the block sum.
template <int NT>
__device__ __forceinline__ int block_scan(int v, int& total, int* tmp) {
    const int lane = threadIdx.x & 63;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if constexpr (NT == 64) {
        total = __shfl(x, 63);
        return x - v;
    } else {
        const int w = threadIdx.x >> 6;
        if (lane == 63) tmp[w] = x;
        __syncthreads();
        int base = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < NT / 64; ++q) {
            const int t = tmp[q];
            base += q < w ? t : 0;
            tot += t;
        }
        __syncthreads();
        total = tot;
        return base + x - v;
    }
}

// pt2index! walk (src/Tools.jl:424-434) of coordinate c from offset s over the faces P of the
// index box [0, span], clamped to the offsets e1, e2 of the edge ends.
__device__ __forceinline__ unsigned walk_clamp(double c, int s, int e1, int e2, const double* P, int span) {
    int I = s;
    while (I < span && c > P[I + 1] + EPS) ++I;
    while (I > 0 && c < P[I] - EPS) --I;
    return (unsigned)min(max(I, min(e1, e2)), max(e1, e2));
}

// One level of mesh-plane cuts along `axis` (src/VFgeom.jl:832-899) of every node of list
// `cur` spanning at least `minspan` cells along it; others pass unchanged. Returns false when
// the pool or the next list would overflow. With `par` (the reverse pass) each new pool
// vertex records how it was made: its edge ends' pool ids (11 bits each), the axis and the
// plane offset, packed by par_pack.
struct Span3 {
    int s[3];
};

__device__ __forceinline__ unsigned par_pack(unsigned a, unsigned b, int axis, int mx) {
    return a | (b << 11) | ((unsigned)axis << 22) | ((unsigned)mx << 24);
}

template <class C>
__device__ bool tree_level(TreeShared<C>& S, int axis, int minspan, int cur, int& n, int& npool, Span3 span,
                           unsigned* par = nullptr) {
    static_assert(C::PV <= 2048, "pool ids are packed in 11 bits");
    static_assert(C::NL <= 2048, "leaf entries hold the node in 11 bits");
    constexpr int NT = C::NT;
    int nout = 0;
    for (int b = 0; b < n; b += NT) {
        const int e = b + (int)threadIdx.x;
        const bool active = e < n;
        TreeNode N{0u, 0u, 0u};
        int mx = 0, cs = 0, nch = 0, nv = 0;
        bool cut = false;
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        if (active) {
            N = node_get(S, cur, e);
            const int o0 = boff(N.bits, 0, axis), o1 = boff(N.bits, 1, axis), o2 = boff(N.bits, 2, axis),
                      o3 = boff(N.bits, 3, axis);
            mx = max(max(o0, o1), max(o2, o3));
            const int mn = min(min(o0, o1), min(o2, o3));
            cut = mx - mn >= minspan;
            if (cut) {
                const double P = S.plane[axis][mx];
                double dd[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    d[q] = pool_c(S, tn_id(N, q), axis) - P;
                    dd[q] = d[q];
                }
                cs = d2case_snap(dd);  // the reference snaps a copy
                nch = S.tab.ntets[cs];
                nv = S.tab.nvert[cs];
            } else {
                nch = 1;
            }
        }
        int tc, tv;
        const int oc = block_scan<NT>(nch, tc, S.scan);
        const int ov = block_scan<NT>(nv, tv, S.scan);
        if (nout + tc > C::NL || npool + tv > C::PV) return false;
        if (active && !cut) node_set(S, cur ^ 1, nout + oc, N);
        if (cut) {
            const int vb = npool + ov;
            unsigned cb[4] = {0u, 0u, 0u, 0u};  // side-1 offsets (6 bits) of the cut vertices
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q < nv) {
                    const unsigned eg = S.tab.edge[cs][q];
                    const int v1 = eg & 15, v2 = eg >> 4;
                    const double d1 = sel4(d, v1), d2 = sel4(d, v2);
                    const double mu = jmin(1.0, jmax(0.0, sdiv(-d1, d2 - d1)));
                    const V3 c = lerp3(pool_v(S, tn_id(N, v1)), pool_v(S, tn_id(N, v2)), mu);
                    S.px[vb + q] = c.x;
                    S.py[vb + q] = c.y;
                    S.pz[vb + q] = c.z;
                    if (par) par[vb + q] = par_pack(tn_id(N, v1), tn_id(N, v2), axis, mx);
                    unsigned bb = 0;
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) {
                        unsigned o;
                        if (ax == axis) {
                            o = (unsigned)(mx - 1);
                        } else {
                            const int e1 = boff(N.bits, v1, ax), e2 = boff(N.bits, v2, ax);
                            o = e1 == e2 ? (unsigned)e1
                                         : walk_clamp(ax == 0 ? c.x : (ax == 1 ? c.y : c.z), boff(N.bits, q, ax),
                                                      e1, e2, S.plane[ax], span.s[ax]);
                        }
                        bb |= o << (2 * ax);
                    }
                    cb[q] = bb;
                }
            }
            const unsigned amask = 3u << (2 * axis);
            const unsigned side2 = S.tab.side2[cs];
            for (int c = 0; c < nch; ++c) {
                const unsigned qv = S.tab.vtet[cs][c];
                const bool s2 = (side2 >> c) & 1;
                unsigned ids[4], bits = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int lv = nib(qv, k);
                    unsigned ob, id;
                    if (lv < 4) {
                        id = tn_id(N, lv);
                        ob = (N.bits >> (6 * lv)) & 63u;
                        const unsigned oa = (ob >> (2 * axis)) & 3u;
                        const unsigned na = s2 ? (unsigned)mx : min(oa, (unsigned)(mx - 1));
                        ob = (ob & ~amask) | (na << (2 * axis));
                    } else {
                        id = (unsigned)(vb + lv - 4);
                        ob = sel4(cb, lv - 4);
                        if (s2) ob = (ob & ~amask) | ((unsigned)mx << (2 * axis));
                    }
                    ids[k] = id;
                    bits |= ob << (6 * k);
                }
                TreeNode ch;
                ch.w0 = ids[0] | (ids[1] << 16);
                ch.w1 = ids[2] | (ids[3] << 16);
                ch.bits = bits;
                node_set(S, cur ^ 1, nout + oc + c, ch);
            }
        }
        nout += tc;
        npool += tv;
    }
    n = nout;
    __syncthreads();
    return true;
}

// Side of cell `c` of the index box relative to its PLIC plane over the cell's extent
// [x0,x1]x[y0,y1]x[z0,z1]: +1 when every point gives n.x - D < 0 (liquid) by a margin far
// above rounding, -1 when every point gives > 0 (gas), 0 otherwise.
__device__ __forceinline__ int box_class(double n0, double n1, double n2, double Dc, double x0, double x1, double y0,
                                         double y1, double z0, double z1) {
    const double a0 = n0 * x0, a1 = n0 * x1, b0 = n1 * y0, b1 = n1 * y1, c0 = n2 * z0, c1 = n2 * z1;
    const double dmax = ((fmax(a0, a1) + fmax(b0, b1)) + fmax(c0, c1)) - Dc;
    const double dmin = ((fmin(a0, a1) + fmin(b0, b1)) + fmin(c0, c1)) - Dc;
    const double mag = fabs(n0) * (fmax(fabs(x0), fabs(x1)) + (x1 - x0)) +
                       fabs(n1) * (fmax(fabs(y0), fabs(y1)) + (y1 - y0)) +
                       fabs(n2) * (fmax(fabs(z0), fabs(z1)) + (z1 - z0)) + fabs(Dc);
    const double tau = 1e-9 * mag;
    return dmax < -tau ? 1 : (dmin > tau ? -1 : 0);
}

// Slot `base + rank` of a wave-aggregated append of the lanes with `flag` to *count.
__device__ __forceinline__ int wave_slot(bool flag, int* count) {
    const unsigned long long mask = __ballot(flag);
    if (mask == 0) return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, leader);
    return flag ? base + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

// The momentum predictor's part of a cut-tree pass: the leaf cells' velocities and phase
// densities, the face velocities and dt that build the correction tets, the predicted cell
// velocities it writes, and the flag of a cell search that left the mesh.
struct TreeNS {
    LeafMomentum lm;
    const double* uf;
    const double* vf;
    const double* wf;
    double dt;
    double* us;
    double* vs;
    double* ws;
    int* ovf;
};

// A band cell to cut: interior id, smallest cell index `lo` of its projected corners, and the
// 6-bit offsets (x, y, z: 2 bits each) of corner c from lo at bits 6c.
struct CutCell {
    int q, lo0, lo1, lo2;
    unsigned long long offs;
};
static_assert(sizeof(CutCell) == 24, "the context allocates 24 bytes per cut-list entry");

// A leaf of the tree of a cell: its four vertices, the cell it lies in (offset l in the index
// box) and how it was made from its node (the node's last z cut, src/VFgeom.jl:853-862, or the
// node itself). The same arithmetic as the leaf loop of k_sl_tree.
struct TreeLeaf {
    V3 Q[4];
    TreeNode N;
    int l;
    bool zcut;
    double P;       // the z-plane of the cut
    int cs;         // its case
    unsigned qv;    // the leaf's vertices among the node's 4 + cut vertices
};

template <class C>
__device__ __forceinline__ void tree_leaf(const TreeShared<C>& S, int cur, int L, const int (&span)[3], int bx, int by,
                                          TreeLeaf& T) {
    const unsigned le = S.lnode[L];
    const int e = (int)(le & 0x7FFu), c = (int)(le >> 11);
    const TreeNode N = node_get(S, cur, e);
    T.N = N;
    const int o0 = boff(N.bits, 0, 2), o1 = boff(N.bits, 1, 2), o2 = boff(N.bits, 2, 2), o3 = boff(N.bits, 3, 2);
    const int mx = max(max(o0, o1), max(o2, o3));
    unsigned ob;
    T.zcut = mx > min(min(o0, o1), min(o2, o3));
    if (T.zcut) {
        const double P = S.plane[2][mx];
        double dd[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) dd[v] = S.pz[tn_id(N, v)] - P;
        const int cs = d2case_snap(dd);
        const unsigned qv = S.tab.vtet[cs][c];
        const bool s2 = (S.tab.side2[cs] >> c) & 1;
        T.P = P;
        T.cs = cs;
        T.qv = qv;
        int e1v = 0, e2v = 0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int lv = nib(qv, kk);
            const bool cutv = lv >= 4;
            const unsigned eg = S.tab.edge[cs][cutv ? lv - 4 : 0];
            const int v1 = cutv ? (int)(eg & 15) : lv, v2 = cutv ? (int)(eg >> 4) : lv;
            const V3 A = pool_v(S, tn_id(N, v1)), B = pool_v(S, tn_id(N, v2));
            const double d1 = A.z - P, d2 = B.z - P;
            const double mu = cutv ? jmin(1.0, jmax(0.0, sdiv(-d1, d2 - d1))) : 0.0;
            T.Q[kk] = lerp3(A, B, mu);
            if (kk == 0) { e1v = v1; e2v = v2; }
        }
        const int lv0 = nib(qv, 0);
        if (lv0 < 4) {
            ob = (N.bits >> (6 * lv0)) & 63u;
            const unsigned oa = ob >> 4;
            ob = (ob & 15u) | ((s2 ? (unsigned)mx : min(oa, (unsigned)(mx - 1))) << 4);
        } else {
            ob = (unsigned)(s2 ? mx : mx - 1) << 4;
#pragma unroll
            for (int ax = 0; ax < 2; ++ax) {
                const int ea = boff(N.bits, e1v, ax), eb = boff(N.bits, e2v, ax);
                const unsigned oo = ea == eb ? (unsigned)ea
                                             : walk_clamp(ax == 0 ? T.Q[0].x : T.Q[0].y, boff(N.bits, lv0 - 4, ax),
                                                          ea, eb, S.plane[ax], span[ax]);
                ob |= oo << (2 * ax);
            }
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) T.Q[v] = pool_v(S, tn_id(N, v));
        ob = N.bits & 63u;
        T.P = 0.0;
        T.cs = 0;
        T.qv = 0;
    }
    const unsigned ox = ob & 3u, oy = (ob >> 2) & 3u, oz = (ob >> 4) & 3u;
    T.l = (int)ox + bx * ((int)oy + by * (int)oz);
}

}  // namespace mist
