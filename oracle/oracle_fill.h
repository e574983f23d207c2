// oracle_fill.h -- TEST INFRASTRUCTURE ONLY (the oracle): the ghost fills Neumann! and the
// single-rank update_borders!, for arrays of any element type (the oracle's doubles and int16,
// and the dual numbers of its forward-mode restatement).
#ifndef MIST_ORACLE_FILL_H
#define MIST_ORACLE_FILL_H

#include <vector>

#include "oracle_core.h"

namespace ora {

// ------------------------------------------------------------------------------------
// Neumann! (src/Tools.jl:4-18): one-layer copy on every wall of the single rank, x then
// y then z, each over the array's full extent in the other two directions.
// ------------------------------------------------------------------------------------
template <class T>
inline void neumann_t(const Mesh& m, Arr3<T>& A) {
    const int64_t lo1 = A.e.lo[1], hi1 = A.hi(1), lo2 = A.e.lo[2], hi2 = A.hi(2), lo0 = A.e.lo[0], hi0 = A.hi(0);
    for (int64_t k = lo2; k <= hi2; ++k)
        for (int64_t j = lo1; j <= hi1; ++j) A(m.imin_ - 1, j, k) = A(m.imin_, j, k);
    for (int64_t k = lo2; k <= hi2; ++k)
        for (int64_t j = lo1; j <= hi1; ++j) A(m.imax_ + 1, j, k) = A(m.imax_, j, k);
    for (int64_t k = lo2; k <= hi2; ++k)
        for (int64_t i = lo0; i <= hi0; ++i) A(i, m.jmin_ - 1, k) = A(i, m.jmin_, k);
    for (int64_t k = lo2; k <= hi2; ++k)
        for (int64_t i = lo0; i <= hi0; ++i) A(i, m.jmax_ + 1, k) = A(i, m.jmax_, k);
    for (int64_t j = lo1; j <= hi1; ++j)
        for (int64_t i = lo0; i <= hi0; ++i) A(i, j, m.kmin_ - 1) = A(i, j, m.kmin_);
    for (int64_t j = lo1; j <= hi1; ++j)
        for (int64_t i = lo0; i <= hi0; ++i) A(i, j, m.kmax_ + 1) = A(i, j, m.kmax_);
}
// update_borders! on one rank (src/Parallel.jl:63-143). A periodic axis is a self
// exchange: the high ghost slab receives a copy of the first nghost interior layers, then
// the low ghost slab receives the last nghost layers (read after the first write). A
// non-periodic axis receives from MPI.PROC_NULL, which leaves the ghosts unchanged.
template <class T>
inline void borders_axis(const Mesh& m, Arr3<T>& A, int d) {
    const int64_t ng = m.ng;
    const int64_t mn = d == 0 ? m.imin_ : (d == 1 ? m.jmin_ : m.kmin_);
    const int64_t mx = d == 0 ? m.imax_ : (d == 1 ? m.jmax_ : m.kmax_);
    const int64_t mno = d == 0 ? m.imino_ : (d == 1 ? m.jmino_ : m.kmino_);
    const int64_t mxo = d == 0 ? m.imaxo_ : (d == 1 ? m.jmaxo_ : m.kmaxo_);
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    auto copy_slab = [&](int64_t src0, int64_t dst0, int64_t len) {
        std::vector<T> buf;
        buf.reserve((size_t)(len * A.e.n[d1] * A.e.n[d2]));
        int64_t id[3];
        for (int64_t s = 0; s < len; ++s)
            for (int64_t b = A.e.lo[d2]; b <= A.hi(d2); ++b)
                for (int64_t a = A.e.lo[d1]; a <= A.hi(d1); ++a) {
                    id[d] = src0 + s; id[d1] = a; id[d2] = b;
                    buf.push_back(A(id[0], id[1], id[2]));
                }
        size_t q = 0;
        for (int64_t s = 0; s < len; ++s)
            for (int64_t b = A.e.lo[d2]; b <= A.hi(d2); ++b)
                for (int64_t a = A.e.lo[d1]; a <= A.hi(d1); ++a) {
                    id[d] = dst0 + s; id[d1] = a; id[d2] = b;
                    A(id[0], id[1], id[2]) = buf[q++];
                }
    };
    copy_slab(mn, mx + 1, mxo - mx);          // A[max+1:maxo] = A[min:min+ng-1]
    copy_slab(mx - ng + 1, mno, mn - mno);    // A[mino:min-1] = A[max-ng+1:max]
}
template <class T>
inline void borders_t(const Ctx& c, Arr3<T>& A) {
    if (c.p.xper) borders_axis(c.m, A, 0);
    if (c.p.yper) borders_axis(c.m, A, 1);
    if (c.p.zper) borders_axis(c.m, A, 2);
}
}  // namespace ora

#endif
