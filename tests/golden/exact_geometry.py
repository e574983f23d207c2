"""Exact rational geometry used to pin the oracle (no reference code involved).

* ``cube_fraction``: volume fraction of a box below a plane, by the inclusion-exclusion
  (truncated power) formula over the box vertices -- the forward volume formula of
  Scardovelli & Zaleski (2000), evaluated in exact rationals.
* ``tet_fraction``: volume fraction of a tetrahedron on the negative side of a linear
  function with distinct vertex values d_i: sum_i (-d_i)_+^3 / prod_{j!=i} (d_j - d_i)
  (divided-difference form of the simplex volume below a hyperplane).
* ``tet_volume``: signed volume -det(v1-v4, v2-v4, v3-v4)/6 with the reference's sign
  convention (src/VFgeom.jl:669-681) evaluated exactly.
"""
from __future__ import annotations

from fractions import Fraction as F
from itertools import product


def _pos(a):
    return a if a > 0 else F(0)


def unit_cube_fraction(m, alpha):
    """Fraction of [0,1]^3 with m.u <= alpha, m >= 0 (exact)."""
    m = [F(v) for v in m]
    alpha = F(alpha)
    nz = [v for v in m if v != 0]
    if len(nz) == 0:
        return F(1) if alpha >= 0 else F(0)
    if len(nz) == 1:
        r = alpha / nz[0]
        return min(F(1), max(F(0), r))
    dim = len(nz)
    s = F(0)
    for v in product((0, 1), repeat=dim):
        t = alpha - sum(c * b for c, b in zip(nz, v))
        s += (-1) ** sum(v) * _pos(t) ** dim
    den = F(1)
    for c in nz:
        den *= c
    den *= 2 if dim == 2 else 6
    return s / den


def cube_fraction(n, D, lo, h):
    """Fraction of the box lo + [0,h] (per axis) with n.x <= D (exact rationals)."""
    n = [F(v) for v in n]
    lo = [F(v) for v in lo]
    h = [F(v) for v in h]
    m = [n[q] * h[q] for q in range(3)]
    alpha = F(D) - sum(n[q] * lo[q] for q in range(3))
    for q in range(3):
        if m[q] < 0:
            alpha -= m[q]
            m[q] = -m[q]
    return unit_cube_fraction(m, alpha)


def tet_volume(v):
    v = [[F(c) for c in p] for p in v]
    a = [v[0][q] - v[3][q] for q in range(3)]
    b = [v[1][q] - v[3][q] for q in range(3)]
    c = [v[2][q] - v[3][q] for q in range(3)]
    det = (a[0] * (b[1] * c[2] - c[1] * b[2]) - a[1] * (b[0] * c[2] - c[0] * b[2])
           + a[2] * (b[0] * c[1] - c[0] * b[1]))
    return -det / 6


def tet_fraction(v, n, D):
    """Fraction of tet v with n.x - D < 0 (vertex values must be distinct)."""
    d = [sum(F(n[q]) * F(p[q]) for q in range(3)) - F(D) for p in v]
    if len(set(d)) != 4:
        raise ValueError("vertex values must be distinct")
    s = F(0)
    for i in range(4):
        den = F(1)
        for j in range(4):
            if j != i:
                den *= d[j] - d[i]
        s += _pos(-d[i]) ** 3 / den
    return s
