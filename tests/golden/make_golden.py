"""Writes tests/golden/geometry_vectors.json: exact-rational known answers for the PLIC
closed form, the cell/tet plane cuts and the cube-plane volume.

Every input is a dyadic rational (exact in FP64); every expected output is the exact
rational result rounded once to FP64. Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import random
from fractions import Fraction as F

from exact_geometry import cube_fraction, tet_fraction, tet_volume, unit_cube_fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def dy(rng, bits=10, lo=-1.0, hi=1.0):
    """Random dyadic rational in [lo, hi] with `bits` fractional bits."""
    span = int((hi - lo) * (1 << bits))
    return F(int(lo * (1 << bits)) + rng.randint(0, span), 1 << bits)


def alpha_cases(rng):
    cases = []
    ms = [(F(1, 8), F(3, 8), F(1, 2)), (F(1, 4), F(1, 4), F(1, 2)), (F(1, 16), F(5, 16), F(5, 8)),
          (F(0), F(1, 4), F(3, 4)), (F(0), F(1, 2), F(1, 2)), (F(0), F(0), F(1)),
          (F(5, 16), F(5, 16), F(3, 8)),
          (F(3, 32), F(13, 32), F(1, 2)), (F(1, 64), F(1, 64), F(31, 32))]
    for m in ms:
        assert sum(m) == 1
        for q in range(1, 24):
            alpha = F(q, 48)  # alpha in (0, 0.5)
            V = unit_cube_fraction(m, alpha)
            if V <= 0 or V > F(1, 2):
                continue
            cases.append({"m": [float(c) for c in m], "V": float(V), "alpha": float(alpha)})
        for _ in range(8):
            alpha = dy(rng, 12, 0.0, 0.5)
            V = unit_cube_fraction(m, alpha)
            if V <= 0 or V > F(1, 2):
                continue
            cases.append({"m": [float(c) for c in m], "V": float(V), "alpha": float(alpha)})
    return cases


def tet_cases(rng, n=200):
    cases = []
    while len(cases) < n:
        v = [[dy(rng, 8) for _ in range(3)] for _ in range(4)]
        vol = tet_volume(v)
        if abs(vol) < F(1, 64):
            continue
        nrm = [dy(rng, 8) for _ in range(3)]
        if all(c == 0 for c in nrm):
            continue
        D = dy(rng, 8, -0.5, 0.5)
        try:
            fr = tet_fraction(v, nrm, D)
        except ValueError:
            continue
        cases.append({"v": [[float(c) for c in p] for p in v], "n": [float(c) for c in nrm],
                      "D": float(D), "frac": float(fr), "vol": float(vol),
                      "liq": float(fr * vol)})
    return cases


def cube_cases(rng, n=300):
    """Cells of the mesh used by the reference's CI test (Lx=3, Nx=Ny=10, Nz=1) are not
    dyadic, so these use a dyadic mesh: cell (i,j,k) of a 16^3 unit cube."""
    cases = []
    h = F(1, 16)
    while len(cases) < n:
        i, j, k = (rng.randint(1, 16) for _ in range(3))
        lo = [(i - 1) * h, (j - 1) * h, (k - 1) * h]
        nrm = [dy(rng, 8) for _ in range(3)]
        if all(c == 0 for c in nrm):
            continue
        centre = [lo[q] + h / 2 for q in range(3)]
        off = dy(rng, 10, -0.06, 0.06)
        D = sum(nrm[q] * centre[q] for q in range(3)) + off
        fr = cube_fraction(nrm, D, lo, [h, h, h])
        cases.append({"ijk": [i, j, k], "n": [float(c) for c in nrm], "D": float(D),
                      "frac": float(fr)})
    return cases


def main():
    rng = random.Random(20240601)
    out = {"alpha": alpha_cases(rng), "tet": tet_cases(rng), "cube": cube_cases(rng)}
    path = os.path.join(HERE, "geometry_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(path, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
