"""Host-side mirror of the reference's ``parameters`` struct and ``mesh_struct``.

``parameters`` follows ``src/Parameters.jl:3-91``: keyword construction, the same field
names and defaults, and the derived defaults ``VFhi = 1 - VFlo`` and ``nband = ceil(CFL)``.
String-valued options keep the reference's strings on the Python side and are turned into
the enums of ``include/mist_hip.h`` when the POD ``MistParams`` is built for the C ABI.
Fields the reference does not have (``vel_period`` and the ``mg_*`` knobs) are keyword
arguments with defaults that reproduce the reference behaviour where one exists.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

# enums of include/mist_hip.h
PRESSURE_SCHEMES = {"semi-lagrangian": 0, "finite-difference": 1}
NORMAL_METHODS = {"ELVIRA": 0}
VF_VELOCITIES = {"Nothing": 0, "Deformation": 1, "Deformation3D": 2, "rotation": 3,
                 "divFlow": 4, "divFlow3D": 5, "airblast": 6}
PROJECTION_METHODS = {"RK4": 0, "Euler": 1, "Midpoint": 2}
TESSELATIONS = {"5_tets": 0}
MG_PRECISIONS = {"fp64": 0, "mixed": 1}
COARSE_OPS = {"face-average": 0, "restricted-vf": 1}

FIELD = {
    "VF": 0, "nx": 1, "ny": 2, "nz": 3, "D": 4, "band": 5, "u": 6, "v": 7, "w": 8,
    "uf": 9, "vf": 10, "wf": 11, "P": 12, "divg": 13, "denx": 14, "deny": 15, "denz": 16,
    "RHS": 17, "us": 18, "vs": 19, "ws": 20, "viscx": 21, "viscy": 22, "viscz": 23,
    "Curve": 24, "sfx": 25, "sfy": 26, "sfz": 27, "mask": 28,
    "VF_bar": 29, "uf_bar": 30, "vf_bar": 31, "wf_bar": 32,
}
# wall sides of the boundary-condition descriptor: x-, x+, y-, y+, z-, z+
SIDES = {"xm": 0, "xp": 1, "ym": 2, "yp": 3, "zm": 4, "zp": 5}
REDUCE = {"sum": 0, "max": 1, "min": 2, "maxabs": 3}
IC_SHAPES = {"circle": 0, "sphere": 1, "bubble3d": 2}


class MistParams(ctypes.Structure):
    """POD ``mist_params`` of include/mist_hip.h (field order must match)."""

    _fields_ = [
        ("mu_liq", ctypes.c_double), ("mu_gas", ctypes.c_double),
        ("rho_liq", ctypes.c_double), ("rho_gas", ctypes.c_double), ("sigma", ctypes.c_double),
        ("grav_x", ctypes.c_double), ("grav_y", ctypes.c_double), ("grav_z", ctypes.c_double),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("tFinal", ctypes.c_double),
        ("Nx", ctypes.c_int64), ("Ny", ctypes.c_int64), ("Nz", ctypes.c_int64),
        ("stepMax", ctypes.c_int64),
        ("max_dt", ctypes.c_double), ("CFL", ctypes.c_double),
        ("xper", ctypes.c_int64), ("yper", ctypes.c_int64), ("zper", ctypes.c_int64),
        ("solveNS", ctypes.c_int64),
        ("tol", ctypes.c_double),
        ("pressure_scheme", ctypes.c_int64),
        ("mg_lvl", ctypes.c_int64),
        ("VFlo", ctypes.c_double), ("VFhi", ctypes.c_double),
        ("nband", ctypes.c_int64),
        ("normal_method", ctypes.c_int64),
        ("vf_velocity", ctypes.c_int64),
        ("projection_method", ctypes.c_int64),
        ("tesselation", ctypes.c_int64),
        ("test_case", ctypes.c_int64),
        ("vel_period", ctypes.c_double),
        ("mg_precision", ctypes.c_int64),
        ("mg_nu1", ctypes.c_int64), ("mg_nu2", ctypes.c_int64),
        ("mg_max_cycles", ctypes.c_int64),
        ("mg_coarse_op", ctypes.c_int64),
        ("mg_coarse_sweeps", ctypes.c_int64),
    ]


@dataclass(kw_only=True)
class parameters:  # noqa: N801 - the reference's type name
    """Mirror of ``@with_kw struct parameters`` (src/Parameters.jl:3-91)."""

    mu_liq: float
    mu_gas: float
    rho_liq: float
    rho_gas: float
    sigma: float
    Lx: float
    Ly: float
    Lz: float
    tFinal: float
    Nx: int
    Ny: int
    Nz: int
    stepMax: int
    CFL: float
    out_period: int
    nprocx: int
    nprocy: int
    nprocz: int
    xper: bool
    yper: bool
    zper: bool
    tol: float
    grav_x: float = 0.0
    grav_y: float = 0.0
    grav_z: float = 0.0
    max_dt: float = math.inf
    std_out_period: float = 1.0
    restart: bool = False
    restart_itr: int = 0
    solveNS: bool = True
    pressureSolver: str = "Secant"
    pressure_scheme: str = "semi-lagrangian"
    hypreSolver: str = "LGMRES"
    pressurePrecond: str = "none"
    mg_lvl: int = 1
    VFlo: float = 1e-10
    VFhi: float | None = None
    nband: int | None = None
    normalMethod: str = "ELVIRA"
    VFVelocity: str = "Nothing"
    projection_method: str = "RK4"
    interpolation_method: str = "trilinear"
    tesselation: str = "5_tets"
    iter_type: str = "floop"
    instability: str = "none"
    test_case: str = "test"
    VTK_dir: str | None = None
    # -- fields with no reference counterpart ------------------------------------------
    vel_period: float = 2.0          # T of cos(pi t / T); the reference hard-codes 2
    mg_precision: str = "fp64"       # "fp64" or "mixed"
    mg_nu1: int = 10                 # src/Multigrid.jl:688
    mg_nu2: int = 10                 # src/Multigrid.jl:689
    mg_max_cycles: int = 10000       # src/Multigrid.jl:347
    mg_coarse_op: str = "face-average"
    mg_coarse_sweeps: int = 64

    def __post_init__(self):
        if self.VFhi is None:
            self.VFhi = 1.0 - self.VFlo
        if self.nband is None:
            self.nband = int(math.ceil(self.CFL))
        if self.VTK_dir is None:
            self.VTK_dir = (f"results/VTK_{self.test_case}_{self.pressure_scheme}_"
                            f"{self.Nx}_{self.Ny}_{self.Nz}")
        if self.nprocx * self.nprocy * self.nprocz != 1:
            # One GPU holds the whole domain; the MPI block decomposition is not built.
            raise ValueError("Wrong number of processors, # procs is 1, example requires "
                             f"{self.nprocx} x {self.nprocy} x {self.nprocz}")

    def to_c(self) -> MistParams:
        def enum(table, key, what):
            if key not in table:
                raise ValueError(f"Unknown {what}: {key}")
            return table[key]

        p = MistParams()
        for name in ("mu_liq", "mu_gas", "rho_liq", "rho_gas", "sigma", "grav_x", "grav_y",
                     "grav_z", "Lx", "Ly", "Lz", "tFinal", "max_dt", "CFL", "tol", "VFlo",
                     "VFhi", "vel_period"):
            setattr(p, name, float(getattr(self, name)))
        for name in ("Nx", "Ny", "Nz", "stepMax", "mg_lvl", "nband", "mg_nu1", "mg_nu2",
                     "mg_max_cycles", "mg_coarse_sweeps"):
            setattr(p, name, int(getattr(self, name)))
        p.xper, p.yper, p.zper = int(self.xper), int(self.yper), int(self.zper)
        p.solveNS = int(self.solveNS)
        p.pressure_scheme = enum(PRESSURE_SCHEMES, self.pressure_scheme, "pressure_scheme")
        p.normal_method = enum(NORMAL_METHODS, self.normalMethod, "method to copute interface normal")
        p.vf_velocity = enum(VF_VELOCITIES, self.VFVelocity, "VFVelocity")
        p.projection_method = enum(PROJECTION_METHODS, self.projection_method, "projection_method")
        p.tesselation = enum(TESSELATIONS, self.tesselation, "tesselation")
        p.test_case = 1 if self.test_case == "Zalesak" else 0
        p.mg_precision = enum(MG_PRECISIONS, self.mg_precision, "mg_precision")
        p.mg_coarse_op = enum(COARSE_OPS, self.mg_coarse_op, "mg_coarse_op")
        return p


@dataclass
class mesh_struct:  # noqa: N801 - the reference's type name
    """Single-rank ``mesh_struct`` built like ``create_mesh`` (src/Mesh.jl:1-196)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    xm: np.ndarray
    ym: np.ndarray
    zm: np.ndarray
    dx: float
    dy: float
    dz: float
    Nx: int
    Ny: int
    Nz: int
    Lx: float
    Ly: float
    Lz: float
    nghost: int
    imin_: int = 1
    jmin_: int = 1
    kmin_: int = 1

    @property
    def imax_(self):
        return self.Nx

    @property
    def jmax_(self):
        return self.Ny

    @property
    def kmax_(self):
        return self.Nz

    @property
    def imino_(self):
        return 1 - self.nghost

    @property
    def jmino_(self):
        return 1 - self.nghost

    @property
    def kmino_(self):
        return 1 - self.nghost

    @property
    def imaxo_(self):
        return self.Nx + self.nghost

    @property
    def jmaxo_(self):
        return self.Ny + self.nghost

    @property
    def kmaxo_(self):
        return self.Nz + self.nghost

    def cell_shape(self):
        """Shape (i, j, k) of a ghosted cell field."""
        g = 2 * self.nghost
        return (self.Nx + g, self.Ny + g, self.Nz + g)

    def vof_shape(self):
        g = 2 * self.nghost + 6
        return (self.Nx + g, self.Ny + g, self.Nz + g)


def create_mesh(param: parameters, level: int = 1) -> mesh_struct:
    """create_mesh / mg_create_mesh for one rank (src/Mesh.jl:55-196, :204-352).

    Coordinate arrays are 0-based numpy vectors: ``x[q]`` is the reference's
    ``x[imino_ + q]``.
    """
    f = 2 ** (level - 1)
    N = [max(param.Nx // f, 1), max(param.Ny // f, 1), max(param.Nz // f, 1)]
    L = [float(param.Lx), float(param.Ly), float(param.Lz)]
    ng = 3 + max(0, int(math.ceil(param.CFL)) - 2)
    faces, centres, h = [], [], []
    for n, ln in zip(N, L):
        fc = np.zeros(n + 2 * ng + 1)
        for q in range(n + 1):
            fc[ng + q] = (float(q) * ln) / float(n)
        hh = fc[ng + 1] - fc[ng]
        for m in range(1, ng + 1):
            fc[ng - m] = fc[ng] - m * hh
            fc[ng + n + m] = fc[ng + n] + m * hh
        faces.append(fc)
        centres.append(0.5 * (fc[:-1] + fc[1:]))
        h.append(hh)
    return mesh_struct(x=faces[0], y=faces[1], z=faces[2], xm=centres[0], ym=centres[1],
                       zm=centres[2], dx=h[0], dy=h[1], dz=h[2], Nx=N[0], Ny=N[1], Nz=N[2],
                       Lx=L[0], Ly=L[1], Lz=L[2], nghost=ng)
