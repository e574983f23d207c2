"""MI355X-native SL-VOF transport and pressure multigrid for Mist.jl / NS_Multiphase."""
from .params import create_mesh, mesh_struct, parameters  # noqa: F401
