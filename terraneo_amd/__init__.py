"""terraneo_amd -- MI355X-native matrix-free Stokes hot path (TerraNeo drop-in).

The compute path is libterraneo_amd.so (host C++ + HIP kernels for gfx950)
behind the C ABI declared in include/terraneo_amd.h.  This Python package is a
thin ctypes binding used by tests and benchmarks; torch is used only for
device memory and streams.  Importing works without a GPU (the library loads
and exports its symbols); compute calls need one.
"""
from ._capi import (  # noqa: F401
    LIB_PATH,
    TN_FGMRES_BLOCKDIAG,
    TN_FGMRES_BLOCKTRI,
    TN_MINRES_BLOCKDIAG,
    GmgOptions,
    KrylovOptions,
    SolveInfo,
    TN_ADD,
    TN_COMM_ADDITIVE,
    TN_COMM_SKIP,
    TN_DIRICHLET,
    TN_F32,
    TN_F64,
    TN_FREESLIP,
    TN_NEUMANN,
    TN_REPLACE,
    TNError,
    lib,
)
from .api import FV, Eps, Gmg, Mesh, Stokes, pcg_solve, stokes_solve  # noqa: F401
from . import api  # noqa: F401
