"""ctypes binding of the oracle's FV energy step (oracle/fv_oracle.c).

Test infrastructure: the checker for the HIP energy-step kernels.
FV scalars are [10][n+1][n+1][nr+1] with one ghost layer; cell centres
[..][3]; antidiffusive fluxes [..][6].
"""
import ctypes as C

import numpy as np

import oracle_lib as O

_done = False


def lib():
    global _done
    L = O.lib()
    if not _done:
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.orc_fv_size.argtypes = [vp]
        L.orc_fv_size.restype = C.c_int64
        L.orc_fv_cell_centers.argtypes = [vp, dp]
        L.orc_fv_ghost_update.argtypes = [vp, dp, C.c_int]
        L.orc_fv_ghost_pairs.argtypes = [vp, C.c_void_p]
        L.orc_fv_ghost_pairs.restype = C.c_int
        L.orc_fv_geometry.argtypes = [vp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.orc_fv_dt_stable.argtypes = [vp, dp, dp, C.c_double]
        L.orc_fv_dt_stable.restype = C.c_double
        L.orc_fct_predictor.argtypes = [vp, dp, dp, dp, C.c_double, C.c_double, C.c_void_p, C.c_int, dp, dp]
        L.orc_fv_dirichlet.argtypes = [vp, dp, C.c_int, C.c_double, C.c_int, C.c_double]
        L.orc_fct_limiter.argtypes = [vp, dp, dp, dp, dp]
        L.orc_fct_correction.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_fct_explicit_step.argtypes = [vp, dp, dp, dp, C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_int,
                                            C.c_double, C.c_int, C.c_double, dp]
        L.orc_fv_to_fe_rhs.argtypes = [vp, dp, dp]
        _done = True
    return L


def fv_shape(m):
    return (10, m.n + 1, m.n + 1, m.nr + 1)


def cell_centers(m):
    c = np.zeros(fv_shape(m) + (3,))
    lib().orc_fv_cell_centers(m.h, c)
    return c


def ghost_update(m, f, ncomp=1):
    f = np.ascontiguousarray(f, dtype=np.float64)
    lib().orc_fv_ghost_update(m.h, f, ncomp)
    return f


def ghost_pairs(m):
    cnt = lib().orc_fv_ghost_pairs(m.h, None)
    out = np.zeros((cnt, 6), np.int32)
    lib().orc_fv_ghost_pairs(m.h, out.ctypes.data)
    return out


def geometry(m, centers, vel, d, xc, yc, rc):
    beta = np.zeros(6)
    M = np.zeros(1)
    S = np.zeros((6, 3))
    lib().orc_fv_geometry(m.h, centers, np.ascontiguousarray(vel, dtype=np.float64), d, xc, yc, rc, beta, M, S)
    return beta, M[0], S


def dt_stable(m, centers, vel, kappa):
    return lib().orc_fv_dt_stable(m.h, centers, np.ascontiguousarray(vel, dtype=np.float64), kappa)


def _src(source):
    return None if source is None else np.ascontiguousarray(source, dtype=np.float64).ctypes.data


def predictor(m, centers, vel, T, dt, kappa=0.0, source=None, subtract_divergence=True):
    """(T with updated ghosts, T_L, antidiff)"""
    T = np.ascontiguousarray(T, dtype=np.float64).copy()
    TL = np.zeros(fv_shape(m))
    ad = np.zeros(fv_shape(m) + (6,))
    src = None if source is None else np.ascontiguousarray(source, dtype=np.float64)
    lib().orc_fct_predictor(m.h, centers, np.ascontiguousarray(vel, dtype=np.float64), T, dt, kappa,
                            None if src is None else src.ctypes.data, int(subtract_divergence), TL, ad)
    return T, TL, ad


def dirichlet(m, f, cmb=None, surf=None):
    f = np.ascontiguousarray(f, dtype=np.float64)
    lib().orc_fv_dirichlet(m.h, f, int(cmb is not None), float(cmb or 0.0), int(surf is not None),
                           float(surf or 0.0))
    return f


def limiter(m, TL, ad):
    TL = np.ascontiguousarray(TL, dtype=np.float64).copy()
    Rp = np.zeros(fv_shape(m))
    Rm = np.zeros(fv_shape(m))
    lib().orc_fct_limiter(m.h, TL, np.ascontiguousarray(ad), Rp, Rm)
    return TL, Rp, Rm


def correction(m, TL, ad, Rp, Rm, T):
    T = np.ascontiguousarray(T, dtype=np.float64).copy()
    lib().orc_fct_correction(m.h, TL, ad, Rp, Rm, T)
    return T


class FctStepper:
    """explicit FCT steps with persistent work buffers (as the reference's FVFCTBuffers)"""

    def __init__(self, m, centers):
        self.m, self.centers = m, centers
        sz = int(np.prod(fv_shape(m)))
        self.work = np.zeros(9 * sz)

    def step(self, vel, T, dt, kappa=0.0, source=None, subtract_divergence=True, cmb=None, surf=None):
        T = np.ascontiguousarray(T, dtype=np.float64)
        src = None if source is None else np.ascontiguousarray(source, dtype=np.float64)
        lib().orc_fct_explicit_step(self.m.h, self.centers, np.ascontiguousarray(vel, dtype=np.float64), T, dt,
                                    kappa, None if src is None else src.ctypes.data, int(subtract_divergence),
                                    int(cmb is not None), float(cmb or 0.0), int(surf is not None),
                                    float(surf or 0.0), self.work)
        return T


def fv_to_fe_rhs(m, Tfv):
    b = np.zeros(m.N)
    lib().orc_fv_to_fe_rhs(m.h, np.ascontiguousarray(Tfv, dtype=np.float64), b)
    return b
