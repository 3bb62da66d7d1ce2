"""Analytic Stokes solutions used by the reference's assess tests (numpy).

Spherical Stokes solutions with l=2, m=2, k=2 on the shell r in [0.5, 1],
nu = 1, g = 1: velocity from the poloidal scalar P_l(r) Y_22 with
P_l = A r^2 + B r^-3 + C r^4 + D r^-1 + E r^5, pressure (G r^2 + H r^-3 +
K r^3) Y_22, body force f = -g r^2 Y_22 x / r.  Coefficient sets:
  zero-slip both shells  tests/test_epsilon_divdiv_stokes_assess.cpp:104-257
  free-slip both shells  tests/test_epsilon_divdiv_stokes_assess_freeslip_freeslip.cpp:104-118
  free-slip CMB, zero-slip surface
                         tests/test_epsilon_divdiv_stokes_assess_freeslip.cpp:104-118
The module-level functions use the zero-slip set; Solution(name) any set.
"""
import numpy as np

A_ = 5.08347602739726064230e-03
B_ = 9.81021689497716840520e-05
C_ = -1.13441780821917810596e-02
D_ = -7.81844558599695542596e-04
E_ = 6.94444444444444405895e-03
G_ = 4.76455479452054797562e-01
H_ = 9.38213470319634694483e-03
K_ = -6.66666666666666629659e-01
CY = 0.38627420202318958  # sqrt(15 / (32 pi))
R_MIN, R_MAX = 0.5, 1.0


def _sph(X):
    r = np.sqrt(X[..., 0] ** 2 + X[..., 1] ** 2 + X[..., 2] ** 2)
    th = np.arccos(X[..., 2] / r)
    ph = np.arctan2(X[..., 1], X[..., 0])
    return r, th, ph


def _P(r):
    return A_ * r ** 2 + B_ / r ** 3 + C_ * r ** 4 + D_ / r + E_ * r ** 5


def _dP(r):
    return 2 * A_ * r - 3 * B_ / r ** 4 + 4 * C_ * r ** 3 - D_ / r ** 2 + 5 * E_ * r ** 4


def velocity(X):
    r, th, ph = _sph(X)
    st, ct = np.sin(th), np.cos(th)
    Y = CY * st * st * np.cos(2 * ph)
    dYt = CY * 2 * st * ct * np.cos(2 * ph)
    dYp = CY * st * st * (-2 * np.sin(2 * ph))
    u_r = -6.0 * _P(r) * Y / r
    rad = _P(r) / r + _dP(r)
    u_t = -rad * dYt
    req = np.sqrt(X[..., 0] ** 2 + X[..., 1] ** 2)
    pole = req < 1e-14 * r
    reqs = np.where(pole, 1.0, req)
    u_p = -rad / np.where(pole, 1.0, st) * dYp
    ux = X[..., 0] / r * u_r + X[..., 0] / reqs * ct * u_t - X[..., 1] / reqs * u_p
    uy = X[..., 1] / r * u_r + X[..., 1] / reqs * ct * u_t + X[..., 0] / reqs * u_p
    uz = X[..., 2] / r * u_r - st * u_t
    U = np.stack([ux, uy, uz])
    U[:, pole] = 0.0
    return U


def pressure(X):
    r, th, ph = _sph(X)
    return (G_ * r ** 2 + H_ / r ** 3 + K_ * r ** 3) * CY * np.sin(th) ** 2 * np.cos(2 * ph)


def body_force(X):
    r, th, ph = _sph(X)
    rho = r * r * CY * np.sin(th) ** 2 * np.cos(2 * ph)
    fr = -rho / r
    return np.stack([fr * X[..., 0], fr * X[..., 1], fr * X[..., 2]])


# L2 errors measured from the reference (Kokkos-OpenMP build of the reference,
# FGMRES rtol 1e-14, Dirichlet/Dirichlet, constant viscosity): level -> (vel, pre)
REFERENCE_ERRORS_DIRICHLET = {3: (2.88619e-05, 4.63833e-04), 4: (7.89908e-06, 1.48111e-04),
                              5: (2.07359e-06, 4.98176e-05), 6: (5.31344e-07, 1.70315e-05)}


# free-slip / free-slip and free-slip CMB + zero-slip surface coefficient sets
COEFFS = {
    "zeroslip": (A_, B_, C_, D_, E_, G_, H_, K_),
    "freeslip": (6.25000000000000034694e-03, 4.92125984251968497107e-05, -1.25492125984251971099e-02,
                 -6.94444444444444470947e-04, 6.94444444444444405895e-03, 5.27066929133858330658e-01,
                 8.33333333333333321769e-03, -6.66666666666666629659e-01),
    "freezeroslip": (4.54414476717381228543e-03, 3.74596588289534304281e-05, -1.10448363301060385744e-02,
                     -4.81212540341170963258e-04, 6.94444444444444405895e-03, 4.63883125864453627063e-01,
                     5.77455048409405112542e-03, -6.66666666666666629659e-01),
}


class Solution:
    """One coefficient set of the spherical Stokes solution family."""

    def __init__(self, name):
        self.c = COEFFS[name]

    def _P(self, r):
        A, B, C, D, E = self.c[:5]
        return A * r ** 2 + B / r ** 3 + C * r ** 4 + D / r + E * r ** 5

    def _dP(self, r):
        A, B, C, D, E = self.c[:5]
        return 2 * A * r - 3 * B / r ** 4 + 4 * C * r ** 3 - D / r ** 2 + 5 * E * r ** 4

    def velocity(self, X):
        r, th, ph = _sph(X)
        st, ct = np.sin(th), np.cos(th)
        Y = CY * st * st * np.cos(2 * ph)
        dYt = CY * 2 * st * ct * np.cos(2 * ph)
        dYp = CY * st * st * (-2 * np.sin(2 * ph))
        u_r = -6.0 * self._P(r) * Y / r
        rad = self._P(r) / r + self._dP(r)
        u_t = -rad * dYt
        req = np.sqrt(X[..., 0] ** 2 + X[..., 1] ** 2)
        pole = req < 1e-14 * r
        reqs = np.where(pole, 1.0, req)
        u_p = -rad / np.where(pole, 1.0, st) * dYp
        ux = X[..., 0] / r * u_r + X[..., 0] / reqs * ct * u_t - X[..., 1] / reqs * u_p
        uy = X[..., 1] / r * u_r + X[..., 1] / reqs * ct * u_t + X[..., 0] / reqs * u_p
        uz = X[..., 2] / r * u_r - st * u_t
        U = np.stack([ux, uy, uz])
        U[:, pole] = 0.0
        return U

    def pressure(self, X):
        G, H, K = self.c[5:]
        r, th, ph = _sph(X)
        return (G * r ** 2 + H / r ** 3 + K * r ** 3) * CY * np.sin(th) ** 2 * np.cos(2 * ph)

    def body_force(self, X):
        return body_force(X)


# free-slip reference runs (same solver stack, FGMRES rtol 1e-14): level -> (vel, pre)
REFERENCE_ERRORS_FREESLIP = {3: (3.09494e-05, 5.53425e-04), 4: (7.78134e-06, 1.50746e-04)}
REFERENCE_ERRORS_FREEZEROSLIP = {3: (3.19586e-05, 4.56753e-04), 4: (8.64578e-06, 1.22424e-04),
                                 5: (2.26962e-06, 3.57488e-05)}
