"""Analytic Stokes solution used by the reference's assess test (numpy).

SphericalStokesSolutionSmoothZeroSlip, l=2, m=2, k=2, shell r in [0.5, 1],
nu = 1, g = 1: coefficients and formulas as stated in the reference test
tests/test_epsilon_divdiv_stokes_assess.cpp:104-257 (velocity from the
poloidal scalar P_l(r) Y_22, pressure (G r^2 + H r^-3 + K r^3) Y_22, body
force f = -g r^2 Y_22 x / r).
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
