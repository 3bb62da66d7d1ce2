"""Synthetic inputs of the benchmark configurations (host side, numpy).

Random-harmonic fields on the thick shell: a sum of 32 real orthonormal
spherical harmonics Y_lm (degree l <= 8) with N(0,1) coefficients, times the
radial profile cos(pi (r - r_min) / (r_max - r_min)), affinely normalised over
the mesh nodes to [0, ln(contrast)].  The 32 (l, m) pairs are drawn without
replacement from the 81 pairs of degree <= 8 (ordered l = 0..8, m = -l..l) by
numpy's default_rng(seed); the coefficients are the next 32 standard normals
of the same generator.

Nodal layout everywhere: [10][n][n][nr] (radius fastest), matching the mesh.
"""
import math

import numpy as np

LMAX = 8


def _lm_pairs(lmax=LMAX):
    return [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]


def _assoc_legendre(lmax, x):
    """P_l^m(x) for 0 <= m <= l <= lmax (no Condon-Shortley phase), by the
    standard three-term recurrences.  Returns dict (l, m) -> array."""
    P = {}
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P[(0, 0)] = np.ones_like(x)
    for m in range(1, lmax + 1):
        P[(m, m)] = (2 * m - 1) * s * P[(m - 1, m - 1)]
    for m in range(0, lmax):
        P[(m + 1, m)] = (2 * m + 1) * x * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * x * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    return P


def real_sph_harm(l, m, xyz, P=None):
    """Real orthonormal spherical harmonic Y_lm at unit vectors xyz [..., 3]:
    m > 0: sqrt(2) N_lm P_l^m(cos t) cos(m p), m < 0: sqrt(2) N_l|m| P_l^|m| sin(|m| p),
    m = 0: N_l0 P_l(cos t), N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)."""
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    ct = np.clip(z / np.sqrt(x * x + y * y + z * z), -1.0, 1.0)
    ph = np.arctan2(y, x)
    if P is None:
        P = _assoc_legendre(l, ct)
    am = abs(m)
    N = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    if m == 0:
        return N * P[(l, 0)]
    if m > 0:
        return math.sqrt(2.0) * N * P[(l, am)] * np.cos(am * ph)
    return math.sqrt(2.0) * N * P[(l, am)] * np.sin(am * ph)


def harmonic_sum(xyz, seed, nterms=32, lmax=LMAX):
    """sum_i c_i Y_{l_i m_i}(xyz) with the seeded selection described above."""
    pairs = _lm_pairs(lmax)
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(len(pairs), size=nterms, replace=False))
    coef = rng.standard_normal(nterms)
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    ct = np.clip(z / np.sqrt(x * x + y * y + z * z), -1.0, 1.0)
    P = _assoc_legendre(lmax, ct)
    out = np.zeros(xyz.shape[:-1])
    for c, i in zip(coef, sel):
        l, m = pairs[i]
        out += c * real_sph_harm(l, m, xyz, P)
    return out


def shell_field(coords, radii, seed, r_min=None, r_max=None):
    """Unnormalised random-harmonic field at the nodes: [10][n][n][nr]."""
    r_min = radii[0] if r_min is None else r_min
    r_max = radii[-1] if r_max is None else r_max
    lat = harmonic_sum(coords, seed)  # [10][n][n]
    prof = np.cos(np.pi * (radii - r_min) / (r_max - r_min))
    return lat[..., None] * prof[None, None, None, :]


def normalise(f, hi):
    lo_v, hi_v = f.min(), f.max()
    return (f - lo_v) / (hi_v - lo_v) * hi


def log_viscosity(coords, radii, seed, contrast):
    """ln eta in [0, ln contrast] (nodal)."""
    return normalise(shell_field(coords, radii, seed), math.log(contrast))


def viscosity_c1(coords, radii):
    """Stokes free-slip configuration: contrast 1e4, seed 1."""
    return np.exp(log_viscosity(coords, radii, 1, 1e4))


def viscosity_c2(coords, radii):
    """Mixed-precision configuration: contrast 1e5 (seed 3), x30 for r > 0.9."""
    eta = np.exp(log_viscosity(coords, radii, 3, 1e5))
    return eta * np.where(radii > 0.9, 30.0, 1.0)[None, None, None, :]


def viscosity_c4(coords, radii):
    """Largest configuration: contrast 1e6 (seed 6), x100 for r > 0.95."""
    eta = np.exp(log_viscosity(coords, radii, 6, 1e6))
    return eta * np.where(radii > 0.95, 100.0, 1.0)[None, None, None, :]


def temperature(coords, radii, seed=2):
    """Buoyancy temperature T in [0, 1] (seed 2)."""
    return normalise(shell_field(coords, radii, seed), 1.0)


def buoyancy(coords, radii, seed=2):
    """Nodal body force f = T r_hat as [3][10][n][n][nr]."""
    T = temperature(coords, radii, seed)
    rh = coords / np.linalg.norm(coords, axis=-1, keepdims=True)
    return np.stack([T * rh[..., k][..., None] for k in range(3)])


def cell_centres(coords, radii):
    """Unit lateral direction and mid radius of every hex cell:
    ([10][n-1][n-1][3], [nr-1])."""
    c = coords[:, :-1, :-1] + coords[:, 1:, :-1] + coords[:, :-1, 1:] + coords[:, 1:, 1:]
    c = c / np.linalg.norm(c, axis=-1, keepdims=True)
    return c, 0.5 * (radii[:-1] + radii[1:])


def cell_log_viscosity(coords, radii, seed, contrast):
    """Per hex cell ln eta in [0, ln contrast]: [10][n-1][n-1][nr-1]."""
    cc, rm = cell_centres(coords, radii)
    f = shell_field(cc, rm, seed, radii[0], radii[-1])
    return normalise(f, math.log(contrast))


def cell_smooth_normal(coords, radii, seed):
    """Smooth per-cell field with zero mean and unit standard deviation
    (random-harmonic construction at the cell centres)."""
    cc, rm = cell_centres(coords, radii)
    f = shell_field(cc, rm, seed, radii[0], radii[-1])
    return (f - f.mean()) / f.std()
