"""Deterministic synthetic fields on the shell (test helpers).

Fields are *consistent*: every copy of a duplicated interface node carries
the same value, as the reference's vectors do after an additive exchange.
"""
import numpy as np


def node_xyz(coords, radii):
    return coords[:, :, :, None, :] * radii[None, None, None, :, None]


def smooth_eta(coords, radii, contrast=100.0):
    """Positive nodal viscosity, smooth in space, max/min ~ contrast."""
    X = node_xyz(coords, radii)
    r = np.linalg.norm(X, axis=-1)
    s = np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]) + X[..., 2] * (r - r.min())
    s = (s - s.min()) / (s.max() - s.min() + 1e-300)
    return np.exp(np.log(contrast) * s)


def radial_eta(radii, shape, r_min=0.55, r_max=1.0, contrast=100.0):
    """eta = exp(ln(contrast) (r - r_min)/(r_max - r_min)) -- the C0 viscosity."""
    e = np.exp(np.log(contrast) * (radii - r_min) / (r_max - r_min))
    return np.broadcast_to(e[None, None, None, :], shape).copy()


def consistent_random(omesh, ncomp, seed):
    """Random values per unique node (shared by all copies)."""
    rng = np.random.default_rng(seed)
    g = omesh.groups.reshape(-1)                       # lateral group id per stored lateral node
    vals = rng.uniform(-1, 1, size=(ncomp, omesh.ngroups, omesh.nr))
    out = vals[:, g, :]                                # [ncomp, 10*n*n, nr]
    return np.ascontiguousarray(out.reshape(ncomp, -1))
