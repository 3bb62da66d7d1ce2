"""Oracle checks for the per-cell viscosity mode and the adjoint gradient.

The gradient g[cell] = -kc[cell] sum_wedges (T lam)^T A_e(1) u is pinned by
central finite differences of the bilinear form lam^T A_N(exp m) u (the
Neumann operator, lam with the trial rule applied), and the cell mode by the
constant-viscosity identity with the nodal operator.
"""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random, smooth_eta


def _mesh():
    return O.Mesh(2, 5, 0.55, 1.0)


def test_cell_viscosity_constant_matches_nodal():
    om = _mesh()
    x = consistent_random(om, 3, 0).reshape(-1)
    cells = 10 * (om.n - 1) ** 2 * (om.nr - 1)
    for bc in [(0, 0), (1, 1), (2, 2), (2, 1)]:
        nod = O.Eps(om, np.full(om.N, 2.5), *bc)
        cel = O.Eps(om, None, *bc, kc=np.full(cells, 2.5))
        a, b = nod.apply(x), cel.apply(x)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
        a, b = nod.apply(x, diagonal=True), cel.apply(x, diagonal=True)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()


def _trial(om, lam, bc):
    """Dirichlet: zero the shell's nodes, free-slip: project them."""
    lam = lam.reshape(3, 10, om.n, om.n, om.nr).copy()
    nrm = om.coords / np.linalg.norm(om.coords, axis=-1, keepdims=True)
    for lev, mode in ((0, bc[0]), (om.nr - 1, bc[1])):
        v = lam[..., lev]
        if mode == O.DIRICHLET:
            v[:] = 0.0
        elif mode == O.FREESLIP:
            ln = np.einsum("cdxy,dxyc->dxy", v, nrm)
            v -= np.moveaxis(nrm, -1, 0) * ln[None]
    return lam.reshape(-1)


@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 2)])
def test_adjoint_gradient_matches_finite_differences(bc):
    om = _mesh()
    cells = 10 * (om.n - 1) ** 2 * (om.nr - 1)
    rng = np.random.default_rng(3)
    m = rng.uniform(-1.0, 1.0, cells)
    u = consistent_random(om, 3, 1).reshape(-1)
    lam = consistent_random(om, 3, 2).reshape(-1)
    g = O.Eps(om, None, *bc, kc=np.exp(m)).grad_log_cell_viscosity(u, lam)
    lt = _trial(om, lam, bc)

    def J(mm):
        return om.dot(lt, O.Eps(om, None, 0, 0, kc=np.exp(mm)).apply(u), 3)

    h = 1e-5
    for c in list(rng.choice(cells, 6, replace=False)) + [0, cells - 1]:
        e = np.zeros(cells)
        e[c] = h
        fd = (J(m + e) - J(m - e)) / (2 * h)
        assert abs(-g[c] - fd) <= 1e-8 * max(abs(fd), 1e-3 * np.abs(g).max()), (c, -g[c], fd)
    dvec = rng.standard_normal(cells)
    fd = (J(m + h * dvec) - J(m - h * dvec)) / (2 * h)
    assert abs(-g @ dvec - fd) <= 1e-8 * abs(fd)
