"""GCACapable side of the operator (operator.hpp:80-112):
get_local_matrix of a matrix-free operator assembles the wedge's Neumann
matrix (kerngen.hpp:481-497, 2051-2108) and equals the oracle's local matrix;
set_stored_matrices (set_stored_matrix_mode Full) reproduces the matrix-free
apply; set_local_matrix / get_local_matrix round-trip and change the apply
exactly as the oracle's stored operator does."""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random, smooth_eta

CELLS = [(0, 0, 0, 0), (3, 1, 2, 1), (9, 3, 3, 3), (5, 2, 0, 2), (7, 0, 3, 0)]


def _setup(cuda, lat=2, rad=2):
    import torch

    import terraneo_amd as tn

    m = tn.Mesh(lat, rad, 0.55, 1.0)
    om = O.Mesh(lat, m.nr, 0.55, 1.0)
    eta = smooth_eta(om.coords, om.radii).reshape(-1)
    return tn, m, om, eta, torch.from_numpy(eta).to(cuda)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_local_matrix_on_the_fly_matches_oracle(cuda, dtype):
    import torch

    tn, m, om, eta, eta_d = _setup(cuda)
    dt = torch.float64 if dtype == "f64" else torch.float32
    op = tn.Eps(m, eta_d.to(dt), tn.TN_DIRICHLET, tn.TN_FREESLIP)
    oref = O.Eps(om, eta)
    tol = 1e-13 if dtype == "f64" else 1e-6
    for (d, x, y, r) in CELLS:
        for w in (0, 1):
            A = op.local_matrix(d, x, y, r, w)
            R = oref.local_matrix(d, x, y, r, w)
            assert np.abs(A - R).max() <= tol * np.abs(R).max(), (d, x, y, r, w)


@pytest.mark.gpu
def test_local_matrix_cell_viscosity(cuda):
    import torch

    tn, m, om, eta, eta_d = _setup(cuda)
    kc = np.exp(np.random.default_rng(4).uniform(-1, 1, 10 * (m.n - 1) ** 2 * (m.nr - 1)))
    op = tn.Eps(m, eta_d, eta_cell=torch.from_numpy(kc).to(cuda))
    lib = O.lib()
    for (d, x, y, r) in CELLS:
        cell = ((d * (m.n - 1) + x) * (m.n - 1) + y) * (m.nr - 1) + r
        for w in (0, 1):
            R = np.zeros(324)
            lib.orc_eps_local_matrix_k(om.h, kc[cell], d, x, y, r, w, R)
            A = op.local_matrix(d, x, y, r, w)
            assert np.abs(A - R.reshape(18, 18)).max() <= 1e-13 * np.abs(R).max()


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 2), (2, 1)])
def test_stored_own_matrices_apply_as_matrix_free(cuda, bc):
    import torch

    tn, m, om, eta, eta_d = _setup(cuda, 3, 3)
    x = torch.from_numpy(consistent_random(om, 3, 2).reshape(-1)).to(cuda)
    mf = tn.Eps(m, eta_d, *bc)
    st = tn.Eps(m, eta_d, *bc)
    st.set_stored_matrices(tn.TN_F64)
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    mf.apply(x, y0)
    st.apply(x, y1)
    err = float((y1 - y0).abs().max() / y0.abs().max())
    assert err <= 1e-12, err


@pytest.mark.gpu
def test_set_local_matrix_round_trip_and_apply(cuda):
    import torch

    tn, m, om, eta, eta_d = _setup(cuda, 2, 2)
    bc = (tn.TN_DIRICHLET, tn.TN_DIRICHLET)
    op = tn.Eps(m, eta_d, *bc)
    with pytest.raises(RuntimeError, match="no stored matrices"):
        op.set_local_matrix(0, 0, 0, 0, 0, np.eye(18))
    op.set_stored_matrices(tn.TN_F64)
    # oracle operator with the same stored matrices
    ncell = 10 * (m.n - 1) ** 2 * (m.nr - 1)
    mats = np.zeros((ncell, 2, 18, 18))
    for cell in range(ncell):
        r = cell % (m.nr - 1)
        rest = cell // (m.nr - 1)
        y, rest = rest % (m.n - 1), rest // (m.n - 1)
        x, d = rest % (m.n - 1), rest // (m.n - 1)
        for w in (0, 1):
            mats[cell, w] = op.local_matrix(d, x, y, r, w)
    rng = np.random.default_rng(9)
    for (d, x, y, r) in CELLS[:3]:
        cell = ((d * (m.n - 1) + x) * (m.n - 1) + y) * (m.nr - 1) + r
        for w in (0, 1):
            B = rng.standard_normal((18, 18))
            M = mats[cell, w] + 0.1 * np.abs(mats[cell, w]).max() * (B + B.T)
            op.set_local_matrix(d, x, y, r, w, M)
            got = op.local_matrix(d, x, y, r, w)
            assert np.abs(got - M).max() == 0.0
            mats[cell, w] = M
    with pytest.raises(RuntimeError, match="not symmetric"):
        op.set_local_matrix(0, 0, 0, 0, 0, np.triu(np.ones((18, 18))))
    oref = O.Eps(om, eta, *bc)
    oref.set_stored(mats.reshape(-1))
    xh = consistent_random(om, 3, 3).reshape(-1)
    ref = oref.apply(xh)
    xd = torch.from_numpy(xh).to(cuda)
    yd = torch.empty_like(xd)
    op.apply(xd, yd)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err
