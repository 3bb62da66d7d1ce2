"""Pins of the linear-transfer oracle (oracle/linear_transfer.py): exactness for
affine fields (the property grid_transfer_linear.hpp:43-45 states), adjointness
of restriction and prolongation in the owned inner product, and the vector
variants' early return.  CPU only."""
import os
import sys

import numpy as np

import oracle_lib as O
from fields import consistent_random

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import linear_transfer as LT  # noqa: E402


def _meshes(lat=2, rad=2):
    mf = O.Mesh(lat, (1 << rad) + 1, 0.55, 1.0)
    mc = O.Mesh(lat - 1, (1 << (rad - 1)) + 1, 0.55, 1.0)
    return mf, mc


def test_affine_fields_are_interpolated_exactly():
    mf, mc = _meshes()
    a, c0 = np.array([0.3, -0.7, 0.5]), 0.2
    xyz_f = mf.node_xyz()                                            # [10, nf, nf, nrf, 3]
    f_fine = xyz_f @ a + c0
    f_coarse = f_fine[:, ::2, ::2, ::2]                              # coarse nodes are the even fine nodes
    p = LT.prolongate(mf.coords, mf.radii, f_coarse[None], mc.n, mc.nr, 1)[0]
    assert np.abs(p - f_fine).max() <= 1e-13 * np.abs(f_fine).max()
    # a plain midpoint average (the constant transfer) is not exact on the curved shell
    mid = 0.5 * (f_coarse[:, 0, 0, 0] + f_coarse[:, 1, 0, 0])
    assert np.abs(mid - f_fine[:, 1, 0, 0]).max() > 1e-6


def test_restriction_is_the_adjoint():
    mf, mc = _meshes()
    for ncomp in (1, 3):
        x = consistent_random(mf, ncomp, 1).reshape(-1)
        y = consistent_random(mc, ncomp, 2).reshape(-1)
        rx = LT.restrict(mf.coords, mf.radii, x, mf.owned, mc.n, mc.nr, ncomp, mc.interface_sum).reshape(-1)
        py = LT.prolongate(mf.coords, mf.radii, y, mc.n, mc.nr, ncomp).reshape(-1)
        lhs, rhs = mc.dot(rx, y, ncomp), mf.dot(x, py, ncomp)
        assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_vector_early_return():
    mf, mc = _meshes()
    y = consistent_random(mc, 3, 3).reshape(3, -1)
    full = LT.prolongate(mf.coords, mf.radii, y, mc.n, mc.nr, 3)
    quirk = LT.prolongate(mf.coords, mf.radii, y, mc.n, mc.nr, 3, vec_early_return=True)
    # component 0 everywhere, and every component at lateral midpoints, as without the early return
    assert np.array_equal(quirk[0], full[0])
    lateral_mid = np.zeros(full.shape[1:], bool)
    lateral_mid[:, 1::2, :, :] = True
    lateral_mid[:, :, 1::2, :] = True
    assert np.array_equal(quirk[1:, lateral_mid], full[1:, lateral_mid])
    # components 1, 2 stay zero at coincident and radially aligned nodes
    assert np.all(quirk[1:, ~lateral_mid] == 0.0) and np.any(full[1:, ~lateral_mid] != 0.0)
