"""GPU parity of the geometry-aware linear transfers (tn_prolongate_linear,
tn_restrict_linear) against the oracle (oracle/linear_transfer.py), plus the
affine exactness on the GPU at a larger level."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import linear_transfer as LT  # noqa: E402


def _pair(lat, rad):
    import terraneo_amd as tn

    mf, mc = tn.Mesh(lat, rad, 0.55, 1.0), tn.Mesh(lat - 1, rad - 1, 0.55, 1.0)
    of, oc = O.Mesh(lat, mf.nr, 0.55, 1.0), O.Mesh(lat - 1, mc.nr, 0.55, 1.0)
    return mf, mc, of, oc


@pytest.mark.gpu
@pytest.mark.parametrize("ncomp,quirk", [(1, False), (3, False), (3, True)])
@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_linear_transfers_match_oracle(cuda, ncomp, quirk, dt):
    import torch

    import terraneo_amd as tn

    mf, mc, of, oc = _pair(3, 3)
    tdt, tol = (torch.float64, 1e-14) if dt == "f64" else (torch.float32, 2e-6)
    y = consistent_random(oc, ncomp, 5).reshape(-1)
    x = consistent_random(of, ncomp, 6).reshape(-1)
    # prolongation, Replace and Add
    ref = LT.prolongate(of.coords, of.radii, y, oc.n, oc.nr, ncomp, vec_early_return=quirk).reshape(-1)
    xf = torch.full((ncomp * mf.N,), 7.0, dtype=tdt, device=cuda)
    tn.api.prolongate_linear(mc, mf, torch.from_numpy(y).to(cuda, tdt), xf, ncomp=ncomp, vec_early_return=quirk)
    assert np.abs(xf.double().cpu().numpy() - ref).max() <= tol * np.abs(ref).max()
    base = consistent_random(of, ncomp, 8).reshape(-1)
    ref_add = LT.prolongate(of.coords, of.radii, y, oc.n, oc.nr, ncomp, xf=base, vec_early_return=quirk).reshape(-1)
    xf = torch.from_numpy(base).to(cuda, tdt)
    tn.api.prolongate_linear(mc, mf, torch.from_numpy(y).to(cuda, tdt), xf, ncomp=ncomp, add=True,
                             vec_early_return=quirk)
    assert np.abs(xf.double().cpu().numpy() - ref_add).max() <= tol * np.abs(ref_add).max()
    # restriction
    ref_r = LT.restrict(of.coords, of.radii, x, of.owned, oc.n, oc.nr, ncomp, oc.interface_sum,
                        vec_early_return=quirk).reshape(-1)
    xc = torch.full((ncomp * mc.N,), 7.0, dtype=tdt, device=cuda)
    tn.api.restrict_linear(mf, mc, torch.from_numpy(x).to(cuda, tdt), xc, ncomp=ncomp, vec_early_return=quirk)
    assert np.abs(xc.double().cpu().numpy() - ref_r).max() <= tol * np.abs(ref_r).max()


@pytest.mark.gpu
def test_linear_prolongation_exact_for_affine_fields(cuda):
    """Level 6 <- 5, 65 radial nodes: P applied to an affine field's coarse samples
    reproduces its fine samples (the transfers' defining property)."""
    import torch

    import terraneo_amd as tn

    mf, mc = tn.Mesh(6, 6, 0.55, 1.0), tn.Mesh(5, 5, 0.55, 1.0)
    lat = np.asarray(mf.lateral_coords()).reshape(10, mf.n, mf.n, 3)
    rad = np.asarray(mf.radii())
    xyz = lat[:, :, :, None, :] * rad[None, None, None, :, None]
    a, c0 = np.array([[0.3, -0.7, 0.5], [1.0, 0.2, -0.4], [-0.6, 0.9, 0.1]]), np.array([0.2, -1.0, 0.5])
    f_fine = np.stack([xyz @ a[k] + c0[k] for k in range(3)])          # [3, 10, nf, nf, nrf]
    f_coarse = np.ascontiguousarray(f_fine[:, :, ::2, ::2, ::2]).reshape(-1)
    out = torch.empty(3 * mf.N, dtype=torch.float64, device=cuda)
    tn.api.prolongate_linear(mc, mf, torch.from_numpy(f_coarse).to(cuda), out, ncomp=3)
    assert np.abs(out.cpu().numpy() - f_fine.reshape(-1)).max() <= 1e-13 * np.abs(f_fine).max()
