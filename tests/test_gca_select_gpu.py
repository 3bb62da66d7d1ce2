"""GPU parity of selective GCA: the element collector (tn_gca_collect_elements)
and the selective operator (tn_eps_gca_select + the stored apply) against the
oracle (oracle/gca_select.py over the C oracle's GCA and local matrices)."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import gca_select as GS  # noqa: E402


def _jump_eta(om):
    """exp of a 30x radial jump at r = 0.8 plus a lateral band: large gradients near the jumps only."""
    xyz = om.node_xyz()
    r = np.linalg.norm(xyz, axis=-1)
    eta = np.where(r > 0.8, 30.0, 1.0) * np.where(xyz[..., 2] > 0.3, 5.0, 1.0)
    return np.ascontiguousarray(eta.reshape(-1))


@pytest.mark.gpu
def test_collect_matches_oracle(cuda):
    import torch

    import terraneo_amd as tn

    mf, m0 = tn.Mesh(3, 3, 0.55, 1.0), tn.Mesh(1, 1, 0.55, 1.0)
    of = O.Mesh(3, mf.nr, 0.55, 1.0)
    k = _jump_eta(of)
    ref = GS.collect(k, mf.n, mf.nr, 2, m0.n, m0.nr).reshape(-1)
    assert 0 < ref.sum() < 10 * 2 * 2 * 2  # some, not all, coarsest cells
    flags = torch.full((m0.N,), 5.0, dtype=torch.float64, device=cuda)
    tn.api.gca_collect_elements(mf, m0, torch.from_numpy(k).to(cuda), flags)
    assert np.array_equal(flags.cpu().numpy(), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 2)])
def test_selective_operator_matches_oracle(cuda, bc):
    import torch

    import terraneo_amd as tn

    mf, mc, m0 = tn.Mesh(3, 3, 0.55, 1.0), tn.Mesh(2, 2, 0.55, 1.0), tn.Mesh(1, 1, 0.55, 1.0)
    of, oc = O.Mesh(3, mf.nr, 0.55, 1.0), O.Mesh(2, mc.nr, 0.55, 1.0)
    eta_f = _jump_eta(of)
    eta_c = np.ascontiguousarray(eta_f.reshape(10, mf.n, mf.n, mf.nr)[:, ::2, ::2, ::2].reshape(-1))
    flags = GS.collect(eta_f, mf.n, mf.nr, 2, m0.n, m0.nr)
    # device side: GCA from the fine operator, then the selective replacement
    fine = tn.Eps(mf, torch.from_numpy(eta_f).to(cuda), *bc)
    coarse = tn.Eps(mc, torch.from_numpy(eta_c).to(cuda), *bc)
    coarse.gca_from(fine)
    fl = torch.from_numpy(flags.reshape(-1)).to(cuda)
    coarse.gca_select(m0, fl)
    # oracle: Galerkin matrices on flagged ancestors, the coarse level's own elsewhere
    oce = O.Eps(oc, eta_c, *bc)
    gca = O.gca_build(O.Eps(of, eta_f), oc)
    mats = GS.select(gca, oce.local_matrix, flags, oc.n, oc.nr, 1)
    mixed = ((mats != gca.reshape(mats.shape)).any(axis=(4, 5, 6))).mean()
    assert 0.0 < mixed < 1.0  # both kinds of cells are present
    oce.set_stored(mats)
    x = consistent_random(oc, 3, 4).reshape(-1)
    ref = oce.apply(x)
    y = torch.empty(3 * mc.N, dtype=torch.float64, device=cuda)
    coarse.apply(torch.from_numpy(x).to(cuda), y)
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
