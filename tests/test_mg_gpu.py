"""GPU parity of the multigrid building blocks against the CPU oracle.

GCA coarse matrices, the stored-matrix (coarse level) apply with every
boundary-condition pair, and the constant prolongation / restriction.
FP64 bar: 1e-12 relative; FP32 storage: 1e-6 relative.
"""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random, smooth_eta

BCS = [(0, 0), (1, 1), (2, 2), (2, 1), (1, 2)]


def _pair(cuda, lat=3, rad=3, dtype=None):
    import torch

    import terraneo_amd as tn

    dtype = dtype or torch.float64
    mf, mc = tn.Mesh(lat, rad, 0.55, 1.0), tn.Mesh(lat - 1, rad - 1, 0.55, 1.0)
    of, oc = O.Mesh(lat, mf.nr, 0.55, 1.0), O.Mesh(lat - 1, mc.nr, 0.55, 1.0)
    eta_f = smooth_eta(of.coords, of.radii).reshape(-1)
    eta_c = smooth_eta(oc.coords, oc.radii).reshape(-1)
    return dict(mf=mf, mc=mc, of=of, oc=oc, eta_f=eta_f, eta_c=eta_c,
                ef=torch.from_numpy(eta_f).to(cuda, dtype), ec=torch.from_numpy(eta_c).to(cuda, dtype))


@pytest.mark.gpu
@pytest.mark.parametrize("store", ["f64", "f32"])
def test_gca_matches_oracle(cuda, store):
    import terraneo_amd as tn

    s = _pair(cuda)
    fine = tn.Eps(s["mf"], s["ef"])
    coarse = tn.Eps(s["mc"], s["ec"])
    coarse.gca_from(fine, tn.TN_F64 if store == "f64" else tn.TN_F32)
    ref = O.gca_build(O.Eps(s["of"], s["eta_f"]), s["oc"]).reshape(10, s["oc"].n - 1, s["oc"].n - 1,
                                                                    s["oc"].nr - 1, 2, 18, 18)
    tol = 1e-12 if store == "f64" else 1e-6
    rng = np.random.default_rng(0)
    cells = [(0, 0, 0, 0), (9, 3, 3, 3), (4, 1, 2, 1)] + [tuple(rng.integers(0, k) for k in (10, 4, 4, 4))
                                                          for _ in range(6)]
    for (d, x, y, r) in cells:
        for w in range(2):
            A = coarse.local_matrix(d, x, y, r, w)
            R = ref[d, x, y, r, w]
            err = np.abs(A - R).max() / np.abs(R).max()
            assert err < tol, (d, x, y, r, w, err)


@pytest.mark.gpu
def test_gca_two_levels_from_stored(cuda):
    """GCA of a GCA level (fine operator itself stored)."""
    import torch

    import terraneo_amd as tn

    s = _pair(cuda, 3, 3)
    m1 = tn.Mesh(1, 1, 0.55, 1.0)
    o1 = O.Mesh(1, m1.nr, 0.55, 1.0)
    fine = tn.Eps(s["mf"], s["ef"])
    mid = tn.Eps(s["mc"], s["ec"])
    mid.gca_from(fine)
    e1 = torch.ones(m1.N, dtype=torch.float64, device=cuda)
    low = tn.Eps(m1, e1)
    low.gca_from(mid)
    omid = O.Eps(s["oc"], s["eta_c"])
    omid.set_stored(O.gca_build(O.Eps(s["of"], s["eta_f"]), s["oc"]))
    ref = O.gca_build(omid, o1).reshape(10, 2, 2, 2, 2, 18, 18)
    for d in (0, 5, 9):
        for x, y, r in [(0, 0, 0), (1, 1, 1), (1, 0, 1)]:
            for w in range(2):
                A = low.local_matrix(d, x, y, r, w)
                R = ref[d, x, y, r, w]
                assert np.abs(A - R).max() / np.abs(R).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("diag", [False, True])
def test_stored_apply_matches_oracle(cuda, bc, diag):
    import torch

    import terraneo_amd as tn

    s = _pair(cuda)
    fine = tn.Eps(s["mf"], s["ef"])
    coarse = tn.Eps(s["mc"], s["ec"], *bc)
    coarse.gca_from(fine)
    coarse.set_diagonal(diag)
    oc = O.Eps(s["oc"], s["eta_c"], *bc)
    oc.set_stored(O.gca_build(O.Eps(s["of"], s["eta_f"]), s["oc"]))
    x = consistent_random(s["oc"], 3, 5).reshape(-1)
    ref = oc.apply(x, diagonal=diag)
    xd = torch.from_numpy(x).to(cuda)
    yd = torch.empty_like(xd)
    coarse.apply(xd, yd)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.gpu
def test_stored_apply_fp32_storage(cuda):
    import torch

    import terraneo_amd as tn

    s = _pair(cuda)
    fine = tn.Eps(s["mf"], s["ef"])
    coarse = tn.Eps(s["mc"], s["ec"].float(), 1, 1)
    coarse.gca_from(fine, tn.TN_F32)
    oc = O.Eps(s["oc"], s["eta_c"], 1, 1)
    oc.set_stored(O.gca_build(O.Eps(s["of"], s["eta_f"]), s["oc"]))
    x = consistent_random(s["oc"], 3, 6).reshape(-1)
    ref = oc.apply(x)
    xd = torch.from_numpy(x).float().to(cuda)
    yd = torch.empty_like(xd)
    coarse.apply(xd, yd)
    err = np.abs(yd.double().cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


@pytest.mark.gpu
def test_transfers_match_oracle(cuda):
    import torch

    import terraneo_amd as tn

    s = _pair(cuda, 3, 3)
    xc = consistent_random(s["oc"], 3, 7).reshape(-1)
    xf0 = consistent_random(s["of"], 3, 8).reshape(-1)
    ref_p = O.prolongate_add(s["oc"], s["of"], xc, xf0)
    xf = torch.from_numpy(xf0).to(cuda)
    tn.api.prolongate_add(s["mc"], s["mf"], torch.from_numpy(xc).to(cuda), xf)
    assert np.abs(xf.cpu().numpy() - ref_p).max() < 1e-14 * np.abs(ref_p).max()
    ref_r = O.restrict(s["of"], s["oc"], xf0)
    xcd = torch.empty(3 * s["mc"].N, dtype=torch.float64, device=cuda)
    tn.api.restrict(s["mf"], s["mc"], torch.from_numpy(xf0).to(cuda), xcd)
    assert np.abs(xcd.cpu().numpy() - ref_r).max() < 1e-13 * np.abs(ref_r).max()
