"""GPU parity of the Stokes block (B, B^T, A) and the mass operators against
the CPU oracle, and the assess Stokes solve against the L2 errors the
reference printed (tests/test_epsilon_divdiv_stokes_assess.cpp).
"""
import numpy as np
import pytest

import assess as AS
import oracle_lib as O
from fields import consistent_random, smooth_eta

BCS = [(0, 0), (1, 1), (2, 2), (2, 1), (1, 2)]


def _meshes(lat=3, rmin=0.55, rmax=1.0):
    import terraneo_amd as tn

    mv, mp = tn.Mesh(lat, lat, rmin, rmax), tn.Mesh(lat - 1, lat - 1, rmin, rmax)
    ov, op = O.Mesh(lat, mv.nr, rmin, rmax), O.Mesh(lat - 1, mp.nr, rmin, rmax)
    return mv, mp, ov, op


@pytest.mark.gpu
@pytest.mark.parametrize("bc", BCS)
def test_div_grad_match_oracle(cuda, bc):
    import torch

    import terraneo_amd as tn

    mv, mp, ov, op = _meshes()
    eta = smooth_eta(ov.coords, ov.radii).reshape(-1)
    A = tn.Eps(mv, torch.from_numpy(eta).to(cuda), *bc)
    K = tn.Stokes(mv, mp, A)
    u = consistent_random(ov, 3, 1).reshape(-1)
    p = consistent_random(op, 1, 2).reshape(-1)
    ref_b = O.div_apply(ov, op, u, *bc)
    pd = torch.empty(mp.N, dtype=torch.float64, device=cuda)
    K.div(torch.from_numpy(u).to(cuda), pd)
    assert np.abs(pd.cpu().numpy() - ref_b).max() <= 1e-12 * np.abs(ref_b).max()
    ref_g = O.grad_apply_add(ov, op, p, np.zeros(3 * ov.N), *bc)
    ud = torch.empty(3 * mv.N, dtype=torch.float64, device=cuda)
    K.grad(torch.from_numpy(p).to(cuda), ud)
    assert np.abs(ud.cpu().numpy() - ref_g).max() <= 1e-12 * np.abs(ref_g).max()
    # Stokes apply = [A u + B^T p ; B u]
    yu, yp = torch.empty_like(ud), torch.empty_like(pd)
    K.apply(torch.from_numpy(u).to(cuda), torch.from_numpy(p).to(cuda), yu, yp)
    ref_u = O.Eps(ov, eta, *bc).apply(u) + ref_g
    assert np.abs(yu.cpu().numpy() - ref_u).max() <= 1e-12 * np.abs(ref_u).max()
    assert np.abs(yp.cpu().numpy() - ref_b).max() <= 1e-12 * np.abs(ref_b).max()


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(1, 1), (2, 2)])
def test_div_two_radial_chunks_repeats_bitwise(cuda, bc):
    """B with 128 coarse radial cells (two 64-cell chunks: the chunk-top sums
    meet the next chunk's first level) against the oracle, and a second apply
    equal bit for bit (no atomics in B)."""
    import torch

    import terraneo_amd as tn

    mv, mp = tn.Mesh(3, 8, 0.55, 1.0), tn.Mesh(2, 7, 0.55, 1.0)
    ov, op = O.Mesh(3, mv.nr, 0.55, 1.0), O.Mesh(2, mp.nr, 0.55, 1.0)
    assert mp.nr - 1 == 128
    eta = smooth_eta(ov.coords, ov.radii).reshape(-1)
    K = tn.Stokes(mv, mp, tn.Eps(mv, torch.from_numpy(eta).to(cuda), *bc))
    u = consistent_random(ov, 3, 5).reshape(-1)
    ud = torch.from_numpy(u).to(cuda)
    p1 = torch.empty(mp.N, dtype=torch.float64, device=cuda)
    p2 = torch.full_like(p1, 7.0)
    K.div(ud, p1)
    K.div(ud, p2)
    assert torch.equal(p1, p2)
    ref = O.div_apply(ov, op, u, *bc)
    assert np.abs(p1.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
def test_mass_operators_match_oracle(cuda):
    import torch

    import terraneo_amd as tn

    mv, mp, ov, op = _meshes()
    x = consistent_random(ov, 3, 3).reshape(-1)
    ref = O.vector_mass(ov, x)
    y = torch.empty(3 * mv.N, dtype=torch.float64, device=cuda)
    tn.api.vector_mass(mv, torch.from_numpy(x).to(cuda), y)
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    k = smooth_eta(op.coords, op.radii).reshape(-1)
    refd = O.kmass_lumped(op, k)
    d = torch.empty(mp.N, dtype=torch.float64, device=cuda)
    tn.api.kmass_lumped(mp, torch.from_numpy(k).to(cuda), d)
    assert np.abs(d.cpu().numpy() - refd).max() <= 1e-13 * np.abs(refd).max()


def assess_solve_gpu(cuda, level, min_level=2, kind="zeroslip", restart=30):
    """Reference assess tests on the GPU: EpsDivDivStokes, FGMRES(30) rtol 1e-14,
    block-triangular preconditioner, V-cycle x2 with Chebyshev(2) x3, constant
    transfers, re-discretised coarse operators, PCG coarsest (1000, 1e-6),
    Schur = lumped mass (eta = 1).
      zeroslip      Dirichlet both shells, lifting of the exact boundary values
                    (test_epsilon_divdiv_stokes_assess.cpp)
      freeslip      free-slip both shells with the rotation penalty, RHS zeroed
                    on both shells (test_epsilon_divdiv_stokes_assess_freeslip_freeslip.cpp:592-606)
      freezeroslip  free-slip CMB (RHS normal part removed), zero-slip surface
                    (RHS zeroed) (test_epsilon_divdiv_stokes_assess_freeslip.cpp:592-610)"""
    import torch

    import terraneo_amd as tn

    sol = AS.Solution(kind)
    bc = {"zeroslip": (tn.TN_DIRICHLET, tn.TN_DIRICHLET), "freeslip": (tn.TN_FREESLIP, tn.TN_FREESLIP),
          "freezeroslip": (tn.TN_FREESLIP, tn.TN_DIRICHLET)}[kind]
    rmin, rmax = AS.R_MIN, AS.R_MAX
    meshes = [tn.Mesh(l, l, rmin, rmax) for l in range(min_level, level + 1)]
    mv, mp = meshes[-1], tn.Mesh(level - 1, level - 1, rmin, rmax)
    ov, op = O.Mesh(level, mv.nr, rmin, rmax), O.Mesh(level - 1, mp.nr, rmin, rmax)
    f64 = dict(dtype=torch.float64, device=cuda)
    uex = torch.from_numpy(np.ascontiguousarray(sol.velocity(ov.node_xyz()).reshape(-1))).to(cuda)
    pex = torch.from_numpy(np.ascontiguousarray(sol.pressure(op.node_xyz()).reshape(-1))).to(cuda)
    fn = torch.from_numpy(np.ascontiguousarray(sol.body_force(ov.node_xyz()).reshape(-1))).to(cuda)
    ones = [torch.ones(m.N, **f64) for m in meshes]
    f = torch.empty(3 * mv.N, **f64)
    tn.api.vector_mass(mv, fn, f)
    zp = torch.zeros(mp.N, **f64)
    if kind == "zeroslip":
        g = torch.zeros(3 * mv.N, **f64)
        tn.api.boundary_assign(mv, g, uex, 3)
        AN = tn.Eps(mv, ones[-1], tn.TN_NEUMANN, tn.TN_NEUMANN)
        KN = tn.Stokes(mv, mp, AN)
        ku, kp = torch.empty_like(f), torch.empty_like(zp)
        KN.apply(g, zp, ku, kp)
        bu = f - ku
        bp = -kp
        AN.set_diagonal(True)
        dg = torch.empty_like(f)
        AN.apply(g, dg)
        tn.api.boundary_assign(mv, bu, dg, 3)
    else:
        bu, bp = f, zp.clone()
        zero = torch.zeros_like(f)
        if kind == "freeslip":
            tn.api.boundary_assign(mv, bu, zero, 3)
        else:
            tn.api.freeslip_project(mv, bu, 1)
            tn.api.boundary_assign(mv, bu, zero, 2)
    ops = [tn.Eps(m, e, *bc) for m, e in zip(meshes, ones)]
    mg = tn.Gmg(ops, tn.api.gmg_options(cheby_order=2, cheby_iterations=3, power_iterations=10, num_cycles=2))
    K = tn.Stokes(mv, mp, ops[-1])
    mass = torch.empty(mp.N, **f64)
    tn.api.kmass_lumped(mp, torch.ones(mp.N, **f64), mass)
    u, p = torch.zeros_like(f), torch.zeros_like(zp)
    info = tn.stokes_solve(K, tn.TN_FGMRES_BLOCKTRI, mg, mass, u, p, bu, bp, max_iterations=200, rtol=1e-14,
                           atol=1e-16, restart=restart)
    own_u = np.tile(mv.owned_mask().reshape(-1).astype(bool), 3)
    own_p = mp.owned_mask().reshape(-1).astype(bool)
    uu, pp, ue, pe = u.cpu().numpy(), p.cpu().numpy(), uex.cpu().numpy(), pex.cpu().numpy()
    pp = pp - pp[own_p].mean()
    pe = pe - pe[own_p].mean()
    ev = np.sqrt(np.sum((uu - ue)[own_u] ** 2) / own_u.sum())
    ep = np.sqrt(np.sum((pp - pe)[own_p] ** 2) / own_p.sum())
    return ev, ep, info


@pytest.mark.gpu
@pytest.mark.parametrize("level", [3, 4])
def test_assess_stokes_solve_matches_reference_errors(cuda, level):
    ev, ep, info = assess_solve_gpu(cuda, level)
    rv, rp = AS.REFERENCE_ERRORS_DIRICHLET[level]
    print(f"level {level}: vel {ev:.6e} (ref {rv}) pre {ep:.6e} (ref {rp}) its {info['iterations']} "
          f"res {info['final_residual'] / info['initial_residual']:.2e} t {info['seconds']:.2f}s")
    assert info["final_residual"] / info["initial_residual"] < 1e-12
    assert abs(ev - rv) / rv < 1e-4, (ev, rv)
    assert abs(ep - rp) / rp < 1e-4, (ep, rp)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,level", [("freeslip", 3), ("freeslip", 4), ("freezeroslip", 3), ("freezeroslip", 4)])
def test_assess_freeslip_solve_matches_reference_errors(cuda, kind, level):
    ev, ep, info = assess_solve_gpu(cuda, level, kind=kind)
    ref = AS.REFERENCE_ERRORS_FREESLIP if kind == "freeslip" else AS.REFERENCE_ERRORS_FREEZEROSLIP
    rv, rp = ref[level]
    print(f"{kind} level {level}: vel {ev:.6e} (ref {rv}) pre {ep:.6e} (ref {rp}) its {info['iterations']} "
          f"res {info['final_residual'] / info['initial_residual']:.2e} t {info['seconds']:.2f}s")
    assert info["final_residual"] / info["initial_residual"] < 1e-12
    assert abs(ev - rv) / rv < 1e-4, (ev, rv)
    assert abs(ep - rp) / rp < 1e-4, (ep, rp)


@pytest.mark.gpu
def test_assess_solve_long_restart(cuda):
    """FGMRES(100): the Hessenberg column kept on the device grows past its initial 64 slots"""
    ev, ep, info = assess_solve_gpu(cuda, 3, restart=100)
    rv, rp = AS.REFERENCE_ERRORS_DIRICHLET[3]
    print(f"restart 100: vel {ev:.6e} (ref {rv}) pre {ep:.6e} (ref {rp}) its {info['iterations']}")
    assert info["final_residual"] / info["initial_residual"] < 1e-12
    assert abs(ev - rv) / rv < 1e-4, (ev, rv)
    assert abs(ep - rp) / rp < 1e-4, (ep, rp)
