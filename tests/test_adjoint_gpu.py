"""GPU parity of the per-cell viscosity mode and the adjoint viscosity-gradient
kernel against the oracle (FP64, 1e-12 relative)."""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random

BCS = [(0, 0), (1, 1), (2, 2), (2, 1), (1, 2)]


def _setup(lat=3, rad=3, seed=0):
    import terraneo_amd as tn

    m = tn.Mesh(lat, rad, 0.55, 1.0)
    om = O.Mesh(lat, m.nr, 0.55, 1.0)
    cells = 10 * (om.n - 1) ** 2 * (om.nr - 1)
    rng = np.random.default_rng(seed)
    kc = np.exp(rng.uniform(-2.0, 2.0, cells))
    return m, om, kc


@pytest.mark.gpu
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("diag", [False, True])
def test_cell_viscosity_apply_matches_oracle(cuda, bc, diag):
    import torch

    import terraneo_amd as tn

    print(f"case bc={bc} diag={diag}", flush=True)
    m, om, kc = _setup()
    x = consistent_random(om, 3, 1).reshape(-1)
    ref = O.Eps(om, None, *bc, kc=kc).apply(x, diagonal=diag)
    kcd = torch.from_numpy(kc).to(cuda)
    op = tn.Eps(m, None, *bc, eta_cell=kcd)
    op.set_diagonal(diag)
    xd = torch.from_numpy(x).to(cuda)
    yd = torch.empty_like(xd)
    op.apply(xd, yd)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


@pytest.mark.gpu
@pytest.mark.parametrize("bc", BCS)
def test_cell_viscosity_stokes_apply_matches_oracle(cuda, bc):
    """[A(eta_cell) B^T; B 0] applied on the GPU vs the oracle's operators."""
    import torch

    import terraneo_amd as tn

    m, om, kc = _setup()
    mp = tn.Mesh(2, 2, 0.55, 1.0)
    op = O.Mesh(2, mp.nr, 0.55, 1.0)
    u = consistent_random(om, 3, 5).reshape(-1)
    p = consistent_random(op, 1, 6).reshape(-1)
    oA = O.Eps(om, None, *bc, kc=kc)
    ref_u = oA.apply(u) + O.grad_apply_add(om, op, p, np.zeros(3 * om.N), *bc)
    ref_p = O.div_apply(om, op, u, *bc)
    A = tn.Eps(m, None, *bc, eta_cell=torch.from_numpy(kc).to(cuda))
    K = tn.Stokes(m, mp, A)
    yu = torch.empty(3 * m.N, dtype=torch.float64, device=cuda)
    yp = torch.empty(mp.N, dtype=torch.float64, device=cuda)
    K.apply(torch.from_numpy(u).to(cuda), torch.from_numpy(p).to(cuda), yu, yp)
    assert np.abs(yu.cpu().numpy() - ref_u).max() <= 1e-12 * np.abs(ref_u).max()
    assert np.abs(yp.cpu().numpy() - ref_p).max() <= 1e-12 * np.abs(ref_p).max()


@pytest.mark.gpu
def test_gca_from_cell_viscosity_matches_oracle(cuda):
    import torch

    import terraneo_amd as tn

    m, om, kc = _setup()
    mc = tn.Mesh(2, 2, 0.55, 1.0)
    oc = O.Mesh(2, mc.nr, 0.55, 1.0)
    fine = tn.Eps(m, None, eta_cell=torch.from_numpy(kc).to(cuda))
    coarse = tn.Eps(mc, torch.ones(mc.N, dtype=torch.float64, device=cuda))
    coarse.gca_from(fine)
    ref = O.gca_build(O.Eps(om, None, kc=kc), oc).reshape(10, oc.n - 1, oc.n - 1, oc.nr - 1, 2, 18, 18)
    for (d, x, y, r) in [(0, 0, 0, 0), (3, 1, 2, 1), (9, 3, 3, 3)]:
        for w in range(2):
            A = coarse.local_matrix(d, x, y, r, w)
            assert np.abs(A - ref[d, x, y, r, w]).max() <= 1e-12 * np.abs(ref[d, x, y, r, w]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 2), (1, 2)])
def test_grad_log_cell_viscosity_matches_oracle(cuda, bc):
    import torch

    import terraneo_amd as tn

    m, om, kc = _setup(seed=2)
    u = consistent_random(om, 3, 3).reshape(-1)
    lam = consistent_random(om, 3, 4).reshape(-1)
    ref = O.Eps(om, None, *bc, kc=kc).grad_log_cell_viscosity(u, lam)
    op = tn.Eps(m, None, *bc, eta_cell=torch.from_numpy(kc).to(cuda))
    g = torch.empty(kc.size, dtype=torch.float64, device=cuda)
    op.grad_log_cell_eta(torch.from_numpy(u).to(cuda), torch.from_numpy(lam).to(cuda), g)
    err = np.abs(g.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def _adjoint_setup(cuda, bc, lat=3, rad=3, levels=2):
    import torch

    from terraneo_amd import api
    from terraneo_amd import workloads as W
    from terraneo_amd.problem import StokesProblem

    pr = StokesProblem(lat, rad, levels, bc=bc, device=cuda,
                       gmg=api.gmg_options(cheby_order=3, cheby_iterations=2, coarse_solver="dense"))
    coords, radii = pr.coords()
    m0 = W.cell_log_viscosity(coords, radii, 4, 1e3)
    xi = W.cell_smooth_normal(coords, radii, 5)
    f = torch.from_numpy(W.buoyancy(coords, radii)).to(cuda)
    pr.set_viscosity(eta_cell=torch.from_numpy(np.exp(m0 + 0.1 * xi)).to(cuda))
    bu, bp = pr.rhs_from_force(f)
    uobs, _, info = pr.solve(bu, bp, rtol=1e-13, max_iterations=1500, pressure_mean_free=True)
    print(f"observation solve: {info['iterations']} its, rel. res "
          f"{info['final_residual'] / info['initial_residual']:.2e}", flush=True)
    own = torch.from_numpy(pr.vel.owned_mask().reshape(10, pr.vel.n, pr.vel.n, pr.vel.nr).astype(bool)).to(cuda)
    surf = torch.zeros_like(own)
    surf[..., -1] = True
    mask = (own & surf).reshape(-1).repeat(3).to(torch.float64)
    return pr, m0, bu, bp, uobs.clone(), mask, coords, radii


def _misfit(u, uobs, mask):
    return 0.5 * (mask * (u - uobs) ** 2).sum()


@pytest.mark.gpu
def test_adjoint_gradient_fd_and_cpu_reference(cuda):
    """dJ/dm of J = 1/2 sum_surface |u - u_obs|^2 (Dirichlet CMB, free-slip
    surface): autograd (adjoint solve + gradient kernel) vs central finite
    differences along smooth directions, and vs the gradient of the oracle's
    assembled system solved directly on the CPU."""
    import torch

    import stokes_direct as SD
    from terraneo_amd import workloads as W
    from terraneo_amd.adjoint import stokes_velocity

    bc = (1, 2)
    pr, m0, bu, bp, uobs, mask, coords, radii = _adjoint_setup(cuda, bc)
    # mean-free pressure: the free-slip surface (nodal normals) leaves the constant
    # pressure a near-null mode of K; the CPU reference borders the same constraint
    kw = dict(rtol=1e-13, max_iterations=1500, pressure_mean_free=True)
    m = torch.from_numpy(m0).to(cuda).requires_grad_(True)
    u = stokes_velocity(m, pr, bu, bp, **kw)
    J = _misfit(u, uobs, mask)
    J.backward()
    g = m.grad.detach().cpu().numpy().reshape(-1)

    # CPU reference: direct solves of the oracle's assembled Stokes system
    om = O.Mesh(3, 9, 0.55, 1.0)
    op = O.Mesh(2, 5, 0.55, 1.0)
    ope = O.Eps(om, None, *bc, kc=np.exp(m0).reshape(-1))
    K, nu, npr = SD.stokes_matrix(om, op, ope, bc)
    buh, bph = bu.cpu().numpy(), bp.cpu().numpy()
    uc, pc = SD.solve_unique(K, nu, npr, om, op, buh, bph)
    res_cpu_on_gpu = pr.residual(torch.from_numpy(uc).to(cuda), torch.from_numpy(pc).to(cuda), bu, bp)
    du = np.abs(uc - u.detach().cpu().numpy()).max() / np.abs(uc).max()
    print(f"GPU operator residual of the CPU direct solution: {res_cpu_on_gpu:.3e}; "
          f"GPU vs CPU velocity: {du:.3e}", flush=True)
    assert np.abs(uc - u.detach().cpu().numpy()).max() <= 1e-8 * np.abs(uc).max()
    mk = mask.cpu().numpy()
    gu = mk * (uc - uobs.cpu().numpy())
    # owned-inner-product representative: sum over node copies
    gu_c = O.from_unique(om, _sum_copies(om, gu), 3)
    lu, _ = SD.solve_unique(K, nu, npr, om, op, gu_c, np.zeros_like(bph))
    gref = ope.grad_log_cell_viscosity(uc, lu)
    err = np.abs(g - gref).max() / np.abs(gref).max()
    print(f"adjoint vs CPU reference: {err:.3e}")
    assert err <= 1e-7, err

    # central finite differences along smooth directions (step 1e-5: the
    # solves stop at 1e-13, so the difference quotient keeps ~1e-8 of noise)
    h = 1e-5
    for s in range(4):
        dvec = W.cell_smooth_normal(coords, radii, 100 + s).reshape(-1)
        Jp, Jm = [], []
        for sgn, acc in ((1.0, Jp), (-1.0, Jm)):
            mm = torch.from_numpy(m0.reshape(-1) + sgn * h * dvec).to(cuda)
            pr.set_viscosity(eta_cell=torch.exp(mm).view(m.shape))
            uu, _, _ = pr.solve(bu, bp, **kw)
            acc.append(_misfit(uu, uobs, mask).item())
        fd = (Jp[0] - Jm[0]) / (2 * h)
        an = float(g @ dvec)
        print(f"direction {s}: adjoint {an:.10e} fd {fd:.10e} rel {abs(an - fd) / abs(fd):.2e}")
        assert abs(an - fd) <= 1e-5 * abs(fd)


def _sum_copies(om, v):
    """unique-dof vector whose entries are the sums over each node's copies"""
    vv = np.asarray(v).reshape(3, 10 * om.n * om.n, om.nr)
    g = om.groups.reshape(-1)
    out = np.zeros((om.ngroups, om.nr, 3))
    np.add.at(out, g, np.moveaxis(vv, 0, -1))
    return out.reshape(-1)
