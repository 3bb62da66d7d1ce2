"""GPU solver recurrences against numpy restatements of the reference's
algorithms over the CPU oracle's operators (tests/solver_restate.py):

  power iteration   the GPU's lambda estimate from its SplitMix64 start
                    vector equals the restatement's (power_iteration.hpp:13-47)
  V-cycle           one V-cycle with fixed Chebyshev bounds (the lambda set
                    through tn_gmg_set_lambda), Chebyshev degree 2 and 3,
                    one and two sweeps, dense coarsest solve, 3 levels, three
                    boundary-condition pairs (chebyshev.hpp:100-157,
                    multigrid.hpp:184-211)
  PMINRES           residual history of the block-diagonally preconditioned
                    Stokes MINRES (V-cycle velocity block, inverse lumped
                    pressure mass) over 12 iterations (pminres.hpp:75-185)
Bars: 1e-10 relative (single applications), 1e-8 (Krylov histories).
"""
import numpy as np
import pytest

import oracle_lib as O
import solver_restate as S
from fields import consistent_random, smooth_eta

LEVELS = [(2, 2), (3, 3), (4, 4)]


def _hierarchy(cuda, bc, levels=LEVELS):
    import torch

    import terraneo_amd as tn

    meshes = [tn.Mesh(l, r, 0.55, 1.0) for l, r in levels]
    oms = [O.Mesh(l, m.nr, 0.55, 1.0) for (l, _), m in zip(levels, meshes)]
    eta = smooth_eta(oms[-1].coords, oms[-1].radii).reshape(-1)
    ops = [None] * len(levels)
    ops[-1] = tn.Eps(meshes[-1], torch.from_numpy(eta).to(cuda), *bc)
    keep = []
    for l in range(len(levels) - 2, -1, -1):
        one = torch.ones(meshes[l].N, dtype=torch.float64, device=cuda)
        keep.append(one)
        ops[l] = tn.Eps(meshes[l], one, *bc)
        ops[l].gca_from(ops[l + 1], tn.TN_F64)
    # oracle: Galerkin matrices from the Neumann cell matrices, boundary rule of the level
    oops = [None] * len(levels)
    oops[-1] = O.Eps(oms[-1], eta, *bc)
    src = O.Eps(oms[-1], eta)
    for l in range(len(levels) - 2, -1, -1):
        mats = O.gca_build(src, oms[l])
        op = O.Eps(oms[l], np.ones(oms[l].N), *bc)
        op.set_stored(mats)
        oops[l] = op
        src = O.Eps(oms[l], np.ones(oms[l].N))
        src.set_stored(mats)
    lv = [S.Level(om, op) for om, op in zip(oms, oops)]
    return meshes, oms, ops, lv, eta, keep


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(1, 1), (2, 2)])
def test_power_iteration_matches_restatement(cuda, bc):
    import torch

    import terraneo_amd as tn

    print(f"case power iteration bc={bc}", flush=True)
    meshes, oms, ops, lv, _, _ = _hierarchy(cuda, bc)
    opts = tn.api.gmg_options(cheby_order=2, cheby_iterations=1, power_iterations=10, coarse_solver="dense",
                              seed=12345)
    mg = tn.Gmg(ops, opts)
    b = torch.from_numpy(consistent_random(oms[-1], 3, 3).reshape(-1)).to(cuda)
    x = torch.zeros_like(b)
    mg.solve(x, b)  # estimates lambda on every smoothed level
    for l in range(1, len(ops)):
        ref = S.power_iteration(lv[l], 12345 + 1000003 * l, 10)
        got = mg.lambda_max(l)
        assert abs(got - ref) <= 1e-10 * abs(ref), (l, got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(1, 1), (2, 2), (2, 1)])
@pytest.mark.parametrize("order,sweeps", [(2, 1), (3, 2)])
def test_vcycle_fixed_lambda_matches_restatement(cuda, bc, order, sweeps):
    import torch

    import terraneo_amd as tn

    print(f"case V-cycle bc={bc} order={order} sweeps={sweeps}", flush=True)
    meshes, oms, ops, lv, _, _ = _hierarchy(cuda, bc)
    opts = tn.api.gmg_options(cheby_order=order, cheby_iterations=sweeps, power_iterations=10, coarse_solver="dense")
    mg = tn.Gmg(ops, opts)
    # fixed bounds near the operators' largest eigenvalues (restated estimate, rounded)
    lams = [0.0] + [round(S.power_iteration(lv[l], 7, 10), 2) for l in (1, 2)]
    for l in (1, 2):
        mg.set_lambda_max(l, lams[l])
    b = consistent_random(oms[-1], 3, 5).reshape(-1)
    x0 = 0.1 * consistent_random(oms[-1], 3, 6).reshape(-1)
    xd = torch.from_numpy(x0.copy()).to(cuda)
    mg.solve(xd, torch.from_numpy(b).to(cuda))
    ref = S.VCycle(lv, lams, order, sweeps)(x0, b)
    got = xd.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= 1e-10, err


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(1, 1), (2, 2)])
def test_pminres_history_matches_restatement(cuda, bc):
    import torch

    import terraneo_amd as tn

    print(f"case PMINRES bc={bc}", flush=True)
    meshes, oms, ops, lv, eta, _ = _hierarchy(cuda, bc, levels=[(1, 1), (2, 2), (3, 3)])
    mf, of = meshes[-1], oms[-1]
    mp = tn.Mesh(2, 2, 0.55, 1.0)
    op_ = O.Mesh(2, mp.nr, 0.55, 1.0)
    K = tn.Stokes(mf, mp, ops[-1])
    opts = tn.api.gmg_options(cheby_order=3, cheby_iterations=1, power_iterations=10, coarse_solver="dense")
    mg = tn.Gmg(ops, opts)
    lams = [0.0] + [round(S.power_iteration(lv[l], 7, 10), 2) for l in (1, 2)]
    for l in (1, 2):
        mg.set_lambda_max(l, lams[l])
    # Schur diagonal: lumped pressure mass weighted by 1/eta at the pressure nodes
    eta_p = eta.reshape(10, of.n, of.n, of.nr)[:, ::2, ::2, ::2].reshape(-1)
    sd = O.kmass_lumped(op_, 1.0 / eta_p)
    bu = consistent_random(of, 3, 8).reshape(3, 10, of.n, of.n, of.nr)
    for shell, mode in ((0, bc[0]), (-1, bc[1])):
        if mode == 1:
            bu[..., shell] = 0.0
    bu = np.ascontiguousarray(bu.reshape(-1))
    if 2 in bc:  # free-slip: no normal component on the shells
        X = of.node_xyz()
        nrm = X / np.linalg.norm(X, axis=-1, keepdims=True)
        u3 = bu.reshape(3, 10, of.n, of.n, of.nr)
        for shell in (0, -1):
            n = nrm[..., shell, :]
            un = sum(u3[c, ..., shell] * n[..., c] for c in range(3))
            for c in range(3):
                u3[c, ..., shell] -= un * n[..., c]
        bu = np.ascontiguousarray(u3.reshape(-1))
    bp = consistent_random(op_, 1, 9).reshape(-1)
    its = 12
    u = torch.zeros(3 * mf.N, dtype=torch.float64, device=cuda)
    p = torch.zeros(mp.N, dtype=torch.float64, device=cuda)
    info = tn.stokes_solve(K, tn.TN_MINRES_BLOCKDIAG, mg, torch.from_numpy(sd).to(cuda), u, p,
                           torch.from_numpy(bu).to(cuda), torch.from_numpy(bp).to(cuda), max_iterations=its,
                           rtol=1e-30)
    vc = S.VCycle(lv, lams, 3, 1)
    nu = 3 * of.N

    def Kop(v):
        vu, vp = v[:nu], v[nu:]
        yu = lv[-1].A(vu) + O.grad_apply_add(of, op_, vp, np.zeros(nu), *bc)
        yp = O.div_apply(of, op_, vu, *bc)
        return np.concatenate([yu, yp])

    def prec(r):
        return np.concatenate([vc(np.zeros(nu), r[:nu]), r[nu:] / sd])

    def dot(a, b):
        return of.dot(a[:nu], b[:nu], 3) + op_.dot(a[nu:], b[nu:], 1)

    xs, hist = S.pminres(Kop, prec, dot, np.zeros(nu + mp.N), np.concatenate([bu, bp]), its)
    h = info["history"]
    assert len(h) == its + 1
    np.testing.assert_allclose(h, hist, rtol=1e-8)
    got = np.concatenate([u.cpu().numpy(), p.cpu().numpy()])
    assert np.abs(got - xs).max() <= 1e-7 * np.abs(xs).max()
