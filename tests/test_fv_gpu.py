"""GPU parity of the FV energy step (tn_fv_*, tn_fct_*) against the oracle
(oracle/fv_oracle.c, pinned in test_oracle_fv.py) on the same seeded inputs:
cell centres with ghost layers, ghost exchange (bit-exact: copies), stable
step, the three FCT stages, the explicit and upwind steps, and the FV -> FE
projection.  Floating-point bars: 1e-12 relative (max norm) for single
stages, 1e-11 after several chained steps; the limiter's ratios are compared
at 1e-10.  Results repeat bit for bit (no atomics)."""
import numpy as np
import pytest

import fv_oracle_lib as F
import oracle_lib as O

SHAPES = [(2, None, 5), (3, None, 9), (2, [0.55, 0.6, 0.68, 0.8, 0.87, 0.95, 1.0], None)]


def _meshes(lat, radii, nr):
    from terraneo_amd import api

    if radii is not None:
        radii = np.asarray(radii, dtype=np.float64)
        return api.Mesh(lat, radii=radii), O.Mesh(lat, radii=radii)
    rad = int(np.log2(nr - 1))
    return api.Mesh(lat, rad, 0.55, 1.0), O.Mesh(lat, nr, 0.55, 1.0)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _velocity(om, seed):
    rng = np.random.default_rng(seed)
    X = om.node_xyz()
    rot = np.cross(np.array([0.5, -0.3, 0.8])[None, None, None, None, :], X)
    u = np.moveaxis(rot, -1, 0) + 0.2 * rng.uniform(-1, 1, (3,) + om.shape)
    return np.ascontiguousarray(u).reshape(-1)


def _real(a):
    return a[:, 1:-1, 1:-1, 1:-1]


@pytest.mark.gpu
@pytest.mark.parametrize("lat,radii,nr", SHAPES)
def test_centres_ghosts_and_stable_step(cuda, lat, radii, nr):
    import torch

    from terraneo_amd import api

    gm, om = _meshes(lat, radii, nr)
    fv = api.FV(gm)
    c = fv.cell_centers().cpu().numpy()
    co = F.cell_centers(om)
    assert np.abs(c - co).max() <= 1e-15 * 4
    # ghost exchange: copies, bit-exact
    T = np.random.default_rng(1).uniform(-1, 1, F.fv_shape(om))
    Tg = torch.from_numpy(T.copy()).to(cuda)
    fv.ghost_update(Tg)
    assert np.array_equal(Tg.cpu().numpy(), F.ghost_update(om, T.copy(), 1))
    vel = _velocity(om, 2)
    vg = torch.from_numpy(vel).to(cuda)
    for kappa in (0.0, 0.05):
        dt = fv.dt_stable(vg, kappa)
        dto = F.dt_stable(om, co, vel, kappa)
        assert abs(dt - dto) <= 1e-13 * dto


@pytest.mark.gpu
@pytest.mark.parametrize("lat,radii,nr", SHAPES)
@pytest.mark.parametrize("kappa,with_source,subtract", [(0.0, False, True), (0.03, True, True), (0.01, False, False)])
def test_fct_stages(cuda, lat, radii, nr, kappa, with_source, subtract):
    import torch

    from terraneo_amd import api

    gm, om = _meshes(lat, radii, nr)
    fv = api.FV(gm)
    co = F.cell_centers(om)
    rng = np.random.default_rng(3)
    T = rng.uniform(0, 1, F.fv_shape(om))
    T = F.dirichlet(om, T, cmb=1.0, surf=0.0)
    src = rng.uniform(-1, 1, F.fv_shape(om)) if with_source else None
    vel = _velocity(om, 4)
    dt = 0.5 * F.dt_stable(om, co, vel, kappa)
    Tg = torch.from_numpy(T.copy()).to(cuda)
    vg = torch.from_numpy(vel).to(cuda)
    sg = torch.from_numpy(src).to(cuda) if src is not None else None
    fv.predictor(Tg, vg, dt, kappa, sg, subtract)
    TL, ad, _, _ = fv.buffers()
    To, TLo, ado = F.predictor(om, co, vel, T, dt, kappa, src, subtract)
    assert np.array_equal(Tg.cpu().numpy(), To)  # ghost exchange of T
    assert _rel(_real(TL), _real(TLo)) <= 1e-12
    ado_f = np.moveaxis(ado, -1, 0)
    assert _rel(ad[:, :, 1:-1, 1:-1, 1:-1], ado_f[:, :, 1:-1, 1:-1, 1:-1]) <= 1e-12
    # limiter on the same T_L / fluxes (the oracle's), then the correction
    fv.limiter()
    TL2, _, Rp, Rm = fv.buffers()
    TLo2, Rpo, Rmo = F.limiter(om, TLo, ado)
    assert _rel(_real(TL2), _real(TLo2)) <= 1e-12
    assert np.abs(Rp - Rpo).max() <= 1e-10 and np.abs(Rm - Rmo).max() <= 1e-10
    fv.correction(Tg)
    Tn = F.correction(om, TLo2, ado, Rpo, Rmo, To)
    assert _rel(_real(Tg.cpu().numpy()), _real(Tn)) <= 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("lat,radii,nr", SHAPES[:2])
def test_explicit_and_upwind_steps(cuda, lat, radii, nr):
    import torch

    from terraneo_amd import api

    gm, om = _meshes(lat, radii, nr)
    fv = api.FV(gm)
    co = F.cell_centers(om)
    T = np.random.default_rng(5).uniform(0, 1, F.fv_shape(om))
    T = F.dirichlet(om, T, cmb=1.0, surf=0.0)
    vel = _velocity(om, 6)
    kappa = 0.02
    dt = 0.8 * F.dt_stable(om, co, vel, kappa)
    src = np.full(F.fv_shape(om), 0.3)
    st = F.FctStepper(om, co)
    To = T.copy()
    Tg = torch.from_numpy(T.copy()).to(cuda)
    vg, sg = torch.from_numpy(vel).to(cuda), torch.from_numpy(src).to(cuda)
    for _ in range(4):
        To = st.step(vel, To, dt, kappa, src, True, cmb=1.0, surf=0.0)
        To = F.dirichlet(om, To, cmb=1.0, surf=0.0)
        fv.explicit_step(Tg, vg, dt, kappa, sg, True, cmb=1.0, surf=0.0)
        fv.apply_dirichlet(Tg, cmb=1.0, surf=0.0)
    assert _rel(_real(Tg.cpu().numpy()), _real(To)) <= 1e-11
    # same inputs, same bits
    T2 = torch.from_numpy(T.copy()).to(cuda)
    fv2 = api.FV(gm)
    for _ in range(4):
        fv2.explicit_step(T2, vg, dt, kappa, sg, True, cmb=1.0, surf=0.0)
        fv2.apply_dirichlet(T2, cmb=1.0, surf=0.0)
    assert torch.equal(T2, Tg)
    # upwind step = the predictor's T_L
    Tu = torch.from_numpy(T.copy()).to(cuda)
    fv.upwind_step(Tu, vg, dt, kappa, None, True)
    _, TLo, _ = F.predictor(om, co, vel, T, dt, kappa, None, True)
    assert _rel(_real(Tu.cpu().numpy()), _real(TLo)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("lat,radii,nr", SHAPES)
def test_fv_to_fe_projection(cuda, lat, radii, nr):
    import torch

    from terraneo_amd import api

    gm, om = _meshes(lat, radii, nr)
    fv = api.FV(gm)
    T = np.random.default_rng(7).uniform(0, 1, F.fv_shape(om))
    Tg = torch.from_numpy(T).to(cuda)
    b = torch.empty(gm.N, dtype=torch.float64, device=cuda)
    fv.to_fe_rhs(Tg, b)
    bo = F.fv_to_fe_rhs(om, T)
    assert _rel(b.cpu().numpy(), bo) <= 1e-13
    x = torch.zeros(gm.N, dtype=torch.float64, device=cuda)
    info = fv.to_fe(Tg, x)
    assert info["final_residual"] <= max(1e-12 * info["initial_residual"], 1e-12)
    # M x = b with the oracle's mass matrix
    xs = np.concatenate([x.cpu().numpy(), np.zeros(2 * gm.N)])
    Mx = O.vector_mass(om, xs)[: gm.N]
    assert _rel(Mx, bo) <= 1e-10
    # a constant field projects to the constant
    one = torch.ones(F.fv_shape(om), dtype=torch.float64, device=cuda)
    x1 = torch.zeros(gm.N, dtype=torch.float64, device=cuda)
    fv.to_fe(one, x1)
    assert float((x1 - 1).abs().max()) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("lat,radii,nr", SHAPES)
def test_fv_to_fe_lumped(cuda, lat, radii, nr):
    """b / D against the oracle's right-hand side (pinned above); D = M 1, the
    consistent mass's row sums (partition of unity); constants are kept; the
    result stays inside the FV values' bounds (a convex combination)."""
    import torch

    from terraneo_amd import api

    print(f"case lumped fv->fe lat={lat} nr={nr}", flush=True)
    gm, om = _meshes(lat, radii, nr)
    fv = api.FV(gm)
    T = np.random.default_rng(11).uniform(0.2, 0.7, F.fv_shape(om))
    u = torch.empty(gm.N, dtype=torch.float64, device=cuda)
    D = torch.empty(gm.N, dtype=torch.float64, device=cuda)
    fv.to_fe_lumped(torch.from_numpy(T).to(cuda), u, D)
    Do = F.fv_to_fe_rhs(om, np.ones(F.fv_shape(om)))
    assert _rel(D.cpu().numpy(), Do) <= 1e-13
    M1 = O.vector_mass(om, np.concatenate([np.ones(gm.N), np.zeros(2 * gm.N)]))[: gm.N]
    assert _rel(Do, M1) <= 1e-12
    uo = F.fv_to_fe_rhs(om, T) * (1.0 / Do)
    un = u.cpu().numpy()
    assert _rel(un, uo) <= 1e-13
    Tr = _real(T)
    assert un.min() >= Tr.min() - 1e-14 and un.max() <= Tr.max() + 1e-14
    one = torch.ones(F.fv_shape(om), dtype=torch.float64, device=cuda)
    fv.to_fe_lumped(one, u, D)
    assert float((u - 1).abs().max()) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("lat,radii,nr", SHAPES)
def test_fe_to_fv_cell_averages(cuda, lat, radii, nr):
    """The nodal coordinate fields average to the cell centres (their Q1
    interpolant is the coordinate itself, so T_K = int_K x / |K|, the centre the
    oracle computes); constants are kept; lateral ghosts equal the exchange;
    linearity in the nodal field."""
    import torch

    from terraneo_amd import api

    print(f"case fe->fv lat={lat} nr={nr}", flush=True)
    gm, om = _meshes(lat, radii, nr)
    fv = api.FV(gm)
    X = om.node_xyz()
    co = F.cell_centers(om)
    for k in range(3):
        Tg = torch.zeros(F.fv_shape(om), dtype=torch.float64, device=cuda)
        fv.from_fe(torch.from_numpy(np.ascontiguousarray(X[..., k]).reshape(-1)).to(cuda), Tg)
        Tn = Tg.cpu().numpy()
        assert np.abs(_real(Tn) - _real(co[..., k])).max() <= 1e-14
        assert np.array_equal(Tn, F.ghost_update(om, Tn.copy(), 1))
    Tg = torch.zeros(F.fv_shape(om), dtype=torch.float64, device=cuda)
    fv.from_fe(torch.full((gm.N,), 3.0, dtype=torch.float64, device=cuda), Tg)
    assert np.abs(_real(Tg.cpu().numpy()) - 3.0).max() <= 1e-14
    rng = np.random.default_rng(12)
    a, b = rng.uniform(-1, 1, gm.N), rng.uniform(-1, 1, gm.N)
    outs = []
    for v in (a, b, 2.0 * a - b):
        Tg = torch.zeros(F.fv_shape(om), dtype=torch.float64, device=cuda)
        fv.from_fe(torch.from_numpy(v).to(cuda), Tg)
        outs.append(_real(Tg.cpu().numpy()))
    assert np.abs(outs[2] - (2.0 * outs[0] - outs[1])).max() <= 1e-13
