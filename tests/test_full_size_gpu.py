"""Property checks at the two full single-GPU sizes, where no oracle runs:

  mixed-precision workload   velocity lateral 7 x 129 radial, 5-level GMG,
                             nodal viscosity contrast 1e5 (seed 3) x30 above
                             r = 0.9, free-slip shells
  forward + adjoint workload velocity lateral 8 x 129 radial, 6-level FP32
                             GMG, per-cell viscosity contrast 1e6 (seed 6)
                             x100 above r = 0.95, free-slip shells

  * symmetry <K x, y> = <x, K y> of the Stokes block and of A, with the
    rotation penalty once per node (the reference's per-copy rule is not
    symmetric), on interface-consistent random vectors with tangential
    velocities on the free-slip shells;
  * the FP32 fine operator (the V-cycle's) against the FP64 one: relative L2
    difference <= 2e-5 on random input (FP32 rounding times the cancellation
    of a random vector's apply, ~ 6e-8 x 100);
  * residual decrease of the mixed FGMRES(30) over a bounded number of
    iterations: monotone restart residuals, the reported residual equal to a
    recomputed true residual;
  * the adjoint gradient kernel against a central difference of l^T A(m) u
    along one smooth direction, on free-slip-projected u and l (there the
    boundary terms of A that the kernel leaves out vanish);
  * a chunked solve (warm-started calls of whole restart cycles) takes the
    single call's iterations to the same solution (lateral 4, FP64 V-cycle;
    the calls' pressure mean projections at their ends differ from the single
    call only by rounding.  With the FP32 V-cycle such rounding-level input
    differences flip FP32 roundings in the preconditioner, and the two
    iterations then agree only to ~1e-6).
"""
import numpy as np
import pytest


def _consistent_random(mesh, seed, ncomp=3, dtype=None, device="cuda:0"):
    """uniform [-1, 1) values, one per node (every copy holds the owned copy's)"""
    import torch

    from terraneo_amd import api

    dtype = dtype or torch.float64
    x = torch.empty(ncomp * mesh.N, dtype=dtype, device=device)
    mesh.randomize(x, seed, ncomp)
    x = 2 * x - 1
    own = torch.from_numpy(mesh.owned_mask().reshape(-1).astype(np.float64)).to(device=device, dtype=dtype)
    x = (x.view(ncomp, -1) * own).reshape(-1).contiguous()
    mesh.interface_sum(x, ncomp)
    return x


def _project(pr, v):
    from terraneo_amd import api

    for which in (1, 2):
        api.freeslip_project(pr.vel, v, which)
    return v


def _symmetry(pr, seed):
    """relative asymmetry of the Stokes block and of A (penalty once per node) on
    free-slip-projected velocities, the space the Krylov iterates live in"""
    import torch

    u = _project(pr, _consistent_random(pr.vel, seed))
    v = _project(pr, _consistent_random(pr.vel, seed + 1))
    p, q = _consistent_random(pr.pre, seed + 2, 1), _consistent_random(pr.pre, seed + 3, 1)
    Ku, Kp = torch.empty_like(u), torch.empty_like(p)
    Kv, Kq = torch.empty_like(v), torch.empty_like(q)
    pr.K.apply(u, p, Ku, Kp)
    pr.K.apply(v, q, Kv, Kq)
    a = pr.vel.dot(Ku, v, 3) + pr.pre.dot(Kp, q, 1)
    b = pr.vel.dot(u, Kv, 3) + pr.pre.dot(p, Kq, 1)
    scale = np.sqrt((pr.vel.dot(Ku, Ku, 3) + pr.pre.dot(Kp, Kp, 1)) * (pr.vel.dot(v, v, 3) + pr.pre.dot(q, q, 1)))
    sK = abs(a - b) / scale
    pr.A.apply(u, Ku)
    pr.A.apply(v, Kv)
    a, b = pr.vel.dot(Ku, v, 3), pr.vel.dot(u, Kv, 3)
    sA = abs(a - b) / np.sqrt(pr.vel.dot(Ku, Ku, 3) * pr.vel.dot(v, v, 3))
    return sK, sA


def _fp32_vs_fp64(pr, seed):
    import torch

    x = _consistent_random(pr.vel, seed)
    y64 = torch.empty_like(x)
    pr.A.apply(x, y64)
    x32 = x.float()
    y32 = torch.empty_like(x32)
    pr.ops[-1].apply(x32, y32)
    d = y32.double() - y64
    return float(np.sqrt(pr.vel.dot(d, d, 3) / pr.vel.dot(y64, y64, 3)))


def _residual_decrease(pr, bu, bp, iterations):
    u, p, info = pr.solve(bu, bp, rtol=1e-8, max_iterations=iterations, restart=30, pressure_mean_free=True)
    hist = np.asarray(info["history"])
    rel = info["final_residual"] / info["initial_residual"]
    true = pr.residual(u, p, bu, bp, pressure_mean_free=True)
    return hist, rel, true, info


def _gradient_check(pr, m, coords, radii, seed):
    import torch

    from terraneo_amd import workloads as W
    from terraneo_amd.forward_adjoint import bilinear_derivative

    u = _project(pr, _consistent_random(pr.vel, seed))
    lam = _project(pr, _consistent_random(pr.vel, seed + 1))
    grad = torch.empty_like(m)
    pr.A.grad_log_cell_eta(u, lam, grad)
    d = torch.from_numpy(np.ascontiguousarray(W.cell_smooth_normal(coords, radii, 1000))).to(m.device)
    fd = bilinear_derivative(pr, m, u, lam, d)
    an = -float((grad * d).sum())
    return an, fd


@pytest.fixture(scope="module")
def mixed_problem(cuda):
    import torch

    from terraneo_amd import api
    from terraneo_amd import workloads as W
    from terraneo_amd.problem import StokesProblem

    pr = StokesProblem(7, 7, 5, device=cuda, prec_dtype="f32",
                       gmg=api.gmg_options(cheby_order=3, cheby_iterations=2, power_iterations=10,
                                           coarse_solver="dense"))
    coords, radii = pr.coords()
    pr.set_viscosity(eta=torch.from_numpy(W.viscosity_c2(coords, radii).reshape(-1)).to(cuda))
    f = torch.from_numpy(W.buoyancy(coords, radii)).to(cuda)
    bu, bp = pr.rhs_from_force(f)
    del f
    yield pr, bu, bp
    del pr, bu, bp
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_mixed_workload_full_size_properties(mixed_problem):
    pr, bu, bp = mixed_problem
    pr.A.set_penalty_once(True)
    sK, sA = _symmetry(pr, 11)
    pr.A.set_penalty_once(False)
    e32 = _fp32_vs_fp64(pr, 21)
    hist, rel, true, info = _residual_decrease(pr, bu, bp, 60)
    print(f"lateral 7 x 129: asymmetry K {sK:.2e} A {sA:.2e}; FP32 vs FP64 apply {e32:.2e}; "
          f"FGMRES(30) 60 iterations rel. res {rel:.3e} (true {true:.3e})", flush=True)
    assert sK <= 1e-12 and sA <= 1e-12
    assert e32 <= 2e-5
    assert info["iterations"] == 60
    assert rel <= 1e-3
    assert abs(true - rel) <= 1e-6 * rel
    assert np.all(hist[1:] <= hist[:-1] * (1 + 1e-6))


@pytest.fixture(scope="module")
def largest_problem(cuda):
    import torch

    from terraneo_amd.forward_adjoint import build_problem

    pr, m, bu, bp, coords, radii = build_problem(8, 7, 6)
    yield pr, m, bu, bp, coords, radii
    del pr, m, bu, bp
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_forward_adjoint_workload_full_size_properties(largest_problem):
    pr, m, bu, bp, coords, radii = largest_problem
    sK, sA = _symmetry(pr, 31)
    e32 = _fp32_vs_fp64(pr, 41)
    an, fd = _gradient_check(pr, m, coords, radii, 51)
    hist, rel, true, info = _residual_decrease(pr, bu, bp, 60)
    print(f"lateral 8 x 129: asymmetry K {sK:.2e} A {sA:.2e}; FP32 vs FP64 apply {e32:.2e}; "
          f"gradient kernel {an:.10e} vs central difference {fd:.10e}; "
          f"FGMRES(30) 60 iterations rel. res {rel:.3e} (true {true:.3e})", flush=True)
    assert sK <= 1e-12 and sA <= 1e-12
    assert e32 <= 2e-5
    assert abs(an - fd) <= 1e-6 * abs(fd)
    assert info["iterations"] == 60
    assert rel <= 1e-1
    assert abs(true - rel) <= 1e-6 * rel
    assert np.all(hist[1:] <= hist[:-1] * (1 + 1e-6))


@pytest.mark.gpu
def test_chunked_solve_equals_single_call(cuda):
    import torch

    from terraneo_amd.forward_adjoint import build_problem, chunked_solve

    pr, m, bu, bp, coords, radii = build_problem(4, 4, 3, prec_dtype="f64")
    u1, p1, info = pr.solve(bu, bp, rtol=1e-8, max_iterations=150, restart=30, pressure_mean_free=True)
    u2, p2, c = chunked_solve(pr, bu, bp, 1e-8, 150, 60, "chunked")
    assert c["iterations"] == info["iterations"]
    assert c["calls"] >= 2
    du = float((u1 - u2).abs().max() / u1.abs().max())
    dp = float((p1 - p2).abs().max() / p1.abs().max())
    print(f"chunked vs single call: {c['iterations']} iterations, velocity {du:.2e} pressure {dp:.2e}", flush=True)
    assert du <= 1e-9 and dp <= 1e-9


@pytest.mark.gpu
def test_adjoint_rotation_part_does_not_change_the_gradient(cuda):
    """fs/fs with the penalty once per node: the rigid rotations are an invariant
    subspace of K and carry no strain, so removing the adjoint right-hand side's
    rotation part leaves the viscosity gradient unchanged (lateral 4, FP64
    V-cycle, the lateral-8 workload's viscosity)"""
    import torch

    from terraneo_amd.forward_adjoint import build_problem, remove_rotations, rotation_content, surface_objective

    pr, m, bu, bp, coords, radii = build_problem(4, 4, 3, prec_dtype="f64")
    kw = dict(rtol=1e-12, max_iterations=4000, restart=30, pressure_mean_free=True)
    u, p, info = pr.solve(bu, bp, **kw)
    J, g = surface_objective(pr, u)
    rot = rotation_content(pr, g)
    gp = torch.zeros(pr.pre.N, dtype=torch.float64, device=cuda)
    grads, its = [], []
    for rhs in (g, remove_rotations(pr, g)):
        lam, _, ai = pr.solve(rhs, gp, **kw)
        gr = torch.empty_like(m)
        pr.A.grad_log_cell_eta(u, lam, gr)
        grads.append(gr)
        its.append((ai["iterations"], ai["final_residual"] / ai["initial_residual"]))
    diff = float((grads[0] - grads[1]).abs().max() / grads[0].abs().max())
    print(f"rotation content {rot:.3e}; adjoint (iterations, rel. res) with / without it {its}; "
          f"gradient difference {diff:.2e}", flush=True)
    assert rot > 1e-3
    assert its[1][1] <= 1e-12 * (1 + 1e-9)
    assert diff <= 1e-6
