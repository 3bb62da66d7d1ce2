"""Forward + adjoint free-slip Stokes run with the log-viscosity gradient, at
full size on one GPU.

Default workload: velocity lateral level 8 (257 x 257 per diamond) x 129
radial nodes (255 M velocity DoF), free-slip on both shells, per-cell
log-viscosity from random harmonics (seed 6) normalised to [0, ln 1e6] plus
ln 100 in the cells above r = 0.95, buoyancy forcing (seed 2).  The velocity
multigrid (6 levels, Chebyshev degree 3, Galerkin coarse matrices) runs in
FP32 with an FP64 dense coarsest solve, inside FP64 FGMRES(30) with the
block-triangular preconditioner, to rel. res 1e-8 for both solves.  The
rotation penalty is counted once per node, so K = K^T and the adjoint system
is K itself.

  forward   K (u, p) = b
  objective J = 1/2 sum over owned outer-shell nodes |u|^2 (u_obs = 0)
  adjoint   K (l, q) = (dJ/du, 0)   (dJ/du summed over node copies, free-slip projected,
                                     its rigid-rotation part removed: see remove_rotations)
  gradient  dJ/dm[cell] = -eta_cell (T l)^T A_cell(1) u   (tn_eps_grad_log_cell_eta)

Each solve runs as warm-started calls of whole FGMRES(30) cycles (a restart
recomputes r = b - K x, so a new call continues the same iteration), with the
stopping test fixed to 1e-8 of the first call's initial residual; every call
prints one line.  Also measured: the FP64 and FP32 apply at this size (HIP
events on the launch stream) and the gradient kernel against a central
difference of l^T A(m) u along one smooth direction (two applies, no solve).

    python -m terraneo_amd.forward_adjoint [--lat 8 --rad 7 --levels 6] [--json out.json]
"""
import argparse
import json
import math
import sys
import time


def _log(msg):
    print(msg, flush=True)


def chunked_solve(pr, bu, bp, rtol, max_iterations, chunk, label, restart=30):
    """FGMRES(restart) solve from zero in calls of `chunk` iterations (a multiple
    of the restart length), stopping at rtol times the first residual."""
    import torch

    u = torch.zeros(3 * pr.vel.N, **pr.f64)
    p = torch.zeros(pr.pre.N, **pr.f64)
    total, seconds, r0, res, calls = 0, 0.0, None, None, 0
    while total < max_iterations:
        n = min(chunk, max_iterations - total)
        kw = dict(rtol=rtol, atol=1e-300) if r0 is None else dict(rtol=1e-300, atol=rtol * r0)
        u, p, info = pr.solve(bu, bp, max_iterations=n, restart=restart, u=u, p=p, pressure_mean_free=True, **kw)
        calls += 1
        if r0 is None:
            r0 = info["initial_residual"]
        total += info["iterations"]
        seconds += info["seconds"]
        res = info["final_residual"] / r0
        _log(f"[{label}] {total} iterations, rel. res {res:.3e}, {seconds:.1f} s")
        if res <= rtol or info["iterations"] < n:
            break
    return u, p, dict(iterations=total, seconds=seconds, rel_residual=res, calls=calls)


def time_apply(op, x, y, reps=10):
    """mean ms of op.apply over reps launches (HIP events on the current stream)"""
    import torch

    for _ in range(2):
        op.apply(x, y)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        op.apply(x, y)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def build_problem(lat=8, rad=7, levels=6, cheby_iterations=2, rtol=1e-8, prec_dtype="f32"):
    """The module's workload set up on cuda:0: (problem, m, bu, bp, coords, radii),
    m the per-cell log-viscosity (torch FP64 on the device); prec_dtype "f64"
    runs the V-cycle in FP64 instead."""
    import numpy as np
    import torch

    from . import api
    from . import workloads as W
    from ._capi import TN_FREESLIP
    from .problem import StokesProblem

    dev = torch.device("cuda:0")
    pr = StokesProblem(lat, rad, levels, bc=(TN_FREESLIP, TN_FREESLIP), device=dev, prec_dtype=prec_dtype,
                       penalty="once",
                       gmg=api.gmg_options(cheby_order=3, cheby_iterations=cheby_iterations, power_iterations=10,
                                           coarse_solver="dense"))
    coords, radii = pr.coords()
    m = W.cell_log_viscosity(coords, radii, 6, 1e6)
    rmid = 0.5 * (radii[:-1] + radii[1:])
    m = m + math.log(100.0) * (rmid > 0.95)[None, None, None, :]
    mt = torch.from_numpy(np.ascontiguousarray(m)).to(dev)
    f = torch.from_numpy(W.buoyancy(coords, radii)).to(dev)
    pr.set_viscosity(eta_cell=torch.exp(mt).contiguous())
    bu, bp = pr.rhs_from_force(f)
    del f
    pr.solve(bu, bp, rtol=rtol, max_iterations=1, pressure_mean_free=True)  # eigenvalue estimates
    torch.cuda.synchronize()
    return pr, mt, bu, bp, coords, radii


def surface_objective(pr, u):
    """J = 1/2 sum over owned outer-shell nodes |u|^2 and its adjoint right-hand
    side (dJ/du summed over the node copies, free-slip projected)."""
    import torch

    from . import api

    vel = pr.vel
    own = torch.from_numpy(vel.owned_mask().reshape(10, vel.n, vel.n, vel.nr).astype(bool)).to(u.device)
    surf = torch.zeros_like(own)
    surf[..., -1] = True
    mask = (own & surf).reshape(-1).repeat(3).to(torch.float64)
    J = float(0.5 * (mask * u * u).sum())
    g = (mask * u).contiguous()
    vel.interface_sum(g, 3)
    for which in (1, 2):
        api.freeslip_project(vel, g, which)
    return J, g


def rotation_projection(pr, v):
    """P v, the owned-inner-product projection of v onto the rigid rotations
    e_k x x (the fs/fs near-null space the rotation penalty controls)"""
    import numpy as np
    import torch

    vel = pr.vel
    xyz = torch.from_numpy(np.asarray(vel.lateral_coords())).to(v.device)[:, :, :, None, :] * \
        torch.from_numpy(np.asarray(vel.radii())).to(v.device)[None, None, None, :, None]
    modes = []
    for k in range(3):
        e = torch.zeros(3, dtype=torch.float64, device=v.device)
        e[k] = 1.0
        modes.append(torch.cross(e.expand_as(xyz), xyz, dim=-1).movedim(-1, 0).reshape(-1).contiguous())
    G = np.array([[vel.dot(a, b, 3) for b in modes] for a in modes])
    c = np.linalg.solve(G, np.array([vel.dot(a, v, 3) for a in modes]))
    return sum(float(ck) * mk for ck, mk in zip(c, modes))


def rotation_content(pr, v):
    """|P v| / |v| (see rotation_projection)"""
    import numpy as np

    pv = rotation_projection(pr, v)
    return float(np.sqrt(pr.vel.dot(pv, pv, 3) / pr.vel.dot(v, v, 3)))


def remove_rotations(pr, g):
    """g - P g.  On free-slip shells with the penalty once per node the rigid
    rotations span an invariant subspace of the symmetric K (A and B annihilate
    them, the penalty maps them into their span), so the adjoint solution's part
    from P g is a rigid rotation, whose strain, and hence whose contribution to
    the viscosity gradient, is zero.  Removing P g leaves the gradient unchanged
    and takes the slowly converging, penalty-scaled component out of the
    adjoint solve."""
    return (g - rotation_projection(pr, g)).contiguous()


def residual_report(pr, u, p, bu, bp, label):
    """Where the residual of a solve sits: velocity and pressure parts, the
    velocity part on the two shells (and its normal component there) and inside,
    its rigid-rotation content, |r_u|^2 and |r_p|^2 by radial eighth, and the
    cell viscosity at the largest velocity residual entries.  Printed lines."""
    import numpy as np
    import torch

    vel, pre = pr.vel, pr.pre
    own_p = torch.from_numpy(pre.owned_mask().reshape(-1).astype(np.float64)).to(u.device)
    p_mf = p - (own_p * p).sum() / own_p.sum()
    ku, kp = torch.empty_like(u), torch.empty_like(p)
    pr.K.apply(u, p_mf, ku, kp)
    ru = (bu - ku).contiguous()
    rp = bp - kp
    rp = (rp - (own_p * rp).sum() / own_p.sum()).contiguous()

    def nrm(v, nc=3, m=vel):
        v = v.contiguous()
        return float(np.sqrt(m.dot(v, v, nc)))

    nb = nrm(bu)
    _log(f"[{label} residual] |b_u| {nb:.3e}, |r_u| {nrm(ru):.3e}, |r_p| {nrm(rp, 1, pre):.3e}")
    R = ru.view(3, 10, vel.n, vel.n, vel.nr)
    c = torch.from_numpy(np.asarray(vel.lateral_coords())).to(u.device).permute(3, 0, 1, 2)
    for name, lev in (("CMB", 0), ("surface", vel.nr - 1)):
        z = torch.zeros_like(R)
        z[..., lev] = R[..., lev]
        zn = torch.zeros_like(R)
        zn[..., lev] = (R[..., lev] * c).sum(0)[None] * c
        _log(f"[{label} residual] r_u on the {name} level {nrm(z.view(-1)):.3e} (normal part {nrm(zn.view(-1)):.3e})")
    zi = R.clone()
    zi[..., 0] = 0
    zi[..., -1] = 0
    _log(f"[{label} residual] r_u inside {nrm(zi.view(-1)):.3e}, rigid-rotation content {rotation_content(pr, ru):.3e}")
    prof = np.array_split((R ** 2).sum(dim=(0, 1, 2, 3)).cpu().numpy(), 8)
    _log(f"[{label} residual] |r_u|^2 by radial eighth (CMB to surface): " + " ".join(f"{b.sum():.2e}" for b in prof))
    P = rp.view(10, pre.n, pre.n, pre.nr)
    pp = np.array_split((P ** 2).sum(dim=(0, 1, 2)).cpu().numpy(), 8)
    _log(f"[{label} residual] |r_p|^2 by radial eighth: " + " ".join(f"{b.sum():.2e}" for b in pp))
    if pr.eta_cell is not None:
        e = pr.eta_cell.view(10, vel.n - 1, vel.n - 1, vel.nr - 1)
        mag = (R ** 2).sum(0).view(-1)
        for i in torch.topk(mag, 3).indices.tolist():
            d, rem = divmod(i, vel.n * vel.n * vel.nr)
            x, rem = divmod(rem, vel.n * vel.nr)
            y, r = divmod(rem, vel.nr)
            ec = e[d, min(x, vel.n - 2), min(y, vel.n - 2), min(r, vel.nr - 2)].item()
            _log(f"[{label} residual] large r_u {mag[i].sqrt().item():.3e} at diamond {d} ({x}, {y}) level {r}, "
                 f"cell viscosity {ec:.2e}")


def bilinear_derivative(pr, m, u, lam, d, h=1e-4):
    """central difference of l^T A(m) u along d (two applies at m +- h d)"""
    import torch

    from . import api

    vel = pr.vel
    y = torch.empty_like(u)
    vals = []
    for sgn in (1.0, -1.0):
        op = api.Eps(vel, None, pr.bc[0], pr.bc[1], eta_cell=torch.exp(m + sgn * h * d).contiguous())
        op.set_penalty_once(pr.penalty == "once")
        op.apply(u, y)
        vals.append(vel.dot(lam, y, 3))
        del op
    return (vals[0] - vals[1]) / (2 * h)


def run(lat=8, rad=7, levels=6, cheby_iterations=2, max_iterations=6000, chunk=240, rtol=1e-8, prec="f32",
        report=False, restart=30):
    import numpy as np
    import torch

    from . import workloads as W

    t0 = time.perf_counter()
    pr, mt, bu, bp, coords, radii = build_problem(lat, rad, levels, cheby_iterations, rtol, prec_dtype=prec)
    vel = pr.vel
    setup = time.perf_counter() - t0
    free, total = torch.cuda.mem_get_info()
    used_setup = (total - free) / 2**30
    _log(f"setup {setup:.1f} s: velocity {3 * vel.N} DoF, pressure {pr.pre.N} DoF, "
         f"device memory in use {used_setup:.1f} GiB")

    u, p, fwd = chunked_solve(pr, bu, bp, rtol, max_iterations, chunk, "forward", restart)
    fwd["true_rel_residual"] = pr.residual(u, p, bu, bp, pressure_mean_free=True)
    _log(f"forward: true rel. res {fwd['true_rel_residual']:.3e}")
    if report:
        residual_report(pr, u, p, bu, bp, "forward")

    J, g = surface_objective(pr, u)
    rot = rotation_content(pr, g)
    g = remove_rotations(pr, g)
    _log(f"adjoint right-hand side: rigid-rotation content {rot:.3e} (removed; it cannot change the gradient)")
    gp = torch.zeros(pr.pre.N, **pr.f64)
    lam, q, adj = chunked_solve(pr, g, gp, rtol, max_iterations, chunk, "adjoint", restart)
    adj["true_rel_residual"] = pr.residual(lam, q, g, gp, pressure_mean_free=True)
    _log(f"adjoint: true rel. res {adj['true_rel_residual']:.3e}")
    if report:
        residual_report(pr, lam, q, g, gp, "adjoint")
    del q, gp

    grad = torch.empty_like(mt)
    pr.A.grad_log_cell_eta(u, lam, grad)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    pr.A.grad_log_cell_eta(u, lam, grad)
    b.record()
    torch.cuda.synchronize()
    grad_ms = a.elapsed_time(b)

    # gradient kernel against a central difference of l^T A(m) u along one direction
    d = torch.from_numpy(np.ascontiguousarray(W.cell_smooth_normal(coords, radii, 1000))).to(mt.device)
    fd = bilinear_derivative(pr, mt, u, lam, d)
    an = -float((grad * d).sum())
    grad_fd_rel = abs(an - fd) / abs(fd)
    _log(f"gradient kernel {grad_ms:.3f} ms; d/dm (l^T A u) along one direction: kernel {an:.10e} "
         f"central difference {fd:.10e} rel {grad_fd_rel:.2e}")

    # apply at this size: FP64 operator, and FP32 (the V-cycle's fine level)
    n_nodes = vel.N
    y = torch.empty_like(u)
    ms64 = time_apply(pr.A, u, y)
    ms32 = None
    if prec == "f32":
        u32 = u.float()
        y32 = torch.empty_like(u32)
        ms32 = time_apply(pr.ops[-1], u32, y32)
        del u32, y32
    _log(f"apply FP64 {ms64:.3f} ms ({3 * n_nodes / ms64 / 1e6:.1f} GDoF/s, "
         f"{56 * n_nodes / ms64 / 1e9:.3f} TB/s algorithmic)" +
         (f", FP32 {ms32:.3f} ms ({3 * n_nodes / ms32 / 1e6:.1f} GDoF/s)" if ms32 else ""))

    n = 2 ** lat + 1
    nr = 2 ** rad + 1
    return {
        "workload": f"free-slip Stokes forward + adjoint, velocity lateral level {lat} ({n}x{n} per diamond) x "
                    f"{nr} radial, {levels}-level {prec.upper()} GMG (Chebyshev degree 3 x{cheby_iterations}, "
                    f"{prec.upper()} Galerkin matrices, FP64 dense coarsest) in FP64 FGMRES({restart}), rel. res {rtol:g}, "
                    "per-cell viscosity "
                    "contrast 1e6 (seed 6) x100 for r > 0.95, penalty once per node",
        "velocity_dof": 3 * n_nodes,
        "setup_seconds": round(setup, 2),
        "forward": fwd,
        "adjoint": adj,
        "objective": J,
        "adjoint_rhs_rotation_content": rot,
        "gradient_kernel_ms": round(grad_ms, 4),
        "gradient_vs_central_difference_of_bilinear_form": float(f"{grad_fd_rel:.3e}"),
        "apply_ms": {"f64": round(ms64, 4), "f32": round(ms32, 4) if ms32 else None},
        "apply_gdof_per_s": {"f64": round(3 * n_nodes / ms64 / 1e6, 2),
                             "f32": round(3 * n_nodes / ms32 / 1e6, 2) if ms32 else None},
        "apply_hbm_fraction_f64": round(56 * n_nodes / ms64 / 1e9 / 8.0, 4),
        "device_memory_after_setup_gib": round(used_setup, 1),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--lat", type=int, default=8)
    ap.add_argument("--rad", type=int, default=7)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--cheby-iterations", type=int, default=2)
    ap.add_argument("--max-iterations", type=int, default=6000)
    ap.add_argument("--chunk", type=int, default=240)
    ap.add_argument("--prec", choices=("f32", "f64"), default="f32", help="V-cycle precision")
    ap.add_argument("--residual-report", action="store_true", help="where each solve's final residual sits")
    ap.add_argument("--restart", type=int, default=30, help="FGMRES restart length (the workload's: 30)")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    if a.chunk % a.restart:
        sys.exit("--chunk must be a multiple of the restart length")
    out = run(a.lat, a.rad, a.levels, a.cheby_iterations, a.max_iterations, a.chunk, prec=a.prec,
              report=a.residual_report, restart=a.restart)
    line = json.dumps(out)
    print(line, flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            fh.write(line + "\n")


if __name__ == "__main__":
    main()
