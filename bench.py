#!/usr/bin/env python3
"""Benchmark of the matrix-free Stokes hot path on one MI355X.

Workload (one "step"): one full epsilon-div-div operator apply -- matrix-free
wedge kernel (writes every node, no separate zero-fill), additive interface
sum -- on the velocity mesh of
the mixed-precision Stokes configuration: lateral level 7 (129 x 129 nodes
per diamond) x 129 radial nodes, 10 diamonds, r in [0.55, 1], 63 406 854
velocity DoF, FP64, Dirichlet on both shells.  Inputs are resident in HBM
before the timed region.  GDoF = 3 x owned velocity nodes (the reference's
convention, apps/benchmarks/performance/benchmark_operators.cpp:207-212).

Prints ONE JSON line (rank 0).  `roofline` is for the dominant kernel (the
wedge kernel alone, timed with HIP events on its own stream); `cpu_baseline`
is the CPU oracle (OpenMP C) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--lat", type=int, default=7)
    p.add_argument("--rad", type=int, default=7)
    p.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-solve", action="store_true", help="skip the FP64 / mixed-precision Stokes solves")
    p.add_argument("--solve-lat", type=int, default=7)
    p.add_argument("--solve-rad", type=int, default=7)
    p.add_argument("--solve-levels", type=int, default=5)
    p.add_argument("--solve-max-iterations", type=int, default=4000)
    p.add_argument("--solve-cheby-iterations", type=int, default=2, help="Chebyshev (degree 3) sweeps per smoothing")
    p.add_argument("--no-tight-reference", dest="tight_reference", action="store_false",
                   help="skip the FP64 solve continued to 1e-12 that the velocity errors are measured against")
    return p.parse_args()


def log(msg):
    """progress line on stderr (the JSON result is the only stdout line)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def solve_bench(args):
    """Free-slip Stokes solve to relative residual 1e-8: FP64 multigrid vs the
    mixed-precision one (FP32 V-cycle, FP32 Galerkin matrices, FP64 coarsest
    solve) inside FP64 FGMRES(30); V-cycle: Chebyshev degree 3 (2 sweeps),
    dense direct coarsest solve; velocity lateral level 7 x 129 radial nodes,
    viscosity contrast 1e5 (random harmonics, seed 3) x30 for r > 0.9, buoyancy
    forcing.  Setup (meshes, GCA, Chebyshev eigenvalue estimates) is excluded;
    the timed region is the Krylov solve from a zero initial guess, in the
    mean-free pressure space (free-slip shells leave the constant pressure a
    near-null mode).  Both solves stop at rel. res 1e-8 or after
    --solve-max-iterations; the residual reached is reported."""
    import gc

    import numpy as np
    import torch

    from terraneo_amd import api
    from terraneo_amd import workloads as W
    from terraneo_amd.forward_adjoint import rotation_projection
    from terraneo_amd.problem import StokesProblem

    dev = torch.device("cuda:0")
    out = {}
    sols = {}
    for prec in ("f64", "f32"):
        pr = StokesProblem(args.solve_lat, args.solve_rad, args.solve_levels, device=dev, prec_dtype=prec,
                           gmg=api.gmg_options(cheby_order=3, cheby_iterations=args.solve_cheby_iterations,
                                               power_iterations=10, coarse_solver="dense"))
        coords, radii = pr.coords()
        eta = torch.from_numpy(W.viscosity_c2(coords, radii).reshape(-1)).to(dev)
        f = torch.from_numpy(W.buoyancy(coords, radii)).to(dev)
        t0 = time.perf_counter()
        pr.set_viscosity(eta=eta)
        bu, bp = pr.rhs_from_force(f)
        pr.solve(bu, bp, rtol=1e-8, max_iterations=1, pressure_mean_free=True)  # eigenvalue estimates (setup)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        log(f"solve {prec}: setup {setup:.2f} s")
        # whole FGMRES(30) cycles in calls of 600 iterations, warm-started, each
        # printing a progress line: a call boundary is a restart, which recomputes
        # the true residual exactly as a restart inside one call does, so the
        # iterates are those of a single call; the stopping test keeps 1e-8 of
        # the first call's initial residual
        u = torch.zeros(3 * pr.vel.N, dtype=torch.float64, device=dev)
        p = torch.zeros(pr.pre.N, dtype=torch.float64, device=dev)
        total, secs, r0 = 0, 0.0, None
        while total < args.solve_max_iterations:
            n_it = min(600, args.solve_max_iterations - total)
            kw = dict(rtol=1e-8, atol=1e-300) if r0 is None else dict(rtol=1e-300, atol=1e-8 * r0)
            u, p, step = pr.solve(bu, bp, max_iterations=n_it, restart=30, u=u, p=p, pressure_mean_free=True, **kw)
            if r0 is None:
                r0 = step["initial_residual"]
            total += step["iterations"]
            secs += step["seconds"]
            log(f"solve {prec}: {total} iterations, rel. res {step['final_residual'] / r0:.3e}, {secs:.1f} s")
            if step["final_residual"] <= 1e-8 * r0 or step["iterations"] < n_it:
                break
        info = {"iterations": total, "seconds": secs, "initial_residual": r0}
        res = pr.residual(u, p, bu, bp, pressure_mean_free=True)
        out[prec] = {"seconds": round(info["seconds"], 4), "iterations": info["iterations"],
                     "ms_per_iteration": round(info["seconds"] / max(info["iterations"], 1) * 1e3, 3),
                     "true_rel_residual": float(f"{res:.3e}"), "setup_seconds": round(setup, 3)}
        if prec == "f64":
            sols["u"] = u.clone()
            sols["pr"] = pr
            if args.tight_reference:
                # untimed: the FP64 solution continued with FGMRES(100) until the
                # residual is 1e-12 of the initial one (the reference velocity the
                # errors below are measured against)
                ut, pt = u.clone(), p.clone()
                t1 = time.perf_counter()
                it = pr.solve(bu, bp, rtol=1e-300, atol=1e-12 * info["initial_residual"], max_iterations=3000,
                              restart=100, u=ut, p=pt, pressure_mean_free=True)[2]
                tight_res = pr.residual(ut, pt, bu, bp, pressure_mean_free=True)
                log(f"tight FP64 reference: {it['iterations']} more iterations, rel. res {tight_res:.2e}, "
                    f"{time.perf_counter() - t1:.1f} s")
                sols["tight"] = ut
                out["tight"] = {"iterations_after_fp64": it["iterations"], "true_rel_residual": float(f"{tight_res:.3e}")}
                del pt
        else:
            m = sols["pr"].vel

            def rel(a, b):
                d = a - b
                return float(f"{np.sqrt(m.dot(d, d, 3) / m.dot(b, b, 3)):.3e}")

            def rel_no_rotation(a, b):
                # the difference without its rigid-rotation part (the fs/fs
                # near-null modes, which only the rotation penalty pins)
                d = (a - b).contiguous()
                d = d - rotation_projection(sols["pr"], d)
                return float(f"{np.sqrt(m.dot(d, d, 3) / m.dot(b, b, 3)):.3e}")

            out["velocity_rel_l2_vs_fp64"] = rel(u, sols["u"])
            if "tight" in sols:
                out["tight"]["fp64_velocity_rel_l2"] = rel(sols["u"], sols["tight"])
                out["tight"]["mixed_velocity_rel_l2"] = rel(u, sols["tight"])
                out["tight"]["fp64_velocity_rel_l2_without_rigid_rotations"] = rel_no_rotation(sols["u"], sols["tight"])
                out["tight"]["mixed_velocity_rel_l2_without_rigid_rotations"] = rel_no_rotation(u, sols["tight"])
        del pr, u, p, bu, bp, eta, f
        gc.collect()
        torch.cuda.empty_cache()
    n = 2 ** args.solve_lat + 1
    nr = 2 ** args.solve_rad + 1
    return {
        "workload": f"free-slip Stokes, velocity lateral level {args.solve_lat} ({n}x{n} per diamond) x {nr} radial, "
                    f"{args.solve_levels}-level GMG (Chebyshev degree 3 x{args.solve_cheby_iterations}, dense "
                    "coarsest), FGMRES(30) "
                    "block-triangular, mean-free pressure, target rel. res 1e-8, "
                    "viscosity contrast 1e5 x30 (r > 0.9)",
        "fp64_gmg": out["f64"],
        "mixed_gmg": out["f32"],
        "mixed_over_fp64_time": round(out["f32"]["seconds"] / out["f64"]["seconds"], 4),
        "mixed_over_fp64_time_per_iteration": round(out["f32"]["ms_per_iteration"] / out["f64"]["ms_per_iteration"], 4),
        "velocity_rel_l2_vs_fp64": out["velocity_rel_l2_vs_fp64"],
        "velocity_vs_fp64_converged_to_1e-12": out.get("tight"),
    }


def _host_cpu():
    """model name and online cpu count of the host (for the baseline's record)"""
    model = "unknown cpu"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(seconds=12.0):
    """Time the CPU baseline port (oracle/cpu_port.c: the matrix-free apply in
    this build's cell formulation, OpenMP, FP64, Dirichlet shells, interface
    sum) on a bounded sample: lateral level 6 x 65 radial nodes (the C1 mesh,
    7 987 590 DoF), repeated for about `seconds`.  Threads: OMP_NUM_THREADS,
    else the CPUs this process may run on."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    import oracle_lib as O
    from fields import consistent_random, radial_eta

    O.build()
    om = O.Mesh(6, 65, 0.55, 1.0)
    eta = radial_eta(om.radii, (10, om.n, om.n, om.nr)).reshape(-1)
    x = consistent_random(om, 3, 0).reshape(-1)
    y = np.zeros_like(x)
    O.cpu_port_apply(om, eta, x, O.DIRICHLET, O.DIRICHLET, y)
    t0 = time.perf_counter()
    reps = 0
    while True:
        O.cpu_port_apply(om, eta, x, O.DIRICHLET, O.DIRICHLET, y)
        reps += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = (time.perf_counter() - t0) / reps
    dofs = 3 * int(om.owned.sum())
    gdofs = dofs / dt / 1e9
    try:
        threads = int(os.environ["OMP_NUM_THREADS"])
    except (KeyError, ValueError):
        threads = len(os.sched_getaffinity(0))
    model, ncpu = _host_cpu()
    return {
        "value": round(gdofs, 5),
        "unit": "GDoF/s",
        "cores": threads,
        "kind": "port",
        "sample": f"matrix-free OpenMP C port of the FP64 eps apply (oracle/cpu_port.c, equal to the oracle to "
                  f"1e-12), Dirichlet shells, lateral 6 x 65 radial ({dofs} DoF), {reps} applies in {dt * reps:.1f} s "
                  f"on {threads} threads of {ncpu} CPUs ({model}); the reference's own Kokkos-OpenMP fast path "
                  f"measured 0.0403 GDoF/s at lateral 5 x 33 on 8 Xeon cores in the build container",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import terraneo_amd as tn

    dev = torch.device("cuda:0")
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    esize = 8 if args.dtype == "f64" else 4

    mesh = tn.Mesh(args.lat, args.rad, 0.55, 1.0)
    radii = torch.from_numpy(mesh.radii()).to(dev)
    shape = mesh.shape
    eta = torch.exp(np.log(100.0) * (radii - 0.55) / 0.45).to(tdt).view(1, 1, 1, -1).expand(shape).contiguous()
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = (torch.rand((3,) + shape, generator=g, device=dev, dtype=tdt) * 2 - 1).contiguous()
    mesh.interface_sum(x, 3)  # make duplicated nodes consistent (value irrelevant for timing)
    y = torch.empty_like(x)
    op = tn.Eps(mesh, eta, tn.TN_DIRICHLET, tn.TN_DIRICHLET)

    stream = torch.cuda.current_stream()
    log(f"apply: lateral {args.lat} x {mesh.nr} radial, {args.dtype}")
    for _ in range(args.warmup):
        op.apply(x, y)
    torch.cuda.synchronize()

    # ---- timed: full apply (fill + kernel + interface sum)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.apply(x, y)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    # ---- dominant kernel alone: Replace mode + skipped communication = one
    # launch of the wedge kernel (it writes every node, so there is no fill)
    op.set_modes(tn.TN_REPLACE, tn.TN_COMM_SKIP)
    for _ in range(3):
        op.apply(x, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kreps = max(10, args.steps)
    e0.record(stream)
    for _ in range(kreps):
        op.apply(x, y)
    e1.record(stream)
    torch.cuda.synchronize()
    kernel_ms = e0.elapsed_time(e1) / kreps
    op.set_modes(tn.TN_REPLACE, tn.TN_COMM_ADDITIVE)

    # plain copy of a known size (x -> y, 2 x 3N elements moved): calibrates the
    # HBM byte counters of a rocprofv3 run of this script
    y.copy_(x)
    torch.cuda.synchronize()

    dofs = 3 * mesh.owned_nodes
    value = dofs / (ms_per_step * 1e-3) / 1e9
    alg_bytes = mesh.N * (3 + 1 + 3) * esize  # read x, read eta, write y
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"eps_apply_L{args.lat}x{(1 << args.rad) + 1}_{args.dtype}_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    log(f"apply: {ms_per_step:.4f} ms per apply, kernel {kernel_ms:.4f} ms")
    solve = None
    if not args.no_solve:
        del x, y, eta, op
        torch.cuda.empty_cache()
        solve = solve_bench(args)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, bounded sample)")
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the baseline never blocks the GPU number
            cpu = {"error": str(e)[:200]}

    out = {
        "metric": "Stokes solve s to rel. res 1e-8 (mixed vs FP64 GMG); operator apply GDoF/s",
        "value": round(value, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"epsilon-div-div operator apply (wedge kernel + interface sum), "
                        f"lateral level {args.lat} x {(1 << args.rad) + 1} radial nodes, 10 diamonds, "
                        f"Dirichlet/Dirichlet, eta = exp(ln100 (r-0.55)/0.45)",
            "velocity_dof": dofs,
            "stored_nodes": mesh.N,
        },
        "kernel_ms": round(kernel_ms, 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "solve": solve,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
