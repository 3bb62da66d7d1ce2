"""Per-operator timing on one GPU, for profiling the kernels one at a time.

    python -m terraneo_amd.kernel_timing stored [--lat 7] [--reps 10]
        the level lat-1 Galerkin (stored-matrix) operator of a 3-level
        free-slip hierarchy (eta = 1), FP64 and FP32 storage, plain and
        diagonal mode
    python -m terraneo_amd.kernel_timing coupling [--lat 7] [--reps 20]
        B, B^T (Replace and Add) and the full Stokes block apply, velocity
        lateral level lat x (2^lat + 1) radial nodes, free-slip shells
    python -m terraneo_amd.kernel_timing solve [--lat 7] [--reps 10]
        the free-slip Stokes solve of the mixed-precision workload (viscosity
        contrast 1e5 x 30, 5-level GMG, Chebyshev degree 3 x 2, FGMRES(30),
        mean-free pressure): after the setup, `reps` FGMRES iterations in FP64
        and then with the FP32 V-cycle.  A small torch sort launch brackets
        each timed solve, so a kernel trace can be cut into per-iteration
        kernel costs
    python -m terraneo_amd.kernel_timing transfers [--lat 7] [--reps 20]
        vector prolongation (Add) and restriction between lateral levels lat-1
        and lat (radial 2^lat + 1 fine nodes): the constant transfers and the
        geometry-aware linear ones
    python -m terraneo_amd.kernel_timing eps [--lat 7] [--reps 20]
        the matrix-free epsilon-div-div apply (Replace, no interface sum) at
        lateral level lat x (2^lat + 1) radial nodes for each boundary-condition
        pair, FP64 and FP32; free-slip on both shells includes the rotation
        penalty (its moment pass and the fused store)

    python -m terraneo_amd.kernel_timing energy [--lat 7] [--reps 10]
        the FCT energy step on the velocity mesh lateral level lat x
        (2^lat + 1) radial nodes (FV cells 10 (2^lat)^3): stable step, the
        predictor, limiter and correction stages (each with its ghost-layer
        exchanges), the explicit step, and the FV -> FE projection (its
        right-hand side, and the mass-matrix CG to 1e-12); velocity a rigid
        rotation plus noise, diffusivity 1, Dirichlet T on the shells

    python -m terraneo_amd.kernel_timing slice TRACE.csv [--reps 10]
        cuts the rocprofv3 kernel trace (CSV) of a `solve` run at the
        bracketing sort launches and prints, per precision, the kernel time
        and calls per FGMRES iteration, the trace span against the kernel-busy
        time, and the idle time before each kernel grouped by the kernel
        before it (host-side gaps)

Each line reports the mean wall time per call over `reps` calls after three
warm-up calls (stream-synchronised), so that a `rocprofv3 --kernel-trace
--stats` run of the same command attributes the time to kernels.
"""
import argparse
import time

import torch

from . import api
from ._capi import TN_ADD, TN_COMM_SKIP, TN_DIRICHLET, TN_FREESLIP, TN_NEUMANN, TN_REPLACE
from .problem import StokesProblem


def _time(fn, reps):
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def stored(lat, reps):
    dev = torch.device("cuda:0")
    for prec in ("f64", "f32"):
        pr = StokesProblem(lat, lat, 3, bc=(TN_FREESLIP, TN_FREESLIP), device=dev, prec_dtype=prec,
                           coarsest_f64=True, gmg=api.gmg_options(coarse_solver="pcg"))
        pr.set_viscosity(eta=torch.ones(pr.vel.N, dtype=torch.float64, device=dev))
        op, m = pr.ops[-2], pr.meshes[-2]
        dt = torch.float64 if prec == "f64" else torch.float32
        x = torch.randn(3 * m.N, dtype=dt, device=dev)
        y = torch.empty_like(x)
        ncell = 10 * (m.n - 1) ** 2 * (m.nr - 1)
        gb = ncell * 342 * (8 if prec == "f64" else 4) / 1e9
        for diag in (False, True):
            op.set_diagonal(diag)
            ms = _time(lambda: op.apply(x, y), reps)
            print(f"level {lat - 1} stored {prec} diagonal={diag}: {ms:.3f} ms per apply "
                  f"(matrix stream {gb:.2f} GB)", flush=True)
        op.set_diagonal(False)
        del pr
        torch.cuda.empty_cache()


def coupling(lat, reps):
    dev = torch.device("cuda:0")
    vel = api.Mesh(lat, lat, 0.55, 1.0)
    pre = api.Mesh(lat - 1, lat - 1, 0.55, 1.0)
    eta = torch.ones(vel.N, dtype=torch.float64, device=dev)
    A = api.Eps(vel, eta, TN_FREESLIP, TN_FREESLIP)
    K = api.Stokes(vel, pre, A)
    u = torch.randn(3 * vel.N, dtype=torch.float64, device=dev)
    p = torch.randn(pre.N, dtype=torch.float64, device=dev)
    yu, yp = torch.empty_like(u), torch.empty_like(p)
    for name, fn in [("B (div)", lambda: K.div(u, yp)), ("B^T (grad, Replace)", lambda: K.grad(p, yu)),
                     ("B^T (grad, Add)", lambda: K.grad(p, yu, apply_mode=TN_ADD)),
                     ("Stokes block", lambda: K.apply(u, p, yu, yp))]:
        print(f"velocity level {lat}: {name}: {_time(fn, reps):.3f} ms per call", flush=True)


def eps(lat, reps):
    dev = torch.device("cuda:0")
    m = api.Mesh(lat, lat, 0.55, 1.0)
    names = {TN_NEUMANN: "neumann", TN_DIRICHLET: "dirichlet", TN_FREESLIP: "free-slip"}
    for dt in (torch.float64, torch.float32):
        eta = torch.ones(m.N, dtype=dt, device=dev)
        x = torch.randn(3 * m.N, dtype=dt, device=dev)
        y = torch.empty_like(x)
        for bc in ((TN_NEUMANN, TN_NEUMANN), (TN_DIRICHLET, TN_DIRICHLET), (TN_FREESLIP, TN_DIRICHLET),
                   (TN_FREESLIP, TN_FREESLIP)):
            A = api.Eps(m, eta, *bc)
            A.set_modes(TN_REPLACE, TN_COMM_SKIP)
            ms = _time(lambda: A.apply(x, y), reps)
            print(f"level {lat} eps {str(dt)[6:]} {names[bc[0]]}/{names[bc[1]]}: {ms:.3f} ms per apply", flush=True)
            del A


def solve(lat, reps):
    import gc

    from . import workloads as W

    dev = torch.device("cuda:0")

    def marker():
        torch.cuda.synchronize()
        torch.sort(torch.rand(100, device=dev))
        torch.cuda.synchronize()

    for prec in ("f64", "f32"):
        pr = StokesProblem(lat, lat, 5, device=dev, prec_dtype=prec,
                           gmg=api.gmg_options(cheby_order=3, cheby_iterations=2, power_iterations=10,
                                               coarse_solver="dense"))
        coords, radii = pr.coords()
        pr.set_viscosity(eta=torch.from_numpy(W.viscosity_c2(coords, radii).reshape(-1)).to(dev))
        bu, bp = pr.rhs_from_force(torch.from_numpy(W.buoyancy(coords, radii)).to(dev))
        pr.solve(bu, bp, rtol=1e-8, max_iterations=1, pressure_mean_free=True)  # setup: eigenvalue estimates
        marker()
        t0 = time.perf_counter()
        u, p, info = pr.solve(bu, bp, rtol=1e-8, max_iterations=reps, restart=30, pressure_mean_free=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        marker()
        # wall time includes allocating and zeroing the Krylov workspace
        # (2 x restart + 4 block vectors); the solve's own timer starts after it
        print(f"velocity level {lat} solve {prec}: {info['iterations']} FGMRES iterations, "
              f"{dt / info['iterations'] * 1e3:.2f} ms per iteration (wall), "
              f"{info['seconds'] / info['iterations'] * 1e3:.2f} ms per iteration (solve timer)", flush=True)
        del pr, u, p, bu, bp
        gc.collect()
        torch.cuda.empty_cache()


def transfers(lat, reps):
    dev = torch.device("cuda:0")
    mf, mc = api.Mesh(lat, lat, 0.55, 1.0), api.Mesh(lat - 1, lat - 1, 0.55, 1.0)
    xf = torch.randn(3 * mf.N, dtype=torch.float64, device=dev)
    xc = torch.randn(3 * mc.N, dtype=torch.float64, device=dev)
    for name, fn in [("constant P (Add)", lambda: api.prolongate_add(mc, mf, xc, xf)),
                     ("constant R", lambda: api.restrict(mf, mc, xf, xc)),
                     ("linear P (Add)", lambda: api.prolongate_linear(mc, mf, xc, xf, add=True)),
                     ("linear R", lambda: api.restrict_linear(mf, mc, xf, xc))]:
        print(f"velocity level {lat}: {name}: {_time(fn, reps):.3f} ms per call", flush=True)


def energy(lat, reps):
    dev = torch.device("cuda:0")
    m = api.Mesh(lat, lat, 0.55, 1.0)
    fv = api.FV(m)
    n, nr = m.n, m.nr
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    xyz = torch.from_numpy(m.lateral_coords()).to(dev)[:, :, :, None, :] * \
        torch.from_numpy(m.radii()).to(dev)[None, None, None, :, None]
    omega = torch.tensor([0.3, -0.2, 0.7], dtype=torch.float64, device=dev)
    u = torch.cross(omega.expand_as(xyz), xyz, dim=-1).movedim(-1, 0).contiguous()
    u += 0.05 * torch.rand(u.shape, generator=g, device=dev, dtype=torch.float64)
    u = u.reshape(-1).contiguous()
    T = torch.rand(fv.fv_shape, generator=g, device=dev, dtype=torch.float64)
    fv.apply_dirichlet(T, cmb=1.0, surf=0.0)
    kappa = 1.0
    dt = 0.5 * fv.dt_stable(u, kappa)
    cells = 10 * (n - 1) ** 2 * (nr - 1)
    b = torch.empty(m.N, dtype=torch.float64, device=dev)
    Tfe = torch.zeros(m.N, dtype=torch.float64, device=dev)
    print(f"energy step, {cells} FV cells (lateral level {lat} x {nr} radial nodes), dt = {dt:.4e}", flush=True)
    for name, fn in [("stable step (reduction)", lambda: fv.dt_stable(u, kappa)),
                     ("predictor (+ ghosts of T)", lambda: fv.predictor(T, u, dt, kappa)),
                     ("limiter (+ ghosts of T_L, R+, R-)", lambda: fv.limiter()),
                     ("correction", lambda: fv.correction(T)),
                     ("explicit step", lambda: fv.explicit_step(T, u, dt, kappa, None, True, cmb=1.0, surf=0.0)),
                     ("FV -> FE right-hand side", lambda: fv.to_fe_rhs(T, b))]:
        ms = _time(fn, reps)
        print(f"  {name}: {ms:.3f} ms per call ({cells / ms / 1e6:.2f} Gcells/s)", flush=True)
    info = fv.to_fe(T, Tfe)
    print(f"  FV -> FE projection (CG on the mass matrix from 0 to 1e-12): {info['iterations']} iterations, "
          f"{info['seconds'] * 1e3:.2f} ms", flush=True)


def slice_trace(path, reps):
    """Per-iteration kernel costs of the two solves in a kernel trace of `solve`."""
    import collections
    import csv

    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    marks = [i for i, r in enumerate(rows) if "sort" in r["Kernel_Name"].lower()]
    clusters = []
    for i in marks:
        if not clusters or i - clusters[-1][-1] > 3:
            clusters.append([i])
        else:
            clusters[-1].append(i)
    if len(clusters) < 4:
        raise SystemExit("trace: expected two bracketed solves (4 sort launches)")
    spans = [(clusters[0][-1], clusters[1][0]), (clusters[2][-1], clusters[3][0])]
    for name, (a, b) in zip(("f64", "mixed"), spans):
        seg = rows[a + 1:b]
        t0, t1 = int(seg[0]["Start_Timestamp"]), int(seg[-1]["End_Timestamp"])
        tot = collections.defaultdict(lambda: [0, 0.0])
        gaps = collections.defaultdict(float)
        busy = 0.0
        for p, r in zip([None] + seg[:-1], seg):
            dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
            k = r["Kernel_Name"][:100]
            tot[k][0] += 1
            tot[k][1] += dur
            busy += dur
            if p is not None:
                g = (int(r["Start_Timestamp"]) - int(p["End_Timestamp"])) / 1e6
                if g > 0:
                    gaps[p["Kernel_Name"][:70]] += g
        span = (t1 - t0) / 1e6
        print(f"== {name}: {len(seg)} kernels, per iteration span {span / reps:.2f} ms, "
              f"kernel-busy {busy / reps:.2f} ms")
        for k, (c, d) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:24]:
            print(f"  {d / reps:8.3f} ms/it {c / reps:7.1f} calls/it {d / c * 1e3:9.1f} us  {k}")
        print(f"  idle before a kernel, by the kernel before it (ms/it):")
        for k, g in sorted(gaps.items(), key=lambda kv: -kv[1])[:4]:
            print(f"  {g / reps:8.3f}  {k}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["stored", "coupling", "eps", "solve", "transfers", "energy", "slice"])
    ap.add_argument("trace", nargs="?", help="slice: rocprofv3 kernel trace CSV of a `solve` run")
    ap.add_argument("--lat", type=int, default=7)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    if a.what == "slice":
        if a.trace is None:
            ap.error("slice needs the kernel trace CSV")
        slice_trace(a.trace, a.reps)
        return
    {"stored": stored, "coupling": coupling, "eps": eps, "solve": solve, "transfers": transfers,
     "energy": energy}[a.what](a.lat, a.reps)


if __name__ == "__main__":
    main()
