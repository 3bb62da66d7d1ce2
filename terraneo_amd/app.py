"""The Stokes half of the reference's mantle-circulation time step, on the C ABI
(apps/mantlecirculation/mantlecirculation.cpp:866-932).

Per time step, as the reference app does:

1. right-hand side from the temperature T: the nodal buoyancy Ra T n (n the
   unit position vector) times the vector mass matrix, with the velocity
   boundary conditions enforced strongly (`tn_stokes_rhs_from_temperature`,
   :876-898);
2. FGMRES with the block-triangular preconditioner (velocity V-cycle, lumped
   pressure mass with 1/eta), warm-started from the previous step's (u, p)
   (:900-903; the solver set-up of :486-720 is `problem.StokesProblem`);
3. pressure "normalisation": the owned mean of p is subtracted
   (`tn_pressure_normalize`, :926-932);
4. one row per step in a timer table (the reference prints its FGMRES table
   and a util::Timer per step).

The energy step that follows (FV/FCT advection-diffusion of T, :736-1012) is
`MantleCirculationEnergy`: FV temperature at the cell centres with Dirichlet
shells (T_cmb = 1, T_surface = 0), the stable step dt_scaling x dt_stable from
the new velocity, `energy_substeps` explicit FCT steps (each followed by the
Dirichlet values), and the L2 projection of T to the nodes, which the next
Stokes right-hand side uses.  `--energy` couples the two halves; without it
the caller supplies T for every step.

Defaults are the app's (apps/mantlecirculation/src/parameters.hpp:14-67): mesh
levels 1..4, r in [0.5, 1], no-slip shells, Ra = 1e5, FGMRES(10) with at most 10
iterations per step to rel. 1e-6 / abs. 1e-12, one V-cycle with Chebyshev order
2 and 2 smoothing steps, 10 power iterations, coarsest-level PCG(50, 1e-6, 1e-16).

    python -m terraneo_amd.app [--lat 4] [--steps 5] [--free-slip] [--prec f32] [--energy] [--xdmf DIR]

With --xdmf, every step also writes the velocity and the temperature as an XDMF
time series / checkpoint (terraneo_amd.xdmf), as the app's output block does.
"""
import argparse
import time
from dataclasses import dataclass

import numpy as np

from . import api
from ._capi import TN_DIRICHLET, TN_FREESLIP
from .problem import StokesProblem


def _torch():
    import torch

    return torch


@dataclass
class AppParameters:
    """apps/mantlecirculation/src/parameters.hpp defaults (refinement levels 1..4)."""

    lat: int = 4                      # finest velocity level (refinement_level_mesh_max)
    rad: int = 4
    levels: int = 4                   # max - min + 1
    r_min: float = 0.5
    r_max: float = 1.0
    bc: tuple = (TN_DIRICHLET, TN_DIRICHLET)
    rayleigh: float = 1e5
    krylov_restart: int = 10
    krylov_max_iterations: int = 10
    krylov_rtol: float = 1e-6
    krylov_atol: float = 1e-12
    cheby_order: int = 2
    smoothing_steps: int = 2
    power_iterations: int = 10
    vcycles: int = 1
    prec_dtype: str = "f64"           # "f32": FP32 V-cycle inside the FP64 FGMRES
    pressure_mean_free: bool = False  # solve in the mean-free pressure space (free-slip shells)


class MantleCirculationStokes:
    """Stokes solves of consecutive time steps with warm starts.  `eta` is the
    nodal viscosity on the velocity mesh (default 1)."""

    def __init__(self, prm=None, eta=None, device="cuda:0"):
        torch = _torch()
        self.prm = prm = prm or AppParameters()
        gmg = api.gmg_options(cheby_order=prm.cheby_order, cheby_iterations=prm.smoothing_steps,
                              power_iterations=prm.power_iterations, num_cycles=prm.vcycles,
                              coarse_solver="pcg", coarse_max_iterations=50, coarse_rtol=1e-6, coarse_atol=1e-16)
        self.pr = StokesProblem(prm.lat, prm.rad, prm.levels, r_min=prm.r_min, r_max=prm.r_max, bc=prm.bc,
                                device=device, prec_dtype=prm.prec_dtype, gmg=gmg)
        f64 = self.pr.f64
        self.pr.set_viscosity(eta=eta if eta is not None else torch.ones(self.pr.vel.N, **f64))
        self.u = torch.zeros(3 * self.pr.vel.N, **f64)
        self.p = torch.zeros(self.pr.pre.N, **f64)
        self.fu = torch.empty_like(self.u)
        self.fp = torch.empty_like(self.p)
        self.timestep = 0
        self.table = []

    def initial_temperature(self):
        """((r_max - r) / (r_max - r_min))^5 at every node copy
        (InitialConditionInterpolator, mantlecirculation.cpp:74-113)."""
        torch = _torch()
        r = torch.from_numpy(np.asarray(self.pr.vel.radii(), dtype=np.float64)).to(self.pr.device)
        frac = ((self.prm.r_max - r) / (self.prm.r_max - self.prm.r_min)) ** 5
        return frac.expand(self.pr.vel.N // self.pr.vel.nr, -1).reshape(-1).contiguous()

    def step(self, temperature):
        """One Stokes solve for the temperature T[10][n][n][nr] (FP64, velocity
        mesh); returns the table row."""
        torch = _torch()
        prm, pr = self.prm, self.pr
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        api.stokes_rhs_from_temperature(pr.K, temperature.contiguous(), prm.rayleigh, self.fu, self.fp)
        _, _, info = pr.solve(self.fu, self.fp, rtol=prm.krylov_rtol, atol=prm.krylov_atol,
                              max_iterations=prm.krylov_max_iterations, restart=prm.krylov_restart,
                              u=self.u, p=self.p, pressure_mean_free=prm.pressure_mean_free)
        api.pressure_normalize(pr.pre, self.p)
        torch.cuda.synchronize()
        self.timestep += 1
        row = {"timestep": self.timestep, "iterations": info["iterations"],
               "initial_residual": info["initial_residual"], "final_residual": info["final_residual"],
               "solve_s": info["seconds"], "step_s": time.perf_counter() - t0}
        self.table.append(row)
        return row

    def table_text(self):
        head = f"{'timestep':>8} {'iterations':>10} {'initial res':>12} {'final res':>12} {'solve s':>9} {'step s':>9}"
        lines = [head]
        for r in self.table:
            lines.append(f"{r['timestep']:>8d} {r['iterations']:>10d} {r['initial_residual']:>12.4e} "
                         f"{r['final_residual']:>12.4e} {r['solve_s']:>9.4f} {r['step_s']:>9.4f}")
        return "\n".join(lines)


class MantleCirculationEnergy:
    """The FCT energy step of the app's time loop on the velocity mesh of
    `stokes` (mantlecirculation.cpp:736-775 set-up, :935-1012 per step).

    Initial temperature: ((r_max - |c|) / (r_max - r_min))^5 at the cell
    centres c (FVInitialConditionInterpolator, :198-216) plus uniform noise in
    [-0.1, 0.1], clamped to [0, 1] (FVNoiseAdder, :218-234; drawn from torch's
    generator with the app's seed 12345, so the values differ from Kokkos'
    random pool), Dirichlet values on the shells, ghost layers, and the L2
    projection to the nodes."""

    def __init__(self, stokes, diffusivity=1.0, dt_scaling=0.5, substeps=1, t_cmb=1.0, t_surface=0.0,
                 internal_heating=None, noise=0.1, seed=12345):
        torch = _torch()
        self.stokes = stokes
        prm = stokes.prm
        vel = stokes.pr.vel
        self.fv = api.FV(vel)
        self.kappa, self.dt_scaling, self.substeps = diffusivity, dt_scaling, substeps
        self.t_cmb, self.t_surface = t_cmb, t_surface
        c = self.fv.cell_centers()
        radius = torch.sqrt((c * c).sum(-1))
        T = ((prm.r_max - radius) / (prm.r_max - prm.r_min)) ** 5
        g = torch.Generator(device=T.device)
        g.manual_seed(seed)
        pert = noise * (2.0 * torch.rand(T.shape, generator=g, device=T.device, dtype=T.dtype) - 1.0)
        real = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        self.T = torch.zeros_like(T)
        self.T[real] = torch.clamp(T[real] + pert[real], 0.0, 1.0)
        self.fv.apply_dirichlet(self.T, cmb=t_cmb, surf=t_surface)
        self.fv.ghost_update(self.T)
        self.source = None
        if internal_heating is not None:
            self.source = torch.full_like(self.T, float(internal_heating))
        self.T_nodal = torch.zeros(vel.N, dtype=torch.float64, device=T.device)
        self.fv.to_fe(self.T, self.T_nodal)
        self.table = []

    def step(self, u):
        """Energy step with the velocity u (FP64 nodal): returns the table row."""
        torch = _torch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dt_stable = self.fv.dt_stable(u, self.kappa)
        dt = self.dt_scaling * dt_stable
        for _ in range(self.substeps):
            self.fv.explicit_step(self.T, u, dt, self.kappa, self.source, True, cmb=self.t_cmb, surf=self.t_surface)
            self.fv.apply_dirichlet(self.T, cmb=self.t_cmb, surf=self.t_surface)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info = self.fv.to_fe(self.T, self.T_nodal)
        torch.cuda.synchronize()
        real = self.T[:, 1:-1, 1:-1, 1:-1]
        row = {"dt": dt, "dt_stable": dt_stable, "fct_s": t1 - t0, "projection_iterations": info["iterations"],
               "projection_s": time.perf_counter() - t1, "T_min": float(real.min()), "T_max": float(real.max())}
        self.table.append(row)
        return row


def main():
    ap = argparse.ArgumentParser(description="Stokes solves of the mantle-circulation time step (warm-started)")
    ap.add_argument("--lat", type=int, default=4)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--free-slip", action="store_true", help="free-slip shells (mean-free pressure)")
    ap.add_argument("--prec", default="f64", choices=["f64", "f32"])
    ap.add_argument("--energy", action="store_true",
                    help="couple the FCT energy step: T for each Stokes solve comes from the previous energy step")
    ap.add_argument("--xdmf", default=None, help="write an XDMF time series / checkpoint to this directory")
    a = ap.parse_args()
    torch = _torch()
    prm = AppParameters(lat=a.lat, rad=a.lat, levels=a.levels, prec_dtype=a.prec)
    if a.free_slip:
        prm.bc = (TN_FREESLIP, TN_FREESLIP)
        prm.pressure_mean_free = True
    app = MantleCirculationStokes(prm)
    T = app.initial_temperature()
    # a temperature that changes a little per step, as the energy step would
    lat = torch.from_numpy(np.asarray(app.pr.vel.lateral_coords()).reshape(-1, 3)).to(T.device)
    out, Tk = None, T
    if a.xdmf:
        from .xdmf import XDMFOutput

        out = XDMFOutput(a.xdmf, app.pr.vel)
        out.add("velocity", lambda: app.u, vec_dim=3)
        out.add("temperature", lambda: Tk)
    energy = MantleCirculationEnergy(app) if a.energy else None
    for k in range(a.steps):
        if energy is not None:
            Tk = energy.T_nodal
        else:
            wobble = 0.02 * torch.sin(3.0 * lat[:, 0] + 0.3 * k)[:, None].expand(-1, app.pr.vel.nr).reshape(-1)
            Tk = T * (1.0 + wobble)
        row = app.step(Tk)
        msg = (f"timestep {row['timestep']}: {row['iterations']} FGMRES iterations, "
               f"residual {row['initial_residual']:.3e} -> {row['final_residual']:.3e}, "
               f"{row['step_s'] * 1e3:.1f} ms")
        if energy is not None:
            e = energy.step(app.u)
            msg += (f"; energy dt {e['dt']:.3e}, FCT {e['fct_s'] * 1e3:.2f} ms, projection "
                    f"{e['projection_iterations']} CG iterations {e['projection_s'] * 1e3:.1f} ms, "
                    f"T in [{e['T_min']:.3f}, {e['T_max']:.3f}]")
        if out is not None:
            out.write()
        print(msg, flush=True)
    print(app.table_text(), flush=True)


if __name__ == "__main__":
    main()
