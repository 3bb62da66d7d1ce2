"""Assembled Stokes problem on the thick shell: the way the reference's apps
wire its pieces together (apps/mantlecirculation/mantlecirculation.cpp:486-930,
tests/test_epsilon_divdiv_stokes_assess*.cpp), on top of the C ABI.

Velocity mesh hierarchy (coarse -> fine) for the multigrid, pressure on the
once-coarsened mesh (Q1-iso-Q2 / Q1), the viscous block on the finest level
matrix-free and Galerkin (GCA) matrices on the coarser ones, lumped
M_p(1/eta) as the Schur-complement preconditioner, and a FP64 Krylov solver
(MINRES with the block-diagonal or FGMRES with the block-triangular
preconditioner).  The velocity multigrid runs in FP64 or FP32 (mixed
precision: FP32 V-cycle and FP32 Galerkin matrices inside the FP64 FGMRES,
whose restarts recompute the true FP64 residual -- iterative refinement --
with the coarsest-level solve kept in FP64).

torch provides device memory only; every operation is a library call.
"""
import numpy as np

from . import api
from ._capi import (TN_DIRICHLET, TN_F32, TN_F64, TN_FGMRES_BLOCKTRI, TN_FREESLIP, TN_MINRES_BLOCKDIAG)


def _torch():
    import torch

    return torch


class StokesProblem:
    """lat/rad: velocity refinement levels of the finest mesh; levels: multigrid
    levels (the coarsest has lateral level lat-levels+1); bc: (CMB, surface)."""

    def __init__(self, lat, rad, levels, r_min=0.55, r_max=1.0, bc=(TN_FREESLIP, TN_FREESLIP), device="cuda:0",
                 prec_dtype="f64", coarsest_f64=True, gmg=None, gca_bracket="reference", coarse_ops="gca",
                 mg_penalty=None, penalty="reference"):
        torch = _torch()
        assert levels >= 2 and lat - levels + 1 >= 0 and rad - levels + 1 >= 1
        self.device = torch.device(device)
        self.bc = bc
        self.levels = [(lat - k, rad - k) for k in range(levels - 1, -1, -1)]  # coarse -> fine
        self.meshes = [api.Mesh(l, r, r_min, r_max) for l, r in self.levels]
        self.vel = self.meshes[-1]
        self.pre = api.Mesh(lat - 1, rad - 1, r_min, r_max)
        self.prec_dtype = prec_dtype
        self.coarsest_f64 = coarsest_f64
        self.gca_bracket = gca_bracket
        assert coarse_ops in ("gca", "rediscretize")
        # "gca": Galerkin matrices (gca.hpp); "rediscretize": matrix-free coarse
        # operators with the nodal viscosity injected (the assess tests' setup)
        self.coarse_ops = coarse_ops
        # fs/fs rotation penalty of the multigrid level operators: None keeps the
        # operator's epsilon (1, the reference's); a number s sets epsilon to s
        # times the level's mean diagonal entry, so that the penalised rotation
        # modes sit inside the spectrum the Chebyshev smoother targets (with
        # epsilon = 1 they leave it as the mesh is refined).  The outer Stokes
        # operator always keeps epsilon = 1.
        self.mg_penalty = mg_penalty
        # fs/fs rotation penalty multiplicity: "reference" (every node copy, the
        # reference's rule: the operator is then not symmetric) or "once" (the
        # owned copy only: a symmetric operator, as MINRES and CG require)
        assert penalty in ("reference", "once")
        self.penalty = penalty
        self.gmg_opts = gmg or api.gmg_options(cheby_order=3, cheby_iterations=1, power_iterations=10, num_cycles=1)
        self.f64 = dict(dtype=torch.float64, device=self.device)
        self.A = None
        self.K = None
        self.gmg = None

    # ---------------------------------------------------------------- viscosity
    def coords(self, mesh=None):
        m = mesh or self.vel
        return m.lateral_coords(), m.radii()

    def set_viscosity(self, eta=None, eta_cell=None):
        """Nodal viscosity on the finest velocity mesh [10][n][n][nr] or per hex
        cell [10][n-1][n-1][nr-1] (torch FP64 on the device).  Rebuilds the
        viscous block, the multigrid hierarchy and the Schur preconditioner."""
        torch = _torch()
        self.eta = eta.contiguous() if eta is not None else None
        self.eta_cell = eta_cell.contiguous() if eta_cell is not None else None
        bc = self.bc
        self.A = api.Eps(self.vel, self.eta, *bc, eta_cell=self.eta_cell)
        self.K = api.Stokes(self.vel, self.pre, self.A)
        # velocity multigrid: finest level matrix-free in the preconditioner dtype
        pd = torch.float64 if self.prec_dtype == "f64" else torch.float32
        st = TN_F64 if self.prec_dtype == "f64" else TN_F32
        if pd == torch.float64:
            self._eta_p = self.eta
            self._eta_cell_p = self.eta_cell
            fine = self.A if self.mg_penalty is None else api.Eps(self.vel, self.eta, *bc, eta_cell=self.eta_cell)
        else:
            self._eta_p = self.eta.to(pd) if self.eta is not None else None
            self._eta_cell_p = self.eta_cell.to(pd) if self.eta_cell is not None else None
            fine = api.Eps(self.vel, self._eta_p, *bc, eta_cell=self._eta_cell_p)
        ops = [fine]
        self._dummies = []
        if self.coarse_ops == "rediscretize":
            assert self.eta is not None, "rediscretized coarse operators need a nodal viscosity"
            eta_l = self.eta.view(10, self.vel.n, self.vel.n, self.vel.nr)
        for k, m in enumerate(reversed(self.meshes[:-1])):
            coarsest = k == len(self.meshes) - 2
            dt = torch.float64 if (coarsest and self.coarsest_f64) else pd
            sdt = TN_F64 if dt == torch.float64 else st
            if self.coarse_ops == "rediscretize":
                eta_l = eta_l[:, ::2, ::2, ::2]
                e = eta_l.contiguous().view(-1).to(dt)
                self._dummies.append(e)
                op = api.Eps(m, e, *bc)
            else:
                one = torch.ones(m.N, dtype=dt, device=self.device)  # placeholder viscosity, unused once stored
                self._dummies.append(one)
                op = api.Eps(m, one, *bc)
                op.gca_from(ops[-1], sdt, bracket=self.gca_bracket)
            ops.append(op)
        self.ops = list(reversed(ops))
        if self.penalty == "once":
            for op in self.ops + [self.A]:
                op.set_penalty_once(True)
        if self.mg_penalty is not None and bc == (TN_FREESLIP, TN_FREESLIP):
            for op, m in zip(self.ops, self.meshes):
                op.set_penalty_epsilon(float(self.mg_penalty) * self._mean_diagonal(op, m))
        self.gmg = api.Gmg(self.ops, self.gmg_opts)
        # Schur preconditioner: lumped M_p(1/eta), eta sampled at the pressure nodes
        self.schur = torch.empty(self.pre.N, **self.f64)
        api.kmass_lumped(self.pre, (1.0 / self.pressure_viscosity()).contiguous(), self.schur)

    def _mean_diagonal(self, op, mesh):
        """Mean over owned velocity dofs of the operator's diagonal."""
        torch = _torch()
        dt = torch.float64 if op.dtype == TN_F64 else torch.float32
        one = torch.ones(3 * mesh.N, dtype=dt, device=self.device)
        d = torch.empty_like(one)
        op.set_diagonal(True)
        op.apply(one, d)
        op.set_diagonal(False)
        return float(mesh.dot(d, one, 3)) / (3 * mesh.owned_nodes)

    def pressure_viscosity(self):
        """Viscosity at the pressure nodes (the velocity nodes of even index)."""
        torch = _torch()
        if self.eta is not None:
            e = self.eta.view(10, self.vel.n, self.vel.n, self.vel.nr)
            return e[:, ::2, ::2, ::2].contiguous().view(-1)
        # per-cell field: geometric mean of the cells around each pressure node
        lc = torch.log(self.eta_cell.view(10, self.vel.n - 1, self.vel.n - 1, self.vel.nr - 1))
        lc = torch.nn.functional.pad(lc[None], (1, 1, 1, 1, 1, 1), mode="replicate")[0]
        node = 0.125 * sum(lc[:, i:i + self.vel.n, j:j + self.vel.n, k:k + self.vel.nr]
                           for i in (0, 1) for j in (0, 1) for k in (0, 1))
        node = node[:, ::2, ::2, ::2].contiguous().view(-1)
        v = torch.exp(node)
        # consistent across the diamonds' duplicated interface nodes
        cnt = torch.ones_like(v)
        self.pre.interface_sum(v, 1)
        self.pre.interface_sum(cnt, 1)
        return v / cnt

    # ---------------------------------------------------------------- right-hand side
    def rhs_from_force(self, f_nodal):
        """bu = M f with the free-slip normal components removed and zero on
        Dirichlet shells (homogeneous data), bp = 0."""
        torch = _torch()
        bu = torch.empty(3 * self.vel.N, **self.f64)
        api.vector_mass(self.vel, f_nodal.contiguous().view(-1), bu)
        zero = None
        for which, mode in ((1, self.bc[0]), (2, self.bc[1])):
            if mode == TN_FREESLIP:
                api.freeslip_project(self.vel, bu, which)
            elif mode == TN_DIRICHLET:
                zero = torch.zeros_like(bu) if zero is None else zero
                api.boundary_assign(self.vel, bu, zero, which)
        bp = torch.zeros(self.pre.N, **self.f64)
        return bu, bp

    # ---------------------------------------------------------------- solve
    def solve(self, bu, bp, method="fgmres", rtol=1e-8, max_iterations=500, restart=30, u=None, p=None,
              pressure_mean_free=False, atol=1e-30):
        """pressure_mean_free: solve with the pressure constrained to zero mean
        (see tn_krylov_options.pressure_mean_free)."""
        torch = _torch()
        u = torch.zeros(3 * self.vel.N, **self.f64) if u is None else u
        p = torch.zeros(self.pre.N, **self.f64) if p is None else p
        meth = TN_FGMRES_BLOCKTRI if method == "fgmres" else TN_MINRES_BLOCKDIAG
        info = api.stokes_solve(self.K, meth, self.gmg, self.schur, u, p, bu, bp, max_iterations=max_iterations,
                                rtol=rtol, atol=atol, restart=restart, pressure_mean_free=pressure_mean_free)
        return u, p, info

    def residual(self, u, p, bu, bp, pressure_mean_free=False):
        """Relative residual ||b - K x|| / ||b|| in the owned-node norm
        (pressure_mean_free: of the mean-projected system, i.e. with the
        pressure residual's owned mean removed)."""
        torch = _torch()
        yu, yp = torch.empty_like(u), torch.empty_like(p)
        self.K.apply(u, p, yu, yp)
        ru, rp = bu - yu, bp - yp
        if pressure_mean_free:
            one = torch.ones_like(rp)
            rp = rp - self.pre.dot(rp, one, 1) / self.pre.dot(one, one, 1)
        num = self.vel.dot(ru, ru, 3) + self.pre.dot(rp, rp, 1)
        den = self.vel.dot(bu, bu, 3) + self.pre.dot(bp, bp, 1)
        return float(np.sqrt(num / den))

    def l2_owned(self, v, mesh, ncomp):
        return float(np.sqrt(mesh.dot(v, v, ncomp) / (ncomp * mesh.owned_nodes)))
