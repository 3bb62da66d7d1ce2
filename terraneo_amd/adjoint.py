"""Differentiation of the Stokes solve with respect to a per-cell
log-viscosity, as a torch.autograd.Function forward/backward pair.

forward:  m [10][n-1][n-1][nr-1] -> eta_cell = exp(m) -> solve K(m) (u, p) = b
backward: given dJ/du (torch's Euclidean gradient over the stored node
          copies), form its owned-inner-product representative (sum over the
          copies of every node), solve the adjoint system K (lu, lp) = (g, 0)
          with the same solver (K is symmetric; on free-slip shells with the
          penalty once per node the rigid-rotation part of g is removed first,
          which leaves the gradient unchanged), then
          dJ/dm[cell] = -eta_cell (T lu)^T A_cell(1) u      (library kernel)

The reference has no adjoint; the discrete gradient is exact for the discrete
solve up to the solver tolerance, which is what the finite-difference checks
measure.
"""
from ._capi import TN_FREESLIP


def _torch():
    import torch

    return torch


def make_function():
    torch = _torch()

    class StokesLogViscosity(torch.autograd.Function):
        @staticmethod
        def forward(ctx, m, problem, bu, bp, solve_kw):
            eta_cell = torch.exp(m.detach()).contiguous()
            problem.set_viscosity(eta_cell=eta_cell)
            u, p, info = problem.solve(bu, bp, **solve_kw)
            ctx.problem = problem
            ctx.solve_kw = solve_kw
            ctx.eta_cell = eta_cell
            ctx.save_for_backward(u)
            ctx.info = info
            return u

        @staticmethod
        def backward(ctx, gu):
            (u,) = ctx.saved_tensors
            problem = ctx.problem
            if problem.eta_cell is None or problem.eta_cell.data_ptr() != ctx.eta_cell.data_ptr():
                problem.set_viscosity(eta_cell=ctx.eta_cell)
            g = gu.detach().contiguous().clone()
            problem.vel.interface_sum(g, 3)
            for which, mode in ((1, problem.bc[0]), (2, problem.bc[1])):
                if mode == TN_FREESLIP:
                    from . import api

                    api.freeslip_project(problem.vel, g, which)
            if problem.bc == (TN_FREESLIP, TN_FREESLIP) and problem.penalty == "once":
                # rigid rotations: an invariant subspace of the symmetric K without
                # strain, so their part of g cannot change the gradient; removing it
                # spares the adjoint solve its penalty-scaled, slowly converging part
                from .forward_adjoint import remove_rotations

                g = remove_rotations(problem, g)
            gp = torch.zeros(problem.pre.N, dtype=torch.float64, device=u.device)
            lu, _, info = problem.solve(g, gp, **ctx.solve_kw)
            ctx.adjoint_info = info
            grad = torch.empty_like(ctx.eta_cell)
            problem.A.grad_log_cell_eta(u, lu, grad)
            return grad.view_as(ctx.eta_cell), None, None, None, None

    return StokesLogViscosity


_FN = None


def stokes_velocity(m, problem, bu, bp, **solve_kw):
    """u(m): differentiable velocity of the Stokes solve (see module docstring)."""
    global _FN
    if _FN is None:
        _FN = make_function()
    return _FN.apply(m, problem, bu, bp, solve_kw)
