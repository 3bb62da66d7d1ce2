"""Direct sparse solve of the oracle's Stokes system (test helper, CPU).

Builds [A B^T; B 0] over unique dofs from the oracle's effective local
matrices, applies the reference's RHS construction (VectorMass of the nodal
body force, strong Dirichlet lifting with the Neumann operator and its
diagonal: strong_algebraic_dirichlet_enforcement.hpp:88-110), and solves with
a zero-mean pressure constraint.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import assess as AS
import oracle_lib as O


def solve_assess_dirichlet(level):
    mv = O.Mesh(level, (1 << level) + 1, AS.R_MIN, AS.R_MAX)
    mp = O.Mesh(level - 1, (1 << (level - 1)) + 1, AS.R_MIN, AS.R_MAX)
    Xv, Xp = mv.node_xyz(), mp.node_xyz()
    uex = AS.velocity(Xv).reshape(-1)
    pex = AS.pressure(Xp).reshape(-1)
    f = O.vector_mass(mv, AS.body_force(Xv).reshape(-1))
    bnd = np.zeros(mv.shape, bool)
    bnd[..., 0] = bnd[..., -1] = True
    bnd3 = np.tile(bnd.reshape(-1), 3)
    g = np.where(bnd3, uex, 0.0)
    k = np.ones(mv.N)
    EN = O.Eps(mv, k, O.NEUMANN, O.NEUMANN)
    bu = f - EN.apply(g)
    bp = -O.div_apply(mv, mp, g, O.NEUMANN, O.NEUMANN)
    bu = np.where(bnd3, EN.apply(g, diagonal=True), bu)
    A = O.eps_assemble(O.Eps(mv, k, O.DIRICHLET, O.DIRICHLET))
    B = O.div_assemble(mv, mp, O.DIRICHLET, O.DIRICHLET)
    nu, npr = A.shape[0], B.shape[0]
    c = sp.csr_matrix(np.ones((1, npr)))
    K = sp.bmat([[A, B.T, None], [B, None, c.T], [None, c, None]], format="csc")
    rhs = np.concatenate([O.to_unique(mv, bu, 3), O.to_unique(mp, bp, 1), [0.0]])
    sol = spl.spsolve(K, rhs)
    u, p = sol[:nu], sol[nu:nu + npr]
    ue, pe = O.to_unique(mv, uex, 3), O.to_unique(mp, pex, 1)
    p, pe = p - p.mean(), pe - pe.mean()
    ev = np.sqrt(np.sum((u - ue) ** 2) / nu)
    ep = np.sqrt(np.sum((p - pe) ** 2) / npr)
    return ev, ep, dict(mv=mv, mp=mp, u=u, p=p, K=K, rhs=rhs, sol=sol)


def stokes_matrix(mv, mp, eps_op, bc):
    """[A B^T 0; B 0 c; 0 c^T 0] over unique dofs (c: mean-zero pressure row),
    A from the oracle operator `eps_op` (its boundary conditions), B with `bc`."""
    A = O.eps_assemble(eps_op)
    B = O.div_assemble(mv, mp, *bc)
    npr = B.shape[0]
    c = sp.csr_matrix(np.ones((1, npr)))
    return sp.bmat([[A, B.T, None], [B, None, c.T], [None, c, None]], format="csc"), A.shape[0], npr


def solve_unique(K, nu, npr, mv, mp, bu, bp):
    """Solve K (u, p, mu) = (bu, bp, 0); inputs/outputs in the duplicated layout."""
    rhs = np.concatenate([O.to_unique(mv, bu, 3), O.to_unique(mp, bp, 1), [0.0]])
    sol = spl.spsolve(K, rhs)
    return O.from_unique(mv, sol[:nu], 3), O.from_unique(mp, sol[nu:nu + npr], 1)
