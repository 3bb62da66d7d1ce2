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
    sol = spl.spsolve(K, rhs, permc_spec="MMD_ATA")
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
    sol = spl.spsolve(K, rhs, permc_spec="MMD_ATA")
    return O.from_unique(mv, sol[:nu], 3), O.from_unique(mp, sol[nu:nu + npr], 1)


def _rotation_fields_unique(m):
    """Phi [3 nu x 3]: columns e_i x p at every unique node (dof 3*node + comp)."""
    X = m.node_xyz().reshape(10 * m.n * m.n, m.nr, 3)
    g = m.groups.reshape(-1)
    P = np.zeros((m.ngroups, m.nr, 3))
    P[g] = X
    P = P.reshape(-1, 3)
    cols = []
    for i in range(3):
        e = np.zeros(3)
        e[i] = 1.0
        cols.append(np.cross(e, P).reshape(-1))
    return np.stack(cols, axis=1)


def _multiplicity_unique(m):
    g = m.groups.reshape(-1)
    cnt = np.bincount(g, minlength=m.ngroups).astype(float)
    return np.repeat(np.repeat(cnt, m.nr), 3)


def solve_assess(kind, level, penalty_eps=1.0):
    """Direct solve of the reference's assess tests for the free-slip variants.

    freeslip      both shells free-slip, rotation penalty (added per node copy
                  before the interface sum: D_mult eps Phi G^-1 Phi^T), RHS
                  zeroed on both shells
    freezeroslip  free-slip CMB (RHS normal part removed), zero-slip surface
    Returns (vel L2 error, pressure L2 error) over owned dofs, pressure mean removed.
    """
    sol = AS.Solution(kind)
    bc = (O.FREESLIP, O.FREESLIP) if kind == "freeslip" else (O.FREESLIP, O.DIRICHLET)
    mv = O.Mesh(level, (1 << level) + 1, AS.R_MIN, AS.R_MAX)
    mp = O.Mesh(level - 1, (1 << (level - 1)) + 1, AS.R_MIN, AS.R_MAX)
    Xv, Xp = mv.node_xyz(), mp.node_xyz()
    uex = sol.velocity(Xv).reshape(-1)
    pex = sol.pressure(Xp).reshape(-1)
    f = O.vector_mass(mv, sol.body_force(Xv).reshape(-1)).reshape(3, 10, mv.n, mv.n, mv.nr)
    nrm = mv.coords / np.linalg.norm(mv.coords, axis=-1, keepdims=True)
    if kind == "freeslip":
        f[..., 0] = 0.0
        f[..., -1] = 0.0
    else:
        v = f[..., 0]
        fn = np.einsum("cdxy,dxyc->dxy", v, nrm)
        f[..., 0] = v - np.moveaxis(nrm, -1, 0) * fn[None]
        f[..., -1] = 0.0
    bu = f.reshape(-1)
    A = O.eps_assemble(O.Eps(mv, np.ones(mv.N), *bc))
    B = O.div_assemble(mv, mp, *bc)
    nu, npr = A.shape[0], B.shape[0]
    c = sp.csr_matrix(np.ones((1, npr)))
    blocks = [[A, B.T, None], [B, None, c.T], [None, c, None]]
    rhs = [O.to_unique(mv, bu, 3), np.zeros(npr), [0.0]]
    ntot = nu + npr + 1
    b = np.concatenate(rhs)
    if kind != "freeslip":
        K = sp.bmat(blocks, format="csc")
        sol_v = spl.spsolve(K, b, permc_spec="MMD_ATA")
    else:
        # K_pen = K0 + U C Phi^T (penalty, rank 3) is factored through a sparse
        # base K_b = K0 + d E E^T (E: the 6 dofs of two nodes, which no rigid
        # rotation leaves fixed) and Woodbury: K_pen = K_b + L M R^T with
        # L = [U, E], R = [Phi, E], M = diag(C, -d I).
        Phi = _rotation_fields_unique(mv)
        G = Phi.T @ Phi
        C = penalty_eps * np.linalg.inv(G)
        U = _multiplicity_unique(mv)[:, None] * Phi
        node_a, node_b = 3 * (mv.nr // 2), 3 * (mv.ngroups // 2 * mv.nr + mv.nr // 2)
        edofs = [node_a, node_a + 1, node_a + 2, node_b, node_b + 1, node_b + 2]
        d = float(A.diagonal().max())
        Ad = A + sp.csr_matrix((np.full(6, d), (edofs, edofs)), shape=A.shape)
        Kb = sp.bmat([[Ad, B.T, None], [B, None, c.T], [None, c, None]], format="csc")
        lu = spl.splu(Kb, permc_spec="MMD_ATA")
        E = np.zeros((ntot, 6))
        E[edofs, np.arange(6)] = 1.0
        L = np.concatenate([np.vstack([U, np.zeros((npr + 1, 3))]), E], axis=1)
        R = np.concatenate([np.vstack([Phi, np.zeros((npr + 1, 3))]), E], axis=1)
        Minv = np.zeros((9, 9))
        Minv[:3, :3] = np.linalg.inv(C)
        Minv[3:, 3:] = -np.eye(6) / d
        y = lu.solve(b)
        Z = lu.solve(L)
        cap = Minv + R.T @ Z
        sol_v = y - Z @ np.linalg.solve(cap, R.T @ y)
    u, p = sol_v[:nu], sol_v[nu:nu + npr]
    ue, pe = O.to_unique(mv, uex, 3), O.to_unique(mp, pex, 1)
    p, pe = p - p.mean(), pe - pe.mean()
    ev = np.sqrt(np.sum((u - ue) ** 2) / nu)
    ep = np.sqrt(np.sum((p - pe) ** 2) / npr)
    return ev, ep


def solve_freeslip_penalized(mv, mp, eps_op, bu, bp, penalty_eps=1.0, once=True):
    """Direct solve of the free-slip / free-slip Stokes system with the rotation
    penalty eps Phi G^-1 Phi^T counted once per node (once=True, the symmetric
    operator of tn_eps_set_penalty_once) or once per node copy (the reference's
    rule), zero-mean pressure; inputs / outputs in the duplicated layout."""
    A = O.eps_assemble(eps_op)
    B = O.div_assemble(mv, mp, O.FREESLIP, O.FREESLIP)
    nu, npr = A.shape[0], B.shape[0]
    c = sp.csr_matrix(np.ones((1, npr)))
    ntot = nu + npr + 1
    b = np.concatenate([O.to_unique(mv, bu, 3), O.to_unique(mp, bp, 1), [0.0]])
    Phi = _rotation_fields_unique(mv)
    C = penalty_eps * np.linalg.inv(Phi.T @ Phi)
    U = Phi if once else _multiplicity_unique(mv)[:, None] * Phi
    node_a, node_b = 3 * (mv.nr // 2), 3 * (mv.ngroups // 2 * mv.nr + mv.nr // 2)
    edofs = [node_a, node_a + 1, node_a + 2, node_b, node_b + 1, node_b + 2]
    d = float(A.diagonal().max())
    Ad = A + sp.csr_matrix((np.full(6, d), (edofs, edofs)), shape=A.shape)
    Kb = sp.bmat([[Ad, B.T, None], [B, None, c.T], [None, c, None]], format="csc")
    lu = spl.splu(Kb, permc_spec="MMD_ATA")
    E = np.zeros((ntot, 6))
    E[edofs, np.arange(6)] = 1.0
    L = np.concatenate([np.vstack([U, np.zeros((npr + 1, 3))]), E], axis=1)
    R = np.concatenate([np.vstack([Phi, np.zeros((npr + 1, 3))]), E], axis=1)
    Minv = np.zeros((9, 9))
    Minv[:3, :3] = np.linalg.inv(C)
    Minv[3:, 3:] = -np.eye(6) / d
    y = lu.solve(b)
    Z = lu.solve(L)
    sol_v = y - Z @ np.linalg.solve(Minv + R.T @ Z, R.T @ y)
    return O.from_unique(mv, sol_v[:nu], 3), O.from_unique(mp, sol_v[nu:nu + npr], 1)
