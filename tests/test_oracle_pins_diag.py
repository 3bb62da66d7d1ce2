"""Pins of the oracle's diagonal mode and of the Galerkin coarsening's
reference radial bracket, by what the mathematics fixes.

Diagonal mode (epsilon_divdiv_kerngen.hpp:1351-1384 and its free-slip
counterpart in :1421-1960):
  * Neumann / Dirichlet shells: y = diag(A) x, with A the oracle's assembled
    operator for the same boundary conditions (built from the local matrices
    and masks, a different code path from the diagonal branch).
  * free-slip shell nodes: y = R^T diag(R B R^T) R x, B the node's assembled
    3x3 block of the Neumann operator and R the orthonormal basis with rows
    (n, t1, t2), n the unit normal and t1 orthogonal to the coordinate axis
    least aligned with n (local_basis_trafo_normal_tangential.hpp:34-79).  R is
    built here from those properties alone, as null spaces by SVD; the result
    does not depend on the signs of t1 and t2.
A dropped Dirichlet diagonal, a Cartesian diagonal at free-slip nodes or a
wrong choice of the reference axis each fail.

Reference bracket of GCA (gca.hpp:213-216): on fields constant along radial
columns both local prolongations give the same fine values, so z^T A_c x must
agree with the exact Galerkin product for such z and x; a wrong lateral
placement of the weights fails.
"""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random


def _basis_from_properties(n):
    """Rows (n, t1, t2): t1 spans the orthogonal complement of {n, e}, e the
    least-aligned axis (first of x, y, z on ties); t2 completes the basis."""
    a = np.abs(n)
    if a[0] <= a[1] and a[0] <= a[2]:
        e = np.array([1.0, 0.0, 0.0])
    elif a[1] <= a[0] and a[1] <= a[2]:
        e = np.array([0.0, 1.0, 0.0])
    else:
        e = np.array([0.0, 0.0, 1.0])
    t1 = np.linalg.svd(np.stack([n, e]))[2][-1]
    t2 = np.linalg.svd(np.stack([n, t1]))[2][-1]
    R = np.stack([n, t1, t2])
    assert np.abs(R @ R.T - np.eye(3)).max() < 1e-13
    assert abs(t1 @ e) < 1e-13
    return R


def _unique_node_normals(om):
    """unit lateral position per unique lateral node group"""
    g = om.groups.reshape(-1)
    c = om.coords.reshape(-1, 3)
    out = np.zeros((om.ngroups, 3))
    out[g] = c
    return out


@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (0, 1), (2, 2), (2, 1), (1, 2), (2, 0)])
def test_eps_diagonal_mode_is_assembled_diagonal(bc):
    om = O.Mesh(2, 5, 0.55, 1.0)
    k = np.exp(np.random.default_rng(11).uniform(-1.5, 1.5, om.N))
    x = consistent_random(om, 3, 12).reshape(-1)
    y = O.Eps(om, k, *bc).apply(x, diagonal=True)
    yu = O.to_unique(om, y, 3).reshape(om.ngroups, om.nr, 3)
    xu = O.to_unique(om, x, 3).reshape(om.ngroups, om.nr, 3)
    A_bc = O.eps_assemble(O.Eps(om, k, *bc))
    A_n = O.eps_assemble(O.Eps(om, k, 0, 0))
    expect = (A_bc.diagonal() * xu.reshape(-1)).reshape(om.ngroups, om.nr, 3)
    normals = _unique_node_normals(om)
    shells = [(0, bc[0]), (om.nr - 1, bc[1])]
    nfs = 0
    for r, mode in shells:
        if mode != 2:
            continue
        for g in range(om.ngroups):
            i = 3 * (g * om.nr + r)
            B = A_n[i:i + 3, i:i + 3].toarray()
            R = _basis_from_properties(normals[g])
            expect[g, r] = R.T @ (np.diag(R @ B @ R.T) * (R @ xu[g, r]))
            nfs += 1
    scale = np.abs(expect).max()
    assert np.abs(yu - expect).max() <= 1e-13 * scale
    if 2 in bc:
        assert nfs > 0
        # the Cartesian diagonal at free-slip nodes is not what the operator does
        cart = (A_n.diagonal() * xu.reshape(-1)).reshape(om.ngroups, om.nr, 3)
        r = 0 if bc[0] == 2 else om.nr - 1
        assert np.abs(cart[:, r] - expect[:, r]).max() > 1e-6 * scale


def test_diagonal_pin_rejects_wrong_reference_axis():
    """the free-slip diagonal depends on the tangent basis: a basis whose t1 is
    orthogonal to a different axis gives a different R^T diag(R B R^T) R"""
    n = np.array([0.2, 0.5, 0.84])
    n /= np.linalg.norm(n)
    B = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.2], [0.5, 0.2, 2.0]])
    R = _basis_from_properties(n)
    e_wrong = np.array([0.0, 0.0, 1.0])
    t1 = np.linalg.svd(np.stack([n, e_wrong]))[2][-1]
    t2 = np.linalg.svd(np.stack([n, t1]))[2][-1]
    Rw = np.stack([n, t1, t2])
    D = R.T @ np.diag(np.diag(R @ B @ R.T)) @ R
    Dw = Rw.T @ np.diag(np.diag(Rw @ B @ Rw.T)) @ Rw
    assert np.abs(D - Dw).max() > 1e-3


def _radially_constant(om, ncomp, seed):
    v = consistent_random(om, ncomp, seed).reshape(ncomp, -1, om.nr)
    v[:] = v[:, :, :1]
    return np.ascontiguousarray(v.reshape(-1))


def test_gca_reference_bracket_on_radially_constant_fields():
    mf, mc = O.Mesh(2, 9, 0.55, 1.0), O.Mesh(1, 5, 0.55, 1.0)
    k = np.exp(np.random.default_rng(5).uniform(-1, 1, mf.N))
    fine = O.Eps(mf, k)
    ops = {}
    for name, geometric in (("ref", False), ("geo", True)):
        op = O.Eps(mc, np.ones(mc.N))
        op.set_stored(O.gca_build(fine, mc, geometric=geometric))
        ops[name] = op
    z = _radially_constant(mc, 3, 21)
    x = _radially_constant(mc, 3, 22)
    a = mc.dot(z, ops["ref"].apply(x), 3)
    b = mc.dot(z, ops["geo"].apply(x), 3)
    assert abs(a - b) <= 1e-12 * abs(b)
    # and the Galerkin product itself: (P z)^T A_f (P x)
    Pz = O.prolongate_add(mc, mf, z, np.zeros(3 * mf.N))
    Px = O.prolongate_add(mc, mf, x, np.zeros(3 * mf.N))
    c = mf.dot(Pz, fine.apply(Px), 3)
    assert abs(a - c) <= 1e-12 * abs(c)
    # general fields see the bracket difference below the top layer
    zr = consistent_random(mc, 3, 23).reshape(-1)
    xr = consistent_random(mc, 3, 24).reshape(-1)
    ar = mc.dot(zr, ops["ref"].apply(xr), 3)
    br = mc.dot(zr, ops["geo"].apply(xr), 3)
    assert abs(ar - br) > 1e-6 * abs(br)


def test_penalty_multiplicity_and_symmetry():
    """free-slip on both shells: the reference's penalty (every node copy,
    before the interface sum) leaves <y, A x> != <x, A y>; counted once per
    node (owned copy) the operator is symmetric, and the two differ exactly by
    (copies - 1) times the penalty at interface nodes"""
    om = O.Mesh(3, 9, 0.55, 1.0)
    k = np.exp(np.random.default_rng(3).uniform(-1, 1, om.N))
    x = consistent_random(om, 3, 1).reshape(-1)
    y = consistent_random(om, 3, 2).reshape(-1)
    ref = O.Eps(om, k, 2, 2)
    once = O.Eps(om, k, 2, 2)
    once.set_penalty_once()
    a, b = om.dot(y, ref.apply(x), 3), om.dot(x, ref.apply(y), 3)
    assert abs(a - b) > 1e-3 * abs(a)
    a1, b1 = om.dot(y, once.apply(x), 3), om.dot(x, once.apply(y), 3)
    assert abs(a1 - b1) <= 1e-12 * abs(a1)
    modes = ref.penalty_modes()
    pen = sum(om.dot(mv, x, 3) * mv for mv in modes)
    cnt = np.zeros(10 * om.n * om.n)
    g = om.groups.reshape(-1)
    per_group = np.bincount(g, minlength=om.ngroups)
    cnt = np.repeat(per_group[g], om.nr)
    diff = ref.apply(x) - once.apply(x)
    expect = (np.tile(cnt, 3) - 1) * pen
    assert np.abs(diff - expect).max() <= 1e-12 * np.abs(pen).max()
    # Dirichlet / free-slip pairs are symmetric either way (no penalty)
    op = O.Eps(om, k, 2, 1)
    assert abs(om.dot(y, op.apply(x), 3) - om.dot(x, op.apply(y), 3)) <= 1e-12 * abs(om.dot(y, op.apply(x), 3))
