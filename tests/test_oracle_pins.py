"""Pins of the CPU oracle by what the mathematics fixes (closed forms,
invariants, brute force), for each oracle function the reference's own
numbers (the assess L2 errors) do not already isolate.

  eps local matrix   symmetric; bilinear form on linear fields u = G x is
                     keval |det J(centroid)| (2 sym G : sym H - 2/3 trG trH)
                     with |det J| from the node coordinates (so rigid motions
                     are in the kernel)
  vector mass        1^T M 1 = volume of the wedge mesh (frustum closed form)
  lumped k-mass      sum = integral of a linear k (closed form)
  B, B^T             <B u, p> = <u, B^T p>; sum(B u) = -tr(G) volume for u = G x
  P, R               <R xf, xc> = <xf, P xc>; P 1 = 1
  GCA                geometric-bracket variant == P^T A_f P (assembled);
                     the reference's bracket differs only below the top layer
  penalty modes      orthonormal in the owned inner product
  owned dot          == masked numpy sum
"""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random


def _wedge_nodes(om, d, x, y, r, w):
    off = [[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
           [(1, 1, 0), (0, 1, 0), (1, 0, 0), (1, 1, 1), (0, 1, 1), (1, 0, 1)]][w]
    X = np.array([om.coords[d, x + a, y + b] * om.radii[r + c] for a, b, c in off])
    L = np.array([om.coords[d, x + a, y + b] for a, b, c in off[:3]])
    return X, L, (om.radii[r], om.radii[r + 1])


def _detJ_centroid(L, r1, r2):
    rm, gr = 0.5 * (r1 + r2), 0.5 * (r2 - r1)
    c = (L[0] + L[1] + L[2]) / 3.0
    J = np.stack([(L[1] - L[0]) * rm, (L[2] - L[0]) * rm, c * gr], axis=1)
    return abs(np.linalg.det(J))


def test_eps_local_matrix_closed_form():
    om = O.Mesh(2, 5, 0.55, 1.0)
    k = np.exp(np.random.default_rng(0).uniform(-1, 1, om.N))
    op = O.Eps(om, k)
    rng = np.random.default_rng(1)
    kk = k.reshape(10, om.n, om.n, om.nr)
    for (d, x, y, r, w) in [(0, 0, 0, 0, 0), (3, 1, 2, 1, 1), (9, 3, 3, 3, 0), (5, 2, 0, 2, 1)]:
        A = op.local_matrix(d, x, y, r, w)
        assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
        X, L, (r1, r2) = _wedge_nodes(om, d, x, y, r, w)
        off = [[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
               [(1, 1, 0), (0, 1, 0), (1, 0, 0), (1, 1, 1), (0, 1, 1), (1, 0, 1)]][w]
        keval = np.mean([kk[d, x + a, y + b, r + c] for a, b, c in off])
        vol = _detJ_centroid(L, r1, r2)
        for _ in range(4):
            G, H = rng.standard_normal((3, 3)), rng.standard_normal((3, 3))
            u = (X @ G.T).T.reshape(-1)  # dimension-wise order: component*6 + node
            v = (X @ H.T).T.reshape(-1)
            sG, sH = 0.5 * (G + G.T), 0.5 * (H + H.T)
            ref = keval * vol * (2 * np.sum(sG * sH) - 2.0 / 3.0 * np.trace(G) * np.trace(H))
            assert abs(v @ A @ u - ref) <= 1e-12 * abs(ref) + 1e-14
        # rigid motions (translations, rotations) are in the kernel
        for Wm in [np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0.0]]), np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0.0]])]:
            u = (X @ Wm.T).T.reshape(-1)
            assert np.abs(A @ u).max() <= 1e-13 * np.abs(A).max()


def _volume(om):
    c = om.coords
    P0, P1, P2, P3 = c[:, :-1, :-1], c[:, 1:, :-1], c[:, :-1, 1:], c[:, 1:, 1:]
    det = np.abs(np.einsum("dxyi,dxyi->dxy", P0, np.cross(P1, P2))) + \
        np.abs(np.einsum("dxyi,dxyi->dxy", P3, np.cross(P2, P1)))
    r = om.radii
    return det.sum() * np.sum(r[1:] ** 3 - r[:-1] ** 3) / 6.0


def test_vector_mass_total_is_volume():
    om = O.Mesh(3, 5, 0.55, 1.0)
    y = O.vector_mass(om, np.ones(3 * om.N))
    vol = _volume(om)
    for c in range(3):
        tot = om.dot(y[c * om.N:(c + 1) * om.N], np.ones(om.N), 1)
        assert abs(tot - vol) <= 1e-12 * vol


def test_kmass_lumped_integrates_linear_k():
    om = O.Mesh(3, 5, 0.55, 1.0)
    a = np.array([0.3, -0.7, 1.1])
    X = om.node_xyz()  # [10][n][n][nr][3]
    k = (X @ a).reshape(-1) + 2.0
    d = O.kmass_lumped(om, k)
    tot = om.dot(d, np.ones(om.N), 1)
    c = om.coords
    r = om.radii
    ref = 0.0
    for tri in ((c[:, :-1, :-1], c[:, 1:, :-1], c[:, :-1, 1:]), (c[:, 1:, 1:], c[:, :-1, 1:], c[:, 1:, :-1])):
        det = np.abs(np.einsum("dxyi,dxyi->dxy", tri[0], np.cross(tri[1], tri[2])))
        cen = (tri[0] + tri[1] + tri[2]) / 3.0
        ref += np.sum(det * (cen @ a)) * np.sum(r[1:] ** 4 - r[:-1] ** 4) / 8.0
        ref += 2.0 * np.sum(det) * np.sum(r[1:] ** 3 - r[:-1] ** 3) / 6.0
    assert abs(tot - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 2), (2, 1)])
def test_div_grad_adjoint_pair(bc):
    mf, mc = O.Mesh(3, 9, 0.55, 1.0), O.Mesh(2, 5, 0.55, 1.0)
    u = consistent_random(mf, 3, 0).reshape(-1)
    p = consistent_random(mc, 1, 1).reshape(-1)
    Bu = O.div_apply(mf, mc, u, *bc)
    Btp = O.grad_apply_add(mf, mc, p, np.zeros(3 * mf.N), *bc)
    lhs, rhs = mc.dot(Bu, p, 1), mf.dot(u, Btp, 3)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1e-300)


def test_div_of_linear_field_closed_form():
    mf, mc = O.Mesh(3, 9, 0.55, 1.0), O.Mesh(2, 5, 0.55, 1.0)
    G = np.random.default_rng(2).standard_normal((3, 3))
    X = mf.node_xyz()
    u = np.moveaxis(X @ G.T, -1, 0).reshape(-1)
    Bu = O.div_apply(mf, mc, u, 0, 0)
    tot = mc.dot(Bu, np.ones(mc.N), 1)
    ref = -np.trace(G) * _volume(mf)
    assert abs(tot - ref) <= 1e-11 * abs(ref)


def test_prolongation_restriction_adjoint_and_constants():
    mf, mc = O.Mesh(3, 9, 0.55, 1.0), O.Mesh(2, 5, 0.55, 1.0)
    xc = consistent_random(mc, 3, 3).reshape(-1)
    xf = consistent_random(mf, 3, 4).reshape(-1)
    Pxc = O.prolongate_add(mc, mf, xc, np.zeros(3 * mf.N))
    Rxf = O.restrict(mf, mc, xf)
    assert abs(mc.dot(Rxf, xc, 3) - mf.dot(xf, Pxc, 3)) <= 1e-12 * abs(mf.dot(xf, Pxc, 3))
    one = O.prolongate_add(mc, mf, np.ones(3 * mc.N), np.zeros(3 * mf.N))
    assert np.abs(one - 1.0).max() <= 1e-15


def _assembled_prolongation(mc, mf):
    """Scalar P over unique nodes, column by column."""
    import scipy.sparse as sp

    nuc = mc.ngroups * mc.nr
    cols = []
    for j in range(nuc):
        e = np.zeros(nuc)
        e[j] = 1.0
        xf = O.prolongate_add(mc, mf, O.from_unique(mc, e, 1), np.zeros(mf.N), ncomp=1)
        cols.append(O.to_unique(mf, xf, 1))
    return sp.csr_matrix(np.stack(cols, axis=1))


def test_gca_is_galerkin_product_and_reference_bracket_differs_only_below_top():
    import scipy.sparse as sp

    mf, mc = O.Mesh(2, 5, 0.55, 1.0), O.Mesh(1, 3, 0.55, 1.0)
    k = np.exp(np.random.default_rng(5).uniform(-1, 1, mf.N))
    fine = O.Eps(mf, k)
    Af = O.eps_assemble(fine)
    P = sp.kron(_assembled_prolongation(mc, mf), sp.identity(3), format="csr")
    ref = (P.T @ Af @ P).toarray()
    geo = O.gca_build(fine, mc, geometric=True)
    cop = O.Eps(mc, np.ones(mc.N))
    cop.set_stored(geo)
    Ac = O.eps_assemble(cop).toarray()
    assert np.abs(Ac - ref).max() <= 1e-12 * np.abs(ref).max()
    # the reference's bracket (gca.hpp:213-216): identical in the top radial layer of coarse cells
    refb = O.gca_build(fine, mc).reshape(10, mc.n - 1, mc.n - 1, mc.nr - 1, 2, 324)
    geob = geo.reshape(refb.shape)
    assert np.abs(refb[..., -1, :, :] - geob[..., -1, :, :]).max() <= 1e-15 * np.abs(geob).max()
    assert np.abs(refb[..., :-1, :, :] - geob[..., :-1, :, :]).max() > 1e-3 * np.abs(geob).max()


def test_penalty_modes_orthonormal():
    om = O.Mesh(2, 5, 0.55, 1.0)
    op = O.Eps(om, np.ones(om.N), 2, 2)
    modes = op.penalty_modes().reshape(3, 3 * om.N)
    G = np.array([[om.dot(modes[i], modes[j], 3) for j in range(3)] for i in range(3)])
    np.testing.assert_allclose(G, np.eye(3), atol=1e-12)


def test_owned_dot_is_masked_sum():
    om = O.Mesh(3, 5, 0.55, 1.0)
    a = consistent_random(om, 3, 7).reshape(-1)
    b = consistent_random(om, 3, 8).reshape(-1)
    own = np.tile(om.owned.reshape(-1).astype(bool), 3)
    assert abs(om.dot(a, b, 3) - np.sum((a * b)[own])) <= 1e-12 * np.sum(np.abs(a * b))
