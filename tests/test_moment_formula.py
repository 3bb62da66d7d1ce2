"""The closed-form per-wedge moment formula used by the B / B^T kernels
(csrc/kernels_stokes.hip), restated in numpy, against the oracle's 6-point
quadrature of divergence.hpp:117-323 / gradient.hpp:115-325.  The Felippa
rule integrates the divergence-against-coarse-shape integrand exactly, so
the two agree to rounding for every boundary-condition pair."""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random

G2 = 1 / np.sqrt(3)
COEF = [[[1.0, -0.5, -0.5], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5]],
        [[0.5, -0.5, -0.5], [0.5, 0.5, 0.0], [0.0, 0.0, 0.5]],
        [[0.5, -0.5, -0.5], [0.0, 0.5, 0.0], [0.5, 0.0, 0.5]],
        [[0.0, 0.5, 0.5], [0.5, -0.5, 0.0], [0.5, 0.0, -0.5]]]
LATM = np.array([[[al / 2 + (be + ga) / 6, al / 6 + be / 12 + ga / 24, al / 6 + be / 24 + ga / 12]
                  for (al, be, ga) in f] for f in COEF])
IND = [[[0, 1], [2, 3]], [[3, 2], [1, 0]]]
WK = [[0, 1, 2], [3, 2, 1]]


def radial(r1, r2, frw):
    g = 0.5 * (r2 - r1)
    A = np.zeros((2, 2)); rr = np.zeros(2)
    for z in (-G2, G2):
        r = r1 + g * (1 + z)
        R = [0.5 * (1 - z), 0.5 * (1 + z)]
        rad = [0.25 * (1 - z) if frw else 0.25 * (3 - z), 0.25 * (3 + z) if frw else 0.25 * (1 + z)]
        for ir in range(2):
            A[ir] += [rad[ir] * r * g * R[0], rad[ir] * r * g * R[1]]
            rr[ir] += rad[ir] * r * r
    C = np.stack([-0.5 * rr, 0.5 * rr], axis=1)
    return A, C


def latw(P0, P1, P2):
    a, b = P1 - P0, P2 - P0
    v1, v3, nv = np.cross(b, P0), np.cross(P0, a), np.cross(a, b)
    sg = -1.0 if a @ v1 > 0 else 1.0
    return v1, v3, nv, sg


def trial(u, mode, P):
    if mode == 1:
        return np.zeros(3)
    if mode == 2:
        nn = P / np.linalg.norm(P)
        return u - (nn @ u) * nn
    return u


def div_moments(mf, mc, u, bcb, bct):
    nf, nrf, nc, nrc = mf.n, mf.nr, mc.n, mc.nr
    U = u.reshape(3, 10, nf, nf, nrf)
    p = np.zeros((10, nc, nc, nrc))
    co = mf.coords
    for d in range(10):
        for x in range(nf - 1):
            for y in range(nf - 1):
                Pk = [co[d, x, y], co[d, x + 1, y], co[d, x, y + 1], co[d, x + 1, y + 1]]
                lk = [(x, y), (x + 1, y), (x, y + 1), (x + 1, y + 1)]
                for c in range(nrf - 1):
                    A, C = radial(mf.radii[c], mf.radii[c + 1], c & 1)
                    for w in range(2):
                        k0, k1, k2 = WK[w]
                        v1, v3, nv, sg = latw(Pk[k0], Pk[k1], Pk[k2])
                        M = LATM[IND[w][y & 1][x & 1]]
                        def uu(k, lev):
                            v = U[:, d, lk[k][0], lk[k][1], lev].copy()
                            if lev == 0: v = trial(v, bcb, Pk[k])
                            if lev == nrf - 1: v = trial(v, bct, Pk[k])
                            return v
                        al, sm, bx, be = [], [], [], []
                        for t in range(2):
                            u0, u1, u2 = uu(k0, c + t), uu(k1, c + t), uu(k2, c + t)
                            du1, du2 = u1 - u0, u2 - u0
                            al.append(v1 @ du1 + v3 @ du2); sm.append(nv @ u0); bx.append(nv @ du1); be.append(nv @ du2)
                        for ir in range(2):
                            P0 = sg * sum(al[t] * A[ir, t] + sm[t] * C[ir, t] for t in range(2))
                            Px = sg * sum(bx[t] * (C[ir, t] - A[ir, t]) for t in range(2))
                            Pe = sg * sum(be[t] * (C[ir, t] - A[ir, t]) for t in range(2))
                            for il in range(3):
                                ck = WK[w][il]
                                X = x // 2 + (ck & 1); Y = y // 2 + (ck >> 1)
                                p[d, X, Y, c // 2 + ir] += P0 * M[il, 0] + Px * M[il, 1] + Pe * M[il, 2]
    return p.reshape(-1)


def grad_moments(mf, mc, p, bcb, bct):
    nf, nrf, nc, nrc = mf.n, mf.nr, mc.n, mc.nr
    Pp = p.reshape(10, nc, nc, nrc)
    u = np.zeros((3, 10, nf, nf, nrf))
    co = mf.coords
    for d in range(10):
        for x in range(nf - 1):
            for y in range(nf - 1):
                Pk = [co[d, x, y], co[d, x + 1, y], co[d, x, y + 1], co[d, x + 1, y + 1]]
                lk = [(x, y), (x + 1, y), (x, y + 1), (x + 1, y + 1)]
                ck = [(x // 2, y // 2), (x // 2 + 1, y // 2), (x // 2, y // 2 + 1), (x // 2 + 1, y // 2 + 1)]
                for c in range(nrf - 1):
                    A, C = radial(mf.radii[c], mf.radii[c + 1], c & 1)
                    R = c // 2
                    for w in range(2):
                        k = WK[w]
                        v1, v3, nv, sg = latw(Pk[k[0]], Pk[k[1]], Pk[k[2]])
                        M = LATM[IND[w][y & 1][x & 1]]
                        pw = np.array([[Pp[d, ck[k[il]][0], ck[k[il]][1], R + ir] for ir in range(2)] for il in range(3)])
                        q0 = sg * (pw[:, :].T @ M[:, 0]); qx = sg * (pw.T @ M[:, 1]); qe = sg * (pw.T @ M[:, 2])
                        for t in range(2):
                            Xv = A[:, t] @ q0; Yv = C[:, t] @ q0
                            ca = C[:, t] - A[:, t]
                            Zx = ca @ qx; Ze = ca @ qe
                            outs = [(Yv - Zx - Ze) * nv - Xv * (v1 + v3), Xv * v1 + Zx * nv, Xv * v3 + Ze * nv]
                            for m in range(3):
                                u[:, d, lk[k[m]][0], lk[k[m]][1], c + t] += outs[m]
    # test side
    for lev, mode in ((0, bcb), (nrf - 1, bct)):
        if mode == 1:
            u[:, :, :, :, lev] = 0
        elif mode == 2:
            nn = co / np.linalg.norm(co, axis=-1, keepdims=True)
            v = u[:, :, :, :, lev]
            vn = np.einsum("cdxy,dxyc->dxy", v, nn)
            u[:, :, :, :, lev] = v - np.moveaxis(nn, -1, 0) * vn
    return u.reshape(-1)




@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (2, 2), (1, 2), (2, 1)])
def test_moment_formula_matches_quadrature(bc):
    mf, mc = O.Mesh(2, 5, 0.55, 1.0), O.Mesh(1, 3, 0.55, 1.0)
    u = consistent_random(mf, 3, 0).reshape(-1)
    p = consistent_random(mc, 1, 1).reshape(-1)
    ref = O.div_apply(mf, mc, u, *bc, comm=False)
    got = div_moments(mf, mc, u, *bc)
    assert np.abs(ref - got).max() <= 1e-14 * np.abs(ref).max()
    refg = O.grad_apply_add(mf, mc, p, np.zeros(3 * mf.N), *bc, comm=False)
    gotg = grad_moments(mf, mc, p, *bc)
    assert np.abs(refg - gotg).max() <= 1e-14 * np.abs(refg).max()
