"""Pins of the numpy solver restatements (tests/solver_restate.py, the checker
of the GPU solvers) by what the mathematics fixes, on small dense SPD systems:

  * power iteration -> the largest eigenvalue of D^-1 A
  * one Chebyshev sweep of order k with bounds [0.1, 1.5] lambda maps the
    error by the residual polynomial T_k((theta - t) / delta) / T_k(theta / delta)
    of D^-1 A (Saad, Iterative Methods, Alg. 12.1)
  * the V-cycle: the exact solution is a fixed point, and the cycle from a zero
    guess is a symmetric operator (equal pre- and post-smoothing polynomials,
    R = P^T)
  * PMINRES: its residual history equals the minimal M^-1-norm residual over
    the Krylov spaces, computed by least squares
A plausible mistake (a wrong rho update, a missing post-smoothing, a sign in
the Lanczos step) fails one of these.
"""
import numpy as np
import pytest

import solver_restate as SR


class _Mesh:
    def __init__(self, n):
        self.N = n // 3


class _Dense:
    """duck-typed level over a dense SPD matrix (Euclidean dots)"""

    def __init__(self, A):
        self.Amat = A
        self.m = _Mesh(A.shape[0])
        self.dinv = 1.0 / np.diag(A)

    def A(self, x):
        return self.Amat @ x

    def dot(self, a, b):
        return float(a @ b)


def _spd(n, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, cond, n)
    D = np.diag(rng.uniform(0.5, 2.0, n))
    return D @ (Q * ev) @ Q.T @ D


def test_power_iteration_converges_to_largest_eigenvalue():
    A = _spd(30, 0)
    lv = _Dense(A)
    lam = SR.power_iteration(lv, 7, 2000)
    Dh = np.sqrt(lv.dinv)
    ev = np.linalg.eigvalsh(Dh[:, None] * A * Dh[None, :])
    assert abs(lam - ev[-1]) <= 1e-8 * ev[-1]


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_one_chebyshev_sweep_is_the_residual_polynomial(order):
    A = _spd(24, 1)
    lv = _Dense(A)
    rng = np.random.default_rng(2)
    xs, x0 = rng.standard_normal(24), rng.standard_normal(24)
    b = A @ xs
    lam = 1.3
    x1 = SR.chebyshev(lv, lam, order, 1, x0, b)
    # error map P(D^-1 A) through the symmetric similar matrix D^1/2 (D^-1 A) D^-1/2
    Dh = np.sqrt(lv.dinv)
    w, V = np.linalg.eigh(Dh[:, None] * A * Dh[None, :])
    lmax, lmin = 1.5 * lam, 0.1 * lam
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    T = np.polynomial.chebyshev.Chebyshev.basis(order)
    p = T((theta - w) / delta) / T(theta / delta)
    e0 = (xs - x0) / Dh
    e1 = V @ (p * (V.T @ e0))
    assert np.abs((xs - x1) / Dh - e1).max() <= 1e-12 * np.abs(e0).max()


def test_vcycle_fixed_point_and_symmetry():
    import oracle_lib as O

    mc, mf = O.Mesh(1, 3, 0.55, 1.0), O.Mesh(2, 5, 0.55, 1.0)
    kf = np.ones(mf.N)
    opf = O.Eps(mf, kf, O.DIRICHLET, O.DIRICHLET)
    opc = O.Eps(mc, np.ones(mc.N), O.DIRICHLET, O.DIRICHLET)
    opc.set_stored(O.gca_build(opf, mc))
    levels = [SR.Level(mc, opc), SR.Level(mf, opf)]
    lam = SR.power_iteration(levels[1], 3, 50)
    V = SR.VCycle(levels, [None, lam], 2, 1)
    rng = np.random.default_rng(4)

    def consistent(v):
        v = v * np.tile(mf.owned, 3)
        return mf.interface_sum(v, 3)

    def dirichlet_zero(v):
        v = v.reshape(3, 10, mf.n, mf.n, mf.nr).copy()
        v[..., 0] = 0.0
        v[..., -1] = 0.0
        return v.reshape(-1)

    xs = dirichlet_zero(consistent(rng.standard_normal(3 * mf.N)))
    b = opf.apply(xs)
    x = V(xs.copy(), b)
    assert np.abs(x - xs).max() <= 1e-10 * np.abs(xs).max()
    b1 = dirichlet_zero(consistent(rng.standard_normal(3 * mf.N)))
    b2 = dirichlet_zero(consistent(rng.standard_normal(3 * mf.N)))
    z1 = V(np.zeros(3 * mf.N), b1)
    z2 = V(np.zeros(3 * mf.N), b2)
    lhs, rhs = mf.dot(z1, b2, 3), mf.dot(b1, z2, 3)
    assert abs(lhs - rhs) <= 1e-10 * np.sqrt(mf.dot(z1, z1, 3) * mf.dot(b2, b2, 3))


def test_pminres_history_is_the_minimal_residual():
    n = 20
    K = _spd(n, 5)
    rng = np.random.default_rng(6)
    Minv = np.linalg.inv(_spd(n, 7, cond=5.0))
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    its = 12
    _, hist = SR.pminres(lambda v: K @ v, lambda v: Minv @ v, lambda a, c: float(a @ c), x0, b, its)
    r0 = b - K @ x0
    z0 = Minv @ r0
    # minimal ||r||_{M^-1} over x0 + span{(M^-1 K)^j M^-1 r0, j < k}
    # (Krylov basis orthonormalised as it grows, so that it stays well conditioned)
    L = np.linalg.cholesky(Minv)
    Q = [z0 / np.linalg.norm(z0)]
    for k in range(1, its + 1):
        B = np.array(Q).T
        c, *_ = np.linalg.lstsq(L.T @ K @ B, L.T @ r0, rcond=None)
        res = L.T @ (r0 - K @ B @ c)
        assert abs(hist[k] - np.linalg.norm(res)) <= 1e-9 * hist[0]
        v = Minv @ (K @ Q[-1])
        for _ in range(2):
            v = v - B @ (B.T @ v)
        Q.append(v / np.linalg.norm(v))
    assert abs(hist[0] - np.sqrt(z0 @ r0)) <= 1e-14 * hist[0]
