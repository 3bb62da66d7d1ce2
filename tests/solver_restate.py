"""numpy restatements of the reference's solver recurrences over the CPU
oracle's operators (test infrastructure: the checker for the GPU solvers).

  power_iteration   linalg/solvers/power_iteration.hpp:13-47
  chebyshev         linalg/solvers/chebyshev.hpp:100-157 (rho carried across
                    the outer iterations of one call)
  vcycle            linalg/solvers/multigrid.hpp:184-211 (coarse correction
                    from zero, additive prolongation)
  pminres           linalg/solvers/pminres.hpp:75-185
Vectors are stored fields (interface copies included); dots and norms run
over owned nodes, as the reference's masked dot products do.
"""
import numpy as np

import oracle_lib as O


def splitmix_unit(seed, count):
    """uniform [0,1) values of SplitMix64 (seed, index), the start vector the
    GPU power iteration draws (its randomize)."""
    i = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (i + np.uint64(1))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Level:
    """One multigrid level: oracle operator, its mesh, inverse diagonal."""

    def __init__(self, om, op):
        self.m, self.op = om, op
        self.dinv = 1.0 / op.apply(np.ones(3 * om.N), diagonal=True)

    def A(self, x):
        return self.op.apply(x)

    def dot(self, a, b):
        return self.m.dot(a, b, 3)


def power_iteration(lv, seed, iterations):
    it = splitmix_unit(seed, 3 * lv.m.N)
    it = it / np.sqrt(lv.dot(it, it))
    aux = lv.dinv * lv.A(it)
    radius = 0.0
    for _ in range(iterations):
        it = aux / np.sqrt(lv.dot(aux, aux))
        aux = lv.dinv * lv.A(it)
        radius = lv.dot(it, aux)
    return radius


def chebyshev(lv, lam, order, iterations, x, b):
    lmax, lmin = 1.5 * lam, 0.1 * lam
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho = 1.0 / sigma
    x = x.copy()
    d = None
    for _ in range(iterations):
        z = lv.dinv * (b - lv.A(x))
        d = z / theta
        x = x + d
        for _i in range(2, order + 1):
            rho_new = 1.0 / (2.0 * sigma - rho)
            beta, alpha = rho_new * rho, 2.0 * rho_new / delta
            z = lv.dinv * (b - lv.A(x))
            d = alpha * z + beta * d
            x = x + d
            rho = rho_new
    return x


def assemble_by_apply(lv):
    """operator over unique dofs, column by column (penalty included)"""
    om = lv.m
    nu = 3 * om.ngroups * om.nr
    M = np.zeros((nu, nu))
    for j in range(nu):
        e = np.zeros(nu)
        e[j] = 1.0
        M[:, j] = O.to_unique(om, lv.A(O.from_unique(om, e, 3)), 3)
    return M


class VCycle:
    """levels coarse -> fine; lams[l] for l >= 1; dense solve on level 0"""

    def __init__(self, levels, lams, order, iterations):
        self.lv, self.lams, self.order, self.its = levels, lams, order, iterations
        self.Ac = assemble_by_apply(levels[0])

    def coarse(self, b):
        om = self.lv[0].m
        return O.from_unique(om, np.linalg.solve(self.Ac, O.to_unique(om, b, 3)), 3)

    def __call__(self, x, b, level=None):
        level = len(self.lv) - 1 if level is None else level
        if level == 0:
            return self.coarse(b)
        lv, lc = self.lv[level], self.lv[level - 1]
        x = chebyshev(lv, self.lams[level], self.order, self.its, x, b)
        r = b - lv.A(x)
        rc = O.restrict(lv.m, lc.m, r)
        ec = self(np.zeros(3 * lc.m.N), rc, level - 1)
        x = O.prolongate_add(lc.m, lv.m, ec, x.copy())
        return chebyshev(lv, self.lams[level], self.order, self.its, x, b)


def pminres(K, prec, dot, x, b, iterations):
    """residual history (|eta|) of the preconditioned MINRES recurrence"""
    x = x.copy()
    v_jm1 = np.zeros_like(b)
    w_j = np.zeros_like(b)
    w_jm1 = np.zeros_like(b)
    v_j = b - K(x)
    z = prec(v_j)
    g_jm1, g_j = 1.0, np.sqrt(dot(z, v_j))
    eta, s_jm1, s_j, c_jm1, c_j = g_j, 0.0, 0.0, 1.0, 1.0
    hist = [g_j]
    for _ in range(iterations):
        z = z / g_j
        az = K(z)
        delta = dot(az, z)
        v_jm1 = az - delta / g_j * v_j - g_j / g_jm1 * v_jm1
        v_jm1, v_j = v_j, v_jm1
        z_jp1 = prec(v_j)
        g_jp1 = np.sqrt(dot(z_jp1, v_j))
        a0 = c_j * delta - c_jm1 * s_j * g_j
        a1 = np.sqrt(a0 * a0 + g_jp1 * g_jp1)
        a2 = s_j * delta + c_jm1 * c_j * g_j
        a3 = s_jm1 * g_j
        c_jp1, s_jp1 = a0 / a1, g_jp1 / a1
        w_jm1 = z / a1 - a3 / a1 * w_jm1 - a2 / a1 * w_j
        w_jm1, w_j = w_j, w_jm1
        x = x + c_jp1 * eta * w_j
        eta = -s_jp1 * eta
        hist.append(abs(eta))
        z = z_jp1
        g_jm1, g_j = g_j, g_jp1
        c_jm1, c_j = c_j, c_jp1
        s_jm1, s_j = s_j, s_jp1
    return x, np.array(hist)
