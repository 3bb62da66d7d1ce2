"""CPU oracle (test infrastructure only) for selective GCA: a numpy
restatement of GCAElementsCollector::operator() (gca_elements_collector.hpp:
58-96) with the wedge shape-function gradients of fe/wedge/integrands.hpp:
118-258 and the one-point Felippa quadrature point (quadrature.hpp:36-39),
and of the selective operator: Galerkin matrices on flagged coarsest
ancestors, the level's own local matrices elsewhere
(local_matrix_storage.hpp:68-100).

The wedge node table is the reference's local ordering (kernel_helpers.hpp:
302-327): wedge 0 (x,y), (x+1,y), (x,y+1) on the bottom and top layer, wedge 1
(x+1,y+1), (x,y+1), (x+1,y).
"""
import numpy as np

WEDGE_NODES = [[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
               [(1, 1, 0), (0, 1, 0), (1, 0, 0), (1, 1, 1), (0, 1, 1), (1, 0, 1)]]


def _shape_lat(j, xi, eta):
    return (1.0 - xi - eta, xi, eta)[j % 3]


def _shape_rad(j, zeta):
    return 0.5 * (1.0 - zeta) if j < 3 else 0.5 * (1.0 + zeta)


def grad_shape(j, xi, eta, zeta):
    g_xi, g_eta, g_rad = (-1.0, 1.0, 0.0)[j % 3], (-1.0, 0.0, 1.0)[j % 3], (-0.5, 0.5)[j // 3]
    return np.array([g_xi * _shape_rad(j, zeta), g_eta * _shape_rad(j, zeta), _shape_lat(j, xi, eta) * g_rad])


def collect(k, n, nr, levels, n0, nr0):
    """flags[10][n0][n0][nr0] from the nodal coefficient k[10][n][n][nr]."""
    k = np.asarray(k, dtype=np.float64).reshape(10, n, n, nr)
    flags = np.zeros((10, n0, n0, nr0))
    qp = (1.0 / 3.0, 1.0 / 3.0, 0.0)
    grads = [grad_shape(j, *qp) for j in range(6)]
    for d in range(10):
        for x in range(n - 1):
            for y in range(n - 1):
                for r in range(nr - 1):
                    for w in range(2):
                        g = sum(grads[j] * k[d, x + ox, y + oy, r + orr]
                                for j, (ox, oy, orr) in enumerate(WEDGE_NODES[w]))
                        if np.linalg.norm(g) > 10:
                            flags[d, x >> levels, y >> levels, r >> levels] = 1.0
    return flags


def select(gca_mats, local_matrix, flags, n, nr, levels):
    """gca_mats [10][n-1][n-1][nr-1][2][18][18]; local_matrix(d, x, y, r, w) -> 18x18."""
    out = np.array(gca_mats, dtype=np.float64).reshape(10, n - 1, n - 1, nr - 1, 2, 18, 18)
    for d in range(10):
        for x in range(n - 1):
            for y in range(n - 1):
                for r in range(nr - 1):
                    if flags[d, x >> levels, y >> levels, r >> levels] == 0:
                        for w in range(2):
                            out[d, x, y, r, w] = local_matrix(d, x, y, r, w)
    return out
