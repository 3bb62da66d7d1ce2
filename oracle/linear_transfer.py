"""CPU oracle (test infrastructure only; never the product path) for the
geometry-aware linear grid transfers: a numpy restatement of

- the weights, grid_transfer_linear.hpp:13-157 (radially aligned case
  :13-45, lateral-midpoint chord case :47-157),
- ProlongationLinear::operator() (prolongation_linear.hpp:117-191) and the
  vector variant with its early return inside the component loop
  (ProlongationVecLinear, :271-330; returns at :284 and :298),
- RestrictionLinear::operator() (restriction_linear.hpp:104-188): the
  transpose, pushed from owned fine nodes, and RestrictionVecLinear's early
  returns (:354, :367), followed by the additive communication.

Arrays: scalar [10][n][n][nr], vectors [3][10][n][n][nr] (component-major).
Pinned in tests/test_oracle_linear_transfer.py by the documented exactness for
affine fields and by adjointness, not by retyping these loops.
"""
import numpy as np


def _pos(coords, radii, d, x, y, r):
    return coords[d, x, y] * radii[r]


def stencil(coords_f, radii_f, d, xf, yf, rf, nrf):
    """[(coarse (x, y, r), weight), ...], case (0 coincident, 1 radial, 2 lateral midpoint)."""
    if xf % 2 == 0 and yf % 2 == 0 and rf % 2 == 0:
        return [((xf // 2, yf // 2, rf // 2), 1.0)], 0
    rb = rf // 2 if rf < nrf - 1 else rf // 2 - 1
    rt = rb + 1
    fn = np.linalg.norm(_pos(coords_f, radii_f, d, xf, yf, rf))
    if xf % 2 == 0 and yf % 2 == 0:
        X, Y = xf // 2, yf // 2
        bn = np.linalg.norm(_pos(coords_f, radii_f, d, 2 * X, 2 * Y, 2 * rb))
        tn = np.linalg.norm(_pos(coords_f, radii_f, d, 2 * X, 2 * Y, 2 * rt))
        nu = (fn - bn) / (tn - bn)
        return [((X, Y, rb), 1.0 - nu), ((X, Y, rt), nu)], 1
    if xf % 2 == 0:
        (x0, y0), (x1, y1) = (xf // 2, yf // 2), (xf // 2, yf // 2 + 1)
    elif yf % 2 == 0:
        (x0, y0), (x1, y1) = (xf // 2, yf // 2), (xf // 2 + 1, yf // 2)
    else:
        (x0, y0), (x1, y1) = (xf // 2 + 1, yf // 2), (xf // 2, yf // 2 + 1)

    def chord(r):
        return np.linalg.norm(0.5 * (_pos(coords_f, radii_f, d, 2 * x0, 2 * y0, 2 * r)
                                     + _pos(coords_f, radii_f, d, 2 * x1, 2 * y1, 2 * r)))

    bn, tn = chord(rb), chord(rt)
    nu = (fn - bn) / (tn - bn)
    wb, wt = 0.5 * (1.0 - nu), 0.5 * nu
    return [((x0, y0, rb), wb), ((x1, y1, rb), wb), ((x0, y0, rt), wt), ((x1, y1, rt), wt)], 2


def prolongate(coords_f, radii_f, xc, nc, nrc, ncomp, xf=None, vec_early_return=False):
    """xf (+)= P xc.  xc: [ncomp, 10, nc, nc, nrc]; returns [ncomp, 10, nf, nf, nrf]."""
    nf, nrf = 2 * nc - 1, 2 * nrc - 1
    xc = np.asarray(xc, dtype=np.float64).reshape(ncomp, 10, nc, nc, nrc)
    out = np.zeros((ncomp, 10, nf, nf, nrf)) if xf is None else np.array(xf, dtype=np.float64).reshape(
        ncomp, 10, nf, nf, nrf)
    for d in range(10):
        for x in range(nf):
            for y in range(nf):
                for r in range(nrf):
                    st, kase = stencil(coords_f, radii_f, d, x, y, r, nrf)
                    comps = 1 if (vec_early_return and ncomp == 3 and kase < 2) else ncomp
                    for c in range(comps):
                        out[c, d, x, y, r] += sum(w * xc[c, d, X, Y, R] for (X, Y, R), w in st)
    return out


def restrict(coords_f, radii_f, xf, owned_f, nc, nrc, ncomp, interface_sum, vec_early_return=False):
    """xc = P^T xf pushed from owned fine nodes, then interface_sum(xc) (additive communication)."""
    nf, nrf = 2 * nc - 1, 2 * nrc - 1
    xf = np.asarray(xf, dtype=np.float64).reshape(ncomp, 10, nf, nf, nrf)
    owned = np.asarray(owned_f).reshape(10, nf, nf, nrf)
    xc = np.zeros((ncomp, 10, nc, nc, nrc))
    for d in range(10):
        for x in range(nf):
            for y in range(nf):
                for r in range(nrf):
                    if not owned[d, x, y, r]:
                        continue
                    st, kase = stencil(coords_f, radii_f, d, x, y, r, nrf)
                    comps = 1 if (vec_early_return and ncomp == 3 and kase < 2) else ncomp
                    for c in range(comps):
                        for (X, Y, R), w in st:
                            xc[c, d, X, Y, R] += w * xf[c, d, x, y, r]
    return interface_sum(xc.reshape(-1), ncomp).reshape(ncomp, 10, nc, nc, nrc)
