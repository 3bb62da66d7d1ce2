/*
 * fv_oracle.c -- CPU restatement of the reference's hex finite-volume energy
 * step (FCT advection-diffusion) and its helpers.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * FV fields have one ghost layer on every side:
 *   scalar [10][n+1][n+1][nr+1], r fastest; real cells (xc, yc, rc) in
 *   [1, n-1]^2 x [1, nr-1]; FV cell (xc, yc, rc) is the FE hex cell
 *   (xc-1, yc-1, rc-1).  Vector FV fields (cell centres) carry 3 components
 *   innermost: [10][n+1][n+1][nr+1][3].
 *   Antidiffusive fluxes: [10][n+1][n+1][nr+1][6] (face innermost).
 * Reference: fv/hex/operators/fct_advection_diffusion.hpp,
 *            fv/hex/operators/geometry_helper.hpp, fv/hex/helpers.hpp,
 *            fv/hex/conversion.hpp:131-211,
 *            communication/shell/fv_communication.hpp.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------------------------------------------------------- indexing */
static inline int64_t fvi(const orc_mesh* m, int d, int x, int y, int r)
{
    const int64_t s = m->n + 1, sr = m->nr + 1;
    return (((int64_t)d * s + x) * s + y) * sr + r;
}

int64_t orc_fv_size(const orc_mesh* m) { return 10LL * (m->n + 1) * (m->n + 1) * (m->nr + 1); }

/* ---------------------------------------------------------------- wedge helpers
 * (fe/wedge/integrands.hpp:17-101, kernel_helpers.hpp:26-55,302-327) */
static const int FWOFF[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                   {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};

static double fshape(int k, double xi, double eta, double zeta)
{
    const double lat[3] = {1.0 - xi - eta, xi, eta};
    const double rad[2] = {0.5 * (1 - zeta), 0.5 * (1 + zeta)};
    return lat[k % 3] * rad[k / 3];
}
static void fgrad_shape(int k, double xi, double eta, double zeta, double g[3])
{
    static const double dxi[3] = {-1.0, 1.0, 0.0}, deta[3] = {-1.0, 0.0, 1.0}, drad[2] = {-0.5, 0.5};
    const double lat[3] = {1.0 - xi - eta, xi, eta};
    const double rad[2] = {0.5 * (1 - zeta), 0.5 * (1 + zeta)};
    g[0] = dxi[k % 3] * rad[k / 3];
    g[1] = deta[k % 3] * rad[k / 3];
    g[2] = lat[k % 3] * drad[k / 3];
}
/* lateral corner positions of the two wedges of FE hex cell (x, y) */
static void fwedge_surf(const orc_mesh* m, int d, int x, int y, double P[2][3][3])
{
    const int n = m->n;
    for (int w = 0; w < 2; ++w)
        for (int i = 0; i < 3; ++i) {
            const int xx = x + FWOFF[w][i][0], yy = y + FWOFF[w][i][1];
            const double* c = m->coords + 3 * (((int64_t)d * n + xx) * n + yy);
            for (int a = 0; a < 3; ++a) P[w][i][a] = c[a];
        }
}
static void fcross(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double fdot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Felippa 3x2 points and weight (fe/wedge/quadrature/quadrature.hpp:67-110) */
static const double FQ6P[6][3] = {{0.6666666666666666, 0.1666666666666667, -0.5773502691896257},
                                  {0.1666666666666667, 0.6666666666666666, -0.5773502691896257},
                                  {0.1666666666666667, 0.1666666666666667, -0.5773502691896257},
                                  {0.6666666666666666, 0.1666666666666667, 0.5773502691896257},
                                  {0.1666666666666667, 0.6666666666666666, 0.5773502691896257},
                                  {0.1666666666666667, 0.1666666666666667, 0.5773502691896257}};
static const double FQ6W = 0.1666666666666667;
/* 3-point triangle rule (fe/triangle/quadrature/quadrature.hpp:8-26) */
static const double TRI_P[3][2] = {{0.6666666666666666, 0.1666666666666667},
                                   {0.1666666666666667, 0.6666666666666666},
                                   {0.1666666666666667, 0.1666666666666667}};
static const double TRI_W = 0.1666666666666667;

/* |det J| of the wedge map at reference point q (integrands.hpp:623-670) */
static double fabs_det(const double P[3][3], double r1, double r2, const double q[3])
{
    const double r = r1 + 0.5 * (r2 - r1) * (1 + q[2]);
    const double gr = 0.5 * (r2 - r1);
    double J[3][3];
    for (int a = 0; a < 3; a++) {
        J[a][0] = (P[1][a] - P[0][a]) * r;
        J[a][1] = (P[2][a] - P[0][a]) * r;
        J[a][2] = ((1 - q[0] - q[1]) * P[0][a] + q[0] * P[1][a] + q[1] * P[2][a]) * gr;
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    return fabs(det);
}

/* ---------------------------------------------------------------- cell centres
 * helpers.hpp:123-190: centre = (int_K x) / |K|, Felippa 3x2 on both wedges;
 * real cells only, then the lateral ghost layers by face exchange; the radial
 * ghost cells at the shells keep zero */
void orc_fv_cell_centers(const orc_mesh* m, double* c)
{
    const int n = m->n, nr = m->nr;
    memset(c, 0, sizeof(double) * 3 * orc_fv_size(m));
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc) {
                double P[2][3][3];
                fwedge_surf(m, d, xc - 1, yc - 1, P);
                for (int rc = 1; rc <= nr - 1; ++rc) {
                    const double r1 = m->radii[rc - 1], r2 = m->radii[rc];
                    double vol = 0, I[3] = {0, 0, 0};
                    for (int w = 0; w < 2; ++w)
                        for (int q = 0; q < 6; ++q) {
                            const double Jw = fabs_det(P[w], r1, r2, FQ6P[q]) * FQ6W;
                            vol += Jw;
                            const double xi = FQ6P[q][0], eta = FQ6P[q][1], zeta = FQ6P[q][2];
                            const double rr = r1 + 0.5 * (r2 - r1) * (1 + zeta);
                            for (int a = 0; a < 3; ++a) {
                                const double lat = (1 - xi - eta) * P[w][0][a] + xi * P[w][1][a] + eta * P[w][2][a];
                                I[a] += rr * lat * Jw;
                            }
                        }
                    for (int a = 0; a < 3; ++a) c[3 * fvi(m, d, xc, yc, rc) + a] = I[a] / vol;
                }
            }
    orc_fv_ghost_update(m, c, 3);
}

/* ---------------------------------------------------------------- ghost layers
 * fv_communication.hpp: every lateral ghost cell of a diamond receives the
 * real cell of the neighbouring diamond across that face.  Restated here by
 * position: the face between ghost and real cell is the FE edge of two
 * boundary nodes; the source is the one real cell of another diamond that
 * has both nodes as corners (found by brute-force search over the other
 * diamonds' boundary cells). */
typedef struct {
    int d, x, y;    /* ghost (FV lateral index) */
    int sd, sx, sy; /* source real cell */
} ghost_pair;

static int same_pos(const double* a, const double* b)
{
    return fabs(a[0] - b[0]) < 1e-12 && fabs(a[1] - b[1]) < 1e-12 && fabs(a[2] - b[2]) < 1e-12;
}

static int ghost_pairs(const orc_mesh* m, ghost_pair** out)
{
    const int n = m->n;
    ghost_pair* gp = (ghost_pair*)malloc(sizeof(ghost_pair) * 40 * (n - 1));
    int cnt = 0;
    for (int d = 0; d < 10; ++d)
        for (int side = 0; side < 4; ++side)
            for (int j = 1; j <= n - 1; ++j) {
                /* the face's two FE nodes (a, b) and the ghost cell */
                int ax, ay, bx, by, gx, gy;
                if (side == 0) { ax = 0; ay = j - 1; bx = 0; by = j; gx = 0; gy = j; }
                else if (side == 1) { ax = n - 1; ay = j - 1; bx = n - 1; by = j; gx = n; gy = j; }
                else if (side == 2) { ax = j - 1; ay = 0; bx = j; by = 0; gx = j; gy = 0; }
                else { ax = j - 1; ay = n - 1; bx = j; by = n - 1; gx = j; gy = n; }
                const double* pa = m->coords + 3 * (((int64_t)d * n + ax) * n + ay);
                const double* pb = m->coords + 3 * (((int64_t)d * n + bx) * n + by);
                int found = 0;
                for (int sd = 0; sd < 10 && !found; ++sd) {
                    if (sd == d) continue;
                    for (int xc = 0; xc < n - 1 && !found; ++xc)
                        for (int yc = 0; yc < n - 1 && !found; ++yc) {
                            if (!(xc == 0 || yc == 0 || xc == n - 2 || yc == n - 2)) continue;
                            int ha = 0, hb = 0;
                            for (int k = 0; k < 4; ++k) {
                                const int xx = xc + (k & 1), yy = yc + (k >> 1);
                                const double* p = m->coords + 3 * (((int64_t)sd * n + xx) * n + yy);
                                ha |= same_pos(p, pa);
                                hb |= same_pos(p, pb);
                            }
                            if (ha && hb) {
                                gp[cnt++] = (ghost_pair){d, gx, gy, sd, xc + 1, yc + 1};
                                found = 1;
                            }
                        }
                }
            }
    *out = gp;
    return cnt;
}

/* copy the neighbours' real cells into the lateral ghost cells (real radial
 * range, fv_communication.hpp:60-176); ncomp components innermost */
void orc_fv_ghost_update(const orc_mesh* m, double* f, int ncomp)
{
    ghost_pair* gp;
    const int cnt = ghost_pairs(m, &gp);
    for (int k = 0; k < cnt; ++k)
        for (int rc = 1; rc <= m->nr - 1; ++rc)
            for (int c = 0; c < ncomp; ++c)
                f[ncomp * fvi(m, gp[k].d, gp[k].x, gp[k].y, rc) + c] =
                    f[ncomp * fvi(m, gp[k].sd, gp[k].sx, gp[k].sy, rc) + c];
    free(gp);
}

int orc_fv_ghost_pairs(const orc_mesh* m, int32_t* pairs /* [cnt][6] or NULL */)
{
    ghost_pair* gp;
    const int cnt = ghost_pairs(m, &gp);
    if (pairs)
        for (int k = 0; k < cnt; ++k) {
            pairs[6 * k + 0] = gp[k].d; pairs[6 * k + 1] = gp[k].x; pairs[6 * k + 2] = gp[k].y;
            pairs[6 * k + 3] = gp[k].sd; pairs[6 * k + 4] = gp[k].sx; pairs[6 * k + 5] = gp[k].sy;
        }
    free(gp);
    return cnt;
}

/* ---------------------------------------------------------------- cell geometry
 * geometry_helper.hpp:53-176 (face maps), :178-360 (compute_geometry):
 * beta[f] = int_face u . n dS and S[f] = int_face n dS per face (0: x-1, 1: x+1,
 * 2: y-1, 3: y+1, 4: r-1, 5: r+1), oriented towards the neighbour's centre;
 * M = cell volume */
void orc_fv_geometry(const orc_mesh* m, const double* centers, const double* vel, int d, int xc, int yc, int rc,
                     double beta[6], double* M, double S[6][3])
{
    static const int OX[6] = {-1, 1, 0, 0, 0, 0}, OY[6] = {0, 0, -1, 1, 0, 0}, OR[6] = {0, 0, 0, 0, -1, 1};
    const int n = m->n, nr = m->nr;
    const int64_t N = 10LL * n * n * nr;
    double P[2][3][3];
    fwedge_surf(m, d, xc - 1, yc - 1, P);
    const double r1 = m->radii[rc - 1], r2 = m->radii[rc];
    double U[2][6][3];
    for (int w = 0; w < 2; ++w)
        for (int i = 0; i < 6; ++i) {
            const int xx = xc - 1 + FWOFF[w][i][0], yy = yc - 1 + FWOFF[w][i][1], rr = rc - 1 + FWOFF[w][i][2];
            const int64_t node = (((int64_t)d * n + xx) * n + yy) * nr + rr;
            for (int a = 0; a < 3; ++a) U[w][i][a] = vel[a * N + node];
        }
    const double gl = 1.0 / sqrt(3.0);
    const double line[2] = {-gl, gl};
    const double* cc = centers + 3 * fvi(m, d, xc, yc, rc);
    for (int f = 0; f < 6; ++f) {
        beta[f] = 0;
        S[f][0] = S[f][1] = S[f][2] = 0;
        for (int w = 0; w < 2; ++w) {
            /* contribution (geometry_helper.hpp:53-89): wedge 0 owns faces 0, 2, wedge 1 faces 1, 3
             * (quads); both wedges' triangles make faces 4, 5 */
            const int quad = (w == 0 && (f == 0 || f == 2)) || (w == 1 && (f == 1 || f == 3));
            const int tri = f >= 4;
            if (!quad && !tri) continue;
            const int np = tri ? 3 : 4;
            for (int q = 0; q < np; ++q) {
                double xi, eta, zeta, wq;
                int cpt; /* 0: xi x eta, 1: xi x zeta, 2: eta x zeta */
                if (tri) {
                    xi = TRI_P[q][0];
                    eta = TRI_P[q][1];
                    zeta = f == 4 ? -1.0 : 1.0;
                    wq = TRI_W;
                    cpt = 0;
                } else {
                    const double s = 0.5 * (line[q / 2] + 1.0), t = line[q % 2];
                    wq = 0.5;
                    zeta = t;
                    if (f == 0 || f == 1) { xi = 0; eta = s; cpt = 2; }
                    else { xi = s; eta = 0; cpt = 1; }
                }
                double dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0}, dzeta[3] = {0, 0, 0}, u[3] = {0, 0, 0};
                for (int i = 0; i < 6; ++i) {
                    const double sh = fshape(i, xi, eta, zeta);
                    double g[3];
                    fgrad_shape(i, xi, eta, zeta, g);
                    const double rad = i < 3 ? r1 : r2;
                    for (int a = 0; a < 3; ++a) {
                        const double xa = P[w][i % 3][a] * rad;
                        dxi[a] += g[0] * xa;
                        deta[a] += g[1] * xa;
                        dzeta[a] += g[2] * xa;
                        u[a] += sh * U[w][i][a];
                    }
                }
                double nv[3];
                if (cpt == 0) fcross(dxi, deta, nv);
                else if (cpt == 1) fcross(dxi, dzeta, nv);
                else fcross(deta, dzeta, nv);
                beta[f] += wq * fdot(u, nv);
                for (int a = 0; a < 3; ++a) S[f][a] += wq * nv[a];
            }
        }
        const double* nc = centers + 3 * fvi(m, d, xc + OX[f], yc + OY[f], rc + OR[f]);
        const double dx[3] = {nc[0] - cc[0], nc[1] - cc[1], nc[2] - cc[2]};
        if (fdot(S[f], dx) < 0) {
            for (int a = 0; a < 3; ++a) S[f][a] = -S[f][a];
            beta[f] = -beta[f];
        }
    }
    double vol = 0;
    for (int w = 0; w < 2; ++w)
        for (int q = 0; q < 6; ++q) vol += fabs_det(P[w], r1, r2, FQ6P[q]) * FQ6W;
    *M = vol;
}

static const int NOX[6] = {-1, 1, 0, 0, 0, 0}, NOY[6] = {0, 0, -1, 1, 0, 0}, NOR[6] = {0, 0, 0, 0, -1, 1};

/* ---------------------------------------------------------------- stable time step
 * fct_advection_diffusion.hpp:110-231 */
double orc_fv_dt_stable(const orc_mesh* m, const double* centers, const double* vel, double kappa)
{
    const int n = m->n, nr = m->nr;
    double best = INFINITY;
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc)
                for (int rc = 1; rc <= nr - 1; ++rc) {
                    double beta[6], M, S[6][3];
                    orc_fv_geometry(m, centers, vel, d, xc, yc, rc, beta, &M, S);
                    const double* ci = centers + 3 * fvi(m, d, xc, yc, rc);
                    double lambda = 0;
                    for (int f = 0; f < 6; ++f) {
                        if (beta[f] < 0) lambda -= beta[f];
                        if (kappa > 0) {
                            const double* cj = centers + 3 * fvi(m, d, xc + NOX[f], yc + NOY[f], rc + NOR[f]);
                            const double dx[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
                            const double den = fdot(dx, S[f]);
                            if (den > 0) lambda += kappa * fdot(S[f], S[f]) / den;
                        }
                    }
                    const double dtc = lambda > 0 ? M / lambda : 1e30;
                    if (dtc < best) best = dtc;
                }
    return best;
}

/* ---------------------------------------------------------------- FCT stages
 * predictor fct_advection_diffusion.hpp:261-402: T's ghosts are updated
 * first; T_L and the pre-scaled antidiffusive fluxes of every real cell */
void orc_fct_predictor(const orc_mesh* m, const double* centers, const double* vel, double* T, double dt,
                       double kappa, const double* source, int subtract_divergence, double* TL, double* antidiff)
{
    const int n = m->n, nr = m->nr;
    orc_fv_ghost_update(m, T, 1);
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc)
                for (int rc = 1; rc <= nr - 1; ++rc) {
                    double beta[6], M, S[6][3];
                    orc_fv_geometry(m, centers, vel, d, xc, yc, rc, beta, &M, S);
                    const int64_t i = fvi(m, d, xc, yc, rc);
                    const double Ti = T[i];
                    const double* ci = centers + 3 * i;
                    double lo_diag = 0, lo_update = 0, bsum = 0;
                    for (int f = 0; f < 6; ++f) {
                        const int64_t j = fvi(m, d, xc + NOX[f], yc + NOY[f], rc + NOR[f]);
                        const double Tj = T[j];
                        if (beta[f] >= 0) lo_diag += beta[f];
                        else lo_update += beta[f] * Tj;
                        bsum += beta[f];
                        const double* cj = centers + 3 * j;
                        const double dx[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
                        const double den = fdot(dx, S[f]);
                        const double dc = kappa * (fdot(S[f], S[f]) / den);
                        lo_diag += dc;
                        lo_update -= dc * Tj;
                        antidiff[6 * i + f] = (dt / M) * (fabs(beta[f]) / 2.0) * (Ti - Tj);
                    }
                    if (subtract_divergence) lo_diag -= bsum;
                    double v = Ti - (dt / M) * (lo_diag * Ti + lo_update);
                    if (source) v += dt * source[i];
                    TL[i] = v;
                }
}

/* Dirichlet values at the shells: real boundary cell and the ghost outside
 * (helpers.hpp:64-96) */
void orc_fv_dirichlet(const orc_mesh* m, double* f, int apply_cmb, double t_cmb, int apply_surf, double t_surf)
{
    const int n = m->n, nr = m->nr;
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc) {
                if (apply_cmb) {
                    f[fvi(m, d, xc, yc, 1)] = t_cmb;
                    f[fvi(m, d, xc, yc, 0)] = t_cmb;
                }
                if (apply_surf) {
                    f[fvi(m, d, xc, yc, nr - 1)] = t_surf;
                    f[fvi(m, d, xc, yc, nr)] = t_surf;
                }
            }
}

/* limiter fct_advection_diffusion.hpp:485-599: T_L's ghosts updated, Zalesak
 * R+ / R-, then R's ghosts updated */
void orc_fct_limiter(const orc_mesh* m, double* TL, const double* antidiff, double* Rp, double* Rm)
{
    const int n = m->n, nr = m->nr;
    orc_fv_ghost_update(m, TL, 1);
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc)
                for (int rc = 1; rc <= nr - 1; ++rc) {
                    const int64_t i = fvi(m, d, xc, yc, rc);
                    const double ti = TL[i];
                    double pp = 0, pm = 0, tmax = ti, tmin = ti;
                    for (int f = 0; f < 6; ++f) {
                        const double fij = antidiff[6 * i + f];
                        if (fij > 0) pp += fij;
                        else pm += fij;
                        const double tj = TL[fvi(m, d, xc + NOX[f], yc + NOY[f], rc + NOR[f])];
                        tmax = fmax(tmax, tj);
                        tmin = fmin(tmin, tj);
                    }
                    const double qp = tmax - ti, qm = tmin - ti;
                    Rp[i] = pp > 0 ? fmin(1.0, qp / pp) : 1.0;
                    Rm[i] = pm < 0 ? fmin(1.0, qm / pm) : 1.0;
                }
    orc_fv_ghost_update(m, Rp, 1);
    orc_fv_ghost_update(m, Rm, 1);
}

/* correction fct_advection_diffusion.hpp:604-715: T = T_L + sum_j alpha_ij f_ij */
void orc_fct_correction(const orc_mesh* m, const double* TL, const double* antidiff, const double* Rp,
                        const double* Rm, double* T)
{
    const int n = m->n, nr = m->nr;
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc)
                for (int rc = 1; rc <= nr - 1; ++rc) {
                    const int64_t i = fvi(m, d, xc, yc, rc);
                    double corr = 0;
                    for (int f = 0; f < 6; ++f) {
                        const int64_t j = fvi(m, d, xc + NOX[f], yc + NOY[f], rc + NOR[f]);
                        const double fij = antidiff[6 * i + f];
                        const double a = fij > 0 ? fmin(Rp[i], Rm[j]) : fmin(Rm[i], Rp[j]);
                        corr += a * fij;
                    }
                    T[i] = TL[i] + corr;
                }
}

/* one explicit FCT step (fct_advection_diffusion.hpp:719-757); work holds
 * T_L, R+, R- (3 FV scalars, zero on the first call) and the fluxes (6 per cell) */
void orc_fct_explicit_step(const orc_mesh* m, const double* centers, const double* vel, double* T, double dt,
                           double kappa, const double* source, int subtract_divergence, int apply_cmb,
                           double t_cmb, int apply_surf, double t_surf, double* work)
{
    const int64_t sz = orc_fv_size(m);
    double *TL = work, *Rp = work + sz, *Rm = work + 2 * sz, *ad = work + 3 * sz;
    orc_fct_predictor(m, centers, vel, T, dt, kappa, source, subtract_divergence, TL, ad);
    if (apply_cmb || apply_surf) orc_fv_dirichlet(m, TL, apply_cmb, t_cmb, apply_surf, t_surf);
    orc_fct_limiter(m, TL, ad, Rp, Rm);
    orc_fct_correction(m, TL, ad, Rp, Rm, T);
}

/* ---------------------------------------------------------------- FV -> FE
 * right-hand side of the L2 projection (conversion.hpp:131-202):
 * b_i = sum_K T_K int_K phi_i, interface summed; b is a nodal scalar [10][n][n][nr] */
void orc_fv_to_fe_rhs(const orc_mesh* m, const double* Tfv, double* b)
{
    const int n = m->n, nr = m->nr;
    memset(b, 0, sizeof(double) * 10 * n * n * nr);
    for (int d = 0; d < 10; ++d)
        for (int xc = 1; xc <= n - 1; ++xc)
            for (int yc = 1; yc <= n - 1; ++yc) {
                double P[2][3][3];
                fwedge_surf(m, d, xc - 1, yc - 1, P);
                for (int rc = 1; rc <= nr - 1; ++rc) {
                    const double r1 = m->radii[rc - 1], r2 = m->radii[rc];
                    const double tv = Tfv[fvi(m, d, xc, yc, rc)];
                    for (int w = 0; w < 2; ++w)
                        for (int i = 0; i < 6; ++i) {
                            double s = 0;
                            for (int q = 0; q < 6; ++q)
                                s += fshape(i, FQ6P[q][0], FQ6P[q][1], FQ6P[q][2]) *
                                     fabs_det(P[w], r1, r2, FQ6P[q]) * FQ6W;
                            s *= tv;
                            const int xx = xc - 1 + FWOFF[w][i][0], yy = yc - 1 + FWOFF[w][i][1],
                                      rr = rc - 1 + FWOFF[w][i][2];
                            b[(((int64_t)d * n + xx) * n + yy) * nr + rr] += s;
                        }
                }
            }
    orc_interface_sum(n, nr, 1, m->group_of, m->ngroups, b);
}
