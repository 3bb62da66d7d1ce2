/*
 * CPU baseline of the benchmark: a matrix-free OpenMP restatement of the
 * epsilon-div-div apply (Dirichlet / Neumann shells), the CPU counterpart of
 * the reference's fast path (epsilon_divdiv_kerngen.hpp:1055-1389) in this
 * build's own cell formulation.  TEST INFRASTRUCTURE: only the benchmark's
 * CPU-baseline leg and the tests call it, never the product.
 *
 * Cell arithmetic (one-point centroid quadrature per wedge): with the wedge's
 * lateral triangle v0, v1, v2 on the unit sphere, A = [v1-v0, v2-v0, c] with
 * c = (v0+v1+v2)/3, l1 = A^-T e0 / 2, l2 = A^-T e1 / 2, z = A^-T e2 / 6, the
 * shape gradients at the centroid are a_i -/+ zz with a_i = l_i / r_mid and
 * zz = z / (dr/2) (a_0 = -(a_1 + a_2)), |det J| = r_mid^2 (dr/2) |det A|.
 * Per wedge: G = sum over the six nodes of u (x) grad N, sigma = kwJ (2 sym G
 * - 2/3 tr G I) with kwJ = mean(eta) |det J|, and node i receives sigma grad N_i.
 * Dirichlet shells: the trial side reads zero there and the boundary rows are
 * replaced by the Neumann diagonal times the source (the reference's masks).
 * Interface duplicates are summed at the end (the additive communication).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IDX(n, nr, d, x, y, r) ((((int64_t)(d) * (n) + (x)) * (n) + (y)) * (nr) + (r))

/* wedge corner offsets in the lateral quad: wedge 0 (x,y),(x+1,y),(x,y+1);
 * wedge 1 (x+1,y+1),(x,y+1),(x+1,y) (kernel_helpers.hpp:302-327) */
static const int QX[2][3] = {{0, 1, 0}, {1, 0, 1}};
static const int QY[2][3] = {{0, 0, 1}, {1, 1, 0}};

static void wedge_geometry(const double* v0, const double* v1, const double* v2, double l1[3], double l2[3],
                           double z[3], double* detA)
{
    double A[3][3];
    for (int k = 0; k < 3; k++) {
        A[k][0] = v1[k] - v0[k];
        A[k][1] = v2[k] - v0[k];
        A[k][2] = (v0[k] + v1[k] + v2[k]) / 3.0;
    }
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    /* rows of A^-1 = columns of A^-T: cof(A)^T / det */
    double inv[3][3];
    inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
    inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
    inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    /* A^-T e_j = j-th row of A^-1 */
    for (int k = 0; k < 3; k++) {
        l1[k] = 0.5 * inv[0][k];
        l2[k] = 0.5 * inv[1][k];
        z[k] = inv[2][k] / 6.0;
    }
    *detA = fabs(det);
}

/* y = A x over all stored node copies (Replace, interface sum included).
 * coords [10][n][n][3], radii [nr], eta [10][n][n][nr], x, y [3][10][n][n][nr],
 * group_of [10][n][n] lateral duplicate groups (ngroups), bc: 0 Neumann, 1 Dirichlet. */
void cpu_eps_apply(int n, int nr, const double* coords, const double* radii, const double* eta, const double* x,
                   double* y, int bc_cmb, int bc_surf, const int32_t* group_of, int ngroups)
{
    const int64_t N = 10LL * n * n * nr, nlat = 10LL * n * n;
    const int ncl = n - 1, ncr = nr - 1;
    memset(y, 0, sizeof(double) * 3 * N);
    /* Neumann diagonal at the two boundary levels, [shell][3][nlat] */
    double* dB = (double*)calloc(2 * 3 * nlat, sizeof(double));
    /* rows of cells of one parity touch disjoint node rows */
    for (int parity = 0; parity < 2; parity++) {
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
        for (int d = 0; d < 10; d++)
            for (int xr = parity; xr < ncl; xr += 2) {
                for (int yc = 0; yc < ncl; yc++) {
                    for (int w = 0; w < 2; w++) {
                        int64_t lat[3];
                        const double* v[3];
                        for (int i = 0; i < 3; i++) {
                            lat[i] = ((int64_t)d * n + xr + QX[w][i]) * n + yc + QY[w][i];
                            v[i] = coords + 3 * lat[i];
                        }
                        double l1[3], l2[3], z[3], detA;
                        wedge_geometry(v[0], v[1], v[2], l1, l2, z, &detA);
                        for (int r = 0; r < ncr; r++) {
                            const double rm = 0.5 * (radii[r] + radii[r + 1]), hd = 0.5 * (radii[r + 1] - radii[r]);
                            const double kwJ_per_eta = rm * rm * hd * detA / 6.0;
                            double a[3][3], zz[3];
                            for (int k = 0; k < 3; k++) {
                                a[1][k] = l1[k] / rm;
                                a[2][k] = l2[k] / rm;
                                a[0][k] = -(a[1][k] + a[2][k]);
                                zz[k] = z[k] / hd;
                            }
                            const int zb = (r == 0 && bc_cmb == 1), zt = (r == ncr - 1 && bc_surf == 1);
                            double G[3][3] = {{0}}, esum = 0.0;
                            for (int i = 0; i < 3; i++) {
                                const int64_t ib = lat[i] * nr + r;
                                esum += eta[ib] + eta[ib + 1];
                                for (int c = 0; c < 3; c++) {
                                    const double xb = zb ? 0.0 : x[c * N + ib], xt = zt ? 0.0 : x[c * N + ib + 1];
                                    for (int k = 0; k < 3; k++) G[c][k] += (xb + xt) * a[i][k] + (xt - xb) * zz[k];
                                }
                            }
                            const double kwJ = esum * kwJ_per_eta;
                            const double tr = G[0][0] + G[1][1] + G[2][2];
                            double s[3][3];
                            for (int c = 0; c < 3; c++)
                                for (int k = 0; k < 3; k++)
                                    s[c][k] = kwJ * (G[c][k] + G[k][c] - (c == k ? 2.0 / 3.0 * tr : 0.0));
                            for (int i = 0; i < 3; i++) {
                                const int64_t ib = lat[i] * nr + r;
                                for (int c = 0; c < 3; c++) {
                                    double p = 0.0, q = 0.0;
                                    for (int k = 0; k < 3; k++) {
                                        p += s[c][k] * a[i][k];
                                        q += s[c][k] * zz[k];
                                    }
                                    y[c * N + ib] += p - q;
                                    y[c * N + ib + 1] += p + q;
                                }
                                /* Neumann diagonal of the boundary-level nodes */
                                for (int lev = 0; lev < 2; lev++) {
                                    const int shell = (lev == 0 && r == 0) ? 0 : ((lev == 1 && r == ncr - 1) ? 1 : -1);
                                    if (shell < 0) continue;
                                    double g[3], gg = 0.0;
                                    for (int k = 0; k < 3; k++) {
                                        g[k] = a[i][k] + (lev ? zz[k] : -zz[k]);
                                        gg += g[k] * g[k];
                                    }
                                    for (int c = 0; c < 3; c++)
                                        dB[((int64_t)shell * 3 + c) * nlat + lat[i]] += kwJ * (gg + g[c] * g[c] / 3.0);
                                }
                            }
                        }
                    }
                }
            }
    }
    /* Dirichlet rows: the Neumann diagonal times the source */
    for (int shell = 0; shell < 2; shell++) {
        if ((shell == 0 ? bc_cmb : bc_surf) != 1) continue;
        const int r = shell == 0 ? 0 : nr - 1;
#pragma omp parallel for
        for (int64_t l = 0; l < nlat; l++)
            for (int c = 0; c < 3; c++)
                y[c * N + l * nr + r] = dB[((int64_t)shell * 3 + c) * nlat + l] * x[c * N + l * nr + r];
    }
    free(dB);
    /* additive interface sum over duplicated lateral nodes */
    int32_t* cnt = (int32_t*)calloc(ngroups, sizeof(int32_t));
    for (int64_t l = 0; l < nlat; l++) cnt[group_of[l]]++;
    double* acc = (double*)calloc((size_t)ngroups * nr, sizeof(double));
    for (int c = 0; c < 3; c++) {
        memset(acc, 0, sizeof(double) * (size_t)ngroups * nr);
        for (int64_t l = 0; l < nlat; l++)
            if (cnt[group_of[l]] > 1)
                for (int r = 0; r < nr; r++) acc[(int64_t)group_of[l] * nr + r] += y[c * N + l * nr + r];
        for (int64_t l = 0; l < nlat; l++)
            if (cnt[group_of[l]] > 1)
                for (int r = 0; r < nr; r++) y[c * N + l * nr + r] = acc[(int64_t)group_of[l] * nr + r];
    }
    free(acc);
    free(cnt);
}
