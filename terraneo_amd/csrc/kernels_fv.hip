// Hex finite-volume energy step: flux-corrected transport (FCT) of a scalar
// by the Stokes velocity, with two-point diffusion (FP64).
//
// Reference: fv/hex/operators/fct_advection_diffusion.hpp (stages),
// fv/hex/operators/geometry_helper.hpp:178-360 (cell geometry),
// fv/hex/helpers.hpp:64-190 (Dirichlet values, cell centres),
// fv/hex/conversion.hpp:131-211 (FV -> FE projection right-hand side),
// communication/shell/fv_communication.hpp (ghost layers).
//
// FV fields carry one ghost layer: [10][n+1][n+1][nr+1], r fastest; real
// cells (xc, yc, rc) in [1, n-1]^2 x [1, nr-1] are the FE hex cells
// (xc-1, yc-1, rc-1).  Every kernel runs one thread per real cell (or per
// FE node for the projection), radially consecutive threads on radially
// consecutive cells, so every field access of a wave is one contiguous
// radial run.  Nothing is accumulated with atomics: every output element is
// written by one thread, so results repeat bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tn_fv.hpp"

namespace tn {

namespace {

constexpr int FV_BLOCK = 256;

__device__ __forceinline__ int64_t fvi(int n, int nr, int d, int x, int y, int r)
{
    return (((int64_t)d * (n + 1) + x) * (n + 1) + y) * (nr + 1) + r;
}

// wedge local node offsets (kernel_helpers.hpp:302-327)
__constant__ int FV_WOFF[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                     {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};

// Felippa 3x2 wedge rule (fe/wedge/quadrature/quadrature.hpp:67-110): triangle
// points (2/3,1/6), (1/6,2/3), (1/6,1/6) at zeta = -+1/sqrt(3), weight 1/6
constexpr double QT_A = 0.6666666666666666, QT_B = 0.1666666666666667, QZ = 0.5773502691896257,
                 QW = 0.1666666666666667;

__device__ __forceinline__ void cell_real(int64_t t, int n, int nr, int& d, int& xc, int& yc, int& rc)
{
    rc = 1 + (int)(t % (nr - 1));
    t /= (nr - 1);
    yc = 1 + (int)(t % (n - 1));
    t /= (n - 1);
    xc = 1 + (int)(t % (n - 1));
    d = (int)(t / (n - 1));
}

// lateral corner positions of the two wedges of FE hex cell (x, y)
__device__ __forceinline__ void wedge_corners(const double* __restrict__ coords, int n, int d, int x, int y,
                                              double P[2][3][3])
{
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double* c = coords + 3 * (((int64_t)d * n + x + FV_WOFF[w][i][0]) * n + y + FV_WOFF[w][i][1]);
            P[w][i][0] = c[0];
            P[w][i][1] = c[1];
            P[w][i][2] = c[2];
        }
}

// |det J| of the wedge map x = r(zeta) L(xi, eta) at (xi, eta, zeta)
// (integrands.hpp:623-670): det [r (p1-p0), r (p2-p0), (dr/2) L]
__device__ __forceinline__ double abs_det(const double P[3][3], double r1, double r2, double xi, double eta,
                                          double zeta)
{
    const double r = r1 + 0.5 * (r2 - r1) * (1 + zeta);
    const double gr = 0.5 * (r2 - r1);
    double J[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        J[a][0] = (P[1][a] - P[0][a]) * r;
        J[a][1] = (P[2][a] - P[0][a]) * r;
        J[a][2] = ((1 - xi - eta) * P[0][a] + xi * P[1][a] + eta * P[2][a]) * gr;
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    return fabs(det);
}

__device__ __forceinline__ void qpoint(int q, double& xi, double& eta, double& zeta)
{
    const int k = q % 3;
    xi = k == 0 ? QT_A : QT_B;
    eta = k == 1 ? QT_A : QT_B;
    zeta = q < 3 ? -QZ : QZ;
}

// Geometry of real FV cell (d, xc, yc, rc): beta[f] = int_face u.n dS and
// S[f] = int_face n dS (faces 0: x-1, 1: x+1, 2: y-1, 3: y+1, 4: r-1, 5: r+1),
// each oriented towards the neighbour's centre, and the cell volume M.
//
// The reference integrates the faces with 2 x 2 Gauss points (lateral faces)
// and the 3-point triangle rule on both wedges (radial faces), and the volume
// with the Felippa 3x2 rule.  With the wedge map X = rho(zeta) L(xi, eta)
// (rho linear from r1 to r2, L on the flat triangle V0 V1 V2, h = (r2-r1)/2)
// these rules are exact, and the integrals have closed forms:
//   lateral face xi = 0 of a wedge:  n = rho h (V2 x V0), nodes V0, V2
//   lateral face eta = 0:            n = rho h (V1 x V0), nodes V0, V1
//     S = h (r1 + r2) (Vb x Va),
//     beta = h (Vb x Va) . [ab (u_a,bot + u_b,bot) + at (u_a,top + u_b,top)] / 2,
//     ab = (2 r1 + r2) / 3, at = (r1 + 2 r2) / 3     (int rho (1 -+ zeta)/2 dzeta)
//   radial face at radius rr (r1 bottom, r2 top), summed over both wedges:
//     n = rr^2 (V1 - V0) x (V2 - V0),  S = n / 2,  beta = n . (u_0 + u_1 + u_2) / 6
//   volume: sum over the wedges of |det(V0, V1, V2)| (r2^3 - r1^3) / 6
// so the results equal the reference's quadrature up to rounding.
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Lateral part of the geometry of hex column (d, x, y), the same at every
// radius: g[0..12) the lateral face normals Vb x Va of faces 0..3, g[12..15)
// and g[15..18) the two triangles' (V1 - V0) x (V2 - V0), g[18] the sum of the
// wedges' |det(V0, V1, V2)|.  Corners c00 = (x, y), c10 = (x+1, y),
// c01 = (x, y+1), c11 = (x+1, y+1); wedge 0: V0 = c00, V1 = c10, V2 = c01;
// wedge 1: W0 = c11, W1 = c01, W2 = c10.  Lateral faces (face, Va, Vb):
// 0 (c00, c01) xi = 0 of wedge 0: n ~ V2 x V0 = c01 x c00; 2 (c00, c10)
// eta = 0 of wedge 0: V1 x V0 = c10 x c00; 1 (c11, c10): W2 x W0 = c10 x c11;
// 3 (c11, c01): W1 x W0 = c01 x c11.
__device__ __forceinline__ void column_geometry(const double* __restrict__ coords, int n, int d, int x, int y,
                                                double g[FV_COLUMN])
{
    double c[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double* p = coords + 3 * (((int64_t)d * n + x + (k & 1)) * n + y + (k >> 1));
        c[k][0] = p[0];
        c[k][1] = p[1];
        c[k][2] = p[2];
    }
    const int FA[4] = {0, 3, 0, 3}, FB[4] = {2, 1, 1, 2};  // corner indices of faces 0, 1, 2, 3
#pragma unroll
    for (int f = 0; f < 4; ++f) cross3(c[FB[f]], c[FA[f]], g + 3 * f);
    double e1[3], e2[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        e1[j] = c[1][j] - c[0][j];
        e2[j] = c[2][j] - c[0][j];
    }
    cross3(e1, e2, g + 12);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        e1[j] = c[2][j] - c[3][j];
        e2[j] = c[1][j] - c[3][j];
    }
    cross3(e1, e2, g + 15);
    // |det(V0, V1, V2)| per wedge (= V0 . (V1 x V2))
    double v12[3], w12[3];
    cross3(c[1], c[2], v12);
    cross3(c[2], c[1], w12);
    const double det0 = fabs(c[0][0] * v12[0] + c[0][1] * v12[1] + c[0][2] * v12[2]);
    const double det1 = fabs(c[3][0] * w12[0] + c[3][1] * w12[1] + c[3][2] * w12[2]);
    g[18] = det0 + det1;
    g[19] = 0.0;
}

// Geometry of real FV cell (d, xc, yc, rc) from its column's lateral part
// (a.columns, see column_geometry) and the radii.
__device__ void cell_geometry(const FvArgs& a, const double* __restrict__ vel, int64_t N, int d, int xc, int yc,
                              int rc, double beta[6], double& M, double S[6][3])
{
    const int n = a.n, nr = a.nr;
    const int x = xc - 1, y = yc - 1, r = rc - 1;
    double g[FV_COLUMN];
    {
        const double2* gp =
            reinterpret_cast<const double2*>(a.columns + FV_COLUMN * (((int64_t)d * (n - 1) + x) * (n - 1) + y));
#pragma unroll
        for (int k = 0; k < FV_COLUMN / 2; ++k) {
            const double2 v = gp[k];
            g[2 * k] = v.x;
            g[2 * k + 1] = v.y;
        }
    }
    // nodal velocity at the 4 corners x {bottom, top}
    double ub[4][3], ut[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t node = (((int64_t)d * n + x + (k & 1)) * n + y + (k >> 1)) * nr + r;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            ub[k][j] = vel[j * N + node];
            ut[k][j] = vel[j * N + node + 1];
        }
    }
    const double r1 = a.radii[r], r2 = a.radii[r + 1];
    const double h = 0.5 * (r2 - r1);
    const double ab = (2.0 * r1 + r2) / 3.0, at = (r1 + 2.0 * r2) / 3.0;
    const int FA[4] = {0, 3, 0, 3}, FB[4] = {2, 1, 1, 2};  // corner indices of faces 0, 1, 2, 3
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double* nv = g + 3 * f;
        const double sS = h * (r1 + r2);
        double w = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            w += nv[j] * (ab * (ub[FA[f]][j] + ub[FB[f]][j]) + at * (ut[FA[f]][j] + ut[FB[f]][j]));
        beta[f] = 0.5 * h * w;
        S[f][0] = sS * nv[0];
        S[f][1] = sS * nv[1];
        S[f][2] = sS * nv[2];
    }
    // radial faces: both wedges' (V1 - V0) x (V2 - V0)
    const double* n0 = g + 12;
    const double* n1 = g + 15;
    double bb = 0, bt = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        bb += n0[j] * (ub[0][j] + ub[1][j] + ub[2][j]) + n1[j] * (ub[3][j] + ub[2][j] + ub[1][j]);
        bt += n0[j] * (ut[0][j] + ut[1][j] + ut[2][j]) + n1[j] * (ut[3][j] + ut[2][j] + ut[1][j]);
    }
    beta[4] = r1 * r1 * bb / 6.0;
    beta[5] = r2 * r2 * bt / 6.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        S[4][j] = 0.5 * r1 * r1 * (n0[j] + n1[j]);
        S[5][j] = 0.5 * r2 * r2 * (n0[j] + n1[j]);
    }
    // orientation: towards the neighbour's centre (ghost centres included)
    const double* cc = a.centers + 3 * fvi(n, nr, d, xc, yc, rc);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int ox = f == 0 ? -1 : (f == 1 ? 1 : 0), oy = f == 2 ? -1 : (f == 3 ? 1 : 0),
                  orr = f == 4 ? -1 : (f == 5 ? 1 : 0);
        const double* nc = a.centers + 3 * fvi(n, nr, d, xc + ox, yc + oy, rc + orr);
        const double s = S[f][0] * (nc[0] - cc[0]) + S[f][1] * (nc[1] - cc[1]) + S[f][2] * (nc[2] - cc[2]);
        if (s < 0) {
            S[f][0] = -S[f][0];
            S[f][1] = -S[f][1];
            S[f][2] = -S[f][2];
            beta[f] = -beta[f];
        }
    }
    M = g[18] * (r2 - r1) * (r1 * r1 + r1 * r2 + r2 * r2) / 6.0;  // (r2^3 - r1^3) / 6 without cancellation
}

__device__ __forceinline__ int64_t neighbour(int n, int nr, int d, int xc, int yc, int rc, int f)
{
    const int ox = f == 0 ? -1 : (f == 1 ? 1 : 0), oy = f == 2 ? -1 : (f == 3 ? 1 : 0),
              orr = f == 4 ? -1 : (f == 5 ? 1 : 0);
    return fvi(n, nr, d, xc + ox, yc + oy, rc + orr);
}

// ---------------------------------------------------------------- kernels

// lateral geometry of every hex column (column_geometry), one thread per column
__global__ __launch_bounds__(FV_BLOCK) void k_fv_columns(FvArgs a, double* __restrict__ columns)
{
    const int n = a.n;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 10LL * (n - 1) * (n - 1)) return;
    const int y = (int)(t % (n - 1)), x = (int)((t / (n - 1)) % (n - 1)), d = (int)(t / ((int64_t)(n - 1) * (n - 1)));
    double g[FV_COLUMN];
    column_geometry(a.coords, n, d, x, y, g);
    double2* o = reinterpret_cast<double2*>(columns + FV_COLUMN * t);
#pragma unroll
    for (int k = 0; k < FV_COLUMN / 2; ++k) o[k] = make_double2(g[2 * k], g[2 * k + 1]);
}

// cell centre = (int_K x) / |K| with the Felippa rule on both wedges (helpers.hpp:123-175)
__global__ __launch_bounds__(FV_BLOCK) void k_fv_centers(FvArgs a, double* __restrict__ c)
{
    const int n = a.n, nr = a.nr;
    const int64_t cells = 10LL * (n - 1) * (n - 1) * (nr - 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    int d, xc, yc, rc;
    cell_real(t, n, nr, d, xc, yc, rc);
    double P[2][3][3];
    wedge_corners(a.coords, n, d, xc - 1, yc - 1, P);
    const double r1 = a.radii[rc - 1], r2 = a.radii[rc];
    double vol = 0, I[3] = {0, 0, 0};
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double xi, eta, zeta;
            qpoint(q, xi, eta, zeta);
            const double Jw = abs_det(P[w], r1, r2, xi, eta, zeta) * QW;
            vol += Jw;
            const double rr = r1 + 0.5 * (r2 - r1) * (1 + zeta);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                I[k] += rr * ((1 - xi - eta) * P[w][0][k] + xi * P[w][1][k] + eta * P[w][2][k]) * Jw;
        }
    double* o = c + 3 * fvi(n, nr, d, xc, yc, rc);
    o[0] = I[0] / vol;
    o[1] = I[1] / vol;
    o[2] = I[2] / vol;
}

// lateral ghost cells <- the neighbouring diamonds' real cells (real radial range)
__global__ __launch_bounds__(FV_BLOCK) void k_fv_ghost(const int32_t* __restrict__ pairs, int cnt, int nr,
                                                       int ncomp, double* __restrict__ f)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)cnt * (nr - 1)) return;
    const int rc = 1 + (int)(t % (nr - 1));
    const int k = (int)(t / (nr - 1));
    const int64_t dst = ((int64_t)pairs[2 * k] * (nr + 1) + rc) * ncomp;
    const int64_t src = ((int64_t)pairs[2 * k + 1] * (nr + 1) + rc) * ncomp;
    for (int c = 0; c < ncomp; ++c) f[dst + c] = f[src + c];
}

// local stable time step M / lambda per cell, minimum per block
// (fct_advection_diffusion.hpp:110-164)
__global__ __launch_bounds__(FV_BLOCK) void k_fv_dt(FvArgs a, const double* __restrict__ vel, double kappa,
                                                    double* __restrict__ part)
{
    const int n = a.n, nr = a.nr;
    const int64_t cells = 10LL * (n - 1) * (n - 1) * (nr - 1);
    const int64_t N = 10LL * n * n * nr;
    double dtc = 1e300;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += (int64_t)gridDim.x * blockDim.x) {
        int d, xc, yc, rc;
        cell_real(t, n, nr, d, xc, yc, rc);
        double beta[6], M, S[6][3];
        cell_geometry(a, vel, N, d, xc, yc, rc, beta, M, S);
        const double* ci = a.centers + 3 * fvi(n, nr, d, xc, yc, rc);
        double lambda = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            if (beta[f] < 0) lambda -= beta[f];
            if (kappa > 0) {
                const double* cj = a.centers + 3 * neighbour(n, nr, d, xc, yc, rc, f);
                const double den = S[f][0] * (cj[0] - ci[0]) + S[f][1] * (cj[1] - ci[1]) + S[f][2] * (cj[2] - ci[2]);
                if (den > 0) lambda += kappa * (S[f][0] * S[f][0] + S[f][1] * S[f][1] + S[f][2] * S[f][2]) / den;
            }
        }
        dtc = fmin(dtc, lambda > 0 ? M / lambda : 1e30);
    }
    // block minimum (order-independent)
    __shared__ double red[FV_BLOCK / 64];
    for (int o = 32; o > 0; o >>= 1) dtc = fmin(dtc, __shfl_xor(dtc, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dtc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < FV_BLOCK / 64; ++k) m = fmin(m, red[k]);
        part[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(256) void k_min_finalize(const double* __restrict__ part, int nblk,
                                                      double* __restrict__ out)
{
    double m = 1e300;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) m = fmin(m, part[b]);
    for (int o = 32; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}

// predictor (fct_advection_diffusion.hpp:261-402): low-order upwind +
// two-point diffusion T_L, and the pre-scaled antidiffusive fluxes
// (dt / M) (|beta| / 2) (T_i - T_j), stored face-major [6][fv cells]
__global__ __launch_bounds__(FV_BLOCK) void k_fct_predictor(FvArgs a, const double* __restrict__ vel,
                                                            const double* __restrict__ T, double dt, double kappa,
                                                            const double* __restrict__ source, int subtract_div,
                                                            double* __restrict__ TL, double* __restrict__ ad,
                                                            int64_t fvsize)
{
    const int n = a.n, nr = a.nr;
    const int64_t cells = 10LL * (n - 1) * (n - 1) * (nr - 1);
    const int64_t N = 10LL * n * n * nr;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    int d, xc, yc, rc;
    cell_real(t, n, nr, d, xc, yc, rc);
    double beta[6], M, S[6][3];
    cell_geometry(a, vel, N, d, xc, yc, rc, beta, M, S);
    const int64_t i = fvi(n, nr, d, xc, yc, rc);
    const double Ti = T[i];
    const double* ci = a.centers + 3 * i;
    const double sc = dt / M;
    double lo_diag = 0, lo_update = 0, bsum = 0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int64_t j = neighbour(n, nr, d, xc, yc, rc, f);
        const double Tj = T[j];
        if (beta[f] >= 0)
            lo_diag += beta[f];
        else
            lo_update += beta[f] * Tj;
        bsum += beta[f];
        const double* cj = a.centers + 3 * j;
        const double den = S[f][0] * (cj[0] - ci[0]) + S[f][1] * (cj[1] - ci[1]) + S[f][2] * (cj[2] - ci[2]);
        const double dc = kappa * ((S[f][0] * S[f][0] + S[f][1] * S[f][1] + S[f][2] * S[f][2]) / den);
        lo_diag += dc;
        lo_update -= dc * Tj;
        ad[f * fvsize + i] = sc * (fabs(beta[f]) / 2.0) * (Ti - Tj);
    }
    if (subtract_div) lo_diag -= bsum;
    double v = Ti - sc * (lo_diag * Ti + lo_update);
    if (source) v += dt * source[i];
    TL[i] = v;
}

// Dirichlet values on the shells: the real boundary cell and the ghost beyond (helpers.hpp:64-96)
__global__ __launch_bounds__(FV_BLOCK) void k_fv_dirichlet(int n, int nr, double* __restrict__ f, int apply_cmb,
                                                           double t_cmb, int apply_surf, double t_surf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 10LL * (n - 1) * (n - 1)) return;
    const int yc = 1 + (int)(t % (n - 1));
    const int xc = 1 + (int)((t / (n - 1)) % (n - 1));
    const int d = (int)(t / ((int64_t)(n - 1) * (n - 1)));
    if (apply_cmb) {
        f[fvi(n, nr, d, xc, yc, 1)] = t_cmb;
        f[fvi(n, nr, d, xc, yc, 0)] = t_cmb;
    }
    if (apply_surf) {
        f[fvi(n, nr, d, xc, yc, nr - 1)] = t_surf;
        f[fvi(n, nr, d, xc, yc, nr)] = t_surf;
    }
}

// limiter (fct_advection_diffusion.hpp:485-537): Zalesak R+ / R-
__global__ __launch_bounds__(FV_BLOCK) void k_fct_limiter(int n, int nr, const double* __restrict__ TL,
                                                          const double* __restrict__ ad, int64_t fvsize,
                                                          double* __restrict__ Rp, double* __restrict__ Rm)
{
    const int64_t cells = 10LL * (n - 1) * (n - 1) * (nr - 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    int d, xc, yc, rc;
    cell_real(t, n, nr, d, xc, yc, rc);
    const int64_t i = fvi(n, nr, d, xc, yc, rc);
    const double ti = TL[i];
    double pp = 0, pm = 0, tmax = ti, tmin = ti;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const double fij = ad[f * fvsize + i];
        if (fij > 0)
            pp += fij;
        else
            pm += fij;
        const double tj = TL[neighbour(n, nr, d, xc, yc, rc, f)];
        tmax = fmax(tmax, tj);
        tmin = fmin(tmin, tj);
    }
    Rp[i] = pp > 0 ? fmin(1.0, (tmax - ti) / pp) : 1.0;
    Rm[i] = pm < 0 ? fmin(1.0, (tmin - ti) / pm) : 1.0;
}

// correction (fct_advection_diffusion.hpp:604-658): T = T_L + sum_j alpha_ij f_ij
__global__ __launch_bounds__(FV_BLOCK) void k_fct_correction(int n, int nr, const double* __restrict__ TL,
                                                             const double* __restrict__ ad, int64_t fvsize,
                                                             const double* __restrict__ Rp,
                                                             const double* __restrict__ Rm, double* __restrict__ T)
{
    const int64_t cells = 10LL * (n - 1) * (n - 1) * (nr - 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    int d, xc, yc, rc;
    cell_real(t, n, nr, d, xc, yc, rc);
    const int64_t i = fvi(n, nr, d, xc, yc, rc);
    const double rpi = Rp[i], rmi = Rm[i];
    double corr = 0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int64_t j = neighbour(n, nr, d, xc, yc, rc, f);
        const double fij = ad[f * fvsize + i];
        const double al = fij > 0 ? fmin(rpi, Rm[j]) : fmin(rmi, Rp[j]);
        corr += al * fij;
    }
    T[i] = TL[i] + corr;
}

// right-hand side of the FV -> FE L2 projection (conversion.hpp:151-200):
// b_i = sum over the cells K around node i of T_K int_K phi_i, gathered per
// node copy (no atomics); the interface sum follows.  The reference evaluates
// int_K phi_i |det J| with the Felippa 3x2 rule, which is exact: per wedge
//   |det(V0, V1, V2)| / 6 x (I_bot or I_top),  I = int rho^2 x radial hat drho,
//   I_bot = dr (r1^2/2 + r1 dr/3 + dr^2/12),  I_top = dr (r1^2/2 + 2 r1 dr/3 + dr^2/4)
// Tfv null: every T_K = 1, which gives the lumped diagonal D_i = int phi_i.
__global__ __launch_bounds__(FV_BLOCK) void k_fv_to_fe_rhs(FvArgs a, const double* __restrict__ Tfv,
                                                           double* __restrict__ b)
{
    const int n = a.n, nr = a.nr;
    const int64_t N = 10LL * n * n * nr;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int r = (int)(t % nr);
    const int y = (int)((t / nr) % n);
    const int x = (int)((t / ((int64_t)nr * n)) % n);
    const int d = (int)(t / ((int64_t)nr * n * n));
    double acc = 0;
    for (int cx = x - 1; cx <= x; ++cx) {
        if (cx < 0 || cx > n - 2) continue;
        for (int cy = y - 1; cy <= y; ++cy) {
            if (cy < 0 || cy > n - 2) continue;
            double P[2][3][3];
            wedge_corners(a.coords, n, d, cx, cy, P);
            double det[2];
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                double v12[3];
                cross3(P[w][1], P[w][2], v12);
                det[w] = fabs(P[w][0][0] * v12[0] + P[w][0][1] * v12[1] + P[w][0][2] * v12[2]) / 6.0;
            }
            // the node's weight in this quad: the wedges that contain its lateral corner
            const int lx = x - cx, ly = y - cy;
            const double wlat = (lx == 0 && ly == 0) ? det[0] : ((lx == 1 && ly == 1) ? det[1] : det[0] + det[1]);
            for (int cr = r - 1; cr <= r; ++cr) {
                if (cr < 0 || cr > nr - 2) continue;
                const double r1 = a.radii[cr], dr = a.radii[cr + 1] - r1;
                const double irad = (r == cr + 1) ? dr * (0.5 * r1 * r1 + r1 * dr * (2.0 / 3.0) + 0.25 * dr * dr)
                                                  : dr * (0.5 * r1 * r1 + r1 * dr / 3.0 + dr * dr / 12.0);
                acc += wlat * irad * (Tfv ? Tfv[fvi(n, nr, d, cx + 1, cy + 1, cr + 1)] : 1.0);
            }
        }
    }
    b[t] = acc;
}

// u_i = b_i (1 / D_i): the lumped FV -> FE transfer's last step
// (conversion.hpp:337-340, invert_entries then scale_in_place)
__global__ __launch_bounds__(FV_BLOCK) void k_fv_lumped_scale(const double* __restrict__ D, double* __restrict__ u,
                                                             int64_t N)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) u[t] = u[t] * (1.0 / D[t]);
}

// FE -> FV cell averages (conversion.hpp:348-416): T_K = int_K u_h / |K| with
// the Felippa 3x2 rule on both wedges; wedge shape i = lateral hat (1 - xi -
// eta, xi, eta)[i % 3] x radial hat ((1 - zeta) / 2 for i < 3, else (1 + zeta) / 2)
__global__ __launch_bounds__(FV_BLOCK) void k_fe_to_fv(FvArgs a, const double* __restrict__ u,
                                                       double* __restrict__ Tfv)
{
    const int n = a.n, nr = a.nr;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 10LL * (n - 1) * (n - 1) * (nr - 1)) return;
    int d, xc, yc, rc;
    cell_real(t, n, nr, d, xc, yc, rc);
    const int x = xc - 1, y = yc - 1, r = rc - 1;
    double P[2][3][3];
    wedge_corners(a.coords, n, d, x, y, P);
    const double r1 = a.radii[r], r2 = a.radii[r + 1];
    double vol = 0, integral = 0;
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        double lu[6];
#pragma unroll
        for (int i = 0; i < 6; ++i)
            lu[i] = u[(((int64_t)d * n + x + FV_WOFF[w][i][0]) * n + y + FV_WOFF[w][i][1]) * nr + r + FV_WOFF[w][i][2]];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double xi, eta, zeta;
            qpoint(q, xi, eta, zeta);
            const double Jw = abs_det(P[w], r1, r2, xi, eta, zeta) * QW;
            vol += Jw;
            const double lat[3] = {1 - xi - eta, xi, eta};
            const double bot = 0.5 * (1 - zeta), top = 0.5 * (1 + zeta);
            double uq = 0;
#pragma unroll
            for (int i = 0; i < 6; ++i) uq += lu[i] * lat[i % 3] * (i < 3 ? bot : top);
            integral += uq * Jw;
        }
    }
    Tfv[fvi(n, nr, d, xc, yc, rc)] = integral / vol;
}

int blocks_for(int64_t work) { return (int)((work + FV_BLOCK - 1) / FV_BLOCK); }

}  // namespace

int64_t fv_real_cells(int n, int nr) { return 10LL * (n - 1) * (n - 1) * (nr - 1); }
int64_t fv_size(int n, int nr) { return 10LL * (n + 1) * (n + 1) * (nr + 1); }
// grid-stride launch of the stable-step kernel: at most FV_DT_BLOCKS partial minima
constexpr int FV_DT_BLOCKS = 2048;
int fv_dt_blocks(int n, int nr) { return (int)std::min<int64_t>(blocks_for(fv_real_cells(n, nr)), FV_DT_BLOCKS); }

void launch_fv_centers(const FvArgs& a, double* c, hipStream_t s)
{
    hipLaunchKernelGGL(k_fv_centers, dim3(blocks_for(fv_real_cells(a.n, a.nr))), dim3(FV_BLOCK), 0, s, a, c);
}

void launch_fv_columns(const FvArgs& a, double* columns, hipStream_t s)
{
    hipLaunchKernelGGL(k_fv_columns, dim3(blocks_for(10LL * (a.n - 1) * (a.n - 1))), dim3(FV_BLOCK), 0, s, a,
                       columns);
}

void launch_fv_ghost(const int32_t* pairs, int cnt, int nr, int ncomp, double* f, hipStream_t s)
{
    if (cnt == 0) return;
    hipLaunchKernelGGL(k_fv_ghost, dim3(blocks_for((int64_t)cnt * (nr - 1))), dim3(FV_BLOCK), 0, s, pairs, cnt, nr,
                       ncomp, f);
}

void launch_fv_dt(const FvArgs& a, const double* vel, double kappa, double* part, double* out, hipStream_t s)
{
    const int nb = fv_dt_blocks(a.n, a.nr);
    hipLaunchKernelGGL(k_fv_dt, dim3(nb), dim3(FV_BLOCK), 0, s, a, vel, kappa, part);
    hipLaunchKernelGGL(k_min_finalize, dim3(1), dim3(256), 0, s, part, nb, out);
}

void launch_fct_predictor(const FvArgs& a, const double* vel, const double* T, double dt, double kappa,
                          const double* source, int subtract_div, double* TL, double* ad, hipStream_t s)
{
    hipLaunchKernelGGL(k_fct_predictor, dim3(blocks_for(fv_real_cells(a.n, a.nr))), dim3(FV_BLOCK), 0, s, a, vel, T,
                       dt, kappa, source, subtract_div, TL, ad, fv_size(a.n, a.nr));
}

void launch_fv_dirichlet(int n, int nr, double* f, int apply_cmb, double t_cmb, int apply_surf, double t_surf,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_fv_dirichlet, dim3(blocks_for(10LL * (n - 1) * (n - 1))), dim3(FV_BLOCK), 0, s, n, nr, f,
                       apply_cmb, t_cmb, apply_surf, t_surf);
}

void launch_fct_limiter(int n, int nr, const double* TL, const double* ad, double* Rp, double* Rm, hipStream_t s)
{
    hipLaunchKernelGGL(k_fct_limiter, dim3(blocks_for(fv_real_cells(n, nr))), dim3(FV_BLOCK), 0, s, n, nr, TL, ad,
                       fv_size(n, nr), Rp, Rm);
}

void launch_fct_correction(int n, int nr, const double* TL, const double* ad, const double* Rp, const double* Rm,
                           double* T, hipStream_t s)
{
    hipLaunchKernelGGL(k_fct_correction, dim3(blocks_for(fv_real_cells(n, nr))), dim3(FV_BLOCK), 0, s, n, nr, TL,
                       ad, fv_size(n, nr), Rp, Rm, T);
}

void launch_fv_lumped_scale(const double* D, double* u, int64_t N, hipStream_t s)
{
    hipLaunchKernelGGL(k_fv_lumped_scale, dim3(blocks_for(N)), dim3(FV_BLOCK), 0, s, D, u, N);
}

void launch_fe_to_fv(const FvArgs& a, const double* u, double* Tfv, hipStream_t s)
{
    hipLaunchKernelGGL(k_fe_to_fv, dim3(blocks_for(fv_real_cells(a.n, a.nr))), dim3(FV_BLOCK), 0, s, a, u, Tfv);
}

void launch_fv_to_fe_rhs(const FvArgs& a, const double* Tfv, double* b, hipStream_t s)
{
    hipLaunchKernelGGL(k_fv_to_fe_rhs, dim3(blocks_for(10LL * a.n * a.n * a.nr)), dim3(FV_BLOCK), 0, s, a, Tfv, b);
}

}  // namespace tn
