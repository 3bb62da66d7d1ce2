// Stokes coupling and mass kernels on wedges.
//
//   B   (Divergence, divergence.hpp:117-323):  p_c += -int div(u_h) q_H
//   B^T (Gradient,   gradient.hpp:115-325):    u   += -int p_H div(v_h)
//   VectorMass (vector_mass.hpp:83-153), lumped KMass (kmass.hpp:97-185)
//
// Velocity lives on the fine mesh, pressure on the mesh one level coarser;
// each fine wedge integrates with the 6-point Felippa 3x2 rule and the coarse
// (P1 on the pressure mesh) shape functions evaluated on the fine wedge
// (integrands.hpp:103-243).  The coarse dofs a fine wedge couples to are
// those of coarse cell (x/2, y/2, r/2) in the node pattern of the FINE wedge
// index, exactly as the reference scatters them (kernel_helpers.hpp:396-424);
// for the two centre fine wedges this is the neighbouring coarse wedge.
//
// Instead of forming the 6x18 element matrix, the apply works per quadrature
// point: B uses div u_h(q) = sum_j grad N_j(q) . u_j, B^T uses
// p_H(q) = sum_i N^H_i(q) p_i.  Dirichlet layers drop their velocity dofs,
// free-slip layers project them tangentially (trial side for B, test side
// for B^T), which is the masked / rotated element matrix of the reference.
#include "tn_cell.hpp"
#include "tn_eps_march.hpp"
#include "tn_internal.hpp"
#include "tn_march.hpp"
#include "tn_stokes.hpp"

namespace tn {
namespace {

using cellops::dot3;
using cellops::V3;

__device__ constexpr int WOFF[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                         {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};
__device__ constexpr double Q6P[6][3] = {{0.6666666666666666, 0.1666666666666667, -0.5773502691896257},
                                        {0.1666666666666667, 0.6666666666666666, -0.5773502691896257},
                                        {0.1666666666666667, 0.1666666666666667, -0.5773502691896257},
                                        {0.6666666666666666, 0.1666666666666667, 0.5773502691896257},
                                        {0.1666666666666667, 0.6666666666666666, 0.5773502691896257},
                                        {0.1666666666666667, 0.1666666666666667, 0.5773502691896257}};
__device__ constexpr double Q6W = 0.1666666666666667;

__device__ inline double shape_lat(int k, double xi, double eta)
{
    const int m = k % 3;
    return m == 0 ? 1.0 - xi - eta : (m == 1 ? xi : eta);
}
__device__ inline double shape_rad(int k, double zeta) { return k < 3 ? 0.5 * (1 - zeta) : 0.5 * (1 + zeta); }

struct WedgeGeo {
    double P[3][3];  // lateral vertices (unit sphere)
    double r1, r2;
};

__device__ inline void load_wedge(const double* coords, const double* radii, int n, int d, int x, int y, int rc, int w,
                                  WedgeGeo& G)
{
    for (int k = 0; k < 3; ++k) {
        const double* c = coords + 3 * (((int64_t)d * n + x + WOFF[w][k][0]) * n + y + WOFF[w][k][1]);
        G.P[k][0] = c[0];
        G.P[k][1] = c[1];
        G.P[k][2] = c[2];
    }
    G.r1 = radii[rc];
    G.r2 = radii[rc + 1];
}

// ---------------------------------------------------------------------------
// B and B^T by per-wedge moments.
//
// A fine wedge with lateral unit-sphere vertices P0, P1, P2 (a = P1 - P0,
// b = P2 - P0) between radii r1 < r2 is x(xi, eta, zeta) = r(zeta) c(xi, eta)
// with c = P0 + xi a + eta b, r = r1 + g (1 + zeta), g = (r2 - r1) / 2.  Its
// Jacobian [r a, r b, g c] has det J = r^2 g D, D = det[a, b, P0], constant
// over the triangle, and for a shape N
//   |det J| grad N = sgn(D) [ r g (N_xi (v1 - xi nv) + N_eta (v3 - eta nv))
//                             + r^2 N_zeta nv ],
//   v1 = b x P0,  v3 = P0 x a,  nv = a x b.
// With the wedge shapes N_(m,t) = L_m(xi, eta) R_t(zeta) (L = 1 - xi - eta,
// xi, eta; R = (1 -/+ zeta) / 2) the divergence of u_h is, per node level t,
//   |det J| div u_h = sgn(D) sum_t [ r g R_t (alpha_t - xi bx_t - eta be_t)
//                                    + r^2 R_t' (sigma_t + xi bx_t + eta be_t) ]
//   du1 = u_1t - u_0t, du2 = u_2t - u_0t,
//   alpha_t = v1.du1 + v3.du2, sigma_t = nv.u_0t, bx_t = nv.du1, be_t = nv.du2.
// Against a coarse shape lat_il(xi, eta) rad_ir(zeta) (integrands.hpp:103-190)
// the integrand has lateral degree <= 2 and radial degree <= 3, which the
// 6-point Felippa rule of the reference (3-point triangle x 2-point Gauss)
// integrates exactly, so the quadrature sum is the moment formula
//   (B u)_i = -sgn(D) [ P0_ir M0_il + Px_ir Mx_il + Pe_ir Me_il ]
//   P0_ir = sum_t alpha_t A_(ir,t) + sigma_t C_(ir,t)
//   Px_ir = sum_t bx_t (C - A)_(ir,t),   Pe_ir = sum_t be_t (C - A)_(ir,t)
// with the lateral moments M0, Mx, Me = int lat_il (1, xi, eta) over the
// reference triangle and the radial moments A_(ir,t) = int rad_ir r g R_t,
// C_(ir,t) = R_t' int rad_ir r^2 over [-1, 1].  B^T is its transpose:
// with q0_ir = sum_il p_i M0_il (qx, qe likewise) and per level t
//   X_t = sum_ir A q0, Y_t = sum_ir C q0, Zx_t = sum_ir (C - A) qx,
//   Ze_t = sum_ir (C - A) qe,
// the node values are -sgn(D) times
//   (0, t): -X_t (v1 + v3) + (Y_t - Zx_t - Ze_t) nv
//   (1, t):  X_t v1 + Zx_t nv
//   (2, t):  X_t v3 + Ze_t nv.
// About 100 operations per fine wedge either way (the per-quadrature-point
// Jacobian inversion of the reference costs ~1700).

__device__ constexpr double GAUSS2 = 0.57735026918962576451;  // 1/sqrt(3)

// coarse lateral shape il on fine lateral wedge flw: alpha + beta xi + gamma eta
// (integrands.hpp:103-190; flw from kernel_helpers.hpp:600-608)
__host__ __device__ constexpr double coarse_lat_coef(int flw, int il, int k)
{
    constexpr double C[4][3][3] = {{{1.0, -0.5, -0.5}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}},
                                   {{0.5, -0.5, -0.5}, {0.5, 0.5, 0.0}, {0.0, 0.0, 0.5}},
                                   {{0.5, -0.5, -0.5}, {0.0, 0.5, 0.0}, {0.5, 0.0, 0.5}},
                                   {{0.0, 0.5, 0.5}, {0.5, -0.5, 0.0}, {0.5, 0.0, -0.5}}};
    return C[flw][il][k];
}

// int over the reference triangle of (alpha + beta xi + gamma eta) * {1, xi, eta}
struct LatMoments {
    double m[4][3][3];  // [flw][il][M0, Mx, Me]
};
__host__ __device__ constexpr LatMoments make_lat_moments()
{
    LatMoments r{};
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < 3; ++i) {
            const double al = coarse_lat_coef(f, i, 0), be = coarse_lat_coef(f, i, 1), ga = coarse_lat_coef(f, i, 2);
            r.m[f][i][0] = al / 2 + (be + ga) / 6;
            r.m[f][i][1] = al / 6 + be / 12 + ga / 24;
            r.m[f][i][2] = al / 6 + be / 24 + ga / 12;
        }
    return r;
}
__device__ constexpr LatMoments LATM = make_lat_moments();

__device__ inline int fine_lateral_wedge_idx(int xp, int yp, int w)
{
    // kernel_helpers.hpp:600-608 (xp, yp: parities of the fine quad indices)
    constexpr int ind[2][2][2] = {{{0, 1}, {2, 3}}, {{3, 2}, {1, 0}}};
    return ind[w][yp][xp];
}

template <typename T>
struct RadMoments {
    T A[2][2];  // [ir][t] int rad_ir r g R_t
    T C[2][2];  // [ir][t] R_t' int rad_ir r^2
};

// fine radial cell (r1, r2) in position frw (0 lower, 1 upper) of its coarse
// cell; scale 0 zeroes every moment (lanes past the last cell)
template <typename T>
__device__ inline RadMoments<T> radial_moments(double r1, double r2, int frw, double scale)
{
    const double g = 0.5 * (r2 - r1);
    double A[2][2] = {{0, 0}, {0, 0}}, rr[2] = {0, 0};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const double z = s ? GAUSS2 : -GAUSS2;
        const double r = r1 + g * (1 + z);
        const double R[2] = {0.5 * (1 - z), 0.5 * (1 + z)};
        const double rad[2] = {frw ? 0.25 * (1 - z) : 0.25 * (3 - z), frw ? 0.25 * (3 + z) : 0.25 * (1 + z)};
#pragma unroll
        for (int ir = 0; ir < 2; ++ir) {
            A[ir][0] += rad[ir] * r * g * R[0];
            A[ir][1] += rad[ir] * r * g * R[1];
            rr[ir] += rad[ir] * r * r;
        }
    }
    RadMoments<T> m;
#pragma unroll
    for (int ir = 0; ir < 2; ++ir) {
        m.A[ir][0] = (T)(scale * A[ir][0]);
        m.A[ir][1] = (T)(scale * A[ir][1]);
        m.C[ir][0] = (T)(-0.5 * scale * rr[ir]);
        m.C[ir][1] = (T)(0.5 * scale * rr[ir]);
    }
    return m;
}

template <typename T>
struct LatWedge {
    V3<T> v1, v3, nv;
    T sg;  // -sgn(D): the sign of B folded in
};

template <typename T>
__device__ inline V3<T> cross3(const V3<T>& a, const V3<T>& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <typename T>
__device__ inline LatWedge<T> lat_wedge(const V3<T>& P0, const V3<T>& P1, const V3<T>& P2)
{
    const V3<T> a{P1.x - P0.x, P1.y - P0.y, P1.z - P0.z}, b{P2.x - P0.x, P2.y - P0.y, P2.z - P0.z};
    LatWedge<T> L;
    L.v1 = cross3(b, P0);
    L.v3 = cross3(P0, a);
    L.nv = cross3(a, b);
    L.sg = dot3(a, L.v1) > T(0) ? T(-1) : T(1);
    return L;
}

template <typename T>
__device__ inline V3<T> load_coord(const double* coords, int64_t lat)
{
    return {(T)coords[3 * lat], (T)coords[3 * lat + 1], (T)coords[3 * lat + 2]};
}

// trial-side boundary treatment of a shell node value: Dirichlet drops it,
// free-slip keeps its tangential part (normal: the node's radial direction)
template <typename T>
__device__ inline void trial_bc(V3<T>& u, int mode, const V3<T>& P)
{
    if (mode == TN_DIRICHLET) {
        u = {T(0), T(0), T(0)};
    } else if (mode == TN_FREESLIP) {
        const V3<T>& nn = P;  // lateral coordinates are unit vectors
        const T un = dot3(nn, u);
        u = {u.x - un * nn.x, u.y - un * nn.y, u.z - un * nn.z};
    }
}

// (B u) contributions of one fine wedge to its six coarse dofs:
// out[il][ir] += ..., u0/u1/u2 at node levels t = 0, 1
template <typename T>
__device__ inline void div_wedge(const LatWedge<T>& L, const T (*M)[3], const RadMoments<T>& rm, const V3<T> u0[2],
                                 const V3<T> u1[2], const V3<T> u2[2], T out[3][2])
{
    T al[2], sgm[2], bx[2], be[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const V3<T> du1{u1[t].x - u0[t].x, u1[t].y - u0[t].y, u1[t].z - u0[t].z};
        const V3<T> du2{u2[t].x - u0[t].x, u2[t].y - u0[t].y, u2[t].z - u0[t].z};
        al[t] = dot3(L.v1, du1) + dot3(L.v3, du2);
        sgm[t] = dot3(L.nv, u0[t]);
        bx[t] = dot3(L.nv, du1);
        be[t] = dot3(L.nv, du2);
    }
#pragma unroll
    for (int ir = 0; ir < 2; ++ir) {
        const T ca0 = rm.C[ir][0] - rm.A[ir][0], ca1 = rm.C[ir][1] - rm.A[ir][1];
        const T P0 = L.sg * (al[0] * rm.A[ir][0] + sgm[0] * rm.C[ir][0] + al[1] * rm.A[ir][1] + sgm[1] * rm.C[ir][1]);
        const T Px = L.sg * (bx[0] * ca0 + bx[1] * ca1);
        const T Pe = L.sg * (be[0] * ca0 + be[1] * ca1);
#pragma unroll
        for (int il = 0; il < 3; ++il) out[il][ir] += P0 * M[il][0] + Px * M[il][1] + Pe * M[il][2];
    }
}

// (B^T p) contributions of one fine wedge to its six velocity nodes:
// p[il][ir] coarse values, y0/y1/y2 at node levels t = 0, 1 (accumulated)
template <typename T>
__device__ inline void grad_wedge(const LatWedge<T>& L, const T (*M)[3], const RadMoments<T>& rm, const T p[3][2],
                                  T y0[2][3], T y1[2][3], T y2[2][3])
{
    T q0[2], qx[2], qe[2];
#pragma unroll
    for (int ir = 0; ir < 2; ++ir) {
        q0[ir] = L.sg * (p[0][ir] * M[0][0] + p[1][ir] * M[1][0] + p[2][ir] * M[2][0]);
        qx[ir] = L.sg * (p[0][ir] * M[0][1] + p[1][ir] * M[1][1] + p[2][ir] * M[2][1]);
        qe[ir] = L.sg * (p[0][ir] * M[0][2] + p[1][ir] * M[1][2] + p[2][ir] * M[2][2]);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const T X = rm.A[0][t] * q0[0] + rm.A[1][t] * q0[1];
        const T Y = rm.C[0][t] * q0[0] + rm.C[1][t] * q0[1];
        const T ca0 = rm.C[0][t] - rm.A[0][t], ca1 = rm.C[1][t] - rm.A[1][t];
        const T Zx = ca0 * qx[0] + ca1 * qx[1];
        const T Ze = ca0 * qe[0] + ca1 * qe[1];
        const T s0 = Y - Zx - Ze;
        y0[t][0] += s0 * L.nv.x - X * (L.v1.x + L.v3.x);
        y0[t][1] += s0 * L.nv.y - X * (L.v1.y + L.v3.y);
        y0[t][2] += s0 * L.nv.z - X * (L.v1.z + L.v3.z);
        y1[t][0] += X * L.v1.x + Zx * L.nv.x;
        y1[t][1] += X * L.v1.y + Zx * L.nv.y;
        y1[t][2] += X * L.v1.z + Zx * L.nv.z;
        y2[t][0] += X * L.v3.x + Ze * L.nv.x;
        y2[t][1] += X * L.v3.y + Ze * L.nv.y;
        y2[t][2] += X * L.v3.z + Ze * L.nv.z;
    }
}

// Lateral corners of a quad: k = 0 (x, y), 1 (x+1, y), 2 (x, y+1), 3 (x+1, y+1);
// wedge w's vertices 0, 1, 2 in that numbering (WOFF[w][0..2]).
__device__ constexpr int WK[2][3] = {{0, 1, 2}, {3, 2, 1}};

// B: a wave per coarse lateral quad (X, Y) and chunk of 64 coarse radial
// cells (one per lane).  A lane sums the 16 fine wedges of its coarse cell
// into the cell's 8 corners; the shared radial level is merged by a lane
// shuffle, so a lane writes its lower 4 corner sums per cell and the chunk's
// top lane the upper 4 once per chunk.  k_div_gather then sums, per pressure
// node copy, the (at most 4) lateral quads' values in a fixed order: no
// atomics, the same result on every run.
template <typename T>
__global__ __launch_bounds__(256) void k_div(DivArgs<T> a)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nqc = a.nc - 1, ncrc = a.nrc - 1;
    const int nch = (ncrc + 63) >> 6;
    if (wave >= 10LL * nqc * nqc * nch) return;  // wave-uniform
    int64_t t = wave;
    const int h = (int)(t % nch);
    t /= nch;
    const int Y = (int)(t % nqc);
    t /= nqc;
    const int X = (int)(t % nqc);
    const int d = (int)(t / nqc);
    const int R = 64 * h + lane;
    const bool valid = R < ncrc;
    const int Rc = valid ? R : ncrc - 1;
    const int toplane = min(63, ncrc - 1 - 64 * h);
    const int nf = a.nf, nrf = a.nrf;
    const int64_t Nf = 10LL * nf * nf * nrf;
    const double sc = valid ? 1.0 : 0.0;
    const RadMoments<T> rm[2] = {radial_moments<T>(a.radii[2 * Rc], a.radii[2 * Rc + 1], 0, sc),
                                 radial_moments<T>(a.radii[2 * Rc + 1], a.radii[2 * Rc + 2], 1, sc)};
    const int bl = Rc == 0 ? a.bc_cmb : TN_NEUMANN;           // fine level 2R
    const int tl = Rc == ncrc - 1 ? a.bc_surf : TN_NEUMANN;  // fine level 2R + 2
    T acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    // one fine quad at a time (all four at once would exceed the register file in FP64)
#pragma unroll 1
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll 1
        for (int dx = 0; dx < 2; ++dx) {
            const int fx = 2 * X + dx, fy = 2 * Y + dy;
            const int64_t l00 = ((int64_t)d * nf + fx) * nf + fy;
            const int64_t latc[4] = {l00, l00 + nf, l00 + 1, l00 + nf + 1};
            V3<T> u[4][3];
            V3<T> P[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t ib = latc[k] * nrf + 2 * Rc;
#pragma unroll
                for (int l = 0; l < 3; ++l) u[k][l] = {a.u[ib + l], a.u[Nf + ib + l], a.u[2 * Nf + ib + l]};
                P[k] = load_coord<T>(a.coords, latc[k]);
            }
            if (bl != TN_NEUMANN || tl != TN_NEUMANN) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (bl != TN_NEUMANN) trial_bc(u[k][0], bl, P[k]);
                    if (tl != TN_NEUMANN) trial_bc(u[k][2], tl, P[k]);
                }
            }
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int k0 = WK[w][0], k1 = WK[w][1], k2 = WK[w][2];
                const LatWedge<T> L = lat_wedge(P[k0], P[k1], P[k2]);
                const int flw = fine_lateral_wedge_idx(dx, dy, w);
                T M[3][3];
#pragma unroll
                for (int il = 0; il < 3; ++il)
#pragma unroll
                    for (int j = 0; j < 3; ++j) M[il][j] = (T)LATM.m[flw][il][j];
                T out[3][2] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
                for (int frw = 0; frw < 2; ++frw) {
                    const V3<T> u0[2] = {u[k0][frw], u[k0][frw + 1]};
                    const V3<T> u1[2] = {u[k1][frw], u[k1][frw + 1]};
                    const V3<T> u2[2] = {u[k2][frw], u[k2][frw + 1]};
                    div_wedge(L, M, rm[frw], u0, u1, u2, out);
                }
                // coarse corners of wedge w: WOFF[w][il] in the k numbering
#pragma unroll
                for (int il = 0; il < 3; ++il) {
                    acc[WK[w][il]][0] += out[il][0];
                    acc[WK[w][il]][1] += out[il][1];
                }
            }
        }
    const int64_t ncol = 10LL * nqc * nqc, ncell = ncol * ncrc;
    const int64_t col = ((int64_t)d * nqc + X) * nqc + Y;
    T* ctop = a.csum + 4 * ncell;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T below = __shfl_up(acc[k][1], 1, 64);
        if (lane == 0) below = T(0);
        if (valid) a.csum[k * ncell + col * ncrc + R] = acc[k][0] + below;
        if (lane == toplane) ctop[(k * ncol + col) * nch + h] = acc[k][1];
    }
}

// Pressure node copy (d, X, Y, level): corner k of quad (X - (k & 1), Y - (k >> 1)).
template <typename T>
__global__ __launch_bounds__(256) void k_div_gather(DivArgs<T> a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = a.nc, nrc = a.nrc, nqc = nc - 1, ncrc = nrc - 1;
    if (tid >= 10LL * nc * nc * nrc) return;
    const int nch = (ncrc + 63) >> 6;
    const int64_t ncol = 10LL * nqc * nqc, ncell = ncol * ncrc;
    const T* ctop = a.csum + 4 * ncell;
    const int lev = (int)(tid % nrc);
    const int64_t lat = tid / nrc;
    const int Yn = (int)(lat % nc);
    const int Xn = (int)((lat / nc) % nc);
    const int d = (int)(lat / ((int64_t)nc * nc));
    T v = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int qx = Xn - (k & 1), qy = Yn - (k >> 1);
        if (qx < 0 || qy < 0 || qx > nqc - 1 || qy > nqc - 1) continue;
        const int64_t col = ((int64_t)d * nqc + qx) * nqc + qy;
        if (lev < ncrc) v += a.csum[k * ncell + col * ncrc + lev];
        // first level of a chunk > 0, or the top level: the previous chunk's top
        if ((lev > 0 && (lev & 63) == 0) || lev == ncrc) v += ctop[(k * ncol + col) * nch + ((lev - 1) >> 6)];
    }
    a.p[tid] = v;
}

// B^T: the atomic-free lateral march (tn_march.hpp) over fine cells: each
// fine cell evaluates its two wedges from the 8 coarse values of its coarse
// cell.  Every velocity node copy is written once (store or read-modify-
// write); Dirichlet shells are left out and free-slip shells projected
// tangentially on the summed value (the test-side masks and rotations of
// gradient.hpp:115-325).
template <typename T>
struct GradCell {
    const GradArgs<T> a;  // by value: a reference would put the kernel argument in scratch
    // the arrays, taken from restrict kernel parameters: loads at wave-uniform
    // addresses (the lateral coordinates) can then go through the scalar unit
    const double* coords;
    const T* p;
    T* u;
    int64_t N;
    RadMoments<T> rm;
    int R;
    __device__ GradCell(const GradArgs<T>& args, const double* c, const T* pp, T* uu)
        : a(args), coords(c), p(pp), u(uu), N(10LL * args.nf * args.nf * args.nrf)
    {
    }
    __device__ void begin_chunk(int, int cc, bool valid)
    {
        rm = radial_moments<T>(a.radii[cc], a.radii[cc + 1], cc & 1, valid ? 1.0 : 0.0);
        R = cc >> 1;
    }
    __device__ void cell(int d, int x, int qy, T yb[4][3], T yt[4][3])
    {
        const int n = a.nf, nc = a.nc, nrc = a.nrc;
        const int64_t c00 = ((int64_t)d * nc + (x >> 1)) * nc + (qy >> 1);
        const int64_t cl[4] = {c00, c00 + nc, c00 + 1, c00 + nc + 1};
        T pc[4][2];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            pc[k][0] = p[cl[k] * nrc + R];
            pc[k][1] = p[cl[k] * nrc + R + 1];
        }
        const int64_t lat00 = ((int64_t)d * n + x) * n + qy;
        const int64_t latc[4] = {lat00, lat00 + n, lat00 + 1, lat00 + n + 1};
        V3<T> P[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = load_coord<T>(coords, latc[k]);
#pragma unroll
        for (int w2 = 0; w2 < 2; ++w2) {
            const int k0 = WK[w2][0], k1 = WK[w2][1], k2 = WK[w2][2];
            const LatWedge<T> L = lat_wedge(P[k0], P[k1], P[k2]);
            const int flw = fine_lateral_wedge_idx(x & 1, qy & 1, w2);
            T M[3][3];
#pragma unroll
            for (int il = 0; il < 3; ++il)
#pragma unroll
                for (int j = 0; j < 3; ++j) M[il][j] = (T)LATM.m[flw][il][j];
            const T p[3][2] = {{pc[k0][0], pc[k0][1]}, {pc[k1][0], pc[k1][1]}, {pc[k2][0], pc[k2][1]}};
            T o0[2][3] = {{0, 0, 0}, {0, 0, 0}}, o1[2][3] = {{0, 0, 0}, {0, 0, 0}}, o2[2][3] = {{0, 0, 0}, {0, 0, 0}};
            grad_wedge(L, M, rm, p, o0, o1, o2);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                yb[k0][j] += o0[0][j];
                yt[k0][j] += o0[1][j];
                yb[k1][j] += o1[0][j];
                yt[k1][j] += o1[1][j];
                yb[k2][j] += o2[0][j];
                yt[k2][j] += o2[1][j];
            }
        }
    }
    __device__ void finish(int64_t lat, int level, V3<T> v)
    {
        const int mode = level == 0 ? a.bc_cmb : (level == a.nrf - 1 ? a.bc_surf : TN_NEUMANN);
        const int64_t idx = lat * a.nrf + level;
        if (mode == TN_DIRICHLET) {
            if (a.add) return;
            v = {T(0), T(0), T(0)};
        } else if (mode == TN_FREESLIP) {
            const V3<T> nn = load_coord<T>(coords, lat);  // unit vector
            const T vn = dot3(nn, v);
            v = {v.x - vn * nn.x, v.y - vn * nn.y, v.z - vn * nn.z};
        }
        if (a.add) {
            u[idx] += v.x;
            u[N + idx] += v.y;
            u[2 * N + idx] += v.z;
        } else {
            u[idx] = v.x;
            u[N + idx] = v.y;
            u[2 * N + idx] = v.z;
        }
    }
};

template <typename T, int W>
// (3 waves per SIMD: the FP64 build fits 168 VGPRs without spills)
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(3))) void k_grad_march(GradArgs<T> a, const double* __restrict__ coords,
                                                         const T* __restrict__ p, T* __restrict__ u)
{
    GradCell<T> op(a, coords, p, u);
    march::run<T, W>(op, a.nf, a.nrf, a.seg, a.guard);
}

// One row sum of a wedge's k-weighted mass matrix (kmass.hpp): vertex j of
// wedge w of hex cell (d, x, y, rc), the reference's 6-point quadrature.
template <typename T>
__device__ double kmass_wedge_row(const MassArgs<T>& a, int d, int x, int y, int rc, int w, int jrow)
{
    WedgeGeo G;
    load_wedge(a.coords, a.radii, a.n, d, x, y, rc, w, G);
    int64_t idx[6];
    for (int j = 0; j < 6; ++j)
        idx[j] = (((int64_t)d * a.n + x + WOFF[w][j][0]) * a.n + y + WOFF[w][j][1]) * a.nr + rc + WOFF[w][j][2];
    const double gr = 0.5 * (G.r2 - G.r1);
    double M[6][6] = {{0}};
    for (int q = 0; q < 6; ++q) {
        const double* qp = Q6P[q];
        double Jl[3][3];  // lateral Jacobian (kernel_helpers.hpp:153-176)
        for (int c = 0; c < 3; ++c) {
            Jl[c][0] = G.P[1][c] - G.P[0][c];
            Jl[c][1] = G.P[2][c] - G.P[0][c];
            Jl[c][2] = (1 - qp[0] - qp[1]) * G.P[0][c] + qp[0] * G.P[1][c] + qp[1] * G.P[2][c];
        }
        const double det = fabs(Jl[0][0] * (Jl[1][1] * Jl[2][2] - Jl[1][2] * Jl[2][1]) -
                                Jl[0][1] * (Jl[1][0] * Jl[2][2] - Jl[1][2] * Jl[2][0]) +
                                Jl[0][2] * (Jl[1][0] * Jl[2][1] - Jl[1][1] * Jl[2][0]));
        const double rr = G.r1 + 0.5 * (G.r2 - G.r1) * (1 + qp[2]);
        double N6[6];
        for (int j = 0; j < 6; ++j) N6[j] = shape_lat(j, qp[0], qp[1]) * shape_rad(j, qp[2]);
        double kq = 0.0;
        for (int j = 0; j < 6; ++j) kq += N6[j] * (double)a.k[idx[j]];
        const double f = Q6W * kq * rr * rr * gr * det;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) M[i][j] += f * N6[i] * N6[j];
    }
    double s = 0.0;
    for (int j = 0; j < 6; ++j) s += M[jrow][j];
    return s;
}

// lumped k-weighted mass (kmass.hpp): row sums of the k-weighted wedge mass
// matrices, as a gather (setup only): one thread per node copy adds the row
// sums of the wedges of its diamond that contain it, in a fixed order, so the
// result repeats bit for bit (the Schur-complement diagonal of the solves)
template <typename T>
__global__ __launch_bounds__(128) void k_kmass_lumped(MassArgs<T> a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n, nr = a.nr, ncr = nr - 1, ncl = n - 1;
    if (tid >= 10LL * n * n * nr) return;
    const int lev = (int)(tid % nr);
    int64_t lat = tid / nr;
    const int Y = (int)(lat % n);
    lat /= n;
    const int X = (int)(lat % n);
    const int d = (int)(lat / n);
    double acc = 0.0;
    for (int rc = lev - 1; rc <= lev; ++rc) {
        if (rc < 0 || rc >= ncr) continue;
        for (int qx = X - 1; qx <= X; ++qx) {
            if (qx < 0 || qx >= ncl) continue;
            for (int qy = Y - 1; qy <= Y; ++qy) {
                if (qy < 0 || qy >= ncl) continue;
                for (int w = 0; w < 2; ++w)
                    for (int j = 0; j < 6; ++j)
                        if (qx + WOFF[w][j][0] == X && qy + WOFF[w][j][1] == Y && rc + WOFF[w][j][2] == lev)
                            acc += kmass_wedge_row(a, d, qx, qy, rc, w, j);
            }
        }
    }
    a.y[tid] = (T)acc;
}

// VectorMass as a gather: one thread per node copy writes y_i = sum over the
// wedges of its diamond that contain it of (M_w x)_i; no atomics, so the
// result repeats bit for bit.  The reference's Felippa 3x2 rule factors into
// the 3-point triangle rule, exact for the lateral products (the P1 triangle
// mass (1 + delta_ij) / 24 times |det(V0, V1, V2)|), and the 2-point Gauss
// rule in zeta for rho^2 hat_a hat_b, which is not exact (degree 4) and is kept:
//   M_w[i][j] = |det_w| (1 + delta_ij) / 24 x h sum_{g = -+1/sqrt 3} rho(g)^2 hat_a(g) hat_b(g).
// Summed over the six wedges around a lateral node this is a 7-point lateral
// stencil (the node and its neighbours (x+-1, y), (x, y+-1), (x-1, y+1),
// (x+1, y-1)) times a 3-point radial stencil; |det_w| comes from the mesh's
// wedge table (geo[..][9]).
template <typename T>
__global__ __launch_bounds__(256) void k_mass_stencil(MassArgs<T> a)
{
    const int n = a.n, nr = a.nr, nc = n - 1;
    const int64_t N = 10LL * n * n * nr;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int r = (int)(t % nr);
    const int y = (int)((t / nr) % n);
    const int x = (int)((t / ((int64_t)nr * n)) % n);
    const int d = (int)(t / ((int64_t)nr * n * n));
    auto det = [&](int qx, int qy, int w) -> double {
        if (qx < 0 || qy < 0 || qx >= nc || qy >= nc) return 0.0;
        return a.geo[((((int64_t)d * nc + qx) * nc + qy) * 2 + w) * GEO_STRIDE + 9];
    };
    // the six wedges around (x, y): quad (x, y) w0; (x-1, y) w0, w1; (x, y-1) w0, w1; (x-1, y-1) w1
    const double a0 = det(x, y, 0), b0 = det(x - 1, y, 0), b1 = det(x - 1, y, 1);
    const double c0 = det(x, y - 1, 0), c1 = det(x, y - 1, 1), d1 = det(x - 1, y - 1, 1);
    constexpr double s = 1.0 / 24.0;
    constexpr int NB = 7;
    const int ox[NB] = {0, 1, 0, -1, 0, -1, 1}, oy[NB] = {0, 0, 1, 0, -1, 1, -1};
    const double wl[NB] = {2.0 * s * (a0 + b0 + b1 + c0 + c1 + d1), s * (a0 + c1), s * (a0 + b1),
                           s * (b0 + d1),                          s * (c0 + d1), s * (b0 + b1), s * (c0 + c1)};
    // radial stencil: the node's own level and the levels below / above
    constexpr double g = 0.57735026918962584;  // 1/sqrt(3)
    constexpr double hm = 0.5 * (1.0 + g), hp = 0.5 * (1.0 - g);  // bottom hat at zeta = -g, +g
    double cm = 0.0, c0r = 0.0, cp = 0.0;
    if (r > 0) {  // layer (r-1, r): the node is its top level
        const double r1 = a.radii[r - 1], h = 0.5 * (a.radii[r] - r1);
        const double rm = r1 + h * (1.0 - g), rp = r1 + h * (1.0 + g);
        c0r += h * (rm * rm * hp * hp + rp * rp * hm * hm);
        cm = h * (rm * rm + rp * rp) * hm * hp;
    }
    if (r < nr - 1) {  // layer (r, r+1): the node is its bottom level
        const double r1 = a.radii[r], h = 0.5 * (a.radii[r + 1] - r1);
        const double rm = r1 + h * (1.0 - g), rp = r1 + h * (1.0 + g);
        c0r += h * (rm * rm * hm * hm + rp * rp * hp * hp);
        cp = h * (rm * rm + rp * rp) * hm * hp;
    }
    double acc[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (wl[k] == 0.0) continue;  // a neighbour outside the diamond (no wedge holds it)
        const int64_t col = (((int64_t)d * n + x + ox[k]) * n + y + oy[k]) * nr + r;
        for (int c = 0; c < a.ncomp; ++c) {
            const T* xc = a.x + c * N + col;
            double v = c0r * (double)xc[0];
            if (r > 0) v += cm * (double)xc[-1];
            if (r < nr - 1) v += cp * (double)xc[1];
            acc[c] += wl[k] * v;
        }
    }
    for (int c = 0; c < a.ncomp; ++c) a.y[c * N + t] = (T)acc[c];
}

}  // namespace

template <typename T>
void launch_div(const DivArgs<T>& a, hipStream_t s)
{
    TN_REQUIRE(a.nf == 2 * a.nc - 1 && a.nrf == 2 * a.nrc - 1 && a.nrc >= 2 && a.nc >= 2,
               "B: velocity mesh must be the pressure mesh refined once");
    TN_REQUIRE(a.csum, "B: corner-sum scratch missing");
    const int64_t waves = 10LL * (a.nc - 1) * (a.nc - 1) * ((a.nrc - 1 + 63) / 64);
    hipLaunchKernelGGL(k_div<T>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, a);
    TN_HIP(hipGetLastError());
    const int64_t Nc = 10LL * a.nc * a.nc * a.nrc;
    hipLaunchKernelGGL(k_div_gather<T>, dim3((unsigned)((Nc + 255) / 256)), dim3(256), 0, s, a);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_grad(const GradArgs<T>& a, hipStream_t s)
{
    TN_REQUIRE(a.nf == 2 * a.nc - 1 && a.nrf == 2 * a.nrc - 1 && a.nrc >= 2 && a.nc >= 2,
               "B^T: velocity mesh must be the pressure mesh refined once");
    TN_REQUIRE(a.seg >= 1 && a.seg <= MARCH_SEG_MAX, "B^T: segment length out of range");
    constexpr int W = grad_march_waves<T>();
    hipLaunchKernelGGL((k_grad_march<T, W>), dim3((unsigned)march::blocks_for(a.nf, a.seg, W)), dim3(64 * W), 0, s,
                       a, a.coords, a.p, a.u);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_mass(const MassArgs<T>& a, bool lumped_k, hipStream_t s)
{
    if (lumped_k) {
        const int64_t N = 10LL * a.n * a.n * a.nr;
        hipLaunchKernelGGL((k_kmass_lumped<T>), dim3((unsigned)((N + 127) / 128)), dim3(128), 0, s, a);
    } else {
        TN_REQUIRE(a.ncomp >= 1 && a.ncomp <= 3, "mass: ncomp must be 1..3");
        const int64_t N = 10LL * a.n * a.n * a.nr;
        TN_REQUIRE(a.geo, "mass: wedge table missing");
        hipLaunchKernelGGL((k_mass_stencil<T>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, a);
    }
    TN_HIP(hipGetLastError());
}

template void launch_div<double>(const DivArgs<double>&, hipStream_t);
template void launch_div<float>(const DivArgs<float>&, hipStream_t);
template void launch_grad<double>(const GradArgs<double>&, hipStream_t);
template void launch_grad<float>(const GradArgs<float>&, hipStream_t);
template void launch_mass<double>(const MassArgs<double>&, bool, hipStream_t);
template void launch_mass<float>(const MassArgs<float>&, bool, hipStream_t);

}  // namespace tn
