// Per-cell epsilon-div-div arithmetic shared by the matrix-free kernels
// (apply, diagonal, adjoint gradient).  Device-only header.
//
// Cell arithmetic (own formulation): with J = A diag(r_mid, r_mid, dr/2) every
// wedge shape gradient is a_i -/+ zz (a_i = l_i / r_mid, zz = z * 2/dr, from
// the per-wedge lateral geometry table).  Per lateral corner c of a hex cell
// S_c = x_bot + x_top, d_c = x_top - x_bot, e_c = eta_bot + eta_top; per wedge
//   grad u = U1 (x) a1 + U2 (x) a2 + D (x) zz   (U1 = S_1 - S_0, U2 = S_2 - S_0,
//                                               D = d_0 + d_1 + d_2)
//   sigma  = kwJ (2 sym(grad u) - 2/3 tr(grad u) I),  kwJ = (sum eta) kr |det A|
//   p_i = sigma a_i, q = sigma zz,  y_(i,bot) = p_i - q, y_(i,top) = p_i + q
// (one-point centroid quadrature, kwJ = mean of the six nodal viscosities
// times |det J|: epsilon_divdiv_kerngen.hpp:1055-1389).
#pragma once

#include "tn_internal.hpp"

namespace tn {
namespace cellops {

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T>
__device__ inline T dot3(const V3<T>& a, const V3<T>& b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Outward unit normal of a shell node: its lateral coordinates, which the mesh
// generator normalises to the unit sphere (normalising again would change
// them by at most an ulp, at the cost of a square root and a division in
// every boundary lane).
template <typename T>
__device__ inline V3<T> unit_normal(const T* coords, int64_t lat)
{
    return {coords[3 * lat], coords[3 * lat + 1], coords[3 * lat + 2]};
}

// Normal-tangential basis (reference local_basis_trafo_normal_tangential.hpp:34-79):
// t1 = normalised (e x n) with e the coordinate axis least aligned with n,
// t2 = normalised (n x t1).
template <typename T>
__device__ inline void nt_basis(const V3<T>& n, V3<T>& t1, V3<T>& t2)
{
    const T ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
    V3<T> e{0, 0, 0};
    if (ax <= ay && ax <= az)
        e.x = 1;
    else if (ay <= ax && ay <= az)
        e.y = 1;
    else
        e.z = 1;
    t1 = {e.y * n.z - e.z * n.y, e.z * n.x - e.x * n.z, e.x * n.y - e.y * n.x};
    T l = sqrt(dot3(t1, t1));
    t1 = {t1.x / l, t1.y / l, t1.z / l};
    t2 = {n.y * t1.z - n.z * t1.y, n.z * t1.x - n.x * t1.z, n.x * t1.y - n.y * t1.x};
    l = sqrt(dot3(t2, t2));
    t2 = {t2.x / l, t2.y / l, t2.z / l};
}

// per lateral corner of a cell: S = x_bot + x_top, d = x_top - x_bot, e = eta_bot + eta_top
template <typename T>
struct CornerIn {
    T S[3], d[3], e;
};

// raw corner values (diagonal mode)
template <typename T>
struct CornerRaw {
    T b[3], t[3], eb, et;
};

template <typename T>
__device__ __forceinline__ void wedge_products(const T U1[3], const T U2[3], const T D[3], T kwJ, const T a1[3],
                                               const T a2[3], const T zz[3], T p1[3], T p2[3], T q[3])
{
    T G[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) G[k][l] = U1[k] * a1[l] + U2[k] * a2[l] + D[k] * zz[l];
    const T tr = G[0][0] + G[1][1] + G[2][2];
    const T k2 = kwJ + kwJ;
    const T kt = (T(2.0 / 3.0) * kwJ) * tr;
    const T s00 = k2 * G[0][0] - kt, s11 = k2 * G[1][1] - kt, s22 = k2 * G[2][2] - kt;
    const T s01 = kwJ * (G[0][1] + G[1][0]), s02 = kwJ * (G[0][2] + G[2][0]), s12 = kwJ * (G[1][2] + G[2][1]);
    p1[0] = s00 * a1[0] + s01 * a1[1] + s02 * a1[2];
    p1[1] = s01 * a1[0] + s11 * a1[1] + s12 * a1[2];
    p1[2] = s02 * a1[0] + s12 * a1[1] + s22 * a1[2];
    p2[0] = s00 * a2[0] + s01 * a2[1] + s02 * a2[2];
    p2[1] = s01 * a2[0] + s11 * a2[1] + s12 * a2[2];
    p2[2] = s02 * a2[0] + s12 * a2[1] + s22 * a2[2];
    q[0] = s00 * zz[0] + s01 * zz[1] + s02 * zz[2];
    q[1] = s01 * zz[0] + s11 * zz[1] + s12 * zz[2];
    q[2] = s02 * zz[0] + s12 * zz[1] + s22 * zz[2];
}

// The same with the radial scalings applied around the wave-uniform geometry
// g (l1, l2, z in scalar registers): a1 = alpha l1, a2 = alpha l2,
// zz = beta z are never formed per lane, which keeps 9 values per wedge out
// of the vector registers (the scalings move onto U1, U2, D and the outputs).
template <typename T>
__device__ __forceinline__ void wedge_products_g(T U1[3], T U2[3], T D[3], T kwJ, const T* __restrict__ g, T alpha,
                                                 T beta, T p1[3], T p2[3], T q[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        U1[k] *= alpha;
        U2[k] *= alpha;
        D[k] *= beta;
    }
    T G[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) G[k][l] = U1[k] * g[l] + U2[k] * g[3 + l] + D[k] * g[6 + l];
    const T tr = G[0][0] + G[1][1] + G[2][2];
    const T k2 = kwJ + kwJ;
    const T kt = (T(2.0 / 3.0) * kwJ) * tr;
    const T s00 = k2 * G[0][0] - kt, s11 = k2 * G[1][1] - kt, s22 = k2 * G[2][2] - kt;
    const T s01 = kwJ * (G[0][1] + G[1][0]), s02 = kwJ * (G[0][2] + G[2][0]), s12 = kwJ * (G[1][2] + G[2][1]);
    p1[0] = alpha * (s00 * g[0] + s01 * g[1] + s02 * g[2]);
    p1[1] = alpha * (s01 * g[0] + s11 * g[1] + s12 * g[2]);
    p1[2] = alpha * (s02 * g[0] + s12 * g[1] + s22 * g[2]);
    p2[0] = alpha * (s00 * g[3] + s01 * g[4] + s02 * g[5]);
    p2[1] = alpha * (s01 * g[3] + s11 * g[4] + s12 * g[5]);
    p2[2] = alpha * (s02 * g[3] + s12 * g[4] + s22 * g[5]);
    q[0] = beta * (s00 * g[6] + s01 * g[7] + s02 * g[8]);
    q[1] = beta * (s01 * g[6] + s11 * g[7] + s12 * g[8]);
    q[2] = beta * (s02 * g[6] + s12 * g[7] + s22 * g[8]);
}

// Neumann cell (two wedges: corners 0,1,2 and 3,2,1 of the quad 00,10,01,11;
// kernel_helpers.hpp:302-327).  geo: the two wedges' lateral geometry.
template <typename T>
__device__ __forceinline__ void cell_neumann(const CornerIn<T> in[4], const T* __restrict__ geo, T alpha, T beta,
                                             T kr, T yb[4][3], T yt[4][3])
{
    T d12[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d12[k] = in[1].d[k] + in[2].d[k];
    const T e12 = in[1].e + in[2].e;
    T p1a[3], p2a[3], qa[3], p1b[3], p2b[3], qb[3];
    {
        const T* g = geo;
        T U1[3], U2[3], D[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            U1[k] = in[1].S[k] - in[0].S[k];
            U2[k] = in[2].S[k] - in[0].S[k];
            D[k] = in[0].d[k] + d12[k];
        }
        const T kwJ = (in[0].e + e12) * (kr * g[9]);
        wedge_products_g(U1, U2, D, kwJ, g, alpha, beta, p1a, p2a, qa);
    }
    {
        const T* g = geo + GEO_STRIDE;
        T U1[3], U2[3], D[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            U1[k] = in[2].S[k] - in[3].S[k];
            U2[k] = in[1].S[k] - in[3].S[k];
            D[k] = in[3].d[k] + d12[k];
        }
        const T kwJ = (in[3].e + e12) * (kr * g[9]);
        wedge_products_g(U1, U2, D, kwJ, g, alpha, beta, p1b, p2b, qb);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T t0 = p1a[k] + p2a[k];
        yb[0][k] = -t0 - qa[k];
        yt[0][k] = qa[k] - t0;
        const T t3 = p1b[k] + p2b[k];
        yb[3][k] = -t3 - qb[k];
        yt[3][k] = qb[k] - t3;
        const T Q = qa[k] + qb[k];
        const T P1 = p1a[k] + p2b[k];
        yb[1][k] = P1 - Q;
        yt[1][k] = P1 + Q;
        const T P2 = p2a[k] + p1b[k];
        yb[2][k] = P2 - Q;
        yt[2][k] = P2 + Q;
    }
}

// Neumann cell, diagonal only: y_n[d] = kwJ (|g|^2 + g_d^2 / 3) x_n[d]
template <typename T>
__device__ __forceinline__ void cell_neumann_diag(const CornerRaw<T> in[4], const T* __restrict__ geo, T alpha,
                                                  T beta, T kr, T yb[4][3], T yt[4][3])
{
    constexpr int WC[2][3] = {{0, 1, 2}, {3, 2, 1}};
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) yb[c][k] = yt[c][k] = T(0);
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const T* gw = geo + w * GEO_STRIDE;
        const int c0 = WC[w][0], c1 = WC[w][1], c2 = WC[w][2];
        const T kwJ = ((in[c0].eb + in[c1].eb + in[c2].eb) + (in[c0].et + in[c1].et + in[c2].et)) * (kr * gw[9]);
        T a[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a[1][k] = alpha * gw[k];
            a[2][k] = alpha * gw[3 + k];
            a[0][k] = -(a[1][k] + a[2][k]);
        }
        const T zz[3] = {beta * gw[6], beta * gw[7], beta * gw[8]};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = WC[w][i];
            T gb[3], gt[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                gb[k] = a[i][k] - zz[k];
                gt[k] = a[i][k] + zz[k];
            }
            const T nb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
            const T nt = gt[0] * gt[0] + gt[1] * gt[1] + gt[2] * gt[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                yb[c][k] += kwJ * in[c].b[k] * (nb + T(1.0 / 3.0) * gb[k] * gb[k]);
                yt[c][k] += kwJ * in[c].t[k] * (nt + T(1.0 / 3.0) * gt[k] * gt[k]);
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ void apply_trial(T v[3], int mode, const T* coords, int64_t lat)
{
    if (mode == TN_DIRICHLET) {
        v[0] = v[1] = v[2] = T(0);
    } else if (mode == TN_FREESLIP) {
        const V3<T> nn = unit_normal(coords, lat);
        const T un = nn.x * v[0] + nn.y * v[1] + nn.z * v[2];
        v[0] -= un * nn.x;
        v[1] -= un * nn.y;
        v[2] -= un * nn.z;
    }
}

}  // namespace cellops
}  // namespace tn
