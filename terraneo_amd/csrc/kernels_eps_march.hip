// epsilon-div-div apply without global atomics: lateral march, CDNA4.
//
// Decomposition.  A block of W waves owns W-1 lateral node rows of one
// diamond and one segment of node columns.  Wave w handles the quad strip
// x = x0 - 1 + w and marches over the quads of that strip in y; its 64
// lanes are radial cells (P passes of 64 cover up to 128 radial cells), so
// every global load and store is radially contiguous.  A quad's
// contributions to its four lateral corners are routed as follows:
//   corner (x,   y)   own node row, column y   -> complete after this step
//   corner (x,   y+1) own node row, column y+1 -> carried in registers
//   corner (x+1, y)   next wave's row          -> LDS, read by wave w+1
//   corner (x+1, y+1)                          -> carried, sent next step
// and radially adjacent cells share their node level through a lane shuffle.
// Every stored node is written exactly once with a plain store (Replace) or
// one read-modify-write (Add); the strip at a block edge is computed twice
// (once as the halo wave 0 of the next block), the column before a segment
// once more -- (W/(W-1)) * (S+1)/S redundant work instead of atomics.
//
// Boundary conditions are algebraic at the two boundary node levels, so the
// per-cell arithmetic is the Neumann formula for every lane:
//   Dirichlet  A x = P_I A_N P_I x + diag(A_N) x   on boundary nodes
//   free-slip  A x = P A_N P x + (n^T B_b n)(n . x) n
// with P_I zeroing and P projecting the boundary-node values, and B_b the
// 3x3 diagonal block of A_N at boundary node b (precomputed once per eta,
// see k_boundary_blocks).  This is the same matrix the reference assembles
// through its masks and rotations (epsilon_divdiv_kerngen.hpp:823-1051,
// 1421-1960).  The fs/fs rotation penalty (kerngen.hpp:583-607) is fused
// into the store.
#include "tn_eps_march.hpp"

namespace tn {
namespace {

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T>
__device__ inline T dot3(const V3<T>& a, const V3<T>& b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Neumann cell: two wedges, lean formulation (see kernels_eps.hip).
// s*/e*: source and viscosity at corners 00,10,01,11 (bottom b, top t).
template <typename T>
__device__ __forceinline__ void cell_neumann(const V3<T> sb[4], const V3<T> st[4], const T eb[4], const T et[4],
                                             const T* __restrict__ geo, T alpha, T beta, T kr, V3<T> yb[4],
                                             V3<T> yt[4])
{
    constexpr int WC[2][3] = {{0, 1, 2}, {3, 2, 1}};
#pragma unroll
    for (int c = 0; c < 4; ++c) yb[c] = yt[c] = V3<T>{0, 0, 0};
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const T* gw = geo + w * GEO_STRIDE;
        const int c0 = WC[w][0], c1 = WC[w][1], c2 = WC[w][2];
        const T kwJ = ((eb[c0] + eb[c1] + eb[c2]) + (et[c0] + et[c1] + et[c2])) * (kr * gw[9]);
        const V3<T> a1{alpha * gw[0], alpha * gw[1], alpha * gw[2]};
        const V3<T> a2{alpha * gw[3], alpha * gw[4], alpha * gw[5]};
        const V3<T> zz{beta * gw[6], beta * gw[7], beta * gw[8]};
        const V3<T> S0{sb[c0].x + st[c0].x, sb[c0].y + st[c0].y, sb[c0].z + st[c0].z};
        const V3<T> S1{sb[c1].x + st[c1].x - S0.x, sb[c1].y + st[c1].y - S0.y, sb[c1].z + st[c1].z - S0.z};
        const V3<T> S2{sb[c2].x + st[c2].x - S0.x, sb[c2].y + st[c2].y - S0.y, sb[c2].z + st[c2].z - S0.z};
        const V3<T> D{(st[c0].x + st[c1].x + st[c2].x) - (sb[c0].x + sb[c1].x + sb[c2].x),
                      (st[c0].y + st[c1].y + st[c2].y) - (sb[c0].y + sb[c1].y + sb[c2].y),
                      (st[c0].z + st[c1].z + st[c2].z) - (sb[c0].z + sb[c1].z + sb[c2].z)};
        const T G00 = S1.x * a1.x + S2.x * a2.x + D.x * zz.x;
        const T G01 = S1.x * a1.y + S2.x * a2.y + D.x * zz.y;
        const T G02 = S1.x * a1.z + S2.x * a2.z + D.x * zz.z;
        const T G10 = S1.y * a1.x + S2.y * a2.x + D.y * zz.x;
        const T G11 = S1.y * a1.y + S2.y * a2.y + D.y * zz.y;
        const T G12 = S1.y * a1.z + S2.y * a2.z + D.y * zz.z;
        const T G20 = S1.z * a1.x + S2.z * a2.x + D.z * zz.x;
        const T G21 = S1.z * a1.y + S2.z * a2.y + D.z * zz.y;
        const T G22 = S1.z * a1.z + S2.z * a2.z + D.z * zz.z;
        const T dd = T(2.0 / 3.0) * (G00 + G11 + G22);
        const T s00 = kwJ * (T(2) * G00 - dd), s11 = kwJ * (T(2) * G11 - dd), s22 = kwJ * (T(2) * G22 - dd);
        const T s01 = kwJ * (G01 + G10), s02 = kwJ * (G02 + G20), s12 = kwJ * (G12 + G21);
        const V3<T> p1{s00 * a1.x + s01 * a1.y + s02 * a1.z, s01 * a1.x + s11 * a1.y + s12 * a1.z,
                       s02 * a1.x + s12 * a1.y + s22 * a1.z};
        const V3<T> p2{s00 * a2.x + s01 * a2.y + s02 * a2.z, s01 * a2.x + s11 * a2.y + s12 * a2.z,
                       s02 * a2.x + s12 * a2.y + s22 * a2.z};
        const V3<T> q{s00 * zz.x + s01 * zz.y + s02 * zz.z, s01 * zz.x + s11 * zz.y + s12 * zz.z,
                      s02 * zz.x + s12 * zz.y + s22 * zz.z};
        const V3<T> p0{-(p1.x + p2.x), -(p1.y + p2.y), -(p1.z + p2.z)};
        yb[c0].x += p0.x - q.x; yb[c0].y += p0.y - q.y; yb[c0].z += p0.z - q.z;
        yt[c0].x += p0.x + q.x; yt[c0].y += p0.y + q.y; yt[c0].z += p0.z + q.z;
        yb[c1].x += p1.x - q.x; yb[c1].y += p1.y - q.y; yb[c1].z += p1.z - q.z;
        yt[c1].x += p1.x + q.x; yt[c1].y += p1.y + q.y; yt[c1].z += p1.z + q.z;
        yb[c2].x += p2.x - q.x; yb[c2].y += p2.y - q.y; yb[c2].z += p2.z - q.z;
        yt[c2].x += p2.x + q.x; yt[c2].y += p2.y + q.y; yt[c2].z += p2.z + q.z;
    }
}

// Neumann cell, diagonal only: y_n[d] = kwJ (|g|^2 + g_d^2 / 3) s_n[d]
template <typename T>
__device__ __forceinline__ void cell_neumann_diag(const V3<T> sb[4], const V3<T> st[4], const T eb[4], const T et[4],
                                                  const T* __restrict__ geo, T alpha, T beta, T kr, V3<T> yb[4],
                                                  V3<T> yt[4])
{
    constexpr int WC[2][3] = {{0, 1, 2}, {3, 2, 1}};
#pragma unroll
    for (int c = 0; c < 4; ++c) yb[c] = yt[c] = V3<T>{0, 0, 0};
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const T* gw = geo + w * GEO_STRIDE;
        const int c0 = WC[w][0], c1 = WC[w][1], c2 = WC[w][2];
        const T kwJ = ((eb[c0] + eb[c1] + eb[c2]) + (et[c0] + et[c1] + et[c2])) * (kr * gw[9]);
        V3<T> a[3];
        a[1] = {alpha * gw[0], alpha * gw[1], alpha * gw[2]};
        a[2] = {alpha * gw[3], alpha * gw[4], alpha * gw[5]};
        a[0] = {-(a[1].x + a[2].x), -(a[1].y + a[2].y), -(a[1].z + a[2].z)};
        const V3<T> zz{beta * gw[6], beta * gw[7], beta * gw[8]};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = WC[w][i];
            const V3<T> gbt{a[i].x - zz.x, a[i].y - zz.y, a[i].z - zz.z};
            const V3<T> gtp{a[i].x + zz.x, a[i].y + zz.y, a[i].z + zz.z};
            const T gb = dot3(gbt, gbt), gt = dot3(gtp, gtp);
            yb[c].x += kwJ * sb[c].x * (gb + T(1.0 / 3.0) * gbt.x * gbt.x);
            yb[c].y += kwJ * sb[c].y * (gb + T(1.0 / 3.0) * gbt.y * gbt.y);
            yb[c].z += kwJ * sb[c].z * (gb + T(1.0 / 3.0) * gbt.z * gbt.z);
            yt[c].x += kwJ * st[c].x * (gt + T(1.0 / 3.0) * gtp.x * gtp.x);
            yt[c].y += kwJ * st[c].y * (gt + T(1.0 / 3.0) * gtp.y * gtp.y);
            yt[c].z += kwJ * st[c].z * (gt + T(1.0 / 3.0) * gtp.z * gtp.z);
        }
    }
}

template <typename T>
__device__ inline V3<T> unit_normal(const T* coords, int64_t lat)
{
    const T cx = coords[3 * lat], cy = coords[3 * lat + 1], cz = coords[3 * lat + 2];
    const T inv = T(1) / sqrt(cx * cx + cy * cy + cz * cz);
    return {cx * inv, cy * inv, cz * inv};
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

// source value at a boundary node as the trial side sees it
template <typename T>
__device__ inline V3<T> trial_value(const T* __restrict__ src, int64_t N, int64_t idx, int mode, const T* coords,
                                    int64_t lat)
{
    V3<T> s{src[idx], src[N + idx], src[2 * N + idx]};
    if (mode == TN_DIRICHLET) {
        s = V3<T>{0, 0, 0};
    } else if (mode == TN_FREESLIP) {
        const V3<T> nn = unit_normal(coords, lat);
        const T un = dot3(nn, s);
        s = {s.x - un * nn.x, s.y - un * nn.y, s.z - un * nn.z};
    }
    return s;
}

// Final value of a boundary node copy from its Neumann-formula value y (of the
// zeroed / projected source).  blk: the copy's three quadratic forms of the
// 3x3 diagonal block B of the Neumann operator, e^T B e for e = x, y, z axes
// (Dirichlet) or e = n, t1, t2 (free-slip).  Linear in y and blk, so summing
// the copies afterwards gives the full node's value.
template <typename T, bool DIAG>
__device__ inline V3<T> test_value(V3<T> y, const T* __restrict__ src, int64_t N, int64_t idx, int mode,
                                   const T* coords, int64_t lat, const T* __restrict__ blk)
{
    const V3<T> x{src[idx], src[N + idx], src[2 * N + idx]};
    if (mode == TN_DIRICHLET) {
        // DIAG: the Neumann diagonal of the raw source is already diag(B) x
        if constexpr (DIAG) return y;
        return {blk[0] * x.x, blk[1] * x.y, blk[2] * x.z};
    }
    // free-slip
    const V3<T> nn = unit_normal(coords, lat);
    if constexpr (DIAG) {
        // R^T diag(R B R^T) R x, rows of R = (n, t1, t2)
        V3<T> t1, t2;
        nt_basis(nn, t1, t2);
        const T v0 = blk[0] * dot3(nn, x), v1 = blk[1] * dot3(t1, x), v2 = blk[2] * dot3(t2, x);
        return {nn.x * v0 + t1.x * v1 + t2.x * v2, nn.y * v0 + t1.y * v1 + t2.y * v2,
                nn.z * v0 + t1.z * v1 + t2.z * v2};
    } else {
        // P y + (n^T B n)(n . x) n
        const T corr = blk[0] * dot3(nn, x) - dot3(nn, y);
        return {y.x + corr * nn.x, y.y + corr * nn.y, y.z + corr * nn.z};
    }
}

template <typename T>
__device__ __forceinline__ T shfl_up1(T v)
{
    return __shfl_up(v, 1, 64);
}
template <typename T>
__device__ __forceinline__ T bcast(T v, int lane)
{
    return __shfl(v, lane, 64);
}

template <typename T, int P, bool DIAG>
__global__ __launch_bounds__(64 * MARCH_WAVES) void k_eps_march(MarchArgs<T> a)
{
    constexpr int W = MARCH_WAVES, OWN = W - 1, L = 64 * P;
    __shared__ T ebuf[2][W][3][L + 1];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n = a.n, nr = a.nr, ncr = nr - 1, ncl = n - 1;
    const int nxb = (n + OWN - 1) / OWN, nys = (n + a.seg - 1) / a.seg;
    int b = blockIdx.x;
    const int ys = b % nys;
    b /= nys;
    const int xb = b % nxb;
    const int d = b / nxb;
    const int x = xb * OWN - 1 + w;
    const bool strip_ok = x >= 0 && x <= n - 2;
    const bool row_own = w >= 1 && x <= n - 1;
    const int y0 = ys * a.seg, y1 = min(y0 + a.seg, n);
    const int64_t N = 10LL * n * n * nr;
    const int bmode = a.bc_cmb, tmode = a.bc_surf;  // of the two boundary node levels
    // fs/fs rotation penalty: y += gamma x p, gamma = S beta (double, as the separate penalty pass)
    double gam[3] = {0, 0, 0};
    if (a.pen) {
        const double* bt = a.pen;
        const double* S = a.pen + 4;
        gam[0] = S[0] * bt[0] + S[1] * bt[1] + S[2] * bt[2];
        gam[1] = S[3] * bt[0] + S[4] * bt[1] + S[5] * bt[2];
        gam[2] = S[6] * bt[0] + S[7] * bt[1] + S[8] * bt[2];
    }

    T c01[P][3], c11[P][3], c01t[3], c11t[3];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k) c01[p][k] = c11[p][k] = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) c01t[k] = c11t[k] = T(0);

    for (int qy = y0 - 1; qy <= y1 - 1; ++qy) {
        const bool quad_ok = strip_ok && qy >= 0 && qy <= n - 2;
        T part[P][3], partt[3];
        T (*E)[L + 1] = ebuf[qy & 1][w];
        T geo[2 * GEO_STRIDE];
        if (quad_ok) {
            const T* gsrc = a.geo + ((((int64_t)d * ncl + x) * ncl + qy) * 2) * GEO_STRIDE;
#pragma unroll
            for (int i = 0; i < 2 * GEO_STRIDE; ++i) geo[i] = gsrc[i];
        }
        const int64_t lat00 = ((int64_t)d * n + x) * n + qy;
        const int64_t latc[4] = {lat00, lat00 + n, lat00 + 1, lat00 + n + 1};
        T carry[4][3];  // top contribution of lane 63 of the previous pass
#pragma unroll
        for (int k = 0; k < 4; ++k) carry[k][0] = carry[k][1] = carry[k][2] = T(0);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int c = 64 * p + lane;
            const bool valid = c < ncr;
            V3<T> yb[4], yt[4];
            if (quad_ok && valid) {
                V3<T> sb[4], st[4];
                T eb[4], et[4];
                const int blev = c == 0 ? bmode : TN_NEUMANN;
                const int tlev = c == ncr - 1 ? tmode : TN_NEUMANN;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int64_t ib = latc[k] * nr + c;
                    if constexpr (DIAG) {
                        sb[k] = {a.src[ib], a.src[N + ib], a.src[2 * N + ib]};
                        st[k] = {a.src[ib + 1], a.src[N + ib + 1], a.src[2 * N + ib + 1]};
                    } else {
                        sb[k] = trial_value(a.src, N, ib, blev, a.coords, latc[k]);
                        st[k] = trial_value(a.src, N, ib + 1, tlev, a.coords, latc[k]);
                    }
                    eb[k] = a.eta[ib];
                    et[k] = a.eta[ib + 1];
                }
                const T r0 = a.radii[c], r1 = a.radii[c + 1];
                const T half_dr = T(0.5) * (r1 - r0), r_mid = T(0.5) * (r0 + r1);
                const T kr = r_mid * r_mid * half_dr * T(1.0 / 6.0);
                if constexpr (DIAG)
                    cell_neumann_diag<T>(sb, st, eb, et, geo, T(1) / r_mid, T(1) / half_dr, kr, yb, yt);
                else
                    cell_neumann<T>(sb, st, eb, et, geo, T(1) / r_mid, T(1) / half_dr, kr, yb, yt);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) yb[k] = yt[k] = V3<T>{0, 0, 0};
            }
            // radial merge: node level c = bottom of cell c + top of cell c-1
            T M[4][3];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T t[3] = {yt[k].x, yt[k].y, yt[k].z};
                const T bb[3] = {yb[k].x, yb[k].y, yb[k].z};
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    T up = shfl_up1(t[j]);
                    if (lane == 0) up = carry[k][j];
                    M[k][j] = bb[j] + up;
                    carry[k][j] = bcast(t[j], 63);
                }
            }
            // route corners
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                part[p][j] = M[0][j] + c01[p][j];
                c01[p][j] = M[2][j];
                if (valid) E[j][c] = M[1][j] + c11[p][j];
                c11[p][j] = M[3][j];
            }
            if (valid && c == ncr - 1) {  // top node level ncr: only the top of the last cell
                const T t0[3] = {yt[0].x, yt[0].y, yt[0].z}, t1[3] = {yt[1].x, yt[1].y, yt[1].z};
                const T t2[3] = {yt[2].x, yt[2].y, yt[2].z}, t3[3] = {yt[3].x, yt[3].y, yt[3].z};
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    partt[j] = t0[j] + c01t[j];
                    c01t[j] = t2[j];
                    E[j][L] = t1[j] + c11t[j];
                    c11t[j] = t3[j];
                }
            }
        }
        __syncthreads();
        if (row_own && qy >= y0) {
            const T (*Ein)[L + 1] = ebuf[qy & 1][w - 1];
            const int64_t lat = ((int64_t)d * n + x) * n + qy;
            const T* blk0 = a.blk ? a.blk + lat * 3 : nullptr;                   // CMB shell
            const T* blk1 = a.blk ? a.blk + (10LL * n * n + lat) * 3 : nullptr;  // surface shell
            double px = 0, py = 0, pz = 0;
            if (a.pen) {
                const double cx = a.coords_d[3 * lat], cy = a.coords_d[3 * lat + 1], cz = a.coords_d[3 * lat + 2];
                px = cx, py = cy, pz = cz;
            }
#pragma unroll
            for (int p = 0; p <= P; ++p) {
                int c;
                V3<T> v;
                if (p < P) {
                    c = 64 * p + lane;
                    if (c >= ncr) continue;
                    v = {part[p][0] + Ein[0][c], part[p][1] + Ein[1][c], part[p][2] + Ein[2][c]};
                } else {
                    c = ncr;
                    if (lane != ((ncr - 1) & 63)) continue;
                    v = {partt[0] + Ein[0][L], partt[1] + Ein[1][L], partt[2] + Ein[2][L]};
                }
                const int64_t idx = lat * nr + c;
                if (c == 0 && bmode != TN_NEUMANN)
                    v = test_value<T, DIAG>(v, a.src, N, idx, bmode, a.coords, lat, blk0);
                else if (c == ncr && tmode != TN_NEUMANN)
                    v = test_value<T, DIAG>(v, a.src, N, idx, tmode, a.coords, lat, blk1);
                if (a.pen) {
                    const double r = a.radii_d[c];
                    const double qx = px * r, qy_ = py * r, qz = pz * r;
                    v.x += (T)(gam[1] * qz - gam[2] * qy_);
                    v.y += (T)(gam[2] * qx - gam[0] * qz);
                    v.z += (T)(gam[0] * qy_ - gam[1] * qx);
                }
                if (a.add) {
                    a.dst[idx] += v.x;
                    a.dst[N + idx] += v.y;
                    a.dst[2 * N + idx] += v.z;
                } else {
                    a.dst[idx] = v.x;
                    a.dst[N + idx] = v.y;
                    a.dst[2 * N + idx] = v.z;
                }
            }
        }
    }
}

// Quadratic forms of the 3x3 diagonal block of the Neumann operator at every
// boundary node copy, from the copy's own diamond's cells (gather over the up
// to four adjacent quads, no atomics):
//   B = sum_wedges kwJ (|g|^2 I + g g^T / 3),  blk = (e^T B e) for the basis of
//   the shell's condition (axes for Dirichlet, n/t1/t2 for free-slip).
template <typename T>
__global__ void k_boundary_blocks(Geom<T> g, const T* __restrict__ eta, int bc_cmb, int bc_surf, T* __restrict__ blk)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = g.n, ncl = n - 1, ncr = g.nr - 1;
    const int64_t nlat = 10LL * n * n;
    if (tid >= 2 * nlat) return;
    const int shell = (int)(tid / nlat);
    const int mode = shell == 0 ? bc_cmb : bc_surf;
    if (mode == TN_NEUMANN) return;
    const int64_t lat = tid % nlat;
    const int y = (int)(lat % n), x = (int)((lat / n) % n), d = (int)(lat / ((int64_t)n * n));
    const int rc = shell == 0 ? 0 : ncr - 1;
    const T r0 = g.radii[rc], r1 = g.radii[rc + 1];
    const T half_dr = T(0.5) * (r1 - r0), r_mid = T(0.5) * (r0 + r1);
    const T alpha = T(1) / r_mid, beta = T(shell == 0 ? -1 : 1) / half_dr;
    const T kr = r_mid * r_mid * half_dr * T(1.0 / 6.0);
    V3<T> e[3];
    if (mode == TN_FREESLIP) {
        e[0] = unit_normal(g.coords, lat);
        nt_basis(e[0], e[1], e[2]);
    } else {
        e[0] = {1, 0, 0};
        e[1] = {0, 1, 0};
        e[2] = {0, 0, 1};
    }
    constexpr int WC[2][3] = {{0, 1, 2}, {3, 2, 1}};
    T acc[3] = {0, 0, 0};
    for (int qx = x - 1; qx <= x; ++qx) {
        for (int qy = y - 1; qy <= y; ++qy) {
            if (qx < 0 || qy < 0 || qx > ncl - 1 || qy > ncl - 1) continue;
            const int k = (x - qx) + 2 * (y - qy);
            const int64_t l00 = ((int64_t)d * n + qx) * n + qy;
            const int64_t latc[4] = {l00, l00 + n, l00 + 1, l00 + n + 1};
            const T* geo = g.geo + ((((int64_t)d * ncl + qx) * ncl + qy) * 2) * GEO_STRIDE;
            for (int w = 0; w < 2; ++w) {
                int i = -1;
                for (int j = 0; j < 3; ++j)
                    if (WC[w][j] == k) i = j;
                if (i < 0) continue;
                const T* gw = geo + w * GEO_STRIDE;
                T ks = 0;
                for (int j = 0; j < 3; ++j) {
                    const int64_t ib = latc[WC[w][j]] * g.nr + rc;
                    ks += eta[ib];
                }
                T kt = 0;
                for (int j = 0; j < 3; ++j) {
                    const int64_t ib = latc[WC[w][j]] * g.nr + rc + 1;
                    kt += eta[ib];
                }
                const T kwJ = (ks + kt) * (kr * gw[9]);
                V3<T> gr;
                const T a1[3] = {alpha * gw[0], alpha * gw[1], alpha * gw[2]};
                const T a2[3] = {alpha * gw[3], alpha * gw[4], alpha * gw[5]};
                const T ai[3] = {i == 0 ? -(a1[0] + a2[0]) : (i == 1 ? a1[0] : a2[0]),
                                 i == 0 ? -(a1[1] + a2[1]) : (i == 1 ? a1[1] : a2[1]),
                                 i == 0 ? -(a1[2] + a2[2]) : (i == 1 ? a1[2] : a2[2])};
                gr = {ai[0] + beta * gw[6], ai[1] + beta * gw[7], ai[2] + beta * gw[8]};
                const T gg = dot3(gr, gr);
                for (int j = 0; j < 3; ++j) {
                    const T eg = dot3(e[j], gr);
                    acc[j] += kwJ * (gg + T(1.0 / 3.0) * eg * eg);
                }
            }
        }
    }
    T* o = blk + ((int64_t)shell * nlat + lat) * 3;
    o[0] = acc[0];
    o[1] = acc[1];
    o[2] = acc[2];
}

}  // namespace

template <typename T>
void launch_boundary_blocks(const Geom<T>& g, const T* eta, int bc_cmb, int bc_surf, T* blk, hipStream_t s)
{
    const int64_t total = 2LL * 10 * g.n * g.n;
    hipLaunchKernelGGL(k_boundary_blocks<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, eta, bc_cmb,
                       bc_surf, blk);
    TN_HIP(hipGetLastError());
}

bool march_supported(int nr) { return nr - 1 >= 2 && nr - 1 <= 128; }

template <typename T>
void launch_eps_march(const MarchArgs<T>& a, bool diagonal, hipStream_t s)
{
    const int ncr = a.nr - 1;
    TN_REQUIRE(ncr >= 2 && ncr <= 128, "march kernel: 2..128 radial cells");
    constexpr int OWN = MARCH_WAVES - 1;
    const int nxb = (a.n + OWN - 1) / OWN, nys = (a.n + a.seg - 1) / a.seg;
    const dim3 gr((unsigned)(10 * nxb * nys)), bl(64 * MARCH_WAVES);
    if (ncr <= 64) {
        if (diagonal)
            hipLaunchKernelGGL((k_eps_march<T, 1, true>), gr, bl, 0, s, a);
        else
            hipLaunchKernelGGL((k_eps_march<T, 1, false>), gr, bl, 0, s, a);
    } else {
        if (diagonal)
            hipLaunchKernelGGL((k_eps_march<T, 2, true>), gr, bl, 0, s, a);
        else
            hipLaunchKernelGGL((k_eps_march<T, 2, false>), gr, bl, 0, s, a);
    }
    TN_HIP(hipGetLastError());
}

template void launch_boundary_blocks<double>(const Geom<double>&, const double*, int, int, double*, hipStream_t);
template void launch_boundary_blocks<float>(const Geom<float>&, const float*, int, int, float*, hipStream_t);
template void launch_eps_march<double>(const MarchArgs<double>&, bool, hipStream_t);
template void launch_eps_march<float>(const MarchArgs<float>&, bool, hipStream_t);

}  // namespace tn
