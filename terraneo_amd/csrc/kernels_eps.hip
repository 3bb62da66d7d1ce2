// epsilon-div-div wedge operator: matrix-free apply and diagonal, CDNA4.
//
// Operator (reference EpsilonDivDivKerngen, epsilon_divdiv_kerngen.hpp:1055-1389
// for Dirichlet/Neumann, :1421-1960 for free-slip):
//   per wedge, one centroid quadrature point, kwJ = mean(eta_6) |det J|,
//   y_n = kwJ * (2 eps(u) - 2/3 div(u) I) grad N_n.
//
// Formulation used here (own derivation): J = A diag(r_mid, r_mid, dr/2) with
// A a purely lateral 3x3 matrix, so every shape gradient is
//   grad N_(i,bottom) = a_i - zz ,  grad N_(i,top) = a_i + zz,
//   a_i = l_i / r_mid,  zz = z * 2/dr,  l0 = -(l1 + l2)
// with l1, l2, z precomputed per lateral wedge (MeshLevel::geo).  Then
//   grad u = (S1-S0) (x) a1 + (S2-S0) (x) a2 + D (x) zz,
//   S_i = u_bot_i + u_top_i,  D = sum_i (u_top_i - u_bot_i),
// and the six outputs are p_i -/+ q with p_i = sigma a_i, q = sigma zz.
// That is ~190 flops per wedge instead of the ~560 of the gather/scatter
// form, which matters because the FP64 kernel sits near the VALU ridge.
//
// Launch: one thread per hex cell, radial index fastest, so a wavefront
// reads radially contiguous node columns (coalesced).  Radially adjacent
// cells of one wavefront combine their shared node level through a lane
// shuffle before the global atomic add, halving the atomics.
#include "tn_internal.hpp"

namespace tn {
namespace {

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T>
__device__ inline V3<T> v3(const T* p)
{
    return {p[0], p[1], p[2]};
}

template <typename T>
__device__ inline T dot3(const V3<T>& a, const V3<T>& b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Normal-tangential basis: rows n, t1, t2 (reference
// local_basis_trafo_normal_tangential.hpp:34-79: reference axis = the
// coordinate axis least aligned with n).
template <typename T>
__device__ inline void nt_basis(V3<T> n, V3<T>& t1, V3<T>& t2)
{
    const T ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
    V3<T> ref{0, 0, 0};
    if (ax <= ay && ax <= az)
        ref.x = 1;
    else if (ay <= ax && ay <= az)
        ref.y = 1;
    else
        ref.z = 1;
    t1 = {ref.y * n.z - ref.z * n.y, ref.z * n.x - ref.x * n.z, ref.x * n.y - ref.y * n.x};
    T l = sqrt(dot3(t1, t1));
    if (l < T(1e-15)) {
        t1 = {n.y, -n.x, T(0)};
        l = sqrt(dot3(t1, t1));
    }
    t1 = {t1.x / l, t1.y / l, t1.z / l};
    t2 = {n.y * t1.z - n.z * t1.y, n.z * t1.x - n.x * t1.z, n.x * t1.y - n.y * t1.x};
    l = sqrt(dot3(t2, t2));
    t2 = {t2.x / l, t2.y / l, t2.z / l};
}

template <typename T>
struct CellIO {
    V3<T> sb[4], st[4];  // source, lateral corners 00,10,01,11 at bottom/top
    T eb[4], et[4];      // viscosity
    V3<T> yb[4], yt[4];  // outputs
};

// corner ids of the three lateral vertices of wedge w (kernel_helpers.hpp:302-327)
__device__ constexpr int WC[2][3] = {{0, 1, 2}, {3, 2, 1}};

template <typename T>
__device__ inline V3<T> normal_of(const T* coords, int64_t lat_idx)
{
    V3<T> c = v3(coords + 3 * lat_idx);
    const T inv = T(1) / sqrt(dot3(c, c));
    return {c.x * inv, c.y * inv, c.z * inv};
}

// Per-cell operator.  bmode/tmode: boundary condition of the bottom/top node
// layer of this cell (TN_NEUMANN when the layer is interior).
template <typename T, bool DIAG>
__device__ inline void eps_cell(CellIO<T>& io, const T* __restrict__ geo, T r0, T r1, int bmode, int tmode,
                                const V3<T> nrm[4])
{
    const T half_dr = T(0.5) * (r1 - r0);
    const T r_mid = T(0.5) * (r0 + r1);
    const T alpha = T(1) / r_mid, beta = T(1) / half_dr;
    const T kr = r_mid * r_mid * half_dr * T(1.0 / 6.0);
    T ann_b[4] = {0, 0, 0, 0}, ann_t[4] = {0, 0, 0, 0};
    for (int c = 0; c < 4; ++c) io.yb[c] = io.yt[c] = V3<T>{0, 0, 0};

#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const T* gw = geo + w * GEO_STRIDE;
        const int c0 = WC[w][0], c1 = WC[w][1], c2 = WC[w][2];
        const T ksum = (io.eb[c0] + io.eb[c1] + io.eb[c2]) + (io.et[c0] + io.et[c1] + io.et[c2]);
        const T kwJ = ksum * kr * gw[9];
        V3<T> a[3];
        a[1] = {alpha * gw[0], alpha * gw[1], alpha * gw[2]};
        a[2] = {alpha * gw[3], alpha * gw[4], alpha * gw[5]};
        a[0] = {-(a[1].x + a[2].x), -(a[1].y + a[2].y), -(a[1].z + a[2].z)};
        const V3<T> zz{beta * gw[6], beta * gw[7], beta * gw[8]};

        if constexpr (DIAG) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int c = WC[w][i];
#pragma unroll
                for (int lev = 0; lev < 2; ++lev) {
                    const int mode = lev ? tmode : bmode;
                    const V3<T> g = lev ? V3<T>{a[i].x + zz.x, a[i].y + zz.y, a[i].z + zz.z}
                                        : V3<T>{a[i].x - zz.x, a[i].y - zz.y, a[i].z - zz.z};
                    const T gg = dot3(g, g);
                    const V3<T> s = lev ? io.st[c] : io.sb[c];
                    V3<T> o;
                    if (mode == TN_FREESLIP) {
                        // R^T diag(R A_cc R^T) R s, rows of R = (n, t1, t2)
                        const V3<T> nn = nrm[c];
                        V3<T> t1, t2;
                        nt_basis(nn, t1, t2);
                        const T g0 = dot3(nn, g), g1 = dot3(t1, g), g2 = dot3(t2, g);
                        const T v0 = kwJ * (gg + T(1.0 / 3.0) * g0 * g0) * dot3(nn, s);
                        const T v1 = kwJ * (gg + T(1.0 / 3.0) * g1 * g1) * dot3(t1, s);
                        const T v2 = kwJ * (gg + T(1.0 / 3.0) * g2 * g2) * dot3(t2, s);
                        o = {nn.x * v0 + t1.x * v1 + t2.x * v2, nn.y * v0 + t1.y * v1 + t2.y * v2,
                             nn.z * v0 + t1.z * v1 + t2.z * v2};
                    } else {
                        o = {kwJ * s.x * (gg + T(1.0 / 3.0) * g.x * g.x), kwJ * s.y * (gg + T(1.0 / 3.0) * g.y * g.y),
                             kwJ * s.z * (gg + T(1.0 / 3.0) * g.z * g.z)};
                    }
                    V3<T>& y = lev ? io.yt[c] : io.yb[c];
                    y.x += o.x;
                    y.y += o.y;
                    y.z += o.z;
                }
            }
        } else {
            // trial side: Dirichlet layers drop out, free-slip layers are projected
            V3<T> ub[3], ut[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int c = WC[w][i];
                ub[i] = io.sb[c];
                ut[i] = io.st[c];
                if (bmode == TN_DIRICHLET) ub[i] = V3<T>{0, 0, 0};
                if (tmode == TN_DIRICHLET) ut[i] = V3<T>{0, 0, 0};
                if (bmode == TN_FREESLIP) {
                    const T un = dot3(nrm[c], ub[i]);
                    ub[i] = {ub[i].x - un * nrm[c].x, ub[i].y - un * nrm[c].y, ub[i].z - un * nrm[c].z};
                }
                if (tmode == TN_FREESLIP) {
                    const T un = dot3(nrm[c], ut[i]);
                    ut[i] = {ut[i].x - un * nrm[c].x, ut[i].y - un * nrm[c].y, ut[i].z - un * nrm[c].z};
                }
            }
            const V3<T> S0{ub[0].x + ut[0].x, ub[0].y + ut[0].y, ub[0].z + ut[0].z};
            const V3<T> S1{ub[1].x + ut[1].x - S0.x, ub[1].y + ut[1].y - S0.y, ub[1].z + ut[1].z - S0.z};
            const V3<T> S2{ub[2].x + ut[2].x - S0.x, ub[2].y + ut[2].y - S0.y, ub[2].z + ut[2].z - S0.z};
            const V3<T> D{(ut[0].x + ut[1].x + ut[2].x) - (ub[0].x + ub[1].x + ub[2].x),
                          (ut[0].y + ut[1].y + ut[2].y) - (ub[0].y + ub[1].y + ub[2].y),
                          (ut[0].z + ut[1].z + ut[2].z) - (ub[0].z + ub[1].z + ub[2].z)};
            // G[p][q] = d u_p / d x_q
            const T G00 = S1.x * a[1].x + S2.x * a[2].x + D.x * zz.x;
            const T G01 = S1.x * a[1].y + S2.x * a[2].y + D.x * zz.y;
            const T G02 = S1.x * a[1].z + S2.x * a[2].z + D.x * zz.z;
            const T G10 = S1.y * a[1].x + S2.y * a[2].x + D.y * zz.x;
            const T G11 = S1.y * a[1].y + S2.y * a[2].y + D.y * zz.y;
            const T G12 = S1.y * a[1].z + S2.y * a[2].z + D.y * zz.z;
            const T G20 = S1.z * a[1].x + S2.z * a[2].x + D.z * zz.x;
            const T G21 = S1.z * a[1].y + S2.z * a[2].y + D.z * zz.y;
            const T G22 = S1.z * a[1].z + S2.z * a[2].z + D.z * zz.z;
            const T div = G00 + G11 + G22;
            const T tkw = T(2) * kwJ;
            const T dd = T(2.0 / 3.0) * kwJ * div;
            const T s00 = tkw * G00 - dd, s11 = tkw * G11 - dd, s22 = tkw * G22 - dd;
            const T s01 = kwJ * (G01 + G10), s02 = kwJ * (G02 + G20), s12 = kwJ * (G12 + G21);
            auto smul = [&](const V3<T>& v) -> V3<T> {
                return {s00 * v.x + s01 * v.y + s02 * v.z, s01 * v.x + s11 * v.y + s12 * v.z,
                        s02 * v.x + s12 * v.y + s22 * v.z};
            };
            V3<T> p[3];
            p[1] = smul(a[1]);
            p[2] = smul(a[2]);
            p[0] = {-(p[1].x + p[2].x), -(p[1].y + p[2].y), -(p[1].z + p[2].z)};
            const V3<T> q = smul(zz);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int c = WC[w][i];
                if (bmode == TN_DIRICHLET) {
                    const V3<T> g{a[i].x - zz.x, a[i].y - zz.y, a[i].z - zz.z};
                    const T gg = dot3(g, g);
                    io.yb[c].x += kwJ * io.sb[c].x * (gg + T(1.0 / 3.0) * g.x * g.x);
                    io.yb[c].y += kwJ * io.sb[c].y * (gg + T(1.0 / 3.0) * g.y * g.y);
                    io.yb[c].z += kwJ * io.sb[c].z * (gg + T(1.0 / 3.0) * g.z * g.z);
                } else {
                    io.yb[c].x += p[i].x - q.x;
                    io.yb[c].y += p[i].y - q.y;
                    io.yb[c].z += p[i].z - q.z;
                    if (bmode == TN_FREESLIP) {
                        const V3<T> g{a[i].x - zz.x, a[i].y - zz.y, a[i].z - zz.z};
                        const T ng = dot3(nrm[c], g);
                        ann_b[c] += kwJ * (dot3(g, g) + T(1.0 / 3.0) * ng * ng);
                    }
                }
                if (tmode == TN_DIRICHLET) {
                    const V3<T> g{a[i].x + zz.x, a[i].y + zz.y, a[i].z + zz.z};
                    const T gg = dot3(g, g);
                    io.yt[c].x += kwJ * io.st[c].x * (gg + T(1.0 / 3.0) * g.x * g.x);
                    io.yt[c].y += kwJ * io.st[c].y * (gg + T(1.0 / 3.0) * g.y * g.y);
                    io.yt[c].z += kwJ * io.st[c].z * (gg + T(1.0 / 3.0) * g.z * g.z);
                } else {
                    io.yt[c].x += p[i].x + q.x;
                    io.yt[c].y += p[i].y + q.y;
                    io.yt[c].z += p[i].z + q.z;
                    if (tmode == TN_FREESLIP) {
                        const V3<T> g{a[i].x + zz.x, a[i].y + zz.y, a[i].z + zz.z};
                        const T ng = dot3(nrm[c], g);
                        ann_t[c] += kwJ * (dot3(g, g) + T(1.0 / 3.0) * ng * ng);
                    }
                }
            }
        }
    }
    if constexpr (!DIAG) {
        // test-side projection of the free-slip layer plus the normal diagonal
        // (reference :1854-1922): y <- P y + Ann (n . u) n
        if (bmode == TN_FREESLIP) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const V3<T> nn = nrm[c];
                const T yn = dot3(nn, io.yb[c]);
                const T corr = ann_b[c] * dot3(nn, io.sb[c]) - yn;
                io.yb[c] = {io.yb[c].x + corr * nn.x, io.yb[c].y + corr * nn.y, io.yb[c].z + corr * nn.z};
            }
        }
        if (tmode == TN_FREESLIP) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const V3<T> nn = nrm[c];
                const T yn = dot3(nn, io.yt[c]);
                const T corr = ann_t[c] * dot3(nn, io.st[c]) - yn;
                io.yt[c] = {io.yt[c].x + corr * nn.x, io.yt[c].y + corr * nn.y, io.yt[c].z + corr * nn.z};
            }
        }
    }
}

template <typename T>
__device__ inline T shfl_up1(T v)
{
    return __shfl_up(v, 1, 64);
}

template <typename T, bool DIAG>
__global__ __launch_bounds__(256) void k_eps_atomic(Geom<T> g, const T* __restrict__ eta, const T* __restrict__ src,
                                                    T* __restrict__ dst, int bc_cmb, int bc_surf, int64_t total)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = tid < total;
    const int64_t t = active ? tid : total - 1;
    const int ncr = g.nr - 1, ncl = g.n - 1;
    const int rc = (int)(t % ncr);
    int64_t lat = t / ncr;
    const int y = (int)(lat % ncl);
    lat /= ncl;
    const int x = (int)(lat % ncl);
    const int d = (int)(lat / ncl);
    const int64_t N = 10LL * g.n * g.n * g.nr;
    const int64_t lat00 = ((int64_t)d * g.n + x) * g.n + y;
    const int64_t latc[4] = {lat00, lat00 + g.n, lat00 + 1, lat00 + g.n + 1};

    CellIO<T> io;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int64_t ib = latc[c] * g.nr + rc;
        io.sb[c] = {src[ib], src[N + ib], src[2 * N + ib]};
        io.st[c] = {src[ib + 1], src[N + ib + 1], src[2 * N + ib + 1]};
        io.eb[c] = eta[ib];
        io.et[c] = eta[ib + 1];
    }
    const int bmode = rc == 0 ? bc_cmb : TN_NEUMANN;
    const int tmode = rc == ncr - 1 ? bc_surf : TN_NEUMANN;
    V3<T> nrm[4];
    if (bmode == TN_FREESLIP || tmode == TN_FREESLIP) {
#pragma unroll
        for (int c = 0; c < 4; ++c) nrm[c] = normal_of(g.coords, latc[c]);
    }
    const T* geo = g.geo + ((((int64_t)d * ncl + x) * ncl + y) * 2) * GEO_STRIDE;
    eps_cell<T, DIAG>(io, geo, g.radii[rc], g.radii[rc + 1], bmode, tmode, nrm);

    // merge the shared level with the cell below (previous lane), then scatter
    const int lane = threadIdx.x & 63;
    const bool take = active && lane > 0 && rc > 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const T vx = shfl_up1(io.yt[c].x), vy = shfl_up1(io.yt[c].y), vz = shfl_up1(io.yt[c].z);
        if (take) {
            io.yb[c].x += vx;
            io.yb[c].y += vy;
            io.yb[c].z += vz;
        }
    }
    if (!active) return;
    const bool top_flush = lane == 63 || rc == ncr - 1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int64_t ib = latc[c] * g.nr + rc;
        unsafeAtomicAdd(dst + ib, io.yb[c].x);
        unsafeAtomicAdd(dst + N + ib, io.yb[c].y);
        unsafeAtomicAdd(dst + 2 * N + ib, io.yb[c].z);
        if (top_flush) {
            unsafeAtomicAdd(dst + ib + 1, io.yt[c].x);
            unsafeAtomicAdd(dst + N + ib + 1, io.yt[c].y);
            unsafeAtomicAdd(dst + 2 * N + ib + 1, io.yt[c].z);
        }
    }
}

}  // namespace

template <typename T>
void launch_eps_apply(const Geom<T>& g, const T* eta, const T* src, T* dst, const EpsLaunch& cfg, hipStream_t s)
{
    const int64_t total = 10LL * (g.n - 1) * (g.n - 1) * (g.nr - 1);
    const int block = 256;
    const int64_t grid = (total + block - 1) / block;
    TN_REQUIRE(grid < (1LL << 31), "grid too large");
    if (cfg.diagonal)
        hipLaunchKernelGGL((k_eps_atomic<T, true>), dim3((unsigned)grid), dim3(block), 0, s, g, eta, src, dst,
                           cfg.bc_cmb, cfg.bc_surf, total);
    else
        hipLaunchKernelGGL((k_eps_atomic<T, false>), dim3((unsigned)grid), dim3(block), 0, s, g, eta, src, dst,
                           cfg.bc_cmb, cfg.bc_surf, total);
    TN_HIP(hipGetLastError());
}

template void launch_eps_apply<double>(const Geom<double>&, const double*, const double*, double*, const EpsLaunch&,
                                       hipStream_t);
template void launch_eps_apply<float>(const Geom<float>&, const float*, const float*, float*, const EpsLaunch&,
                                      hipStream_t);

}  // namespace tn
