// Multigrid kernels: constant (linear-on-wedges) prolongation/restriction,
// Galerkin coarse-grid (GCA) assembly and the stored-matrix operator apply.
//
// Reference counterparts:
//   ProlongationVecConstant / RestrictionVecConstant   prolongation_constant.hpp:103-193,
//                                                      restriction_constant.hpp:134-250,
//                                                      weights grid_transfer_constant.hpp:25-89
//   TwoGridGCA (constant interpolation)                linalg/solvers/gca/gca.hpp:154-531
//   slow (stored local matrix) operator path           epsilon_divdiv_kerngen.hpp:823-1051
//
// Storage of coarse operators: the two 18x18 wedge matrices of every hex cell
// are symmetric (P^T A P of a symmetric A), so only the upper triangle is
// kept: 171 entries per wedge, dimension-wise dof order (d*6 + node), packed
// row-major.  Cells (diamond, x, y, r order, r fastest) are grouped in tiles
// of 64, interleaved in 16-byte vectors inside a tile (stored_at in tn_mg.hpp): neighbouring
// lanes (radially adjacent cells) read neighbouring addresses and one wave's
// quad reads one contiguous block.
// The boundary masks and free-slip rotations are applied at apply time, as
// the reference does for its stored matrices.
#include "tn_internal.hpp"
#include "tn_mg.hpp"

namespace tn {
namespace {

__device__ constexpr int WOFF[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                         {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};

struct V3d {
    double x, y, z;
};

// 16-byte vectors of stored entries
template <typename S>
struct VecSel;
template <>
struct VecSel<float> {
    using type = float4;
};
template <>
struct VecSel<double> {
    using type = double2;
};
template <typename S>
using VecOf = typename VecSel<S>::type;
__device__ __forceinline__ float vec_get(const float4& v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
__device__ __forceinline__ double vec_get(const double2& v, int i) { return i == 0 ? v.x : v.y; }
// 16-byte matrix vectors in flight per wave in the stored-matrix apply, and
// the occupancy its register budget is sized for (3 waves per SIMD leave room
// for the vectors in flight; at 4 the compiler waits for every load)
constexpr int STORED_PREFETCH = 6;
constexpr int STORED_WAVES_PER_EU = 3;
// The matrix stream is read once per apply and is larger than the L2 and the
// last-level cache: non-temporal loads keep it from evicting the vectors.
__device__ __forceinline__ float4 ld_stream(const float4* p)
{
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 r = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return make_float4(r.x, r.y, r.z, r.w);
}
__device__ __forceinline__ double2 ld_stream(const double2* p)
{
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 r = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_double2(r.x, r.y);
}

// A wave-uniform pointer, made opaque through readfirstlane so that a load
// p[off32] becomes one global_load with an SGPR base and a 32-bit VGPR offset
// (otherwise the address arithmetic is folded into a 64-bit VGPR address per
// load, which for a 171-entry matrix row stream costs ~340 VGPRs of addresses).
__host__ __device__ constexpr int pk(int i, int j)  // packed upper index, i <= j
{
    return i * 18 - (i * (i - 1)) / 2 + (j - i);
}

// ---------------------------------------------------------------- transfers

// Prolongation (Add): every fine node gathers from its coarse parents.
template <typename T>
__global__ void k_prolongate_add(const T* __restrict__ xc, T* __restrict__ xf, int nf, int nrf, int nc, int nrc,
                                 int64_t Nf, int64_t Nc, int ncomp)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= Nf) return;
    const int rf = (int)(tid % nrf);
    int64_t lat = tid / nrf;
    const int yf = (int)(lat % nf);
    lat /= nf;
    const int xf_ = (int)(lat % nf);
    const int d = (int)(lat / nf);
    // lateral parents: vertex, edge midpoint (x, y or anti-diagonal)
    int px[2], py[2], np;
    T lw;
    if (!(xf_ & 1) && !(yf & 1)) {
        px[0] = xf_ >> 1; py[0] = yf >> 1; np = 1; lw = T(1);
    } else if ((xf_ & 1) && !(yf & 1)) {
        px[0] = xf_ >> 1; py[0] = yf >> 1; px[1] = px[0] + 1; py[1] = py[0]; np = 2; lw = T(0.5);
    } else if (!(xf_ & 1) && (yf & 1)) {
        px[0] = xf_ >> 1; py[0] = yf >> 1; px[1] = px[0]; py[1] = py[0] + 1; np = 2; lw = T(0.5);
    } else {
        px[0] = (xf_ >> 1) + 1; py[0] = yf >> 1; px[1] = xf_ >> 1; py[1] = (yf >> 1) + 1; np = 2; lw = T(0.5);
    }
    int pr[2], nq;
    T rw;
    if (!(rf & 1)) {
        pr[0] = rf >> 1; nq = 1; rw = T(1);
    } else {
        pr[0] = rf >> 1; pr[1] = pr[0] + 1; nq = 2; rw = T(0.5);
    }
    const T w = lw * rw;
    for (int c = 0; c < ncomp; ++c) {
        T s = 0;
        for (int a = 0; a < np; ++a)
            for (int b = 0; b < nq; ++b)
                s += xc[c * Nc + ((((int64_t)d * nc + px[a]) * nc + py[a]) * nrc + pr[b])];
        xf[c * Nf + tid] += w * s;
    }
}

// ---------------------------------------------------------------- linear transfers
// Geometry-aware linear prolongation (ProlongationLinear / ProlongationVecLinear,
// prolongation_linear.hpp; weights grid_transfer_linear.hpp:13-157): a fine
// node coincident with a coarse node takes its value; a fine node between two
// radially aligned coarse nodes interpolates linearly in |x|; a lateral
// midpoint interpolates between the chord midpoints of its coarse edge on the
// bottom and top coarse shells (negative weights near the surface are part of
// the scheme).  Positions are unit lateral coordinates times radii of the fine
// mesh, in FP64.
struct LinStencil {
    int n;          // 1, 2 or 4 coarse nodes
    int kase;       // 0 coincident, 1 radially aligned, 2 lateral midpoint
    int64_t c[4];   // coarse node index (lat * nrc + r)
    double w[4];
};

__device__ inline double lin_norm_at(const double* __restrict__ cf, const double* __restrict__ radf, int d, int nf,
                                     int x, int y, int r)
{
    const double* u = cf + 3 * (((int64_t)d * nf + x) * nf + y);
    const double R = radf[r];
    const double p0 = u[0] * R, p1 = u[1] * R, p2 = u[2] * R;
    return sqrt(p0 * p0 + p1 * p1 + p2 * p2);
}

__device__ inline double lin_chord_norm(const double* __restrict__ cf, const double* __restrict__ radf, int d,
                                        int nf, int x0, int y0, int x1, int y1, int r)
{
    const double* u0 = cf + 3 * (((int64_t)d * nf + x0) * nf + y0);
    const double* u1 = cf + 3 * (((int64_t)d * nf + x1) * nf + y1);
    const double R = radf[r];
    const double p0 = 0.5 * (u0[0] * R + u1[0] * R), p1 = 0.5 * (u0[1] * R + u1[1] * R),
                 p2 = 0.5 * (u0[2] * R + u1[2] * R);
    return sqrt(p0 * p0 + p1 * p1 + p2 * p2);
}

__device__ inline LinStencil lin_stencil(const double* __restrict__ cf, const double* __restrict__ radf, int d, int xf,
                                         int yf, int rf, int nf, int nrf, int nc, int nrc)
{
    LinStencil s;
    auto cidx = [&](int x, int y, int r) { return (((int64_t)d * nc + x) * nc + y) * nrc + r; };
    if (!(xf & 1) && !(yf & 1) && !(rf & 1)) {
        s.n = 1;
        s.kase = 0;
        s.c[0] = cidx(xf >> 1, yf >> 1, rf >> 1);
        s.w[0] = 1.0;
        return s;
    }
    const int rb = rf < nrf - 1 ? rf / 2 : rf / 2 - 1, rt = rb + 1;
    const double fn = lin_norm_at(cf, radf, d, nf, xf, yf, rf);
    if (!(xf & 1) && !(yf & 1)) {
        const int X = xf >> 1, Y = yf >> 1;
        const double bn = lin_norm_at(cf, radf, d, nf, 2 * X, 2 * Y, 2 * rb);
        const double tn = lin_norm_at(cf, radf, d, nf, 2 * X, 2 * Y, 2 * rt);
        const double nu = (fn - bn) / (tn - bn);
        s.n = 2;
        s.kase = 1;
        s.c[0] = cidx(X, Y, rb);
        s.c[1] = cidx(X, Y, rt);
        s.w[0] = 1.0 - nu;
        s.w[1] = nu;
        return s;
    }
    int x0, x1, y0, y1;
    if (!(xf & 1)) {  // "vertical" edge
        x0 = x1 = xf / 2;
        y0 = yf / 2;
        y1 = yf / 2 + 1;
    } else if (!(yf & 1)) {  // "horizontal" edge
        x0 = xf / 2;
        x1 = xf / 2 + 1;
        y0 = y1 = yf / 2;
    } else {  // "diagonal" edge
        x0 = xf / 2 + 1;
        x1 = xf / 2;
        y0 = yf / 2;
        y1 = yf / 2 + 1;
    }
    const double bn = lin_chord_norm(cf, radf, d, nf, 2 * x0, 2 * y0, 2 * x1, 2 * y1, 2 * rb);
    const double tn = lin_chord_norm(cf, radf, d, nf, 2 * x0, 2 * y0, 2 * x1, 2 * y1, 2 * rt);
    const double nu = (fn - bn) / (tn - bn);
    s.n = 4;
    s.kase = 2;
    s.c[0] = cidx(x0, y0, rb);
    s.c[1] = cidx(x1, y1, rb);
    s.c[2] = cidx(x0, y0, rt);
    s.c[3] = cidx(x1, y1, rt);
    s.w[0] = s.w[1] = 0.5 * (1.0 - nu);
    s.w[2] = s.w[3] = 0.5 * nu;
    return s;
}

// xf (+)= P_lin xc, thread per fine node.  quirk: ProlongationVecLinear's
// early return (prolongation_linear.hpp:284,298) -- at coincident and radially
// aligned fine nodes only component 0 is transferred.
template <typename T>
__global__ void k_prolongate_linear(const T* __restrict__ xc, T* __restrict__ xf, const double* __restrict__ cf,
                                    const double* __restrict__ radf, int nf, int nrf, int nc, int nrc, int64_t Nf,
                                    int64_t Nc, int ncomp, int add, int quirk)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= Nf) return;
    const int rf = (int)(tid % nrf);
    int64_t lat = tid / nrf;
    const int yf = (int)(lat % nf);
    lat /= nf;
    const int xf_ = (int)(lat % nf);
    const int d = (int)(lat / nf);
    const LinStencil st = lin_stencil(cf, radf, d, xf_, yf, rf, nf, nrf, nc, nrc);
    const int nc_used = (quirk && st.kase < 2) ? 1 : ncomp;
    for (int c = 0; c < ncomp; ++c) {
        double v = 0.0;
        if (c < nc_used)
            for (int k = 0; k < st.n; ++k) v += st.w[k] * (double)xc[c * Nc + st.c[k]];
        if (add)
            xf[c * Nf + tid] += (T)v;
        else
            xf[c * Nf + tid] = (T)v;
    }
}

// xc = R_lin xf (the transpose over owned fine nodes; interface sum follows):
// thread per coarse node, gathering from the fine nodes whose stencil holds
// it (lateral offsets -1..1, radial -2..2), no atomics.  quirk: the early
// return of RestrictionVecLinear (restriction_linear.hpp:354,367).
template <typename T>
__global__ void k_restrict_linear(const T* __restrict__ xf, T* __restrict__ xc, const uint8_t* __restrict__ owned_lat_f,
                                  const double* __restrict__ cf, const double* __restrict__ radf, int nf, int nrf,
                                  int nc, int nrc, int64_t Nf, int64_t Nc, int ncomp, int quirk)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= Nc) return;
    const int R = (int)(tid % nrc);
    int64_t lat = tid / nrc;
    const int Y = (int)(lat % nc);
    lat /= nc;
    const int X = (int)(lat % nc);
    const int d = (int)(lat / nc);
    double acc[3] = {0, 0, 0};
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int fx = 2 * X + dx, fy = 2 * Y + dy;
            if (fx < 0 || fx >= nf || fy < 0 || fy >= nf) continue;
            const int64_t flat = ((int64_t)d * nf + fx) * nf + fy;
            if (!owned_lat_f[flat]) continue;
            for (int dr = -2; dr <= 2; ++dr) {
                const int fr = 2 * R + dr;
                if (fr < 0 || fr >= nrf) continue;
                const LinStencil st = lin_stencil(cf, radf, d, fx, fy, fr, nf, nrf, nc, nrc);
                double w = 0.0;
                for (int k = 0; k < st.n; ++k)
                    if (st.c[k] == tid) w += st.w[k];
                if (w == 0.0) continue;
                const int nc_used = (quirk && st.kase < 2) ? 1 : ncomp;
                for (int c = 0; c < nc_used; ++c) acc[c] += w * (double)xf[c * Nf + flat * nrf + fr];
            }
        }
    for (int c = 0; c < ncomp; ++c) xc[c * Nc + tid] = (T)acc[c];
}

__device__ inline double pc_lat_weight(int dx, int dy)
{
    if (dx == 0 && dy == 0) return 1.0;
    if ((dy == 0 && (dx == 1 || dx == -1)) || (dx == 0 && (dy == 1 || dy == -1)) || (dx == -1 && dy == 1) ||
        (dx == 1 && dy == -1))
        return 0.5;
    return 0.0;
}

// Restriction (Replace): coarse node = owned-masked weighted sum over the
// 21-point stencil of fine nodes; interface sum follows on the coarse mesh.
template <typename T>
__global__ void k_restrict(const T* __restrict__ xf, T* __restrict__ xc, const uint8_t* __restrict__ owned_lat_f,
                           int nf, int nrf, int nc, int nrc, int64_t Nf, int64_t Nc, int ncomp)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= Nc) return;
    const int r = (int)(tid % nrc);
    int64_t lat = tid / nrc;
    const int y = (int)(lat % nc);
    lat /= nc;
    const int x = (int)(lat % nc);
    const int d = (int)(lat / nc);
    constexpr int ST[7][2] = {{-1, 0}, {-1, 1}, {0, 1}, {1, 0}, {1, -1}, {0, -1}, {0, 0}};
    double acc[3] = {0, 0, 0};
    for (int s = 0; s < 7; ++s) {
        const int fx = 2 * x + ST[s][0], fy = 2 * y + ST[s][1];
        if (fx < 0 || fx >= nf || fy < 0 || fy >= nf) continue;
        const int64_t flat = ((int64_t)d * nf + fx) * nf + fy;
        if (!owned_lat_f[flat]) continue;
        const double lw = pc_lat_weight(ST[s][0], ST[s][1]);
        for (int dr = -1; dr <= 1; ++dr) {
            const int fr = 2 * r + dr;
            if (fr < 0 || fr >= nrf) continue;
            const double w = lw * (dr == 0 ? 1.0 : 0.5);
            for (int c = 0; c < ncomp; ++c) acc[c] += w * (double)xf[c * Nf + flat * nrf + fr];
        }
    }
    for (int c = 0; c < ncomp; ++c) xc[c * Nc + tid] = (T)acc[c];
}

// ---------------------------------------------------------------- GCA assembly

// Interpolation from the 6 coarse wedge dofs onto the 6 dofs of one fine
// wedge (gca.hpp:195-345, constant mode).  (fx, fy, fr) = fine hex cell,
// fine wedge w, nrf = fine radial node count.  Reproduces the coded radial
// bracket r_bot = r/2 (gca.hpp:213-216); geometric != 0 takes the bracket of
// the coarse cell R the fine dof lies in instead (exact Galerkin product).
__device__ inline void gca_local_P(int fx, int fy, int fr, int w, int nrf, int R, int geometric, double P[6][6])
{
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) P[i][j] = 0.0;
    for (int fl = 0; fl < 6; ++fl) {
        const int ix = fx + WOFF[w][fl][0], iy = fy + WOFF[w][fl][1], ir = fr + WOFF[w][fl][2];
        if (!(ix & 1) && !(iy & 1) && !(ir & 1)) {
            P[fl][fl] = 1.0;
            continue;
        }
        const int rb = geometric ? R : (ir < nrf - 1 ? ir / 2 : ir / 2 - 1), rt = rb + 1;
        if (!(ix & 1) && !(iy & 1)) {
            const int a = fl % 3;
            P[fl][a] = 0.5;
            P[fl][a + 3] = 0.5;
            continue;
        }
        int x0, x1, y0, y1, cl[4];
        if (!(ix & 1)) {
            x0 = x1 = ix / 2; y0 = iy / 2; y1 = iy / 2 + 1;
            cl[0] = 0; cl[1] = 2; cl[2] = 3; cl[3] = 5;
        } else if (!(iy & 1)) {
            x0 = ix / 2; x1 = ix / 2 + 1; y0 = y1 = iy / 2;
            cl[0] = 0; cl[1] = 1; cl[2] = 3; cl[3] = 4;
        } else {
            x0 = ix / 2 + 1; x1 = ix / 2; y0 = iy / 2; y1 = iy / 2 + 1;
            cl[0] = 1; cl[1] = 2; cl[2] = 4; cl[3] = 5;
        }
        auto wgt = [&](int xc, int yc, int rc) {
            const int dr = ir - 2 * rc;
            const double lw = pc_lat_weight(ix - 2 * xc, iy - 2 * yc);
            return dr == 0 ? lw : ((dr == 1 || dr == -1) ? 0.5 * lw : 0.0);
        };
        P[fl][cl[0]] = wgt(x0, y0, rb);
        P[fl][cl[1]] = wgt(x1, y1, rb);
        P[fl][cl[2]] = wgt(x0, y0, rt);
        P[fl][cl[3]] = wgt(x1, y1, rt);
    }
}

// Fine-level 6x6 block (row component q, column component p) of the raw
// epsilon wedge matrix assembled from geometry (one-point quadrature):
//   A_(i,q),(j,p) = kwJ [ delta_pq g_i.g_j + g_i[p] g_j[q] - 2/3 g_i[q] g_j[p] ]
template <typename T>
__device__ inline void eps_block_geo(const T* __restrict__ geo_w, T r0, T r1, double ksum, int q, int p,
                                     double A[6][6])
{
    const double half_dr = 0.5 * ((double)r1 - (double)r0), r_mid = 0.5 * ((double)r0 + (double)r1);
    const double alpha = 1.0 / r_mid, beta = 1.0 / half_dr;
    const double kwJ = ksum / 6.0 * r_mid * r_mid * half_dr * (double)geo_w[9];
    double g[6][3];
    for (int a = 0; a < 3; ++a) {
        const double a1 = alpha * (double)geo_w[a], a2 = alpha * (double)geo_w[3 + a], z = beta * (double)geo_w[6 + a];
        const double a0 = -(a1 + a2);
        g[0][a] = a0 - z; g[1][a] = a1 - z; g[2][a] = a2 - z;
        g[3][a] = a0 + z; g[4][a] = a1 + z; g[5][a] = a2 + z;
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const double gg = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
            A[i][j] = kwJ * ((p == q ? gg : 0.0) + g[i][p] * g[j][q] - (2.0 / 3.0) * g[i][q] * g[j][p]);
        }
}

template <typename S>
__device__ inline void stored_block(const S* __restrict__ st, int64_t ncells, int64_t cell, int w, int q, int p,
                                    double A[6][6])
{
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            int r = i + 6 * q, c = j + 6 * p;
            if (r > c) {
                const int t = r;
                r = c;
                c = t;
            }
            A[i][j] = (double)st[stored_at<S>(cell, w, pk(r, c))];
        }
}

// One thread per (coarse cell, coarse wedge, component block q <= p).
// FINE_GEO: fine operator is matrix-free (raw local matrices from geometry);
// otherwise it is read from the fine level's packed storage.
template <typename T, typename S, typename SF, bool FINE_GEO>
__global__ void k_gca_build(GcaArgs<T, S, SF> a)
{
    constexpr int QP[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncc = 10LL * (a.nc - 1) * (a.nc - 1) * (a.nrc - 1);
    if (tid >= ncc * 12) return;
    const int blk = (int)(tid % 6);
    const int W = (int)((tid / 6) % 2);
    const int64_t cell = tid / 12;
    const int q = QP[blk][0], p = QP[blk][1];
    const int R = (int)(cell % (a.nrc - 1));
    int64_t lat = cell / (a.nrc - 1);
    const int Y = (int)(lat % (a.nc - 1));
    lat /= (a.nc - 1);
    const int X = (int)(lat % (a.nc - 1));
    const int d = (int)(lat / (a.nc - 1));
    constexpr int SH[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    // fine wedges of coarse wedge W (gca.hpp:353-370)
    constexpr int LIST[2][8][2] = {{{0, 0}, {1, 0}, {2, 0}, {4, 0}, {5, 0}, {6, 0}, {0, 1}, {4, 1}},
                                   {{1, 1}, {2, 1}, {3, 1}, {5, 1}, {6, 1}, {7, 1}, {3, 0}, {7, 0}}};
    double out[6][6];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) out[i][j] = 0.0;
    const int ncf = a.nf - 1, ncrf = a.nrf - 1;
    for (int k = 0; k < 8; ++k) {
        const int h = LIST[W][k][0], w = LIST[W][k][1];
        const int fx = 2 * X + SH[h][0], fy = 2 * Y + SH[h][1], fr = 2 * R + SH[h][2];
        double P[6][6], A[6][6];
        gca_local_P(fx, fy, fr, w, a.nrf, R, a.geometric, P);
        const int64_t fcell = (((int64_t)d * ncf + fx) * ncf + fy) * ncrf + fr;
        if constexpr (FINE_GEO) {
            const double* geo_w = a.geo_f + ((((int64_t)d * ncf + fx) * ncf + fy) * 2 + w) * GEO_STRIDE;
            double ksum = 0.0;
            if (a.eta_cell_f) {
                ksum = 6.0 * (double)a.eta_cell_f[fcell];
            } else {
                for (int j = 0; j < 6; ++j)
                    ksum += (double)a.eta_f[(((int64_t)d * a.nf + fx + WOFF[w][j][0]) * a.nf + fy + WOFF[w][j][1]) *
                                                a.nrf +
                                            fr + WOFF[w][j][2]];
            }
            eps_block_geo<double>(geo_w, a.radii_f[fr], a.radii_f[fr + 1], ksum, q, p, A);
        } else {
            const int64_t ncellf = 10LL * ncf * ncf * ncrf;
            stored_block<SF>(a.stored_f, ncellf, fcell, w, q, p, A);
        }
        // out += P^T A P
        double T1[6][6];
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                double s = 0.0;
                for (int l = 0; l < 6; ++l) s += A[i][l] * P[l][j];
                T1[i][j] = s;
            }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                double s = 0.0;
                for (int l = 0; l < 6; ++l) s += P[l][i] * T1[l][j];
                out[i][j] += s;
            }
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const int r = i + 6 * q, c = j + 6 * p;
            if (r <= c) a.stored_c[stored_at<S>(cell, W, pk(r, c))] = (S)out[i][j];
        }
}

// ---------------------------------------------------------------- selective GCA
// GCAElementsCollector (gca_elements_collector.hpp:32-96): a fine cell whose
// wedge has |grad k| > 10 at the centroid, the gradient taken in reference
// coordinates (grad_shape of integrands.hpp at the one-point Felippa point
// (1/3, 1/3, 0)), marks its ancestor cell on the coarsest level (indices
// shifted right by `levels`).  Every writer stores the same 1: the race is benign.
template <typename T>
__global__ void k_gca_collect(const T* __restrict__ k, int n, int nr, int levels, T* __restrict__ flags, int n0,
                              int nr0)
{
    const int ncl = n - 1, ncr = nr - 1;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= 10LL * ncl * ncl * ncr) return;
    const int r = (int)(cell % ncr);
    int64_t lat = cell / ncr;
    const int y = (int)(lat % ncl);
    lat /= ncl;
    const int x = (int)(lat % ncl);
    const int d = (int)(lat / ncl);
    bool mark = false;
    for (int w = 0; w < 2; ++w) {
        double kv[6];
        for (int j = 0; j < 6; ++j)
            kv[j] = (double)k[(((int64_t)d * n + x + WOFF[w][j][0]) * n + y + WOFF[w][j][1]) * nr + r + WOFF[w][j][2]];
        // grad_shape at (1/3, 1/3, 0): lateral parts times shape_rad = 1/2, radial
        // part shape_lat = 1/3 times grad_shape_rad = -+1/2
        const double gx = 0.5 * ((kv[1] - kv[0]) + (kv[4] - kv[3]));
        const double gy = 0.5 * ((kv[2] - kv[0]) + (kv[5] - kv[3]));
        const double gz = (1.0 / 6.0) * ((kv[3] + kv[4] + kv[5]) - (kv[0] + kv[1] + kv[2]));
        if (sqrt(gx * gx + gy * gy + gz * gz) > 10.0) mark = true;
    }
    if (mark)
        flags[(((int64_t)d * n0 + (x >> levels)) * n0 + (y >> levels)) * nr0 + (r >> levels)] = T(1);
}

// Selective storage (local_matrix_storage.hpp:68-100 with the operator's
// on-the-fly assembly for cells without a stored matrix): the cells of a
// Galerkin level whose coarsest ancestor is not flagged get the level's own
// rediscretized wedge matrices (geometry and nodal viscosity of the level).
// One thread per (cell, wedge, component block q <= p), as k_gca_build.
template <typename T, typename S>
__global__ void k_gca_select(const T* __restrict__ flags, int n0, int nr0, int levels, const double* __restrict__ geo,
                             const double* __restrict__ radii, const T* __restrict__ eta,
                             const T* __restrict__ eta_cell, int n, int nr, S* __restrict__ stored)
{
    constexpr int QP[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    const int ncl = n - 1, ncr = nr - 1;
    const int64_t ncells = 10LL * ncl * ncl * ncr;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= ncells * 12) return;
    const int blk = (int)(tid % 6);
    const int w = (int)((tid / 6) % 2);
    const int64_t cell = tid / 12;
    const int r = (int)(cell % ncr);
    int64_t lat = cell / ncr;
    const int y = (int)(lat % ncl);
    lat /= ncl;
    const int x = (int)(lat % ncl);
    const int d = (int)(lat / ncl);
    if (flags && flags[(((int64_t)d * n0 + (x >> levels)) * n0 + (y >> levels)) * nr0 + (r >> levels)] != T(0)) return;
    const int q = QP[blk][0], p = QP[blk][1];
    const double* geo_w = geo + ((((int64_t)d * ncl + x) * ncl + y) * 2 + w) * GEO_STRIDE;
    double ksum = 0.0;
    if (eta_cell) {
        ksum = 6.0 * (double)eta_cell[cell];
    } else {
        for (int j = 0; j < 6; ++j)
            ksum +=
                (double)eta[(((int64_t)d * n + x + WOFF[w][j][0]) * n + y + WOFF[w][j][1]) * nr + r + WOFF[w][j][2]];
    }
    double A[6][6];
    eps_block_geo<double>(geo_w, radii[r], radii[r + 1], ksum, q, p, A);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const int rr = i + 6 * q, cc = j + 6 * p;
            if (rr <= cc) stored[stored_at<S>(cell, w, pk(rr, cc))] = (S)A[i][j];
        }
}

// ---------------------------------------------------------------- stored-matrix apply
//
// Two launches, every node copy finished once, no atomics, deterministic:
//  * k_stored_cells: a wave per lateral quad and chunk of 64 radial cells
//    (one per lane).  Each cell applies its two wedge matrices (plain
//    product, or their diagonal) read as coalesced non-temporal 16-byte
//    vectors of the tile layout (stored_at); the cell's 8 corner sums are merged radially with a
//    lane shuffle (node level c = bottom of cell c + top of cell c-1) and the
//    4 lateral corners x 3 components at level c are written per cell, the
//    chunk's top level once per chunk.  No barriers: the matrix stream runs at
//    the occupancy the registers allow.
//  * k_stored_gather: a thread per node copy sums its (at most 4) lateral
//    quads' corner values and stores (Replace) or adds (Add) the result.
// Boundary conditions are algebraic, as in the matrix-free apply: on a
// Dirichlet / free-slip shell the cells mask / project the trial side, the
// gather completes the test side from the node's boundary forms (k_stored_blk,
// cached per operator) and adds the fs/fs penalty.  This is the matrix the
// reference applies to its stored matrices through masks and rotations
// (epsilon_divdiv_kerngen.hpp:823-1051).
// Trial side on the boundary level of this lane's cell, branch-free (a
// divergent branch here makes the compiler hoist the whole matrix stream):
// keep[t] = 0 zeroes level t (Dirichlet), proj[t] = 1 projects it tangentially
// with the corner normals nrm (free-slip).
template <typename T, typename S, bool DIAG>
__device__ __forceinline__ void stored_cell_sums(const StoredArgs<T, S>& a, int64_t cell, const int64_t latc[4],
                                                 int cc, T scale, const T keep[2], const T proj[2],
                                                 const T nrm[4][3], T yb[4][3], T yt[4][3])
{
    constexpr int WCORN[2][3] = {{0, 1, 2}, {3, 2, 1}};
    const int64_t N = 10LL * a.n * a.n * a.nr;
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        T xs[18];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int64_t ib = latc[WCORN[w][j % 3]] * a.nr + cc + (j >= 3);
#pragma unroll
            for (int k = 0; k < 3; ++k) xs[k * 6 + j] = scale * a.src[k * N + ib];
        }
        if constexpr (!DIAG) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                const int t = j >= 3;
                const T* nn = nrm[WCORN[w][j % 3]];
                const T un = proj[t] * (nn[0] * xs[j] + nn[1] * xs[6 + j] + nn[2] * xs[12 + j]);
#pragma unroll
                for (int k = 0; k < 3; ++k) xs[k * 6 + j] = keep[t] * (xs[k * 6 + j] - un * nn[k]);
            }
        }
        // dof i of the wedge (comp*6 + node) accumulates into its corner sum
        auto Y = [&](int i) -> T& {
            const int node = i % 6, k = i / 6, cr = WCORN[w][node % 3];
            return node < 3 ? yb[cr][k] : yt[cr][k];
        };
        if constexpr (DIAG) {
#pragma unroll
            for (int i = 0; i < 18; ++i) Y(i) += (T)a.stored[stored_at<S>(cell, w, pk(i, i))] * xs[i];
        } else {
            // packed entries (row-major upper triangle), V per 16-byte vector load.
            // Vector q is loaded when the arithmetic reaches vector q - PD, so
            // PD loads per wave are in flight (latency, not bandwidth, bounds a
            // stream with one load in flight per wave).
            constexpr int V = stored_vec<S>(), NV = WEDGE_SLOTS / V;
            constexpr int PD = STORED_PREFETCH;
            using Vec = VecOf<S>;
            const Vec* vp = reinterpret_cast<const Vec*>(a.stored + stored_at<S>(cell, w, 0));
            Vec buf[NV];
#pragma unroll
            for (int q = 0; q < PD; ++q) buf[q] = ld_stream(vp + q * 64);
            int k = 0;
#pragma unroll
            for (int i = 0; i < 18; ++i) {
#pragma unroll
                for (int j = i; j < 18; ++j, ++k) {
                    if (k % V == 0 && k / V + PD < NV) buf[k / V + PD] = ld_stream(vp + (k / V + PD) * 64);
                    const T aij = (T)vec_get(buf[k / V], k % V);
                    Y(i) += aij * xs[j];
                    if (j != i) Y(j) += aij * xs[i];
                }
            }
        }
    }
}

template <typename T, typename S, bool DIAG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STORED_WAVES_PER_EU, STORED_WAVES_PER_EU))) void k_stored_cells(StoredArgs<T, S> a)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int ncr = a.nr - 1, ncl = a.n - 1;
    const int nch = (ncr + 63) >> 6;
    const int64_t ncol = 10LL * ncl * ncl;
    if (wave >= ncol * nch) return;  // wave-uniform
    const int h = (int)(wave % nch);
    const int64_t col = wave / nch;  // (d, x, y) quad
    const int y = (int)(col % ncl);
    const int x = (int)((col / ncl) % ncl);
    const int d = (int)(col / ((int64_t)ncl * ncl));
    const int c = 64 * h + lane;
    const bool valid = c < ncr;
    const int cc = valid ? c : ncr - 1;
    const T scale = valid ? T(1) : T(0);
    // trial side of the boundary levels: bottom of cell 0, top of cell ncr - 1
    const int mb = c == 0 ? a.bc_cmb : TN_NEUMANN, mt = c == ncr - 1 ? a.bc_surf : TN_NEUMANN;
    const T keep[2] = {mb == TN_DIRICHLET ? T(0) : T(1), mt == TN_DIRICHLET ? T(0) : T(1)};
    const T proj[2] = {mb == TN_FREESLIP ? T(1) : T(0), mt == TN_FREESLIP ? T(1) : T(0)};
    const int64_t ncell = ncol * ncr;
    const int toplane = min(63, ncr - 1 - 64 * h);
    const int64_t lat00 = ((int64_t)d * a.n + x) * a.n + y;
    const int64_t latc[4] = {lat00, lat00 + a.n, lat00 + 1, lat00 + a.n + 1};
    T nrm[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (a.bc_cmb == TN_FREESLIP || a.bc_surf == TN_FREESLIP) {  // wave-uniform; coordinates are unit vectors
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < 3; ++j) nrm[k][j] = (T)a.coords[3 * latc[k] + j];
    }
    T yb[4][3], yt[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) yb[k][j] = yt[k][j] = T(0);
    // (the lane condition also keeps the compiler from hoisting the whole
    // matrix stream of both wedges into registers ahead of the arithmetic)
    if (valid) stored_cell_sums<T, S, DIAG>(a, col * ncr + cc, latc, cc, scale, keep, proj, nrm, yb, yt);
    // radial merge (level c = bottom of cell c + top of cell c-1) and write-out
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T up = __shfl_up(yt[k][j], 1, 64);
            if (lane == 0) up = T(0);
            if (valid) a.csum[(int64_t)(k * 3 + j) * ncell + col * ncr + c] = yb[k][j] + up;
            if (lane == toplane) a.ctop[(int64_t)(k * 3 + j) * ncol * nch + col * nch + h] = yt[k][j];
        }
}

// Node copy (d, X, Y, level): sum of its lateral quads' corner values, then on
// Dirichlet / free-slip shells the test side (boundary-node forms blk):
//   Dirichlet  y = diag(B) x            (diagonal mode: already diag(A) x)
//   free-slip  y = P y + (n^T B n)(n . x) n
//              diagonal mode: R^T diag(R B R^T) R x, rows of R = (n, t1, t2)
// and the fs/fs rotation penalty gamma x p (FP64) -- the same algebraic
// boundary treatment as the matrix-free apply; stores (Replace) or adds (Add).
template <typename T, typename S, bool DIAG>
__global__ __launch_bounds__(256) void k_stored_gather(StoredArgs<T, S> a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n, nr = a.nr;
    const int64_t N = 10LL * n * n * nr;
    if (tid >= N) return;
    const int ncr = nr - 1, ncl = n - 1;
    const int nch = (ncr + 63) >> 6;
    const int64_t ncol = 10LL * ncl * ncl, ncell = ncol * ncr;
    const int lev = (int)(tid % nr);
    const int64_t lat = tid / nr;
    const int Yn = (int)(lat % n);
    const int Xn = (int)((lat / n) % n);
    const int d = (int)(lat / ((int64_t)n * n));
    T v[3] = {T(0), T(0), T(0)};
    // node = corner k of quad (Xn - (k & 1), Yn - (k >> 1))
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int qx = Xn - (k & 1), qy = Yn - (k >> 1);
        if (qx < 0 || qy < 0 || qx > ncl - 1 || qy > ncl - 1) continue;
        const int64_t col = ((int64_t)d * ncl + qx) * ncl + qy;
        if (lev < ncr) {
#pragma unroll
            for (int j = 0; j < 3; ++j) v[j] += a.csum[(int64_t)(k * 3 + j) * ncell + col * ncr + lev];
        }
        // first level of a chunk > 0, or the top level: the previous chunk's top
        if ((lev > 0 && (lev & 63) == 0) || lev == ncr) {
            const int hp = (lev - 1) >> 6;
#pragma unroll
            for (int j = 0; j < 3; ++j) v[j] += a.ctop[(int64_t)(k * 3 + j) * ncol * nch + col * nch + hp];
        }
    }
    const int mode = lev == 0 ? a.bc_cmb : (lev == ncr ? a.bc_surf : TN_NEUMANN);
    if (mode != TN_NEUMANN) {
        const T* bk = a.blk + ((lev == 0 ? 0 : 10LL * n * n) + lat) * 3;
        const T x[3] = {a.src[tid], a.src[N + tid], a.src[2 * N + tid]};
        if (mode == TN_DIRICHLET) {
            if constexpr (!DIAG) {
#pragma unroll
                for (int j = 0; j < 3; ++j) v[j] = bk[j] * x[j];
            }
        } else {
            const double* cp = a.coords + 3 * lat;  // unit vector
            double n3[3] = {cp[0], cp[1], cp[2]};
            if constexpr (DIAG) {
                double t1[3], t2[3];
                nt_basis_d(n3, t1, t2);
                const double* rows[3] = {n3, t1, t2};
                double out[3] = {0, 0, 0};
                for (int b = 0; b < 3; ++b) {
                    const double* e = rows[b];
                    const double ex = (double)bk[b] * (e[0] * x[0] + e[1] * x[1] + e[2] * x[2]);
                    for (int j = 0; j < 3; ++j) out[j] += ex * e[j];
                }
#pragma unroll
                for (int j = 0; j < 3; ++j) v[j] = (T)out[j];
            } else {
                const T nx = (T)n3[0], ny = (T)n3[1], nz = (T)n3[2];
                const T corr = bk[0] * (nx * x[0] + ny * x[1] + nz * x[2]) - (nx * v[0] + ny * v[1] + nz * v[2]);
                v[0] += corr * nx;
                v[1] += corr * ny;
                v[2] += corr * nz;
            }
        }
    }
    if (a.pen) {
        // gamma = S beta, formed by the moment pass's finalize
        const double g0 = a.pen[13], g1 = a.pen[14], g2 = a.pen[15];
        const double r = a.radii[lev];
        const double px = a.pen_coords[3 * lat] * r, py = a.pen_coords[3 * lat + 1] * r,
                     pz = a.pen_coords[3 * lat + 2] * r;
        v[0] += (T)(g1 * pz - g2 * py);
        v[1] += (T)(g2 * px - g0 * pz);
        v[2] += (T)(g0 * py - g1 * px);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if (a.add)
            a.dst[j * N + tid] += v[j];
        else
            a.dst[j * N + tid] = v[j];
    }
}

// Boundary-node forms of the stored operator: one thread per (shell, lateral
// node); B sums the node's 3x3 diagonal blocks over the boundary-layer
// wedges of its diamond that contain it.
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_stored_blk(StoredArgs<T, S> a, T* __restrict__ blk)
{
    constexpr int WCORN[2][3] = {{0, 1, 2}, {3, 2, 1}};
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n, ncl = n - 1, ncr = a.nr - 1;
    const int64_t nlat = 10LL * n * n;
    if (tid >= 2 * nlat) return;
    const int shell = (int)(tid / nlat);
    const int mode = shell == 0 ? a.bc_cmb : a.bc_surf;
    if (mode == TN_NEUMANN) return;
    const int64_t lat = tid % nlat;
    const int Yn = (int)(lat % n), Xn = (int)((lat / n) % n), d = (int)(lat / ((int64_t)n * n));
    const int rc = shell == 0 ? 0 : ncr - 1;
    double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 4; ++k) {
        const int qx = Xn - (k & 1), qy = Yn - (k >> 1);
        if (qx < 0 || qy < 0 || qx > ncl - 1 || qy > ncl - 1) continue;
        const int64_t cell = (((int64_t)d * ncl + qx) * ncl + qy) * ncr + rc;
        for (int w = 0; w < 2; ++w) {
            int m = -1;
            for (int i = 0; i < 3; ++i)
                if (WCORN[w][i] == k) m = i;
            if (m < 0) continue;
            const int j = m + (shell == 0 ? 0 : 3);
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    const int r = p * 6 + j, cq = q * 6 + j;
                    B[p][q] += (double)a.stored[stored_at<S>(cell, w, r <= cq ? pk(r, cq) : pk(cq, r))];
                }
        }
    }
    double e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (mode == TN_FREESLIP) {
        const double* cp = a.coords + 3 * lat;  // unit vector
        for (int i = 0; i < 3; ++i) e[0][i] = cp[i];
        nt_basis_d(e[0], e[1], e[2]);
    }
    for (int b = 0; b < 3; ++b) {
        double q = 0.0;
        for (int p = 0; p < 3; ++p)
            for (int r = 0; r < 3; ++r) q += e[b][p] * B[p][r] * e[b][r];
        blk[(shell * nlat + lat) * 3 + b] = (T)q;
    }
}

// Jacobi scaling with 3x3 blocks on free-slip shells: z = dinv * z entrywise,
// except at free-slip boundary node copies, where z = R^T diag(1 / d) R z with
// rows of R = (n, t1, t2) and d = the node's forms (n^T B n, t1^T B t1,
// t2^T B t2) -- the inverse of the operator's free-slip diagonal
// R^T diag(R B R^T) R, which entrywise inversion of that diagonal is not.
template <typename T>
__global__ __launch_bounds__(256) void k_fs_block_scale(T* __restrict__ z, const T* __restrict__ dinv,
                                                        const T* __restrict__ blk, const double* __restrict__ coords,
                                                        int n, int nr, int bc_cmb, int bc_surf)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = 10LL * n * n * nr;
    if (tid >= N) return;
    const int lev = (int)(tid % nr);
    const int mode = lev == 0 ? bc_cmb : (lev == nr - 1 ? bc_surf : TN_NEUMANN);
    if (mode != TN_FREESLIP) {
#pragma unroll
        for (int j = 0; j < 3; ++j) z[j * N + tid] *= dinv[j * N + tid];
        return;
    }
    const int64_t lat = tid / nr;
    const T* bk = blk + ((lev == 0 ? 0 : 10LL * n * n) + lat) * 3;
    const double* cp = coords + 3 * lat;  // unit vector
    double n3[3] = {cp[0], cp[1], cp[2]}, t1[3], t2[3];
    nt_basis_d(n3, t1, t2);
    const double* rows[3] = {n3, t1, t2};
    const double r[3] = {(double)z[tid], (double)z[N + tid], (double)z[2 * N + tid]};
    double out[3] = {0, 0, 0};
    for (int b = 0; b < 3; ++b) {
        const double* e = rows[b];
        const double c = (e[0] * r[0] + e[1] * r[1] + e[2] * r[2]) / (double)bk[b];
        for (int j = 0; j < 3; ++j) out[j] += c * e[j];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) z[j * N + tid] = (T)out[j];
}

}  // namespace

// ---------------------------------------------------------------- host launchers
template <typename T>
void launch_prolongate_add(const MeshLevel& mc, const MeshLevel& mf, int ncomp, const T* xc, T* xf, hipStream_t s)
{
    hipLaunchKernelGGL(k_prolongate_add<T>, dim3((unsigned)((mf.N + 255) / 256)), dim3(256), 0, s, xc, xf, mf.n, mf.nr,
                       mc.n, mc.nr, mf.N, mc.N, ncomp);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_prolongate_linear(const MeshLevel& mc, const MeshLevel& mf, int ncomp, const T* xc, T* xf, bool add,
                              bool quirk, hipStream_t s)
{
    hipLaunchKernelGGL(k_prolongate_linear<T>, dim3((unsigned)((mf.N + 255) / 256)), dim3(256), 0, s, xc, xf,
                       mf.d_coords.p, mf.d_radii.p, mf.n, mf.nr, mc.n, mc.nr, mf.N, mc.N, ncomp, add ? 1 : 0,
                       quirk ? 1 : 0);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_restrict_linear(const MeshLevel& mf, const MeshLevel& mc, int ncomp, const T* xf, T* xc, bool quirk,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_restrict_linear<T>, dim3((unsigned)((mc.N + 255) / 256)), dim3(256), 0, s, xf, xc,
                       mf.d_owned_lat.p, mf.d_coords.p, mf.d_radii.p, mf.n, mf.nr, mc.n, mc.nr, mf.N, mc.N, ncomp,
                       quirk ? 1 : 0);
    TN_HIP(hipGetLastError());
    launch_interface_sum<T>(mc, ncomp, xc, s);
}

template <typename T>
void launch_restrict(const MeshLevel& mf, const MeshLevel& mc, int ncomp, const T* xf, T* xc, hipStream_t s)
{
    hipLaunchKernelGGL(k_restrict<T>, dim3((unsigned)((mc.N + 255) / 256)), dim3(256), 0, s, xf, xc,
                       mf.d_owned_lat.p, mf.n, mf.nr, mc.n, mc.nr, mf.N, mc.N, ncomp);
    TN_HIP(hipGetLastError());
    launch_interface_sum<T>(mc, ncomp, xc, s);
}

template <typename T>
void launch_gca_collect(const MeshLevel& fine, const MeshLevel& coarsest, int levels, const T* k, T* flags,
                        hipStream_t s)
{
    launch_fill<T>(flags, coarsest.N, T(0), s);
    const int64_t cells = 10LL * (fine.n - 1) * (fine.n - 1) * (fine.nr - 1);
    hipLaunchKernelGGL(k_gca_collect<T>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, k, fine.n, fine.nr,
                       levels, flags, coarsest.n, coarsest.nr);
    TN_HIP(hipGetLastError());
}

template <typename T, typename S>
void launch_gca_select(const MeshLevel& coarsest, int levels, const T* flags, const MeshLevel& m, const T* eta,
                       S* stored, hipStream_t s)
{
    const int64_t total = 12LL * 10 * (m.n - 1) * (m.n - 1) * (m.nr - 1);
    hipLaunchKernelGGL((k_gca_select<T, S>), dim3((unsigned)((total + 127) / 128)), dim3(128), 0, s, flags,
                       coarsest.n, coarsest.nr, levels, m.d_geo.p, m.d_radii.p, eta, (const T*)nullptr, m.n, m.nr,
                       stored);
    TN_HIP(hipGetLastError());
}

template <typename T, typename S>
void launch_stored_own(const MeshLevel& m, const T* eta, const T* eta_cell, S* stored, hipStream_t s)
{
    const int64_t total = 12LL * 10 * (m.n - 1) * (m.n - 1) * (m.nr - 1);
    hipLaunchKernelGGL((k_gca_select<T, S>), dim3((unsigned)((total + 127) / 128)), dim3(128), 0, s,
                       (const T*)nullptr, 0, 0, 0, m.d_geo.p, m.d_radii.p, eta, eta_cell, m.n, m.nr, stored);
    TN_HIP(hipGetLastError());
}

namespace {
// one wedge's 18x18 Neumann matrix from the geometry (dimension-wise dof
// order component * 6 + node), a thread per 6x6 component block
template <typename T>
__global__ void k_local_matrix(const double* __restrict__ geo, const double* __restrict__ radii,
                               const T* __restrict__ eta, const T* __restrict__ eta_cell, int n, int nr, int d, int x,
                               int y, int r, int w, double* __restrict__ out)
{
    const int blk = threadIdx.x;
    if (blk >= 9) return;
    const int q = blk / 3, p = blk % 3, ncl = n - 1, ncr = nr - 1;
    const int64_t cell = (((int64_t)d * ncl + x) * ncl + y) * ncr + r;
    const double* geo_w = geo + ((((int64_t)d * ncl + x) * ncl + y) * 2 + w) * GEO_STRIDE;
    double ksum = 0.0;
    if (eta_cell) {
        ksum = 6.0 * (double)eta_cell[cell];
    } else {
        for (int j = 0; j < 6; ++j)
            ksum +=
                (double)eta[(((int64_t)d * n + x + WOFF[w][j][0]) * n + y + WOFF[w][j][1]) * nr + r + WOFF[w][j][2]];
    }
    double A[6][6];
    eps_block_geo<double>(geo_w, radii[r], radii[r + 1], ksum, q, p, A);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) out[(i + 6 * q) * 18 + j + 6 * p] = A[i][j];
}

// full 18x18 matrix of one stored wedge <-> a dense buffer (packed symmetric storage)
template <typename S>
__global__ void k_stored_get(const S* __restrict__ stored, int64_t cell, int w, double* __restrict__ out)
{
    const int t = threadIdx.x;
    if (t >= 324) return;
    const int i = t / 18, j = t % 18;
    out[t] = (double)stored[stored_at<S>(cell, w, i <= j ? pk(i, j) : pk(j, i))];
}
template <typename S>
__global__ void k_stored_set(S* __restrict__ stored, int64_t cell, int w, const double* __restrict__ in)
{
    const int t = threadIdx.x;
    if (t >= 324) return;
    const int i = t / 18, j = t % 18;
    if (i <= j) stored[stored_at<S>(cell, w, pk(i, j))] = (S)in[t];
}
}  // namespace

template <typename T>
void launch_local_matrix(const MeshLevel& m, const T* eta, const T* eta_cell, int d, int x, int y, int r, int w,
                         double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_local_matrix<T>, dim3(1), dim3(64), 0, s, m.d_geo.p, m.d_radii.p, eta, eta_cell, m.n, m.nr,
                       d, x, y, r, w, out);
    TN_HIP(hipGetLastError());
}

template <typename S>
void launch_stored_get(const S* stored, int64_t cell, int w, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_stored_get<S>, dim3(1), dim3(384), 0, s, stored, cell, w, out);
    TN_HIP(hipGetLastError());
}

template <typename S>
void launch_stored_set(S* stored, int64_t cell, int w, const double* in, hipStream_t s)
{
    hipLaunchKernelGGL(k_stored_set<S>, dim3(1), dim3(384), 0, s, stored, cell, w, in);
    TN_HIP(hipGetLastError());
}
template void launch_local_matrix<double>(const MeshLevel&, const double*, const double*, int, int, int, int, int,
                                          double*, hipStream_t);
template void launch_local_matrix<float>(const MeshLevel&, const float*, const float*, int, int, int, int, int,
                                         double*, hipStream_t);
template void launch_stored_get<double>(const double*, int64_t, int, double*, hipStream_t);
template void launch_stored_get<float>(const float*, int64_t, int, double*, hipStream_t);
template void launch_stored_set<double>(double*, int64_t, int, const double*, hipStream_t);
template void launch_stored_set<float>(float*, int64_t, int, const double*, hipStream_t);
template void launch_stored_own<double, double>(const MeshLevel&, const double*, const double*, double*, hipStream_t);
template void launch_stored_own<double, float>(const MeshLevel&, const double*, const double*, float*, hipStream_t);
template void launch_stored_own<float, double>(const MeshLevel&, const float*, const float*, double*, hipStream_t);
template void launch_stored_own<float, float>(const MeshLevel&, const float*, const float*, float*, hipStream_t);

template <typename T, typename S, typename SF>
void launch_gca_build(const GcaArgs<T, S, SF>& a, bool fine_geo, hipStream_t s)
{
    const int64_t ncc = 10LL * (a.nc - 1) * (a.nc - 1) * (a.nrc - 1);
    const int64_t total = ncc * 12;
    const dim3 gr((unsigned)((total + 127) / 128)), bl(128);
    if (fine_geo)
        hipLaunchKernelGGL((k_gca_build<T, S, SF, true>), gr, bl, 0, s, a);
    else
        hipLaunchKernelGGL((k_gca_build<T, S, SF, false>), gr, bl, 0, s, a);
    TN_HIP(hipGetLastError());
}

template <typename T, typename S>
void launch_stored_apply(const StoredArgs<T, S>& a, bool diagonal, hipStream_t s)
{
    TN_REQUIRE(a.nr >= 3, "stored apply: at least two radial cell layers");
    TN_REQUIRE(a.csum && a.ctop, "stored apply: no corner-sum buffers");
    TN_REQUIRE(a.blk || (a.bc_cmb == TN_NEUMANN && a.bc_surf == TN_NEUMANN), "stored apply: no boundary forms");
    const int ncr = a.nr - 1, ncl = a.n - 1, nch = (ncr + 63) / 64;
    const int64_t waves = 10LL * ncl * ncl * nch;
    const dim3 gr((unsigned)((waves + 3) / 4)), bl(256);
    const int64_t N = 10LL * a.n * a.n * a.nr;
    const dim3 gg((unsigned)((N + 255) / 256)), bg(256);
    if (diagonal) {
        hipLaunchKernelGGL((k_stored_cells<T, S, true>), gr, bl, 0, s, a);
        TN_HIP(hipGetLastError());
        hipLaunchKernelGGL((k_stored_gather<T, S, true>), gg, bg, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_stored_cells<T, S, false>), gr, bl, 0, s, a);
        TN_HIP(hipGetLastError());
        hipLaunchKernelGGL((k_stored_gather<T, S, false>), gg, bg, 0, s, a);
    }
    TN_HIP(hipGetLastError());
}

template <typename T, typename S>
void launch_stored_blk(const StoredArgs<T, S>& a, T* blk, hipStream_t s)
{
    const int64_t total = 2LL * 10 * a.n * a.n;
    hipLaunchKernelGGL((k_stored_blk<T, S>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, blk);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_fs_block_scale(const MeshLevel& m, T* z, const T* dinv, const T* blk, int bc_cmb, int bc_surf,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_fs_block_scale<T>, dim3((unsigned)((m.N + 255) / 256)), dim3(256), 0, s, z, dinv, blk,
                       m.d_coords.p, m.n, m.nr, bc_cmb, bc_surf);
    TN_HIP(hipGetLastError());
}
template void launch_fs_block_scale<double>(const MeshLevel&, double*, const double*, const double*, int, int,
                                            hipStream_t);
template void launch_fs_block_scale<float>(const MeshLevel&, float*, const float*, const float*, int, int,
                                           hipStream_t);

// elements of the corner-sum buffers: 12 per cell + 12 per (quad, chunk)
int64_t stored_cellsum_elems(int n, int nr)
{
    const int64_t ncol = 10LL * (n - 1) * (n - 1);
    return 12 * (ncol * (nr - 1) + ncol * ((nr - 1 + 63) / 64));
}

#define TN_INST_T(T)                                                                                           \
    template void launch_prolongate_add<T>(const MeshLevel&, const MeshLevel&, int, const T*, T*, hipStream_t); \
    template void launch_restrict<T>(const MeshLevel&, const MeshLevel&, int, const T*, T*, hipStream_t);       \
    template void launch_prolongate_linear<T>(const MeshLevel&, const MeshLevel&, int, const T*, T*, bool, bool,  \
                                              hipStream_t);                                                     \
    template void launch_restrict_linear<T>(const MeshLevel&, const MeshLevel&, int, const T*, T*, bool, hipStream_t);
TN_INST_T(double)
TN_INST_T(float)
template void launch_gca_collect<double>(const MeshLevel&, const MeshLevel&, int, const double*, double*,
                                         hipStream_t);
template void launch_gca_collect<float>(const MeshLevel&, const MeshLevel&, int, const float*, float*, hipStream_t);
template void launch_gca_select<double, double>(const MeshLevel&, int, const double*, const MeshLevel&,
                                                const double*, double*, hipStream_t);
template void launch_gca_select<double, float>(const MeshLevel&, int, const double*, const MeshLevel&,
                                               const double*, float*, hipStream_t);
template void launch_gca_select<float, double>(const MeshLevel&, int, const float*, const MeshLevel&, const float*,
                                               double*, hipStream_t);
template void launch_gca_select<float, float>(const MeshLevel&, int, const float*, const MeshLevel&, const float*,
                                              float*, hipStream_t);
#define TN_INST_GCA(T, S, SF) template void launch_gca_build<T, S, SF>(const GcaArgs<T, S, SF>&, bool, hipStream_t);
TN_INST_GCA(double, double, double)
TN_INST_GCA(double, double, float)
TN_INST_GCA(double, float, double)
TN_INST_GCA(double, float, float)
TN_INST_GCA(float, double, double)
TN_INST_GCA(float, double, float)
TN_INST_GCA(float, float, double)
TN_INST_GCA(float, float, float)
#define TN_INST_ST(T, S)                                                                   \
    template void launch_stored_apply<T, S>(const StoredArgs<T, S>&, bool, hipStream_t); \
    template void launch_stored_blk<T, S>(const StoredArgs<T, S>&, T*, hipStream_t);
TN_INST_ST(double, double)
TN_INST_ST(double, float)
TN_INST_ST(float, float)
TN_INST_ST(float, double)

}  // namespace tn
