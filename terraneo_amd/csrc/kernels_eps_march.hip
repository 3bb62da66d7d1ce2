// epsilon-div-div apply without global atomics: lateral march, CDNA4.
//
// Decomposition.  A block of W waves owns W-1 lateral node rows of one
// diamond and one segment of node columns.  Wave w handles the quad strip
// x = x0 - 1 + w and marches over the quads of that strip in y; its 64 lanes
// are 64 radial cells (a chunk; chunks of one column are processed one after
// the other), so every global load and store is radially contiguous.  A
// quad's contributions to its four lateral corners are routed as follows:
//   corner (x,   y)   own node row, column y   -> complete after this step
//   corner (x,   y+1) own node row, column y+1 -> carried (own LDS slots)
//   corner (x+1, y)   next wave's row          -> LDS ring, read by wave w+1
//   corner (x+1, y+1)                          -> carried, sent next step
// and radially adjacent cells share their node level through a lane shuffle.
// Waves hand off through LDS counters (written / read steps per wave), not a
// block barrier, so the waves of a SIMD drift apart by up to the ring depth
// and hide each other's load latency.
// The node level above a chunk's last cell is routed the same way by that
// lane; it is either the surface (stored) or the first level of the next
// chunk (kept in LDS until that chunk reaches the column).  Every node copy
// is written exactly once with a plain store (Replace) or one
// read-modify-write (Add).  The strip at a block edge is computed twice (once
// as the halo wave 0 of the next block) and the column before a segment once
// more: (W/(W-1)) * (S+1)/S redundant cell work instead of atomics.
//
// Boundary conditions are algebraic at the two boundary node levels, so the
// per-cell arithmetic is the Neumann formula for every lane:
//   Dirichlet  A x = P_I A_N P_I x + diag(A_N) x   on boundary nodes
//   free-slip  A x = P A_N P x + (n^T B_b n)(n . x) n
// with P_I zeroing and P projecting the boundary-node values and B_b the 3x3
// diagonal block of A_N at boundary node b (its quadratic forms precomputed
// per viscosity by k_boundary_blocks).  This is the matrix the reference
// assembles through its masks and rotations (epsilon_divdiv_kerngen.hpp:
// 823-1051, 1421-1960).  The fs/fs rotation penalty (kerngen.hpp:583-607) is
// fused into the store.
//
// Cell arithmetic: tn_cell.hpp.
// FP64 VALU operations per hex cell: 28 (corner sums) + 4 (shared) +
// 2 x 89 (wedges: 9 scalings, 9 sums, 27 gradient, 14 stress, 27 products,
// 3 kwJ) + 39 (outputs) + 15 (merge, routing) = 264.
#include <cstdlib>

#include "tn_cell.hpp"
#include "tn_eps_march.hpp"
#include "tn_march.hpp"

namespace tn {
namespace {

using namespace cellops;

// Final value of a boundary node copy from its Neumann-formula value y (of the
// zeroed / projected source) and its source value x.  blk: the copy's three quadratic forms of the
// 3x3 diagonal block B of the Neumann operator, e^T B e for e = x, y, z axes
// (Dirichlet) or e = n, t1, t2 (free-slip).  Linear in y and blk, so summing
// the copies afterwards gives the full node's value.
template <typename T, bool DIAG>
__device__ inline V3<T> test_value(V3<T> y, V3<T> x, int mode, const T* coords, int64_t lat,
                                   const T* __restrict__ blk)
{
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

template <typename T, bool DIAG, int W>
__global__ __launch_bounds__(64 * W) void k_eps_march(MarchArgs<T> a, const T* __restrict__ geo_tab, const T* __restrict__ coords,
                                                          const T* __restrict__ blk_tab,
                                                          const double* __restrict__ coords_d,
                                                          const double* __restrict__ pen)
{
    // the read-only tables indexed by the (wave-uniform) lateral node are
    // separate restrict parameters, so their loads can go through the scalar unit
    constexpr int OWN = W - 1;
    // hand-off slots per wave: a producer may lead its consumer by RING - 1 steps
    // (FP64 at 16 waves: 3, to stay within 160 KB of LDS)
    constexpr int RING = (sizeof(T) == 8 && W > 12) ? 3 : 4;
    __shared__ T ebuf[RING][W][3][65];       // corner (x+1, y) hand-off; slot 64: chunk-top level
    __shared__ int done[W], used[W];         // steps whose hand-off wave w has written / has read
    __shared__ T lvl[W][MARCH_SEG_MAX][3];   // chunk-top level values awaiting the next chunk
    // carried corner sums (x, y+1) and (x+1, y+1) of the previous quad, kept in
    // LDS (each lane reads back only its own slots) to free registers
    __shared__ T carry[W][2][3][64];
    __shared__ T carry_top[W][2][3];        // the same for the chunk-top level (one lane)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n = a.n, nr = a.nr, ncr = nr - 1, ncl = n - 1;
    const int nxb = (n + OWN - 1) / OWN, nys = (n + a.seg - 1) / a.seg;
    int b = xcd_tile(blockIdx.x, gridDim.x);
    const int ys = b % nys;
    b /= nys;
    const int xb = b % nxb;
    const int d = b / nxb;
    const int x = xb * OWN - 1 + w;
    const bool strip_ok = x >= 0 && x <= n - 2;
    const bool row_own = w >= 1 && x <= n - 1;
    const int y0 = ys * a.seg, y1 = min(y0 + a.seg, n);
    const int64_t N = 10LL * n * n * nr;
    const int bmode = a.bc_cmb, tmode = a.bc_surf;
    const int nchunks = (ncr + 63) >> 6;
    // fs/fs rotation penalty: y += gamma x p, gamma = S beta (double, as the separate penalty pass)
    double gam[3] = {0, 0, 0};
    if (pen) {
        const double* bt = pen;
        const double* S = pen + 4;
        gam[0] = S[0] * bt[0] + S[1] * bt[1] + S[2] * bt[2];
        gam[1] = S[3] * bt[0] + S[4] * bt[1] + S[5] * bt[2];
        gam[2] = S[6] * bt[0] + S[7] * bt[1] + S[8] * bt[2];
    }
    if (lane == 0) done[w] = used[w] = 0;
    __syncthreads();
    int step = 0;  // hand-off ring position, continuous over chunks

    for (int h = 0; h < nchunks; ++h) {
        const int c = 64 * h + lane;
        const bool valid = c < ncr;
        const int toplane = min(63, ncr - 1 - 64 * h);
        const bool last = h == nchunks - 1;
        const int cc = valid ? c : ncr - 1;
        const T r0 = a.radii[cc], r1 = a.radii[cc + 1];
        const T half_dr = T(0.5) * (r1 - r0), r_mid = T(0.5) * (r0 + r1);
        const T alpha = T(1) / r_mid, beta = T(1) / half_dr;
        const T kr = r_mid * r_mid * half_dr * T(1.0 / 6.0);
        const int blev = c == 0 ? bmode : TN_NEUMANN;
        const int tlev = c == ncr - 1 ? tmode : TN_NEUMANN;
        // rotation penalty: radii of this lane's node level and of the top level
        const double prad = pen ? a.radii_d[cc] : 0.0;
        const double prad_top = pen ? a.radii_d[ncr] : 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            carry[w][0][j][lane] = T(0);
            carry[w][1][j][lane] = T(0);
            if (lane == 0) carry_top[w][0][j] = carry_top[w][1][j] = T(0);
        }

        for (int qy = y0 - 1; qy <= y1 - 1; ++qy, ++step) {
            const bool quad_ok = strip_ok && qy >= 0 && qy <= n - 2;
            T(*E)[65] = ebuf[step % RING][w];
            const int64_t lat00 = ((int64_t)d * n + x) * n + qy;
            T yb[4][3], yt[4][3];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int j = 0; j < 3; ++j) yb[k][j] = yt[k][j] = T(0);
            // the source value at node (x, qy) of a boundary lane (level 0 or the top
            // level), kept from the cell's own loads for the test side of the store
            T xn[3] = {0, 0, 0};
            const bool have_xn = quad_ok && ncr > 1;
            if (quad_ok) {
                // wave-uniform: scalar loads into SGPRs
                const int goff = __builtin_amdgcn_readfirstlane(((d * ncl + x) * ncl + qy) * 2 * GEO_STRIDE);
                const T* gsrc = geo_tab + goff;
                T geo[2 * GEO_STRIDE];
#pragma unroll
                for (int i = 0; i < 2 * GEO_STRIDE; ++i) geo[i] = gsrc[i];
                // (the lane condition also keeps the compiler's load hoisting in check)
                if (valid) {
                    const int64_t latc[4] = {lat00, lat00 + n, lat00 + 1, lat00 + n + 1};
                    const T ecell = a.eta_cell ? a.eta_cell[(((int64_t)d * ncl + x) * ncl + qy) * ncr + cc] : T(0);
                    if constexpr (DIAG) {
                        CornerRaw<T> in[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int64_t ib = latc[k] * nr + cc;
                            in[k].b[0] = a.src[ib];
                            in[k].b[1] = a.src[N + ib];
                            in[k].b[2] = a.src[2 * N + ib];
                            in[k].t[0] = a.src[ib + 1];
                            in[k].t[1] = a.src[N + ib + 1];
                            in[k].t[2] = a.src[2 * N + ib + 1];
                            if (a.eta_cell) {
                                in[k].eb = in[k].et = ecell;
                            } else {
                                in[k].eb = a.eta[ib];
                                in[k].et = a.eta[ib + 1];
                            }
                        }
#pragma unroll
                        for (int j = 0; j < 3; ++j) xn[j] = c == 0 ? in[0].b[j] : in[0].t[j];
                        cell_neumann_diag<T>(in, geo, alpha, beta, kr, yb, yt);
                    } else {
                        // issue every load of the cell before any lane-divergent boundary work
                        T xb[4][3], xt[4][3], eb[4], et[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int64_t ib = latc[k] * nr + cc;
#pragma unroll
                            for (int j = 0; j < 3; ++j) {
                                xb[k][j] = a.src[j * N + ib];
                                xt[k][j] = a.src[j * N + ib + 1];
                            }
                            if (a.eta_cell) {
                                eb[k] = et[k] = ecell;
                            } else {
                                eb[k] = a.eta[ib];
                                et[k] = a.eta[ib + 1];
                            }
                        }
#pragma unroll
                        for (int j = 0; j < 3; ++j) xn[j] = c == 0 ? xb[0][j] : xt[0][j];
                        if (blev != TN_NEUMANN || tlev != TN_NEUMANN) {
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                if (blev != TN_NEUMANN) apply_trial(xb[k], blev, coords, latc[k]);
                                if (tlev != TN_NEUMANN) apply_trial(xt[k], tlev, coords, latc[k]);
                            }
                        }
                        CornerIn<T> in[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
#pragma unroll
                            for (int j = 0; j < 3; ++j) {
                                in[k].S[j] = xb[k][j] + xt[k][j];
                                in[k].d[j] = xt[k][j] - xb[k][j];
                            }
                            in[k].e = eb[k] + et[k];
                        }
                        cell_neumann<T>(in, geo, alpha, beta, kr, yb, yt);
                    }
                }
            }
            // radial merge (node level c = bottom of cell c + top of cell c-1) and routing;
            // the slot is free once the next wave has read what it held RING steps ago
            if (w + 1 < W) march::wait_counter(&used[w + 1], step - RING + 1, a.guard);
            T part[3], tp[3] = {0, 0, 0};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T M[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    T up = shfl_up1(yt[k][j]);
                    if (lane == 0) up = T(0);
                    M[k] = yb[k][j] + up;
                }
                part[j] = M[0] + carry[w][0][j][lane];
                carry[w][0][j][lane] = M[2];
                E[j][lane] = M[1] + carry[w][1][j][lane];
                carry[w][1][j][lane] = M[3];
            }
            if (lane == toplane) {  // level above this chunk's last cell
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    tp[j] = yt[0][j] + carry_top[w][0][j];
                    carry_top[w][0][j] = yt[2][j];
                    E[j][64] = yt[1][j] + carry_top[w][1][j];
                    carry_top[w][1][j] = yt[3][j];
                }
            }
            march::post_counter(&done[w], step + 1, lane);
            if (w >= 1) march::wait_counter(&done[w - 1], step + 1, a.guard);
            if (row_own && qy >= y0) {
                const T(*Ein)[65] = ebuf[step % RING][w - 1];
                const int64_t lat = lat00;
                const T* blk0 = blk_tab ? blk_tab + lat * 3 : nullptr;                   // CMB shell
                const T* blk1 = blk_tab ? blk_tab + (10LL * n * n + lat) * 3 : nullptr;  // surface shell
                const int slot = qy - y0;
                auto finish = [&](V3<T> v, int level, int mode, const T* blk, double r) {
                    const int64_t idx = lat * nr + level;
                    if (mode != TN_NEUMANN) {
                        const V3<T> xv = have_xn ? V3<T>{xn[0], xn[1], xn[2]}
                                                 : V3<T>{a.src[idx], a.src[N + idx], a.src[2 * N + idx]};
                        v = test_value<T, DIAG>(v, xv, mode, coords, lat, blk);
                    }
                    if (pen) {
                        const double px = coords_d[3 * lat] * r, py = coords_d[3 * lat + 1] * r,
                                     pz = coords_d[3 * lat + 2] * r;
                        v.x += (T)(gam[1] * pz - gam[2] * py);
                        v.y += (T)(gam[2] * px - gam[0] * pz);
                        v.z += (T)(gam[0] * py - gam[1] * px);
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
                };
                if (valid) {
                    V3<T> v{part[0] + Ein[0][lane], part[1] + Ein[1][lane], part[2] + Ein[2][lane]};
                    if (lane == 0 && h > 0) {
                        v.x += lvl[w][slot][0];
                        v.y += lvl[w][slot][1];
                        v.z += lvl[w][slot][2];
                    }
                    finish(v, c, c == 0 ? bmode : TN_NEUMANN, blk0, prad);
                }
                if (lane == toplane) {
                    const V3<T> v{tp[0] + Ein[0][64], tp[1] + Ein[1][64], tp[2] + Ein[2][64]};
                    if (last) {
                        finish(v, ncr, tmode, blk1, prad_top);
                    } else {
                        lvl[w][slot][0] = v.x;
                        lvl[w][slot][1] = v.y;
                        lvl[w][slot][2] = v.z;
                    }
                }
            }
            march::post_counter(&used[w], step + 1, lane);
        }
    }
}

// ---------------------------------------------------------------------------
// Matrix apply (not the diagonal mode) with the corner data carried along the
// march.  The quads (x, y) and (x, y+1) of a strip share the corners
// (x, y+1), (x+1, y+1), so a step loads only its two new corners and keeps
// their per-corner sums S = x_bot + x_top, d = gamma (x_top - x_bot),
// e = eta_bot + eta_top for the next step (half the loads of the form above).
// Fields are read and written through buffer resources: a radial column is
// addressed by a 32-bit wave-uniform byte offset (scalar register) plus the
// lane's byte offset, so no load or store carries per-lane address
// arithmetic, and the two node levels of a lane are one 16-byte load (FP64).
// Accesses past a resource's size return zero / are dropped by the hardware.
// The radial scalings are folded out of the wedge:
//   with a_i = l_i / r_mid, zz = z / (dr/2), G = U1 (x) a1 + U2 (x) a2 + D (x) zz
//   = G' / r_mid with G' = U1 (x) l1 + U2 (x) l2 + gamma D (x) z, gamma = r_mid / (dr/2),
//   and s = K (2 sym G' - 2/3 tr G' I), K = (sum of the six nodal eta) (dr/2)/6 |det A|
//   (= kwJ / r_mid^2), the outputs are p_i = s l_i and q = gamma s z, so the
//   bottom / top node values p_i -/+ q are single fused multiply-adds.
// FP64 VALU operations per hex cell: 20 (two new corners) + 22 (wedge inputs)
// + 2 x 75 (wedges) + 42 (outputs) + 36 (radial merge) + 9 (routing) = 279.
template <typename T>
struct CornerSum {
    T S[3], d[3], e;
};

template <typename T>
struct CornerRawCol {
    T xb[3], xt[3], eb, et;
};

// 64-bit buffer access through a resource, byte offsets (lane: voff, uniform: soff)
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// Output stores are non-temporal (cache policy nt): the result is not read by
// this kernel again, and keeping it out of L2 leaves room for the corner
// columns the neighbouring wave reads a step later (L2 fetch -5 %, time -1.5 %
// at level 7).
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 2);
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 2);
}
// buffer resource over `bytes` bytes at p (raw, unswizzled; gfx9 word 3)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

// one lane's two node levels of a radial column: cb = column byte offset
// (lat * nr * sizeof(T)), Nb = component stride in bytes, voff = lane's level byte offset
template <typename T>
__device__ __forceinline__ void load_corner(CornerRawCol<T>& r, __amdgpu_buffer_rsrc_t src,
                                            __amdgpu_buffer_rsrc_t eta, uint32_t cb, uint32_t Nb, uint32_t voff,
                                            bool nodal)
{
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        r.xb[j] = buf_ld(src, voff, cb + j * Nb, T(0));
        r.xt[j] = buf_ld(src, voff + (uint32_t)sizeof(T), cb + j * Nb, T(0));
    }
    if (nodal) {
        r.eb = buf_ld(eta, voff, cb, T(0));
        r.et = buf_ld(eta, voff + (uint32_t)sizeof(T), cb, T(0));
    } else {
        r.eb = r.et = T(0);
    }
}

template <typename T>
__device__ __forceinline__ void make_corner(CornerSum<T>& c, const CornerRawCol<T>& r, T gam)
{
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        c.S[j] = r.xb[j] + r.xt[j];
        c.d[j] = gam * (r.xt[j] - r.xb[j]);
    }
    c.e = r.eb + r.et;
}

// one wedge: s = K (2 sym G' - 2/3 tr G' I) with G' = U1 (x) l1 + U2 (x) l2 + D (x) z
// (D carries gamma); p1 = s l1, p2 = s l2, q = s z (q without gamma)
template <typename T>
__device__ __forceinline__ void wedge_folded(const T U1[3], const T U2[3], const T D[3], T K, const T* __restrict__ g,
                                             T p1[3], T p2[3], T q[3])
{
    T G[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) G[k][l] = U1[k] * g[l] + U2[k] * g[3 + l] + D[k] * g[6 + l];
    const T tr = G[0][0] + G[1][1] + G[2][2];
    const T K2 = K + K;
    const T kt = (T(2.0 / 3.0) * K) * tr;
    const T s00 = K2 * G[0][0] - kt, s11 = K2 * G[1][1] - kt, s22 = K2 * G[2][2] - kt;
    const T s01 = K * (G[0][1] + G[1][0]), s02 = K * (G[0][2] + G[2][0]), s12 = K * (G[1][2] + G[2][1]);
    p1[0] = s00 * g[0] + s01 * g[1] + s02 * g[2];
    p1[1] = s01 * g[0] + s11 * g[1] + s12 * g[2];
    p1[2] = s02 * g[0] + s12 * g[1] + s22 * g[2];
    p2[0] = s00 * g[3] + s01 * g[4] + s02 * g[5];
    p2[1] = s01 * g[3] + s11 * g[4] + s12 * g[5];
    p2[2] = s02 * g[3] + s12 * g[4] + s22 * g[5];
    q[0] = s00 * g[6] + s01 * g[7] + s02 * g[8];
    q[1] = s01 * g[6] + s11 * g[7] + s12 * g[8];
    q[2] = s02 * g[6] + s12 * g[7] + s22 * g[8];
}

// Neumann cell from its four corner sums (0 (x,y), 1 (x+1,y), 2 (x,y+1),
// 3 (x+1,y+1); wedges (0,1,2) and (3,2,1), kernel_helpers.hpp:302-327).
// hd6 = (dr/2)/6; ecell6 = 6 eta_cell, or a negative value for the nodal viscosity.
template <typename T, bool CELL>
__device__ __forceinline__ void cell_folded(const CornerSum<T>& c0, const CornerSum<T>& c1, const CornerSum<T>& c2,
                                            const CornerSum<T>& c3, const T* __restrict__ geo, T hd6, T gam,
                                            T ecell6, T yb[4][3], T yt[4][3])
{
    T d12[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d12[k] = c1.d[k] + c2.d[k];
    const T e12 = c1.e + c2.e;
    T p1a[3], p2a[3], qa[3], p1b[3], p2b[3], qb[3];
    {
        // the wedge's geometry (wave-uniform: scalar registers)
        T g[GEO_STRIDE];
#pragma unroll
        for (int i = 0; i < GEO_STRIDE; ++i) g[i] = geo[i];
        T U1[3], U2[3], D[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            U1[k] = c1.S[k] - c0.S[k];
            U2[k] = c2.S[k] - c0.S[k];
            D[k] = c0.d[k] + d12[k];
        }
        const T es = CELL ? ecell6 : c0.e + e12;
        wedge_folded(U1, U2, D, es * (hd6 * g[9]), g, p1a, p2a, qa);
    }
    // the second wedge's geometry is read only now, so the two wedges'
    // tables do not occupy scalar registers at the same time
    asm volatile("" ::: "memory");
    {
        T g[GEO_STRIDE];
#pragma unroll
        for (int i = 0; i < GEO_STRIDE; ++i) g[i] = geo[GEO_STRIDE + i];
        T U1[3], U2[3], D[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            U1[k] = c2.S[k] - c3.S[k];
            U2[k] = c1.S[k] - c3.S[k];
            D[k] = c3.d[k] + d12[k];
        }
        const T es = CELL ? ecell6 : c3.e + e12;
        wedge_folded(U1, U2, D, es * (hd6 * g[9]), g, p1b, p2b, qb);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T t0 = p1a[k] + p2a[k];
        yb[0][k] = -t0 - gam * qa[k];
        yt[0][k] = gam * qa[k] - t0;
        const T t3 = p1b[k] + p2b[k];
        yb[3][k] = -t3 - gam * qb[k];
        yt[3][k] = gam * qb[k] - t3;
        const T Q = qa[k] + qb[k];
        const T P1 = p1a[k] + p2b[k];
        yb[1][k] = P1 - gam * Q;
        yt[1][k] = P1 + gam * Q;
        const T P2 = p2a[k] + p1b[k];
        yb[2][k] = P2 - gam * Q;
        yt[2][k] = P2 + gam * Q;
    }
}

// lane i receives lane i-1's value, lane 0 receives 0 (DPP wavefront shift right by one)
__device__ __forceinline__ float shift_up1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double shift_up1(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// CFG bits: 1 a shell is not Neumann (boundary lanes), 2 fs/fs rotation
// penalty, 4 per-cell viscosity; each case is compiled on its own so the
// common ones carry no dead state in registers
template <typename T, int W, int CFG>
__global__ __launch_bounds__(64 * W) void k_eps_flow(MarchArgs<T> a, const T* __restrict__ geo_tab,
                                                     const T* __restrict__ coords, const T* __restrict__ blk_tab,
                                                     const double* __restrict__ coords_d,
                                                     const double* __restrict__ pen)
{
    constexpr bool BND = (CFG & 1) != 0, PEN = (CFG & 2) != 0, CELL = (CFG & 4) != 0;
    // PF: a step's two new corners are loaded during the step before (their
    // latency is then covered by that step's arithmetic; more registers)
    constexpr bool PF = (CFG & 8) != 0;
    constexpr int OWN = W - 1;
    constexpr int RING = (sizeof(T) == 8 && W > 12) ? 3 : 4;
    constexpr uint32_t SZ = sizeof(T);
    __shared__ T ebuf[RING][W][3][65];       // corner (x+1, y) hand-off; slot 64: chunk-top level
    __shared__ int done[W], used[W];
    __shared__ T lvl[W][MARCH_SEG_MAX][3];   // chunk-top level values awaiting the next chunk
    __shared__ T carry[W][2][3][64];         // carried output sums (x, y+1), (x+1, y+1)
    __shared__ T carry_top[W][2][3];
    __shared__ T xkeep[W][2][2][3];          // raw boundary-level source of node (x, y): [parity][cmb / surface]
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n = a.n, nr = a.nr, ncr = nr - 1, ncl = n - 1;
    const int nxb = (n + OWN - 1) / OWN, nys = (n + a.seg - 1) / a.seg;
    int b = xcd_tile(blockIdx.x, gridDim.x);
    const int ys = b % nys;
    b /= nys;
    const int xb = b % nxb;
    const int d = b / nxb;
    const int x = xb * OWN - 1 + w;
    const bool strip_ok = x >= 0 && x <= n - 2;
    const bool row_own = w >= 1 && x <= n - 1;
    const int y0 = ys * a.seg, y1 = min(y0 + a.seg, n);
    const uint32_t Nn = 10u * n * n * nr;    // nodes per component
    const uint32_t Nb = Nn * SZ;
    const __amdgpu_buffer_rsrc_t rs_src = buf_rsrc(a.src, 3 * Nb);
    const __amdgpu_buffer_rsrc_t rs_dst = buf_rsrc(a.dst, 3 * Nb);
    constexpr bool nodal = !CELL;
    const __amdgpu_buffer_rsrc_t rs_eta = buf_rsrc(nodal ? (const void*)a.eta : (const void*)a.src, nodal ? Nb : 0u);
    const int bmode = a.bc_cmb, tmode = a.bc_surf;
    const int nchunks = (ncr + 63) >> 6;
    // byte offsets of radial columns: node (xc, yc) of this diamond at ((d n + xc) n + yc) nr
    const uint32_t colstep = (uint32_t)nr * SZ;             // next y
    const uint32_t rowstep = (uint32_t)n * nr * SZ;         // next x
    const uint32_t colx = (uint32_t)((d * n + x) * n) * nr * SZ;  // column (x, 0) (x >= -1: wraps for x = -1, unused)
    if (lane == 0) done[w] = used[w] = 0;
    __syncthreads();
    int step = 0;

    for (int h = 0; h < nchunks; ++h) {
        const int c = 64 * h + lane;
        const bool valid = c < ncr;
        const int toplane = min(63, ncr - 1 - 64 * h);
        const bool last = h == nchunks - 1;
        const int cc = valid ? c : ncr - 1;
        const uint32_t voff = (uint32_t)cc * SZ;
        const T r0 = a.radii[cc], r1 = a.radii[cc + 1];
        const T half_dr = T(0.5) * (r1 - r0), r_mid = T(0.5) * (r0 + r1);
        const T gam = r_mid / half_dr, hd6 = half_dr * T(1.0 / 6.0);
        // radius of the lane's node level and of the chunk's top level (fs/fs penalty, FP64)
        const double rad_lane = PEN ? a.radii_d[cc] : 0.0;
        const double rad_top = PEN ? a.radii_d[64 * h + toplane + 1] : 0.0;
        // boundary lanes of this chunk (trial side on load, test side at the store)
        const bool bot_b = BND && h == 0 && bmode != TN_NEUMANN && lane == 0;
        const bool top_b = BND && last && tmode != TN_NEUMANN && c == ncr - 1;
        const bool any_b = BND && ((h == 0 && bmode != TN_NEUMANN) || (last && tmode != TN_NEUMANN));
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            carry[w][0][j][lane] = T(0);
            carry[w][1][j][lane] = T(0);
            if (lane == 0) carry_top[w][0][j] = carry_top[w][1][j] = T(0);
        }
        // corner data: raw columns -> sums (trial rule of the boundary levels applied)
        auto process = [&](CornerSum<T>& cs, CornerRawCol<T>& r, int xc, int yc, bool keep) {
            if (any_b) {
                const int64_t lat = ((int64_t)d * n + xc) * n + yc;
                if (keep) {
                    if (bot_b) {
#pragma unroll
                        for (int j = 0; j < 3; ++j) xkeep[w][yc & 1][0][j] = r.xb[j];
                    }
                    if (top_b) {
#pragma unroll
                        for (int j = 0; j < 3; ++j) xkeep[w][yc & 1][1][j] = r.xt[j];
                    }
                }
                if (bot_b) apply_trial(r.xb, bmode, coords, lat);
                if (top_b) apply_trial(r.xt, tmode, coords, lat);
            }
            make_corner(cs, r, gam);
        };
        // Steps of this segment: quads qy = qs .. y1 - 1 of the strip, with
        // qs = y0 - 1 (the column before the segment, computed again for its
        // carried sums) or 0 when the segment starts the diamond (the step
        // before it would carry zeros only; every wave of the block has the
        // same y0, so all of them skip it).  Quads exist for qy <= n - 2; the
        // step qy = n - 1 only finishes the last column.  Waves whose strip
        // has no quads (x = -1, x = n - 1) take empty steps throughout.
        const int qs = y0 == 0 ? 0 : y0 - 1;
        const int goff0 = ((d * ncl + x) * ncl + qs) * 2 * GEO_STRIDE;
        const int step0 = step;
        const int qe = min(y1 - 1, n - 2);  // last quad of the segment
        CornerSum<T> C0, C1, N0, N1;
        CornerRawCol<T> R0, R1;  // loaded corner columns (with PF: those of the next step)
        if (strip_ok) {
            const uint32_t cb = colx + (uint32_t)qs * colstep;
            load_corner(R0, rs_src, rs_eta, cb, Nb, voff, nodal);
            load_corner(R1, rs_src, rs_eta, cb + rowstep, Nb, voff, nodal);
            process(C0, R0, x, qs, true);
            process(C1, R1, x + 1, qs, false);
            if (PF && qs <= qe) {
                load_corner(R0, rs_src, rs_eta, cb + colstep, Nb, voff, nodal);
                load_corner(R1, rs_src, rs_eta, cb + colstep + rowstep, Nb, voff, nodal);
            }
        }

        // one step: quad (x, qy) from its corners (C0, C1) at qy and (N0, N1) at
        // qy + 1, which are loaded here (quad = true); quad = false: an empty
        // step (no quad, nothing loaded).  The loop below alternates the roles
        // of the two pairs, so no corner data is copied between steps.
        auto body = [&](const int qy, CornerSum<T>& C0, CornerSum<T>& C1, CornerSum<T>& N0, CornerSum<T>& N1,
                        const bool quad) __attribute__((always_inline)) {
            const int step = step0 + (qy - qs);
            T(*E)[65] = ebuf[step % RING][w];
            T yb[4][3], yt[4][3];
            if (quad) {
                const uint32_t cb = colx + (uint32_t)(qy + 1) * colstep;
                if (!PF) {
                    load_corner(R0, rs_src, rs_eta, cb, Nb, voff, nodal);
                    load_corner(R1, rs_src, rs_eta, cb + rowstep, Nb, voff, nodal);
                }
                process(N0, R0, x, qy + 1, true);
                process(N1, R1, x + 1, qy + 1, false);
                if (PF && qy + 1 <= qe) {
                    load_corner(R0, rs_src, rs_eta, cb + colstep, Nb, voff, nodal);
                    load_corner(R1, rs_src, rs_eta, cb + colstep + rowstep, Nb, voff, nodal);
                }
                const T* gsrc = geo_tab + __builtin_amdgcn_readfirstlane(goff0 + (qy - qs) * 2 * GEO_STRIDE);
                const T ecell6 = CELL ? T(6) * a.eta_cell[(((int64_t)d * ncl + x) * ncl + qy) * ncr + cc] : T(0);
                cell_folded<T, CELL>(C0, C1, N0, N1, gsrc, hd6, gam, ecell6, yb, yt);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int j = 0; j < 3; ++j) yb[k][j] = yt[k][j] = T(0);
            }
            // fs/fs penalty y += gamma x p = r (gamma x c): the wave-uniform vector
            // w = gamma x c for the node (x, qy) stored by this step (c its unit
            // position), formed after the cell arithmetic, whose geometry holds the
            // scalar registers; the merge, the LDS routing and the hand-off waits
            // hide the loads' latency.  gamma = S beta is re-read from the moment
            // pass's finalize output (a scalar-cache hit) rather than kept for the
            // whole kernel.
            double pw0 = 0.0, pw1 = 0.0, pw2 = 0.0;
            if (PEN && row_own && qy >= y0) {
                asm volatile("" ::: "memory");
                const int o = __builtin_amdgcn_readfirstlane(3 * ((d * n + x) * n + qy));
                const double c0 = coords_d[o], c1 = coords_d[o + 1], c2 = coords_d[o + 2];
                const double g0 = pen[13], g1 = pen[14], g2 = pen[15];
                pw0 = g1 * c2 - g2 * c1;
                pw1 = g2 * c0 - g0 * c2;
                pw2 = g0 * c1 - g1 * c0;
            }
            if (w + 1 < W) march::wait_counter(&used[w + 1], step - RING + 1, a.guard);
            T part[3], tp[3] = {0, 0, 0};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T M[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) M[k] = yb[k][j] + shift_up1(yt[k][j]);
                part[j] = M[0] + carry[w][0][j][lane];
                carry[w][0][j][lane] = M[2];
                E[j][lane] = M[1] + carry[w][1][j][lane];
                carry[w][1][j][lane] = M[3];
            }
            if (lane == toplane) {
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    tp[j] = yt[0][j] + carry_top[w][0][j];
                    carry_top[w][0][j] = yt[2][j];
                    E[j][64] = yt[1][j] + carry_top[w][1][j];
                    carry_top[w][1][j] = yt[3][j];
                }
            }
            march::post_counter(&done[w], step + 1, lane);
            if (w >= 1) march::wait_counter(&done[w - 1], step + 1, a.guard);
            if (row_own && qy >= y0) {
                const T(*Ein)[65] = ebuf[step % RING][w - 1];
                const uint32_t cb = colx + (uint32_t)qy * colstep;
                const int slot = qy - y0;
                auto finish = [&](V3<T> v, int level, int mode, int side) {
                    if ((BND && mode != TN_NEUMANN) || PEN) {
                        const int64_t lat = ((int64_t)d * n + x) * n + qy;
                        if (BND && mode != TN_NEUMANN) {
                            // the node's raw source: kept when its column was loaded as a corner;
                            // the last node row (no quads in this wave's strip) reads it
                            V3<T> xv;
                            if (strip_ok) {
                                const T* xk = xkeep[w][qy & 1][side];
                                xv = {xk[0], xk[1], xk[2]};
                            } else {
                                const uint32_t lo = (uint32_t)level * SZ;
                                xv = {buf_ld(rs_src, lo, cb, T(0)), buf_ld(rs_src, lo, cb + Nb, T(0)),
                                      buf_ld(rs_src, lo, cb + 2 * Nb, T(0))};
                            }
                            const T* blk = blk_tab + ((int64_t)side * 10 * n * n + lat) * 3;
                            v = test_value<T, false>(v, xv, mode, coords, lat, blk);
                        }
                        if (PEN) {
                            const double r = level == c ? rad_lane : rad_top;
                            v.x += (T)(r * pw0);
                            v.y += (T)(r * pw1);
                            v.z += (T)(r * pw2);
                        }
                    }
                    const uint32_t lo = (uint32_t)level * SZ;
                    if (a.add) {
                        v.x += buf_ld(rs_dst, lo, cb, T(0));
                        v.y += buf_ld(rs_dst, lo, cb + Nb, T(0));
                        v.z += buf_ld(rs_dst, lo, cb + 2 * Nb, T(0));
                    }
                    buf_st(v.x, rs_dst, lo, cb);
                    buf_st(v.y, rs_dst, lo, cb + Nb);
                    buf_st(v.z, rs_dst, lo, cb + 2 * Nb);
                };
                if (valid) {
                    V3<T> v{part[0] + Ein[0][lane], part[1] + Ein[1][lane], part[2] + Ein[2][lane]};
                    if (lane == 0 && h > 0) {
                        v.x += lvl[w][slot][0];
                        v.y += lvl[w][slot][1];
                        v.z += lvl[w][slot][2];
                    }
                    finish(v, c, c == 0 ? bmode : TN_NEUMANN, 0);
                }
                if (lane == toplane) {
                    const V3<T> v{tp[0] + Ein[0][64], tp[1] + Ein[1][64], tp[2] + Ein[2][64]};
                    if (last) {
                        finish(v, ncr, tmode, 1);
                    } else {
                        lvl[w][slot][0] = v.x;
                        lvl[w][slot][1] = v.y;
                        lvl[w][slot][2] = v.z;
                    }
                }
            }
            march::post_counter(&used[w], step + 1, lane);
        };
        if (strip_ok) {
            int qy = qs;
            for (; qy + 1 <= qe; qy += 2) {
                body(qy, C0, C1, N0, N1, true);
                body(qy + 1, N0, N1, C0, C1, true);
            }
            if (qy <= qe) body(qy++, C0, C1, N0, N1, true);
            if (qy <= y1 - 1) body(qy, C0, C1, N0, N1, false);  // qy = n - 1
        } else {
            for (int qy = qs; qy <= y1 - 1; ++qy) body(qy, C0, C1, N0, N1, false);
        }
        step += y1 - qs;
    }
}

// Quadratic forms of the 3x3 diagonal block of the Neumann operator at every
// boundary node copy, from the copy's own diamond's cells (gather over the up
// to four adjacent quads, no atomics):
//   B = sum_wedges kwJ (|g|^2 I + g g^T / 3),  blk = (e^T B e) for the basis of
//   the shell's condition (axes for Dirichlet, n/t1/t2 for free-slip).
template <typename T>
__global__ void k_boundary_blocks(Geom<T> g, const T* __restrict__ eta, const T* __restrict__ eta_cell, int bc_cmb,
                                  int bc_surf, T* __restrict__ blk)
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
                T ks = 0, kt = 0;
                if (eta_cell) {
                    const T ec = eta_cell[(((int64_t)d * ncl + qx) * ncl + qy) * ncr + rc];
                    ks = T(3) * ec;
                    kt = T(3) * ec;
                } else {
                    for (int j = 0; j < 3; ++j) {
                        const int64_t ib = latc[WC[w][j]] * g.nr + rc;
                        ks += eta[ib];
                    }
                    for (int j = 0; j < 3; ++j) {
                        const int64_t ib = latc[WC[w][j]] * g.nr + rc + 1;
                        kt += eta[ib];
                    }
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
void launch_boundary_blocks(const Geom<T>& g, const T* eta, const T* eta_cell, int bc_cmb, int bc_surf, T* blk,
                            hipStream_t s)
{
    const int64_t total = 2LL * 10 * g.n * g.n;
    hipLaunchKernelGGL(k_boundary_blocks<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, eta, eta_cell,
                       bc_cmb, bc_surf, blk);
    TN_HIP(hipGetLastError());
}

// Waves per block (W-1 node rows per block), one block per CU (LDS).  With
// the counter hand-off the kernel needs 115 (FP64) / 72 (FP32) VGPRs; measured
// at level 7 x 129 (Dirichlet apply): FP64 0.665 ms at W = 16 against 0.725 ms
// at 12, FP32 0.422 ms at W = 12 against 0.489 ms at 16.  More waves per block
// mean a smaller share of redundant halo strips (W/(W-1)).  TN_MARCH_WAVES
// overrides (tuning: 4, 6, 8, 12, 16).
template <typename T>
int march_waves()
{
    static const int w = [] {
        const char* e = std::getenv("TN_MARCH_WAVES");
        const int v = e ? std::atoi(e) : 0;
        if (v == 4 || v == 6 || v == 8 || v == 12 || v == 16) return v;
        return sizeof(T) == 8 ? 16 : 12;
    }();
    return w;
}

int march_block_rows(int dtype_size) { return (dtype_size == 8 ? march_waves<double>() : march_waves<float>()) - 1; }

int march_waves_per_cu(int dtype_size) { return march_block_rows(dtype_size) + 1; }

// matrix-apply kernel variant: 0 the per-quad-load march, 1 the carried-corner
// march, 2 the same with each step's loads issued a step ahead (TN_EPS_KERNEL
// overrides; tuning)
int eps_kernel_variant()
{
    static const int v = [] {
        const char* e = std::getenv("TN_EPS_KERNEL");
        const int k = e ? std::atoi(e) : 1;
        return (k >= 0 && k <= 2) ? k : 1;
    }();
    return v;
}

template <typename T, int W>
void launch_march_w(const MarchArgs<T>& a, bool diagonal, hipStream_t s)
{
    constexpr int OWN = W - 1;
    const int nxb = (a.n + OWN - 1) / OWN, nys = (a.n + a.seg - 1) / a.seg;
    const dim3 gr((unsigned)(10 * nxb * nys)), bl(64 * W);
    // the carried-corner kernel is built for 8, 12 and 16 waves per block
    const int var = W >= 8 ? eps_kernel_variant() : 0;
    if (diagonal) {
        hipLaunchKernelGGL((k_eps_march<T, true, W>), gr, bl, 0, s, a, a.geo, a.coords, a.blk, a.coords_d, a.pen);
    } else if (var == 0) {
        hipLaunchKernelGGL((k_eps_march<T, false, W>), gr, bl, 0, s, a, a.geo, a.coords, a.blk, a.coords_d, a.pen);
    } else {
        if constexpr (W >= 8) {
            const int cfg = ((a.bc_cmb != TN_NEUMANN || a.bc_surf != TN_NEUMANN) ? 1 : 0) | (a.pen ? 2 : 0) |
                            (a.eta_cell ? 4 : 0) | (var == 2 ? 8 : 0);
            TN_REQUIRE(3ULL * 10 * a.n * a.n * a.nr * sizeof(T) < (1ULL << 32),
                       "apply: field larger than 4 GB (buffer resources address 32-bit byte offsets)");
            switch (cfg) {
#define TN_FLOW(C)                                                                                                \
    case C:                                                                                                      \
        hipLaunchKernelGGL((k_eps_flow<T, W, C>), gr, bl, 0, s, a, a.geo, a.coords, a.blk, a.coords_d, a.pen); \
        break;
                TN_FLOW(0) TN_FLOW(1) TN_FLOW(2) TN_FLOW(3) TN_FLOW(4) TN_FLOW(5) TN_FLOW(6) TN_FLOW(7)
                TN_FLOW(8) TN_FLOW(9) TN_FLOW(10) TN_FLOW(11) TN_FLOW(12) TN_FLOW(13) TN_FLOW(14) TN_FLOW(15)
#undef TN_FLOW
            }
        }
    }
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_eps_march(const MarchArgs<T>& a, bool diagonal, hipStream_t s)
{
    TN_REQUIRE(a.nr - 1 >= 2, "march kernel: at least 2 radial cells");
    TN_REQUIRE(a.seg >= 1 && a.seg <= MARCH_SEG_MAX, "march kernel: segment length out of range");
    TN_REQUIRE(10LL * (a.n - 1) * (a.n - 1) * 2 * GEO_STRIDE < (1LL << 31), "march kernel: mesh too large");
    // at lateral level 7 x 129 the kernel measured 0.64 ms (FP64, 16 waves)
    // and 0.42 ms (FP32, 12 waves); an LDS-tiled variant that staged the
    // source columns in LDS with a barrier per step was slower (0.98 / 0.50 ms
    // against 0.78 / 0.49 ms for the barrier form of this kernel) and is gone
    switch (march_waves<T>()) {
        case 4: launch_march_w<T, 4>(a, diagonal, s); break;
        case 6: launch_march_w<T, 6>(a, diagonal, s); break;
        case 8: launch_march_w<T, 8>(a, diagonal, s); break;
        case 12: launch_march_w<T, 12>(a, diagonal, s); break;
        default: launch_march_w<T, 16>(a, diagonal, s); break;
            break;
    }
}

template void launch_boundary_blocks<double>(const Geom<double>&, const double*, const double*, int, int, double*,
                                             hipStream_t);
template void launch_boundary_blocks<float>(const Geom<float>&, const float*, const float*, int, int, float*,
                                            hipStream_t);
template void launch_eps_march<double>(const MarchArgs<double>&, bool, hipStream_t);
template void launch_eps_march<float>(const MarchArgs<float>&, bool, hipStream_t);

}  // namespace tn
