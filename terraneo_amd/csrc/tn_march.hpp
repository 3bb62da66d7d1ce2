// Atomic-free lateral march over the hex cells of a diamond, shared by the
// kernels whose per-cell work produces sums at the cell's 8 corners (stored-
// matrix apply, B^T).  Device-only header.
//
// Decomposition.  A block of W waves owns W-1 lateral node rows of one
// diamond and one segment of `seg` node columns.  Wave w handles the quad
// strip x = x0 - 1 + w and marches over the quads of that strip in y; its 64
// lanes are 64 radial cells (a chunk; chunks of one column follow one after
// the other).  A quad's sums at its four lateral corners are routed
//   corner (x,   y)   own node row, column y   -> complete after this step
//   corner (x,   y+1) own node row, column y+1 -> carried in registers
//   corner (x+1, y)   next wave's row          -> LDS ring, read by wave w+1
//                                                 (LDS counters order producer and consumer)
//   corner (x+1, y+1)                          -> carried, sent next step
// and radially adjacent cells share their node level through a lane shuffle;
// the level above a chunk's last cell goes the same way through that lane and
// is either the top shell (finished) or the first level of the next chunk
// (kept in LDS until that chunk reaches the column).  Every node copy of the
// diamond is finished exactly once; the strip at a block edge is computed
// twice (as the halo wave 0 of the next block).
//
// The per-cell work is a functor Op with
//   void begin_chunk(int c, int cc, bool valid)      lane's radial cell c (cc: clamped)
//   void cell(int d, int x, int y, T yb[4][3], T yt[4][3])
//        adds the quad's corner sums (k = 0 (x,y), 1 (x+1,y), 2 (x,y+1),
//        3 (x+1,y+1); yb at node level c, yt at c+1); called for valid quads
//        only, lanes past the last cell must add zeros
//   void finish(int64_t lat, int level, V3<T> v)     the node copy's final sum
#pragma once

#include "tn_cell.hpp"
#include "tn_eps_march.hpp"

namespace tn {
namespace march {

using cellops::V3;

// number of blocks of a launch over a mesh of n x n lateral nodes
__host__ __device__ inline int blocks_for(int n, int seg, int waves)
{
    const int own = waves - 1;
    return 10 * ((n + own - 1) / own) * ((n + seg - 1) / seg);
}

// Point-to-point hand-off between neighbouring waves through LDS counters
// instead of a block barrier per step: the waves of a SIMD drift apart and hide
// each other's memory latency.  The waves of a workgroup are co-resident and
// every wave posts both counters at every step, so each wait is satisfied
// after at most RING - 1 steps of a neighbour; the bound (g.limit polls, by
// default 2^20, about 40 ms at 2.4 GHz) only guarantees that a wait ends
// whatever happens.  A wait that gives up flags it in the host-visible word
// g.fail (a vector store; the host turns it into an error).
__device__ __forceinline__ void wait_counter(int* c, int target, const MarchGuard& g)
{
    for (int i = 0; i < g.limit; ++i) {
        if (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) return;
        __builtin_amdgcn_s_sleep(1);
    }
    if (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) return;
    if (g.fail) __hip_atomic_store(g.fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void post_counter(int* c, int value, int lane)
{
    if (lane == 0) __hip_atomic_store(c, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, int W, class Op>
__device__ __forceinline__ void run(Op& op, int n, int nr, int seg, const MarchGuard& guard)
{
    constexpr int OWN = W - 1;
    constexpr int RING = 4;                  // hand-off slots: a producer may lead by RING - 1 steps
    __shared__ T ebuf[RING][W][3][65];       // corner (x+1, y) hand-off; slot 64: chunk-top level
    __shared__ T lvl[W][MARCH_SEG_MAX][3];   // chunk-top level values awaiting the next chunk
    __shared__ int done[W], used[W];         // steps whose hand-off wave w has written / has read
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ncr = nr - 1;
    const int nxb = (n + OWN - 1) / OWN, nys = (n + seg - 1) / seg;
    int b = xcd_tile(blockIdx.x, gridDim.x);
    const int ys = b % nys;
    b /= nys;
    const int xb = b % nxb;
    const int d = b / nxb;
    const int x = xb * OWN - 1 + w;
    const bool strip_ok = x >= 0 && x <= n - 2;
    const bool row_own = w >= 1 && x <= n - 1;
    const int y0 = ys * seg, y1 = min(y0 + seg, n);
    const int nchunks = (ncr + 63) >> 6;
    if (lane == 0) done[w] = used[w] = 0;
    __syncthreads();
    int step = 0;  // hand-off ring position, continuous over chunks

    for (int h = 0; h < nchunks; ++h) {
        const int c = 64 * h + lane;
        const bool valid = c < ncr;
        const int toplane = min(63, ncr - 1 - 64 * h);
        const bool last = h == nchunks - 1;
        op.begin_chunk(c, valid ? c : ncr - 1, valid);
        T c01[3] = {0, 0, 0}, c11[3] = {0, 0, 0}, c01t[3] = {0, 0, 0}, c11t[3] = {0, 0, 0};

        for (int qy = y0 - 1; qy <= y1 - 1; ++qy, ++step) {
            const bool quad_ok = strip_ok && qy >= 0 && qy <= n - 2;
            T(*E)[65] = ebuf[step % RING][w];
            T yb[4][3], yt[4][3];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int j = 0; j < 3; ++j) yb[k][j] = yt[k][j] = T(0);
            if (quad_ok) op.cell(d, x, qy, yb, yt);
            // radial merge (node level c = bottom of cell c + top of cell c-1) and routing;
            // the slot is free once the next wave has read what it held RING steps ago
            if (w + 1 < W) wait_counter(&used[w + 1], step - RING + 1, guard);
            T part[3], tp[3] = {0, 0, 0};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T Mg[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    T up = __shfl_up(yt[k][j], 1, 64);
                    if (lane == 0) up = T(0);
                    Mg[k] = yb[k][j] + up;
                }
                part[j] = Mg[0] + c01[j];
                c01[j] = Mg[2];
                E[j][lane] = Mg[1] + c11[j];
                c11[j] = Mg[3];
            }
            if (lane == toplane) {  // level above this chunk's last cell
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    tp[j] = yt[0][j] + c01t[j];
                    c01t[j] = yt[2][j];
                    E[j][64] = yt[1][j] + c11t[j];
                    c11t[j] = yt[3][j];
                }
            }
            post_counter(&done[w], step + 1, lane);
            if (w >= 1) wait_counter(&done[w - 1], step + 1, guard);
            if (row_own && qy >= y0) {
                const T(*Ein)[65] = ebuf[step % RING][w - 1];
                const int64_t lat = ((int64_t)d * n + x) * n + qy;
                const int slot = qy - y0;
                if (valid) {
                    V3<T> v{part[0] + Ein[0][lane], part[1] + Ein[1][lane], part[2] + Ein[2][lane]};
                    if (lane == 0 && h > 0) {
                        v.x += lvl[w][slot][0];
                        v.y += lvl[w][slot][1];
                        v.z += lvl[w][slot][2];
                    }
                    op.finish(lat, c, v);
                }
                if (lane == toplane) {
                    const V3<T> v{tp[0] + Ein[0][64], tp[1] + Ein[1][64], tp[2] + Ein[2][64]};
                    if (last) {
                        op.finish(lat, ncr, v);
                    } else {
                        lvl[w][slot][0] = v.x;
                        lvl[w][slot][1] = v.y;
                        lvl[w][slot][2] = v.z;
                    }
                }
            }
            post_counter(&used[w], step + 1, lane);
        }
    }
}

}  // namespace march
}  // namespace tn
