// Atomic-free epsilon-div-div apply (lateral march), launch interface.
#pragma once

#include "tn_internal.hpp"

namespace tn {

constexpr int MARCH_SEG_MAX = 64; // node columns per block segment, at most

// XCD-aware tile order.  The dispatcher hands workgroup b to XCD b % 8, and
// each XCD has its own L2.  Tile t (segments fastest, then row blocks, then
// diamonds) is given to workgroup b such that every XCD runs one contiguous
// run of tiles: neighbouring tiles, which share halo rows and columns, then
// share an L2.  A bijection of [0, nblocks).
__device__ __forceinline__ int xcd_tile(int b, int nblocks)
{
    constexpr int XCDS = 8;
    const int xcd = b % XCDS, local = b / XCDS;
    const int per = nblocks / XCDS, rem = nblocks % XCDS;
    return xcd * per + min(xcd, rem) + local;
}

// Bounded hand-off waits of the march kernels.  A wait that polls `limit`
// times without seeing its neighbour's post gives up; the wave then stores 1
// to *fail, a host-visible word, and finishes normally (its output is not to
// be trusted).  The host reports the word through check_march_status().
struct MarchGuard {
    int* fail;  // device address of the host-visible word
    int limit;  // polls per wait
};
// host: the process's guard (the word is allocated on first use)
MarchGuard march_guard();
// host: throws tn::Error if a march wait gave up since the last check (and clears the word)
void check_march_status();

template <typename T>
struct MarchArgs {
    const T* src;             // [3][N]
    T* dst;                   // [3][N]
    const T* eta;             // [N] nodal viscosity
    const T* eta_cell;        // [10][n-1][n-1][nr-1] per-cell viscosity, or null (nodal)
    const T* geo;             // [10][n-1][n-1][2][GEO_STRIDE]
    const T* coords;          // [10][n][n][3] unit lateral coordinates
    const T* radii;           // [nr]
    const double* coords_d;   // FP64 lateral positions where the rotation penalty is added (0: not added)
    const double* radii_d;
    const T* blk;             // [2][10 n n][3] boundary quadratic forms, or null (both shells Neumann)
    const double* pen;        // fs/fs penalty: [0..3) beta, [4..13) S, [13..16) gamma = S beta, or null
    int n, nr;
    int seg;                  // node columns per block in the march direction
    int bc_cmb, bc_surf;
    int add;                  // 1: dst += A x, 0: dst = A x
    MarchGuard guard;
};

// node rows owned per block (waves per block - 1) for an element size in bytes
int march_block_rows(int dtype_size);
// resident waves per CU at the kernel's register use (element size in bytes)
int march_waves_per_cu(int dtype_size);
// segment length of a march kernel (own node rows per block, resident waves per CU)
int march_seg_for(int n, int own, int waves_per_cu);

// boundary-node quadratic forms e^T B e (see k_boundary_blocks)
template <typename T>
void launch_boundary_blocks(const Geom<T>& g, const T* eta, const T* eta_cell, int bc_cmb, int bc_surf, T* blk,
                            hipStream_t s);

template <typename T>
void launch_eps_march(const MarchArgs<T>& a, bool diagonal, hipStream_t s);

}  // namespace tn
