// Argument blocks of the multigrid kernels and their host launchers.
#pragma once

#include "tn_internal.hpp"

namespace tn {

// normal-tangential basis (rows n, t1, t2), double precision: the reference
// axis is the coordinate axis least aligned with n
// (local_basis_trafo_normal_tangential.hpp:34-79)
__device__ inline void nt_basis_d(const double n[3], double t1[3], double t2[3])
{
    const double ax = fabs(n[0]), ay = fabs(n[1]), az = fabs(n[2]);
    double ref[3] = {0, 0, 0};
    if (ax <= ay && ax <= az)
        ref[0] = 1;
    else if (ay <= ax && ay <= az)
        ref[1] = 1;
    else
        ref[2] = 1;
    t1[0] = ref[1] * n[2] - ref[2] * n[1];
    t1[1] = ref[2] * n[0] - ref[0] * n[2];
    t1[2] = ref[0] * n[1] - ref[1] * n[0];
    double l = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    if (l < 1e-15) {
        t1[0] = n[1];
        t1[1] = -n[0];
        t1[2] = 0;
        l = sqrt(t1[0] * t1[0] + t1[1] * t1[1]);
    }
    for (int k = 0; k < 3; ++k) t1[k] /= l;
    t2[0] = n[1] * t1[2] - n[2] * t1[1];
    t2[1] = n[2] * t1[0] - n[0] * t1[2];
    t2[2] = n[0] * t1[1] - n[1] * t1[0];
    l = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
    for (int k = 0; k < 3; ++k) t2[k] /= l;
}

template <typename T, typename S>
struct StoredArgs {
    const S* stored;       // stored_at() layout
    const T* src;
    T* dst;
    const double* coords;  // [10][n][n][3]
    int n, nr;
    int bc_cmb, bc_surf;
    int add;     // 1: dst += A x, 0: dst = A x
    T* csum;     // [12][cells] merged corner sums (pass 1 -> pass 2)
    T* ctop;     // [12][10 (n-1)^2 chunks] chunk-top level sums
    const T* blk;        // [2][10 n n][3] boundary-node quadratic forms (launch_stored_blk), or null
    const double* pen;   // fs/fs penalty: [0..3) beta, [4..13) S, [13..16) gamma = S beta, or null
    const double* radii; // [nr] (penalty positions)
    const double* pen_coords;  // [10][n][n][3] lateral positions where the penalty is added (0: not added)
};

// boundary-node quadratic forms e^T B e of the stored operator (B: the 3x3
// diagonal block at the node copy, summed over its diamond's boundary-layer
// wedges; e = axes for Dirichlet, n / t1 / t2 for free-slip), layout of the
// matrix-free operator's forms
template <typename T, typename S>
void launch_stored_blk(const StoredArgs<T, S>& a, T* blk, hipStream_t s);

template <typename T, typename S, typename SF>
struct GcaArgs {
    int nc, nrc, nf, nrf;
    const double* geo_f;    // fine geometry table (fine level matrix-free)
    const double* radii_f;
    const T* eta_f;         // fine nodal viscosity
    const T* eta_cell_f;    // fine per-cell viscosity (overrides eta_f when set)
    const SF* stored_f; // fine packed storage (fine level stored)
    S* stored_c;        // coarse packed storage (output)
    int geometric;      // 0: the reference's radial bracket, 1: the coarse cell's own
};

constexpr int PACKED = 171;

// Stored-matrix layout.  Each wedge's 171 packed entries are padded to
// WEDGE_SLOTS = 172 slots (slot w*172 + k), cells (diamond, x, y, r order, r
// fastest) are grouped in tiles of 64, and inside a tile the slots are
// interleaved in 16-byte vectors of V = 16 / sizeof(S) slots per lane:
//   element (cell c, slot e) at ((c/64 * G + e/V) * 64 + c%64) * V + e%V,
//   G = 344 / V vectors per cell.
// A wave whose 64 lanes are 64 radially consecutive cells reads its quad's
// matrices as G coalesced 1 KB vector loads (one per lane per instruction).
constexpr int WEDGE_SLOTS = 172;
template <typename S>
__host__ __device__ constexpr int stored_vec()
{
    return 16 / (int)sizeof(S);
}
template <typename S>
__host__ __device__ inline int64_t stored_at(int64_t cell, int w, int k)
{
    constexpr int V = stored_vec<S>(), G = 2 * WEDGE_SLOTS / V;
    const int e = w * WEDGE_SLOTS + k;
    return (((cell >> 6) * G + e / V) * 64 + (cell & 63)) * V + e % V;
}
// elements to allocate for `cells` cells
inline int64_t stored_elems(int64_t cells) { return ((cells + 63) >> 6) * 64 * 2 * WEDGE_SLOTS; }

template <typename T>
void launch_prolongate_add(const MeshLevel& mc, const MeshLevel& mf, int ncomp, const T* xc, T* xf, hipStream_t s);
template <typename T>
void launch_restrict(const MeshLevel& mf, const MeshLevel& mc, int ncomp, const T* xf, T* xc, hipStream_t s);
// geometry-aware linear transfers (quirk: the Vec variants' early return)
template <typename T>
void launch_prolongate_linear(const MeshLevel& mc, const MeshLevel& mf, int ncomp, const T* xc, T* xf, bool add,
                              bool quirk, hipStream_t s);
template <typename T>
void launch_restrict_linear(const MeshLevel& mf, const MeshLevel& mc, int ncomp, const T* xf, T* xc, bool quirk,
                            hipStream_t s);
template <typename T, typename S, typename SF>
void launch_gca_build(const GcaArgs<T, S, SF>& a, bool fine_geo, hipStream_t s);
// selective GCA: flags on the coarsest mesh (zeroed, then marked) from the fine nodal coefficient k
template <typename T>
void launch_gca_collect(const MeshLevel& fine, const MeshLevel& coarsest, int levels, const T* k, T* flags,
                        hipStream_t s);
// rediscretized matrices for the cells of m whose coarsest ancestor is not flagged
template <typename T, typename S>
void launch_gca_select(const MeshLevel& coarsest, int levels, const T* flags, const MeshLevel& m, const T* eta,
                       S* stored, hipStream_t s);
// every cell's own (rediscretized) wedge matrices into the packed storage
template <typename T, typename S>
void launch_stored_own(const MeshLevel& m, const T* eta, const T* eta_cell, S* stored, hipStream_t s);
// one wedge's 18x18 Neumann matrix computed from the geometry into out[324] (device)
template <typename T>
void launch_local_matrix(const MeshLevel& m, const T* eta, const T* eta_cell, int d, int x, int y, int r, int w,
                         double* out, hipStream_t s);
// one stored wedge matrix to / from a dense device buffer of 324 doubles
template <typename S>
void launch_stored_get(const S* stored, int64_t cell, int w, double* out, hipStream_t s);
template <typename S>
void launch_stored_set(S* stored, int64_t cell, int w, const double* in, hipStream_t s);
template <typename T, typename S>
void launch_stored_apply(const StoredArgs<T, S>& a, bool diagonal, hipStream_t s);
// elements of StoredArgs::csum + ctop for an n x n x nr mesh
int64_t stored_cellsum_elems(int n, int nr);
// z = dinv * z, with 3x3 block inverses R^T diag(1/blk) R on free-slip shells
template <typename T>
void launch_fs_block_scale(const MeshLevel& m, T* z, const T* dinv, const T* blk, int bc_cmb, int bc_surf,
                           hipStream_t s);

}  // namespace tn
