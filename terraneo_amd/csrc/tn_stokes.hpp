// Argument blocks of the Stokes coupling / mass kernels.
#pragma once

#include "tn_eps_march.hpp"
#include "tn_internal.hpp"

namespace tn {

template <typename T>
struct DivArgs {
    const double* coords;  // fine lateral coords
    const double* radii;   // fine radii
    int nf, nrf, nc, nrc;
    int bc_cmb, bc_surf;
    const T* u;  // fine velocity [3][Nf]
    T* p;        // coarse pressure [Nc] (every copy written)
    T* csum;     // scratch, div_scratch_elems(nc, nrc): per-cell corner sums, then the chunk tops
};
// scratch elements of B: 4 corners x (coarse cells + one top level per 64-cell chunk)
inline int64_t div_scratch_elems(int nc, int nrc)
{
    const int64_t ncol = 10LL * (nc - 1) * (nc - 1);
    return 4 * ncol * ((int64_t)(nrc - 1) + (nrc - 1 + 63) / 64);
}

template <typename T>
struct GradArgs {
    const double* coords;
    const double* radii;
    int nf, nrf, nc, nrc;
    int bc_cmb, bc_surf;
    const T* p;  // coarse pressure
    T* u;        // fine velocity
    int seg;     // node columns per block of the march
    int add;     // 1: u += B^T p, 0: u = B^T p
    MarchGuard guard;
};

// waves per block of the B^T march (node rows per block + 1): FP64 12 (one
// block per CU at 3 waves per SIMD, <= 168 VGPRs), FP32 8 (two blocks per CU)
template <typename T>
constexpr int grad_march_waves()
{
    return sizeof(T) == 8 ? 12 : 8;
}

template <typename T>
struct MassArgs {
    const double* coords;
    const double* radii;
    int n, nr, ncomp;
    const T* x;  // input (vector mass)
    const T* k;  // coefficient (lumped k-mass)
    T* y;        // output: written (vector mass), accumulated (lumped k-mass)
    const double* geo = nullptr;  // the mesh's wedge table [10][n-1][n-1][2][GEO_STRIDE] (vector mass)
};

template <typename T>
void launch_div(const DivArgs<T>& a, hipStream_t s);
template <typename T>
void launch_grad(const GradArgs<T>& a, hipStream_t s);
template <typename T>
void launch_mass(const MassArgs<T>& a, bool lumped_k, hipStream_t s);

}  // namespace tn
