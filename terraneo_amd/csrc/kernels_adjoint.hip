// Adjoint viscosity gradient of the epsilon-div-div operator, CDNA4.
//
// With a per-hex-cell viscosity eta_c = exp(m_c) the viscous block is
// A(m) = sum_c eta_c A_c(1), A_c(1) the cell's two wedge matrices at unit
// viscosity.  For a misfit J(u(m)) and the adjoint state lam of the Stokes
// system (K lam = dJ/du, K symmetric),
//   dJ/dm_c = -lam_c^T (dK/dm_c) u_c = -eta_c (T lam)_c^T A_c(1) u_c,
// T the trial rule of the boundary shells (Dirichlet: zero, free-slip: project
// onto the tangent plane); u is the converged solution itself (its Dirichlet
// values are the boundary data, its free-slip normal component vanishes), so
// the boundary diagonal and normal terms of K drop out.  There is no
// counterpart in the reference.
//
// One thread per hex cell, radial index fastest (radially contiguous loads).
// A_c(1) u_c is the Neumann cell arithmetic of the apply kernel with unit
// viscosity; the result is dotted with lam at the cell's eight nodes.
#include "tn_cell.hpp"
#include "tn_objects.hpp"

namespace tn {
namespace {

using namespace cellops;

template <typename T>
__global__ __launch_bounds__(256) void k_grad_log_cell_eta(Geom<T> g, const T* __restrict__ eta_cell,
                                                           const T* __restrict__ u, const T* __restrict__ lam,
                                                           int bc_cmb, int bc_surf, T* __restrict__ grad,
                                                           int64_t total)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= total) return;
    const int ncr = g.nr - 1, ncl = g.n - 1;
    const int rc = (int)(tid % ncr);
    int64_t rest = tid / ncr;
    const int y = (int)(rest % ncl);
    rest /= ncl;
    const int x = (int)(rest % ncl);
    const int d = (int)(rest / ncl);
    const int64_t N = 10LL * g.n * g.n * g.nr;
    const int64_t lat00 = ((int64_t)d * g.n + x) * g.n + y;
    const int64_t latc[4] = {lat00, lat00 + g.n, lat00 + 1, lat00 + g.n + 1};
    const int bmode = rc == 0 ? bc_cmb : TN_NEUMANN;
    const int tmode = rc == ncr - 1 ? bc_surf : TN_NEUMANN;

    CornerIn<T> in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t ib = latc[k] * g.nr + rc;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T xb = u[j * N + ib], xt = u[j * N + ib + 1];
            in[k].S[j] = xb + xt;
            in[k].d[j] = xt - xb;
        }
        in[k].e = T(2);  // unit viscosity at both nodes
    }
    const T r0 = g.radii[rc], r1 = g.radii[rc + 1];
    const T half_dr = T(0.5) * (r1 - r0), r_mid = T(0.5) * (r0 + r1);
    const T kr = r_mid * r_mid * half_dr * T(1.0 / 6.0);
    const T* geo = g.geo + ((((int64_t)d * ncl + x) * ncl + y) * 2) * GEO_STRIDE;
    T yb[4][3], yt[4][3];
    cell_neumann<T>(in, geo, T(1) / r_mid, T(1) / half_dr, kr, yb, yt);

    T acc = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t ib = latc[k] * g.nr + rc;
        T lb[3], lt[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            lb[j] = lam[j * N + ib];
            lt[j] = lam[j * N + ib + 1];
        }
        if (bmode != TN_NEUMANN) apply_trial(lb, bmode, g.coords, latc[k]);
        if (tmode != TN_NEUMANN) apply_trial(lt, tmode, g.coords, latc[k]);
#pragma unroll
        for (int j = 0; j < 3; ++j) acc += lb[j] * yb[k][j] + lt[j] * yt[k][j];
    }
    grad[tid] = -eta_cell[tid] * acc;
}

}  // namespace

template <typename T>
void launch_grad_log_cell_eta(const Geom<T>& g, const T* eta_cell, const T* u, const T* lam, int bc_cmb, int bc_surf,
                              T* grad, hipStream_t s)
{
    const int64_t total = 10LL * (g.n - 1) * (g.n - 1) * (g.nr - 1);
    const int64_t blocks = (total + 255) / 256;
    TN_REQUIRE(blocks < (1LL << 31), "grid too large");
    hipLaunchKernelGGL(k_grad_log_cell_eta<T>, dim3((unsigned)blocks), dim3(256), 0, s, g, eta_cell, u, lam, bc_cmb,
                       bc_surf, grad, total);
    TN_HIP(hipGetLastError());
}

template void launch_grad_log_cell_eta<double>(const Geom<double>&, const double*, const double*, const double*, int,
                                               int, double*, hipStream_t);
template void launch_grad_log_cell_eta<float>(const Geom<float>&, const float*, const float*, const float*, int, int,
                                              float*, hipStream_t);

}  // namespace tn
