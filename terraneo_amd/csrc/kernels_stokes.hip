// Stokes coupling and mass kernels on wedges.
//
//   B   (Divergence, divergence.hpp:117-323):  p_c += -int div(u_h) q_H
//   B^T (Gradient,   gradient.hpp:115-325):    u   += -int p_H div(v_h)
//   VectorMass (vector_mass.hpp:83-153), lumped KMass (kmass.hpp:97-185)
//
// Velocity lives on the fine mesh, pressure on the mesh one level coarser;
// each fine wedge integrates with the 6-point Felippa 3x2 rule and the coarse
// (P1 on the pressure mesh) shape functions evaluated on the fine wedge
// (integrands.hpp:103-243).  The coarse dofs a fine wedge couples to are
// those of coarse cell (x/2, y/2, r/2) in the node pattern of the FINE wedge
// index, exactly as the reference scatters them (kernel_helpers.hpp:396-424);
// for the two centre fine wedges this is the neighbouring coarse wedge.
//
// Instead of forming the 6x18 element matrix, the apply works per quadrature
// point: B uses div u_h(q) = sum_j grad N_j(q) . u_j, B^T uses
// p_H(q) = sum_i N^H_i(q) p_i.  Dirichlet layers drop their velocity dofs,
// free-slip layers project them tangentially (trial side for B, test side
// for B^T), which is the masked / rotated element matrix of the reference.
#include "tn_internal.hpp"
#include "tn_stokes.hpp"

namespace tn {
namespace {

__device__ constexpr int WOFF[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                         {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};
__device__ constexpr double Q6P[6][3] = {{0.6666666666666666, 0.1666666666666667, -0.5773502691896257},
                                        {0.1666666666666667, 0.6666666666666666, -0.5773502691896257},
                                        {0.1666666666666667, 0.1666666666666667, -0.5773502691896257},
                                        {0.6666666666666666, 0.1666666666666667, 0.5773502691896257},
                                        {0.1666666666666667, 0.6666666666666666, 0.5773502691896257},
                                        {0.1666666666666667, 0.1666666666666667, 0.5773502691896257}};
__device__ constexpr double Q6W = 0.1666666666666667;

__device__ inline double shape_lat(int k, double xi, double eta)
{
    const int m = k % 3;
    return m == 0 ? 1.0 - xi - eta : (m == 1 ? xi : eta);
}
__device__ inline double shape_rad(int k, double zeta) { return k < 3 ? 0.5 * (1 - zeta) : 0.5 * (1 + zeta); }

__device__ inline void grad_shape(int k, const double q[3], double g[3])
{
    const int m = k % 3;
    const double nr = shape_rad(k, q[2]);
    g[0] = (m == 0 ? -1.0 : (m == 1 ? 1.0 : 0.0)) * nr;
    g[1] = (m == 0 ? -1.0 : (m == 2 ? 1.0 : 0.0)) * nr;
    g[2] = shape_lat(k, q[0], q[1]) * (k < 3 ? -0.5 : 0.5);
}

// coarse shape i of the wedge containing the fine wedge (frw: fine radial
// index r%2, flw: fine lateral index 0..3), integrands.hpp:103-190
__device__ inline double shape_coarse(int i, int frw, int flw, const double q[3])
{
    const double xi = q[0], eta = q[1], z = q[2];
    double lat;
    switch (i % 3) {
    case 0: lat = flw == 0 ? 1.0 - 0.5 * (xi + eta) : (flw == 3 ? 0.5 * (xi + eta) : 0.5 - 0.5 * (xi + eta)); break;
    case 1: lat = flw == 0 ? 0.5 * xi : (flw == 1 ? 0.5 * xi + 0.5 : (flw == 2 ? 0.5 * xi : 0.5 - 0.5 * xi)); break;
    default: lat = flw == 0 ? 0.5 * eta : (flw == 1 ? 0.5 * eta : (flw == 2 ? 0.5 * eta + 0.5 : 0.5 - 0.5 * eta)); break;
    }
    const double rad = i < 3 ? (frw == 0 ? 0.25 * (3 - z) : 0.25 * (1 - z)) : (frw == 0 ? 0.25 * (1 + z) : 0.25 * (3 + z));
    return lat * rad;
}

__device__ inline int fine_lateral_wedge_idx(int x, int y, int w)
{
    // kernel_helpers.hpp:600-608
    constexpr int ind[2][2][2] = {{{0, 1}, {2, 3}}, {{3, 2}, {1, 0}}};
    return ind[w][y & 1][x & 1];
}

struct WedgeGeo {
    double P[3][3];  // lateral vertices (unit sphere)
    double r1, r2;
};

// J^{-T} grad N_j at quad point q and |det J| (integrands.hpp:623-670)
__device__ inline double wedge_qp(const WedgeGeo& G, const double q[3], double g[6][3])
{
    const double r = G.r1 + 0.5 * (G.r2 - G.r1) * (1 + q[2]);
    const double gr = 0.5 * (G.r2 - G.r1);
    double J[3][3];
    for (int a = 0; a < 3; ++a) {
        J[a][0] = (G.P[1][a] - G.P[0][a]) * r;
        J[a][1] = (G.P[2][a] - G.P[0][a]) * r;
        J[a][2] = ((1 - q[0] - q[1]) * G.P[0][a] + q[0] * G.P[1][a] + q[1] * G.P[2][a]) * gr;
    }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    const double K[3][3] = {{c00 * id, c01 * id, c02 * id},
                            {(J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
                             (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id},
                            {(J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id,
                             (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
    for (int j = 0; j < 6; ++j) {
        double gr3[3];
        grad_shape(j, q, gr3);
        for (int a = 0; a < 3; ++a) g[j][a] = K[a][0] * gr3[0] + K[a][1] * gr3[1] + K[a][2] * gr3[2];
    }
    return fabs(det);
}

__device__ inline void load_wedge(const double* coords, const double* radii, int n, int d, int x, int y, int rc, int w,
                                  WedgeGeo& G)
{
    for (int k = 0; k < 3; ++k) {
        const double* c = coords + 3 * (((int64_t)d * n + x + WOFF[w][k][0]) * n + y + WOFF[w][k][1]);
        G.P[k][0] = c[0];
        G.P[k][1] = c[1];
        G.P[k][2] = c[2];
    }
    G.r1 = radii[rc];
    G.r2 = radii[rc + 1];
}

__device__ inline void unit_normal(const double* coords, int64_t lat, double nn[3])
{
    const double* c = coords + 3 * lat;
    const double inv = 1.0 / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    nn[0] = c[0] * inv;
    nn[1] = c[1] * inv;
    nn[2] = c[2] * inv;
}

// B: one thread per fine cell, atomics into the coarse pressure field.
template <typename T>
__global__ __launch_bounds__(128) void k_div(DivArgs<T> a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ncr = a.nrf - 1, ncl = a.nf - 1;
    const int64_t ncell = 10LL * ncl * ncl * ncr;
    if (tid >= ncell) return;
    const int rc = (int)(tid % ncr);
    int64_t lat = tid / ncr;
    const int y = (int)(lat % ncl);
    lat /= ncl;
    const int x = (int)(lat % ncl);
    const int d = (int)(lat / ncl);
    const int64_t Nf = 10LL * a.nf * a.nf * a.nrf;
    const int bmode = rc == 0 ? a.bc_cmb : TN_NEUMANN;
    const int tmode = rc == ncr - 1 ? a.bc_surf : TN_NEUMANN;
    const int frw = rc & 1;
    for (int w = 0; w < 2; ++w) {
        WedgeGeo G;
        load_wedge(a.coords, a.radii, a.nf, d, x, y, rc, w, G);
        double u[6][3];
        for (int j = 0; j < 6; ++j) {
            const int64_t lj = ((int64_t)d * a.nf + x + WOFF[w][j][0]) * a.nf + y + WOFF[w][j][1];
            const int64_t ij = lj * a.nrf + rc + WOFF[w][j][2];
            for (int c = 0; c < 3; ++c) u[j][c] = (double)a.u[c * Nf + ij];
            const int mode = j < 3 ? bmode : tmode;
            if (mode == TN_DIRICHLET) {
                u[j][0] = u[j][1] = u[j][2] = 0.0;
            } else if (mode == TN_FREESLIP) {
                double nn[3];
                unit_normal(a.coords, lj, nn);
                const double un = nn[0] * u[j][0] + nn[1] * u[j][1] + nn[2] * u[j][2];
                for (int c = 0; c < 3; ++c) u[j][c] -= un * nn[c];
            }
        }
        const int flw = fine_lateral_wedge_idx(x, y, w);
        double out[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 6; ++q) {
            double g[6][3];
            const double adet = wedge_qp(G, Q6P[q], g);
            double div = 0.0;
            for (int j = 0; j < 6; ++j) div += g[j][0] * u[j][0] + g[j][1] * u[j][1] + g[j][2] * u[j][2];
            const double f = -Q6W * adet * div;
            for (int i = 0; i < 6; ++i) out[i] += f * shape_coarse(i, frw, flw, Q6P[q]);
        }
        for (int i = 0; i < 6; ++i) {
            const int64_t ic = (((int64_t)d * a.nc + x / 2 + WOFF[w][i][0]) * a.nc + y / 2 + WOFF[w][i][1]) * a.nrc +
                               rc / 2 + WOFF[w][i][2];
            unsafeAtomicAdd(a.p + ic, (T)out[i]);
        }
    }
}

// B^T: one thread per fine cell, atomics into the fine velocity field.
template <typename T>
__global__ __launch_bounds__(128) void k_grad(GradArgs<T> a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ncr = a.nrf - 1, ncl = a.nf - 1;
    const int64_t ncell = 10LL * ncl * ncl * ncr;
    if (tid >= ncell) return;
    const int rc = (int)(tid % ncr);
    int64_t lat = tid / ncr;
    const int y = (int)(lat % ncl);
    lat /= ncl;
    const int x = (int)(lat % ncl);
    const int d = (int)(lat / ncl);
    const int64_t Nf = 10LL * a.nf * a.nf * a.nrf;
    const int bmode = rc == 0 ? a.bc_cmb : TN_NEUMANN;
    const int tmode = rc == ncr - 1 ? a.bc_surf : TN_NEUMANN;
    const int frw = rc & 1;
    for (int w = 0; w < 2; ++w) {
        WedgeGeo G;
        load_wedge(a.coords, a.radii, a.nf, d, x, y, rc, w, G);
        double pc[6];
        for (int i = 0; i < 6; ++i) {
            const int64_t ic = (((int64_t)d * a.nc + x / 2 + WOFF[w][i][0]) * a.nc + y / 2 + WOFF[w][i][1]) * a.nrc +
                               rc / 2 + WOFF[w][i][2];
            pc[i] = (double)a.p[ic];
        }
        const int flw = fine_lateral_wedge_idx(x, y, w);
        double out[6][3] = {{0}};
        for (int q = 0; q < 6; ++q) {
            double g[6][3];
            const double adet = wedge_qp(G, Q6P[q], g);
            double ph = 0.0;
            for (int i = 0; i < 6; ++i) ph += shape_coarse(i, frw, flw, Q6P[q]) * pc[i];
            const double f = -Q6W * adet * ph;
            for (int j = 0; j < 6; ++j)
                for (int c = 0; c < 3; ++c) out[j][c] += f * g[j][c];
        }
        for (int j = 0; j < 6; ++j) {
            const int mode = j < 3 ? bmode : tmode;
            if (mode == TN_DIRICHLET) continue;
            const int64_t lj = ((int64_t)d * a.nf + x + WOFF[w][j][0]) * a.nf + y + WOFF[w][j][1];
            if (mode == TN_FREESLIP) {
                double nn[3];
                unit_normal(a.coords, lj, nn);
                const double on = nn[0] * out[j][0] + nn[1] * out[j][1] + nn[2] * out[j][2];
                for (int c = 0; c < 3; ++c) out[j][c] -= on * nn[c];
            }
            const int64_t ij = lj * a.nrf + rc + WOFF[w][j][2];
            for (int c = 0; c < 3; ++c) unsafeAtomicAdd(a.u + c * Nf + ij, (T)out[j][c]);
        }
    }
}

// VectorMass (ncomp components) / lumped k-weighted mass (LUMPED_K).
template <typename T, bool LUMPED_K>
__global__ __launch_bounds__(128) void k_mass(MassArgs<T> a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ncr = a.nr - 1, ncl = a.n - 1;
    const int64_t ncell = 10LL * ncl * ncl * ncr;
    if (tid >= ncell) return;
    const int rc = (int)(tid % ncr);
    int64_t lat = tid / ncr;
    const int y = (int)(lat % ncl);
    lat /= ncl;
    const int x = (int)(lat % ncl);
    const int d = (int)(lat / ncl);
    const int64_t N = 10LL * a.n * a.n * a.nr;
    for (int w = 0; w < 2; ++w) {
        WedgeGeo G;
        load_wedge(a.coords, a.radii, a.n, d, x, y, rc, w, G);
        int64_t idx[6];
        for (int j = 0; j < 6; ++j)
            idx[j] = (((int64_t)d * a.n + x + WOFF[w][j][0]) * a.n + y + WOFF[w][j][1]) * a.nr + rc + WOFF[w][j][2];
        const double gr = 0.5 * (G.r2 - G.r1);
        double M[6][6] = {{0}};
        for (int q = 0; q < 6; ++q) {
            const double* qp = Q6P[q];
            double Jl[3][3];  // lateral Jacobian (kernel_helpers.hpp:153-176)
            for (int c = 0; c < 3; ++c) {
                Jl[c][0] = G.P[1][c] - G.P[0][c];
                Jl[c][1] = G.P[2][c] - G.P[0][c];
                Jl[c][2] = (1 - qp[0] - qp[1]) * G.P[0][c] + qp[0] * G.P[1][c] + qp[1] * G.P[2][c];
            }
            const double det = fabs(Jl[0][0] * (Jl[1][1] * Jl[2][2] - Jl[1][2] * Jl[2][1]) -
                                    Jl[0][1] * (Jl[1][0] * Jl[2][2] - Jl[1][2] * Jl[2][0]) +
                                    Jl[0][2] * (Jl[1][0] * Jl[2][1] - Jl[1][1] * Jl[2][0]));
            const double rr = G.r1 + 0.5 * (G.r2 - G.r1) * (1 + qp[2]);
            double N6[6];
            for (int j = 0; j < 6; ++j) N6[j] = shape_lat(j, qp[0], qp[1]) * shape_rad(j, qp[2]);
            double kq = 1.0;
            if (LUMPED_K) {
                kq = 0.0;
                for (int j = 0; j < 6; ++j) kq += N6[j] * (double)a.k[idx[j]];
            }
            const double f = Q6W * kq * rr * rr * gr * det;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) M[i][j] += f * N6[i] * N6[j];
        }
        if (LUMPED_K) {
            for (int i = 0; i < 6; ++i) {
                double s = 0.0;
                for (int j = 0; j < 6; ++j) s += M[i][j];
                unsafeAtomicAdd(a.y + idx[i], (T)s);
            }
        } else {
            for (int c = 0; c < a.ncomp; ++c)
                for (int i = 0; i < 6; ++i) {
                    double s = 0.0;
                    for (int j = 0; j < 6; ++j) s += M[i][j] * (double)a.x[c * N + idx[j]];
                    unsafeAtomicAdd(a.y + c * N + idx[i], (T)s);
                }
        }
    }
}

}  // namespace

template <typename T>
void launch_div(const DivArgs<T>& a, hipStream_t s)
{
    const int64_t ncell = 10LL * (a.nf - 1) * (a.nf - 1) * (a.nrf - 1);
    hipLaunchKernelGGL(k_div<T>, dim3((unsigned)((ncell + 127) / 128)), dim3(128), 0, s, a);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_grad(const GradArgs<T>& a, hipStream_t s)
{
    const int64_t ncell = 10LL * (a.nf - 1) * (a.nf - 1) * (a.nrf - 1);
    hipLaunchKernelGGL(k_grad<T>, dim3((unsigned)((ncell + 127) / 128)), dim3(128), 0, s, a);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_mass(const MassArgs<T>& a, bool lumped_k, hipStream_t s)
{
    const int64_t ncell = 10LL * (a.n - 1) * (a.n - 1) * (a.nr - 1);
    const dim3 gr((unsigned)((ncell + 127) / 128)), bl(128);
    if (lumped_k)
        hipLaunchKernelGGL((k_mass<T, true>), gr, bl, 0, s, a);
    else
        hipLaunchKernelGGL((k_mass<T, false>), gr, bl, 0, s, a);
    TN_HIP(hipGetLastError());
}

template void launch_div<double>(const DivArgs<double>&, hipStream_t);
template void launch_div<float>(const DivArgs<float>&, hipStream_t);
template void launch_grad<double>(const GradArgs<double>&, hipStream_t);
template void launch_grad<float>(const GradArgs<float>&, hipStream_t);
template void launch_mass<double>(const MassArgs<double>&, bool, hipStream_t);
template void launch_mass<float>(const MassArgs<float>&, bool, hipStream_t);

}  // namespace tn
