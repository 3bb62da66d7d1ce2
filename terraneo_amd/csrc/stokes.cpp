// C ABI: Stokes block operator, mass operators, grid transfers and the
// boundary-condition helpers used to build right-hand sides.
#include <cstring>
#include <memory>

#include "tn_eps_march.hpp"
#include "tn_mg.hpp"
#include "tn_objects.hpp"
#include "tn_stokes.hpp"

namespace tn {
void set_last_error(const std::string& msg);
}

#define TN_API_BEGIN try {
#define TN_API_END                    \
    }                                 \
    catch (const std::exception& e)   \
    {                                 \
        tn::set_last_error(e.what()); \
        return -1;                    \
    }                                 \
    return 0;

using tn::MeshLevel;

namespace tn {

template <typename T>
void div_t(tn_stokes* K, const void* u, void* p, hipStream_t s)
{
    auto& mf = const_cast<tn_mesh*>(K->vel)->lvl.dev();
    auto& mc = const_cast<tn_mesh*>(K->pre)->lvl.dev();
    TN_REQUIRE(K->div_tmp.n == sizeof(T) * (size_t)div_scratch_elems(mc.n, mc.nr), "B: scratch not allocated");
    DivArgs<T> a{mf.d_coords.p, mf.d_radii.p, mf.n,  mf.nr,         mc.n,         mc.nr,
                 K->A->bc_cmb,  K->A->bc_surf, (const T*)u, (T*)p, (T*)K->div_tmp.p};
    launch_div<T>(a, s);
    launch_interface_sum<T>(mc, 1, (T*)p, s);
}

template <typename T>
void grad_t(tn_stokes* K, const void* p, void* u, bool add, bool comm, hipStream_t s)
{
    auto& mf = const_cast<tn_mesh*>(K->vel)->lvl.dev();
    auto& mc = const_cast<tn_mesh*>(K->pre)->lvl.dev();
    constexpr int W = grad_march_waves<T>();
    const int seg = march_seg_for(mf.n, W - 1, sizeof(T) == 8 ? W : 2 * W);
    GradArgs<T> a{mf.d_coords.p, mf.d_radii.p, mf.n,          mf.nr,        mc.n, mc.nr, K->A->bc_cmb,
                  K->A->bc_surf, (const T*)p,   (T*)u, seg, add ? 1 : 0, march_guard()};
    launch_grad<T>(a, s);
    if (comm) launch_interface_sum<T>(mf, 3, (T*)u, s);
}

// EpsDivDivStokes::apply_impl (epsilon_divdiv_stokes.hpp:55-84):
// A in Replace + skipped communication, B^T in Add + additive communication
// (which also completes A's partial sums), B in Replace + communication.
void stokes_apply(tn_stokes* K, const void* u, const void* p, void* yu, void* yp, hipStream_t s)
{
    tn_eps* A = K->A;
    const int am = A->apply_mode, cm = A->comm_mode;
    const bool dg = A->diagonal;
    A->apply_mode = TN_REPLACE;
    A->comm_mode = TN_COMM_SKIP;
    A->diagonal = false;
    try {
        eps_apply(A, u, yu, s);
    } catch (...) {
        A->apply_mode = am;
        A->comm_mode = cm;
        A->diagonal = dg;
        throw;
    }
    A->apply_mode = am;
    A->comm_mode = cm;
    A->diagonal = dg;
    if (A->dtype == TN_F64) {
        grad_t<double>(K, p, yu, true, true, s);
        div_t<double>(K, u, yp, s);
    } else {
        grad_t<float>(K, p, yu, true, true, s);
        div_t<float>(K, u, yp, s);
    }
}

template void grad_t<double>(tn_stokes*, const void*, void*, bool, bool, hipStream_t);
template void grad_t<float>(tn_stokes*, const void*, void*, bool, bool, hipStream_t);

}  // namespace tn

namespace {

template <typename T>
__global__ void k_boundary_assign(T* dst, const T* src, int64_t N, int nr, int ncomp, int which)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int r = (int)(i % nr);
    if (!(((which & 1) && r == 0) || ((which & 2) && r == nr - 1))) return;
    for (int c = 0; c < ncomp; ++c) dst[c * N + i] = src ? src[c * N + i] : T(0);
}

// strong_algebraic_freeslip_enforcement_in_place
// (strong_algebraic_freeslip_enforcement.hpp:21-37): remove the normal
// component of v on the selected shells
template <typename T>
__global__ void k_freeslip_project(T* v, const double* coords, int64_t N, int nr, int which)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int r = (int)(i % nr);
    if (!(((which & 1) && r == 0) || ((which & 2) && r == nr - 1))) return;
    const double* c = coords + 3 * (i / nr);
    const double inv = 1.0 / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double n0 = c[0] * inv, n1 = c[1] * inv, n2 = c[2] * inv;
    const double un = n0 * (double)v[i] + n1 * (double)v[N + i] + n2 * (double)v[2 * N + i];
    v[i] = (T)((double)v[i] - un * n0);
    v[N + i] = (T)((double)v[N + i] - un * n1);
    v[2 * N + i] = (T)((double)v[2 * N + i] - un * n2);
}

// RHSVelocityInterpolator (mantlecirculation.cpp:117-145): the nodal
// buoyancy force Ra T n at every node copy, n the unit position vector
template <typename T>
__global__ void k_buoyancy(T* f, const T* temp, const double* coords, int64_t N, int nr, double rayleigh)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double* c = coords + 3 * (i / nr);
    const double inv = 1.0 / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double a = rayleigh * (double)temp[i] * inv;
    f[i] = (T)(a * c[0]);
    f[N + i] = (T)(a * c[1]);
    f[2 * N + i] = (T)(a * c[2]);
}

}  // namespace

namespace tn {
template <typename T>
void launch_freeslip_project(const MeshLevel& m, T* v, int which, hipStream_t s)
{
    const unsigned g = (unsigned)((m.N + 255) / 256);
    hipLaunchKernelGGL(k_freeslip_project<T>, dim3(g), dim3(256), 0, s, v, m.d_coords.p, m.N, m.nr, which);
    TN_HIP(hipGetLastError());
}
template void launch_freeslip_project<double>(const MeshLevel&, double*, int, hipStream_t);
template void launch_freeslip_project<float>(const MeshLevel&, float*, int, hipStream_t);
}  // namespace tn

extern "C" {

int tn_stokes_create(const tn_mesh* vel, const tn_mesh* pre, tn_eps* A, tn_stokes** out)
{
    TN_API_BEGIN
    TN_REQUIRE(vel && pre && A && out, "null pointer");
    TN_REQUIRE(A->mesh == vel, "viscous operator must live on the velocity mesh");
    TN_REQUIRE(2 * (pre->lvl.n - 1) == vel->lvl.n - 1 && 2 * (pre->lvl.nr - 1) == vel->lvl.nr - 1,
               "pressure mesh must be one uniform coarsening of the velocity mesh");
    auto K = std::make_unique<tn_stokes>();
    K->vel = vel;
    K->pre = pre;
    K->A = A;
    // B's scratch (per-cell corner sums) is allocated here, not in the apply path
    const size_t esize = A->dtype == TN_F64 ? sizeof(double) : sizeof(float);
    K->div_tmp.alloc(esize * (size_t)tn::div_scratch_elems(pre->lvl.n, pre->lvl.nr));
    *out = K.release();
    TN_API_END
}

int tn_stokes_destroy(tn_stokes* K)
{
    TN_API_BEGIN
    delete K;
    TN_API_END
}

int tn_stokes_apply(tn_stokes* K, const void* u, const void* p, void* yu, void* yp, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(K && u && p && yu && yp, "null pointer");
    tn::check_march_status();
    tn::stokes_apply(K, u, p, yu, yp, (hipStream_t)s);
    TN_API_END
}

int tn_div_apply(tn_stokes* K, const void* u, void* p, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(K && u && p, "null pointer");
    if (K->A->dtype == TN_F64)
        tn::div_t<double>(K, u, p, (hipStream_t)s);
    else
        tn::div_t<float>(K, u, p, (hipStream_t)s);
    TN_API_END
}

int tn_grad_apply(tn_stokes* K, const void* p, void* u, int apply_mode, int comm_mode, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(K && u && p, "null pointer");
    tn::check_march_status();
    const bool add = apply_mode == TN_ADD, comm = comm_mode == TN_COMM_ADDITIVE;
    if (K->A->dtype == TN_F64)
        tn::grad_t<double>(K, p, u, add, comm, (hipStream_t)s);
    else
        tn::grad_t<float>(K, p, u, add, comm, (hipStream_t)s);
    TN_API_END
}

int tn_vector_mass_apply(const tn_mesh* m, int dtype, const void* x, void* y, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && x && y, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    auto go = [&](auto* yy) {
        using T = std::remove_pointer_t<decltype(yy)>;
        tn::launch_fill<T>(yy, 3 * L.N, T(0), (hipStream_t)s);
        tn::MassArgs<T> a{L.d_coords.p, L.d_radii.p, L.n, L.nr, 3, (const T*)x, nullptr, yy, L.d_geo.p};
        tn::launch_mass<T>(a, false, (hipStream_t)s);
        tn::launch_interface_sum<T>(L, 3, yy, (hipStream_t)s);
    };
    if (dtype == TN_F64)
        go((double*)y);
    else
        go((float*)y);
    TN_API_END
}

int tn_kmass_lumped(const tn_mesh* m, int dtype, const void* k, void* d, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && k && d, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    auto go = [&](auto* dd) {
        using T = std::remove_pointer_t<decltype(dd)>;
        tn::launch_fill<T>(dd, L.N, T(0), (hipStream_t)s);
        tn::MassArgs<T> a{L.d_coords.p, L.d_radii.p, L.n, L.nr, 1, nullptr, (const T*)k, dd};
        tn::launch_mass<T>(a, true, (hipStream_t)s);
        tn::launch_interface_sum<T>(L, 1, dd, (hipStream_t)s);
    };
    if (dtype == TN_F64)
        go((double*)d);
    else
        go((float*)d);
    TN_API_END
}

int tn_prolongate_add(const tn_mesh* coarse, const tn_mesh* fine, int dtype, int ncomp, const void* xc, void* xf,
                      tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(coarse && fine && xc && xf, "null pointer");
    auto& mc = const_cast<tn_mesh*>(coarse)->lvl.dev();
    auto& mf = const_cast<tn_mesh*>(fine)->lvl.dev();
    TN_REQUIRE(2 * (mc.n - 1) == mf.n - 1 && 2 * (mc.nr - 1) == mf.nr - 1, "meshes are not a refinement pair");
    if (dtype == TN_F64)
        tn::launch_prolongate_add<double>(mc, mf, ncomp, (const double*)xc, (double*)xf, (hipStream_t)s);
    else
        tn::launch_prolongate_add<float>(mc, mf, ncomp, (const float*)xc, (float*)xf, (hipStream_t)s);
    TN_API_END
}

int tn_restrict(const tn_mesh* fine, const tn_mesh* coarse, int dtype, int ncomp, const void* xf, void* xc,
                tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(coarse && fine && xc && xf, "null pointer");
    auto& mc = const_cast<tn_mesh*>(coarse)->lvl.dev();
    auto& mf = const_cast<tn_mesh*>(fine)->lvl.dev();
    TN_REQUIRE(2 * (mc.n - 1) == mf.n - 1 && 2 * (mc.nr - 1) == mf.nr - 1, "meshes are not a refinement pair");
    if (dtype == TN_F64)
        tn::launch_restrict<double>(mf, mc, ncomp, (const double*)xf, (double*)xc, (hipStream_t)s);
    else
        tn::launch_restrict<float>(mf, mc, ncomp, (const float*)xf, (float*)xc, (hipStream_t)s);
    TN_API_END
}

int tn_prolongate_linear(const tn_mesh* coarse, const tn_mesh* fine, int dtype, int ncomp, const void* xc, void* xf,
                         int apply_mode, int vec_early_return, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(coarse && fine && xc && xf, "null pointer");
    TN_REQUIRE(ncomp == 1 || ncomp == 3, "linear prolongation: ncomp must be 1 or 3");
    auto& mc = const_cast<tn_mesh*>(coarse)->lvl.dev();
    auto& mf = const_cast<tn_mesh*>(fine)->lvl.dev();
    TN_REQUIRE(2 * (mc.n - 1) == mf.n - 1 && 2 * (mc.nr - 1) == mf.nr - 1, "meshes are not a refinement pair");
    const bool add = apply_mode == TN_ADD, quirk = vec_early_return != 0 && ncomp == 3;
    if (dtype == TN_F64)
        tn::launch_prolongate_linear<double>(mc, mf, ncomp, (const double*)xc, (double*)xf, add, quirk,
                                             (hipStream_t)s);
    else
        tn::launch_prolongate_linear<float>(mc, mf, ncomp, (const float*)xc, (float*)xf, add, quirk, (hipStream_t)s);
    TN_API_END
}

int tn_restrict_linear(const tn_mesh* fine, const tn_mesh* coarse, int dtype, int ncomp, const void* xf, void* xc,
                       int vec_early_return, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(coarse && fine && xc && xf, "null pointer");
    TN_REQUIRE(ncomp == 1 || ncomp == 3, "linear restriction: ncomp must be 1 or 3");
    auto& mc = const_cast<tn_mesh*>(coarse)->lvl.dev();
    auto& mf = const_cast<tn_mesh*>(fine)->lvl.dev();
    TN_REQUIRE(2 * (mc.n - 1) == mf.n - 1 && 2 * (mc.nr - 1) == mf.nr - 1, "meshes are not a refinement pair");
    const bool quirk = vec_early_return != 0 && ncomp == 3;
    if (dtype == TN_F64)
        tn::launch_restrict_linear<double>(mf, mc, ncomp, (const double*)xf, (double*)xc, quirk, (hipStream_t)s);
    else
        tn::launch_restrict_linear<float>(mf, mc, ncomp, (const float*)xf, (float*)xc, quirk, (hipStream_t)s);
    TN_API_END
}

int tn_boundary_assign(const tn_mesh* m, int dtype, int ncomp, void* dst, const void* src, int which, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && dst, "null pointer");
    const auto& L = m->lvl;
    const unsigned g = (unsigned)((L.N + 255) / 256);
    if (dtype == TN_F64)
        hipLaunchKernelGGL(k_boundary_assign<double>, dim3(g), dim3(256), 0, (hipStream_t)s, (double*)dst,
                           (const double*)src, L.N, L.nr, ncomp, which);
    else
        hipLaunchKernelGGL(k_boundary_assign<float>, dim3(g), dim3(256), 0, (hipStream_t)s, (float*)dst,
                           (const float*)src, L.N, L.nr, ncomp, which);
    TN_HIP(hipGetLastError());
    TN_API_END
}

int tn_stokes_rhs_from_temperature(tn_stokes* K, const void* temperature, double rayleigh, void* f_u, void* f_p,
                                   tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(K && temperature && f_u && f_p, "null pointer");
    TN_REQUIRE((const void*)f_u != temperature, "rhs: f_u must not alias the temperature");
    auto& L = const_cast<tn_mesh*>(K->vel)->lvl.dev();
    auto& P = const_cast<tn_mesh*>(K->pre)->lvl.dev();
    const int bc[2] = {K->A->bc_cmb, K->A->bc_surf};
    auto go = [&](auto* fu) {
        using T = std::remove_pointer_t<decltype(fu)>;
        const hipStream_t st = (hipStream_t)s;
        const size_t bytes = sizeof(T) * 3 * (size_t)L.N;
        if (K->rhs_tmp.n != bytes) K->rhs_tmp.alloc(bytes);
        T* g = (T*)K->rhs_tmp.p;
        hipLaunchKernelGGL(k_buoyancy<T>, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0, st, g,
                           (const T*)temperature, L.d_coords.p, L.N, L.nr, rayleigh);
        TN_HIP(hipGetLastError());
        // f_u = M g, additive communication (vector_mass.hpp:59-153)
        tn::launch_fill<T>(fu, 3 * L.N, T(0), st);
        tn::MassArgs<T> a{L.d_coords.p, L.d_radii.p, L.n, L.nr, 3, g, nullptr, fu, L.d_geo.p};
        tn::launch_mass<T>(a, false, st);
        tn::launch_interface_sum<T>(L, 3, fu, st);
        // strong enforcement of the velocity BCs on the right-hand side
        // (mantlecirculation.cpp:890-898): homogeneous Dirichlet rows zero,
        // free-slip normal components removed
        for (int shell = 0; shell < 2; ++shell) {
            const int which = shell == 0 ? 1 : 2;
            if (bc[shell] == TN_DIRICHLET) {
                hipLaunchKernelGGL(k_boundary_assign<T>, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0, st, fu,
                                   (const T*)nullptr, L.N, L.nr, 3, which);
                TN_HIP(hipGetLastError());
            } else if (bc[shell] == TN_FREESLIP) {
                tn::launch_freeslip_project<T>(L, fu, which, st);
            }
        }
        tn::launch_fill<T>((T*)f_p, P.N, T(0), st);
    };
    if (K->A->dtype == TN_F64)
        go((double*)f_u);
    else
        go((float*)f_u);
    TN_API_END
}

int tn_pressure_normalize(const tn_mesh* m, int dtype, void* p, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && p, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    const hipStream_t st = (hipStream_t)s;
    double* sum = L.d_scratch.p;
    tn::launch_fill<double>(sum, 1, 0.0, st);
    const double inv_owned = 1.0 / (double)L.owned;
    if (dtype == TN_F64) {
        tn::launch_sum_owned<double>(L, (const double*)p, sum, L.d_part.p, st);
        tn::launch_sub_scaled<double>((double*)p, L.N, sum, inv_owned, st);
    } else {
        tn::launch_sum_owned<float>(L, (const float*)p, sum, L.d_part.p, st);
        tn::launch_sub_scaled<float>((float*)p, L.N, sum, inv_owned, st);
    }
    TN_API_END
}

int tn_freeslip_project(const tn_mesh* m, int dtype, void* v, int which, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && v, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    if (dtype == TN_F64)
        tn::launch_freeslip_project<double>(L, (double*)v, which, (hipStream_t)s);
    else
        tn::launch_freeslip_project<float>(L, (float*)v, which, (hipStream_t)s);
    TN_API_END
}

}  // extern "C"
