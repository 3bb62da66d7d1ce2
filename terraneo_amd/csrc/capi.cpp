// C ABI: mesh, level-1 operations and the epsilon-div-div operator.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "tn_internal.hpp"
#include "tn_eps_march.hpp"
#include "tn_mg.hpp"
#include "tn_objects.hpp"

namespace tn {
namespace {
thread_local std::string g_last_error;
// host-visible word of the march kernels' bounded waits (pinned, coherent)
int* g_wait_fail_host = nullptr;
int* g_wait_fail_dev = nullptr;
std::atomic<int> g_wait_limit{1 << 20};
std::mutex g_wait_mutex;
}  // namespace
void set_last_error(const std::string& msg) { g_last_error = msg; }

MarchGuard march_guard()
{
    std::lock_guard<std::mutex> lk(g_wait_mutex);
    if (!g_wait_fail_host) {
        int* h = nullptr;
        TN_HIP(hipHostMalloc((void**)&h, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
        __atomic_store_n(h, 0, __ATOMIC_RELEASE);
        void* d = nullptr;
        TN_HIP(hipHostGetDevicePointer(&d, h, 0));
        g_wait_fail_host = h;
        g_wait_fail_dev = (int*)d;
    }
    return {g_wait_fail_dev, g_wait_limit.load()};
}

void check_march_status()
{
    int* h = g_wait_fail_host;
    if (h && __atomic_exchange_n(h, 0, __ATOMIC_ACQ_REL))
        throw Error("march kernel: a hand-off wait between waves gave up (output of an earlier apply is invalid)");
}
}  // namespace tn

#define TN_API_BEGIN try {
// closes a TN_API_BEGIN block without returning on success
#define TN_API_END_SILENT           \
    }                               \
    catch (const std::exception& e) \
    {                               \
        tn::set_last_error(e.what()); \
        return -1;                  \
    }
#define TN_API_END                          \
    }                                       \
    catch (const std::exception& e)         \
    {                                       \
        tn::set_last_error(e.what());       \
        return -1;                          \
    }                                       \
    return 0;

using tn::MeshLevel;

namespace {

void build_penalty(tn_eps* op)
{
    const MeshLevel& m = op->mesh->lvl;
    // Gram matrix of the rotation fields e_j x p over owned nodes:
    // G = sum (|p|^2 I - p p^T)   (Gram-Schmidt of these fields, as the
    // reference does, equals multiplication by the Cholesky inverse, so the
    // projector sum_i n_i n_i^T reduces to G^{-1} in mode coordinates)
    double G[3][3] = {{0}};
    const int n = m.n, nr = m.nr;
    for (int64_t lat = 0; lat < 10LL * n * n; ++lat) {
        if (!m.owned_lat[lat]) continue;
        const double* c = &m.coords[3 * lat];
        for (int r = 0; r < nr; ++r) {
            const double p[3] = {c[0] * m.radii[r], c[1] * m.radii[r], c[2] * m.radii[r]};
            const double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) G[a][b] += (a == b ? pp : 0.0) - p[a] * p[b];
        }
    }
    const double det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                       G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                       G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    double S[9];
    S[0] = (G[1][1] * G[2][2] - G[1][2] * G[2][1]) / det;
    S[1] = (G[0][2] * G[2][1] - G[0][1] * G[2][2]) / det;
    S[2] = (G[0][1] * G[1][2] - G[0][2] * G[1][1]) / det;
    S[3] = (G[1][2] * G[2][0] - G[1][0] * G[2][2]) / det;
    S[4] = (G[0][0] * G[2][2] - G[0][2] * G[2][0]) / det;
    S[5] = (G[0][2] * G[1][0] - G[0][0] * G[1][2]) / det;
    S[6] = (G[1][0] * G[2][1] - G[1][1] * G[2][0]) / det;
    S[7] = (G[0][1] * G[2][0] - G[0][0] * G[2][1]) / det;
    S[8] = (G[0][0] * G[1][1] - G[0][1] * G[1][0]) / det;
    // [0..3) beta (rotational moments of the input), [4..13) eps S, [16..) the moment pass's partial sums
    std::vector<double> h(16 + 3 * tn::PENALTY_MAX_BLOCKS, 0.0);
    for (int i = 0; i < 9; ++i) h[4 + i] = op->penalty_eps * S[i];
    op->pen.upload(h);
    op->penalty_ready = true;
}

template <typename T>
tn::Geom<T> geom_of(const MeshLevel& m);
template <>
tn::Geom<double> geom_of<double>(const MeshLevel& m)
{
    return {m.d_coords.p, m.d_radii.p, m.d_geo.p, m.n, m.nr};
}
template <>
tn::Geom<float> geom_of<float>(const MeshLevel& m)
{
    return {m.d_coords_f.p, m.d_radii_f.p, m.d_geo_f.p, m.n, m.nr};
}

int device_cus()
{
    static int cache[64] = {0};
    int dev = 0;
    TN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cache[dev]) {
        int cus = 0;
        TN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cache[dev] = cus > 0 ? cus : 1;
    }
    return cache[dev];
}

// Segment length (node columns per block) of a march kernel with `own` node
// rows per block and `waves_per_cu` resident waves per CU: the kernel takes
// ceil(blocks / resident blocks) rounds of (seg + 1) quad columns each; pick
// the segment that minimises that product.  TN_MARCH_SEG overrides (tuning).
}  // namespace
int tn::march_seg_for(int n, int own, int waves_per_cu)
{
    static const int env = [] {
        const char* e = std::getenv("TN_MARCH_SEG");
        return e ? std::atoi(e) : 0;
    }();
    if (env > 0) return std::min(std::min(env, n), tn::MARCH_SEG_MAX);
    const int64_t per_cu = std::max(1, waves_per_cu / (own + 1));
    const int64_t nxb = (n + own - 1) / own, cus = device_cus() * per_cu;
    int best = std::min(n, tn::MARCH_SEG_MAX);
    int64_t best_cost = -1;
    for (int seg = std::min(n, tn::MARCH_SEG_MAX); seg >= std::min(n, 6); --seg) {
        const int64_t blocks = 10 * nxb * ((n + seg - 1) / seg);
        const int64_t cost = ((blocks + cus - 1) / cus) * (seg + 1);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = seg;
        }
    }
    return best;
}
namespace {

int march_seg(int n, int dtype_size)
{
    return tn::march_seg_for(n, tn::march_block_rows(dtype_size), tn::march_waves_per_cu(dtype_size));
}

template <typename T>
void eps_apply_t(tn_eps* op, const void* src_v, void* dst_v, hipStream_t s)
{
    const MeshLevel& m = const_cast<tn_mesh*>(op->mesh)->lvl.dev();
    const T* src = static_cast<const T*>(src_v);
    T* dst = static_cast<T*>(dst_v);
    const T* eta = static_cast<const T*>(op->eta);
    TN_REQUIRE(src && dst, "eps apply: null pointer");
    TN_REQUIRE((const void*)src != (void*)dst, "eps apply: src and dst must differ");
    const bool penalty = !op->diagonal && op->bc_cmb == TN_FREESLIP && op->bc_surf == TN_FREESLIP;
    const bool add = op->apply_mode == TN_ADD;
    if (penalty) {
        if (!op->penalty_ready) build_penalty(op);
        tn::launch_penalty_moment<T>(m, src, op->pen.p, op->pen.p + 16, s);
    }
    if (op->stored_dtype < 0) {
        // matrix-free, atomic-free: the kernel writes every node copy once
        TN_REQUIRE(eta || op->eta_cell, "eps apply: viscosity not set");
        const T* eta_cell = static_cast<const T*>(op->eta_cell);
        const bool need_blk = op->bc_cmb != TN_NEUMANN || op->bc_surf != TN_NEUMANN;
        if (need_blk && !op->blk_ready) {
            const size_t bytes = sizeof(T) * 2 * 10 * (size_t)m.n * m.n * 3;
            if (op->blk.n != bytes) op->blk.alloc(bytes);
            tn::launch_boundary_blocks<T>(geom_of<T>(m), eta, eta_cell, op->bc_cmb, op->bc_surf, (T*)op->blk.p, s);
            op->blk_ready = true;
        }
        const tn::Geom<T> g = geom_of<T>(m);
        // penalty positions: every copy (the reference's rule) or the owned copy only
        const double* pen_coords = op->penalty_once ? m.d_coords_owned.p : m.d_coords.p;
        tn::MarchArgs<T> a{src, dst, eta, eta_cell, g.geo, g.coords, g.radii, pen_coords, m.d_radii.p,
                           need_blk ? (const T*)op->blk.p : nullptr, penalty ? op->pen.p : nullptr,
                           m.n, m.nr, march_seg(m.n, (int)sizeof(T)), op->bc_cmb, op->bc_surf, add ? 1 : 0,
                           tn::march_guard()};
        tn::launch_eps_march<T>(a, op->diagonal, s);
    } else {
        // the gather writes every node copy once (boundary test side and penalty fused)
        const int64_t ncs = tn::stored_cellsum_elems(m.n, m.nr);
        if (op->cellsum.n != (size_t)ncs * sizeof(T)) op->cellsum.alloc((size_t)ncs * sizeof(T));
        T* csum = (T*)op->cellsum.p;
        T* ctop = csum + 12 * (10LL * (m.n - 1) * (m.n - 1) * (m.nr - 1));
        const bool need_blk = op->bc_cmb != TN_NEUMANN || op->bc_surf != TN_NEUMANN;
        const size_t blk_bytes = sizeof(T) * 2 * 10 * (size_t)m.n * m.n * 3;
        if (need_blk && op->blk.n != blk_bytes) {
            op->blk.alloc(blk_bytes);
            op->blk_ready = false;
        }
        auto run = [&](auto* stored) {
            using S = std::remove_const_t<std::remove_pointer_t<decltype(stored)>>;
            tn::StoredArgs<T, S> a{stored, src, dst, m.d_coords.p, m.n, m.nr, op->bc_cmb, op->bc_surf, add ? 1 : 0,
                                   csum, ctop, need_blk ? (const T*)op->blk.p : nullptr,
                                   penalty ? op->pen.p : nullptr, m.d_radii.p,
                                   op->penalty_once ? m.d_coords_owned.p : m.d_coords.p};
            if (need_blk && !op->blk_ready) {
                tn::launch_stored_blk<T, S>(a, (T*)op->blk.p, s);
                op->blk_ready = true;
            }
            tn::launch_stored_apply<T, S>(a, op->diagonal, s);
        };
        if (op->stored_dtype == TN_F64)
            run((const double*)op->stored.p);
        else
            run((const float*)op->stored.p);
    }
    if (op->comm_mode == TN_COMM_ADDITIVE) tn::launch_interface_sum<T>(m, 3, dst, s);
}

}  // namespace

namespace tn {
void eps_apply(tn_eps* op, const void* src, void* dst, hipStream_t s)
{
    if (op->dtype == TN_F64)
        eps_apply_t<double>(op, src, dst, s);
    else
        eps_apply_t<float>(op, src, dst, s);
}
}  // namespace tn

extern "C" {

const char* tn_last_error(void) { return tn::g_last_error.c_str(); }
int tn_version(void) { return 1; }

int tn_malloc(void** ptr, size_t bytes)
{
    TN_API_BEGIN
    TN_REQUIRE(ptr, "null out pointer");
    TN_HIP(hipMalloc(ptr, bytes));
    TN_API_END
}
int tn_free(void* ptr)
{
    TN_API_BEGIN
    TN_HIP(hipFree(ptr));
    TN_API_END
}
int tn_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    TN_API_BEGIN
    TN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    TN_API_END
}
int tn_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    TN_API_BEGIN
    TN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    TN_API_END
}
int tn_sync(void)
{
    TN_API_BEGIN
    TN_HIP(hipDeviceSynchronize());
    tn::check_march_status();
    TN_API_END
}

int tn_set_wait_limit(int polls)
{
    TN_API_BEGIN
    TN_REQUIRE(polls >= 0, "wait limit must be >= 0");
    tn::g_wait_limit.store(polls == 0 ? 0 : polls);
    TN_API_END
}

// ------------------------------------------------------------------ mesh
static tn_mesh* new_mesh(int lat_level, const std::vector<double>& radii)
{
    auto m = std::make_unique<tn_mesh>();
    m->lvl.build(lat_level, radii);
    return m.release();
}

int tn_mesh_create(int lat_level, int rad_level, double r_min, double r_max, tn_mesh** out)
{
    TN_API_BEGIN
    TN_REQUIRE(out, "null out pointer");
    TN_REQUIRE(rad_level >= 1 && rad_level <= 12, "radial level out of range [1,12]");
    TN_REQUIRE(r_min < r_max, "r_min must be < r_max");
    const int nr = (1 << rad_level) + 1;
    std::vector<double> radii(nr);
    // uniform_shell_radii, spherical_shell.hpp:27-52
    const double step = (r_max - r_min) / (nr - 1);
    for (int i = 0; i < nr; ++i) radii[i] = r_min + i * step;
    radii[nr - 1] = r_max;
    *out = new_mesh(lat_level, radii);
    TN_API_END
}

int tn_mesh_create_subdomains(int lat_level, int rad_level, double r_min, double r_max, int s_lat, int s_rad,
                              tn_mesh** out)
{
    TN_API_BEGIN
    TN_REQUIRE(s_lat >= 0 && s_rad >= 0, "subdomain refinement levels must be >= 0");
    TN_REQUIRE(s_lat == 0 && s_rad == 0,
               "diamonds split into subdomains (s_lat, s_rad > 0) are not supported: one subdomain per diamond");
    TN_API_END_SILENT
    return tn_mesh_create(lat_level, rad_level, r_min, r_max, out);
}

int tn_mesh_create_radii(int lat_level, int nr, const double* radii, tn_mesh** out)
{
    TN_API_BEGIN
    TN_REQUIRE(out && radii, "null pointer");
    *out = new_mesh(lat_level, std::vector<double>(radii, radii + nr));
    TN_API_END
}

int tn_mesh_destroy(tn_mesh* m)
{
    TN_API_BEGIN
    delete m;
    TN_API_END
}

int tn_mesh_dims(const tn_mesh* m, int* n, int* nr, int64_t* num_nodes)
{
    TN_API_BEGIN
    TN_REQUIRE(m, "null mesh");
    if (n) *n = m->lvl.n;
    if (nr) *nr = m->lvl.nr;
    if (num_nodes) *num_nodes = m->lvl.N;
    TN_API_END
}

int64_t tn_mesh_owned_nodes(const tn_mesh* m) { return m ? m->lvl.owned : -1; }

int tn_mesh_lateral_coords(const tn_mesh* m, double* out)
{
    TN_API_BEGIN
    TN_REQUIRE(m && out, "null pointer");
    std::memcpy(out, m->lvl.coords.data(), sizeof(double) * m->lvl.coords.size());
    TN_API_END
}

int tn_mesh_radii(const tn_mesh* m, double* out)
{
    TN_API_BEGIN
    TN_REQUIRE(m && out, "null pointer");
    std::memcpy(out, m->lvl.radii.data(), sizeof(double) * m->lvl.radii.size());
    TN_API_END
}

int tn_mesh_owned_mask(const tn_mesh* m, uint8_t* out)
{
    TN_API_BEGIN
    TN_REQUIRE(m && out, "null pointer");
    const auto& L = m->lvl;
    for (int64_t lat = 0; lat < 10LL * L.n * L.n; ++lat)
        for (int r = 0; r < L.nr; ++r) out[lat * L.nr + r] = L.owned_lat[lat];
    TN_API_END
}

int tn_mesh_interface_info(const tn_mesh* m, int* ngroups, int* nmembers)
{
    TN_API_BEGIN
    TN_REQUIRE(m, "null mesh");
    if (ngroups) *ngroups = m->lvl.ngroups();
    if (nmembers) *nmembers = (int)m->lvl.grp_mem.size();
    TN_API_END
}

// ------------------------------------------------------------------ BLAS-1
int tn_interface_sum(const tn_mesh* m, int dtype, int ncomp, void* v, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && v, "null pointer");
    TN_REQUIRE(ncomp == 1 || ncomp == 3, "ncomp must be 1 or 3");
    const MeshLevel& L = const_cast<tn_mesh*>(m)->lvl.dev();
    if (dtype == TN_F64)
        tn::launch_interface_sum<double>(L, ncomp, (double*)v, (hipStream_t)s);
    else
        tn::launch_interface_sum<float>(L, ncomp, (float*)v, (hipStream_t)s);
    TN_API_END
}

int tn_lincomb(const tn_mesh* m, int dtype, int ncomp, void* y, int nx, const double* c, const void* const* x,
               double c0, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && y, "null pointer");
    TN_REQUIRE(nx == 0 || (c && x), "null coefficient/input arrays");
    const int64_t cnt = m->lvl.N * ncomp;
    if (dtype == TN_F64)
        tn::launch_lincomb<double>(cnt, (double*)y, nx, c, (const double* const*)x, c0, (hipStream_t)s);
    else
        tn::launch_lincomb<float>(cnt, (float*)y, nx, c, (const float* const*)x, c0, (hipStream_t)s);
    TN_API_END
}

int tn_dot(const tn_mesh* m, int dtype, int ncomp, const void* a, const void* b, double* out, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && a && b && out, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    tn::launch_fill<double>(L.d_scratch.p, 1, 0.0, (hipStream_t)s);
    if (dtype == TN_F64)
        tn::launch_dot_owned<double>(L, ncomp, (const double*)a, (const double*)b, L.d_scratch.p, L.d_part.p,
                                     (hipStream_t)s);
    else
        tn::launch_dot_owned<float>(L, ncomp, (const float*)a, (const float*)b, L.d_scratch.p, L.d_part.p,
                                    (hipStream_t)s);
    TN_HIP(hipMemcpyAsync(out, L.d_scratch.p, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)s));
    TN_HIP(hipStreamSynchronize((hipStream_t)s));
    tn::check_march_status();
    TN_API_END
}

int tn_invert_entries(const tn_mesh* m, int dtype, int ncomp, void* v, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && v, "null pointer");
    const int64_t cnt = m->lvl.N * ncomp;
    if (dtype == TN_F64)
        tn::launch_invert<double>((double*)v, cnt, (hipStream_t)s);
    else
        tn::launch_invert<float>((float*)v, cnt, (hipStream_t)s);
    TN_API_END
}

int tn_scale_with_vector(const tn_mesh* m, int dtype, int ncomp, void* y, const void* d, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && y && d, "null pointer");
    const int64_t cnt = m->lvl.N * ncomp;
    if (dtype == TN_F64)
        tn::launch_mul<double>((double*)y, (const double*)d, cnt, (hipStream_t)s);
    else
        tn::launch_mul<float>((float*)y, (const float*)d, cnt, (hipStream_t)s);
    TN_API_END
}

int tn_randomize(const tn_mesh* m, int dtype, int ncomp, void* y, uint64_t seed, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && y, "null pointer");
    const int64_t cnt = m->lvl.N * ncomp;
    if (dtype == TN_F64)
        tn::launch_randomize<double>((double*)y, cnt, seed, (hipStream_t)s);
    else
        tn::launch_randomize<float>((float*)y, cnt, seed, (hipStream_t)s);
    TN_API_END
}

int tn_max_abs_entry(const tn_mesh* m, int dtype, int ncomp, const void* v, double* out, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && v && out, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    const int64_t cnt = L.N * ncomp;
    tn::launch_fill<double>(L.d_scratch.p, 1, 0.0, (hipStream_t)s);
    if (dtype == TN_F64)
        tn::launch_max_abs<double>((const double*)v, cnt, L.d_scratch.p, (hipStream_t)s);
    else
        tn::launch_max_abs<float>((const float*)v, cnt, L.d_scratch.p, (hipStream_t)s);
    TN_HIP(hipMemcpyAsync(out, L.d_scratch.p, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)s));
    TN_HIP(hipStreamSynchronize((hipStream_t)s));
    tn::check_march_status();
    TN_API_END
}

int tn_has_nan_or_inf(const tn_mesh* m, int dtype, int ncomp, const void* v, int* out, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && v && out, "null pointer");
    auto& L = const_cast<tn_mesh*>(m)->lvl.dev();
    const int64_t cnt = L.N * ncomp;
    int* flag = reinterpret_cast<int*>(L.d_scratch.p);
    tn::launch_fill<int>(flag, 1, 0, (hipStream_t)s);
    if (dtype == TN_F64)
        tn::launch_nonfinite<double>((const double*)v, cnt, flag, (hipStream_t)s);
    else
        tn::launch_nonfinite<float>((const float*)v, cnt, flag, (hipStream_t)s);
    TN_HIP(hipMemcpyAsync(out, flag, sizeof(int), hipMemcpyDeviceToHost, (hipStream_t)s));
    TN_HIP(hipStreamSynchronize((hipStream_t)s));
    tn::check_march_status();
    TN_API_END
}

// ------------------------------------------------------------------ epsilon operator
int tn_eps_create(const tn_mesh* m, int dtype, const void* eta, int bc_cmb, int bc_surf, tn_eps** out)
{
    TN_API_BEGIN
    TN_REQUIRE(m && out, "null pointer");
    TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "dtype must be TN_F64 or TN_F32");
    TN_REQUIRE(bc_cmb >= 0 && bc_cmb <= 2 && bc_surf >= 0 && bc_surf <= 2, "invalid boundary condition");
    auto op = std::make_unique<tn_eps>();
    op->mesh = m;
    op->dtype = dtype;
    op->eta = eta;
    op->bc_cmb = bc_cmb;
    op->bc_surf = bc_surf;
    *out = op.release();
    TN_API_END
}

int tn_eps_destroy(tn_eps* op)
{
    TN_API_BEGIN
    delete op;
    TN_API_END
}

int tn_eps_set_modes(tn_eps* op, int apply_mode, int comm_mode)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    TN_REQUIRE(apply_mode == TN_REPLACE || apply_mode == TN_ADD, "invalid apply mode");
    TN_REQUIRE(comm_mode == TN_COMM_SKIP || comm_mode == TN_COMM_ADDITIVE, "invalid communication mode");
    op->apply_mode = apply_mode;
    op->comm_mode = comm_mode;
    TN_API_END
}

int tn_eps_set_diagonal(tn_eps* op, int diagonal)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    op->diagonal = diagonal != 0;
    TN_API_END
}

int tn_eps_set_boundary_conditions(tn_eps* op, int bc_cmb, int bc_surf)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    TN_REQUIRE(bc_cmb >= 0 && bc_cmb <= 2 && bc_surf >= 0 && bc_surf <= 2, "invalid boundary condition");
    op->bc_cmb = bc_cmb;
    op->bc_surf = bc_surf;
    op->blk_ready = false;
    TN_API_END
}

int tn_eps_set_eta(tn_eps* op, const void* eta)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    op->eta = eta;
    op->blk_ready = false;
    TN_API_END
}

int tn_eps_set_cell_eta(tn_eps* op, const void* eta_cell)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    op->eta_cell = eta_cell;
    op->blk_ready = false;
    TN_API_END
}

int tn_eps_grad_log_cell_eta(tn_eps* op, const void* u, const void* lambda, void* grad, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(op && u && lambda && grad, "null pointer");
    TN_REQUIRE(op->eta_cell, "gradient: the operator has no per-cell viscosity (tn_eps_set_cell_eta)");
    TN_REQUIRE(op->stored_dtype < 0, "gradient: the operator must be matrix-free");
    const MeshLevel& m = const_cast<tn_mesh*>(op->mesh)->lvl.dev();
    if (op->dtype == TN_F64)
        tn::launch_grad_log_cell_eta<double>(geom_of<double>(m), (const double*)op->eta_cell, (const double*)u,
                                             (const double*)lambda, op->bc_cmb, op->bc_surf, (double*)grad,
                                             (hipStream_t)s);
    else
        tn::launch_grad_log_cell_eta<float>(geom_of<float>(m), (const float*)op->eta_cell, (const float*)u,
                                            (const float*)lambda, op->bc_cmb, op->bc_surf, (float*)grad,
                                            (hipStream_t)s);
    TN_API_END
}

int tn_eps_apply(tn_eps* op, const void* src, void* dst, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    tn::check_march_status();
    tn::eps_apply(op, src, dst, (hipStream_t)s);
    TN_API_END
}

int tn_eps_set_penalty_once(tn_eps* op, int once)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    op->penalty_once = once ? 1 : 0;
    TN_API_END
}

int tn_eps_set_penalty_epsilon(tn_eps* op, double eps)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    op->penalty_eps = eps;
    op->penalty_ready = false;
    TN_API_END
}

int tn_eps_gca_from(tn_eps* coarse, const tn_eps* fine, int store_dtype, tn_stream_t s)
{
    return tn_eps_gca_from_bracket(coarse, fine, store_dtype, TN_GCA_REFERENCE, s);
}

int tn_eps_gca_from_bracket(tn_eps* coarse, const tn_eps* fine, int store_dtype, int bracket, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(bracket == TN_GCA_REFERENCE || bracket == TN_GCA_GEOMETRIC, "invalid GCA bracket");
    TN_REQUIRE(coarse && fine, "null operator");
    TN_REQUIRE(store_dtype == TN_F64 || store_dtype == TN_F32, "store dtype must be TN_F64 or TN_F32");
    auto& mc = const_cast<tn_mesh*>(coarse->mesh)->lvl.dev();
    auto& mf = const_cast<tn_mesh*>(fine->mesh)->lvl.dev();
    TN_REQUIRE(2 * (mc.n - 1) == mf.n - 1 && 2 * (mc.nr - 1) == mf.nr - 1,
               "GCA: fine mesh must be the uniform refinement of the coarse mesh");
    const int64_t bytes = tn::stored_elems(coarse->cells()) * (store_dtype == TN_F64 ? 8 : 4);
    coarse->stored.alloc((size_t)bytes);
    coarse->stored_dtype = store_dtype;
    coarse->blk_ready = false;  // boundary forms follow the new matrices
    const bool fine_geo = fine->stored_dtype < 0;
    TN_REQUIRE(!fine_geo || fine->eta || fine->eta_cell, "GCA: fine operator has no viscosity");
    auto go = [&](auto* storec, auto* eta_f, auto* stored_f) {
        using S = std::remove_pointer_t<decltype(storec)>;
        using ET = std::remove_const_t<std::remove_pointer_t<decltype(eta_f)>>;
        using SF = std::remove_const_t<std::remove_pointer_t<decltype(stored_f)>>;
        tn::GcaArgs<ET, S, SF> a{mc.n,  mc.nr,  mf.n, mf.nr, mf.d_geo.p, mf.d_radii.p, eta_f,
                                 static_cast<const ET*>(fine->eta_cell), stored_f, storec, bracket};
        tn::launch_gca_build<ET, S, SF>(a, fine_geo, (hipStream_t)s);
    };
    const bool f64e = fine->dtype == TN_F64;
    const bool f64sf = fine->stored_dtype == TN_F64 || fine->stored_dtype < 0;
    if (store_dtype == TN_F64) {
        double* sc = (double*)coarse->stored.p;
        if (f64e && f64sf) go(sc, (const double*)fine->eta, (const double*)fine->stored.p);
        else if (f64e) go(sc, (const double*)fine->eta, (const float*)fine->stored.p);
        else if (f64sf) go(sc, (const float*)fine->eta, (const double*)fine->stored.p);
        else go(sc, (const float*)fine->eta, (const float*)fine->stored.p);
    } else {
        float* sc = (float*)coarse->stored.p;
        if (f64e && f64sf) go(sc, (const double*)fine->eta, (const double*)fine->stored.p);
        else if (f64e) go(sc, (const double*)fine->eta, (const float*)fine->stored.p);
        else if (f64sf) go(sc, (const float*)fine->eta, (const double*)fine->stored.p);
        else go(sc, (const float*)fine->eta, (const float*)fine->stored.p);
    }
    TN_API_END
}

namespace {
// levels between a mesh and the coarsest mesh of its hierarchy (both directions refined alike)
int levels_between(const MeshLevel& m, const MeshLevel& coarsest)
{
    for (int L = 0; L < 20; ++L)
        if ((int64_t)(coarsest.n - 1) << L == m.n - 1 && (int64_t)(coarsest.nr - 1) << L == m.nr - 1) return L;
    TN_REQUIRE(false, "selective GCA: the mesh is not a uniform refinement of the coarsest mesh");
    return -1;
}
}  // namespace

int tn_gca_collect_elements(const tn_mesh* fine, const tn_mesh* coarsest, int dtype, const void* k, void* flags,
                            tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fine && coarsest && k && flags, "null pointer");
    auto& mf = const_cast<tn_mesh*>(fine)->lvl.dev();
    auto& m0 = const_cast<tn_mesh*>(coarsest)->lvl.dev();
    const int L = levels_between(mf, m0);
    if (dtype == TN_F64)
        tn::launch_gca_collect<double>(mf, m0, L, (const double*)k, (double*)flags, (hipStream_t)s);
    else
        tn::launch_gca_collect<float>(mf, m0, L, (const float*)k, (float*)flags, (hipStream_t)s);
    TN_API_END
}

int tn_eps_gca_select(tn_eps* coarse, const tn_mesh* coarsest, const void* flags, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(coarse && coarsest && flags, "null pointer");
    TN_REQUIRE(coarse->stored_dtype >= 0, "selective GCA: the operator has no Galerkin matrices (tn_eps_gca_from first)");
    TN_REQUIRE(coarse->eta && !coarse->eta_cell, "selective GCA: the operator needs a nodal viscosity of its level");
    auto& m = const_cast<tn_mesh*>(coarse->mesh)->lvl.dev();
    auto& m0 = const_cast<tn_mesh*>(coarsest)->lvl.dev();
    const int L = levels_between(m, m0);
    auto go = [&](auto* flg, auto* eta) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(flg)>>;
        if (coarse->stored_dtype == TN_F64)
            tn::launch_gca_select<T, double>(m0, L, flg, m, eta, (double*)coarse->stored.p, (hipStream_t)s);
        else
            tn::launch_gca_select<T, float>(m0, L, flg, m, eta, (float*)coarse->stored.p, (hipStream_t)s);
    };
    if (coarse->dtype == TN_F64)
        go((const double*)flags, (const double*)coarse->eta);
    else
        go((const float*)flags, (const float*)coarse->eta);
    coarse->blk_ready = false;  // boundary forms follow the new matrices
    TN_API_END
}

namespace {
int64_t local_cell(const tn_eps* op, int sd, int x, int y, int r, int wedge)
{
    const auto& m = op->mesh->lvl;
    TN_REQUIRE(sd >= 0 && sd < 10 && x >= 0 && x < m.n - 1 && y >= 0 && y < m.n - 1 && r >= 0 && r < m.nr - 1 &&
                   (wedge == 0 || wedge == 1),
               "local matrix index out of range");
    return (((int64_t)sd * (m.n - 1) + x) * (m.n - 1) + y) * (m.nr - 1) + r;
}
}  // namespace

// GCACapable::get_local_matrix (epsilon_divdiv_kerngen.hpp:481-497): the stored
// matrix when the operator has stored matrices, else the wedge's matrix
// assembled from the geometry and the viscosity (assemble_local_matrix,
// :2051-2108, the Neumann matrix without boundary masks); one bulk copy.
int tn_eps_get_local_matrix(const tn_eps* op_c, int sd, int x, int y, int r, int wedge, double* out)
{
    TN_API_BEGIN
    TN_REQUIRE(op_c && out, "null pointer");
    tn_eps* op = const_cast<tn_eps*>(op_c);
    const int64_t cell = local_cell(op, sd, x, y, r, wedge);
    const MeshLevel& m = const_cast<tn_mesh*>(op->mesh)->lvl.dev();
    if (op->lm.n != 324) op->lm.alloc(324);
    const hipStream_t st = nullptr;
    if (op->stored_dtype == TN_F64) {
        tn::launch_stored_get<double>((const double*)op->stored.p, cell, wedge, op->lm.p, st);
    } else if (op->stored_dtype == TN_F32) {
        tn::launch_stored_get<float>((const float*)op->stored.p, cell, wedge, op->lm.p, st);
    } else {
        TN_REQUIRE(op->eta || op->eta_cell, "local matrix: viscosity not set");
        if (op->dtype == TN_F64)
            tn::launch_local_matrix<double>(m, (const double*)op->eta, (const double*)op->eta_cell, sd, x, y, r, wedge,
                                            op->lm.p, st);
        else
            tn::launch_local_matrix<float>(m, (const float*)op->eta, (const float*)op->eta_cell, sd, x, y, r, wedge,
                                           op->lm.p, st);
    }
    TN_HIP(hipMemcpy(out, op->lm.p, sizeof(double) * 324, hipMemcpyDeviceToHost));
    TN_API_END
}

// GCACapable::set_local_matrix (kerngen.hpp:468-479): needs stored matrices.
// The storage is packed symmetric, so the matrix must be symmetric.
int tn_eps_set_local_matrix(tn_eps* op, int sd, int x, int y, int r, int wedge, const double* in)
{
    TN_API_BEGIN
    TN_REQUIRE(op && in, "null pointer");
    TN_REQUIRE(op->stored_dtype >= 0,
               "set_local_matrix: the operator has no stored matrices (tn_eps_set_stored_matrices or GCA first)");
    const int64_t cell = local_cell(op, sd, x, y, r, wedge);
    double amax = 0.0, asym = 0.0;
    for (int i = 0; i < 18; ++i)
        for (int j = 0; j < 18; ++j) {
            amax = std::max(amax, std::fabs(in[i * 18 + j]));
            asym = std::max(asym, std::fabs(in[i * 18 + j] - in[j * 18 + i]));
        }
    TN_REQUIRE(asym <= 1e-12 * amax, "set_local_matrix: matrix is not symmetric (storage is packed symmetric)");
    if (op->lm.n != 324) op->lm.alloc(324);
    TN_HIP(hipMemcpy(op->lm.p, in, sizeof(double) * 324, hipMemcpyHostToDevice));
    if (op->stored_dtype == TN_F64)
        tn::launch_stored_set<double>((double*)op->stored.p, cell, wedge, op->lm.p, nullptr);
    else
        tn::launch_stored_set<float>((float*)op->stored.p, cell, wedge, op->lm.p, nullptr);
    TN_HIP(hipDeviceSynchronize());
    op->blk_ready = false;  // boundary forms follow the matrices
    TN_API_END
}

// set_stored_matrix_mode(Full) (kerngen.hpp:443-464, local_matrix_storage.hpp:28-243):
// switch the operator to its stored-matrix path with every cell's own
// matrices (the on-the-fly assembly of all cells), stored in store_dtype.
int tn_eps_set_stored_matrices(tn_eps* op, int store_dtype, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(op, "null operator");
    TN_REQUIRE(store_dtype == TN_F64 || store_dtype == TN_F32, "store dtype must be TN_F64 or TN_F32");
    TN_REQUIRE(op->eta || op->eta_cell, "stored matrices: viscosity not set");
    const MeshLevel& m = const_cast<tn_mesh*>(op->mesh)->lvl.dev();
    const int64_t bytes = tn::stored_elems(op->cells()) * (store_dtype == TN_F64 ? 8 : 4);
    op->stored.alloc((size_t)bytes);
    op->stored_dtype = store_dtype;
    op->blk_ready = false;
    auto go = [&](auto* eta, auto* stored) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(eta)>>;
        using S = std::remove_pointer_t<decltype(stored)>;
        tn::launch_stored_own<T, S>(m, eta, static_cast<const T*>(op->eta_cell), stored, (hipStream_t)s);
    };
    if (op->dtype == TN_F64) {
        if (store_dtype == TN_F64) go((const double*)op->eta, (double*)op->stored.p);
        else go((const double*)op->eta, (float*)op->stored.p);
    } else {
        if (store_dtype == TN_F64) go((const float*)op->eta, (double*)op->stored.p);
        else go((const float*)op->eta, (float*)op->stored.p);
    }
    TN_API_END
}

}  // extern "C"
