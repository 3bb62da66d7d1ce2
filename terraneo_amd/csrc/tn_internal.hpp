// Internal declarations shared by the host C++ and the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/terraneo_amd.h"

namespace tn {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);

#define TN_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::tn::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
                              std::to_string(__LINE__) + " in " #call);                               \
    } while (0)

#define TN_REQUIRE(cond, msg)                              \
    do {                                                   \
        if (!(cond)) throw ::tn::Error(std::string(msg)); \
    } while (0)

// ---------------------------------------------------------------- device buffers
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count)
    {
        release();
        n = count;
        if (count) TN_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const std::vector<T>& h)
    {
        alloc(h.size());
        if (n) TN_HIP(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    }
};

// ---------------------------------------------------------------- mesh level
// Lateral geometry of one wedge, radius-independent part of J^{-T}:
//   J = A diag(r_mid, r_mid, dr/2),  A = [v1-v0, v2-v0, (v0+v1+v2)/3]
//   grad N (lateral vertex i, bottom/top) = l_i / r_mid  -/+  z * 2/dr
//   l1 = A^{-T}e0 / 2, l2 = A^{-T}e1 / 2, l0 = -(l1+l2), z = A^{-T}e2 / 6
//   |det J| = r_mid^2 (dr/2) |det A|
constexpr int GEO_STRIDE = 10;  // l1[3] l2[3] z[3] |detA|

struct MeshLevel {
    int lat_level = 0;
    int n = 0, nr = 0;
    int64_t N = 0;        // stored nodes
    int64_t owned = 0;    // owned nodes
    std::vector<double> coords;     // [10][n][n][3]
    std::vector<double> radii;      // [nr]
    std::vector<uint8_t> owned_lat; // [10][n][n]
    std::vector<int32_t> grp_off;   // CSR over duplicated lateral nodes (>= 2 copies)
    std::vector<int32_t> grp_mem;   // lateral index d*n*n + x*n + y
    std::vector<double> geo;        // [10][n-1][n-1][2][GEO_STRIDE]

    DevBuf<double> d_coords;
    DevBuf<double> d_coords_owned;  // coords of owned lateral nodes, 0 at the other copies
    DevBuf<float> d_coords_f;
    DevBuf<double> d_radii;
    DevBuf<float> d_radii_f;
    DevBuf<double> d_geo;
    DevBuf<float> d_geo_f;
    DevBuf<uint8_t> d_owned_lat;
    DevBuf<int32_t> d_grp_off;
    DevBuf<int32_t> d_grp_mem;
    DevBuf<double> d_scratch;  // reduction results (doubles)
    DevBuf<double> d_part;     // per-block partial sums of the reductions (REDUCE_PART_DOUBLES)
    bool uploaded = false;

    int ngroups() const { return (int)grp_off.size() - 1; }
    void build(int lat_level, const std::vector<double>& radii_in);
    void upload();
    // device copies are made on first use, so host-only queries work without a GPU
    MeshLevel& dev()
    {
        if (!uploaded) upload();
        return *this;
    }
};

// ---------------------------------------------------------------- kernel entry points (HIP files)
template <typename T>
struct Geom {
    const T* coords;     // [10][n][n][3]
    const T* radii;      // [nr]
    const T* geo;        // [10][n-1][n-1][2][GEO_STRIDE]
    int n, nr;
};

// adjoint gradient w.r.t. per-cell log-viscosity: grad[cell] = -eta_c (T lam)^T A_c(1) u
template <typename T>
void launch_grad_log_cell_eta(const Geom<T>& g, const T* eta_cell, const T* u, const T* lam, int bc_cmb, int bc_surf,
                              T* grad, hipStream_t s);

template <typename T>
void launch_fill(T* y, int64_t count, T value, hipStream_t s);

template <typename T>
void launch_interface_sum(const MeshLevel& m, int ncomp, T* v, hipStream_t s);

template <typename T>
void launch_lincomb(int64_t count, T* y, int nx, const double* c, const T* const* x, double c0, hipStream_t s);

// Reductions over owned nodes accumulate in double and are deterministic: the
// kernel leaves one partial sum per block in part_dev (REDUCE_PART_DOUBLES
// doubles of device scratch) and a one-block pass adds them to out_dev in a
// fixed order.  Calls that share part_dev must be ordered (one stream).
constexpr int REDUCE_MAX_BLOCKS = 2048;
constexpr int REDUCE_PART_DOUBLES = 16 * (REDUCE_MAX_BLOCKS / 2);
// blocks (= partial sums) of a dot over ncomp fields of mesh m
int reduce_blocks(const MeshLevel& m, int ncomp);

// out_dev[0] += sum over owned nodes of a . b; ncomp fields of N each
template <typename T>
void launch_dot_owned(const MeshLevel& m, int ncomp, const T* a, const T* b, double* out_dev, double* part_dev,
                      hipStream_t s);
// tangential projection of a vector field's nodes on shell `which` (1 CMB, 2 surface)
template <typename T>
void launch_freeslip_project(const MeshLevel& m, T* v, int which, hipStream_t s);

// w += alpha x; *out_dev += sum_owned w . y (y == w: squared norm of the new w).
// alpha_dev (optional): the coefficient is alpha * alpha_dev[0], read on the device.
template <typename T>
void launch_axpy_dot(const MeshLevel& m, int ncomp, T* w, double alpha, const T* x, const T* y, double* out_dev,
                     double* part_dev, hipStream_t s, const double* alpha_dev = nullptr);
// y = x where flag_dev[0] != 0 (count entries)
template <typename T>
void launch_copy_if_flag(T* y, const T* x, int64_t count, const int* flag_dev, hipStream_t s);
// Chebyshev step: z = dinv (b - az); d = alpha z + beta d; x += d.
// mode 0: general; 1: first step (d = alpha z); 2: first step from x = 0
// (az and x not read, x = d written; az may be null)
template <typename T>
void launch_cheby_update(T* x, T* d, const T* az, const T* b, const T* dinv, int64_t count, double alpha, double beta,
                         int mode, hipStream_t s);

// y[i] -= scale * sum_dev[0], i < count
template <typename T>
void launch_sub_scaled(T* y, int64_t count, const double* sum_dev, double scale, hipStream_t s);
// out_dev[0] += sum of a one-component field over owned nodes
template <typename T>
void launch_sum_owned(const MeshLevel& m, const T* a, double* out_dev, double* part_dev, hipStream_t s);

// fs/fs penalty: beta[0..3) = sum_owned r x u  ;  y (+)= gamma x r with gamma = S beta
template <typename T>
void launch_penalty_moment(const MeshLevel& m, const T* u, double* beta_dev, double* partial_dev, hipStream_t s);
// blocks of the moment pass = partial sums it leaves in partial_dev (3 doubles each)
constexpr int PENALTY_MAX_BLOCKS = 2048;
template <typename T>
void launch_penalty_init(const MeshLevel& m, const double* beta_dev, const double* S_dev, T* y, bool add,
                         hipStream_t s);

template <typename T>
struct MultiDotArgs {
    static constexpr int MAX = 16;
    const T* b[MAX];
    int nb;
};

template <typename T>
void launch_mul(T* y, const T* d, int64_t count, hipStream_t s);
template <typename T>
void launch_invert(T* y, int64_t count, hipStream_t s);
template <typename TD, typename TS>
void launch_cast(TD* y, const TS* x, int64_t count, hipStream_t s);
template <typename T>
void launch_randomize(T* y, int64_t count, uint64_t seed, hipStream_t s);
// *out_dev = max |y[i]| (out_dev must hold 0.0 on entry)
template <typename T>
void launch_max_abs(const T* y, int64_t count, double* out_dev, hipStream_t s);
template <typename T>
void launch_nonfinite(const T* y, int64_t count, int* flag_dev, hipStream_t s);
// out_dev[k] += sum_owned a . bs.b[k], k < bs.nb
template <typename T>
void launch_multidot(const MeshLevel& m, int ncomp, const T* a, const MultiDotArgs<T>& bs, double* out_dev,
                     double* part_dev, hipStream_t s);

}  // namespace tn

struct tn_mesh {
    tn::MeshLevel lvl;
};
