// Level-1 kernels on mesh fields: fill, linear combinations, owned-node dots,
// the additive interface sum over duplicated diamond-edge nodes, and the
// free-slip rotation-penalty helpers.
//
// Reference counterparts: kernels/common/grid_operations.hpp:44-785
// (set/lincomb/masked_dot_product), communication/shell/communication.hpp:241-739
// (additive interface reduction, single-process form), epsilon_divdiv_kerngen.hpp:287-408,583-607
// (fs/fs penalty).
#include "tn_coarse.hpp"
#include "tn_internal.hpp"

namespace tn {
namespace {

template <typename T>
__global__ void k_fill(T* __restrict__ y, int64_t count, T v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = v;
}

template <typename T>
struct LinArgs {
    const T* x[3];
    double c[3];
    double c0;
    int nx;
};

template <typename T, int NX>
__global__ void k_lincomb(T* __restrict__ y, int64_t count, LinArgs<T> a)
{
    const T c0 = (T)a.c0;
    T c[3] = {(T)a.c[0], (T)a.c[1], (T)a.c[2]};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        T v = c0;
        if (NX > 0) v += c[0] * a.x[0][i];
        if (NX > 1) v += c[1] * a.x[1][i];
        if (NX > 2) v += c[2] * a.x[2][i];
        y[i] = v;
    }
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block total of `v`, valid in thread 0 (no atomics).
__device__ inline double block_total(double v)
{
    __shared__ double part[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) part[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (wv == 0) {
        const int nw = blockDim.x >> 6;
        s = lane < nw ? part[lane] : 0.0;
        s = wave_sum(s);
    }
    __syncthreads();
    return s;
}

// The block's total of `v` into part[blockIdx.x].  Reductions leave one
// partial per block and k_reduce_finalize adds them in a fixed order, so a
// reduction gives the same bits on every run (no floating-point atomics).
__device__ inline void block_partial(double v, double* __restrict__ part)
{
    const double t = block_total(v);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[k] += sum over b < nblk of part[k * nblk + b], k < nk (one block)
__global__ __launch_bounds__(256) void k_reduce_finalize(const double* __restrict__ part, int nblk, int nk,
                                                         double* __restrict__ out)
{
    for (int k = 0; k < nk; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(int64_t)k * nblk + b];
        const double t = block_total(s);
        if (threadIdx.x == 0) out[k] += t;
    }
}

// Owned-node reductions walk an [ncomp][nlat][nr] field by radial columns:
// one wave per column (c, lat), lanes over the column's nr contiguous entries
// (coalesced), the lateral ownership flag read once per column.  No per-entry
// index division.
struct ColumnWalk {
    int64_t col, ncol, step;
    int64_t nlat;
    __device__ ColumnWalk(int64_t nlat_, int ncomp)
        : col((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)),
          ncol(nlat_ * ncomp),
          step((int64_t)gridDim.x * (blockDim.x >> 6)),
          nlat(nlat_)
    {
    }
};

template <typename T>
__global__ __launch_bounds__(256) void k_dot_owned(const T* __restrict__ a, const T* __restrict__ b,
                                                   const uint8_t* __restrict__ owned_lat, int64_t N, int nr,
                                                   int ncomp, double* __restrict__ part)
{
    double s = 0.0;
    const int lane = threadIdx.x & 63;
    for (ColumnWalk cw(N / nr, ncomp); cw.col < cw.ncol; cw.col += cw.step) {
        if (!owned_lat[cw.col % cw.nlat]) continue;
        const T* ac = a + cw.col * nr;
        const T* bc = b + cw.col * nr;
        for (int r = lane; r < nr; r += 64) s += (double)ac[r] * (double)bc[r];
    }
    block_partial(s, part);
}

// sum over the owned entries of a one-component field (column walk), per block
template <typename T>
__global__ __launch_bounds__(256) void k_sum_owned(const T* __restrict__ a, const uint8_t* __restrict__ owned_lat,
                                                   int64_t N, int nr, double* __restrict__ part)
{
    double s = 0.0;
    const int lane = threadIdx.x & 63;
    for (ColumnWalk cw(N / nr, 1); cw.col < cw.ncol; cw.col += cw.step) {
        if (!owned_lat[cw.col]) continue;
        const T* ac = a + cw.col * nr;
        for (int r = lane; r < nr; r += 64) s += (double)ac[r];
    }
    block_partial(s, part);
}

// One Chebyshev step fused (chebyshev.hpp:100-168 does it in four passes):
// z = dinv (b - Az); d = alpha z + beta d; x += d.  az holds A x on entry and
// is not written.  MODE 1 (first step): beta = 0 and d is not read (it may
// hold anything).  MODE 2 (first step from x = 0): also A x = 0, so neither
// az nor x is read and x = d is written -- the values of MODE 1 on a zeroed
// x, up to the sign of zero.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_cheby_update(T* __restrict__ x, T* __restrict__ d, const T* __restrict__ az,
                                                      const T* __restrict__ b, const T* __restrict__ dinv,
                                                      int64_t count, T alpha, T beta)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const T z = MODE == 2 ? dinv[i] * b[i] : dinv[i] * (b[i] - az[i]);
        const T dn = MODE == 0 ? alpha * z + beta * d[i] : alpha * z;
        d[i] = dn;
        if constexpr (MODE == 2)
            x[i] = dn;
        else
            x[i] += dn;
    }
}

// w += alpha x, then the per-block sums of sum_owned w . y (y == w: the new w's squared norm):
// one modified Gram-Schmidt step fused with the next step's dot.  With alpha_dev
// the coefficient is alpha * alpha_dev[0], read on the device (a preceding
// reduction's result), so the host does not wait for it.
template <typename T, bool SELF>
__global__ __launch_bounds__(256) void k_axpy_dot(T* __restrict__ w, double alpha, const double* __restrict__ alpha_dev,
                                                  const T* __restrict__ x, const T* __restrict__ y,
                                                  const uint8_t* __restrict__ owned_lat, int64_t N, int nr, int ncomp,
                                                  double* __restrict__ part)
{
    double s = 0.0;
    const T a = alpha_dev ? (T)(alpha * alpha_dev[0]) : (T)alpha;
    const int lane = threadIdx.x & 63;
    for (ColumnWalk cw(N / nr, ncomp); cw.col < cw.ncol; cw.col += cw.step) {
        const bool own = owned_lat[cw.col % cw.nlat];
        T* wc = w + cw.col * nr;
        const T* xc = x + cw.col * nr;
        const T* yc = y + cw.col * nr;
        for (int r = lane; r < nr; r += 64) {
            const T wn = wc[r] + a * xc[r];
            wc[r] = wn;
            if (own) s += (double)wn * (double)(SELF ? wn : yc[r]);
        }
    }
    block_partial(s, part);
}

// y[i] -= scale * sum[0] (sum: a device scalar from a preceding reduction)
template <typename T>
__global__ __launch_bounds__(256) void k_sub_scaled(T* __restrict__ y, int64_t count, const double* __restrict__ sum,
                                                    double scale)
{
    const T v = (T)(scale * sum[0]);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] -= v;
}

// One thread per (duplicated lateral group, radial index), r fastest.
template <typename T>
__global__ void k_interface_sum(const int32_t* __restrict__ off, const int32_t* __restrict__ mem, int ngroups,
                                int nr, int64_t N, int ncomp, T* __restrict__ v)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)ngroups * nr) return;
    const int g = (int)(tid / nr), r = (int)(tid % nr);
    const int b = off[g], e = off[g + 1];
    for (int c = 0; c < ncomp; ++c) {
        T* vc = v + c * N;
        double s = 0.0;
        for (int k = b; k < e; ++k) s += (double)vc[(int64_t)mem[k] * nr + r];
        for (int k = b; k < e; ++k) vc[(int64_t)mem[k] * nr + r] = (T)s;
    }
}

// beta += sum over owned nodes of (x_node) cross u_node.  A wave walks radial
// columns (lateral nodes) two at a time, lanes over the radial nodes, up to
// three per lane per pass.  coords_owned holds the column's unit position at owned
// lateral nodes and 0 at the other copies, so every load of a pass (the
// coordinates, radii, the three components) is unconditional and independent
// of the others: a column costs one memory round trip.
template <typename T>
__global__ __launch_bounds__(256) void k_penalty_moment(const T* __restrict__ u,
                                                        const double* __restrict__ coords_owned,
                                                        const double* __restrict__ radii, int64_t N, int nr,
                                                        double* __restrict__ partial)
{
    constexpr int COLS = 2;  // columns in flight per wave
    double m0 = 0, m1 = 0, m2 = 0;
    const int lane = threadIdx.x & 63;
    const int64_t nlat = N / nr;
    // wave-uniform column index (scalar registers: scalar loads of the coordinates)
    const int64_t wave0 =
        (int64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t lat0 = wave0; lat0 < nlat; lat0 += COLS * waves) {
        for (int r0 = 0; r0 < nr; r0 += 192) {
            // clamped indices keep every load unconditional; columns past the mesh
            // and slots past the column get weight 0
            double c[COLS][3], rr[3], wt[COLS][3];
            T vx[COLS][3], vy[COLS][3], vz[COLS][3];
#pragma unroll
            for (int q = 0; q < COLS; ++q) {
                const int64_t lat = lat0 + q * waves;
                const int64_t latc = lat < nlat ? lat : nlat - 1;
                const double lw = lat < nlat ? 1.0 : 0.0;
                c[q][0] = coords_owned[3 * latc];
                c[q][1] = coords_owned[3 * latc + 1];
                c[q][2] = coords_owned[3 * latc + 2];
                const T* ux = u + latc * nr;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int r = r0 + 64 * k + lane;
                    const int rc = r < nr ? r : nr - 1;
                    wt[q][k] = r < nr ? lw : 0.0;
                    if (q == 0) rr[k] = radii[rc];
                    vx[q][k] = ux[rc];
                    vy[q][k] = ux[N + rc];
                    vz[q][k] = ux[2 * N + rc];
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // every load of the pass is issued before the arithmetic
#pragma unroll
            for (int q = 0; q < COLS; ++q)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double w = wt[q][k] * rr[k];
                    const double px = c[q][0] * w, py = c[q][1] * w, pz = c[q][2] * w;
                    const double ax = vx[q][k], ay = vy[q][k], az = vz[q][k];
                    m0 += py * az - pz * ay;
                    m1 += pz * ax - px * az;
                    m2 += px * ay - py * ax;
                }
        }
    }
    // per-block partial sums, added up by k_penalty_finalize: one store per
    // block instead of three same-address atomics, which serialise in L2
    m0 = block_total(m0);
    m1 = block_total(m1);
    m2 = block_total(m2);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = m0;
        partial[3 * blockIdx.x + 1] = m1;
        partial[3 * blockIdx.x + 2] = m2;
    }
}

// beta[k] = sum over the nblk partial sums (one block); then the penalty's
// rotation vector gamma = S beta into beta[13..16) (S = beta[4..13), row-major),
// which the apply kernels read instead of forming it per node
__global__ __launch_bounds__(256) void k_penalty_finalize(const double* __restrict__ partial, int nblk,
                                                          double* __restrict__ beta)
{
    double m[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < 3; ++k) m[k] += partial[3 * b + k];
    double bk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        bk[k] = block_total(m[k]);
        if (threadIdx.x == 0) beta[k] = bk[k];
    }
    if (threadIdx.x == 0) {
        const double* S = beta + 4;
#pragma unroll
        for (int i = 0; i < 3; ++i) beta[13 + i] = S[3 * i] * bk[0] + S[3 * i + 1] * bk[1] + S[3 * i + 2] * bk[2];
    }
}

// y (+)= gamma x p, gamma = S beta (S: 3x3 row-major)
template <typename T>
__global__ void k_penalty_init(const double* __restrict__ beta, const double* __restrict__ S,
                               const double* __restrict__ coords, const double* __restrict__ radii, int64_t N, int nr,
                               T* __restrict__ y, int add)
{
    const double g0 = S[0] * beta[0] + S[1] * beta[1] + S[2] * beta[2];
    const double g1 = S[3] * beta[0] + S[4] * beta[1] + S[5] * beta[2];
    const double g2 = S[6] * beta[0] + S[7] * beta[1] + S[8] * beta[2];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lat = i / nr;
        const double r = radii[i % nr];
        const double px = coords[3 * lat] * r, py = coords[3 * lat + 1] * r, pz = coords[3 * lat + 2] * r;
        const T vx = (T)(g1 * pz - g2 * py), vy = (T)(g2 * px - g0 * pz), vz = (T)(g0 * py - g1 * px);
        if (add) {
            y[i] += vx;
            y[N + i] += vy;
            y[2 * N + i] += vz;
        } else {
            y[i] = vx;
            y[N + i] = vy;
            y[2 * N + i] = vz;
        }
    }
}

template <typename T>
__global__ void k_mul(T* __restrict__ y, const T* __restrict__ d, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] *= d[i];
}

template <typename T>
__global__ void k_invert(T* __restrict__ y, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = T(1) / y[i];
}

template <typename TD, typename TS>
__global__ void k_cast(TD* __restrict__ y, const TS* __restrict__ x, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (TD)x[i];
}

// counter-based uniform [0,1) values (SplitMix64 of seed and global index):
// the same value for every run, independent of the launch shape
__device__ inline double splitmix_unit(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ void k_randomize(T* __restrict__ y, int64_t count, uint64_t seed)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (T)splitmix_unit(seed, (uint64_t)i);
}

template <typename T>
__global__ void k_nonfinite(const T* __restrict__ y, int64_t count, int* flag)
{
    int bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        bad |= !isfinite((double)y[i]);
    if (bad) atomicOr(flag, 1);
}

// y = x where the device flag is set (the decision stays on the device)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_if_flag(T* __restrict__ y, const T* __restrict__ x, int64_t count,
                                                      const int* __restrict__ flag)
{
    if (flag[0] == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x[i];
}

// out = max |y[i]| (non-negative doubles order like their bit patterns)
template <typename T>
__global__ __launch_bounds__(256) void k_max_abs(const T* __restrict__ y, int64_t count, unsigned long long* out)
{
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs((double)y[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// per-block sums of sum_owned a . b_k into part[k * gridDim.x + block], k < nb
// (multi-dot, one pass over a)
template <typename T, int NB>
__global__ __launch_bounds__(256) void k_multidot(const T* __restrict__ a, MultiDotArgs<T> bs,
                                                  const uint8_t* __restrict__ owned_lat, int64_t N, int nr, int ncomp,
                                                  double* __restrict__ part)
{
    double s[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) s[k] = 0.0;
    const int lane = threadIdx.x & 63;
    for (ColumnWalk cw(N / nr, ncomp); cw.col < cw.ncol; cw.col += cw.step) {
        if (!owned_lat[cw.col % cw.nlat]) continue;
        const int64_t base = cw.col * nr;
        for (int r = lane; r < nr; r += 64) {
            const double av = (double)a[base + r];
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (k < bs.nb) s[k] += av * (double)bs.b[k][base + r];
        }
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k >= bs.nb) break;
        block_partial(s[k], part + (int64_t)k * gridDim.x);
    }
}

inline unsigned grid_for(int64_t count, int block = 256, int64_t cap = 16384)
{
    int64_t g = (count + block - 1) / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

template <typename T>
void launch_fill(T* y, int64_t count, T value, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(count)), dim3(256), 0, s, y, count, value);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_lincomb(int64_t count, T* y, int nx, const double* c, const T* const* x, double c0, hipStream_t s)
{
    TN_REQUIRE(nx >= 0 && nx <= 3, "lincomb supports at most 3 inputs");
    LinArgs<T> a{};
    a.nx = nx;
    a.c0 = c0;
    for (int i = 0; i < nx; ++i) {
        a.x[i] = x[i];
        a.c[i] = c[i];
    }
    const dim3 gr(grid_for(count)), bl(256);
    switch (nx) {
    case 0: hipLaunchKernelGGL((k_lincomb<T, 0>), gr, bl, 0, s, y, count, a); break;
    case 1: hipLaunchKernelGGL((k_lincomb<T, 1>), gr, bl, 0, s, y, count, a); break;
    case 2: hipLaunchKernelGGL((k_lincomb<T, 2>), gr, bl, 0, s, y, count, a); break;
    default: hipLaunchKernelGGL((k_lincomb<T, 3>), gr, bl, 0, s, y, count, a); break;
    }
    TN_HIP(hipGetLastError());
}

static void launch_reduce_finalize(const double* part, int nblk, int nk, double* out_dev, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(256), 0, s, part, nblk, nk, out_dev);
    TN_HIP(hipGetLastError());
}

int reduce_blocks(const MeshLevel& m, int ncomp) { return (int)grid_for(m.N / m.nr * ncomp, 4, REDUCE_MAX_BLOCKS); }

template <typename T>
void launch_dot_owned(const MeshLevel& m, int ncomp, const T* a, const T* b, double* out_dev, double* part_dev,
                      hipStream_t s)
{
    const int g = reduce_blocks(m, ncomp);
    hipLaunchKernelGGL(k_dot_owned<T>, dim3(g), dim3(256), 0, s, a, b, m.d_owned_lat.p, m.N, m.nr, ncomp, part_dev);
    TN_HIP(hipGetLastError());
    launch_reduce_finalize(part_dev, g, 1, out_dev, s);
}

template <typename T>
void launch_sum_owned(const MeshLevel& m, const T* a, double* out_dev, double* part_dev, hipStream_t s)
{
    const int g = reduce_blocks(m, 1);
    hipLaunchKernelGGL(k_sum_owned<T>, dim3(g), dim3(256), 0, s, a, m.d_owned_lat.p, m.N, m.nr, part_dev);
    TN_HIP(hipGetLastError());
    launch_reduce_finalize(part_dev, g, 1, out_dev, s);
}

template <typename T>
void launch_cheby_update(T* x, T* d, const T* az, const T* b, const T* dinv, int64_t count, double alpha, double beta,
                         int mode, hipStream_t s)
{
    const dim3 gr(grid_for(count)), bl(256);
    if (mode == 2)
        hipLaunchKernelGGL((k_cheby_update<T, 2>), gr, bl, 0, s, x, d, az, b, dinv, count, (T)alpha, (T)beta);
    else if (mode == 1)
        hipLaunchKernelGGL((k_cheby_update<T, 1>), gr, bl, 0, s, x, d, az, b, dinv, count, (T)alpha, (T)beta);
    else
        hipLaunchKernelGGL((k_cheby_update<T, 0>), gr, bl, 0, s, x, d, az, b, dinv, count, (T)alpha, (T)beta);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_axpy_dot(const MeshLevel& m, int ncomp, T* w, double alpha, const T* x, const T* y, double* out_dev,
                     double* part_dev, hipStream_t s, const double* alpha_dev)
{
    const int g = reduce_blocks(m, ncomp);
    const dim3 gr(g), bl(256);
    if (y == w)
        hipLaunchKernelGGL((k_axpy_dot<T, true>), gr, bl, 0, s, w, alpha, alpha_dev, x, y, m.d_owned_lat.p, m.N, m.nr,
                           ncomp, part_dev);
    else
        hipLaunchKernelGGL((k_axpy_dot<T, false>), gr, bl, 0, s, w, alpha, alpha_dev, x, y, m.d_owned_lat.p, m.N, m.nr,
                           ncomp, part_dev);
    TN_HIP(hipGetLastError());
    launch_reduce_finalize(part_dev, g, 1, out_dev, s);
}

template <typename T>
void launch_copy_if_flag(T* y, const T* x, int64_t count, const int* flag_dev, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_if_flag<T>, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, y, x, count, flag_dev);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_sub_scaled(T* y, int64_t count, const double* sum_dev, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_scaled<T>, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, y, count, sum_dev, scale);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_interface_sum(const MeshLevel& m, int ncomp, T* v, hipStream_t s)
{
    const int ng = m.ngroups();
    if (ng <= 0) return;
    const int64_t total = (int64_t)ng * m.nr;
    hipLaunchKernelGGL(k_interface_sum<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, m.d_grp_off.p,
                       m.d_grp_mem.p, ng, m.nr, m.N, ncomp, v);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_penalty_moment(const MeshLevel& m, const T* u, double* beta_dev, double* partial_dev, hipStream_t s)
{
    const unsigned g = grid_for(m.N / m.nr, 4, PENALTY_MAX_BLOCKS);
    hipLaunchKernelGGL(k_penalty_moment<T>, dim3(g), dim3(256), 0, s, u, m.d_coords_owned.p, m.d_radii.p, m.N, m.nr,
                       partial_dev);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_penalty_finalize, dim3(1), dim3(256), 0, s, partial_dev, (int)g, beta_dev);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_penalty_init(const MeshLevel& m, const double* beta_dev, const double* S_dev, T* y, bool add,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_penalty_init<T>, dim3(grid_for(m.N)), dim3(256), 0, s, beta_dev, S_dev, m.d_coords.p,
                       m.d_radii.p, m.N, m.nr, y, (int)add);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_mul(T* y, const T* d, int64_t count, hipStream_t s)
{
    hipLaunchKernelGGL(k_mul<T>, dim3(grid_for(count)), dim3(256), 0, s, y, d, count);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_invert(T* y, int64_t count, hipStream_t s)
{
    hipLaunchKernelGGL(k_invert<T>, dim3(grid_for(count)), dim3(256), 0, s, y, count);
    TN_HIP(hipGetLastError());
}

template <typename TD, typename TS>
void launch_cast(TD* y, const TS* x, int64_t count, hipStream_t s)
{
    hipLaunchKernelGGL((k_cast<TD, TS>), dim3(grid_for(count)), dim3(256), 0, s, y, x, count);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_randomize(T* y, int64_t count, uint64_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(k_randomize<T>, dim3(grid_for(count)), dim3(256), 0, s, y, count, seed);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_nonfinite(const T* y, int64_t count, int* flag_dev, hipStream_t s)
{
    hipLaunchKernelGGL(k_nonfinite<T>, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, y, count, flag_dev);
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_max_abs(const T* y, int64_t count, double* out_dev, hipStream_t s)
{
    hipLaunchKernelGGL(k_max_abs<T>, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, y, count,
                       reinterpret_cast<unsigned long long*>(out_dev));
    TN_HIP(hipGetLastError());
}

template <typename T>
void launch_multidot(const MeshLevel& m, int ncomp, const T* a, const MultiDotArgs<T>& bs, double* out_dev,
                     double* part_dev, hipStream_t s)
{
    TN_REQUIRE(bs.nb >= 1 && bs.nb <= MultiDotArgs<T>::MAX, "multidot: 1..16 vectors");
    const int g = (int)grid_for(m.N / m.nr * ncomp, 4, REDUCE_MAX_BLOCKS / 2);
    static_assert(MultiDotArgs<T>::MAX * (REDUCE_MAX_BLOCKS / 2) <= REDUCE_PART_DOUBLES, "multidot partials");
    const dim3 gr(g), bl(256);
    if (bs.nb <= 4)
        hipLaunchKernelGGL((k_multidot<T, 4>), gr, bl, 0, s, a, bs, m.d_owned_lat.p, m.N, m.nr, ncomp, part_dev);
    else if (bs.nb <= 8)
        hipLaunchKernelGGL((k_multidot<T, 8>), gr, bl, 0, s, a, bs, m.d_owned_lat.p, m.N, m.nr, ncomp, part_dev);
    else
        hipLaunchKernelGGL((k_multidot<T, 16>), gr, bl, 0, s, a, bs, m.d_owned_lat.p, m.N, m.nr, ncomp, part_dev);
    TN_HIP(hipGetLastError());
    launch_reduce_finalize(part_dev, g, bs.nb, out_dev, s);
}

template void launch_fill<int>(int*, int64_t, int, hipStream_t);
template void launch_cast<double, float>(double*, const float*, int64_t, hipStream_t);
template void launch_cast<float, double>(float*, const double*, int64_t, hipStream_t);

#define TN_INST(T)                                                                                               \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                                   \
    template void launch_lincomb<T>(int64_t, T*, int, const double*, const T* const*, double, hipStream_t);     \
    template void launch_dot_owned<T>(const MeshLevel&, int, const T*, const T*, double*, double*, hipStream_t); \
    template void launch_sub_scaled<T>(T*, int64_t, const double*, double, hipStream_t);                         \
    template void launch_sum_owned<T>(const MeshLevel&, const T*, double*, double*, hipStream_t);                 \
    template void launch_axpy_dot<T>(const MeshLevel&, int, T*, double, const T*, const T*, double*, double*,    \
                                     hipStream_t, const double*);                                                \
    template void launch_copy_if_flag<T>(T*, const T*, int64_t, const int*, hipStream_t);                        \
    template void launch_cheby_update<T>(T*, T*, const T*, const T*, const T*, int64_t, double, double, int,    \
                                         hipStream_t);                                                           \
    template void launch_interface_sum<T>(const MeshLevel&, int, T*, hipStream_t);                               \
    template void launch_penalty_moment<T>(const MeshLevel&, const T*, double*, double*, hipStream_t);                    \
    template void launch_penalty_init<T>(const MeshLevel&, const double*, const double*, T*, bool, hipStream_t); \
    template void launch_mul<T>(T*, const T*, int64_t, hipStream_t);                                             \
    template void launch_max_abs<T>(const T*, int64_t, double*, hipStream_t);                                    \
    template void launch_invert<T>(T*, int64_t, hipStream_t);                                                    \
    template void launch_randomize<T>(T*, int64_t, uint64_t, hipStream_t);                                       \
    template void launch_nonfinite<T>(const T*, int64_t, int*, hipStream_t);                                     \
    template void launch_multidot<T>(const MeshLevel&, int, const T*, const MultiDotArgs<T>&, double*, double*, \
                                     hipStream_t);
TN_INST(double)
TN_INST(float)
#undef TN_INST

}  // namespace tn

// ---------------------------------------------------------------- owned-copy gather / scatter (dense coarse solve)
namespace tn {
namespace {
template <typename T>
__global__ void k_gather_owned(const T* __restrict__ v, int64_t N, const int64_t* __restrict__ idx, int64_t nown,
                               double* __restrict__ u)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nown) return;
    const int64_t k = i / 3;
    const int c = (int)(i % 3);
    u[i] = (double)v[c * N + idx[k]];
}
template <typename T>
__global__ void k_scatter_owned(T* __restrict__ v, int64_t N, const int64_t* __restrict__ idx, int64_t nown,
                                const double* __restrict__ u)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nown) return;
    const int64_t k = i / 3;
    const int c = (int)(i % 3);
    v[c * N + idx[k]] = (T)u[i];
}
template <typename T>
__global__ void k_set_owned_unit(T* __restrict__ v, int64_t N, const int64_t* __restrict__ idx, int c)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) v[c * N + idx[0]] = T(1);
}
}  // namespace

template <typename T>
void launch_gather_owned(const T* v, int64_t N, const int64_t* owned_idx, int64_t nown, double* u, hipStream_t s)
{
    const int64_t n = 3 * nown;
    hipLaunchKernelGGL(k_gather_owned<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, N, owned_idx, nown, u);
    TN_HIP(hipGetLastError());
}
template <typename T>
void launch_scatter_owned(T* v, int64_t N, const int64_t* owned_idx, int64_t nown, const double* u, hipStream_t s)
{
    const int64_t n = 3 * nown;
    hipLaunchKernelGGL(k_scatter_owned<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, N, owned_idx, nown,
                       u);
    TN_HIP(hipGetLastError());
}
template <typename T>
void launch_set_owned_unit(T* v, int64_t N, const int64_t* owned_idx, int c, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_owned_unit<T>, dim3(1), dim3(64), 0, s, v, N, owned_idx, c);
    TN_HIP(hipGetLastError());
}
namespace {
// y = A x for the dense column-major n x n coarsest-level inverse, in two
// passes.  Pass 1: a thread per row over a chunk of DENSE_GEMV_CHUNK columns
// (the grid's y index) writes partial[chunk][row]; the matrix is streamed once
// with non-temporal loads, x is read at wave-uniform addresses.  Pass 2 sums
// the chunks of a row in a fixed order (deterministic, no atomics).
__global__ __launch_bounds__(256) void k_dense_gemv_partial(const double* __restrict__ A, const double* __restrict__ x,
                                                            int n, double* __restrict__ partial)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int j0 = blockIdx.y * DENSE_GEMV_CHUNK;
    const int j1 = min(n, j0 + DENSE_GEMV_CHUNK);
    const double* a = A + (int64_t)j0 * n + i;
    double acc = 0.0;
    int j = j0;
    for (; j + 8 <= j1; j += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(a + (int64_t)(j - j0 + k) * n);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k] * x[j + k];
    }
    for (; j < j1; ++j) acc += __builtin_nontemporal_load(a + (int64_t)(j - j0) * n) * x[j];
    partial[(int64_t)blockIdx.y * n + i] = acc;
}
__global__ __launch_bounds__(256) void k_dense_gemv_sum(const double* __restrict__ partial, int n, int nchunk,
                                                        double* __restrict__ y)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[(int64_t)c * n + i];
    y[i] = s;
}
}  // namespace

void launch_dense_gemv(const double* A, const double* x, int n, double* partial, double* y, hipStream_t s)
{
    TN_REQUIRE(n > 0 && dense_gemv_chunks(n) <= 65535, "dense gemv: size out of range");
    const int nchunk = dense_gemv_chunks(n);
    const unsigned rows = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_dense_gemv_partial, dim3(rows, (unsigned)nchunk), dim3(256), 0, s, A, x, n, partial);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_dense_gemv_sum, dim3(rows), dim3(256), 0, s, partial, n, nchunk, y);
    TN_HIP(hipGetLastError());
}
template void launch_gather_owned<double>(const double*, int64_t, const int64_t*, int64_t, double*, hipStream_t);
template void launch_gather_owned<float>(const float*, int64_t, const int64_t*, int64_t, double*, hipStream_t);
template void launch_scatter_owned<double>(double*, int64_t, const int64_t*, int64_t, const double*, hipStream_t);
template void launch_scatter_owned<float>(float*, int64_t, const int64_t*, int64_t, const double*, hipStream_t);
template void launch_set_owned_unit<double>(double*, int64_t, const int64_t*, int, hipStream_t);
template void launch_set_owned_unit<float>(float*, int64_t, const int64_t*, int, hipStream_t);
}  // namespace tn
