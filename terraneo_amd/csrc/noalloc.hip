// Entry point that allocates nothing: the owned-node dot product with its
// per-block partial sums and its result slot in a file-scope device array
// instead of the mesh's allocated reduction buffers.  For callers that cannot
// call hipMalloc (graph capture, tight serving loops, harnesses that forbid
// allocation).  No counterpart in the reference.
//
// The array `scratch` is one per GPU (each device holds its own copy at its
// own address until the process ends).  In host code its name is not a device
// address: the address is taken once per device with hipGetSymbolAddress and
// only that pointer is handed to kernels.  Every call on a device shares that
// device's array and leaves its contents behind: the result slot is cleared
// by a fill kernel at every call and each block overwrites its partial-sum
// slot, so no state carries over; calls that overlap on one device (two
// streams, two host threads) race on the array.
#include <map>
#include <mutex>

#include "tn_eps_march.hpp"
#include "tn_internal.hpp"

namespace tn {
namespace {

// slot 0: the result; [NOALLOC_PART_OFFSET, +REDUCE_MAX_BLOCKS): per-block partial sums
constexpr int NOALLOC_PART_OFFSET = 32;  // 256 bytes
constexpr int NOALLOC_DOUBLES = NOALLOC_PART_OFFSET + REDUCE_MAX_BLOCKS;
static_assert(NOALLOC_DOUBLES * sizeof(double) == 16640, "scratch size");
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of that device's scratch

double* scratch_for_current_device()
{
    int dev = 0;
    TN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void* p = nullptr;
    TN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TN_REQUIRE(p != nullptr, "no-alloc dot: device scratch has no address");
    g_scratch_ptr[dev] = static_cast<double*>(p);
    return static_cast<double*>(p);
}

}  // namespace
}  // namespace tn

extern "C" int tn_dot_noalloc(const tn_mesh* m, int dtype, int ncomp, const void* a, const void* b, double* out,
                              tn_stream_t s)
{
    try {
        TN_REQUIRE(m && a && b && out, "null pointer");
        TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "no-alloc dot: dtype must be TN_F64 or TN_F32");
        TN_REQUIRE(ncomp >= 1 && ncomp <= 3, "no-alloc dot: ncomp must be 1, 2 or 3");
        const tn::MeshLevel& L = m->lvl;
        TN_REQUIRE(L.uploaded,
                   "no-alloc dot: the mesh is not resident on the device yet (any other device call on it uploads it)");
        const int nblk = tn::reduce_blocks(L, ncomp);
        TN_REQUIRE(nblk >= 1 && nblk <= tn::REDUCE_MAX_BLOCKS,
                   "no-alloc dot: more partial sums than the device scratch holds (2048)");
        double* sp = tn::scratch_for_current_device();
        double* res = sp;
        double* part = sp + tn::NOALLOC_PART_OFFSET;
        const hipStream_t st = (hipStream_t)s;
        tn::launch_fill<double>(res, 1, 0.0, st);  // a kernel, not a memset node
        if (dtype == TN_F64)
            tn::launch_dot_owned<double>(L, ncomp, (const double*)a, (const double*)b, res, part, st);
        else
            tn::launch_dot_owned<float>(L, ncomp, (const float*)a, (const float*)b, res, part, st);
        TN_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, st));
        TN_HIP(hipStreamSynchronize(st));
        tn::check_march_status();
    } catch (const std::exception& e) {
        tn::set_last_error(e.what());
        return -1;
    }
    return 0;
}
