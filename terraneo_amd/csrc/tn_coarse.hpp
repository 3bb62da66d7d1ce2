// Dense direct coarsest-level solver of the multigrid (optional).
#pragma once

#include <rocblas/rocblas.h>

#include "tn_solvers.hpp"

struct tn_eps;

namespace tn {

// owned-copy gather/scatter between a 3-component field and the unique-dof
// vector u[3 k + c] (k: owned node in storage order), FP64 unique values
template <typename T>
void launch_gather_owned(const T* v, int64_t N, const int64_t* owned_idx, int64_t nown, double* u, hipStream_t s);
template <typename T>
void launch_scatter_owned(T* v, int64_t N, const int64_t* owned_idx, int64_t nown, const double* u, hipStream_t s);
// v[c N + *owned_idx] = 1
template <typename T>
void launch_set_owned_unit(T* v, int64_t N, const int64_t* owned_idx, int c, hipStream_t s);

// y = A x, A dense column-major n x n (n < 65535 * DENSE_GEMV_CHUNK);
// partial holds dense_gemv_chunks(n) * n doubles of scratch
constexpr int DENSE_GEMV_CHUNK = 128;
inline int dense_gemv_chunks(int n) { return (n + DENSE_GEMV_CHUNK - 1) / DENSE_GEMV_CHUNK; }
void launch_dense_gemv(const double* A, const double* x, int n, double* partial, double* y, hipStream_t s);

class DenseCoarse : public Solver {
  public:
    DenseCoarse(tn_eps* op, hipStream_t s);
    ~DenseCoarse() override;
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;

  private:
    tn_eps* op_;
    int64_t nown_ = 0, n_ = 0;
    DevBuf<int64_t> owned_idx_;
    DevBuf<double> inv_, xu_, yu_, part_;
    rocblas_handle handle_ = nullptr;
};

}  // namespace tn
