// Dense direct solve of the coarsest multigrid level (optional; the
// reference solves it with unpreconditioned PCG, multigrid.hpp:119-211 +
// pcg.hpp:78-153, which is also available and is the default).
//
// The coarse operator is assembled column by column over the unique degrees
// of freedom (one per owned node and component) by applying it to unit
// vectors -- so every boundary treatment and the free-slip penalty are
// exactly those of the operator -- then inverted once with rocSOLVER
// (LU with partial pivoting, getrf + getri).  A coarse solve is then one
// gather of the owned values, one FP64 product with the inverse
// (launch_dense_gemv: one pass over the matrix) and one scatter plus
// interface sum, instead of hundreds of host-synchronised PCG iterations.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "tn_coarse.hpp"
#include "tn_objects.hpp"

namespace tn {

namespace {
void rb_check(rocblas_status st, const char* what)
{
    if (st != rocblas_status_success)
        throw std::runtime_error(std::string("rocBLAS/rocSOLVER failure in ") + what + ": " +
                                 rocblas_status_to_string(st));
}
}  // namespace

DenseCoarse::DenseCoarse(tn_eps* op, hipStream_t s) : op_(op)
{
    MeshLevel& m = const_cast<tn_mesh*>(op->mesh)->lvl.dev();
    std::vector<int64_t> idx;
    for (int64_t lat = 0; lat < 10LL * m.n * m.n; ++lat)
        if (m.owned_lat[lat])
            for (int r = 0; r < m.nr; ++r) idx.push_back(lat * m.nr + r);
    nown_ = (int64_t)idx.size();
    n_ = 3 * nown_;
    TN_REQUIRE(n_ <= 24000, "dense coarse solve: coarsest level too large (at most 8000 owned nodes)");
    owned_idx_.upload(idx);
    inv_.alloc((size_t)n_ * n_);
    xu_.alloc(n_);
    yu_.alloc(n_);
    part_.alloc((size_t)dense_gemv_chunks((int)n_) * n_);
    rb_check(rocblas_create_handle(&handle_), "create_handle");

    // assemble: column j = A e_j restricted to the owned copies
    const int esize = op->dtype == TN_F64 ? 8 : 4;
    const int64_t N = m.N;
    DevBuf<unsigned char> e((size_t)3 * N * esize), y((size_t)3 * N * esize);
    const int am = op->apply_mode, cm = op->comm_mode;
    const bool dg = op->diagonal;
    op->apply_mode = TN_REPLACE;
    op->comm_mode = TN_COMM_ADDITIVE;
    op->diagonal = false;
    for (int64_t j = 0; j < n_; ++j) {
        const int64_t k = j / 3;
        const int c = (int)(j % 3);
        if (op->dtype == TN_F64) {
            launch_fill<double>((double*)e.p, 3 * N, 0.0, s);
            launch_set_owned_unit<double>((double*)e.p, N, owned_idx_.p + k, c, s);
            launch_interface_sum<double>(m, 3, (double*)e.p, s);
        } else {
            launch_fill<float>((float*)e.p, 3 * N, 0.0f, s);
            launch_set_owned_unit<float>((float*)e.p, N, owned_idx_.p + k, c, s);
            launch_interface_sum<float>(m, 3, (float*)e.p, s);
        }
        eps_apply(op, e.p, y.p, s);
        if (op->dtype == TN_F64)
            launch_gather_owned<double>((const double*)y.p, N, owned_idx_.p, nown_, inv_.p + j * n_, s);
        else
            launch_gather_owned<float>((const float*)y.p, N, owned_idx_.p, nown_, inv_.p + j * n_, s);
    }
    op->apply_mode = am;
    op->comm_mode = cm;
    op->diagonal = dg;

    // invert (column-major n x n)
    rb_check(rocblas_set_stream(handle_, s), "set_stream");
    DevBuf<rocblas_int> ipiv(n_), info(1);
    rb_check(rocsolver_dgetrf(handle_, (rocblas_int)n_, (rocblas_int)n_, inv_.p, (rocblas_int)n_, ipiv.p, info.p),
             "dgetrf");
    rb_check(rocsolver_dgetri(handle_, (rocblas_int)n_, inv_.p, (rocblas_int)n_, ipiv.p, info.p), "dgetri");
    rocblas_int h_info = 0;
    TN_HIP(hipMemcpyAsync(&h_info, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    TN_HIP(hipStreamSynchronize(s));
    TN_REQUIRE(h_info == 0, "dense coarse solve: the coarsest operator is singular");
}

DenseCoarse::~DenseCoarse()
{
    if (handle_) rocblas_destroy_handle(handle_);
}

void DenseCoarse::solve(Ctx& c, LinOp&, const BVec& x, const BVec& b)
{
    MeshLevel& m = const_cast<tn_mesh*>(op_->mesh)->lvl.dev();
    const Field& fb = b.b[0];
    const Field& fx = x.b[0];
    if (fb.dtype == TN_F64)
        launch_gather_owned<double>((const double*)fb.p, m.N, owned_idx_.p, nown_, xu_.p, c.s);
    else
        launch_gather_owned<float>((const float*)fb.p, m.N, owned_idx_.p, nown_, xu_.p, c.s);
    launch_dense_gemv(inv_.p, xu_.p, (int)n_, part_.p, yu_.p, c.s);
    if (fx.dtype == TN_F64) {
        launch_fill<double>((double*)fx.p, 3 * m.N, 0.0, c.s);
        launch_scatter_owned<double>((double*)fx.p, m.N, owned_idx_.p, nown_, yu_.p, c.s);
        launch_interface_sum<double>(m, 3, (double*)fx.p, c.s);
    } else {
        launch_fill<float>((float*)fx.p, 3 * m.N, 0.0f, c.s);
        launch_scatter_owned<float>((float*)fx.p, m.N, owned_idx_.p, nown_, yu_.p, c.s);
        launch_interface_sum<float>(m, 3, (float*)fx.p, c.s);
    }
}

}  // namespace tn
