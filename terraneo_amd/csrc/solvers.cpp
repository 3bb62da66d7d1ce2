// Krylov, Chebyshev and multigrid drivers (host C++ over HIP level-1 kernels).
//
// Each algorithm follows the reference recurrence step by step, including
// its conventions (initial guesses, restart and stopping rules), so residual
// histories are comparable:
//   Chebyshev + power iteration  linalg/solvers/chebyshev.hpp:100-168, power_iteration.hpp:13-47
//   PCG                          linalg/solvers/pcg.hpp:78-153
//   PMINRES                      linalg/solvers/pminres.hpp:75-185
//   FGMRES(m)                    linalg/solvers/fgmres.hpp:108-324
//   Multigrid V-cycle            linalg/solvers/multigrid.hpp:119-211
//   block preconditioners        linalg/solvers/block_preconditioner_2x2.hpp:91-95, 205-221
//   DiagonalSolver               linalg/solvers/diagonal_solver.hpp:34-49
#include "tn_solvers.hpp"
#include "tn_eps_march.hpp"
#include "tn_mg.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace tn {

// ------------------------------------------------------------------ vectors
BVec VecStore::make_like(const BVec& like, int dtype)
{
    BVec v = like;
    for (int i = 0; i < like.nb; ++i) {
        v.b[i].dtype = dtype;
        void* p = nullptr;
        TN_HIP(hipMalloc(&p, (size_t)v.b[i].count() * v.b[i].esize()));
        TN_HIP(hipMemset(p, 0, (size_t)v.b[i].count() * v.b[i].esize()));
        ptrs_.push_back(p);
        v.b[i].p = p;
    }
    return v;
}

BVec VecStore::make(const BVec& like) { return make_like(like, like.b[0].dtype); }

VecStore::~VecStore()
{
    for (void* p : ptrs_) (void)hipFree(p);
}

Ctx::Ctx()
{
    red.alloc(32);
    part.alloc(REDUCE_PART_DOUBLES);
    flag.alloc(1);
    col.alloc(64);
}

void Ctx::fetch_col(int n, double* out, int* flag_out)
{
    TN_REQUIRE(n >= 1 && (size_t)n <= col.n, "fetch_col: slot count");
    TN_HIP(hipMemcpyAsync(out, col.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (flag_out) TN_HIP(hipMemcpyAsync(flag_out, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TN_HIP(hipStreamSynchronize(s));
    check_march_status();
}

double Ctx::fetch(int slots, double* out)
{
    double h[32];
    TN_HIP(hipMemcpyAsync(h, red.p, sizeof(double) * slots, hipMemcpyDeviceToHost, s));
    TN_HIP(hipStreamSynchronize(s));
    check_march_status();
    if (out)
        for (int i = 0; i < slots; ++i) out[i] = h[i];
    return h[0];
}

void v_fill(Ctx& c, const BVec& y, double v)
{
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        if (f.dtype == TN_F64)
            launch_fill<double>((double*)f.p, f.count(), v, c.s);
        else
            launch_fill<float>((float*)f.p, f.count(), (float)v, c.s);
    }
}

void v_lincomb(Ctx& c, const BVec& y, std::initializer_list<double> coef, std::initializer_list<const BVec*> xs,
               double c0)
{
    const int nx = (int)xs.size();
    TN_REQUIRE(nx <= 3 && (int)coef.size() == nx, "lincomb: at most 3 inputs");
    double cc[3];
    int k = 0;
    for (double v : coef) cc[k++] = v;
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        const void* xp[3];
        k = 0;
        for (const BVec* x : xs) {
            TN_REQUIRE(x->b[i].dtype == f.dtype, "lincomb: dtype mismatch");
            xp[k++] = x->b[i].p;
        }
        if (f.dtype == TN_F64)
            launch_lincomb<double>(f.count(), (double*)f.p, nx, cc, (const double* const*)xp, c0, c.s);
        else
            launch_lincomb<float>(f.count(), (float*)f.p, nx, cc, (const float* const*)xp, c0, c.s);
    }
}

void v_assign(Ctx& c, const BVec& y, const BVec& x) { v_lincomb(c, y, {1.0}, {&x}); }

double v_dot(Ctx& c, const BVec& a, const BVec& b)
{
    launch_fill<double>(c.red.p, 1, 0.0, c.s);
    for (int i = 0; i < a.nb; ++i) {
        const Field& f = a.b[i];
        if (f.dtype == TN_F64)
            launch_dot_owned<double>(*f.m, f.ncomp, (const double*)f.p, (const double*)b.b[i].p, c.red.p, c.part.p,
                                     c.s);
        else
            launch_dot_owned<float>(*f.m, f.ncomp, (const float*)f.p, (const float*)b.b[i].p, c.red.p, c.part.p, c.s);
    }
    return c.fetch(1);
}

double v_axpy_dot(Ctx& c, const BVec& w, double alpha, const BVec& x, const BVec& y)
{
    launch_fill<double>(c.red.p, 1, 0.0, c.s);
    for (int i = 0; i < w.nb; ++i) {
        const Field& f = w.b[i];
        if (f.dtype == TN_F64)
            launch_axpy_dot<double>(*f.m, f.ncomp, (double*)f.p, alpha, (const double*)x.b[i].p,
                                    (const double*)y.b[i].p, c.red.p, c.part.p, c.s);
        else
            launch_axpy_dot<float>(*f.m, f.ncomp, (float*)f.p, alpha, (const float*)x.b[i].p, (const float*)y.b[i].p,
                                   c.red.p, c.part.p, c.s);
    }
    return c.fetch(1);
}

void v_dot_to(Ctx& c, const BVec& a, const BVec& b, double* out_dev)
{
    for (int i = 0; i < a.nb; ++i) {
        const Field& f = a.b[i];
        if (f.dtype == TN_F64)
            launch_dot_owned<double>(*f.m, f.ncomp, (const double*)f.p, (const double*)b.b[i].p, out_dev, c.part.p,
                                     c.s);
        else
            launch_dot_owned<float>(*f.m, f.ncomp, (const float*)f.p, (const float*)b.b[i].p, out_dev, c.part.p, c.s);
    }
}

void v_sub_axpy_dot_to(Ctx& c, const BVec& w, const double* coef_dev, const BVec& x, const BVec& y, double* out_dev)
{
    for (int i = 0; i < w.nb; ++i) {
        const Field& f = w.b[i];
        if (f.dtype == TN_F64)
            launch_axpy_dot<double>(*f.m, f.ncomp, (double*)f.p, -1.0, (const double*)x.b[i].p,
                                    (const double*)y.b[i].p, out_dev, c.part.p, c.s, coef_dev);
        else
            launch_axpy_dot<float>(*f.m, f.ncomp, (float*)f.p, -1.0, (const float*)x.b[i].p, (const float*)y.b[i].p,
                                   out_dev, c.part.p, c.s, coef_dev);
    }
}

void v_multidot(Ctx& c, const BVec& a, const std::vector<const BVec*>& bs, double* out)
{
    const int n = (int)bs.size();
    TN_REQUIRE(n >= 1 && n <= 16, "multidot: 1..16 vectors");
    launch_fill<double>(c.red.p, n, 0.0, c.s);
    for (int i = 0; i < a.nb; ++i) {
        const Field& f = a.b[i];
        if (f.dtype == TN_F64) {
            MultiDotArgs<double> args{};
            args.nb = n;
            for (int k = 0; k < n; ++k) args.b[k] = (const double*)bs[k]->b[i].p;
            launch_multidot<double>(*f.m, f.ncomp, (const double*)f.p, args, c.red.p, c.part.p, c.s);
        } else {
            MultiDotArgs<float> args{};
            args.nb = n;
            for (int k = 0; k < n; ++k) args.b[k] = (const float*)bs[k]->b[i].p;
            launch_multidot<float>(*f.m, f.ncomp, (const float*)f.p, args, c.red.p, c.part.p, c.s);
        }
    }
    c.fetch(n, out);
}

void v_mul(Ctx& c, const BVec& y, const BVec& d)
{
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        if (f.dtype == TN_F64)
            launch_mul<double>((double*)f.p, (const double*)d.b[i].p, f.count(), c.s);
        else
            launch_mul<float>((float*)f.p, (const float*)d.b[i].p, f.count(), c.s);
    }
}

void v_invert(Ctx& c, const BVec& y)
{
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        if (f.dtype == TN_F64)
            launch_invert<double>((double*)f.p, f.count(), c.s);
        else
            launch_invert<float>((float*)f.p, f.count(), c.s);
    }
}

void v_randomize(Ctx& c, const BVec& y, uint64_t seed)
{
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        if (f.dtype == TN_F64)
            launch_randomize<double>((double*)f.p, f.count(), seed + 7919ull * i, c.s);
        else
            launch_randomize<float>((float*)f.p, f.count(), seed + 7919ull * i, c.s);
    }
}

void v_replace_nonfinite(Ctx& c, const BVec& y, const BVec& x)
{
    launch_fill<int>(c.flag.p, 1, 0, c.s);
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        if (f.dtype == TN_F64)
            launch_nonfinite<double>((const double*)f.p, f.count(), c.flag.p, c.s);
        else
            launch_nonfinite<float>((const float*)f.p, f.count(), c.flag.p, c.s);
    }
    for (int i = 0; i < y.nb; ++i) {
        const Field& f = y.b[i];
        const Field& g = x.b[i];
        TN_REQUIRE(f.dtype == g.dtype && f.count() == g.count(), "replace_nonfinite: block mismatch");
        if (f.dtype == TN_F64)
            launch_copy_if_flag<double>((double*)f.p, (const double*)g.p, f.count(), c.flag.p, c.s);
        else
            launch_copy_if_flag<float>((float*)f.p, (const float*)g.p, f.count(), c.flag.p, c.s);
    }
}

void v_convert(Ctx& c, const BVec& dst, const BVec& src)
{
    for (int i = 0; i < dst.nb; ++i) {
        const Field& d = dst.b[i];
        const Field& s = src.b[i];
        if (d.dtype == s.dtype) {
            TN_HIP(hipMemcpyAsync(d.p, s.p, (size_t)d.count() * d.esize(), hipMemcpyDeviceToDevice, c.s));
        } else if (d.dtype == TN_F32) {
            launch_cast<float, double>((float*)d.p, (const double*)s.p, d.count(), c.s);
        } else {
            launch_cast<double, float>((double*)d.p, (const float*)s.p, d.count(), c.s);
        }
    }
}

// ------------------------------------------------------------------ simple solvers
void DiagonalSolver::solve(Ctx& c, LinOp&, const BVec& x, const BVec& b)
{
    v_assign(c, x, b);
    v_mul(c, x, d_);
}

void Solver::solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b)
{
    v_fill(c, x, 0.0);
    solve(c, A, x, b);
}

void BlockDiagonal::solve(Ctx& c, LinOp&, const BVec& x, const BVec& b) { run(c, x, b, false); }

void BlockDiagonal::solve_zero(Ctx& c, LinOp&, const BVec& x, const BVec& b) { run(c, x, b, true); }

void BlockDiagonal::run(Ctx& c, const BVec& x, const BVec& b, bool zero)
{
    BVec xu, xp, bu, bp;
    xu.b[0] = x.b[0];
    xp.b[0] = x.b[1];
    bu.b[0] = b.b[0];
    bp.b[0] = b.b[1];
    if (zero) {
        p11_->solve_zero(c, *A11_, xu, bu);
        p22_->solve_zero(c, *A22_, xp, bp);
    } else {
        p11_->solve(c, *A11_, xu, bu);
        p22_->solve(c, *A22_, xp, bp);
    }
}

void BlockTriangular::solve(Ctx& c, LinOp&, const BVec& x, const BVec& b) { run(c, x, b, false); }

void BlockTriangular::solve_zero(Ctx& c, LinOp&, const BVec& x, const BVec& b) { run(c, x, b, true); }

void BlockTriangular::run(Ctx& c, const BVec& x, const BVec& b, bool zero)
{
    BVec xu, xp, bu, bp, tu;
    xu.b[0] = x.b[0];
    xp.b[0] = x.b[1];
    bu.b[0] = b.b[0];
    bp.b[0] = b.b[1];
    tu.b[0] = tmp_.b[0];
    if (zero)
        p22_->solve_zero(c, *A22_, xp, bp);  // x2 = S^-1 b2
    else
        p22_->solve(c, *A22_, xp, bp);
    B12_->apply(c, xp, tu);              // t1 = B^T x2
    v_lincomb(c, tu, {1.0, 1.0}, {&tu, &bu});
    v_lincomb(c, xp, {-1.0}, {&xp});
    if (zero)
        p11_->solve_zero(c, *A11_, xu, tu);  // x1 = A^-1 (b1 + B^T x2)
    else
        p11_->solve(c, *A11_, xu, tu);
}

void PrecisionBridge::solve(Ctx& c, LinOp&, const BVec& x, const BVec& b)
{
    v_convert(c, xi_, x);
    v_convert(c, bi_, b);
    inner_->solve(c, *op_, xi_, bi_);
    v_convert(c, x, xi_);
}

void PrecisionBridge::solve_zero(Ctx& c, LinOp&, const BVec& x, const BVec& b)
{
    v_convert(c, bi_, b);
    inner_->solve_zero(c, *op_, xi_, bi_);
    v_convert(c, x, xi_);
}

// ------------------------------------------------------------------ Chebyshev
Chebyshev::Chebyshev(int order, int iterations, BVec inv_diag, BVec t0, BVec t1, int power_iterations, uint64_t seed)
    : order_(order), iters_(iterations), power_its_(power_iterations), seed_(seed), dinv_(inv_diag), d_(t0), z_(t1)
{
    TN_REQUIRE(order >= 1, "Chebyshev order must be at least 1");
}

void Chebyshev::apply_scale(Ctx& c, const BVec& z)
{
    if (scale)
        scale(c, z);
    else
        v_mul(c, z, dinv_);
}

namespace {
struct DiagScaled : LinOp {
    LinOp& A;
    std::function<void(Ctx&, const BVec&)> sc;
    DiagScaled(LinOp& a, std::function<void(Ctx&, const BVec&)> s) : A(a), sc(std::move(s)) {}
    void apply(Ctx& c, const BVec& x, const BVec& y) override
    {
        A.apply(c, x, y);
        sc(c, y);
    }
};
}  // namespace

void Chebyshev::solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) { run(c, A, x, b, false); }

void Chebyshev::solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) { run(c, A, x, b, true); }

void Chebyshev::run(Ctx& c, LinOp& A, const BVec& x, const BVec& b, bool zero)
{
    if (need_lambda_) {
        // power iteration on D^-1 A from a deterministic random start
        DiagScaled op(A, [this](Ctx& cc, const BVec& z) { apply_scale(cc, z); });
        BVec& it = d_;
        BVec& aux = z_;
        v_randomize(c, it, seed_);
        double nrm = std::sqrt(v_dot(c, it, it));
        v_lincomb(c, it, {1.0 / nrm}, {&it});
        op.apply(c, it, aux);
        double radius = 0.0;
        for (int k = 0; k < power_its_; ++k) {
            nrm = std::sqrt(v_dot(c, aux, aux));
            v_lincomb(c, it, {1.0 / nrm}, {&aux});
            op.apply(c, it, aux);
            radius = v_dot(c, it, aux);
        }
        lambda_ = radius;
        need_lambda_ = false;
    }
    const double lmax = 1.5 * lambda_, lmin = 0.1 * lambda_;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;  // carried across outer iterations, as the reference does
    // one step: z = D^-1 (b - A x); d = alpha z + beta d; x += d (fused into one
    // pass over the vectors unless a custom scaling is set).  from_zero (the
    // first step of solve_zero): A x = 0 is not computed and x = d is written.
    auto step = [&](double alpha, double beta, bool first, bool from_zero) {
        if (!from_zero) A.apply(c, x, z_);
        const int mode = from_zero ? 2 : (first ? 1 : 0);
        if (!scale) {
            for (int i = 0; i < x.nb; ++i) {
                const Field& f = x.b[i];
                if (f.dtype == TN_F64)
                    launch_cheby_update<double>((double*)f.p, (double*)d_.b[i].p, (const double*)z_.b[i].p,
                                                (const double*)b.b[i].p, (const double*)dinv_.b[i].p, f.count(), alpha,
                                                beta, mode, c.s);
                else
                    launch_cheby_update<float>((float*)f.p, (float*)d_.b[i].p, (const float*)z_.b[i].p,
                                               (const float*)b.b[i].p, (const float*)dinv_.b[i].p, f.count(), alpha,
                                               beta, mode, c.s);
            }
            return;
        }
        if (from_zero)
            v_assign(c, z_, b);
        else
            v_lincomb(c, z_, {1.0, -1.0}, {&b, &z_});
        apply_scale(c, z_);
        if (first)
            v_lincomb(c, d_, {alpha}, {&z_});
        else
            v_lincomb(c, d_, {alpha, beta}, {&z_, &d_});
        if (from_zero)
            v_assign(c, x, d_);
        else
            v_lincomb(c, x, {1.0, 1.0}, {&x, &d_});
    };
    for (int it = 0; it < iters_; ++it) {
        step(1.0 / theta, 0.0, true, zero && it == 0);
        for (int i = 2; i <= order_; ++i) {
            const double rho_new = 1.0 / (2.0 * sigma - rho);
            const double beta = rho_new * rho;
            const double alpha = 2.0 * rho_new / delta;
            step(alpha, beta, false, false);
            rho = rho_new;
        }
    }
}

// ------------------------------------------------------------------ PCG
PCG::PCG(KrylovParams p, std::vector<BVec> tmps, Solver* prec) : p_(p), t_(std::move(tmps)), prec_(prec)
{
    TN_REQUIRE(t_.size() >= 4, "PCG needs 4 temporaries");
}

void PCG::solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b)
{
    Solver& P = prec_ ? *prec_ : (Solver&)id_;
    BVec &r = t_[0], &pp = t_[1], &ap = t_[2], &z = t_[3];
    info = IterInfo{};
    A.apply(c, x, r);
    v_lincomb(c, r, {1.0, -1.0}, {&b, &r});
    // The reference's PCG hands the previous z to the preconditioner
    // (pcg.hpp:93,145).  Its PCG callers only use stateless preconditioners
    // (identity, Jacobi), for which that is immaterial; a multigrid V-cycle
    // starts from the content of its output, so z is cleared first -- the
    // convention PMINRES and FGMRES follow (pminres.hpp:130, fgmres.hpp:177).
    P.solve_zero(c, A, z, r);
    v_assign(c, pp, z);
    const double r0 = std::sqrt(v_dot(c, r, r));
    info.initial_residual = info.final_residual = r0;
    info.history.push_back(r0);
    if (r0 < p_.atol) return;
    for (int it = 1; it <= p_.max_iterations; ++it) {
        const double alpha_num = v_dot(c, z, r);
        A.apply(c, pp, ap);
        const double alpha = alpha_num / v_dot(c, ap, pp);
        v_lincomb(c, x, {1.0, alpha}, {&x, &pp});
        v_lincomb(c, r, {1.0, -alpha}, {&r, &ap});
        const double res = std::sqrt(v_dot(c, r, r));
        info.iterations = it;
        info.final_residual = res;
        info.history.push_back(res);
        if (res / r0 <= p_.rtol || res < p_.atol) return;
        P.solve_zero(c, A, z, r);
        const double beta = v_dot(c, z, r) / alpha_num;
        v_lincomb(c, pp, {1.0, beta}, {&z, &pp});
    }
}

// ------------------------------------------------------------------ PMINRES
PMINRES::PMINRES(KrylovParams p, std::vector<BVec> tmps, Solver* prec) : p_(p), t_(std::move(tmps)), prec_(prec)
{
    TN_REQUIRE(t_.size() >= 7, "PMINRES needs 7 temporaries");
}

void PMINRES::solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b)
{
    Solver& P = prec_ ? *prec_ : (Solver&)id_;
    BVec az = t_[0], vjm1 = t_[1], vj = t_[2], wjm1 = t_[3], wj = t_[4], zjp1 = t_[5], z = t_[6];
    info = IterInfo{};
    v_fill(c, vjm1, 0.0);
    v_fill(c, wj, 0.0);
    v_fill(c, wjm1, 0.0);
    A.apply(c, x, vj);
    v_lincomb(c, vj, {1.0, -1.0}, {&b, &vj});
    P.solve(c, A, z, vj);
    double gamma_jm1 = 1.0, gamma_j = std::sqrt(v_dot(c, z, vj));
    double eta = gamma_j, s_jm1 = 0, s_j = 0, c_jm1 = 1, c_j = 1;
    const double r0 = gamma_j;
    info.initial_residual = info.final_residual = r0;
    info.history.push_back(r0);
    if (r0 < p_.atol) return;
    for (int it = 1; it <= p_.max_iterations; ++it) {
        v_lincomb(c, z, {1.0 / gamma_j}, {&z});
        A.apply(c, z, az);
        const double delta = v_dot(c, az, z);
        v_lincomb(c, vjm1, {1.0, -delta / gamma_j, -gamma_j / gamma_jm1}, {&az, &vj, &vjm1});
        std::swap(vjm1, vj);
        P.solve_zero(c, A, zjp1, vj);
        const double gamma_jp1 = std::sqrt(v_dot(c, zjp1, vj));
        const double a0 = c_j * delta - c_jm1 * s_j * gamma_j;
        const double a1 = std::sqrt(a0 * a0 + gamma_jp1 * gamma_jp1);
        const double a2 = s_j * delta + c_jm1 * c_j * gamma_j;
        const double a3 = s_jm1 * gamma_j;
        const double c_jp1 = a0 / a1, s_jp1 = gamma_jp1 / a1;
        v_lincomb(c, wjm1, {1.0 / a1, -a3 / a1, -a2 / a1}, {&z, &wjm1, &wj});
        std::swap(wjm1, wj);
        v_lincomb(c, x, {1.0, c_jp1 * eta}, {&x, &wj});
        eta = -s_jp1 * eta;
        const double res = std::fabs(eta);
        info.iterations = it;
        info.final_residual = res;
        info.history.push_back(res);
        if (res / r0 <= p_.rtol || res < p_.atol) return;
        std::swap(z, zjp1);
        gamma_jm1 = gamma_j;
        gamma_j = gamma_jp1;
        c_jm1 = c_j;
        c_j = c_jp1;
        s_jm1 = s_j;
        s_j = s_jp1;
    }
}

// ------------------------------------------------------------------ FGMRES
FGMRES::FGMRES(KrylovParams p, std::vector<BVec> tmps, Solver* prec, bool skip)
    : p_(p), t_(std::move(tmps)), prec_(prec), skip_nonfinite_(skip)
{
    TN_REQUIRE((int)t_.size() >= 2 * p_.restart + 4, "FGMRES needs 2*restart+4 temporaries");
}

void FGMRES::solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b)
{
    Solver& P = prec_ ? *prec_ : (Solver&)id_;
    const int m = p_.restart;
    BVec& r = t_[0];
    auto V = [&](int i) -> BVec& { return t_[1 + i]; };
    auto Z = [&](int i) -> BVec& { return t_[1 + (m + 1) + i]; };
    BVec& w = t_[1 + (m + 1) + m];
    BVec& acc = t_[2 + (m + 1) + m];
    info = IterInfo{};
    A.apply(c, x, r);
    v_lincomb(c, r, {1.0, -1.0}, {&b, &r});
    double beta0 = std::sqrt(v_dot(c, r, r));
    const double r0 = beta0;
    info.initial_residual = info.final_residual = r0;
    info.history.push_back(r0);
    if (beta0 <= p_.atol) return;
    std::vector<double> H((m + 1) * m), g(m + 1), cs(m + 1), sn(m + 1), y(m), hv(m + 2);
    if (c.col.n < (size_t)m + 2) c.col.alloc(m + 2);
    auto h = [&](int i, int j) -> double& { return H[i * m + j]; };
    const double eps = std::numeric_limits<double>::epsilon();
    int total = 0;
    while (total < p_.max_iterations) {
        v_lincomb(c, V(0), {1.0 / beta0}, {&r});
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        std::fill(sn.begin(), sn.end(), 0.0);
        g[0] = beta0;
        int inner = 0;
        for (int j = 0; j < m && total < p_.max_iterations; ++j, ++total) {
            P.solve_zero(c, A, Z(j), V(j));
            // a non-finite preconditioned vector is replaced by V_j, i.e. the
            // preconditioner is skipped for this step (fgmres.hpp:180-193); the
            // test and the copy run on the device, the flag is read with the column
            if (skip_nonfinite_) v_replace_nonfinite(c, Z(j), V(j));
            A.apply(c, Z(j), w);
            // modified Gram-Schmidt: h_ij = w . V_i, w -= h_ij V_i for i = 0..j in
            // turn; each subtraction is fused with the next dot (last: the norm).
            // The coefficients stay on the device (col[i] = h_ij, col[j+1] = |w|^2),
            // so the column costs one host synchronisation, not j + 2.
            double* hc = c.col.p;
            launch_fill<double>(hc, j + 2, 0.0, c.s);
            v_dot_to(c, w, V(0), hc);
            for (int i = 0; i <= j; ++i) v_sub_axpy_dot_to(c, w, hc + i, V(i), i < j ? V(i + 1) : w, hc + i + 1);
            int skipped = 0;
            c.fetch_col(j + 2, hv.data(), skip_nonfinite_ ? &skipped : nullptr);
            if (skipped) ++prec_skipped;
            for (int i = 0; i <= j; ++i) h(i, j) = hv[i];
            const double hn2 = hv[j + 1];
            const double hn = std::sqrt(std::max(hn2, 0.0));
            if (hn < eps * r0) {
                h(j + 1, j) = 0;
                inner = j + 1;
                ++restarts_breakdown;
                break;
            }
            h(j + 1, j) = hn;
            if (hn > 0) v_lincomb(c, V(j + 1), {1.0 / hn}, {&w});
            for (int i = 0; i < j; ++i) {
                const double tmp = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
                h(i + 1, j) = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
                h(i, j) = tmp;
            }
            const double a = h(j, j), bb = h(j + 1, j), rab = std::hypot(a, bb);
            cs[j] = rab == 0 ? 1.0 : a / rab;
            sn[j] = rab == 0 ? 0.0 : bb / rab;
            h(j, j) = rab;
            h(j + 1, j) = 0;
            if (std::fabs(h(j, j)) < eps) {
                inner = j;
                ++restarts_stagnation;
                break;
            }
            const double gj = g[j], gj1 = g[j + 1];
            g[j] = cs[j] * gj + sn[j] * gj1;
            g[j + 1] = -sn[j] * gj + cs[j] * gj1;
            const double ares = std::fabs(g[j + 1]);
            info.history.push_back(ares);
            info.iterations = total + 1;
            info.final_residual = ares;
            inner = j + 1;
            if (ares / r0 <= p_.rtol || ares < p_.atol) break;
        }
        for (int row = inner - 1; row >= 0; --row) {
            double s = g[row];
            for (int col = row + 1; col < inner; ++col) s -= h(row, col) * y[col];
            y[row] = s / h(row, row);
        }
        v_fill(c, acc, 0.0);
        for (int i = 0; i < inner; ++i) v_lincomb(c, acc, {1.0, y[i]}, {&acc, &Z(i)});
        v_lincomb(c, x, {1.0, 1.0}, {&x, &acc});
        A.apply(c, x, r);
        v_lincomb(c, r, {1.0, -1.0}, {&b, &r});
        beta0 = std::sqrt(v_dot(c, r, r));
        info.final_residual = beta0;
        if (beta0 <= p_.atol || beta0 / r0 <= p_.rtol) return;
    }
}

// ------------------------------------------------------------------ Multigrid
Multigrid::Multigrid(std::vector<LinOp*> ops, std::vector<Transfer> tr, std::vector<Solver*> sm, Solver* coarse,
                     std::vector<BVec> tmp, std::vector<BVec> r, std::vector<BVec> e, int cycles)
    : ops_(std::move(ops)), tr_(std::move(tr)), sm_(std::move(sm)), coarse_(coarse), tmp_(std::move(tmp)),
      r_(std::move(r)), e_(std::move(e)), cycles_(cycles)
{
}

namespace {
template <typename T>
void do_restrict(const Transfer& t, const BVec& xf, const BVec& xc, hipStream_t s)
{
    launch_restrict<T>(*t.fine, *t.coarse, xf.b[0].ncomp, (const T*)xf.b[0].p, (T*)xc.b[0].p, s);
}
template <typename T>
void do_prolongate(const Transfer& t, const BVec& xc, const BVec& xf, hipStream_t s)
{
    launch_prolongate_add<T>(*t.coarse, *t.fine, xc.b[0].ncomp, (const T*)xc.b[0].p, (T*)xf.b[0].p, s);
}
template <typename T>
void fine_project(const Transfer& t, const BVec& v, hipStream_t s)
{
    if (t.project_cmb) launch_freeslip_project<T>(*t.fine, (T*)v.b[0].p, 1, s);
    if (t.project_surf) launch_freeslip_project<T>(*t.fine, (T*)v.b[0].p, 2, s);
}
}  // namespace

void Multigrid::solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b)
{
    const int L = (int)tr_.size();
    for (int cyc = 0; cyc < cycles_; ++cyc) recurse(c, A, x, b, L, false);
}

void Multigrid::solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b)
{
    const int L = (int)tr_.size();
    for (int cyc = 0; cyc < cycles_; ++cyc) recurse(c, A, x, b, L, cyc == 0);
}

// zero: x is to be taken as zero on entry (its contents are ignored)
void Multigrid::recurse(Ctx& c, LinOp& A, const BVec& x, const BVec& b, int level, bool zero)
{
    if (level == 0) {
        if (zero)
            coarse_->solve_zero(c, A, x, b);
        else
            coarse_->solve(c, A, x, b);
        return;
    }
    if (zero)
        sm_[level]->solve_zero(c, A, x, b);
    else
        sm_[level]->solve(c, A, x, b);
    A.apply(c, x, tmp_[level]);
    v_lincomb(c, tmp_[level], {1.0, -1.0}, {&b, &tmp_[level]});
    const Transfer& t = tr_[level - 1];
    const bool f64 = x.b[0].dtype == TN_F64;
    if (f64) {
        fine_project<double>(t, tmp_[level], c.s);
        do_restrict<double>(t, tmp_[level], r_[level - 1], c.s);
    } else {
        fine_project<float>(t, tmp_[level], c.s);
        do_restrict<float>(t, tmp_[level], r_[level - 1], c.s);
    }
    // the coarse correction starts from zero
    recurse(c, *ops_[level - 1], e_[level - 1], r_[level - 1], level - 1, true);
    if (f64) {
        do_prolongate<double>(t, e_[level - 1], x, c.s);
        fine_project<double>(t, x, c.s);
    } else {
        do_prolongate<float>(t, e_[level - 1], x, c.s);
        fine_project<float>(t, x, c.s);
    }
    sm_[level]->solve(c, A, x, b);
}

}  // namespace tn
