// Host-side linear algebra over device fields: block vectors, level-1
// operations, operator/solver interfaces and the Krylov / multigrid drivers.
//
// Mirrors the reference's concept layer (linalg/operator.hpp:56-159,
// linalg/vector.hpp:8-170, linalg/solvers/solver.hpp:19-54) with runtime
// polymorphism: the solvers are host code, every vector operation is a HIP
// kernel on the context's stream, and every scalar that a recurrence needs
// on the host (dots, norms) is one device reduction plus one 8-byte copy.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "tn_internal.hpp"

namespace tn {

struct Field {
    void* p = nullptr;
    int dtype = TN_F64;
    int ncomp = 3;
    MeshLevel* m = nullptr;
    int64_t count() const { return m->N * (int64_t)ncomp; }
    size_t esize() const { return dtype == TN_F64 ? 8 : 4; }
};

// One block (a velocity or scalar field) or two (Stokes velocity + pressure).
struct BVec {
    Field b[2];
    int nb = 1;
    Field& u() { return b[0]; }
    Field& p() { return b[1]; }
    const Field& u() const { return b[0]; }
    const Field& p() const { return b[1]; }
};

// Device memory owner for vectors shaped like a template.
class VecStore {
  public:
    BVec make(const BVec& like);
    BVec make_like(const BVec& like, int dtype);
    ~VecStore();

  private:
    std::vector<void*> ptrs_;
};

struct Ctx {
    hipStream_t s = nullptr;
    DevBuf<double> red;   // reduction slots
    DevBuf<double> part;  // per-block partial sums of the reductions
    DevBuf<int> flag;
    DevBuf<double> col;   // reduction results kept on the device (a Hessenberg column; grown on demand)
    Ctx();
    double fetch(int slots, double* out = nullptr);
    // col[0..n) to out, and the flag to *flag_out when given: one host synchronisation
    void fetch_col(int n, double* out, int* flag_out = nullptr);
};

// level-1 (all blocks)
void v_fill(Ctx& c, const BVec& y, double v);
void v_lincomb(Ctx& c, const BVec& y, std::initializer_list<double> coef, std::initializer_list<const BVec*> xs,
               double c0 = 0.0);
void v_assign(Ctx& c, const BVec& y, const BVec& x);
double v_dot(Ctx& c, const BVec& a, const BVec& b);
// w += alpha x, then return w . y (owned nodes; y may be w itself)
double v_axpy_dot(Ctx& c, const BVec& w, double alpha, const BVec& x, const BVec& y);
// Device-resident forms (no host synchronisation): *out_dev += a . b, and
// w -= coef_dev[0] x followed by *out_dev += w . y
void v_dot_to(Ctx& c, const BVec& a, const BVec& b, double* out_dev);
void v_sub_axpy_dot_to(Ctx& c, const BVec& w, const double* coef_dev, const BVec& x, const BVec& y, double* out_dev);
// y = x if y holds a non-finite entry; the test and the copy stay on the device
// (c.flag holds the outcome until the next fetch_col)
void v_replace_nonfinite(Ctx& c, const BVec& y, const BVec& x);
void v_mul(Ctx& c, const BVec& y, const BVec& d);
void v_invert(Ctx& c, const BVec& y);
void v_randomize(Ctx& c, const BVec& y, uint64_t seed);
void v_convert(Ctx& c, const BVec& dst, const BVec& src);
// out[k] = a . bs[k] (k < n <= 16), one pass over a
void v_multidot(Ctx& c, const BVec& a, const std::vector<const BVec*>& bs, double* out);

struct LinOp {
    virtual void apply(Ctx& c, const BVec& x, const BVec& y) = 0;
    virtual ~LinOp() = default;
};

struct Solver {
    // x is the initial guess on entry (reference semantics: solvers start from x)
    virtual void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) = 0;
    // the same from the zero initial guess; x's contents on entry are ignored.
    // Default: zero x, then solve.  Solvers that can skip work for a zero
    // guess (the product A 0, the fill itself) override it with a result
    // identical to the default's.
    virtual void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b);
    virtual ~Solver() = default;
};

struct IterInfo {
    int iterations = 0;
    double initial_residual = 0.0;
    double final_residual = 0.0;
    std::vector<double> history;  // absolute residual per iteration (incl. 0)
};

struct KrylovParams {
    int max_iterations = 100;
    double rtol = 1e-8;
    double atol = 1e-30;
    int restart = 30;
};

// ---- solvers (reference files cited in solvers.cpp)
class IdentitySolver : public Solver {
  public:
    void solve(Ctx& c, LinOp&, const BVec& x, const BVec& b) override { v_assign(c, x, b); }
    // x is overwritten either way
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override { solve(c, A, x, b); }
};

class DiagonalSolver : public Solver {  // x = b * inv_diag
  public:
    explicit DiagonalSolver(BVec inv_diag) : d_(inv_diag) {}
    void solve(Ctx& c, LinOp&, const BVec& x, const BVec& b) override;
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override { solve(c, A, x, b); }

  private:
    BVec d_;
};

class Chebyshev : public Solver {
  public:
    Chebyshev(int order, int iterations, BVec inv_diag, BVec t0, BVec t1, int power_iterations, uint64_t seed);
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    // first step without A x (= 0) and with x written, not updated
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    double lambda_max() const { return lambda_; }
    void set_lambda_max(double l)
    {
        lambda_ = l;
        need_lambda_ = false;
    }
    // z <- D^-1 z; default: entrywise with the inverse diagonal (the reference's)
    std::function<void(Ctx&, const BVec&)> scale;

  private:
    void run(Ctx& c, LinOp& A, const BVec& x, const BVec& b, bool zero);
    void apply_scale(Ctx& c, const BVec& z);
    int order_, iters_, power_its_;
    uint64_t seed_;
    BVec dinv_, d_, z_;
    bool need_lambda_ = true;
    double lambda_ = 0.0;
};

class PCG : public Solver {
  public:
    PCG(KrylovParams p, std::vector<BVec> tmps, Solver* prec);
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    IterInfo info;

  private:
    KrylovParams p_;
    std::vector<BVec> t_;
    Solver* prec_;
    IdentitySolver id_;
};

class PMINRES : public Solver {
  public:
    PMINRES(KrylovParams p, std::vector<BVec> tmps, Solver* prec);
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    IterInfo info;

  private:
    KrylovParams p_;
    std::vector<BVec> t_;
    Solver* prec_;
    IdentitySolver id_;
};

class FGMRES : public Solver {
  public:
    FGMRES(KrylovParams p, std::vector<BVec> tmps, Solver* prec, bool skip_nonfinite_prec = true);
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    IterInfo info;
    int restarts_breakdown = 0, restarts_stagnation = 0, prec_skipped = 0;

  private:
    KrylovParams p_;
    std::vector<BVec> t_;
    Solver* prec_;
    IdentitySolver id_;
    bool skip_nonfinite_;
};

// One multigrid level pair: transfers between level l (coarse) and l+1.
struct Transfer {
    MeshLevel* coarse;
    MeshLevel* fine;
    // free-slip projected transfer on the fine level's shells (1 CMB, 2 surface)
    bool project_cmb = false, project_surf = false;
};

class Multigrid : public Solver {
  public:
    // ops[l], smoothers[l] for l = 0..L (0 coarsest, L finest operator = the
    // operator passed to solve); coarse solver acts on level 0.
    Multigrid(std::vector<LinOp*> coarse_ops, std::vector<Transfer> transfers, std::vector<Solver*> smoothers,
              Solver* coarse, std::vector<BVec> tmp, std::vector<BVec> r, std::vector<BVec> e, int cycles);
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    // the first cycle's pre-smoothing starts from zero on every level, so the
    // corrections need no fill and the smoothers skip their first product
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;

  private:
    void recurse(Ctx& c, LinOp& A, const BVec& x, const BVec& b, int level, bool zero);
    std::vector<LinOp*> ops_;
    std::vector<Transfer> tr_;
    std::vector<Solver*> sm_;
    Solver* coarse_;
    std::vector<BVec> tmp_, r_, e_;
    int cycles_;
};

// 2x2 block preconditioners on Stokes vectors (block 0 velocity, 1 pressure)
class BlockDiagonal : public Solver {
  public:
    BlockDiagonal(LinOp* A11, Solver* p11, LinOp* A22, Solver* p22) : A11_(A11), p11_(p11), A22_(A22), p22_(p22) {}
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;

  private:
    void run(Ctx& c, const BVec& x, const BVec& b, bool zero);

  private:
    LinOp *A11_;
    Solver* p11_;
    LinOp* A22_;
    Solver* p22_;
};

class BlockTriangular : public Solver {
  public:
    BlockTriangular(LinOp* A11, Solver* p11, LinOp* A22, Solver* p22, LinOp* B12, BVec tmp)
        : A11_(A11), p11_(p11), A22_(A22), p22_(p22), B12_(B12), tmp_(tmp)
    {
    }
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;

  private:
    void run(Ctx& c, const BVec& x, const BVec& b, bool zero);
    LinOp* A11_;
    Solver* p11_;
    LinOp* A22_;
    Solver* p22_;
    LinOp* B12_;
    BVec tmp_;
};

// Runs an inner solver in another precision: x, b are converted to the inner
// dtype, the inner solver starts from the converted x, the result is
// converted back.
class PrecisionBridge : public Solver {
  public:
    PrecisionBridge(Solver* inner, LinOp* inner_op, BVec xi, BVec bi) : inner_(inner), op_(inner_op), xi_(xi), bi_(bi) {}
    void solve(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;
    // no conversion of x: the inner solver starts from zero
    void solve_zero(Ctx& c, LinOp& A, const BVec& x, const BVec& b) override;

  private:
    Solver* inner_;
    LinOp* op_;
    BVec xi_, bi_;
};

}  // namespace tn
