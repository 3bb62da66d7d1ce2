// C ABI: geometric multigrid for the viscous block and the Krylov drivers
// (PCG for the viscous operator; MINRES / FGMRES for the Stokes system).
#include <chrono>
#include <memory>

#include "tn_coarse.hpp"
#include "tn_eps_march.hpp"
#include "tn_fv.hpp"
#include "tn_mg.hpp"
#include "tn_objects.hpp"
#include "tn_solvers.hpp"
#include "tn_stokes.hpp"

namespace tn {
void set_last_error(const std::string& msg);
}  // namespace tn

#define TN_API_BEGIN try {
#define TN_API_END                    \
    }                                 \
    catch (const std::exception& e)   \
    {                                 \
        tn::set_last_error(e.what()); \
        return -1;                    \
    }                                 \
    return 0;

using namespace tn;

namespace {

Field field_of(const tn_mesh* m, int dtype, int ncomp, const void* p)
{
    Field f;
    f.p = const_cast<void*>(p);
    f.dtype = dtype;
    f.ncomp = ncomp;
    f.m = &const_cast<tn_mesh*>(m)->lvl.dev();
    return f;
}

BVec vec1(const Field& f)
{
    BVec v;
    v.b[0] = f;
    v.nb = 1;
    return v;
}

BVec vec2(const Field& u, const Field& p)
{
    BVec v;
    v.b[0] = u;
    v.b[1] = p;
    v.nb = 2;
    return v;
}

// Removes the owned-node mean of a scalar (pressure) field in place:
// p -= (sum over owned nodes of p) / (number of owned nodes), the orthogonal
// projection onto the complement of the constants in the owned inner product.
// Two launches, no host synchronisation.
struct MeanFree {
    MeshLevel* m;
    int dtype;
    VecStore st;
    BVec ones;
    DevBuf<double> sum, part;
    double inv_count = 0.0;
    MeanFree(Ctx& c, const Field& like) : m(like.m), dtype(like.dtype), sum(1), part(REDUCE_PART_DOUBLES)
    {
        ones = st.make(vec1(like));
        v_fill(c, ones, 1.0);
        inv_count = 1.0 / v_dot(c, ones, ones);
    }
    void project(Ctx& c, void* p)
    {
        launch_fill<double>(sum.p, 1, 0.0, c.s);
        if (dtype == TN_F64) {
            launch_dot_owned<double>(*m, 1, (const double*)p, (const double*)ones.b[0].p, sum.p, part.p, c.s);
            launch_sub_scaled<double>((double*)p, m->N, sum.p, inv_count, c.s);
        } else {
            launch_dot_owned<float>(*m, 1, (const float*)p, (const float*)ones.b[0].p, sum.p, part.p, c.s);
            launch_sub_scaled<float>((float*)p, m->N, sum.p, inv_count, c.s);
        }
    }
};

// Pi K Pi with Pi the mean removal of the pressure block (the Stokes system
// bordered by the mean-pressure constraint).  The input's pressure is
// projected into a private copy, so Krylov basis vectors stay untouched.
struct MeanFreeStokes : LinOp {
    LinOp* K;
    MeanFree* mf;
    BVec ptmp;  // one scalar block
    void apply(Ctx& c, const BVec& x, const BVec& y) override
    {
        v_assign(c, ptmp, vec1(x.b[1]));
        mf->project(c, ptmp.b[0].p);
        K->apply(c, vec2(x.b[0], ptmp.b[0]), y);
        mf->project(c, y.b[1].p);
    }
};

struct EpsOp : LinOp {
    tn_eps* op;
    explicit EpsOp(tn_eps* o) : op(o) {}
    void apply(Ctx& c, const BVec& x, const BVec& y) override { eps_apply(op, x.b[0].p, y.b[0].p, c.s); }
};

struct StokesOp : LinOp {
    tn_stokes* K;
    explicit StokesOp(tn_stokes* k) : K(k) {}
    void apply(Ctx& c, const BVec& x, const BVec& y) override
    {
        stokes_apply(K, x.b[0].p, x.b[1].p, y.b[0].p, y.b[1].p, c.s);
    }
};

struct NullOp : LinOp {
    void apply(Ctx& c, const BVec&, const BVec& y) override { v_fill(c, y, 0.0); }
};

}  // namespace

// Geometric multigrid on the epsilon operator (viscous block): one tn_eps
// per level (coarse .. fine), inverse diagonals, Chebyshev smoothers,
// constant transfers, PCG on the coarsest level (reference app/test
// construction: mantlecirculation.cpp:556-645, test_epsilon_divdiv_stokes_assess.cpp:554-660).
struct tn_gmg {
    int dtype = TN_F64;
    std::vector<tn_eps*> ops;
    std::vector<std::unique_ptr<EpsOp>> lops;
    VecStore store;
    std::vector<BVec> invd, ta, tb, tmp, r, e, ctmp;
    std::vector<std::unique_ptr<Chebyshev>> sm;
    std::unique_ptr<PCG> coarse;
    std::unique_ptr<PrecisionBridge> coarse_bridge;  // coarsest level in another precision
    std::unique_ptr<DenseCoarse> dense;              // TN_COARSE_DENSE
    std::unique_ptr<Multigrid> mg;
    Ctx ctx;
};

namespace {
// A tn_gmg as a preconditioner: one multigrid solve on the finest operator.
struct GmgPrec : Solver {
    tn_gmg* g = nullptr;
    void solve(Ctx& c, LinOp&, const BVec& x, const BVec& b) override
    {
        g->ctx.s = c.s;
        g->mg->solve(g->ctx, *g->lops.back(), x, b);
    }
    void solve_zero(Ctx& c, LinOp&, const BVec& x, const BVec& b) override
    {
        g->ctx.s = c.s;
        g->mg->solve_zero(g->ctx, *g->lops.back(), x, b);
    }
};
}  // namespace

namespace {

void gmg_setup(tn_gmg* g, const tn_gmg_options& o, hipStream_t s)
{
    g->ctx.s = s;
    Ctx& c = g->ctx;
    const int L = (int)g->ops.size();
    for (int l = 0; l < L; ++l) {
        tn_eps* op = g->ops[l];
        g->lops.emplace_back(new EpsOp(op));
        const BVec like = vec1(field_of(op->mesh, op->dtype, 3, nullptr));
        BVec ones = g->store.make(like);
        g->invd.push_back(g->store.make(like));
        v_fill(c, ones, 1.0);
        const bool dg = op->diagonal;
        const int am = op->apply_mode, cm = op->comm_mode;
        op->diagonal = true;
        op->apply_mode = TN_REPLACE;
        op->comm_mode = TN_COMM_ADDITIVE;
        eps_apply(op, ones.b[0].p, g->invd.back().b[0].p, s);
        op->diagonal = dg;
        op->apply_mode = am;
        op->comm_mode = cm;
        v_invert(c, g->invd.back());
        g->ta.push_back(g->store.make(like));
        g->tb.push_back(g->store.make(like));
        g->tmp.push_back(g->store.make(like));
        if (l < L - 1) {
            // level vectors between transfers carry the hierarchy's precision
            const BVec hl = vec1(field_of(op->mesh, g->dtype, 3, nullptr));
            g->r.push_back(g->store.make(hl));
            g->e.push_back(g->store.make(hl));
        }
        g->sm.emplace_back(new Chebyshev(o.cheby_order, o.cheby_iterations, g->invd.back(), g->ta.back(), g->tb.back(),
                                         o.power_iterations, o.seed + 1000003ull * l));
        if (o.freeslip_block_jacobi && (op->bc_cmb == TN_FREESLIP || op->bc_surf == TN_FREESLIP)) {
            // the diagonal apply above computed the operator's boundary forms (op->blk)
            TN_REQUIRE(op->blk_ready, "free-slip block Jacobi: boundary forms missing");
            MeshLevel* ml = &const_cast<tn_mesh*>(op->mesh)->lvl.dev();
            const BVec dinv = g->invd.back();
            g->sm.back()->scale = [ml, op, dinv](Ctx& cc, const BVec& z) {
                if (z.b[0].dtype == TN_F64)
                    launch_fs_block_scale<double>(*ml, (double*)z.b[0].p, (const double*)dinv.b[0].p,
                                                  (const double*)op->blk.p, op->bc_cmb, op->bc_surf, cc.s);
                else
                    launch_fs_block_scale<float>(*ml, (float*)z.b[0].p, (const float*)dinv.b[0].p,
                                                 (const float*)op->blk.p, op->bc_cmb, op->bc_surf, cc.s);
            };
        }
    }
    const BVec like0 = vec1(field_of(g->ops[0]->mesh, g->ops[0]->dtype, 3, nullptr));
    for (int i = 0; i < 4; ++i) g->ctmp.push_back(g->store.make(like0));
    KrylovParams kp;
    kp.max_iterations = o.coarse_max_iterations;
    kp.rtol = o.coarse_rtol;
    kp.atol = o.coarse_atol;
    g->coarse.reset(new PCG(kp, g->ctmp, nullptr));
    Solver* coarse = g->coarse.get();
    if (o.coarse_solver == TN_COARSE_DENSE) {
        // gathers / scatters convert any precision: no bridge needed
        g->dense.reset(new DenseCoarse(g->ops[0], s));
        coarse = g->dense.get();
    } else if (g->ops[0]->dtype != g->dtype) {
        // e.g. FP32 V-cycle with an FP64 coarsest-grid solve
        g->coarse_bridge.reset(new PrecisionBridge(g->coarse.get(), g->lops[0].get(), g->store.make(like0),
                                                   g->store.make(like0)));
        coarse = g->coarse_bridge.get();
    }
    std::vector<LinOp*> lo;
    std::vector<Transfer> tr;
    std::vector<Solver*> sms;
    for (int l = 0; l < L; ++l) {
        lo.push_back(g->lops[l].get());
        sms.push_back(g->sm[l].get());
        if (l < L - 1) {
            Transfer t{&const_cast<tn_mesh*>(g->ops[l]->mesh)->lvl, &const_cast<tn_mesh*>(g->ops[l + 1]->mesh)->lvl};
            if (o.freeslip_projected_transfer) {
                t.project_cmb = g->ops[l + 1]->bc_cmb == TN_FREESLIP;
                t.project_surf = g->ops[l + 1]->bc_surf == TN_FREESLIP;
            }
            tr.push_back(t);
        }
    }
    g->mg.reset(new Multigrid(lo, tr, sms, coarse, g->tmp, g->r, g->e, o.num_cycles));
}

void fill_info(tn_solve_info* info, const IterInfo& it, double secs, double* hist, int cap)
{
    if (!info) return;
    info->iterations = it.iterations;
    info->initial_residual = it.initial_residual;
    info->final_residual = it.final_residual;
    info->seconds = secs;
    info->history_len = (int)it.history.size();
    if (hist)
        for (int i = 0; i < cap && i < (int)it.history.size(); ++i) hist[i] = it.history[i];
}

}  // namespace

extern "C" {

int tn_gmg_create(int nlevels, tn_eps* const* ops, const tn_gmg_options* opts, tn_stream_t s, tn_gmg** out)
{
    TN_API_BEGIN
    TN_REQUIRE(ops && opts && out && nlevels >= 2, "need at least 2 levels");
    TN_REQUIRE(opts->coarse_solver == TN_COARSE_PCG || opts->coarse_solver == TN_COARSE_DENSE,
               "invalid coarse solver");
    auto g = std::make_unique<tn_gmg>();
    TN_REQUIRE(ops[nlevels - 1], "null operator");
    g->dtype = ops[nlevels - 1]->dtype;
    for (int l = 0; l < nlevels; ++l) {
        // the coarsest level alone may use another precision (its solve is bridged)
        TN_REQUIRE(ops[l] && (l == 0 || ops[l]->dtype == g->dtype), "levels above the coarsest must share one dtype");
        if (l > 0) {
            const auto& mc = ops[l - 1]->mesh->lvl;
            const auto& mf = ops[l]->mesh->lvl;
            TN_REQUIRE(2 * (mc.n - 1) == mf.n - 1 && 2 * (mc.nr - 1) == mf.nr - 1,
                       "levels must be uniform refinements (coarse first)");
        }
        g->ops.push_back(ops[l]);
    }
    gmg_setup(g.get(), *opts, (hipStream_t)s);
    *out = g.release();
    TN_API_END
}

int tn_gmg_destroy(tn_gmg* g)
{
    TN_API_BEGIN
    delete g;
    TN_API_END
}

// x <- V-cycle(s) applied to A x = b starting from the content of x
int tn_gmg_solve(tn_gmg* g, void* x, const void* b, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(g && x && b, "null pointer");
    g->ctx.s = (hipStream_t)s;
    tn_eps* A = g->ops.back();
    const BVec X = vec1(field_of(A->mesh, A->dtype, 3, x)), B = vec1(field_of(A->mesh, A->dtype, 3, b));
    g->mg->solve(g->ctx, *g->lops.back(), X, B);
    TN_API_END
}

int tn_gmg_get_lambda(const tn_gmg* g, int level, double* lambda)
{
    TN_API_BEGIN
    TN_REQUIRE(g && lambda && level >= 0 && level < (int)g->sm.size(), "bad level");
    *lambda = g->sm[level]->lambda_max();
    TN_API_END
}

int tn_gmg_set_lambda(tn_gmg* g, int level, double lambda)
{
    TN_API_BEGIN
    TN_REQUIRE(g && level >= 0 && level < (int)g->sm.size(), "bad level");
    g->sm[level]->set_lambda_max(lambda);
    TN_API_END
}

int tn_pcg_solve(tn_eps* A, tn_gmg* prec, void* x, const void* b, const tn_krylov_options* o, tn_solve_info* info,
                 double* history, int history_cap, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(A && x && b && o, "null pointer");
    Ctx c;
    c.s = (hipStream_t)s;
    VecStore st;
    const BVec X = vec1(field_of(A->mesh, A->dtype, 3, x)), B = vec1(field_of(A->mesh, A->dtype, 3, b));
    std::vector<BVec> t;
    for (int i = 0; i < 4; ++i) t.push_back(st.make(X));
    KrylovParams kp{o->max_iterations, o->rtol, o->atol, o->restart};
    EpsOp op(A);
    GmgPrec gp;
    gp.g = prec;
    PCG pcg(kp, t, prec ? &gp : nullptr);
    TN_HIP(hipStreamSynchronize(c.s));
    const auto t0 = std::chrono::steady_clock::now();
    pcg.solve(c, op, X, B);
    TN_HIP(hipStreamSynchronize(c.s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    tn::check_march_status();
    fill_info(info, pcg.info, secs, history, history_cap);
    TN_API_END
}

// Stokes solve with a velocity-block multigrid (any dtype; a precision bridge
// is inserted when it differs from the Stokes operator) and a diagonal Schur
// complement preconditioner (inverse of schur_diag, e.g. lumped M_p(1/eta)).
int tn_stokes_solve(tn_stokes* K, int method, tn_gmg* prec_u, const void* schur_diag, void* u, void* p,
                    const void* bu, const void* bp, const tn_krylov_options* o, tn_solve_info* info, double* history,
                    int history_cap, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(K && prec_u && schur_diag && u && p && bu && bp && o, "null pointer");
    Ctx c;
    c.s = (hipStream_t)s;
    VecStore st;
    const int dt = K->A->dtype;
    const Field fu = field_of(K->vel, dt, 3, u), fp = field_of(K->pre, dt, 1, p);
    const BVec X = vec2(fu, fp);
    const BVec B = vec2(field_of(K->vel, dt, 3, bu), field_of(K->pre, dt, 1, bp));
    // Schur: DiagonalSolver inverts its diagonal (diagonal_solver.hpp:34-40)
    BVec sinv = st.make(vec1(fp));
    {
        BVec sd = vec1(field_of(K->pre, dt, 1, schur_diag));
        v_assign(c, sinv, sd);
        v_invert(c, sinv);
    }
    DiagonalSolver schur(sinv);
    StokesOp Kop(K);
    NullOp null;
    // velocity block preconditioner, possibly in another precision
    GmgPrec gp;
    gp.g = prec_u;
    std::unique_ptr<PrecisionBridge> bridge;
    Solver* pu = &gp;
    if (prec_u->dtype != dt) {
        const BVec ui = st.make_like(vec1(fu), prec_u->dtype), bi = st.make_like(vec1(fu), prec_u->dtype);
        bridge.reset(new PrecisionBridge(&gp, prec_u->lops.back().get(), ui, bi));
        pu = bridge.get();
    }
    EpsOp A11(K->A);
    // optional mean-free pressure: project b and the initial guess, solve with Pi K Pi
    std::unique_ptr<MeanFree> mf;
    MeanFreeStokes Kmf;
    LinOp* Kuse = &Kop;
    BVec Buse = B;
    if (o->pressure_mean_free) {
        mf.reset(new MeanFree(c, fp));
        Kmf.K = &Kop;
        Kmf.mf = mf.get();
        Kmf.ptmp = st.make(vec1(fp));
        Kuse = &Kmf;
        Buse = st.make(X);
        v_assign(c, Buse, B);
        mf->project(c, Buse.b[1].p);
        mf->project(c, X.b[1].p);
    }
    KrylovParams kp{o->max_iterations, o->rtol, o->atol, o->restart};
    TN_REQUIRE(method == TN_MINRES_BLOCKDIAG || method == TN_FGMRES_BLOCKTRI || method == TN_FGMRES_BLOCKDIAG,
               "unknown Stokes solver method");
    // Krylov workspace, allocated (and zeroed) before the timed solve
    std::vector<BVec> t;
    const int nwork = method == TN_MINRES_BLOCKDIAG ? 7 : 2 * kp.restart + 4;
    for (int i = 0; i < nwork; ++i) t.push_back(st.make(X));
    const BVec ttmp = method == TN_FGMRES_BLOCKTRI ? st.make(X) : BVec{};
    IterInfo it;
    int restarts = 0, skipped = 0;
    TN_HIP(hipStreamSynchronize(c.s));
    const auto t0 = std::chrono::steady_clock::now();
    if (method == TN_MINRES_BLOCKDIAG) {
        BlockDiagonal P(&A11, pu, &null, &schur);
        PMINRES solver(kp, t, &P);
        solver.solve(c, *Kuse, X, Buse);
        it = solver.info;
    } else {
        std::unique_ptr<Solver> P;
        struct BtOp : LinOp {  // block_12: B^T p, replace + additive communication
            tn_stokes* K;
            void apply(Ctx& c, const BVec& x, const BVec& y) override
            {
                if (K->A->dtype == TN_F64)
                    grad_t<double>(K, x.b[0].p, y.b[0].p, false, true, c.s);
                else
                    grad_t<float>(K, x.b[0].p, y.b[0].p, false, true, c.s);
            }
        } bt;
        bt.K = K;
        if (method == TN_FGMRES_BLOCKTRI)
            P.reset(new BlockTriangular(&A11, pu, &null, &schur, &bt, ttmp));
        else
            P.reset(new BlockDiagonal(&A11, pu, &null, &schur));
        FGMRES solver(kp, t, P.get(), true);
        solver.solve(c, *Kuse, X, Buse);
        it = solver.info;
        restarts = solver.restarts_breakdown + solver.restarts_stagnation;
        skipped = solver.prec_skipped;
    }
    if (mf) mf->project(c, X.b[1].p);
    TN_HIP(hipStreamSynchronize(c.s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    tn::check_march_status();
    fill_info(info, it, secs, history, history_cap);
    if (info) {
        info->restarts = restarts;
        info->prec_skipped = skipped;
    }
    TN_API_END
}

// L2 projection of an FV scalar onto the nodal (Q1) space: M x = b with
// b_i = sum_K T_K int_K phi_i, unpreconditioned CG on the consistent mass
// matrix from x's content (conversion.hpp:131-211; the reference's defaults
// are 1000 iterations, rtol = atol = 1e-12)
int tn_fv_to_fe(tn_fv* fv, const double* Tfv, double* Tfe, const tn_krylov_options* o, tn_solve_info* info,
                double* history, int history_cap, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && Tfv && Tfe && o, "null pointer");
    Ctx c;
    c.s = (hipStream_t)s;
    VecStore st;
    const BVec X = vec1(field_of(fv->mesh, TN_F64, 1, Tfe));
    const BVec B = st.make(X);
    launch_fv_to_fe_rhs(fv->args(), Tfv, (double*)B.b[0].p, c.s);
    MeshLevel& L = fv->mesh->lvl.dev();
    launch_interface_sum<double>(L, 1, (double*)B.b[0].p, c.s);
    struct MassOp : LinOp {
        MeshLevel* L;
        void apply(Ctx& c, const BVec& x, const BVec& y) override
        {
            double* yy = (double*)y.b[0].p;
            launch_fill<double>(yy, L->N, 0.0, c.s);
            MassArgs<double> a{L->d_coords.p, L->d_radii.p, L->n, L->nr, 1, (const double*)x.b[0].p, nullptr, yy, L->d_geo.p};
            launch_mass<double>(a, false, c.s);
            launch_interface_sum<double>(*L, 1, yy, c.s);
        }
    } op;
    op.L = &L;
    std::vector<BVec> t;
    for (int i = 0; i < 4; ++i) t.push_back(st.make(X));
    KrylovParams kp{o->max_iterations, o->rtol, o->atol, o->restart};
    PCG pcg(kp, t, nullptr);
    TN_HIP(hipStreamSynchronize(c.s));
    const auto t0 = std::chrono::steady_clock::now();
    pcg.solve(c, op, X, B);
    TN_HIP(hipStreamSynchronize(c.s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_info(info, pcg.info, secs, history, history_cap);
    TN_API_END
}

}  // extern "C"
