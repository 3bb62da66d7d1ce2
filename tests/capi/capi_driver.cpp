// C++ caller of the C ABI through include/terraneo_amd.h alone (no HIP or
// torch headers): the Dirichlet Stokes problem at velocity lateral level 4 x 17
// radial nodes, smooth variable viscosity, buoyancy forcing.
//   eps apply + symmetry check, GCA coarse operators, GMG V-cycle,
//   block-triangular FGMRES Stokes solve, true residual recomputed.
// Exit status 0 and a line "capi_driver OK ..." on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "terraneo_amd.h"

#define CHECK(call)                                                                       \
    do {                                                                                  \
        if ((call) != 0) {                                                                \
            std::fprintf(stderr, "%s failed: %s\n", #call, tn_last_error());              \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

namespace {
struct Dev {
    void* p = nullptr;
    explicit Dev(size_t bytes) { CHECK(tn_malloc(&p, bytes)); }
    ~Dev() { tn_free(p); }
};

struct Mesh {
    tn_mesh* m = nullptr;
    int n = 0, nr = 0;
    int64_t N = 0;
    std::vector<double> coords, radii;
    Mesh(int lat, int rad)
    {
        CHECK(tn_mesh_create(lat, rad, 0.55, 1.0, &m));
        CHECK(tn_mesh_dims(m, &n, &nr, &N));
        coords.resize(10 * (size_t)n * n * 3);
        radii.resize(nr);
        CHECK(tn_mesh_lateral_coords(m, coords.data()));
        CHECK(tn_mesh_radii(m, radii.data()));
    }
    ~Mesh() { tn_mesh_destroy(m); }
    double dot(const void* a, const void* b, int ncomp) const
    {
        double v = 0;
        CHECK(tn_dot(m, TN_F64, ncomp, a, b, &v, nullptr));
        return v;
    }
};

// a smooth field of the node position, identical on every copy of a node
double smooth(const double* c, double r, double k) { return std::sin(k * c[0] * r) + std::cos(1.3 * k * c[1]) * r + c[2]; }

std::vector<double> nodal(const Mesh& M, int ncomp, double k)
{
    std::vector<double> v((size_t)ncomp * M.N);
    for (int c = 0; c < ncomp; ++c)
        for (int64_t lat = 0; lat < 10LL * M.n * M.n; ++lat)
            for (int r = 0; r < M.nr; ++r)
                v[c * M.N + lat * M.nr + r] = smooth(&M.coords[3 * lat], M.radii[r], k + c);
    return v;
}
}  // namespace

int main()
{
    const int L = 4;
    Mesh Mf(L, L), Mm(L - 1, L - 1), Mc(L - 2, L - 2);  // velocity levels (fine .. coarse)
    Mesh& Mp = Mm;                                       // pressure lives on the once-coarsened mesh
    const size_t vb = sizeof(double) * 3 * Mf.N;
    // viscosity exp(smooth) on every velocity level (placeholder ones on the Galerkin levels)
    std::vector<double> eh(Mf.N);
    {
        const std::vector<double> s = nodal(Mf, 1, 2.0);
        for (int64_t i = 0; i < Mf.N; ++i) eh[i] = std::exp(0.5 * s[i]);
    }
    Dev eta(sizeof(double) * Mf.N), onesm(sizeof(double) * Mm.N), onesc(sizeof(double) * Mc.N);
    CHECK(tn_memcpy_h2d(eta.p, eh.data(), sizeof(double) * Mf.N));
    CHECK(tn_lincomb(Mm.m, TN_F64, 1, onesm.p, 0, nullptr, nullptr, 1.0, nullptr));
    CHECK(tn_lincomb(Mc.m, TN_F64, 1, onesc.p, 0, nullptr, nullptr, 1.0, nullptr));
    tn_eps *Af = nullptr, *Am = nullptr, *Ac = nullptr;
    CHECK(tn_eps_create(Mf.m, TN_F64, eta.p, TN_DIRICHLET, TN_DIRICHLET, &Af));
    CHECK(tn_eps_create(Mm.m, TN_F64, onesm.p, TN_DIRICHLET, TN_DIRICHLET, &Am));
    CHECK(tn_eps_create(Mc.m, TN_F64, onesc.p, TN_DIRICHLET, TN_DIRICHLET, &Ac));

    // apply: <y, A x> = <x, A y> for consistent x, y
    Dev x(vb), y(vb), Ax(vb), Ay(vb);
    CHECK(tn_randomize(Mf.m, TN_F64, 3, x.p, 1, nullptr));
    CHECK(tn_randomize(Mf.m, TN_F64, 3, y.p, 2, nullptr));
    CHECK(tn_interface_sum(Mf.m, TN_F64, 3, x.p, nullptr));  // copies of a node agree (sums)
    CHECK(tn_interface_sum(Mf.m, TN_F64, 3, y.p, nullptr));
    CHECK(tn_eps_apply(Af, x.p, Ax.p, nullptr));
    CHECK(tn_eps_apply(Af, y.p, Ay.p, nullptr));
    const double s1 = Mf.dot(y.p, Ax.p, 3), s2 = Mf.dot(x.p, Ay.p, 3);
    const double sym = std::fabs(s1 - s2) / std::fabs(s1);
    std::printf("apply symmetry |<y,Ax> - <x,Ay>| / |<y,Ax>| = %.3e\n", sym);
    if (sym > 1e-12) return 2;

    // Galerkin coarse operators and the multigrid
    CHECK(tn_eps_gca_from(Am, Af, TN_F64, nullptr));
    CHECK(tn_eps_gca_from(Ac, Am, TN_F64, nullptr));
    tn_gmg_options go = {3, 2, 10, 1, 1000, 1e-6, 1e-16, 12345, TN_COARSE_DENSE, 0, 0};
    tn_eps* ops[3] = {Ac, Am, Af};
    tn_gmg* G = nullptr;
    CHECK(tn_gmg_create(3, ops, &go, nullptr, &G));

    // Stokes: buoyancy T r_hat with T smooth, zero on both shells
    tn_stokes* K = nullptr;
    CHECK(tn_stokes_create(Mf.m, Mp.m, Af, &K));
    const size_t pb = sizeof(double) * Mp.N;
    Dev bu(vb), bp(pb), u(vb), p(pb), schur(pb), inv_eta_p(pb), Tn(sizeof(double) * Mf.N);
    {
        const std::vector<double> T = nodal(Mf, 1, 3.0);
        CHECK(tn_memcpy_h2d(Tn.p, T.data(), sizeof(double) * Mf.N));
        std::vector<double> iep(Mp.N);
        for (int64_t lat = 0; lat < 10LL * Mp.n * Mp.n; ++lat) {
            const int64_t d = lat / ((int64_t)Mp.n * Mp.n), xx = (lat / Mp.n) % Mp.n, yy = lat % Mp.n;
            const int64_t latf = (d * Mf.n + 2 * xx) * Mf.n + 2 * yy;
            for (int r = 0; r < Mp.nr; ++r) iep[lat * Mp.nr + r] = 1.0 / eh[latf * Mf.nr + 2 * r];
        }
        CHECK(tn_memcpy_h2d(inv_eta_p.p, iep.data(), pb));
    }
    CHECK(tn_stokes_rhs_from_temperature(K, Tn.p, 1e5, bu.p, bp.p, nullptr));
    CHECK(tn_kmass_lumped(Mp.m, TN_F64, inv_eta_p.p, schur.p, nullptr));
    CHECK(tn_lincomb(Mf.m, TN_F64, 3, u.p, 0, nullptr, nullptr, 0.0, nullptr));
    CHECK(tn_lincomb(Mp.m, TN_F64, 1, p.p, 0, nullptr, nullptr, 0.0, nullptr));
    tn_krylov_options ko = {400, 1e-8, 1e-30, 30, 1};
    tn_solve_info info{};
    CHECK(tn_stokes_solve(K, TN_FGMRES_BLOCKTRI, G, schur.p, u.p, p.p, bu.p, bp.p, &ko, &info, nullptr, 0, nullptr));
    // true residual of the mean-free system: b - K x, pressure mean removed
    Dev ru(vb), rp(pb), onep(pb);
    CHECK(tn_stokes_apply(K, u.p, p.p, ru.p, rp.p, nullptr));
    {
        const double c[2] = {1.0, -1.0};
        const void* xu[2] = {bu.p, ru.p};
        const void* xp[2] = {bp.p, rp.p};
        CHECK(tn_lincomb(Mf.m, TN_F64, 3, ru.p, 2, c, xu, 0.0, nullptr));
        CHECK(tn_lincomb(Mp.m, TN_F64, 1, rp.p, 2, c, xp, 0.0, nullptr));
        CHECK(tn_pressure_normalize(Mp.m, TN_F64, rp.p, nullptr));
    }
    const double res = std::sqrt((Mf.dot(ru.p, ru.p, 3) + Mp.dot(rp.p, rp.p, 1)) /
                                 (Mf.dot(bu.p, bu.p, 3) + Mp.dot(bp.p, bp.p, 1)));
    std::printf("FGMRES: %d iterations, reported %.3e, true relative residual %.3e, %.3f s\n", info.iterations,
                info.final_residual / info.initial_residual, res, info.seconds);
    CHECK(tn_sync());
    tn_gmg_destroy(G);
    tn_stokes_destroy(K);
    tn_eps_destroy(Af);
    tn_eps_destroy(Am);
    tn_eps_destroy(Ac);
    if (!(info.final_residual <= 1e-8 * info.initial_residual * (1 + 1e-12)) || !(res <= 2e-8)) return 3;
    std::printf("capi_driver OK iterations=%d residual=%.3e\n", info.iterations, res);
    return 0;
}
