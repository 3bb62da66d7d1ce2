// C ABI: the hex finite-volume energy step (FCT advection-diffusion) on the
// velocity mesh -- cell centres, ghost layers, stable time step, the three
// FCT stages and the explicit step that chains them.
//
// Reference: fv/hex/operators/fct_advection_diffusion.hpp:17-948,
// fv/hex/helpers.hpp:17-211, communication/shell/fv_communication.hpp.
#include <cstring>
#include <map>
#include <vector>

#include "tn_fv.hpp"
#include "tn_objects.hpp"

namespace tn {
void set_last_error(const std::string& msg);
}

#define TN_API_BEGIN try {
#define TN_API_END                    \
    }                                 \
    catch (const std::exception& e)   \
    {                                 \
        tn::set_last_error(e.what()); \
        return -1;                    \
    }                                 \
    return 0;

using tn::MeshLevel;

tn::FvArgs tn_fv::args() const
{
    const MeshLevel& L = mesh->lvl;
    return tn::FvArgs{L.d_coords.p, L.d_radii.p, centers.p, columns.p, n, nr};
}

namespace {

// Lateral ghost map from the mesh's interface groups.  The face between
// ghost cell and real cell of diamond d is the FE edge of two boundary nodes
// (a, b); exactly one other diamond holds copies of both as an edge of its
// own boundary, and its real cell on that edge is the ghost's source
// (fv_communication.hpp exchanges the same faces by neighbourhood tables).
std::vector<int32_t> ghost_pairs(const MeshLevel& L)
{
    const int n = L.n;
    const int ng = L.ngroups();
    std::vector<int> gid(10 * n * n, -1);
    for (int g = 0; g < ng; ++g)
        for (int k = L.grp_off[g]; k < L.grp_off[g + 1]; ++k) gid[L.grp_mem[k]] = g;
    auto fvlat = [n](int d, int x, int y) { return (int32_t)(((int64_t)d * (n + 1) + x) * (n + 1) + y); };
    std::vector<int32_t> out;
    for (int d = 0; d < 10; ++d)
        for (int side = 0; side < 4; ++side)
            for (int j = 1; j <= n - 1; ++j) {
                int ax, ay, bx, by, gx, gy;
                if (side == 0) {
                    ax = bx = 0; ay = j - 1; by = j; gx = 0; gy = j;
                } else if (side == 1) {
                    ax = bx = n - 1; ay = j - 1; by = j; gx = n; gy = j;
                } else if (side == 2) {
                    ay = by = 0; ax = j - 1; bx = j; gx = j; gy = 0;
                } else {
                    ay = by = n - 1; ax = j - 1; bx = j; gx = j; gy = n;
                }
                const int ga = gid[(d * n + ax) * n + ay], gb = gid[(d * n + bx) * n + by];
                TN_REQUIRE(ga >= 0 && gb >= 0, "FV ghost map: diamond edge node without copies");
                int found = 0;
                for (int ka = L.grp_off[ga]; ka < L.grp_off[ga + 1]; ++ka) {
                    const int ma = L.grp_mem[ka];
                    const int da = ma / (n * n), xa = (ma / n) % n, ya = ma % n;
                    if (da == d) continue;
                    for (int kb = L.grp_off[gb]; kb < L.grp_off[gb + 1]; ++kb) {
                        const int mb = L.grp_mem[kb];
                        const int db = mb / (n * n), xb = (mb / n) % n, yb = mb % n;
                        if (db != da || std::abs(xa - xb) + std::abs(ya - yb) != 1) continue;
                        int cx, cy;
                        if (xa == xb) {
                            if (xa != 0 && xa != n - 1) continue;
                            cx = xa == 0 ? 0 : n - 2;
                            cy = std::min(ya, yb);
                        } else {
                            if (ya != 0 && ya != n - 1) continue;
                            cy = ya == 0 ? 0 : n - 2;
                            cx = std::min(xa, xb);
                        }
                        out.push_back(fvlat(d, gx, gy));
                        out.push_back(fvlat(da, cx + 1, cy + 1));
                        ++found;
                    }
                }
                TN_REQUIRE(found == 1, "FV ghost map: a diamond face without exactly one neighbour cell");
            }
    return out;
}

void ghost_update(tn_fv* fv, double* f, int ncomp, hipStream_t s)
{
    tn::launch_fv_ghost(fv->pairs.p, fv->npairs, fv->nr, ncomp, f, s);
}

}  // namespace

extern "C" {

int tn_fv_create(tn_mesh* m, tn_fv** out)
{
    TN_API_BEGIN
    TN_REQUIRE(m && out, "null pointer");
    MeshLevel& L = m->lvl.dev();
    TN_REQUIRE(L.n >= 3 && L.nr >= 3, "FV energy step: at least 2 cells per direction");
    auto* fv = new tn_fv;
    try {
        fv->mesh = m;
        fv->n = L.n;
        fv->nr = L.nr;
        fv->size = tn::fv_size(L.n, L.nr);
        fv->centers.alloc(3 * fv->size);
        fv->columns.alloc((size_t)tn::FV_COLUMN * 10 * (L.n - 1) * (L.n - 1));
        fv->TL.alloc(fv->size);
        fv->Rp.alloc(fv->size);
        fv->Rm.alloc(fv->size);
        fv->ad.alloc(6 * fv->size);
        const std::vector<int32_t> pairs = ghost_pairs(L);
        fv->npairs = (int)(pairs.size() / 2);
        fv->pairs.upload(pairs);
        fv->part.alloc(tn::fv_dt_blocks(L.n, L.nr));
        fv->result.alloc(1);
        // zero: the radial ghost cells at the shells are never written by the
        // ghost exchange (cell centres and R+/R- keep zero there, as in the reference)
        TN_HIP(hipMemset(fv->centers.p, 0, sizeof(double) * fv->centers.n));
        TN_HIP(hipMemset(fv->TL.p, 0, sizeof(double) * fv->size));
        TN_HIP(hipMemset(fv->Rp.p, 0, sizeof(double) * fv->size));
        TN_HIP(hipMemset(fv->Rm.p, 0, sizeof(double) * fv->size));
        TN_HIP(hipMemset(fv->ad.p, 0, sizeof(double) * fv->ad.n));
        tn::launch_fv_columns(fv->args(), fv->columns.p, nullptr);
        tn::launch_fv_centers(fv->args(), fv->centers.p, nullptr);
        ghost_update(fv, fv->centers.p, 3, nullptr);
        TN_HIP(hipStreamSynchronize(nullptr));
    } catch (...) {
        delete fv;
        throw;
    }
    *out = fv;
    TN_API_END
}

int tn_fv_destroy(tn_fv* fv)
{
    TN_API_BEGIN
    delete fv;
    TN_API_END
}

int tn_fv_size(const tn_fv* fv, int64_t* elems)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && elems, "null pointer");
    *elems = fv->size;
    TN_API_END
}

int tn_fv_cell_centers(const tn_fv* fv, double* out, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && out, "null pointer");
    TN_HIP(hipMemcpyAsync(out, fv->centers.p, sizeof(double) * 3 * fv->size, hipMemcpyDeviceToDevice,
                          (hipStream_t)s));
    TN_API_END
}

int tn_fv_ghost_update(tn_fv* fv, double* f, int ncomp, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && f, "null pointer");
    TN_REQUIRE(ncomp >= 1 && ncomp <= 3, "ncomp must be 1..3");
    ghost_update(fv, f, ncomp, (hipStream_t)s);
    TN_API_END
}

int tn_fv_dt_stable(tn_fv* fv, const double* vel, double kappa, double* dt, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && vel && dt, "null pointer");
    tn::launch_fv_dt(fv->args(), vel, kappa, fv->part.p, fv->result.p, (hipStream_t)s);
    TN_HIP(hipMemcpyAsync(dt, fv->result.p, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)s));
    TN_HIP(hipStreamSynchronize((hipStream_t)s));
    TN_API_END
}

int tn_fv_apply_dirichlet(tn_fv* fv, double* f, int apply_cmb, double t_cmb, int apply_surf, double t_surf,
                          tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && f, "null pointer");
    tn::launch_fv_dirichlet(fv->n, fv->nr, f, apply_cmb, t_cmb, apply_surf, t_surf, (hipStream_t)s);
    TN_API_END
}

int tn_fct_predictor(tn_fv* fv, double* T, const double* vel, double dt, double kappa, const double* source,
                     int subtract_divergence, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && T && vel, "null pointer");
    ghost_update(fv, T, 1, (hipStream_t)s);
    tn::launch_fct_predictor(fv->args(), vel, T, dt, kappa, source, subtract_divergence, fv->TL.p, fv->ad.p,
                             (hipStream_t)s);
    TN_API_END
}

int tn_fct_limiter(tn_fv* fv, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv, "null pointer");
    ghost_update(fv, fv->TL.p, 1, (hipStream_t)s);
    tn::launch_fct_limiter(fv->n, fv->nr, fv->TL.p, fv->ad.p, fv->Rp.p, fv->Rm.p, (hipStream_t)s);
    ghost_update(fv, fv->Rp.p, 1, (hipStream_t)s);
    ghost_update(fv, fv->Rm.p, 1, (hipStream_t)s);
    TN_API_END
}

int tn_fct_correction(tn_fv* fv, double* T, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && T, "null pointer");
    tn::launch_fct_correction(fv->n, fv->nr, fv->TL.p, fv->ad.p, fv->Rp.p, fv->Rm.p, T, (hipStream_t)s);
    TN_API_END
}

int tn_fct_buffers(tn_fv* fv, double** TL, double** antidiff, double** Rp, double** Rm)
{
    TN_API_BEGIN
    TN_REQUIRE(fv, "null pointer");
    if (TL) *TL = fv->TL.p;
    if (antidiff) *antidiff = fv->ad.p;
    if (Rp) *Rp = fv->Rp.p;
    if (Rm) *Rm = fv->Rm.p;
    TN_API_END
}

int tn_fct_explicit_step(tn_fv* fv, double* T, const double* vel, double dt, double kappa, const double* source,
                         int subtract_divergence, int apply_cmb, double t_cmb, int apply_surf, double t_surf,
                         tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && T && vel, "null pointer");
    const hipStream_t st = (hipStream_t)s;
    ghost_update(fv, T, 1, st);
    tn::launch_fct_predictor(fv->args(), vel, T, dt, kappa, source, subtract_divergence, fv->TL.p, fv->ad.p, st);
    if (apply_cmb || apply_surf)
        tn::launch_fv_dirichlet(fv->n, fv->nr, fv->TL.p, apply_cmb, t_cmb, apply_surf, t_surf, st);
    ghost_update(fv, fv->TL.p, 1, st);
    tn::launch_fct_limiter(fv->n, fv->nr, fv->TL.p, fv->ad.p, fv->Rp.p, fv->Rm.p, st);
    ghost_update(fv, fv->Rp.p, 1, st);
    ghost_update(fv, fv->Rm.p, 1, st);
    tn::launch_fct_correction(fv->n, fv->nr, fv->TL.p, fv->ad.p, fv->Rp.p, fv->Rm.p, T, st);
    TN_API_END
}

int tn_fct_upwind_step(tn_fv* fv, double* T, const double* vel, double dt, double kappa, const double* source,
                       int subtract_divergence, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && T && vel, "null pointer");
    const hipStream_t st = (hipStream_t)s;
    ghost_update(fv, T, 1, st);
    tn::launch_fct_predictor(fv->args(), vel, T, dt, kappa, source, subtract_divergence, fv->TL.p, fv->ad.p, st);
    TN_HIP(hipMemcpyAsync(T, fv->TL.p, sizeof(double) * fv->size, hipMemcpyDeviceToDevice, st));
    TN_API_END
}

int tn_fv_to_fe_lumped(tn_fv* fv, const double* Tfv, double* Tfe, double* tmp, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && Tfv && Tfe && tmp, "null pointer");
    TN_REQUIRE(Tfe != tmp, "tn_fv_to_fe_lumped: Tfe and tmp must be different arrays");
    const hipStream_t hs = (hipStream_t)s;
    tn::launch_fv_to_fe_rhs(fv->args(), Tfv, Tfe, hs);
    tn::launch_interface_sum<double>(fv->mesh->lvl, 1, Tfe, hs);
    tn::launch_fv_to_fe_rhs(fv->args(), nullptr, tmp, hs);
    tn::launch_interface_sum<double>(fv->mesh->lvl, 1, tmp, hs);
    tn::launch_fv_lumped_scale(tmp, Tfe, 10LL * fv->n * fv->n * fv->nr, hs);
    TN_API_END
}

int tn_fe_to_fv(tn_fv* fv, const double* Tfe, double* Tfv, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && Tfe && Tfv, "null pointer");
    tn::launch_fe_to_fv(fv->args(), Tfe, Tfv, (hipStream_t)s);
    ghost_update(fv, Tfv, 1, (hipStream_t)s);
    TN_API_END
}

int tn_fv_to_fe_rhs(tn_fv* fv, const double* Tfv, double* b, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(fv && Tfv && b, "null pointer");
    tn::launch_fv_to_fe_rhs(fv->args(), Tfv, b, (hipStream_t)s);
    tn::launch_interface_sum<double>(fv->mesh->lvl, 1, b, (hipStream_t)s);
    TN_API_END
}

}  // extern "C"
