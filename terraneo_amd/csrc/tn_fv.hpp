// Hex finite-volume energy step (FCT advection-diffusion): argument block,
// launchers and the state object behind tn_fv.
#pragma once

#include "tn_internal.hpp"

namespace tn {

struct FvArgs {
    const double* coords;   // [10][n][n][3] unit lateral positions
    const double* radii;    // [nr]
    const double* centers;  // [10][n+1][n+1][nr+1][3] cell centres (ghost layers included)
    const double* columns;  // [10][n-1][n-1][FV_COLUMN] lateral geometry of each hex column
    int n, nr;
};

// Per hex column (d, x, y): the four lateral face normals (12), the two
// triangles' normals (6), the sum of the two wedges' |det(V0, V1, V2)| (1),
// one pad entry (rows of 160 bytes, 16-byte aligned)
constexpr int FV_COLUMN = 20;

int64_t fv_real_cells(int n, int nr);
int64_t fv_size(int n, int nr);  // FV scalar elements, ghost layers included
int fv_dt_blocks(int n, int nr);

void launch_fv_centers(const FvArgs& a, double* c, hipStream_t s);
void launch_fv_columns(const FvArgs& a, double* columns, hipStream_t s);
// pairs[k] = (ghost lateral index, source lateral index), lateral index (d (n+1) + x) (n+1) + y
void launch_fv_ghost(const int32_t* pairs, int cnt, int nr, int ncomp, double* f, hipStream_t s);
void launch_fv_dt(const FvArgs& a, const double* vel, double kappa, double* part, double* out, hipStream_t s);
void launch_fct_predictor(const FvArgs& a, const double* vel, const double* T, double dt, double kappa,
                          const double* source, int subtract_div, double* TL, double* ad, hipStream_t s);
void launch_fv_dirichlet(int n, int nr, double* f, int apply_cmb, double t_cmb, int apply_surf, double t_surf,
                         hipStream_t s);
void launch_fct_limiter(int n, int nr, const double* TL, const double* ad, double* Rp, double* Rm, hipStream_t s);
void launch_fct_correction(int n, int nr, const double* TL, const double* ad, const double* Rp, const double* Rm,
                           double* T, hipStream_t s);
void launch_fv_to_fe_rhs(const FvArgs& a, const double* Tfv, double* b, hipStream_t s);  // Tfv null: all ones
void launch_fv_lumped_scale(const double* D, double* u, int64_t N, hipStream_t s);      // u *= 1 / D
void launch_fe_to_fv(const FvArgs& a, const double* u, double* Tfv, hipStream_t s);    // real cells only

}  // namespace tn

// FV state on one velocity mesh: cell centres, the ghost-layer map and the
// FCT work arrays (the reference's FVFCTBuffers), allocated once
struct tn_fv {
    tn_mesh* mesh = nullptr;
    int n = 0, nr = 0;
    int64_t size = 0;                 // FV scalar elements (ghosts included)
    tn::DevBuf<double> centers;       // [size][3]
    tn::DevBuf<double> columns;       // [10 (n-1)^2][FV_COLUMN] lateral column geometry
    tn::DevBuf<double> TL, Rp, Rm;    // FV scalars; their radial ghost cells at the shells stay zero
    tn::DevBuf<double> ad;            // [6][size] antidiffusive fluxes
    tn::DevBuf<int32_t> pairs;        // [npairs][2]
    int npairs = 0;
    tn::DevBuf<double> part;          // per-block minima of the stable step
    tn::DevBuf<double> result;        // [1]
    tn::FvArgs args() const;
};
