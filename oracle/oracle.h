/*
 * oracle.h -- CPU restatement of TerraNeo's matrix-free Stokes hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the shipped
 * terraneo_amd library never links or calls it.
 *
 * Plain C11.  Every function follows a reference file:line, cited beside it.
 * It deliberately restates the reference's *slow* (assembled 18x18 local
 * matrix) formulation, so that it is independent of the matrix-free
 * reformulation used by the HIP kernels.
 *
 * Layouts (shared with the product, see include/terraneo_amd.h):
 *   lateral coords  [10][n][n][3]          unit-sphere xyz (double)
 *   radii           [nr]
 *   scalar field    [10][n][n][nr]         r fastest
 *   vector field    [3][10][n][n][nr]      component-major SoA
 */
#ifndef TN_ORACLE_H
#define TN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_NEUMANN = 0, ORC_DIRICHLET = 1, ORC_FREESLIP = 2 };

/* ---------------------------------------------------------------- mesh */
int  orc_lateral_coords(int lat_level, double* coords);
void orc_uniform_radii(int nr, double r_min, double r_max, double* radii);
/* group id of every lateral node (duplicates share an id); returns #groups */
int  orc_lateral_groups(int n, const double* coords, int32_t* group_of);
void orc_ownership(int n, int nr, const int32_t* group_of, int ngroups, uint8_t* owned);
void orc_interface_sum(int n, int nr, int ncomp, const int32_t* group_of, int ngroups, double* v);

/* ---------------------------------------------------------------- mesh handle */
typedef struct orc_mesh {
    int n, nr;            /* nodes per diamond side, radial nodes   */
    double* coords;       /* [10][n][n][3]                          */
    double* radii;        /* [nr]                                   */
    int32_t* group_of;    /* [10][n][n]                             */
    int ngroups;
    uint8_t* owned;       /* [10][n][n][nr]                         */
} orc_mesh;

orc_mesh* orc_mesh_create(int lat_level, int nr, const double* radii);
void      orc_mesh_destroy(orc_mesh* m);
int64_t   orc_mesh_nodes(const orc_mesh* m);

/* ---------------------------------------------------------------- epsilon-div-div */
/* raw wedge-local 18x18 matrix (dimension-wise dof order d*6+node), no BC */
void orc_eps_local_matrix(const orc_mesh* m, const double* k, int d, int x, int y, int r,
                          int wedge, double* A);
/* same with the quadrature-point viscosity given directly (keval) */
void orc_eps_local_matrix_k(const orc_mesh* m, double keval, int d, int x, int y, int r, int wedge, double* A);

typedef struct orc_eps {
    const orc_mesh* m;
    const double* k;      /* nodal viscosity [10][n][n][nr] */
    const double* kc;     /* optional per hex cell viscosity [10][n-1][n-1][nr-1]; overrides k */
    int bc_cmb, bc_surf;
    double* stored;       /* optional [10][n-1][n-1][nr-1][2][18*18] (GCA) */
    double* null_modes;   /* fs/fs penalty modes [3][3*N] */
    double penalty_eps;
    int penalty_once;     /* 0: added at every node copy before the interface sum (reference, F5);
                             1: added at the owned copy only (once per node: a symmetric operator) */
} orc_eps;

orc_eps* orc_eps_create(const orc_mesh* m, const double* k, int bc_cmb, int bc_surf);
void     orc_eps_destroy(orc_eps* op);
/* y = A x ; replace, penalty (matvec only), interface sum when comm != 0 */
void     orc_eps_apply(const orc_eps* op, const double* x, double* y, int diagonal, int comm);
/* install GCA/stored local matrices ([cells][2][324]); copies them */
void     orc_eps_set_stored(orc_eps* op, const double* mats);
/* effective local maps on Cartesian dimension-wise dofs (masks, rotations) */
void     orc_eps_cell_effective(const orc_eps* op, int d, int x, int y, int r, int diagonal, double* E);
/* COO over unique dofs 3*(group*nr + r) + comp, no penalty; returns count */
int64_t  orc_eps_assemble_coo(const orc_eps* op, int64_t* rows, int64_t* cols, double* vals);
/* gather the fs/fs penalty modes (each 3*N doubles); returns 0 when inactive */
int      orc_eps_penalty_modes(const orc_eps* op, double* out);
/* fs/fs penalty multiplicity: once = 0 reference (per copy), 1 once per node (owned copy) */
void     orc_eps_set_penalty_once(orc_eps* op, int once);
/* fs/fs penalty epsilon (set_penalty_epsilon, epsilon_divdiv_kerngen.hpp:279) */
void     orc_eps_set_penalty_epsilon(orc_eps* op, double eps);
/* per hex cell viscosity (both wedges): kwJ = kc[cell] |det J| */
void     orc_eps_set_cell_viscosity(orc_eps* op, const double* kc);
/* adjoint gradient w.r.t. per-cell log-viscosity m = ln kc:
 *   g[cell] = -kc[cell] sum_wedges (T lam)_e^T A_e(1) u_e
 * T: trial rule of the boundary shells (Dirichlet: zero, free-slip: project) */
void     orc_eps_grad_log_cell_viscosity(const orc_eps* op, const double* u, const double* lam, double* g);

/* ---------------------------------------------------------------- Stokes coupling */
/* B: fine velocity -> coarse pressure (coarse mesh level one lower) */
void orc_div_apply(const orc_mesh* mf, const orc_mesh* mc, int bc_cmb, int bc_surf,
                   const double* u, double* p, int comm);
/* B^T: coarse pressure -> fine velocity, ADDED into u */
void orc_grad_apply_add(const orc_mesh* mf, const orc_mesh* mc, int bc_cmb, int bc_surf,
                        const double* p, double* u, int comm);
void    orc_div_cell_effective(const orc_mesh* mf, int bc_cmb, int bc_surf, int d, int x, int y, int r, double* E);
int64_t orc_div_assemble_coo(const orc_mesh* mf, const orc_mesh* mc, int bc_cmb, int bc_surf, int64_t* rows,
                             int64_t* cols, double* vals);
/* y = M x per component (6-point quadrature), replace + interface sum */
void orc_vector_mass_apply(const orc_mesh* m, const double* x, double* y);
/* lumped k-weighted mass diagonal: d = rowsum(M_k) (interface summed) */
void orc_kmass_lumped(const orc_mesh* m, const double* k, double* d);

/* ---------------------------------------------------------------- transfers + GCA */
void orc_prolongate_add(const orc_mesh* mc, const orc_mesh* mf, int ncomp, const double* xc, double* xf);
void orc_restrict(const orc_mesh* mf, const orc_mesh* mc, int ncomp, const double* xf, double* xc);
/* coarse stored matrices from fine op: out [10][nc-1][nc-1][nrc-1][2][324] */
void orc_gca_build(const orc_eps* fine, const orc_mesh* mc, double* out);
/* variant with the geometric radial bracket (the exact Galerkin product P^T A P) */
void orc_gca_build_geometric(const orc_eps* fine, const orc_mesh* mc, double* out);

/* ---------------------------------------------------------------- FV energy step (fv_oracle.c)
 * FV scalar [10][n+1][n+1][nr+1] with one ghost layer (real cells 1..n-1,
 * 1..nr-1), cell centres [..][3], antidiffusive fluxes [..][6] */
int64_t orc_fv_size(const orc_mesh* m);
void    orc_fv_cell_centers(const orc_mesh* m, double* c);
void    orc_fv_ghost_update(const orc_mesh* m, double* f, int ncomp);
/* lateral ghost -> source real cell pairs (d, x, y, sd, sx, sy); returns the count */
int     orc_fv_ghost_pairs(const orc_mesh* m, int32_t* pairs);
void    orc_fv_geometry(const orc_mesh* m, const double* centers, const double* vel, int d, int xc, int yc, int rc,
                        double beta[6], double* M, double S[6][3]);
double  orc_fv_dt_stable(const orc_mesh* m, const double* centers, const double* vel, double kappa);
void    orc_fct_predictor(const orc_mesh* m, const double* centers, const double* vel, double* T, double dt,
                          double kappa, const double* source, int subtract_divergence, double* TL, double* antidiff);
void    orc_fv_dirichlet(const orc_mesh* m, double* f, int apply_cmb, double t_cmb, int apply_surf, double t_surf);
void    orc_fct_limiter(const orc_mesh* m, double* TL, const double* antidiff, double* Rp, double* Rm);
void    orc_fct_correction(const orc_mesh* m, const double* TL, const double* antidiff, const double* Rp,
                           const double* Rm, double* T);
/* work: 3 FV scalars + 6 per FV cell, zero on the first call */
void    orc_fct_explicit_step(const orc_mesh* m, const double* centers, const double* vel, double* T, double dt,
                              double kappa, const double* source, int subtract_divergence, int apply_cmb,
                              double t_cmb, int apply_surf, double t_surf, double* work);
/* b = int phi_i T (nodal scalar, interface summed) */
void    orc_fv_to_fe_rhs(const orc_mesh* m, const double* Tfv, double* b);

/* ---------------------------------------------------------------- BLAS-1 */
double orc_dot_owned(const orc_mesh* m, int ncomp, const double* a, const double* b);

#ifdef __cplusplus
}
#endif
#endif
