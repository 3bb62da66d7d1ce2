/*
 * terraneo_amd.h -- C ABI of the MI355X-native matrix-free Stokes hot path.
 *
 * Drop-in boundary for TerraNeo's operator/solver API (reference: the
 * `OperatorLike` / `GCACapable` / `SolverLike` concepts in
 * src/terra/linalg/operator.hpp:56-141 and src/terra/linalg/solvers/solver.hpp:19-54).
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an int status (0 = ok, < 0 = error; tn_last_error() gives text).
 *
 * Memory layouts (device pointers unless stated otherwise):
 *   scalar field   T[10][n][n][nr]            r fastest (Grid4DDataScalar,
 *                                              src/terra/grid/grid_types.hpp:20)
 *   vector field   T[3][10][n][n][nr]         component-major SoA (Grid4DDataVec,
 *                                              grid_types.hpp:30-75), stride N
 *   N = 10 * n * n * nr  (all stored nodes, interface duplicates included)
 * T is double (TN_F64) or float (TN_F32), chosen per operator.
 *
 * Streams: every compute call takes a hipStream_t (NULL = default stream)
 * and is asynchronous with respect to the host unless documented otherwise.
 * Device-side failures of the matrix-free march kernels (a bounded hand-off
 * wait between waves that gave up) set a host-visible word; the next
 * synchronising call (tn_sync, tn_dot, tn_max_abs_entry, tn_has_nan_or_inf,
 * the solvers) or the next apply returns -1 with that message, and the word
 * is cleared.
 * Objects are not re-entrant: one stream at a time per object (the
 * reference's operators keep mutable src/dst views, epsilon_divdiv_kerngen.hpp:94-96).
 */
#ifndef TERRANEO_AMD_H
#define TERRANEO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tn_stream_t; /* == hipStream_t */

enum tn_dtype { TN_F64 = 0, TN_F32 = 1 };
/* BoundaryConditionFlag, src/terra/grid/shell/bit_masks.hpp:24-29 */
enum tn_bc { TN_NEUMANN = 0, TN_DIRICHLET = 1, TN_FREESLIP = 2 };
/* OperatorApplyMode / OperatorCommunicationMode, src/terra/linalg/operator.hpp:29-45 */
enum tn_apply_mode { TN_REPLACE = 0, TN_ADD = 1 };
enum tn_comm_mode { TN_COMM_SKIP = 0, TN_COMM_ADDITIVE = 1 };

typedef struct tn_mesh tn_mesh;
typedef struct tn_eps tn_eps;
typedef struct tn_stokes tn_stokes;
typedef struct tn_gmg tn_gmg;
typedef struct tn_fv tn_fv;
typedef struct tn_xdmf tn_xdmf;

/* Stokes solver variants (block preconditioners of block_preconditioner_2x2.hpp) */
enum tn_stokes_method { TN_MINRES_BLOCKDIAG = 0, TN_FGMRES_BLOCKTRI = 1, TN_FGMRES_BLOCKDIAG = 2 };

/* geometric multigrid options (Chebyshev parameters chebyshev.hpp:50-80,
 * Multigrid cycles multigrid.hpp:90-104, coarsest PCG pcg.hpp:45-66) */
typedef struct tn_gmg_options {
    int cheby_order;           /* Chebyshev degree                           */
    int cheby_iterations;      /* smoothing steps per pre/post smoothing     */
    int power_iterations;      /* lambda_max(D^-1 A) estimate                */
    int num_cycles;            /* V-cycles per application                   */
    int coarse_max_iterations; /* coarsest-level PCG                         */
    double coarse_rtol;
    double coarse_atol;
    uint64_t seed;             /* power-iteration start vector               */
    int coarse_solver;         /* TN_COARSE_PCG (the reference's) or TN_COARSE_DENSE */
    /* 0: the reference's V-cycle.  1: on levels with free-slip shells the
     * fine residual is projected tangentially before restriction and the
     * iterate after the prolongated correction (P' = Pi P, R' = P^T Pi with
     * Pi the free-slip projection): the constant prolongation of a coarse
     * tangential field is not tangential at the fine boundary nodes. */
    int freeslip_projected_transfer;
    /* 0: the reference's Jacobi scaling, the entrywise inverse of the
     * operator's diagonal.  1: at free-slip node copies, the inverse of the
     * operator's 3x3 free-slip diagonal R^T diag(R B R^T) R instead (its
     * entrywise inversion is not that inverse). */
    int freeslip_block_jacobi;
} tn_gmg_options;

/* coarsest-level solver of the multigrid: unpreconditioned PCG as the
 * reference, or a dense LU inverse of the assembled coarsest operator
 * (rocSOLVER; at most 8000 owned coarse nodes) */
enum tn_coarse_solver { TN_COARSE_PCG = 0, TN_COARSE_DENSE = 1 };
/* radial bracket of the Galerkin coarsening's local prolongation:
 * TN_GCA_REFERENCE reproduces gca.hpp:213-216 (r_fine/2), TN_GCA_GEOMETRIC
 * uses the coarse cell the fine node lies in (exact P^T A P) */
enum tn_gca_bracket { TN_GCA_REFERENCE = 0, TN_GCA_GEOMETRIC = 1 };

/* IterativeSolverParameters / FGMRESOptions (iterative_solver_info.hpp, fgmres.hpp:20-44) */
typedef struct tn_krylov_options {
    int max_iterations;
    double rtol;
    double atol;
    int restart; /* FGMRES only */
    /* tn_stokes_solve only.  0: solve K x = b as given (reference behaviour).
     * 1: solve in the pressure space orthogonal to constants: the operator
     * becomes Pi K Pi with Pi removing the owned-node mean of the pressure
     * block, b and the result are projected the same way.  This is the
     * system bordered by the mean-pressure constraint; it drops the
     * near-null constant-pressure mode that free-slip shells with nodal
     * normals leave in K (B^T 1 is then small but not zero). */
    int pressure_mean_free;
} tn_krylov_options;

typedef struct tn_solve_info {
    int iterations;
    double initial_residual;
    double final_residual;
    double seconds;
    int history_len;  /* number of residual entries produced (iteration 0 included) */
    int restarts;     /* FGMRES breakdown / stagnation restarts */
    int prec_skipped; /* FGMRES preconditioner outputs rejected as non-finite */
} tn_solve_info;

/* ------------------------------------------------------------------ misc */
const char* tn_last_error(void);
int tn_version(void);
/* device memory helpers (hipMalloc / hipFree / memcpy); bytes */
int tn_malloc(void** ptr, size_t bytes);
int tn_free(void* ptr);
int tn_memcpy_h2d(void* dst, const void* src, size_t bytes);
int tn_memcpy_d2h(void* dst, const void* src, size_t bytes);
int tn_sync(void);
/* Test hook, no counterpart in the reference: polls per bounded hand-off wait
 * of the march kernels (default 2^20).  A small value forces the waits to give
 * up, so the error path above can be exercised; every kernel still exits. */
int tn_set_wait_limit(int polls);

/* ------------------------------------------------------------------ mesh
 * One refinement level of the thick spherical shell, 10 diamonds, one
 * subdomain per diamond.  Replaces DistributedDomain::create_uniform +
 * subdomain_unit_sphere_single_shell_coords + subdomain_shell_radii +
 * setup_node_ownership_mask_data + setup_boundary_mask_data
 * (src/terra/grid/shell/spherical_shell.hpp:2547-2588, 2723-2784;
 *  src/terra/grid/bit_masks.hpp:29-85; src/terra/grid/shell/bit_masks.hpp:97-136).
 * Coordinates are bit-identical to the reference generator.
 */
int tn_mesh_create(int lat_level, int rad_level, double r_min, double r_max, tn_mesh** out);
/* DistributedDomain::create_uniform(lat, rad, r_min, r_max, s_lat, s_rad)
 * (spherical_shell.hpp:2547-2588): diamonds split into 4^s_lat lateral and
 * 2^s_rad radial subdomains.  Only s_lat = s_rad = 0 (one subdomain per
 * diamond, the layout of every entry point here) is accepted; other splits
 * return an error (a caller remaps split-subdomain fields to the per-diamond
 * layout, which holds the same nodes). */
int tn_mesh_create_subdomains(int lat_level, int rad_level, double r_min, double r_max, int s_lat, int s_rad,
                              tn_mesh** out);
/* explicit radii (host array of nr values, increasing) */
int tn_mesh_create_radii(int lat_level, int nr, const double* radii, tn_mesh** out);
int tn_mesh_destroy(tn_mesh* m);
/* n (lateral nodes per diamond side), nr (radial nodes), N (stored nodes) */
int tn_mesh_dims(const tn_mesh* m, int* n, int* nr, int64_t* num_nodes);
/* number of OWNED nodes (each duplicated interface node counted once) */
int64_t tn_mesh_owned_nodes(const tn_mesh* m);
/* host copies: coords [10][n][n][3], radii [nr], owned [10][n][n][nr] (0/1) */
int tn_mesh_lateral_coords(const tn_mesh* m, double* host_out);
int tn_mesh_radii(const tn_mesh* m, double* host_out);
int tn_mesh_owned_mask(const tn_mesh* m, uint8_t* host_out);
/* number of duplicated lateral node groups and their total membership */
int tn_mesh_interface_info(const tn_mesh* m, int* ngroups, int* nmembers);

/* ------------------------------------------------------------------ BLAS-1 on mesh fields
 * (grid_operations.hpp:44-785, vector_q1.hpp:204-292).  ncomp = 1 or 3.
 * Dots count OWNED nodes only and accumulate in double. */
int tn_interface_sum(const tn_mesh* m, int dtype, int ncomp, void* v, tn_stream_t s);
/* y = c0 + sum_i c[i] * x[i], i < nx <= 3 (x[i] may alias y) */
int tn_lincomb(const tn_mesh* m, int dtype, int ncomp, void* y, int nx, const double* c,
               const void* const* x, double c0, tn_stream_t s);
/* result written to host *out (synchronises the stream) */
int tn_dot(const tn_mesh* m, int dtype, int ncomp, const void* a, const void* b, double* out, tn_stream_t s);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * tn_dot with its per-block partial sums and result slot in a static per-GPU
 * device array (16 640 bytes: at most 2048 partial sums, which the launch
 * never exceeds) instead of the mesh's allocated buffers; calls no hipMalloc.
 * The mesh must already be resident on the device (any earlier device call on
 * it uploads it; this entry point does not).  ncomp 1..3.  Calls that overlap
 * on one GPU (two streams or threads) race on the shared array.  Result on the
 * host as tn_dot; bit-identical to tn_dot. */
int tn_dot_noalloc(const tn_mesh* m, int dtype, int ncomp, const void* a, const void* b, double* out,
                   tn_stream_t s);
/* v = 1 / v entrywise (invert_entries, grid_operations.hpp:252-270) */
int tn_invert_entries(const tn_mesh* m, int dtype, int ncomp, void* v, tn_stream_t s);
/* y = y * d entrywise (scale_with_vector / mult, grid_operations.hpp:272-294) */
int tn_scale_with_vector(const tn_mesh* m, int dtype, int ncomp, void* y, const void* d, tn_stream_t s);
/* uniform [0, 1) entries, each stored node copy drawn independently like the
 * reference's rand (grid_operations.hpp:758-783); counter-based (SplitMix64 of
 * seed and index), so the values differ from Kokkos' random pool */
int tn_randomize(const tn_mesh* m, int dtype, int ncomp, void* y, uint64_t seed, tn_stream_t s);
/* max |v| over all stored entries (max_abs_entry, grid_operations.hpp:336-368); host *out */
int tn_max_abs_entry(const tn_mesh* m, int dtype, int ncomp, const void* v, double* out, tn_stream_t s);
/* *out = 1 if any entry is NaN or infinite (has_nan_or_inf, grid_operations.hpp:713-743) */
int tn_has_nan_or_inf(const tn_mesh* m, int dtype, int ncomp, const void* v, int* out, tn_stream_t s);

/* ------------------------------------------------------------------ epsilon-div-div operator
 * EpsilonDivDivKerngen<T> (src/terra/fe/wedge/operators/shell/epsilon_divdiv_kerngen.hpp:61-2112):
 * y = A(eta) x with a(u,v) = int 2 eta eps(u):eps(v) - 2/3 eta div u div v on wedges,
 * one-point (centroid) quadrature, eta nodal (mean of the 6 wedge nodes).
 * Dirichlet rows/columns keep only their diagonal; free-slip applies the
 * tangential projection P A P plus the normal diagonal; with free-slip on
 * both shells the rotation null-space penalty is added before the
 * interface sum (as the reference does, :583-613).
 * The operator keeps the eta pointer (a view), like the reference. */
int tn_eps_create(const tn_mesh* m, int dtype, const void* eta, int bc_cmb, int bc_surf, tn_eps** out);
int tn_eps_destroy(tn_eps* op);
/* set_operator_apply_and_communication_modes, kerngen.hpp:410-416 */
int tn_eps_set_modes(tn_eps* op, int apply_mode, int comm_mode);
/* set_diagonal, kerngen.hpp:418 */
int tn_eps_set_diagonal(tn_eps* op, int diagonal);
/* set_boundary_conditions, kerngen.hpp:420-427 */
int tn_eps_set_boundary_conditions(tn_eps* op, int bc_cmb, int bc_surf);
/* Replace the viscosity buffer (the reference operator's k_ view, kerngen.hpp:94-96).
 * The operator keeps a pointer, not a copy, but caches per-viscosity data for
 * the two boundary shells: after changing viscosity values in place, call this
 * again (the same pointer is fine) before the next apply. */
int tn_eps_set_eta(tn_eps* op, const void* eta);
/* apply_impl, kerngen.hpp:513-614 */
int tn_eps_apply(tn_eps* op, const void* src, void* dst, tn_stream_t s);
/* penalty epsilon of the fs/fs rotation penalty (set_penalty_epsilon, kerngen.hpp:279) */
int tn_eps_set_penalty_epsilon(tn_eps* op, double eps);
/* No counterpart in the reference.  Multiplicity of the fs/fs rotation penalty:
 * once = 0 (default) adds it at every node copy before the interface sum, as
 * the reference does (kerngen.hpp:583-613), so an interface node receives it
 * once per copy and the operator is not symmetric; once = 1 adds it at the
 * owned copy only (once per node), which makes the operator symmetric, as
 * CG and MINRES (and a symmetric multigrid preconditioner) require. */
int tn_eps_set_penalty_once(tn_eps* op, int once);
/* Galerkin coarse operator from the next finer operator (TwoGridGCA, gca.hpp:72-531;
 * switches `coarse` to its stored-matrix path, set_stored_matrix_mode kerngen.hpp:443-464).
 * Matrices are stored symmetric-packed in TN_F64 or TN_F32. */
int tn_eps_gca_from(tn_eps* coarse, const tn_eps* fine, int store_dtype, tn_stream_t s);
/* same with an explicit radial bracket (tn_gca_bracket); tn_eps_gca_from == TN_GCA_REFERENCE */
int tn_eps_gca_from_bracket(tn_eps* coarse, const tn_eps* fine, int store_dtype, int bracket, tn_stream_t s);
/* GCAElementsCollector (gca_elements_collector.hpp:32-96): flags (a scalar field on `coarsest`, dtype `dtype`)
 * is set to 1 at the coarsest-level cells (x, y, r) that have a descendant cell on `fine` with a wedge where the
 * reference-coordinate gradient of the nodal coefficient k exceeds 10 at the centroid, and to 0 elsewhere. */
int tn_gca_collect_elements(const tn_mesh* fine, const tn_mesh* coarsest, int dtype, const void* k, void* flags,
                            tn_stream_t s);
/* Selective stored-matrix mode (OperatorStoredMatrixMode::Selective, local_matrix_storage.hpp:68-100): after
 * tn_eps_gca_from, the cells of `coarse` whose coarsest ancestor is not flagged use the rediscretized wedge
 * matrices of `coarse` (its geometry and nodal viscosity, as the reference's on-the-fly assembly for cells
 * without a stored matrix); flagged cells keep their Galerkin matrices.  flags in `coarse`'s dtype. */
int tn_eps_gca_select(tn_eps* coarse, const tn_mesh* coarsest, const void* flags, tn_stream_t s);
/* GCACapable::get_local_matrix (kerngen.hpp:481-497), host out[18*18],
 * dimension-wise dof order (component*6 + wedge node); synchronous.  Stored
 * operators return the stored matrix; matrix-free ones assemble the wedge's
 * Neumann matrix from the geometry and viscosity (assemble_local_matrix,
 * kerngen.hpp:2051-2108; computed in FP64 for either operator dtype). */
int tn_eps_get_local_matrix(const tn_eps* op, int sd, int x, int y, int r, int wedge, double* out);
/* GCACapable::set_local_matrix (kerngen.hpp:468-479, operator.hpp:80-112): host in[18*18],
 * same order; the operator must have stored matrices; storage is packed
 * symmetric, so a nonsymmetric matrix (|A - A^T| > 1e-12 max|A|) is rejected.
 * The boundary forms are rebuilt at the next apply.  Synchronous. */
int tn_eps_set_local_matrix(tn_eps* op, int sd, int x, int y, int r, int wedge, const double* in);
/* set_stored_matrix_mode(Full) (kerngen.hpp:443-464, local_matrix_storage.hpp:28-243):
 * the operator switches to its stored-matrix path, every cell holding its own
 * wedge matrices (those get_local_matrix assembles), stored in store_dtype. */
int tn_eps_set_stored_matrices(tn_eps* op, int store_dtype, tn_stream_t s);

/* ------------------------------------------------------------------ adjoint viscosity gradient
 * No counterpart in the reference (its viscosity is nodal, kerngen.hpp:94-96).
 * Per-hex-cell viscosity eta_cell [10][n-1][n-1][nr-1] (device, operator dtype):
 * both wedges of a cell use kwJ = eta_cell |det J|; overrides the nodal
 * viscosity; NULL returns to it.  Same caching rule as tn_eps_set_eta. */
int tn_eps_set_cell_eta(tn_eps* op, const void* eta_cell);
/* Backward half of the per-cell log-viscosity parameterisation m = ln eta_cell:
 *   grad[cell] = -eta_cell[cell] (T lambda)_cell^T A_cell(1) u_cell
 * (u: converged velocity, lambda: velocity part of the adjoint solution,
 * T: the boundary shells' trial rule).  grad: [10][n-1][n-1][nr-1]. */
int tn_eps_grad_log_cell_eta(tn_eps* op, const void* u, const void* lambda, void* grad, tn_stream_t s);

/* ------------------------------------------------------------------ Stokes block
 * EpsDivDivStokes (epsilon_divdiv_stokes.hpp:19-130): [A B^T; B 0] with the
 * Q1-iso-Q2 velocity on `vel` and Q1 pressure on `pre` (one level coarser).
 * Pressure fields are T[10][n_p][n_p][nr_p]. */
int tn_stokes_create(const tn_mesh* vel, const tn_mesh* pre, tn_eps* A, tn_stokes** out);
int tn_stokes_destroy(tn_stokes* K);
int tn_stokes_apply(tn_stokes* K, const void* u, const void* p, void* yu, void* yp, tn_stream_t s);
/* Divergence B (divergence.hpp:90-115): p = B u, replace + additive communication */
int tn_div_apply(tn_stokes* K, const void* u, void* p, tn_stream_t s);
/* Gradient B^T (gradient.hpp:87-112) */
int tn_grad_apply(tn_stokes* K, const void* p, void* u, int apply_mode, int comm_mode, tn_stream_t s);

/* ------------------------------------------------------------------ mass, transfers, boundary helpers */
/* VectorMass (vector_mass.hpp:59-153): y = M x, replace + communication */
int tn_vector_mass_apply(const tn_mesh* m, int dtype, const void* x, void* y, tn_stream_t s);
/* lumped KMass (kmass.hpp:17-185, set_lumped_diagonal): d = rowsum(M_k) */
int tn_kmass_lumped(const tn_mesh* m, int dtype, const void* k, void* d, tn_stream_t s);
/* ProlongationVecConstant in Add mode (prolongation_constant.hpp:103-193): xf += P xc */
int tn_prolongate_add(const tn_mesh* coarse, const tn_mesh* fine, int dtype, int ncomp, const void* xc, void* xf,
                      tn_stream_t s);
/* RestrictionVecConstant (restriction_constant.hpp:134-250): xc = R xf */
int tn_restrict(const tn_mesh* fine, const tn_mesh* coarse, int dtype, int ncomp, const void* xf, void* xc,
                tn_stream_t s);
/* ProlongationLinear / ProlongationVecLinear (prolongation_linear.hpp:16-191 / 193-370; weights
 * grid_transfer_linear.hpp:13-157): geometry-aware linear interpolation, exact for affine fields.
 * xf = P xc (apply_mode TN_REPLACE) or xf += P xc (TN_ADD); ncomp 1 (scalar) or 3 (vector).
 * vec_early_return = 1 reproduces ProlongationVecLinear's return inside its component loop
 * (:284,298): at fine nodes that coincide with or lie radially between coarse nodes only
 * component 0 is transferred. */
int tn_prolongate_linear(const tn_mesh* coarse, const tn_mesh* fine, int dtype, int ncomp, const void* xc, void* xf,
                         int apply_mode, int vec_early_return, tn_stream_t s);
/* RestrictionLinear / RestrictionVecLinear (restriction_linear.hpp:19-251 / 253-449): xc = P^T xf over owned
 * fine nodes, then additive communication.  vec_early_return as above (:354,367). */
int tn_restrict_linear(const tn_mesh* fine, const tn_mesh* coarse, int dtype, int ncomp, const void* xf, void* xc,
                       int vec_early_return, tn_stream_t s);
/* assign_masked_else_keep_old on shells: which = 1 (CMB) | 2 (SURFACE);
 * src NULL assigns zero (strong_algebraic_dirichlet_enforcement.hpp:88-123) */
int tn_boundary_assign(const tn_mesh* m, int dtype, int ncomp, void* dst, const void* src, int which, tn_stream_t s);
/* strong_algebraic_freeslip_enforcement_in_place (strong_algebraic_freeslip_enforcement.hpp:21-37) */
int tn_freeslip_project(const tn_mesh* m, int dtype, void* v, int which, tn_stream_t s);

/* ------------------------------------------------------------------ app caller (mantle circulation)
 * Stokes right-hand side of one time step of the mantle-circulation app
 * (mantlecirculation.cpp:876-898 with RHSVelocityInterpolator :117-145):
 * f_u = M (Ra T n), n the unit position vector, then the velocity boundary
 * conditions of K's viscous block enforced strongly (Dirichlet shells zero,
 * free-slip shells without normal component); f_p = 0.  T[10][n][n][nr] on the
 * velocity mesh, in K's dtype. */
int tn_stokes_rhs_from_temperature(tn_stokes* K, const void* temperature, double rayleigh, void* f_u, void* f_p,
                                   tn_stream_t s);
/* pressure "normalisation" after each Stokes solve (mantlecirculation.cpp:926-932):
 * p -= (sum of p over owned nodes) / (number of owned nodes) */
int tn_pressure_normalize(const tn_mesh* m, int dtype, void* p, tn_stream_t s);

/* ------------------------------------------------------------------ solvers
 * Multigrid with Chebyshev smoothing over `nlevels` operators ordered coarse
 * to fine (Multigrid multigrid.hpp:24-212, Chebyshev chebyshev.hpp:50-177).
 * The operators are referenced, not owned.  All levels share the finest
 * level's dtype, except the coarsest, which may be TN_F64 under a TN_F32
 * hierarchy (its PCG solve then runs in FP64 behind a precision bridge). */
int tn_gmg_create(int nlevels, tn_eps* const* ops, const tn_gmg_options* opts, tn_stream_t s, tn_gmg** out);
int tn_gmg_destroy(tn_gmg* g);
/* x <- V-cycle(s) for A_finest x = b, starting from x */
int tn_gmg_solve(tn_gmg* g, void* x, const void* b, tn_stream_t s);
int tn_gmg_get_lambda(const tn_gmg* g, int level, double* lambda);
int tn_gmg_set_lambda(tn_gmg* g, int level, double lambda);
/* PCG (pcg.hpp:78-153) on the viscous operator, optional multigrid
 * preconditioner; history (if not NULL) receives up to history_cap residuals */
int tn_pcg_solve(tn_eps* A, tn_gmg* prec, void* x, const void* b, const tn_krylov_options* o, tn_solve_info* info,
                 double* history, int history_cap, tn_stream_t s);
/* Stokes solve: PMINRES with block-diagonal preconditioner or FGMRES(m) with
 * block-triangular / block-diagonal preconditioner; velocity block by
 * `prec_u` (its dtype may differ: FP32 multigrid inside FP64 Krylov), Schur
 * block by the inverse of `schur_diag` (pressure field, e.g. lumped M_p(1/eta)) */
int tn_stokes_solve(tn_stokes* K, int method, tn_gmg* prec_u, const void* schur_diag, void* u, void* p,
                    const void* bu, const void* bp, const tn_krylov_options* o, tn_solve_info* info, double* history,
                    int history_cap, tn_stream_t s);

/* ---------------------------------------------------------------- energy step (hex finite volumes)
 * FV scalars are FP64 [10][n+1][n+1][nr+1] (r fastest) with one ghost layer,
 * the layout of the reference's VectorFVScalar (linalg/vector_fv.hpp): real
 * cells (xc, yc, rc) in [1, n-1]^2 x [1, nr-1] are the FE hex cells
 * (xc-1, yc-1, rc-1) of the velocity mesh.  Velocity: the FP64 nodal vector
 * field of that mesh, [3][10][n][n][nr].  Results repeat bit for bit (no
 * atomics). */
/* FV state on the velocity mesh: cell centres with their ghost layers
 * (fv/hex/helpers.hpp:177-190 initialize_cell_centers), the FCT work
 * arrays (FVFCTBuffers, fct_advection_diffusion.hpp:27-76) and a table of the
 * lateral cell geometry per hex column (160 bytes per column) */
int tn_fv_create(tn_mesh* m, tn_fv** out);
int tn_fv_destroy(tn_fv* fv);
/* FV scalar elements including ghosts: 10 (n+1)^2 (nr+1) */
int tn_fv_size(const tn_fv* fv, int64_t* elems);
/* device copy of the cell centres [size][3] (ghost layers filled; the radial
 * ghosts at the shells hold zero, as in the reference) */
int tn_fv_cell_centers(const tn_fv* fv, double* out, tn_stream_t s);
/* lateral ghost cells <- the neighbouring diamonds' real cells
 * (communication/shell/fv_communication.hpp update_fv_ghost_layers); ncomp
 * components innermost (1 for scalars, 3 for centres) */
int tn_fv_ghost_update(tn_fv* fv, double* f, int ncomp, tn_stream_t s);
/* largest stable explicit step min_K M_K / lambda_K (compute_dt_stable,
 * fct_advection_diffusion.hpp:110-231); result on the host */
int tn_fv_dt_stable(tn_fv* fv, const double* vel, double kappa, double* dt, tn_stream_t s);
/* Dirichlet values on the shells: real boundary cells and the ghosts beyond
 * (apply_dirichlet_bcs, fv/hex/helpers.hpp:64-106) */
int tn_fv_apply_dirichlet(tn_fv* fv, double* f, int apply_cmb, double t_cmb, int apply_surf, double t_surf,
                          tn_stream_t s);
/* FCT stages (fct_predictor :404-483, fct_limiter :547-599, fct_correction
 * :660-715): T's ghosts updated, then T_L and the antidiffusive fluxes into
 * the state's buffers; the limiter updates T_L's ghosts, forms R+ / R- and
 * their ghosts; the correction writes T = T_L + sum alpha f.  source: FV
 * scalar or NULL */
int tn_fct_predictor(tn_fv* fv, double* T, const double* vel, double dt, double kappa, const double* source,
                     int subtract_divergence, tn_stream_t s);
int tn_fct_limiter(tn_fv* fv, tn_stream_t s);
int tn_fct_correction(tn_fv* fv, double* T, tn_stream_t s);
/* device pointers of the state's T_L, fluxes [6][size], R+, R- (any may be NULL) */
int tn_fct_buffers(tn_fv* fv, double** TL, double** antidiff, double** Rp, double** Rm);
/* one explicit FCT step (fct_explicit_step, :719-757): predictor, Dirichlet
 * values on T_L when apply_cmb / apply_surf, limiter, correction; T in place */
int tn_fct_explicit_step(tn_fv* fv, double* T, const double* vel, double dt, double kappa, const double* source,
                         int subtract_divergence, int apply_cmb, double t_cmb, int apply_surf, double t_surf,
                         tn_stream_t s);
/* first-order upwind step, no correction (upwind_explicit_step, :775-797) */
int tn_fct_upwind_step(tn_fv* fv, double* T, const double* vel, double dt, double kappa, const double* source,
                       int subtract_divergence, tn_stream_t s);
/* right-hand side of the FV -> FE L2 projection, b_i = sum_K T_K int_K phi_i,
 * interface summed (nodal scalar of the velocity mesh) */
int tn_fv_to_fe_rhs(tn_fv* fv, const double* Tfv, double* b, tn_stream_t s);
/* bound-preserving FV -> FE transfer through the lumped mass
 * (conversion.hpp:242-341, l2_project_fv_to_fe_lumped): Tfe_i = b_i / D_i with
 * b_i = sum_K T_K int_K phi_i and D_i = int phi_i; tmp: one nodal FP64 scratch
 * vector (receives D), distinct from Tfe */
int tn_fv_to_fe_lumped(tn_fv* fv, const double* Tfv, double* Tfe, double* tmp, tn_stream_t s);
/* FE -> FV cell averages T_K = int_K u_h / |K| with the Felippa rule, lateral
 * ghost layers updated (conversion.hpp:348-416, l2_project_fe_to_fv); the
 * radial ghost cells at the shells are left as they are */
int tn_fe_to_fv(tn_fv* fv, const double* Tfe, double* Tfv, tn_stream_t s);
/* L2 projection FV -> FE (l2_project_fv_to_fe, fv/hex/conversion.hpp:131-211):
 * CG on the consistent mass matrix, no preconditioner, from Tfe's content;
 * the reference uses 1000 iterations and rtol = atol = 1e-12 */
int tn_fv_to_fe(tn_fv* fv, const double* Tfv, double* Tfe, const tn_krylov_options* o, tn_solve_info* info,
                double* history, int history_cap, tn_stream_t s);

/* ---------------------------------------------------------------- XDMF output and checkpoints
 * The reference's directory format (io/xdmf.hpp:12-146) for one subdomain per
 * diamond: geometry.bin, topology.bin, checkpoint_metadata.bin (version 1),
 * <name>_<step>.bin and step_<step>.xmf.  XDMFOutput (:147-1090): create,
 * add device fields (this library's layout, dtype TN_F32 / TN_F64, vec_dim 1
 * or 3, written as out_bytes 4 / 8 per value), then write() once per step
 * (copies the fields to the host; the first call also writes the mesh,
 * topology and metadata).  points_bytes / connectivity_bytes: 4 or 8 (the
 * reference's defaults 4 / 8). */
int tn_xdmf_create(const tn_mesh* m, const char* directory, int points_bytes, int connectivity_bytes,
                   tn_xdmf** out);
int tn_xdmf_destroy(tn_xdmf* x);
int tn_xdmf_add(tn_xdmf* x, const char* name, int dtype, int vec_dim, int out_bytes, const void* field);
int tn_xdmf_set_write_counter(tn_xdmf* x, int step);
int tn_xdmf_write(tn_xdmf* x, tn_stream_t s);
/* read_xdmf_checkpoint_grid (:1282-1420): field `name` of `step` into a device
 * array of this library's layout */
int tn_xdmf_read_checkpoint_grid(const tn_mesh* m, const char* directory, const char* name, int step, int dtype,
                                 int vec_dim, void* field, tn_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
