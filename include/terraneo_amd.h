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

/* ------------------------------------------------------------------ misc */
const char* tn_last_error(void);
int tn_version(void);
/* device memory helpers (hipMalloc / hipFree / memcpy); bytes */
int tn_malloc(void** ptr, size_t bytes);
int tn_free(void* ptr);
int tn_memcpy_h2d(void* dst, const void* src, size_t bytes);
int tn_memcpy_d2h(void* dst, const void* src, size_t bytes);
int tn_sync(void);

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
int tn_eps_set_eta(tn_eps* op, const void* eta);
/* apply_impl, kerngen.hpp:513-614 */
int tn_eps_apply(tn_eps* op, const void* src, void* dst, tn_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
