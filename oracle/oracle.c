/*
 * oracle.c -- CPU restatement of TerraNeo's matrix-free Stokes hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke entry
 * point and bench.py's cpu_baseline leg load this library; the product
 * (libterraneo_amd.so) never links it.
 *
 * Each function cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).  The epsilon operator
 * is restated through the reference's assembled-local-matrix ("slow") path,
 * epsilon_divdiv_kerngen.hpp:823-1051 + 2051-2108, which the reference's own
 * fast-vs-slow test (tests/test_epsilon_divdiv_freeslip_fast_vs_slow.cpp)
 * shows equal to its matrix-free kernels to 1e-16 from one refinement on.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   mesh            closed forms (|x|=1, icosahedron vertices, midpoint rule)
 *   interface/owner invariants (every group owned once; 5-/3-valence poles)
 *   eps local       brute force (FD of the energy), symmetry, rigid-body kernel
 *   eps apply       = sum of local matrices; Stokes assess L2 errors measured
 *                   from the reference (level 3/4, Dirichlet and free-slip)
 *   div/grad        B^T == transpose(B) by assembly; Stokes assess errors
 *   mass            sum(M 1) == shell volume of the wedge mesh (closed form)
 *   P/R             R == P^T by assembly; P exact on constants and on the
 *                   piecewise-linear interpolant
 *   GCA             == P^T A P by brute force with the assembled P
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX4(n, nr, d, x, y, r) ((((int64_t)(d) * (n) + (x)) * (n) + (y)) * (nr) + (r))

/* ====================================================================== mesh */

/* Icosahedral base nodes and diamond->node map:
 * grid/shell/spherical_shell.hpp:441-527. */
static void icosahedron(double node[12][3], int dnode[10][4])
{
    const double fifthpi = 0.4 * asin(1.0);
    const double w = 2.0 * acos(1.0 / (2.0 * sin(fifthpi)));
    const double cosw = cos(w), sinw = sin(w);
    node[0][0] = 0.0; node[0][1] = 0.0; node[0][2] = 1.0;
    node[11][0] = 0.0; node[11][1] = 0.0; node[11][2] = -1.0;
    for (int k = 1; k <= 5; k++) {
        const double phi = 2.0 * ((double)k - 0.5) * fifthpi;
        node[k][0] = sinw * cos(phi);
        node[k][1] = sinw * sin(phi);
        node[k][2] = cosw;
    }
    for (int k = 1; k <= 5; k++) {
        const double phi = 2.0 * ((double)k - 1) * fifthpi;
        node[k + 5][0] = sinw * cos(phi);
        node[k + 5][1] = sinw * sin(phi);
        node[k + 5][2] = -cosw;
    }
    static const int t[10][4] = {{0, 5, 6, 1},  {0, 1, 7, 2},  {0, 2, 8, 3},  {0, 3, 9, 4},
                                 {0, 4, 10, 5}, {11, 7, 1, 6}, {11, 8, 2, 7}, {11, 9, 3, 8},
                                 {11, 10, 4, 9}, {11, 6, 5, 10}};
    memcpy(dnode, t, sizeof(t));
}

typedef struct {
    int N;
    double* val;   /* [(N+1)*(N+1)][3] */
    uint8_t* done; /* memo flags       */
} memo_t;

/* Recursive midpoint-normalise node rule with memoisation:
 * grid/shell/spherical_shell.hpp:230-346 (creation level search :287-306,
 * rules :316-338, normalisation dense/vec.hpp:76-107). */
static const double* node_rec(memo_t* mm, int i, int j)
{
    const int N = mm->N;
    double* out = mm->val + 3 * ((int64_t)i * (N + 1) + j);
    if (mm->done[(int64_t)i * (N + 1) + j]) return out;
    int l2 = 1, l = 2;
    while (l <= N) {
        if (i % l != 0 || j % l != 0) break;
        l2 *= 2;
        l = 2 * l2;
    }
    int i1, j1, i2, j2;
    if (i % l == 0 && j % l == l2) {
        i1 = i; j1 = j - l2; i2 = i; j2 = j + l2;
    } else if (i % l == l2 && j % l == 0) {
        i1 = i - l2; j1 = j; i2 = i + l2; j2 = j;
    } else {
        i1 = i - l2; j1 = j + l2; i2 = i + l2; j2 = j - l2;
    }
    const double* p1 = node_rec(mm, i1, j1);
    const double* p2 = node_rec(mm, i2, j2);
    double m0 = p1[0] + p2[0], m1 = p1[1] + p2[1], m2 = p1[2] + p2[2];
    double s = 0.0;
    s += m0 * m0;
    s += m1 * m1;
    s += m2 * m2;
    const double mag = sqrt(s);
    if (mag > 1e-15) { m0 /= mag; m1 /= mag; m2 /= mag; }
    out[0] = m0; out[1] = m1; out[2] = m2;
    mm->done[(int64_t)i * (N + 1) + j] = 1;
    return out;
}

/* Lateral unit-sphere coordinates of all 10 diamonds, one subdomain per
 * diamond: spherical_shell.hpp:419-587 (corner choice :528-585) and
 * :590-594 (element ranges). */
int orc_lateral_coords(int lat_level, double* coords)
{
    double node[12][3];
    int dnode[10][4];
    icosahedron(node, dnode);
    const int N = 1 << lat_level, n = N + 1;
    memo_t mm;
    mm.N = N;
    mm.val = (double*)malloc(sizeof(double) * 3 * n * n);
    mm.done = (uint8_t*)malloc((size_t)n * n);
    for (int d = 0; d < 10; d++) {
        memset(mm.done, 0, (size_t)n * n);
        const int L = d < 5 ? 1 : 3, R = d < 5 ? 3 : 1;
        const double* p00 = node[dnode[d][0]];
        const double* pN0 = node[dnode[d][L]];
        const double* pNN = node[dnode[d][2]];
        const double* p0N = node[dnode[d][R]];
        memcpy(mm.val + 3 * (0), p00, 3 * sizeof(double));
        memcpy(mm.val + 3 * ((int64_t)0 * n + N), p0N, 3 * sizeof(double));
        memcpy(mm.val + 3 * ((int64_t)N * n + 0), pN0, 3 * sizeof(double));
        memcpy(mm.val + 3 * ((int64_t)N * n + N), pNN, 3 * sizeof(double));
        mm.done[0] = mm.done[N] = mm.done[(int64_t)N * n] = mm.done[(int64_t)N * n + N] = 1;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double* p = node_rec(&mm, i, j);
                double* o = coords + 3 * (((int64_t)d * n + i) * n + j);
                o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
            }
    }
    free(mm.val);
    free(mm.done);
    return 0;
}

/* uniform_shell_radii: spherical_shell.hpp:27-52 */
void orc_uniform_radii(int nr, double r_min, double r_max, double* radii)
{
    const double step = (r_max - r_min) / (nr - 1);
    for (int i = 0; i < nr; i++) radii[i] = r_min + i * step;
    radii[nr - 1] = r_max;
}

typedef struct { double c[3]; int32_t idx; } keyed_t;

static int cmp_keyed(const void* a, const void* b)
{
    const keyed_t* p = (const keyed_t*)a;
    const keyed_t* q = (const keyed_t*)b;
    for (int k = 0; k < 3; k++) {
        if (p->c[k] < q->c[k]) return -1;
        if (p->c[k] > q->c[k]) return 1;
    }
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

/* Duplicated interface nodes.  The reference derives them from its subdomain
 * neighbourhood tables (spherical_shell.hpp:1050-2504); the oracle derives
 * them independently, from bitwise-equal coordinates (shared diamond edges
 * are generated from the same parents, so their coordinates are identical). */
int orc_lateral_groups(int n, const double* coords, int32_t* group_of)
{
    const int64_t cnt = 10LL * n * n;
    keyed_t* ks = (keyed_t*)malloc(sizeof(keyed_t) * cnt);
    for (int64_t i = 0; i < cnt; i++) {
        memcpy(ks[i].c, coords + 3 * i, 3 * sizeof(double));
        ks[i].idx = (int32_t)i;
    }
    qsort(ks, cnt, sizeof(keyed_t), cmp_keyed);
    int g = -1;
    for (int64_t i = 0; i < cnt; i++) {
        if (i == 0 || memcmp(ks[i].c, ks[i - 1].c, 3 * sizeof(double)) != 0) g++;
        group_of[ks[i].idx] = g;
    }
    free(ks);
    return g + 1;
}

/* NodeOwnershipFlag: owner = smallest global subdomain id among copies
 * (MIN reduction), grid/bit_masks.hpp:29-79; with one subdomain per diamond
 * the global id orders like the diamond id (spherical_shell.hpp:658-672). */
void orc_ownership(int n, int nr, const int32_t* group_of, int ngroups, uint8_t* owned)
{
    int* minD = (int*)malloc(sizeof(int) * ngroups);
    for (int g = 0; g < ngroups; g++) minD[g] = 99;
    for (int d = 0; d < 10; d++)
        for (int i = 0; i < n * n; i++) {
            const int g = group_of[(int64_t)d * n * n + i];
            if (d < minD[g]) minD[g] = d;
        }
    for (int d = 0; d < 10; d++)
        for (int i = 0; i < n * n; i++) {
            const int g = group_of[(int64_t)d * n * n + i];
            for (int r = 0; r < nr; r++)
                owned[((int64_t)d * n * n + i) * nr + r] = (minD[g] == d);
        }
    free(minD);
}

/* Additive interface reduction: every copy of a duplicated node ends up with
 * the sum over all copies, communication/shell/communication.hpp:241-739
 * (SUM semantics, same-rank branch). */
void orc_interface_sum(int n, int nr, int ncomp, const int32_t* group_of, int ngroups, double* v)
{
    const int64_t lat = 10LL * n * n, N = lat * nr;
    int* cnt = (int*)calloc(ngroups, sizeof(int));
    for (int64_t i = 0; i < lat; i++) cnt[group_of[i]]++;
    int* start = (int*)malloc(sizeof(int) * (ngroups + 1));
    start[0] = 0;
    for (int g = 0; g < ngroups; g++) start[g + 1] = start[g] + cnt[g];
    int* mem = (int*)malloc(sizeof(int) * lat);
    int* fill = (int*)calloc(ngroups, sizeof(int));
    for (int64_t i = 0; i < lat; i++) {
        const int g = group_of[i];
        mem[start[g] + fill[g]++] = (int)i;
    }
    for (int g = 0; g < ngroups; g++) {
        if (cnt[g] < 2) continue;
        for (int c = 0; c < ncomp; c++)
            for (int r = 0; r < nr; r++) {
                double s = 0.0;
                for (int k = start[g]; k < start[g + 1]; k++) s += v[c * N + (int64_t)mem[k] * nr + r];
                for (int k = start[g]; k < start[g + 1]; k++) v[c * N + (int64_t)mem[k] * nr + r] = s;
            }
    }
    free(cnt); free(start); free(mem); free(fill);
}

orc_mesh* orc_mesh_create(int lat_level, int nr, const double* radii)
{
    orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
    m->n = (1 << lat_level) + 1;
    m->nr = nr;
    m->coords = (double*)malloc(sizeof(double) * 30 * m->n * m->n);
    orc_lateral_coords(lat_level, m->coords);
    m->radii = (double*)malloc(sizeof(double) * nr);
    memcpy(m->radii, radii, sizeof(double) * nr);
    m->group_of = (int32_t*)malloc(sizeof(int32_t) * 10 * m->n * m->n);
    m->ngroups = orc_lateral_groups(m->n, m->coords, m->group_of);
    m->owned = (uint8_t*)malloc(orc_mesh_nodes(m));
    orc_ownership(m->n, nr, m->group_of, m->ngroups, m->owned);
    return m;
}

void orc_mesh_destroy(orc_mesh* m)
{
    if (!m) return;
    free(m->coords); free(m->radii); free(m->group_of); free(m->owned); free(m);
}

int64_t orc_mesh_nodes(const orc_mesh* m) { return 10LL * m->n * m->n * m->nr; }

double orc_dot_owned(const orc_mesh* m, int ncomp, const double* a, const double* b)
{
    /* masked dot over OWNED nodes, kernels/common/grid_operations.hpp:643-684 */
    const int64_t N = orc_mesh_nodes(m);
    double s = 0.0;
    for (int c = 0; c < ncomp; c++) {
        double sc = 0.0;
        for (int64_t i = 0; i < N; i++)
            if (m->owned[i]) sc += a[c * N + i] * b[c * N + i];
        s += sc;
    }
    return s;
}

/* ====================================================================== wedge FE core */

/* wedge local node offsets (dx,dy,dr): kernel_helpers.hpp:302-327 */
static const int WOFF[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                  {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};

/* shape functions: fe/wedge/integrands.hpp:17-101 */
static double shape_lat(int k, double xi, double eta)
{
    const double v[3] = {1.0 - xi - eta, xi, eta};
    return v[k % 3];
}
static double shape_rad(int k, double zeta)
{
    const double v[2] = {0.5 * (1 - zeta), 0.5 * (1 + zeta)};
    return v[k / 3];
}
static double shape(int k, const double q[3]) { return shape_lat(k, q[0], q[1]) * shape_rad(k, q[2]); }
static void grad_shape(int k, const double q[3], double g[3])
{
    static const double dxi[3] = {-1.0, 1.0, 0.0}, deta[3] = {-1.0, 0.0, 1.0}, drad[2] = {-0.5, 0.5};
    g[0] = dxi[k % 3] * shape_rad(k, q[2]);
    g[1] = deta[k % 3] * shape_rad(k, q[2]);
    g[2] = shape_lat(k, q[0], q[1]) * drad[k / 3];
}

/* coarse-on-fine shapes: integrands.hpp:103-243 */
static double shape_rad_coarse(int cn, int frw, double zeta)
{
    if (cn / 3 == 0) return frw == 0 ? 0.25 * (3 - zeta) : 0.25 * (1 - zeta);
    return frw == 0 ? 0.25 * (1 + zeta) : 0.25 * (3 + zeta);
}
static double shape_lat_coarse(int cn, int flw, double xi, double eta)
{
    switch (cn % 3) {
    case 0:
        switch (flw) {
        case 0: return -0.5 * eta - 0.5 * xi + 1.0;
        case 1: return -0.5 * eta - 0.5 * xi + 0.5;
        case 2: return -0.5 * eta - 0.5 * xi + 0.5;
        default: return 0.5 * eta + 0.5 * xi;
        }
    case 1:
        switch (flw) {
        case 0: return 0.5 * xi;
        case 1: return 0.5 * xi + 0.5;
        case 2: return 0.5 * xi;
        default: return 0.5 - 0.5 * xi;
        }
    default:
        switch (flw) {
        case 0: return 0.5 * eta;
        case 1: return 0.5 * eta;
        case 2: return 0.5 * eta + 0.5;
        default: return 0.5 - 0.5 * eta;
        }
    }
}
static double shape_coarse(int cn, int frw, int flw, const double q[3])
{
    return shape_lat_coarse(cn, flw, q[0], q[1]) * shape_rad_coarse(cn, frw, q[2]);
}

/* fine_lateral_wedge_idx: kernel_helpers.hpp:600-608 */
static int fine_lateral_wedge_idx(int x, int y, int w)
{
    static const int ind[2][2][2] = {{{0, 1}, {2, 3}}, {{3, 2}, {1, 0}}};
    return ind[w][y % 2][x % 2];
}

/* Felippa 3x2 rule: quadrature/quadrature.hpp:67-110 */
static const double Q6P[6][3] = {{0.6666666666666666, 0.1666666666666667, -0.5773502691896257},
                                 {0.1666666666666667, 0.6666666666666666, -0.5773502691896257},
                                 {0.1666666666666667, 0.1666666666666667, -0.5773502691896257},
                                 {0.6666666666666666, 0.1666666666666667, 0.5773502691896257},
                                 {0.1666666666666667, 0.6666666666666666, 0.5773502691896257},
                                 {0.1666666666666667, 0.1666666666666667, 0.5773502691896257}};
static const double Q6W = 0.1666666666666667;

/* wedge_surface_physical_coords: kernel_helpers.hpp:26-55 */
static void wedge_surf(const orc_mesh* m, int d, int x, int y, double P[2][3][3])
{
    const int n = m->n;
    const double* c00 = m->coords + 3 * (((int64_t)d * n + x) * n + y);
    const double* c10 = m->coords + 3 * (((int64_t)d * n + x + 1) * n + y);
    const double* c01 = m->coords + 3 * (((int64_t)d * n + x) * n + y + 1);
    const double* c11 = m->coords + 3 * (((int64_t)d * n + x + 1) * n + y + 1);
    for (int a = 0; a < 3; a++) {
        P[0][0][a] = c00[a]; P[0][1][a] = c10[a]; P[0][2][a] = c01[a];
        P[1][0][a] = c11[a]; P[1][1][a] = c01[a]; P[1][2][a] = c10[a];
    }
}

/* jac = jac_lat * diag(r, r, dr/dzeta): integrands.hpp:623-670 */
static void jac(const double P[3][3], double r1, double r2, const double q[3], double J[3][3])
{
    const double r = r1 + 0.5 * (r2 - r1) * (1 + q[2]);
    const double gr = 0.5 * (r2 - r1);
    for (int a = 0; a < 3; a++) {
        const double c0 = P[1][a] - P[0][a];
        const double c1 = P[2][a] - P[0][a];
        const double c2 = (1 - q[0] - q[1]) * P[0][a] + q[0] * P[1][a] + q[1] * P[2][a];
        J[a][0] = c0 * r;
        J[a][1] = c1 * r;
        J[a][2] = c2 * gr;
    }
}

static double det3(const double J[3][3])
{
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

/* inverse transposed: K[a][b] = (J^-1)[b][a] */
static void inv_t3(const double J[3][3], double det, double K[3][3])
{
    K[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    K[0][1] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    K[0][2] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    K[1][0] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    K[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    K[1][2] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    K[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    K[2][1] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    K[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
}

static void matvec3(const double K[3][3], const double v[3], double o[3])
{
    for (int a = 0; a < 3; a++) o[a] = K[a][0] * v[0] + K[a][1] * v[1] + K[a][2] * v[2];
}

/* trafo_mat_cartesian_to_normal_tangential:
 * linalg/trafo/local_basis_trafo_normal_tangential.hpp:34-79 */
static void normalize3(double v[3])
{
    double s = 0.0;
    s += v[0] * v[0];
    s += v[1] * v[1];
    s += v[2] * v[2];
    const double mag = sqrt(s);
    if (mag > 1e-15) { v[0] /= mag; v[1] /= mag; v[2] /= mag; }
}
static void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void trafo_nt(const double nin[3], double R[3][3])
{
    double nn[3] = {nin[0], nin[1], nin[2]};
    normalize3(nn);
    const double ax = fabs(nn[0]), ay = fabs(nn[1]), az = fabs(nn[2]);
    double ref[3] = {0, 0, 0};
    if (ax <= ay && ax <= az) ref[0] = 1;
    else if (ay <= ax && ay <= az) ref[1] = 1;
    else ref[2] = 1;
    double t1[3], t2[3];
    cross3(ref, nn, t1);
    normalize3(t1);
    if (sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]) < 1e-15) {
        t1[0] = nn[1]; t1[1] = -nn[0]; t1[2] = 0;
        normalize3(t1);
    }
    cross3(nn, t1, t2);
    normalize3(t2);
    for (int a = 0; a < 3; a++) { R[0][a] = nn[a]; R[1][a] = t1[a]; R[2][a] = t2[a]; }
}

/* ====================================================================== epsilon-div-div */

/* assemble_local_matrix + assemble_trial_test_vecs:
 * epsilon_divdiv_kerngen.hpp:2016-2108; Felippa 1x1 rule quadrature.hpp:32-46
 * (point (1/3,1/3,0), weight 1); symmetric_grad integrands.hpp:672-683. */
void orc_eps_local_matrix(const orc_mesh* m, const double* k, int d, int x, int y, int r, int wedge, double* A)
{
    const int n = m->n, nr = m->nr;
    const double q[3] = {1.0 / 3.0, 1.0 / 3.0, 0.0};
    double keval = 0.0;
    for (int j = 0; j < 6; j++)
        keval += shape(j, q) * k[IDX4(n, nr, d, x + WOFF[wedge][j][0], y + WOFF[wedge][j][1], r + WOFF[wedge][j][2])];
    orc_eps_local_matrix_k(m, keval, d, x, y, r, wedge, A);
}

/* assemble_local_matrix, epsilon_divdiv_kerngen.hpp:2051-2108, with the
 * quadrature-point viscosity keval given */
void orc_eps_local_matrix_k(const orc_mesh* m, double keval, int d, int x, int y, int r, int wedge, double* A)
{
    double P[2][3][3];
    wedge_surf(m, d, x, y, P);
    const double q[3] = {1.0 / 3.0, 1.0 / 3.0, 0.0};
    const double r1 = m->radii[r], r2 = m->radii[r + 1];
    double J[3][3], K[3][3];
    jac(P[wedge], r1, r2, q, J);
    const double det = det3(J);
    inv_t3(J, det, K);
    const double w = 1.0 * keval * fabs(det);
    double g[6][3];
    for (int j = 0; j < 6; j++) {
        double gr[3];
        grad_shape(j, q, gr);
        matvec3(K, gr, g[j]);
    }
    memset(A, 0, sizeof(double) * 324);
    for (int di = 0; di < 3; di++)
        for (int dj = 0; dj < 3; dj++)
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    /* S(g, dim) = (g e_dim^T + e_dim g^T)/2 */
                    double Si[3][3] = {{0}}, Sj[3][3] = {{0}};
                    for (int a = 0; a < 3; a++) {
                        Si[a][di] += 0.5 * g[i][a];
                        Si[di][a] += 0.5 * g[i][a];
                        Sj[a][dj] += 0.5 * g[j][a];
                        Sj[dj][a] += 0.5 * g[j][a];
                    }
                    double dc = 0.0;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) dc += Sj[a][b] * Si[a][b];
                    A[(i + di * 6) * 18 + j + dj * 6] += w * (2 * dc - 2.0 / 3.0 * Sj[dj][dj] * Si[di][di]);
                }
}

static int bnd_cmb(const orc_mesh* m, int r) { (void)m; return r == 0; }
static int bnd_surf(const orc_mesh* m, int r) { return r == m->nr - 1; }

/* fs/fs rotation null-space modes: epsilon_divdiv_kerngen.hpp:287-408.
 * The reference's free-slip projection of the modes tests has_flag(mask,
 * CMB|SURFACE) (util/bit_masking.hpp:43-53 requires all bits), which no node
 * of a multi-layer shell carries, so the modes stay unprojected.  Gram-Schmidt
 * uses OWNED-masked dots. */
static void penalty_init(orc_eps* op)
{
    const orc_mesh* m = op->m;
    const int n = m->n, nr = m->nr;
    const int64_t N = orc_mesh_nodes(m);
    op->null_modes = (double*)calloc(9 * N, sizeof(double));
    for (int axis = 0; axis < 3; axis++) {
        double* nm = op->null_modes + axis * 3 * N;
        const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
        for (int d = 0; d < 10; d++)
            for (int x = 0; x < n; x++)
                for (int y = 0; y < n; y++)
                    for (int r = 0; r < nr; r++) {
                        const int64_t i = IDX4(n, nr, d, x, y, r);
                        const double* c = m->coords + 3 * (((int64_t)d * n + x) * n + y);
                        double cc[3] = {c[0] * m->radii[r], c[1] * m->radii[r], c[2] * m->radii[r]};
                        nm[a1 * N + i] = cc[a2];
                        nm[a2 * N + i] = -cc[a1];
                    }
    }
    for (int i = 0; i < 3; i++) {
        double* ni = op->null_modes + i * 3 * N;
        for (int j = 0; j < i; j++) {
            const double* nj = op->null_modes + j * 3 * N;
            const double dij = orc_dot_owned(m, 3, ni, nj);
            for (int64_t t = 0; t < 3 * N; t++) ni[t] -= dij * nj[t];
        }
        const double inv = 1.0 / sqrt(orc_dot_owned(m, 3, ni, ni));
        for (int64_t t = 0; t < 3 * N; t++) ni[t] *= inv;
    }
}

orc_eps* orc_eps_create(const orc_mesh* m, const double* k, int bc_cmb, int bc_surf)
{
    orc_eps* op = (orc_eps*)calloc(1, sizeof(orc_eps));
    op->m = m;
    op->k = k;
    op->bc_cmb = bc_cmb;
    op->bc_surf = bc_surf;
    op->penalty_eps = 1.0;
    if (bc_cmb == ORC_FREESLIP && bc_surf == ORC_FREESLIP) penalty_init(op);
    return op;
}

void orc_eps_destroy(orc_eps* op)
{
    if (!op) return;
    free(op->stored);
    free(op->null_modes);
    free(op);
}

/* Not in the reference: its penalty is added at every node copy before the
 * interface sum (F5), so an interface node receives it once per copy and the
 * operator is not symmetric.  once = 1 adds it at the owned copy only. */
void orc_eps_set_penalty_once(orc_eps* op, int once) { op->penalty_once = once ? 1 : 0; }

void orc_eps_set_penalty_epsilon(orc_eps* op, double eps) { op->penalty_eps = eps; }

void orc_eps_set_stored(orc_eps* op, const double* mats)
{
    const orc_mesh* m = op->m;
    const int64_t cells = 10LL * (m->n - 1) * (m->n - 1) * (m->nr - 1);
    free(op->stored);
    op->stored = (double*)malloc(sizeof(double) * cells * 2 * 324);
    memcpy(op->stored, mats, sizeof(double) * cells * 2 * 324);
}

int orc_eps_penalty_modes(const orc_eps* op, double* out)
{
    if (!op->null_modes) return 0;
    memcpy(out, op->null_modes, sizeof(double) * 9 * orc_mesh_nodes(op->m));
    return 1;
}

static void get_local(const orc_eps* op, int d, int x, int y, int r, int w, double* A)
{
    const orc_mesh* m = op->m;
    if (op->stored) {
        const int64_t c = (((int64_t)d * (m->n - 1) + x) * (m->n - 1) + y) * (m->nr - 1) + r;
        memcpy(A, op->stored + (c * 2 + w) * 324, sizeof(double) * 324);
    } else if (op->kc) {
        const int64_t c = (((int64_t)d * (m->n - 1) + x) * (m->n - 1) + y) * (m->nr - 1) + r;
        orc_eps_local_matrix_k(m, op->kc[c], d, x, y, r, w, A);
    } else {
        orc_eps_local_matrix(m, op->k, d, x, y, r, w, A);
    }
}

/* Effective wedge matrices of one hex cell acting on Cartesian, dimension-wise
 * ordered dofs (d*6 + node): operator_slow_path, epsilon_divdiv_kerngen.hpp:823-1051
 * (Dirichlet mask :899-918, free-slip rotation :919-996, Hadamard :999-1002,
 * diagonal :1004-1008, back rotation :1014-1026).  Free-slip cells return
 * R^T (M o (R A R^T)) R re-ordered to dimension-wise dofs, which is the map
 * the reference applies to the Cartesian source. */
void orc_eps_cell_effective(const orc_eps* op, int d, int xc, int yc, int rc, int diagonal, double* E /*[2][324]*/)
{
    const orc_mesh* m = op->m;
    const int n = m->n;
    double A[2][324], M[324], R[2][324], Atmp[324];
    get_local(op, d, xc, yc, rc, 0, A[0]);
    get_local(op, d, xc, yc, rc, 1, A[1]);
    for (int t = 0; t < 324; t++) M[t] = 1.0;
    const int at_cmb = bnd_cmb(m, rc), at_surf = bnd_surf(m, rc + 1);
    int fs = 0;
    if (at_cmb || at_surf) {
        const int bcf = at_cmb ? op->bc_cmb : op->bc_surf;
        if (bcf == ORC_DIRICHLET) {
            for (int di = 0; di < 3; di++)
                for (int dj = 0; dj < 3; dj++)
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) {
                            const int hit = (at_cmb && ((di == dj && i != j && (i < 3 || j < 3)) ||
                                                        (di != dj && (i < 3 || j < 3)))) ||
                                            (at_surf && ((di == dj && i != j && (i >= 3 || j >= 3)) ||
                                                         (di != dj && (i >= 3 || j >= 3))));
                            if (hit) M[(i + di * 6) * 18 + j + dj * 6] = 0.0;
                        }
        } else if (bcf == ORC_FREESLIP) {
            fs = 1;
            for (int w = 0; w < 2; w++) {
                for (int ni = 0; ni < 6; ni++) /* to node-wise order */
                    for (int di = 0; di < 3; di++)
                        for (int nj = 0; nj < 6; nj++)
                            for (int dj = 0; dj < 3; dj++)
                                Atmp[(ni * 3 + di) * 18 + nj * 3 + dj] = A[w][(ni + di * 6) * 18 + nj + dj * 6];
                memset(R[w], 0, sizeof(double) * 324);
                for (int i = 0; i < 18; i++) R[w][i * 18 + i] = 1.0;
                for (int b = 0; b < 3; b++) {
                    const int gx = xc + WOFF[w][b][0], gy = yc + WOFF[w][b][1];
                    const int gr = rc + (at_cmb ? 0 : 1);
                    const double* c = m->coords + 3 * (((int64_t)d * n + gx) * n + gy);
                    const double nv[3] = {c[0] * m->radii[gr], c[1] * m->radii[gr], c[2] * m->radii[gr]};
                    double Ri[3][3];
                    trafo_nt(nv, Ri);
                    const int off = at_cmb ? 0 : 9;
                    for (int di = 0; di < 3; di++)
                        for (int dj = 0; dj < 3; dj++) R[w][(off + b * 3 + di) * 18 + off + b * 3 + dj] = Ri[di][dj];
                }
                double T[324]; /* A = R Atmp R^T */
                for (int i = 0; i < 18; i++)
                    for (int j = 0; j < 18; j++) {
                        double s = 0;
                        for (int l = 0; l < 18; l++) s += R[w][i * 18 + l] * Atmp[l * 18 + j];
                        T[i * 18 + j] = s;
                    }
                for (int i = 0; i < 18; i++)
                    for (int j = 0; j < 18; j++) {
                        double s = 0;
                        for (int l = 0; l < 18; l++) s += T[i * 18 + l] * R[w][j * 18 + l];
                        A[w][i * 18 + j] = s;
                    }
                const int nb = at_surf ? 3 : 0, ne = at_surf ? 6 : 3;
                for (int node = nb; node < ne; node++) {
                    const int idx = node * 3;
                    for (int t = 0; t < 18; t++)
                        if (t != idx) { M[idx * 18 + t] = 0.0; M[t * 18 + idx] = 0.0; }
                }
            }
        }
    }
    for (int w = 0; w < 2; w++) {
        for (int t = 0; t < 324; t++) A[w][t] *= M[t];
        if (diagonal)
            for (int i = 0; i < 18; i++)
                for (int j = 0; j < 18; j++)
                    if (i != j) A[w][i * 18 + j] = 0.0;
        double* Ew = E + w * 324;
        if (!fs) {
            memcpy(Ew, A[w], sizeof(double) * 324);
            continue;
        }
        double T[324], F[324]; /* F = R^T A R (node-wise), then dimension-wise */
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) {
                double s = 0;
                for (int l = 0; l < 18; l++) s += A[w][i * 18 + l] * R[w][l * 18 + j];
                T[i * 18 + j] = s;
            }
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) {
                double s = 0;
                for (int l = 0; l < 18; l++) s += R[w][l * 18 + i] * T[l * 18 + j];
                F[i * 18 + j] = s;
            }
        for (int ni = 0; ni < 6; ni++)
            for (int di = 0; di < 3; di++)
                for (int nj = 0; nj < 6; nj++)
                    for (int dj = 0; dj < 3; dj++) Ew[(ni + di * 6) * 18 + nj + dj * 6] = F[(ni * 3 + di) * 18 + nj * 3 + dj];
    }
}

/* apply = sum over cells of the effective local maps; penalty
 * epsilon_divdiv_kerngen.hpp:583-607 (added before the additive
 * communication, F5); communication :609-613. */
void orc_eps_apply(const orc_eps* op, const double* x, double* yv, int diagonal, int comm)
{
    const orc_mesh* m = op->m;
    const int n = m->n, nr = m->nr;
    const int64_t N = orc_mesh_nodes(m);
    memset(yv, 0, sizeof(double) * 3 * N);
    /* diamonds write disjoint arrays before the interface sum */
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < 10; d++) {
        double E[2 * 324];
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++)
                for (int rc = 0; rc < nr - 1; rc++) {
                    orc_eps_cell_effective(op, d, xc, yc, rc, diagonal, E);
                    for (int w = 0; w < 2; w++) {
                        double src[18];
                        int64_t gid[6];
                        for (int j = 0; j < 6; j++) {
                            gid[j] = IDX4(n, nr, d, xc + WOFF[w][j][0], yc + WOFF[w][j][1], rc + WOFF[w][j][2]);
                            for (int c = 0; c < 3; c++) src[c * 6 + j] = x[c * N + gid[j]];
                        }
                        for (int i = 0; i < 18; i++) {
                            double s = 0;
                            for (int j = 0; j < 18; j++) s += E[w * 324 + i * 18 + j] * src[j];
                            yv[(i / 6) * N + gid[i % 6]] += s;
                        }
                    }
                }
    }
    if (op->null_modes && !diagonal) {
        for (int i = 0; i < 3; i++) {
            const double* nm = op->null_modes + i * 3 * N;
            const double alpha = op->penalty_eps * orc_dot_owned(m, 3, nm, x);
            for (int64_t t = 0; t < 3 * N; t++)
                if (!op->penalty_once || m->owned[t % N]) yv[t] += alpha * nm[t];
        }
    }
    if (comm) orc_interface_sum(n, nr, 3, m->group_of, m->ngroups, yv);
}

/* Global sparse assembly (COO over unique dofs u = 3*(group*nr + r) + comp)
 * of the operator without the penalty; used by tests to solve small systems
 * directly.  Returns the number of triplets written (cells*2*324). */
int64_t orc_eps_assemble_coo(const orc_eps* op, int64_t* rows, int64_t* cols, double* vals)
{
    const orc_mesh* m = op->m;
    const int n = m->n, nr = m->nr;
    const int64_t per_d = (int64_t)(n - 1) * (n - 1) * (nr - 1) * 2 * 324;
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < 10; d++) {
        double E[2 * 324];
        int64_t k = per_d * d;
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++)
                for (int rc = 0; rc < nr - 1; rc++) {
                    orc_eps_cell_effective(op, d, xc, yc, rc, 0, E);
                    for (int w = 0; w < 2; w++) {
                        int64_t uid[6];
                        for (int j = 0; j < 6; j++)
                            uid[j] = (int64_t)m->group_of[((int64_t)d * n + xc + WOFF[w][j][0]) * n + yc + WOFF[w][j][1]] * nr +
                                     rc + WOFF[w][j][2];
                        for (int i = 0; i < 18; i++)
                            for (int j = 0; j < 18; j++, k++) {
                                rows[k] = 3 * uid[i % 6] + i / 6;
                                cols[k] = 3 * uid[j % 6] + j / 6;
                                vals[k] = E[w * 324 + i * 18 + j];
                            }
                    }
                }
    }
    return per_d * 10;
}

/* ====================================================================== divergence / gradient */

/* local 6x18 divergence matrix of fine wedge w with BC handling:
 * divergence.hpp:117-323 (quadrature :139-166, Dirichlet :199-213,
 * free-slip :214-295).  Returns (A6x18 acting on the possibly rotated
 * src, R used for the rotation, fs flag). */
static void div_local(const orc_mesh* mf, int bc_cmb, int bc_surf, int d, int xc, int yc, int rc, double A[2][108],
                      double R[2][324], int* fs_out)
{
    const int n = mf->n;
    double P[2][3][3];
    wedge_surf(mf, d, xc, yc, P);
    const double r1 = mf->radii[rc], r2 = mf->radii[rc + 1];
    const int frw = rc % 2;
    memset(A, 0, sizeof(double) * 216);
    for (int q = 0; q < 6; q++)
        for (int w = 0; w < 2; w++) {
            const int flw = fine_lateral_wedge_idx(xc, yc, w);
            double J[3][3], K[3][3];
            jac(P[w], r1, r2, Q6P[q], J);
            const double dt = det3(J);
            const double adet = fabs(dt);
            inv_t3(J, dt, K);
            for (int i = 0; i < 6; i++) {
                const double si = shape_coarse(i, frw, flw, Q6P[q]);
                for (int j = 0; j < 6; j++) {
                    double gr[3], g[3];
                    grad_shape(j, Q6P[q], gr);
                    matvec3(K, gr, g);
                    for (int c = 0; c < 3; c++) A[w][i * 18 + c * 6 + j] += Q6W * (-g[c] * si * adet);
                }
            }
        }
    const int at_cmb = bnd_cmb(mf, rc), at_surf = bnd_surf(mf, rc + 1);
    double M[108];
    for (int t = 0; t < 108; t++) M[t] = 1.0;
    *fs_out = 0;
    if (at_cmb || at_surf) {
        const int bcf = at_cmb ? bc_cmb : bc_surf;
        if (bcf == ORC_DIRICHLET) {
            for (int c = 0; c < 3; c++)
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++)
                        if ((at_cmb && j < 3) || (at_surf && j >= 3)) M[i * 18 + c * 6 + j] = 0.0;
        } else if (bcf == ORC_FREESLIP) {
            *fs_out = 1;
            for (int w = 0; w < 2; w++) {
                double T[108];
                for (int i = 0; i < 6; i++)
                    for (int c = 0; c < 3; c++)
                        for (int j = 0; j < 6; j++) T[i * 18 + j * 3 + c] = A[w][i * 18 + j + c * 6];
                memset(R[w], 0, sizeof(double) * 324);
                for (int i = 0; i < 18; i++) R[w][i * 18 + i] = 1.0;
                for (int b = 0; b < 3; b++) {
                    const int gx = xc + WOFF[w][b][0], gy = yc + WOFF[w][b][1];
                    const int gr = rc + (at_cmb ? 0 : 1);
                    const double* cc = mf->coords + 3 * (((int64_t)d * n + gx) * n + gy);
                    const double nv[3] = {cc[0] * mf->radii[gr], cc[1] * mf->radii[gr], cc[2] * mf->radii[gr]};
                    double Ri[3][3];
                    trafo_nt(nv, Ri);
                    const int off = at_cmb ? 0 : 9;
                    for (int a = 0; a < 3; a++)
                        for (int bb = 0; bb < 3; bb++) R[w][(off + b * 3 + a) * 18 + off + b * 3 + bb] = Ri[a][bb];
                }
                /* A = T R^T */
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 18; j++) {
                        double s = 0;
                        for (int l = 0; l < 18; l++) s += T[i * 18 + l] * R[w][j * 18 + l];
                        A[w][i * 18 + j] = s;
                    }
                const int nb = at_surf ? 3 : 0, ne = at_surf ? 6 : 3;
                for (int node = nb; node < ne; node++)
                    for (int kk = 0; kk < 6; kk++) M[kk * 18 + node * 3] = 0.0;
            }
        }
    }
    for (int w = 0; w < 2; w++)
        for (int t = 0; t < 108; t++) A[w][t] *= M[t];
}

/* Effective 6x18 divergence blocks acting on Cartesian dimension-wise
 * velocity dofs (free-slip: (M o (A R^T)) R, re-ordered). */
void orc_div_cell_effective(const orc_mesh* mf, int bc_cmb, int bc_surf, int d, int xc, int yc, int rc,
                            double* E /*[2][108]*/)
{
    double A[2][108], R[2][324];
    int fs;
    div_local(mf, bc_cmb, bc_surf, d, xc, yc, rc, A, R, &fs);
    for (int w = 0; w < 2; w++) {
        double* Ew = E + w * 108;
        if (!fs) {
            memcpy(Ew, A[w], sizeof(double) * 108);
            continue;
        }
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 18; j++) { /* column j node-wise */
                double s = 0;
                for (int l = 0; l < 18; l++) s += A[w][i * 18 + l] * R[w][l * 18 + j];
                Ew[i * 18 + (j % 3) * 6 + j / 3] = s;
            }
    }
}

/* coarse wedge nodes for the scatter of fine wedge w into coarse cell
 * (x/2,y/2,r/2): atomically_add_local_wedge_scalar_coefficients,
 * kernel_helpers.hpp:396-424 -- keyed by the FINE wedge index (F4). */
static int64_t coarse_gid(const orc_mesh* mc, int d, int xc, int yc, int rc, int w, int i)
{
    return IDX4(mc->n, mc->nr, d, xc / 2 + WOFF[w][i][0], yc / 2 + WOFF[w][i][1], rc / 2 + WOFF[w][i][2]);
}

void orc_div_apply(const orc_mesh* mf, const orc_mesh* mc, int bc_cmb, int bc_surf, const double* u, double* p,
                   int comm)
{
    const int n = mf->n, nr = mf->nr;
    const int64_t N = orc_mesh_nodes(mf), Nc = orc_mesh_nodes(mc);
    memset(p, 0, sizeof(double) * Nc);
    double A[2][108], R[2][324];
    for (int d = 0; d < 10; d++)
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++)
                for (int rc = 0; rc < nr - 1; rc++) {
                    int fs;
                    div_local(mf, bc_cmb, bc_surf, d, xc, yc, rc, A, R, &fs);
                    for (int w = 0; w < 2; w++) {
                        double src[18];
                        for (int j = 0; j < 6; j++)
                            for (int c = 0; c < 3; c++) {
                                const int64_t g = IDX4(n, nr, d, xc + WOFF[w][j][0], yc + WOFF[w][j][1], rc + WOFF[w][j][2]);
                                src[c * 6 + j] = u[c * N + g];
                            }
                        if (fs) {
                            double s2[18], s3[18];
                            for (int j = 0; j < 6; j++)
                                for (int c = 0; c < 3; c++) s2[j * 3 + c] = src[j + c * 6];
                            for (int i = 0; i < 18; i++) {
                                double s = 0;
                                for (int l = 0; l < 18; l++) s += R[w][i * 18 + l] * s2[l];
                                s3[i] = s;
                            }
                            memcpy(src, s3, sizeof(s3));
                        }
                        for (int i = 0; i < 6; i++) {
                            double s = 0;
                            for (int j = 0; j < 18; j++) s += A[w][i * 18 + j] * src[j];
                            p[coarse_gid(mc, d, xc, yc, rc, w, i)] += s;
                        }
                    }
                }
    if (comm) orc_interface_sum(mc->n, mc->nr, 1, mc->group_of, mc->ngroups, p);
}

/* Gradient = B^T with the same routing: gradient.hpp:115-325.  The
 * reference builds the 18x6 block and rotates rows; the oracle applies the
 * transpose of the masked divergence block, which is the same matrix. */
void orc_grad_apply_add(const orc_mesh* mf, const orc_mesh* mc, int bc_cmb, int bc_surf, const double* p, double* u,
                        int comm)
{
    const int n = mf->n, nr = mf->nr;
    const int64_t N = orc_mesh_nodes(mf);
    double* tmp = (double*)calloc(3 * N, sizeof(double));
    double A[2][108], R[2][324];
    for (int d = 0; d < 10; d++)
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++)
                for (int rc = 0; rc < nr - 1; rc++) {
                    int fs;
                    div_local(mf, bc_cmb, bc_surf, d, xc, yc, rc, A, R, &fs);
                    for (int w = 0; w < 2; w++) {
                        double pc[6], out[18];
                        for (int i = 0; i < 6; i++) pc[i] = p[coarse_gid(mc, d, xc, yc, rc, w, i)];
                        for (int j = 0; j < 18; j++) {
                            double s = 0;
                            for (int i = 0; i < 6; i++) s += A[w][i * 18 + j] * pc[i];
                            out[j] = s;
                        }
                        if (fs) { /* out is node-wise rotated: back-rotate, reorder */
                            double t2[18];
                            for (int i = 0; i < 18; i++) {
                                double s = 0;
                                for (int l = 0; l < 18; l++) s += R[w][l * 18 + i] * out[l];
                                t2[i] = s;
                            }
                            for (int j = 0; j < 6; j++)
                                for (int c = 0; c < 3; c++) out[j + c * 6] = t2[j * 3 + c];
                        }
                        for (int j = 0; j < 6; j++)
                            for (int c = 0; c < 3; c++) {
                                const int64_t g = IDX4(n, nr, d, xc + WOFF[w][j][0], yc + WOFF[w][j][1], rc + WOFF[w][j][2]);
                                tmp[c * N + g] += out[c * 6 + j];
                            }
                    }
                }
    for (int64_t t = 0; t < 3 * N; t++) u[t] += tmp[t];
    free(tmp);
    if (comm) orc_interface_sum(n, nr, 3, mf->group_of, mf->ngroups, u);
}

/* COO of B over unique dofs: rows = coarse pressure group*nrc + r,
 * cols = 3*(fine group*nr + r) + comp; scatter keyed by the fine wedge (F4). */
int64_t orc_div_assemble_coo(const orc_mesh* mf, const orc_mesh* mc, int bc_cmb, int bc_surf, int64_t* rows,
                             int64_t* cols, double* vals)
{
    const int n = mf->n, nr = mf->nr;
    const int64_t per_d = (int64_t)(n - 1) * (n - 1) * (nr - 1) * 2 * 108;
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < 10; d++) {
        double E[2 * 108];
        int64_t k = per_d * d;
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++)
                for (int rc = 0; rc < nr - 1; rc++) {
                    orc_div_cell_effective(mf, bc_cmb, bc_surf, d, xc, yc, rc, E);
                    for (int w = 0; w < 2; w++) {
                        int64_t uid[6], pid[6];
                        for (int j = 0; j < 6; j++) {
                            uid[j] = (int64_t)mf->group_of[((int64_t)d * n + xc + WOFF[w][j][0]) * n + yc + WOFF[w][j][1]] * nr +
                                     rc + WOFF[w][j][2];
                            pid[j] = (int64_t)mc->group_of[((int64_t)d * mc->n + xc / 2 + WOFF[w][j][0]) * mc->n + yc / 2 +
                                                           WOFF[w][j][1]] * mc->nr + rc / 2 + WOFF[w][j][2];
                        }
                        for (int i = 0; i < 6; i++)
                            for (int j = 0; j < 18; j++, k++) {
                                rows[k] = pid[i];
                                cols[k] = 3 * uid[j % 6] + j / 6;
                                vals[k] = E[w * 108 + i * 18 + j];
                            }
                    }
                }
    }
    return per_d * 10;
}

/* ====================================================================== mass operators */

/* lateral Jacobian determinant at quadrature points: kernel_helpers.hpp:153-176
 * with jac_lat integrands.hpp:623-637 */
static double det_jac_lat(const double P[3][3], const double q[3])
{
    double J[3][3];
    for (int a = 0; a < 3; a++) {
        J[a][0] = P[1][a] - P[0][a];
        J[a][1] = P[2][a] - P[0][a];
        J[a][2] = (1 - q[0] - q[1]) * P[0][a] + q[0] * P[1][a] + q[1] * P[2][a];
    }
    return fabs(det3(J));
}

/* VectorMass: vector_mass.hpp:83-153 */
void orc_vector_mass_apply(const orc_mesh* m, const double* x, double* yv)
{
    const int n = m->n, nr = m->nr;
    const int64_t N = orc_mesh_nodes(m);
    memset(yv, 0, sizeof(double) * 3 * N);
    for (int d = 0; d < 10; d++)
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++) {
                double P[2][3][3];
                wedge_surf(m, d, xc, yc, P);
                double djl[2][6];
                for (int w = 0; w < 2; w++)
                    for (int q = 0; q < 6; q++) djl[w][q] = det_jac_lat(P[w], Q6P[q]);
                for (int rc = 0; rc < nr - 1; rc++) {
                    const double r1 = m->radii[rc], r2 = m->radii[rc + 1], gr = 0.5 * (r2 - r1);
                    for (int w = 0; w < 2; w++) {
                        double A[36] = {0};
                        for (int q = 0; q < 6; q++) {
                            const double rr = r1 + 0.5 * (r2 - r1) * (1 + Q6P[q][2]);
                            for (int i = 0; i < 6; i++)
                                for (int j = 0; j < 6; j++)
                                    A[i * 6 + j] += Q6W * (shape(i, Q6P[q]) * shape(j, Q6P[q]) * rr * rr * gr * djl[w][q]);
                        }
                        for (int c = 0; c < 3; c++)
                            for (int i = 0; i < 6; i++) {
                                double s = 0;
                                for (int j = 0; j < 6; j++) {
                                    const int64_t g = IDX4(n, nr, d, xc + WOFF[w][j][0], yc + WOFF[w][j][1], rc + WOFF[w][j][2]);
                                    s += A[i * 6 + j] * x[c * N + g];
                                }
                                const int64_t gi = IDX4(n, nr, d, xc + WOFF[w][i][0], yc + WOFF[w][i][1], rc + WOFF[w][i][2]);
                                yv[c * N + gi] += s;
                            }
                    }
                }
            }
    orc_interface_sum(n, nr, 3, m->group_of, m->ngroups, yv);
}

/* KMass lumped diagonal: kmass.hpp:97-185 (k evaluated at quad points,
 * lumped = row sums) */
void orc_kmass_lumped(const orc_mesh* m, const double* k, double* dg)
{
    const int n = m->n, nr = m->nr;
    const int64_t N = orc_mesh_nodes(m);
    memset(dg, 0, sizeof(double) * N);
    for (int d = 0; d < 10; d++)
        for (int xc = 0; xc < n - 1; xc++)
            for (int yc = 0; yc < n - 1; yc++) {
                double P[2][3][3];
                wedge_surf(m, d, xc, yc, P);
                for (int rc = 0; rc < nr - 1; rc++) {
                    const double r1 = m->radii[rc], r2 = m->radii[rc + 1], gr = 0.5 * (r2 - r1);
                    for (int w = 0; w < 2; w++) {
                        double A[36] = {0};
                        int64_t g[6];
                        for (int j = 0; j < 6; j++)
                            g[j] = IDX4(n, nr, d, xc + WOFF[w][j][0], yc + WOFF[w][j][1], rc + WOFF[w][j][2]);
                        for (int q = 0; q < 6; q++) {
                            const double rr = r1 + 0.5 * (r2 - r1) * (1 + Q6P[q][2]);
                            double ke = 0.0;
                            for (int j = 0; j < 6; j++) ke += shape(j, Q6P[q]) * k[g[j]];
                            const double djl = det_jac_lat(P[w], Q6P[q]);
                            for (int i = 0; i < 6; i++)
                                for (int j = 0; j < 6; j++)
                                    A[i * 6 + j] +=
                                        Q6W * ke * (shape(i, Q6P[q]) * shape(j, Q6P[q]) * rr * rr * gr * djl);
                        }
                        for (int i = 0; i < 6; i++) {
                            double s = 0;
                            for (int j = 0; j < 6; j++) s += A[i * 6 + j];
                            dg[g[i]] += s;
                        }
                    }
                }
            }
    orc_interface_sum(n, nr, 1, m->group_of, m->ngroups, dg);
}

/* ====================================================================== transfers */

/* prolongation_constant_weight: fe/wedge/shell/grid_transfer_constant.hpp:25-89 */
static double pc_weight(int xf, int yf, int rf, int xc, int yc, int rc)
{
    const int dx = xf - 2 * xc, dy = yf - 2 * yc, dr = rf - 2 * rc;
    double lat = 0.0;
    if (dx == 0 && dy == 0) lat = 1.0;
    else if (dy == 0 && (dx == 1 || dx == -1)) lat = 0.5;
    else if (dx == 0 && (dy == 1 || dy == -1)) lat = 0.5;
    else if (dx == -1 && dy == 1) lat = 0.5;
    else if (dx == 1 && dy == -1) lat = 0.5;
    if (dr == 0) return lat;
    if (dr == 1 || dr == -1) return 0.5 * lat;
    return 0.0;
}

static const int STEN[7][2] = {{-1, 0}, {-1, 1}, {0, 1}, {1, 0}, {1, -1}, {0, -1}, {0, 0}};

/* ProlongationVecConstant (Add mode): prolongation_constant.hpp:103-193 */
void orc_prolongate_add(const orc_mesh* mc, const orc_mesh* mf, int ncomp, const double* xc, double* xf)
{
    const int64_t Nc = orc_mesh_nodes(mc), Nf = orc_mesh_nodes(mf);
    for (int d = 0; d < 10; d++)
        for (int x = 0; x < mc->n; x++)
            for (int y = 0; y < mc->n; y++)
                for (int r = 0; r < mc->nr; r++)
                    for (int dr = -1; dr <= 1; dr++)
                        for (int s = 0; s < 7; s++) {
                            const int fx = 2 * x + STEN[s][0], fy = 2 * y + STEN[s][1], fr = 2 * r + dr;
                            if (fx < 0 || fx >= mf->n || fy < 0 || fy >= mf->n || fr < 0 || fr >= mf->nr) continue;
                            const double w = pc_weight(fx, fy, fr, x, y, r);
                            for (int c = 0; c < ncomp; c++)
                                xf[c * Nf + IDX4(mf->n, mf->nr, d, fx, fy, fr)] +=
                                    w * xc[c * Nc + IDX4(mc->n, mc->nr, d, x, y, r)];
                        }
}

/* RestrictionVecConstant: restriction_constant.hpp:134-250 (owned-masked
 * gather then additive communication) */
void orc_restrict(const orc_mesh* mf, const orc_mesh* mc, int ncomp, const double* xf, double* xc)
{
    const int64_t Nc = orc_mesh_nodes(mc), Nf = orc_mesh_nodes(mf);
    memset(xc, 0, sizeof(double) * ncomp * Nc);
    for (int d = 0; d < 10; d++)
        for (int x = 0; x < mc->n; x++)
            for (int y = 0; y < mc->n; y++)
                for (int r = 0; r < mc->nr; r++)
                    for (int dr = -1; dr <= 1; dr++)
                        for (int s = 0; s < 7; s++) {
                            const int fx = 2 * x + STEN[s][0], fy = 2 * y + STEN[s][1], fr = 2 * r + dr;
                            if (fx < 0 || fx >= mf->n || fy < 0 || fy >= mf->n || fr < 0 || fr >= mf->nr) continue;
                            const int64_t fi = IDX4(mf->n, mf->nr, d, fx, fy, fr);
                            const double w = pc_weight(fx, fy, fr, x, y, r) * (mf->owned[fi] ? 1.0 : 0.0);
                            for (int c = 0; c < ncomp; c++)
                                xc[c * Nc + IDX4(mc->n, mc->nr, d, x, y, r)] += w * xf[c * Nf + fi];
                        }
    orc_interface_sum(mc->n, mc->nr, ncomp, mc->group_of, mc->ngroups, xc);
}

/* ====================================================================== GCA */

/* TwoGridGCA, constant interpolation: linalg/solvers/gca/gca.hpp:154-531.
 * The 6x6 local P follows the coded rules literally, including the radial
 * bracket r_bot = r_fine/2 (gca.hpp:213-216) that, for in-plane midpoints on
 * the top plane of an interior coarse cell, places the weights in the bottom
 * coarse dofs (recorded as finding F11). */
static void gca_build(const orc_eps* fine, const orc_mesh* mc, double* out, int geometric_bracket);

void orc_gca_build(const orc_eps* fine, const orc_mesh* mc, double* out) { gca_build(fine, mc, out, 0); }

/* same with the radial bracket of every fine dof taken as the coarse cell it
 * lies in (the Galerkin product of the global constant prolongation); test
 * infrastructure for the brute-force check of the element loop */
void orc_gca_build_geometric(const orc_eps* fine, const orc_mesh* mc, double* out) { gca_build(fine, mc, out, 1); }

static void gca_build(const orc_eps* fine, const orc_mesh* mc, double* out, int geometric_bracket)
{
    const orc_mesh* mf = fine->m;
    static const int SH[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const int ncl = mc->n - 1, ncr = mc->nr - 1;
    for (int d = 0; d < 10; d++)
        for (int X = 0; X < ncl; X++)
            for (int Y = 0; Y < ncl; Y++)
                for (int Rr = 0; Rr < ncr; Rr++) {
                    double Ac[2][324];
                    memset(Ac, 0, sizeof(Ac));
                    for (int h = 0; h < 8; h++) {
                        const int fx = 2 * X + SH[h][0], fy = 2 * Y + SH[h][1], fr = 2 * Rr + SH[h][2];
                        for (int w = 0; w < 2; w++) {
                            double P[36] = {0};
                            for (int fl = 0; fl < 6; fl++) {
                                const int ix = fx + WOFF[w][fl][0], iy = fy + WOFF[w][fl][1], ir = fr + WOFF[w][fl][2];
                                if (ix % 2 == 0 && iy % 2 == 0 && ir % 2 == 0) { P[fl * 6 + fl] = 1.0; continue; }
                                const int rb = geometric_bracket ? Rr : (ir < mf->nr - 1 ? ir / 2 : ir / 2 - 1);
                                const int rt = rb + 1;
                                if (ix % 2 == 0 && iy % 2 == 0) {
                                    const int a = fl % 3;
                                    P[fl * 6 + a] = 0.5;
                                    P[fl * 6 + a + 3] = 0.5;
                                    continue;
                                }
                                int x0, x1, y0, y1, cl[4];
                                if (ix % 2 == 0) {
                                    x0 = x1 = ix / 2; y0 = iy / 2; y1 = iy / 2 + 1;
                                    cl[0] = 0; cl[1] = 2; cl[2] = 3; cl[3] = 5;
                                } else if (iy % 2 == 0) {
                                    x0 = ix / 2; x1 = ix / 2 + 1; y0 = y1 = iy / 2;
                                    cl[0] = 0; cl[1] = 1; cl[2] = 3; cl[3] = 4;
                                } else {
                                    x0 = ix / 2 + 1; x1 = ix / 2; y0 = iy / 2; y1 = iy / 2 + 1;
                                    cl[0] = 1; cl[1] = 2; cl[2] = 4; cl[3] = 5;
                                }
                                P[fl * 6 + cl[0]] = pc_weight(ix, iy, ir, x0, y0, rb);
                                P[fl * 6 + cl[1]] = pc_weight(ix, iy, ir, x1, y1, rb);
                                P[fl * 6 + cl[2]] = pc_weight(ix, iy, ir, x0, y0, rt);
                                P[fl * 6 + cl[3]] = pc_weight(ix, iy, ir, x1, y1, rt);
                            }
                            double Af[324];
                            get_local(fine, d, fx, fy, fr, w, Af);
                            /* PTAP with P_vec = blockdiag(P,P,P) (dimension-wise) */
                            double T[324], PT[324];
                            double Pv[324] = {0};
                            for (int c = 0; c < 3; c++)
                                for (int i = 0; i < 6; i++)
                                    for (int j = 0; j < 6; j++) Pv[(i + c * 6) * 18 + j + c * 6] = P[i * 6 + j];
                            for (int i = 0; i < 18; i++)
                                for (int j = 0; j < 18; j++) {
                                    double s = 0;
                                    for (int l = 0; l < 18; l++) s += Af[i * 18 + l] * Pv[l * 18 + j];
                                    T[i * 18 + j] = s;
                                }
                            for (int i = 0; i < 18; i++)
                                for (int j = 0; j < 18; j++) {
                                    double s = 0;
                                    for (int l = 0; l < 18; l++) s += Pv[l * 18 + i] * T[l * 18 + j];
                                    PT[i * 18 + j] = s;
                                }
                            int tgt;
                            if ((w == 0 && (h == 0 || h == 1 || h == 2 || h == 4 || h == 5 || h == 6)) ||
                                (w == 1 && (h == 0 || h == 4)))
                                tgt = 0;
                            else
                                tgt = 1;
                            for (int t = 0; t < 324; t++) Ac[tgt][t] += PT[t];
                        }
                    }
                    const int64_t c = (((int64_t)d * ncl + X) * ncl + Y) * ncr + Rr;
                    memcpy(out + (c * 2) * 324, Ac[0], sizeof(double) * 324);
                    memcpy(out + (c * 2 + 1) * 324, Ac[1], sizeof(double) * 324);
                }
}

/* ---------------------------------------------------------------- cell viscosity + adjoint gradient */
void orc_eps_set_cell_viscosity(orc_eps* op, const double* kc) { op->kc = kc; }

void orc_eps_grad_log_cell_viscosity(const orc_eps* op, const double* u, const double* lam, double* g)
{
    const orc_mesh* m = op->m;
    const int n = m->n, nr = m->nr, ncl = n - 1, ncr = nr - 1;
    const int64_t N = orc_mesh_nodes(m);
    const int64_t cells = 10LL * ncl * ncl * ncr;
#pragma omp parallel for schedule(static)
    for (int64_t cell = 0; cell < cells; cell++) {
        const int r = (int)(cell % ncr);
        const int y = (int)((cell / ncr) % ncl);
        const int x = (int)((cell / ((int64_t)ncr * ncl)) % ncl);
        const int d = (int)(cell / ((int64_t)ncr * ncl * ncl));
        double acc = 0.0;
        for (int w = 0; w < 2; w++) {
            double A[324], ue[18], le[18];
            orc_eps_local_matrix_k(m, 1.0, d, x, y, r, w, A);
            for (int j = 0; j < 6; j++) {
                const int gx = x + WOFF[w][j][0], gy = y + WOFF[w][j][1], gr = r + WOFF[w][j][2];
                const int64_t idx = IDX4(n, nr, d, gx, gy, gr);
                double lv[3];
                for (int c = 0; c < 3; c++) {
                    ue[j + 6 * c] = u[c * N + idx];
                    lv[c] = lam[c * N + idx];
                }
                const int mode = gr == 0 ? op->bc_cmb : (gr == nr - 1 ? op->bc_surf : ORC_NEUMANN);
                if (mode == ORC_DIRICHLET) {
                    lv[0] = lv[1] = lv[2] = 0.0;
                } else if (mode == ORC_FREESLIP) {
                    const double* cc = m->coords + 3 * (((int64_t)d * n + gx) * n + gy);
                    const double nn = sqrt(cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
                    const double nv[3] = {cc[0] / nn, cc[1] / nn, cc[2] / nn};
                    const double ln = nv[0] * lv[0] + nv[1] * lv[1] + nv[2] * lv[2];
                    for (int c = 0; c < 3; c++) lv[c] -= ln * nv[c];
                }
                for (int c = 0; c < 3; c++) le[j + 6 * c] = lv[c];
            }
            for (int i = 0; i < 18; i++) {
                double s = 0.0;
                for (int j = 0; j < 18; j++) s += A[i * 18 + j] * ue[j];
                acc += le[i] * s;
            }
        }
        g[cell] = -op->kc[cell] * acc;
    }
}
