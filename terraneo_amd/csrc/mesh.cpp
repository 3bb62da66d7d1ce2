// Thick spherical shell mesh: 10 icosahedral diamonds, lateral refinement by
// great-circle midpoint bisection, radial shells; interface topology and
// ownership; per-wedge lateral geometry used by the matrix-free kernels.
//
// Reference behaviour reproduced (bit-exact node coordinates):
//   icosahedron + diamond corner table  spherical_shell.hpp:441-585
//   midpoint-normalise creation rule    spherical_shell.hpp:230-346
//   uniform radii                       spherical_shell.hpp:27-52
//   ownership (lowest subdomain owns)   grid/bit_masks.hpp:29-79
// The generator itself is a breadth-first sweep over strides (each node is
// created from its two parents one stride up) instead of the reference's
// memoised recursion; parents and arithmetic are identical, so the bits are.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "tn_internal.hpp"

namespace tn {
namespace {

struct V3 {
    double x, y, z;
};

// Midpoint of two unit vectors pushed back to the sphere.  Sequential sum of
// squares then three divisions: the rounding sequence the reference uses
// (dense/vec.hpp:76-98), which keeps shared diamond edges bitwise equal.
inline V3 sphere_mid(const V3& a, const V3& b)
{
    V3 m{a.x + b.x, a.y + b.y, a.z + b.z};
    double s = 0.0;
    s += m.x * m.x;
    s += m.y * m.y;
    s += m.z * m.z;
    const double mag = std::sqrt(s);
    if (mag > 1e-15) {
        m.x /= mag;
        m.y /= mag;
        m.z /= mag;
    }
    return m;
}

// The 12 icosahedron vertices (two poles and two staggered pentagonal rings)
// and, per diamond, its (pole, left, far, right) vertices.
void base_icosahedron(V3 vtx[12], int dia[10][4])
{
    const double fifthpi = 0.4 * std::asin(1.0);
    const double w = 2.0 * std::acos(1.0 / (2.0 * std::sin(fifthpi)));
    const double cw = std::cos(w), sw = std::sin(w);
    vtx[0] = {0.0, 0.0, 1.0};
    vtx[11] = {0.0, 0.0, -1.0};
    for (int k = 1; k <= 5; ++k) {
        const double phu = 2.0 * (double(k) - 0.5) * fifthpi;
        const double phl = 2.0 * (double(k) - 1) * fifthpi;
        vtx[k] = {sw * std::cos(phu), sw * std::sin(phu), cw};
        vtx[k + 5] = {sw * std::cos(phl), sw * std::sin(phl), -cw};
    }
    static const int table[10][4] = {{0, 5, 6, 1},  {0, 1, 7, 2},  {0, 2, 8, 3},  {0, 3, 9, 4},  {0, 4, 10, 5},
                                     {11, 7, 1, 6}, {11, 8, 2, 7}, {11, 9, 3, 8}, {11, 10, 4, 9}, {11, 6, 5, 10}};
    std::memcpy(dia, table, sizeof(table));
}

void diamond_nodes(int d, int N, const V3 vtx[12], const int dia[10][4], V3* out /*[(N+1)^2]*/)
{
    const int n = N + 1;
    auto at = [&](int i, int j) -> V3& { return out[(size_t)i * n + j]; };
    // northern diamonds run (pole, left=ring idx 1, far, right=idx 3); southern swap left/right
    const int L = d < 5 ? 1 : 3, R = d < 5 ? 3 : 1;
    at(0, 0) = vtx[dia[d][0]];
    at(N, 0) = vtx[dia[d][L]];
    at(N, N) = vtx[dia[d][2]];
    at(0, N) = vtx[dia[d][R]];
    for (int l = N; l >= 2; l /= 2) {
        const int h = l / 2;
        for (int i = 0; i <= N; i += h)
            for (int j = 0; j <= N; j += h) {
                const int ri = i % l, rj = j % l;
                if (ri == 0 && rj == 0) continue;  // exists already
                if (ri == 0)
                    at(i, j) = sphere_mid(at(i, j - h), at(i, j + h));
                else if (rj == 0)
                    at(i, j) = sphere_mid(at(i - h, j), at(i + h, j));
                else
                    at(i, j) = sphere_mid(at(i - h, j + h), at(i + h, j - h));
            }
    }
}

struct KeyHash {
    size_t operator()(const std::array<uint64_t, 3>& k) const
    {
        uint64_t h = k[0] * 0x9E3779B97F4A7C15ull;
        h ^= k[1] + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= k[2] + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

void cof_inv_t(const double A[3][3], double K[3][3], double& det)
{
    // K = A^{-T} = cofactor(A) / det(A)
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const double c10 = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    const double c11 = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    const double c12 = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    const double c20 = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    const double c21 = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    const double c22 = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const double id = 1.0 / det;
    K[0][0] = c00 * id; K[0][1] = c01 * id; K[0][2] = c02 * id;
    K[1][0] = c10 * id; K[1][1] = c11 * id; K[1][2] = c12 * id;
    K[2][0] = c20 * id; K[2][1] = c21 * id; K[2][2] = c22 * id;
}

}  // namespace

void MeshLevel::build(int lat_level_in, const std::vector<double>& radii_in)
{
    TN_REQUIRE(lat_level_in >= 0 && lat_level_in <= 12, "lateral level out of range [0,12]");
    TN_REQUIRE(radii_in.size() >= 3, "need at least 3 radial shells (2 layers)");
    for (size_t i = 1; i < radii_in.size(); ++i) TN_REQUIRE(radii_in[i] > radii_in[i - 1], "radii must increase");
    lat_level = lat_level_in;
    const int N = 1 << lat_level;
    n = N + 1;
    nr = (int)radii_in.size();
    radii = radii_in;
    this->N = 10LL * n * n * nr;

    V3 vtx[12];
    int dia[10][4];
    base_icosahedron(vtx, dia);
    std::vector<V3> dn((size_t)n * n);
    coords.assign((size_t)30 * n * n, 0.0);
    for (int d = 0; d < 10; ++d) {
        diamond_nodes(d, N, vtx, dia, dn.data());
        for (int i = 0; i < n * n; ++i) {
            double* o = &coords[3 * ((size_t)d * n * n + i)];
            o[0] = dn[i].x;
            o[1] = dn[i].y;
            o[2] = dn[i].z;
        }
    }

    // interface topology: copies of one lateral node carry identical bits
    const int64_t lat = 10LL * n * n;
    std::unordered_map<std::array<uint64_t, 3>, int, KeyHash> key2g;
    key2g.reserve((size_t)lat);
    std::vector<int> gid(lat);
    std::vector<int> gcount;
    for (int64_t i = 0; i < lat; ++i) {
        std::array<uint64_t, 3> k;
        std::memcpy(k.data(), &coords[3 * i], 24);
        auto it = key2g.find(k);
        if (it == key2g.end()) {
            it = key2g.emplace(k, (int)gcount.size()).first;
            gcount.push_back(0);
        }
        gid[i] = it->second;
        gcount[it->second]++;
    }
    // lowest diamond owns; duplicated groups go to the CSR used by the interface sum
    std::vector<int> gmin(gcount.size(), 1 << 30);
    for (int64_t i = 0; i < lat; ++i) gmin[gid[i]] = std::min<int>(gmin[gid[i]], (int)(i / ((int64_t)n * n)));
    owned_lat.assign(lat, 0);
    owned = 0;
    for (int64_t i = 0; i < lat; ++i)
        if (gmin[gid[i]] == (int)(i / ((int64_t)n * n))) {
            owned_lat[i] = 1;
            owned += nr;
        }
    std::vector<int> dup_index(gcount.size(), -1);
    int ndup = 0;
    for (size_t g = 0; g < gcount.size(); ++g)
        if (gcount[g] > 1) dup_index[g] = ndup++;
    grp_off.assign(ndup + 1, 0);
    for (size_t g = 0; g < gcount.size(); ++g)
        if (dup_index[g] >= 0) grp_off[dup_index[g] + 1] = gcount[g];
    for (int g = 0; g < ndup; ++g) grp_off[g + 1] += grp_off[g];
    grp_mem.assign(grp_off[ndup], 0);
    std::vector<int> fill(ndup, 0);
    for (int64_t i = 0; i < lat; ++i) {
        const int g = dup_index[gid[i]];
        if (g >= 0) grp_mem[grp_off[g] + fill[g]++] = (int32_t)i;
    }

    // per-wedge lateral geometry
    const int nc = n - 1;
    geo.assign((size_t)10 * nc * nc * 2 * GEO_STRIDE, 0.0);
    static const int WV[2][3][2] = {{{0, 0}, {1, 0}, {0, 1}}, {{1, 1}, {0, 1}, {1, 0}}};
    for (int d = 0; d < 10; ++d)
        for (int x = 0; x < nc; ++x)
            for (int y = 0; y < nc; ++y)
                for (int w = 0; w < 2; ++w) {
                    const double* v[3];
                    for (int k = 0; k < 3; ++k)
                        v[k] = &coords[3 * (((size_t)d * n + x + WV[w][k][0]) * n + y + WV[w][k][1])];
                    double A[3][3], K[3][3], det;
                    for (int a = 0; a < 3; ++a) {
                        A[a][0] = v[1][a] - v[0][a];
                        A[a][1] = v[2][a] - v[0][a];
                        A[a][2] = (v[0][a] + v[1][a] + v[2][a]) / 3.0;
                    }
                    cof_inv_t(A, K, det);
                    double* g = &geo[((((size_t)d * nc + x) * nc + y) * 2 + w) * GEO_STRIDE];
                    for (int a = 0; a < 3; ++a) {
                        g[0 + a] = 0.5 * K[a][0];
                        g[3 + a] = 0.5 * K[a][1];
                        g[6 + a] = K[a][2] / 6.0;
                    }
                    g[9] = std::fabs(det);
                }
}

void MeshLevel::upload()
{
    d_coords.upload(coords);
    std::vector<double> co(coords);
    for (size_t lat = 0; lat < owned_lat.size(); ++lat)
        if (!owned_lat[lat]) co[3 * lat] = co[3 * lat + 1] = co[3 * lat + 2] = 0.0;
    d_coords_owned.upload(co);
    std::vector<float> cf(coords.begin(), coords.end());
    d_coords_f.upload(cf);
    d_radii.upload(radii);
    std::vector<float> rf(radii.begin(), radii.end());
    d_radii_f.upload(rf);
    d_geo.upload(geo);
    std::vector<float> gf(geo.begin(), geo.end());
    d_geo_f.upload(gf);
    d_owned_lat.upload(owned_lat);
    d_grp_off.upload(grp_off);
    d_grp_mem.upload(grp_mem);
    d_scratch.alloc(64);
    d_part.alloc(REDUCE_PART_DOUBLES);
    uploaded = true;
}

}  // namespace tn
