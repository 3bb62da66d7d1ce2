// C ABI: XDMF time series and checkpoints in the reference's file format
// (io/xdmf.hpp: format :12-146, XDMFOutput :147-1090, checkpoint readers
// :1136-1420) for the one-subdomain-per-diamond layout of this library.
//
// Directory contents: geometry.bin (node coordinates, per diamond r-, y-,
// x-major), topology.bin (6 node indices per wedge, two wedges per hex cell:
// (v0 v1 v2 v4 v5 v6) and (v3 v2 v1 v7 v6 v5)), checkpoint_metadata.bin
// (version 1, written with the first step), <name>_<step>.bin per field and
// step (node order as the geometry, vector components innermost), and
// step_<step>.xmf, written last so that a complete .xmf marks a complete step.
// Fields are device arrays in this library's layout (scalar [10][n][n][nr],
// vector [3][10][n][n][nr]); write() copies them to the host and reorders.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "tn_internal.hpp"

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

namespace {

constexpr int XDMF_VERSION = 1;
constexpr int SCALAR_FLOAT = 2;  // scalar_data_type code of floating point data

// SubdomainInfo::global_id (grid/shell/spherical_shell.hpp:594-680) for (diamond, 0, 0, 0)
int64_t global_subdomain_id(int diamond) { return (int64_t)diamond << 57; }

// util::XML (util/xml.hpp): attributes in sorted order, two-space indent
struct Xml {
    std::string name, content;
    std::map<std::string, std::string> attr;
    std::vector<Xml> children;
    static std::string escape(const std::string& s)
    {
        std::string o;
        for (char c : s) {
            switch (c) {
                case '&': o += "&amp;"; break;
                case '"': o += "&quot;"; break;
                case '\'': o += "&apos;"; break;
                case '<': o += "&lt;"; break;
                case '>': o += "&gt;"; break;
                default: o += c;
            }
        }
        return o;
    }
    std::string str(int indent = 0) const
    {
        const std::string pad(indent, ' ');
        std::string o = pad + "<" + name;
        for (const auto& kv : attr) o += " " + kv.first + "=\"" + escape(kv.second) + "\"";
        if (children.empty() && content.empty()) return o + " />\n";
        o += ">";
        if (!content.empty()) o += escape(content);
        if (!children.empty()) {
            o += "\n";
            for (const auto& c : children) o += c.str(indent + 2);
            o += pad;
        }
        return o + "</" + name + ">\n";
    }
};

Xml data_item(const std::string& type, int precision, const std::string& dims, const std::string& file)
{
    Xml d;
    d.name = "DataItem";
    d.attr = {{"Format", "Binary"}, {"DataType", type}, {"Precision", std::to_string(precision)},
              {"Endian", "Little"}, {"Dimensions", dims}};
    d.content = file;
    return d;
}

void write_file(const std::string& path, const void* data, size_t bytes)
{
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    TN_REQUIRE(f.good(), "XDMF: cannot open " + path + " for writing");
    f.write((const char*)data, (std::streamsize)bytes);
    TN_REQUIRE(f.good(), "XDMF: write to " + path + " failed");
}

std::vector<char> read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    TN_REQUIRE(f.good(), "XDMF: cannot open " + path);
    const std::streamsize n = f.tellg();
    std::vector<char> b((size_t)n);
    f.seekg(0);
    f.read(b.data(), n);
    TN_REQUIRE(f.good(), "XDMF: read of " + path + " failed");
    return b;
}

template <typename V>
void put(std::vector<char>& b, V v)
{
    const char* p = (const char*)&v;
    b.insert(b.end(), p, p + sizeof(V));
}

struct Field {
    std::string name;
    const void* ptr;
    int dtype;  // of the device array
    int vec;
    int out_bytes;
};

// field in this library's layout [vec][10][n(x)][n(y)][nr] -> file order [10][nr][n(y)][n(x)][vec]
template <typename TI, typename TO>
void to_file_order(const TI* in, TO* out, int n, int nr, int vec)
{
    const int64_t N = 10LL * n * n * nr;
    int64_t k = 0;
    for (int d = 0; d < 10; ++d)
        for (int r = 0; r < nr; ++r)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    const int64_t node = (((int64_t)d * n + x) * n + y) * nr + r;
                    for (int c = 0; c < vec; ++c) out[k++] = (TO)in[c * N + node];
                }
}

}  // namespace

struct tn_xdmf {
    const tn_mesh* mesh = nullptr;
    std::string dir;
    int points_bytes = 4, conn_bytes = 8;
    std::vector<Field> fields;
    int counter = 0;
    bool first_write_done = false;
};

namespace {

// the reference writes scalar-float, scalar-double, vector-float, vector-double grids, each in add order
std::vector<const Field*> ordered(const tn_xdmf* x)
{
    std::vector<const Field*> o;
    for (int key = 0; key < 4; ++key)
        for (const auto& f : x->fields) {
            const int k = (f.vec == 1 ? 0 : 2) + (f.dtype == TN_F32 ? 0 : 1);
            if (k == key) o.push_back(&f);
        }
    return o;
}

void write_mesh_files(const tn_xdmf* x)
{
    const MeshLevel& L = x->mesh->lvl;
    const int n = L.n, nr = L.nr;
    // geometry: per diamond r-, y-, x-major
    {
        std::vector<char> b;
        b.reserve(10ULL * n * n * nr * 3 * x->points_bytes);
        for (int d = 0; d < 10; ++d)
            for (int r = 0; r < nr; ++r)
                for (int yy = 0; yy < n; ++yy)
                    for (int xx = 0; xx < n; ++xx)
                        for (int c = 0; c < 3; ++c) {
                            const double v = L.coords[3 * (((int64_t)d * n + xx) * n + yy) + c] * L.radii[r];
                            if (x->points_bytes == 4)
                                put(b, (float)v);
                            else
                                put(b, v);
                        }
        write_file(x->dir + "/geometry.bin", b.data(), b.size());
    }
    // topology: two wedges per hex cell
    {
        const int64_t s0 = (int64_t)n * n * nr, s1 = (int64_t)n * n, s2 = n;
        std::vector<char> b;
        b.reserve(10ULL * (n - 1) * (n - 1) * (nr - 1) * 12 * x->conn_bytes);
        for (int d = 0; d < 10; ++d)
            for (int r = 0; r < nr - 1; ++r)
                for (int yy = 0; yy < n - 1; ++yy)
                    for (int xx = 0; xx < n - 1; ++xx) {
                        const int64_t v0 = d * s0 + r * s1 + yy * s2 + xx;
                        const int64_t v1 = v0 + 1, v2 = v0 + n, v3 = v0 + n + 1;
                        const int64_t v4 = v0 + s1, v5 = v4 + 1, v6 = v4 + n, v7 = v4 + n + 1;
                        const int64_t w[12] = {v0, v1, v2, v4, v5, v6, v3, v2, v1, v7, v6, v5};
                        for (int64_t v : w) {
                            if (x->conn_bytes == 4)
                                put(b, (int32_t)v);
                            else
                                put(b, v);
                        }
                    }
        write_file(x->dir + "/topology.bin", b.data(), b.size());
    }
    // checkpoint metadata (version 1)
    {
        std::vector<char> b;
        for (int v : {XDMF_VERSION, 1, 1, n, n, nr}) put(b, (int32_t)v);
        for (double r : L.radii) put(b, r);
        put(b, (int32_t)x->points_bytes);
        const auto fs = ordered(x);
        put(b, (int32_t)fs.size());
        for (const Field* f : fs) {
            put(b, (int32_t)f->name.size());
            b.insert(b.end(), f->name.begin(), f->name.end());
            put(b, (int32_t)SCALAR_FLOAT);
            put(b, (int32_t)f->out_bytes);
            put(b, (int32_t)f->vec);
        }
        put(b, (int32_t)0);  // checkpoint_subdomain_ordering
        for (int d = 0; d < 10; ++d) put(b, global_subdomain_id(d));
        write_file(x->dir + "/checkpoint_metadata.bin", b.data(), b.size());
    }
    const char* readme =
        "# Output of the XDMF writer (visualisation and checkpoints)\n\n"
        "- `geometry.bin`: node coordinates\n- `topology.bin`: wedge connectivity\n"
        "- `checkpoint_metadata.bin`: checkpoint metadata\n"
        "- `step_*.xmf`: XDMF file per write() step (open with Paraview)\n"
        "- `<name>_*.bin`: binary field data per write() step\n";
    write_file(x->dir + "/README.md", readme, std::strlen(readme));
}

}  // namespace

extern "C" {

int tn_xdmf_create(const tn_mesh* m, const char* directory, int points_bytes, int connectivity_bytes, tn_xdmf** out)
{
    TN_API_BEGIN
    TN_REQUIRE(m && directory && out, "null pointer");
    TN_REQUIRE(points_bytes == 4 || points_bytes == 8, "XDMF: points_bytes must be 4 or 8");
    TN_REQUIRE(connectivity_bytes == 4 || connectivity_bytes == 8, "XDMF: connectivity_bytes must be 4 or 8");
    auto* x = new tn_xdmf;
    x->mesh = m;
    x->dir = directory;
    x->points_bytes = points_bytes;
    x->conn_bytes = connectivity_bytes;
    *out = x;
    TN_API_END
}

int tn_xdmf_destroy(tn_xdmf* x)
{
    TN_API_BEGIN
    delete x;
    TN_API_END
}

int tn_xdmf_add(tn_xdmf* x, const char* name, int dtype, int vec_dim, int out_bytes, const void* field)
{
    TN_API_BEGIN
    TN_REQUIRE(x && name && field, "null pointer");
    TN_REQUIRE(!x->first_write_done, "XDMF add(): cannot add data after write() has been called");
    TN_REQUIRE(vec_dim == 1 || vec_dim == 3, "XDMF add(): vec_dim must be 1 or 3");
    TN_REQUIRE(out_bytes == 4 || out_bytes == 8, "XDMF add(): out_bytes must be 4 or 8");
    TN_REQUIRE(dtype == TN_F32 || dtype == TN_F64, "XDMF add(): dtype must be TN_F32 or TN_F64");
    for (const auto& f : x->fields)
        TN_REQUIRE(f.name != name, std::string("XDMF add(): data with label '") + name + "' was added before");
    x->fields.push_back(Field{name, field, dtype, vec_dim, out_bytes});
    TN_API_END
}

int tn_xdmf_set_write_counter(tn_xdmf* x, int step)
{
    TN_API_BEGIN
    TN_REQUIRE(x && step >= 0, "bad argument");
    x->counter = step;
    TN_API_END
}

int tn_xdmf_write(tn_xdmf* x, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(x, "null pointer");
    const MeshLevel& L = x->mesh->lvl;
    const int n = L.n, nr = L.nr;
    const int64_t nnodes = 10LL * n * n * nr;
    const int64_t nelem = 10LL * (n - 1) * (n - 1) * (nr - 1) * 2;
    if (!x->first_write_done) {
        mkdir(x->dir.c_str(), 0755);  // an existing directory is fine
        write_mesh_files(x);
    }
    Xml grid;
    grid.name = "Grid";
    grid.attr = {{"Name", "Grid"}, {"GridType", "Uniform"}};
    Xml geo;
    geo.name = "Geometry";
    geo.attr = {{"Type", "XYZ"}};
    geo.children.push_back(data_item("Float", x->points_bytes, std::to_string(nnodes) + " 3", "geometry.bin"));
    grid.children.push_back(geo);
    Xml top;
    top.name = "Topology";
    top.attr = {{"Type", "Wedge"}, {"NumberOfElements", std::to_string(nelem)}};
    top.children.push_back(data_item("Int", x->conn_bytes, std::to_string(nelem * 6), "topology.bin"));
    grid.children.push_back(top);
    TN_HIP(hipStreamSynchronize((hipStream_t)s));
    for (const Field* f : ordered(x)) {
        const std::string fname = f->name + "_" + std::to_string(x->counter) + ".bin";
        const size_t in_bytes = (f->dtype == TN_F32 ? 4 : 8) * (size_t)(f->vec * nnodes);
        std::vector<char> host(in_bytes);
        TN_HIP(hipMemcpy(host.data(), f->ptr, in_bytes, hipMemcpyDeviceToHost));
        std::vector<char> out((size_t)f->out_bytes * f->vec * nnodes);
        if (f->dtype == TN_F64 && f->out_bytes == 8)
            to_file_order((const double*)host.data(), (double*)out.data(), n, nr, f->vec);
        else if (f->dtype == TN_F64)
            to_file_order((const double*)host.data(), (float*)out.data(), n, nr, f->vec);
        else if (f->out_bytes == 8)
            to_file_order((const float*)host.data(), (double*)out.data(), n, nr, f->vec);
        else
            to_file_order((const float*)host.data(), (float*)out.data(), n, nr, f->vec);
        write_file(x->dir + "/" + fname, out.data(), out.size());
        Xml at;
        at.name = "Attribute";
        at.attr = {{"Name", f->name}, {"AttributeType", f->vec == 1 ? "Scalar" : "Vector"}, {"Center", "Node"}};
        at.children.push_back(data_item("Float", f->out_bytes,
                                        f->vec == 1 ? std::to_string(nnodes)
                                                    : std::to_string(nnodes) + " " + std::to_string(f->vec),
                                        fname));
        grid.children.push_back(at);
    }
    Xml dom;
    dom.name = "Domain";
    dom.children.push_back(grid);
    Xml root;
    root.name = "Xdmf";
    root.attr = {{"Version", "2.0"}};
    root.children.push_back(dom);
    const std::string text = "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n" + root.str();
    write_file(x->dir + "/step_" + std::to_string(x->counter) + ".xmf", text.data(), text.size());
    x->counter += 1;
    x->first_write_done = true;
    TN_API_END
}

// read_xdmf_checkpoint_grid (io/xdmf.hpp:1282-1420): field `name` of `step`
// into a device array of this library's layout (dtype TN_F32 / TN_F64;
// vec_dim components as recorded in the checkpoint)
int tn_xdmf_read_checkpoint_grid(const tn_mesh* m, const char* directory, const char* name, int step, int dtype,
                                 int vec_dim, void* field, tn_stream_t s)
{
    TN_API_BEGIN
    TN_REQUIRE(m && directory && name && field, "null pointer");
    TN_REQUIRE(dtype == TN_F32 || dtype == TN_F64, "dtype must be TN_F32 or TN_F64");
    const MeshLevel& L = m->lvl;
    const std::string dir = directory;
    const std::vector<char> md = read_file(dir + "/checkpoint_metadata.bin");
    size_t off = 0;
    auto take_i32 = [&]() {
        TN_REQUIRE(off + 4 <= md.size(), "XDMF: truncated checkpoint metadata");
        int32_t v;
        std::memcpy(&v, md.data() + off, 4);
        off += 4;
        return v;
    };
    auto take_f64 = [&]() {
        TN_REQUIRE(off + 8 <= md.size(), "XDMF: truncated checkpoint metadata");
        double v;
        std::memcpy(&v, md.data() + off, 8);
        off += 8;
        return v;
    };
    const int version = take_i32(), nsl = take_i32(), nsr = take_i32();
    const int sx = take_i32(), sy = take_i32(), sr = take_i32();
    TN_REQUIRE(version == 0 || version == 1, "XDMF: unsupported checkpoint version " + std::to_string(version));
    TN_REQUIRE(nsl == 1 && nsr == 1, "XDMF: only one subdomain per diamond is supported");
    TN_REQUIRE(sx == L.n && sy == L.n && sr == L.nr, "XDMF: grid extents do not match the checkpoint");
    std::vector<double> radii(sr);
    for (auto& r : radii) r = take_f64();
    for (int i = 1; i < sr; ++i) TN_REQUIRE(radii[i] > radii[i - 1], "XDMF: radii are not sorted in the checkpoint");
    if (version > 0) (void)take_i32();  // grid scalar bytes
    const int nfiles = take_i32();
    int out_bytes = 0, vec = 0;
    for (int k = 0; k < nfiles; ++k) {
        const int len = take_i32();
        TN_REQUIRE(len >= 0 && off + (size_t)len <= md.size(), "XDMF: truncated checkpoint metadata");
        const std::string nm(md.data() + off, (size_t)len);
        off += (size_t)len;
        (void)take_i32();  // scalar data type
        const int nb = take_i32(), vd = take_i32();
        if (nm == name) {
            out_bytes = nb;
            vec = vd;
        }
    }
    TN_REQUIRE(vec > 0, std::string("XDMF: no data '") + name + "' in the checkpoint");
    TN_REQUIRE(vec == vec_dim, "XDMF: the checkpoint's vec_dim differs from the requested one");
    TN_REQUIRE(take_i32() == 0, "XDMF: checkpoint ordering is not 0");
    std::vector<int64_t> ids(10);
    for (auto& id : ids) {
        TN_REQUIRE(off + 8 <= md.size(), "XDMF: truncated checkpoint metadata");
        std::memcpy(&id, md.data() + off, 8);
        off += 8;
    }
    const std::vector<char> raw = read_file(dir + "/" + name + "_" + std::to_string(step) + ".bin");
    const int64_t chunk = (int64_t)sx * sy * sr * vec;
    TN_REQUIRE((int64_t)raw.size() == 10 * chunk * out_bytes, "XDMF: data file size does not match the metadata");
    const int64_t N = 10LL * L.n * L.n * L.nr;
    std::vector<char> host((size_t)(dtype == TN_F32 ? 4 : 8) * vec * N);
    for (int d = 0; d < 10; ++d) {
        const auto it = std::find(ids.begin(), ids.end(), global_subdomain_id(d));
        TN_REQUIRE(it != ids.end(), "XDMF: a diamond is missing from the checkpoint");
        const int64_t base = (int64_t)(it - ids.begin()) * chunk;
        int64_t k = base;
        for (int r = 0; r < sr; ++r)
            for (int yy = 0; yy < sy; ++yy)
                for (int xx = 0; xx < sx; ++xx) {
                    const int64_t node = (((int64_t)d * L.n + xx) * L.n + yy) * L.nr + r;
                    for (int c = 0; c < vec; ++c, ++k) {
                        double v;
                        if (out_bytes == 4) {
                            float f;
                            std::memcpy(&f, raw.data() + 4 * k, 4);
                            v = f;
                        } else {
                            std::memcpy(&v, raw.data() + 8 * k, 8);
                        }
                        if (dtype == TN_F32)
                            ((float*)host.data())[c * N + node] = (float)v;
                        else
                            ((double*)host.data())[c * N + node] = v;
                    }
                }
    }
    TN_HIP(hipMemcpyAsync(field, host.data(), host.size(), hipMemcpyHostToDevice, (hipStream_t)s));
    TN_HIP(hipStreamSynchronize((hipStream_t)s));
    TN_API_END
}

}  // extern "C"
