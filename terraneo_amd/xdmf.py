"""XDMF output and checkpoints in the reference's file format
(src/terra/io/xdmf.hpp: format :12-146, writer XDMFOutput :147-1090, readers
read_xdmf_checkpoint_metadata :1136-1280 and read_xdmf_checkpoint_grid
:1282-1420), for the single-process, one-subdomain-per-diamond layout of this
library.

A directory holds
- `geometry.bin`    node coordinates, per diamond r-, y-, x-major, xyz per node
                    (Float32 by default);
- `topology.bin`    6 node indices per wedge, two wedges per hex cell
                    (v0 v1 v2 v4 v5 v6) and (v3 v2 v1 v7 v6 v5) (Int64 by default);
- `checkpoint_metadata.bin`  version 1 metadata (sizes, radii, grid list,
                    subdomain ids), written with the first step;
- `<name>_<step>.bin`  one file per field and step, same node order, vector
                    components innermost;
- `step_<step>.xmf` the XDMF description of the step (written last, so a
                    complete .xmf marks a complete step);
- `README.md`.

Field arrays use this library's layout: scalar [10][n][n][nr], vector
[3][10][n][n][nr] (component-major), as numpy arrays or torch tensors (copied
to the host).  The file order is the reference's data(subdomain, x, y, r, d)
written r-, y-, x-, d-major.
"""
import os
import struct

import numpy as np

VERSION = 1
_FLOAT = 2  # scalar_data_type code for floating point


def _np(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _to_file_order(field, n, nr, vec_dim):
    """[vec][10][n(x)][n(y)][nr] -> [10][nr][n(y)][n(x)][vec] (file order)."""
    a = field.reshape(vec_dim, 10, n, n, nr)
    return np.ascontiguousarray(a.transpose(1, 4, 3, 2, 0))


def _from_file_order(data, n, nr, vec_dim):
    a = data.reshape(10, nr, n, n, vec_dim)
    return np.ascontiguousarray(a.transpose(4, 0, 3, 2, 1))


def global_subdomain_id(diamond, sx=0, sy=0, sr=0):
    """SubdomainInfo::global_id (grid/shell/spherical_shell.hpp:594-680)."""
    return (diamond << 57) | (sr << 38) | (sy << 19) | sx


class _XML:
    """util::XML (util/xml.hpp): attributes in sorted order, two-space indent."""

    def __init__(self, name, attributes=None, content=""):
        self.name, self.attributes, self.content, self.children = name, dict(attributes or {}), content, []

    def add_child(self, child):
        self.children.append(child)
        return self

    @staticmethod
    def _escape(s):
        return (s.replace("&", "&amp;").replace('"', "&quot;").replace("'", "&apos;")
                .replace("<", "&lt;").replace(">", "&gt;"))

    def to_string(self, indent=0):
        pad = " " * indent
        out = pad + "<" + self.name
        for k in sorted(self.attributes):
            out += f' {k}="{self._escape(self.attributes[k])}"'
        if not self.children and not self.content:
            return out + " />\n"
        out += ">"
        if self.content:
            out += self._escape(self.content)
        if self.children:
            out += "\n" + "".join(c.to_string(indent + 2) for c in self.children) + pad
        return out + "</" + self.name + ">\n"


class XDMFOutput:
    """Writes the time series of `mesh` (an api.Mesh) to `directory`.

    points_bytes: 4 or 8 (geometry precision); connectivity_bytes: 4 or 8."""

    def __init__(self, directory, mesh, points_bytes=4, connectivity_bytes=8):
        assert points_bytes in (4, 8) and connectivity_bytes in (4, 8)
        self.dir, self.mesh = directory, mesh
        self.points_bytes, self.conn_bytes = points_bytes, connectivity_bytes
        self.fields = []  # (name, getter, vec_dim, out_bytes, input is float32)
        self.counter = 0
        self.first_write_done = False

    def set_write_counter(self, k):
        self.counter = int(k)

    def add(self, name, field, vec_dim=1, out_bytes=4):
        """Registers a field (array or a callable returning one) written at every write()."""
        if self.first_write_done:
            raise RuntimeError("XDMF add(): cannot add data after write() has been called")
        if any(f[0] == name for f in self.fields):
            raise ValueError(f"XDMF add(): data with label '{name}' was added before")
        assert vec_dim in (1, 3) and out_bytes in (4, 8)
        get = field if callable(field) else (lambda f=field: f)
        self.fields.append((name, get, vec_dim, out_bytes, str(_np(get()).dtype) == "float32"))

    def _ordered(self):
        # the reference writes scalar-float, scalar-double, vector-float, vector-double grids, each in add order
        key = {(1, True): 0, (1, False): 1, (3, True): 2, (3, False): 3}
        return sorted(self.fields, key=lambda f: key[(f[2], f[4])])

    def _geometry(self):
        m = self.mesh
        lat = np.asarray(m.lateral_coords(), dtype=np.float64).reshape(10, m.n, m.n, 3)
        rad = np.asarray(m.radii(), dtype=np.float64)
        xyz = lat.transpose(0, 2, 1, 3)[:, None, :, :, :] * rad[None, :, None, None, None]  # [10][r][y][x][3]
        return np.ascontiguousarray(xyz, dtype=np.float32 if self.points_bytes == 4 else np.float64)

    def _topology(self):
        m = self.mesh
        nx = ny = m.n
        nr = m.nr
        s0, s1, s2 = nx * ny * nr, nx * ny, nx
        sd, r, y, x = np.meshgrid(np.arange(10), np.arange(nr - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
        v0 = (sd * s0 + r * s1 + y * s2 + x).reshape(-1)
        v1, v2, v3 = v0 + 1, v0 + nx, v0 + nx + 1
        v4 = v0 + s1
        v5, v6, v7 = v4 + 1, v4 + nx, v4 + nx + 1
        w = np.stack([v0, v1, v2, v4, v5, v6, v3, v2, v1, v7, v6, v5], axis=1)
        return np.ascontiguousarray(w, dtype=np.int32 if self.conn_bytes == 4 else np.int64)

    def _metadata(self):
        m = self.mesh
        b = bytearray()
        b += struct.pack("<iiiiii", VERSION, 1, 1, m.n, m.n, m.nr)
        b += np.asarray(m.radii(), dtype="<f8").tobytes()
        b += struct.pack("<i", self.points_bytes)
        fields = self._ordered()
        b += struct.pack("<i", len(fields))
        for name, _, vec_dim, out_bytes, _ in fields:
            nb = name.encode()
            b += struct.pack("<i", len(nb)) + nb + struct.pack("<iii", _FLOAT, out_bytes, vec_dim)
        b += struct.pack("<i", 0)  # checkpoint_subdomain_ordering
        b += np.array([global_subdomain_id(d) for d in range(10)], dtype="<i8").tobytes()
        return bytes(b)

    def write(self):
        m = self.mesh
        nnodes = 10 * m.n * m.n * m.nr
        nelem = 10 * (m.n - 1) * (m.n - 1) * (m.nr - 1) * 2
        if not self.first_write_done:
            os.makedirs(self.dir, exist_ok=True)
            with open(os.path.join(self.dir, "README.md"), "w") as f:
                f.write("# Output of the XDMF writer (visualisation and checkpoints)\n\n"
                        "- `geometry.bin`: node coordinates\n- `topology.bin`: wedge connectivity\n"
                        "- `checkpoint_metadata.bin`: checkpoint metadata\n"
                        "- `step_*.xmf`: XDMF file per write() step (open with Paraview)\n"
                        "- `<name>_*.bin`: binary field data per write() step\n")
            self._geometry().tofile(os.path.join(self.dir, "geometry.bin"))
            self._topology().tofile(os.path.join(self.dir, "topology.bin"))
            with open(os.path.join(self.dir, "checkpoint_metadata.bin"), "wb") as f:
                f.write(self._metadata())
        grid = _XML("Grid", {"Name": "Grid", "GridType": "Uniform"})
        grid.add_child(_XML("Geometry", {"Type": "XYZ"}).add_child(
            _XML("DataItem", {"Format": "Binary", "DataType": "Float", "Precision": str(self.points_bytes),
                              "Endian": "Little", "Dimensions": f"{nnodes} 3"}, "geometry.bin")))
        grid.add_child(_XML("Topology", {"Type": "Wedge", "NumberOfElements": str(nelem)}).add_child(
            _XML("DataItem", {"Format": "Binary", "DataType": "Int", "Precision": str(self.conn_bytes),
                              "Endian": "Little", "Dimensions": str(nelem * 6)}, "topology.bin")))
        for name, get, vec_dim, out_bytes, _ in self._ordered():
            fname = f"{name}_{self.counter}.bin"
            a = _to_file_order(_np(get()), m.n, m.nr, vec_dim)
            a.astype("<f4" if out_bytes == 4 else "<f8").tofile(os.path.join(self.dir, fname))
            dims = str(nnodes) if vec_dim == 1 else f"{nnodes} {vec_dim}"
            grid.add_child(_XML("Attribute", {"Name": name, "AttributeType": "Scalar" if vec_dim == 1 else "Vector",
                                              "Center": "Node"}).add_child(
                _XML("DataItem", {"Format": "Binary", "DataType": "Float", "Precision": str(out_bytes),
                                  "Endian": "Little", "Dimensions": dims}, fname)))
        xml = _XML("Xdmf", {"Version": "2.0"}).add_child(_XML("Domain").add_child(grid))
        with open(os.path.join(self.dir, f"step_{self.counter}.xmf"), "w") as f:
            f.write('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n' + xml.to_string())
        self.counter += 1
        self.first_write_done = True


def read_checkpoint_metadata(directory):
    """read_xdmf_checkpoint_metadata: the metadata as a dict."""
    with open(os.path.join(directory, "checkpoint_metadata.bin"), "rb") as f:
        b = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, b, off)
        off += struct.calcsize(fmt)
        return v

    md = dict(zip(["version", "num_subdomains_lateral", "num_subdomains_radial", "size_x", "size_y", "size_r"],
                  take("<iiiiii")))
    nrad = md["num_subdomains_radial"] * (md["size_r"] - 1) + 1
    md["radii"] = list(take(f"<{nrad}d"))
    if any(r1 <= r0 for r0, r1 in zip(md["radii"], md["radii"][1:])):
        raise ValueError("radii are not sorted in the checkpoint")
    md["grid_scalar_bytes"] = take("<i")[0] if md["version"] > 0 else None
    files = []
    for _ in range(take("<i")[0]):
        ln = take("<i")[0]
        name = b[off:off + ln].decode()
        off += ln
        dtype, nbytes, vec = take("<iii")
        files.append({"name": name, "scalar_data_type": dtype, "scalar_bytes": nbytes, "vec_dim": vec})
    md["grid_data_files"] = files
    md["ordering"] = take("<i")[0]
    if md["ordering"] == 0:
        nsub = 10 * md["num_subdomains_lateral"] ** 2 * md["num_subdomains_radial"]
        md["subdomain_ids"] = list(take(f"<{nsub}q"))
    return md


def read_checkpoint_grid(directory, name, step, mesh):
    """read_xdmf_checkpoint_grid: field `name` of `step` in this library's layout
    (float64 numpy array, scalar [10*n*n*nr] or vector [3*10*n*n*nr])."""
    md = read_checkpoint_metadata(directory)
    if md["version"] not in (0, 1):
        raise ValueError(f"unsupported checkpoint version {md['version']}")
    entry = next((g for g in md["grid_data_files"] if g["name"] == name), None)
    if entry is None:
        raise KeyError(f"no data '{name}' in the checkpoint")
    if (md["size_x"], md["size_y"], md["size_r"]) != (mesh.n, mesh.n, mesh.nr):
        raise ValueError("grid extents do not match the checkpoint")
    if md["ordering"] != 0:
        raise ValueError("checkpoint ordering is not 0")
    if md["num_subdomains_lateral"] != 1 or md["num_subdomains_radial"] != 1:
        raise ValueError("only one subdomain per diamond is supported")
    vec = entry["vec_dim"]
    dt = "<f4" if entry["scalar_bytes"] == 4 else "<f8"
    raw = np.fromfile(os.path.join(directory, f"{name}_{step}.bin"), dtype=dt).astype(np.float64)
    chunk = mesh.n * mesh.n * mesh.nr * vec
    if raw.size != 10 * chunk:
        raise ValueError("data file size does not match the metadata")
    ids = md["subdomain_ids"]
    ordered = np.concatenate([raw[ids.index(global_subdomain_id(d)) * chunk:][:chunk] for d in range(10)])
    return _from_file_order(ordered, mesh.n, mesh.nr, vec).reshape(-1)


def main():
    """Throughput of a checkpoint write + read on the GPU box:
    python -m terraneo_amd.xdmf [--lat 7] [--dir /tmp/xdmf_bench]"""
    import argparse
    import shutil
    import time

    import torch

    from . import api

    ap = argparse.ArgumentParser()
    ap.add_argument("--lat", type=int, default=7)
    ap.add_argument("--dir", default="/tmp/xdmf_bench")
    a = ap.parse_args()
    dev = torch.device("cuda:0")
    m = api.Mesh(a.lat, a.lat, 0.55, 1.0)
    u = torch.randn(3 * m.N, dtype=torch.float64, device=dev)
    eta = torch.rand(m.N, dtype=torch.float64, device=dev)
    shutil.rmtree(a.dir, ignore_errors=True)
    out = XDMFOutput(a.dir, m)
    out.add("velocity", lambda: u, vec_dim=3, out_bytes=8)
    out.add("viscosity", lambda: eta, vec_dim=1, out_bytes=8)
    out.write()  # first step: mesh, topology, metadata
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out.write()
    t1 = time.perf_counter()
    back = read_checkpoint_grid(a.dir, "velocity", 1, m)
    t2 = time.perf_counter()
    assert np.array_equal(back, u.cpu().numpy())
    gb = 4 * m.N * 8 / 1e9
    print(f"xdmf level {a.lat}: step write {gb:.2f} GB in {t1 - t0:.3f} s ({gb / (t1 - t0):.2f} GB/s), "
          f"velocity read {3 * m.N * 8 / 1e9:.2f} GB in {t2 - t1:.3f} s", flush=True)
    shutil.rmtree(a.dir, ignore_errors=True)


if __name__ == "__main__":
    main()
