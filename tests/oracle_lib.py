"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.  Arrays use the layouts documented in oracle/oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
PORT_SO = os.path.join(ORACLE_DIR, "_build", "libcpuport.so")

NEUMANN, DIRICHLET, FREESLIP = 0, 1, 2

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


_port = None


def cpu_port():
    """the benchmark's CPU baseline (oracle/cpu_port.c): matrix-free OpenMP apply"""
    global _port
    if _port is None:
        if not os.path.exists(PORT_SO):
            build()
        L = C.CDLL(PORT_SO)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.cpu_eps_apply.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, C.c_int, C.c_int, ip, C.c_int]
        _port = L
    return _port


def cpu_port_apply(mesh, eta, x, bc_cmb, bc_surf, y=None):
    """y = A x by the CPU baseline port (Dirichlet / Neumann shells)"""
    y = np.zeros_like(x) if y is None else y
    cpu_port().cpu_eps_apply(mesh.n, mesh.nr, np.ascontiguousarray(mesh.coords.reshape(-1)), mesh.radii,
                             np.ascontiguousarray(eta, dtype=np.float64).reshape(-1),
                             np.ascontiguousarray(x, dtype=np.float64).reshape(-1), y, bc_cmb, bc_surf,
                             np.ascontiguousarray(mesh.groups.reshape(-1)), mesh.ngroups)
    return y


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.orc_lateral_coords.argtypes = [C.c_int, dp]
        L.orc_uniform_radii.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.orc_lateral_groups.argtypes = [C.c_int, dp, ip]
        L.orc_lateral_groups.restype = C.c_int
        L.orc_mesh_create.argtypes = [C.c_int, C.c_int, dp]
        L.orc_mesh_create.restype = vp
        L.orc_mesh_destroy.argtypes = [vp]
        L.orc_eps_local_matrix.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_eps_create.argtypes = [vp, dp, C.c_int, C.c_int]
        L.orc_eps_create.restype = vp
        L.orc_eps_destroy.argtypes = [vp]
        L.orc_eps_apply.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        L.orc_eps_set_stored.argtypes = [vp, dp]
        L.orc_eps_local_matrix_k.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_eps_set_cell_viscosity.argtypes = [vp, dp]
        L.orc_eps_set_penalty_once.argtypes = [vp, C.c_int]
        L.orc_eps_set_penalty_epsilon.argtypes = [vp, C.c_double]
        L.orc_eps_grad_log_cell_viscosity.argtypes = [vp, dp, dp, dp]
        L.orc_eps_penalty_modes.argtypes = [vp, dp]
        L.orc_eps_penalty_modes.restype = C.c_int
        L.orc_div_apply.argtypes = [vp, vp, C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_grad_apply_add.argtypes = [vp, vp, C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_vector_mass_apply.argtypes = [vp, dp, dp]
        L.orc_kmass_lumped.argtypes = [vp, dp, dp]
        L.orc_prolongate_add.argtypes = [vp, vp, C.c_int, dp, dp]
        L.orc_restrict.argtypes = [vp, vp, C.c_int, dp, dp]
        L.orc_gca_build.argtypes = [vp, vp, dp]
        L.orc_gca_build_geometric.argtypes = [vp, vp, dp]
        L.orc_dot_owned.argtypes = [vp, C.c_int, dp, dp]
        L.orc_dot_owned.restype = C.c_double
        L.orc_interface_sum.argtypes = [C.c_int, C.c_int, C.c_int, ip, C.c_int, dp]
        L.orc_ownership.argtypes = [C.c_int, C.c_int, ip, C.c_int, up]
        _lib = L
    return _lib


def lateral_coords(lat_level):
    n = (1 << lat_level) + 1
    c = np.zeros((10, n, n, 3))
    lib().orc_lateral_coords(lat_level, c)
    return c


def uniform_radii(nr, rmin, rmax):
    r = np.zeros(nr)
    lib().orc_uniform_radii(nr, rmin, rmax, r)
    return r


class Mesh:
    """Oracle mesh: 10 diamonds x n x n lateral nodes x nr radial nodes."""

    def __init__(self, lat_level, nr=None, rmin=0.55, rmax=1.0, radii=None):
        if radii is None:
            radii = uniform_radii(nr, rmin, rmax)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        self.lat_level = lat_level
        self.n = (1 << lat_level) + 1
        self.nr = len(radii)
        self.radii = radii
        self.h = lib().orc_mesh_create(lat_level, self.nr, radii)
        self.coords = lateral_coords(lat_level)
        self.groups = np.zeros((10, self.n, self.n), np.int32)
        self.ngroups = lib().orc_lateral_groups(self.n, self.coords, self.groups)
        self.owned = np.zeros(self.N, np.uint8)
        lib().orc_ownership(self.n, self.nr, self.groups, self.ngroups, self.owned)

    @property
    def N(self):
        return 10 * self.n * self.n * self.nr

    @property
    def shape(self):
        return (10, self.n, self.n, self.nr)

    def node_xyz(self):
        """Physical coordinates [10,n,n,nr,3]."""
        return self.coords[:, :, :, None, :] * self.radii[None, None, None, :, None]

    def interface_sum(self, v, ncomp):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_interface_sum(self.n, self.nr, ncomp, self.groups, self.ngroups, v)
        return v

    def dot(self, a, b, ncomp):
        return lib().orc_dot_owned(self.h, ncomp, np.ascontiguousarray(a), np.ascontiguousarray(b))

    def __del__(self):
        try:
            lib().orc_mesh_destroy(self.h)
        except Exception:
            pass


class Eps:
    """Oracle epsilon-div-div operator (assembled local-matrix formulation)."""

    def __init__(self, mesh, k, bc_cmb=NEUMANN, bc_surf=NEUMANN, kc=None):
        """k: nodal viscosity; kc (optional): per hex cell viscosity, overriding k."""
        self.m = mesh
        if k is None:
            k = np.ones(10 * mesh.n * mesh.n * mesh.nr)
        self.k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1)
        self.h = lib().orc_eps_create(mesh.h, self.k, bc_cmb, bc_surf)
        self.kc = None
        if kc is not None:
            self.kc = np.ascontiguousarray(kc, dtype=np.float64).reshape(-1)
            lib().orc_eps_set_cell_viscosity(self.h, self.kc)

    def grad_log_cell_viscosity(self, u, lam):
        """g[cell] = -kc[cell] sum_wedges (T lam)_e^T A_e(1) u_e."""
        assert self.kc is not None
        g = np.zeros(10 * (self.m.n - 1) ** 2 * (self.m.nr - 1))
        lib().orc_eps_grad_log_cell_viscosity(self.h, np.ascontiguousarray(u, dtype=np.float64).reshape(-1),
                                              np.ascontiguousarray(lam, dtype=np.float64).reshape(-1), g)
        return g

    def apply(self, x, diagonal=False, comm=True):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.zeros_like(x)
        lib().orc_eps_apply(self.h, x, y, int(diagonal), int(comm))
        return y

    def local_matrix(self, d, x, y, r, w):
        A = np.zeros(324)
        lib().orc_eps_local_matrix(self.m.h, self.k, d, x, y, r, w, A)
        return A.reshape(18, 18)

    def set_penalty_once(self, once=True):
        """fs/fs penalty at the owned copy only (symmetric operator) instead of every copy"""
        lib().orc_eps_set_penalty_once(self.h, int(bool(once)))

    def set_penalty_epsilon(self, eps):
        lib().orc_eps_set_penalty_epsilon(self.h, float(eps))

    def set_stored(self, mats):
        lib().orc_eps_set_stored(self.h, np.ascontiguousarray(mats, dtype=np.float64).reshape(-1))

    def penalty_modes(self):
        out = np.zeros(9 * self.m.N)
        if not lib().orc_eps_penalty_modes(self.h, out):
            return None
        return out.reshape(3, 3 * self.m.N)

    def __del__(self):
        try:
            lib().orc_eps_destroy(self.h)
        except Exception:
            pass


def div_apply(mf, mc, u, bc_cmb, bc_surf, comm=True):
    p = np.zeros(mc.N)
    lib().orc_div_apply(mf.h, mc.h, bc_cmb, bc_surf, np.ascontiguousarray(u, dtype=np.float64), p, int(comm))
    return p


def grad_apply_add(mf, mc, p, u, bc_cmb, bc_surf, comm=True):
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    lib().orc_grad_apply_add(mf.h, mc.h, bc_cmb, bc_surf, np.ascontiguousarray(p, dtype=np.float64), u, int(comm))
    return u


def vector_mass(m, x):
    y = np.zeros(3 * m.N)
    lib().orc_vector_mass_apply(m.h, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def kmass_lumped(m, k):
    d = np.zeros(m.N)
    lib().orc_kmass_lumped(m.h, np.ascontiguousarray(k, dtype=np.float64).reshape(-1), d)
    return d


def prolongate_add(mc, mf, xc, xf, ncomp=3):
    xf = np.ascontiguousarray(xf, dtype=np.float64).copy()
    lib().orc_prolongate_add(mc.h, mf.h, ncomp, np.ascontiguousarray(xc, dtype=np.float64), xf)
    return xf


def restrict(mf, mc, xf, ncomp=3):
    xc = np.zeros(ncomp * mc.N)
    lib().orc_restrict(mf.h, mc.h, ncomp, np.ascontiguousarray(xf, dtype=np.float64), xc)
    return xc


def gca_build(fine_eps, mc, geometric=False):
    """Stored coarse matrices; geometric=True: exact Galerkin variant (test only)."""
    cells = 10 * (mc.n - 1) ** 2 * (mc.nr - 1)
    out = np.zeros(cells * 2 * 324)
    (lib().orc_gca_build_geometric if geometric else lib().orc_gca_build)(fine_eps.h, mc.h, out)
    return out


def _coo_sizes_eps(m):
    return 10 * (m.n - 1) ** 2 * (m.nr - 1) * 2 * 324


def eps_assemble(eps):
    """scipy CSR of the operator over unique dofs 3*(group*nr + r) + comp (no penalty)."""
    import scipy.sparse as sp

    L = lib()
    L.orc_eps_assemble_coo.argtypes = [C.c_void_p] + [np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")] * 2 + \
        [np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
    L.orc_eps_assemble_coo.restype = C.c_int64
    m = eps.m
    cnt = _coo_sizes_eps(m)
    r = np.empty(cnt, np.int64)
    c = np.empty(cnt, np.int64)
    v = np.empty(cnt)
    L.orc_eps_assemble_coo(eps.h, r, c, v)
    nu = 3 * m.ngroups * m.nr
    return sp.csr_matrix((v, (r, c)), shape=(nu, nu))


def div_assemble(mf, mc, bc_cmb, bc_surf):
    import scipy.sparse as sp

    L = lib()
    L.orc_div_assemble_coo.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int] + \
        [np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")] * 2 + \
        [np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
    L.orc_div_assemble_coo.restype = C.c_int64
    cnt = 10 * (mf.n - 1) ** 2 * (mf.nr - 1) * 2 * 108
    r = np.empty(cnt, np.int64)
    c = np.empty(cnt, np.int64)
    v = np.empty(cnt)
    L.orc_div_assemble_coo(mf.h, mc.h, bc_cmb, bc_surf, r, c, v)
    return sp.csr_matrix((v, (r, c)), shape=(mc.ngroups * mc.nr, 3 * mf.ngroups * mf.nr))


def to_unique(m, v, ncomp):
    """Stored (consistent) field [ncomp*N] -> unique dofs [ (group*nr + r)*ncomp + comp ]."""
    v = np.asarray(v).reshape(ncomp, 10 * m.n * m.n, m.nr)
    g = m.groups.reshape(-1)
    out = np.zeros((m.ngroups, m.nr, ncomp))
    out[g] = np.moveaxis(v, 0, -1)
    return out.reshape(-1)


def from_unique(m, u, ncomp):
    u = np.asarray(u).reshape(m.ngroups, m.nr, ncomp)
    g = m.groups.reshape(-1)
    return np.ascontiguousarray(np.moveaxis(u[g], -1, 0).reshape(-1))
