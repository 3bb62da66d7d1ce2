"""Thin Python handles over the C ABI (plumbing for tests and benchmarks).

Device arrays are torch tensors on the HIP device; their data pointers are
passed straight to the C ABI.  Layouts: scalar fields [10, n, n, nr], vector
fields [3, 10, n, n, nr] (contiguous).
"""
import ctypes as C

import numpy as np

from ._capi import TN_COMM_ADDITIVE, TN_F32, TN_F64, TN_NEUMANN, TN_REPLACE, check, lib


def _stream_ptr(stream=None):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _ptr(t):
    if t is None:
        return None
    assert t.is_contiguous(), "tensor must be contiguous"
    return C.c_void_p(t.data_ptr())


def dtype_code(torch_dtype):
    import torch

    if torch_dtype == torch.float64:
        return TN_F64
    if torch_dtype == torch.float32:
        return TN_F32
    raise ValueError(f"unsupported dtype {torch_dtype}")


class Mesh:
    """One refinement level of the icosahedral thick shell (10 diamonds)."""

    def __init__(self, lat_level, rad_level=None, r_min=0.55, r_max=1.0, radii=None):
        h = C.c_void_p()
        if radii is not None:
            radii = np.ascontiguousarray(radii, dtype=np.float64)
            check(lib().tn_mesh_create_radii(lat_level, len(radii), radii.ctypes.data_as(C.POINTER(C.c_double)),
                                             C.byref(h)))
        else:
            check(lib().tn_mesh_create(lat_level, rad_level, r_min, r_max, C.byref(h)))
        self.h = h
        self.lat_level = lat_level
        n, nr, N = C.c_int(), C.c_int(), C.c_int64()
        check(lib().tn_mesh_dims(h, C.byref(n), C.byref(nr), C.byref(N)))
        self.n, self.nr, self.N = n.value, nr.value, N.value
        self.owned_nodes = lib().tn_mesh_owned_nodes(h)

    @property
    def shape(self):
        return (10, self.n, self.n, self.nr)

    def lateral_coords(self):
        out = np.zeros((10, self.n, self.n, 3))
        check(lib().tn_mesh_lateral_coords(self.h, out.ctypes.data))
        return out

    def radii(self):
        out = np.zeros(self.nr)
        check(lib().tn_mesh_radii(self.h, out.ctypes.data))
        return out

    def owned_mask(self):
        out = np.zeros(self.shape, np.uint8)
        check(lib().tn_mesh_owned_mask(self.h, out.ctypes.data))
        return out

    def interface_info(self):
        g, m = C.c_int(), C.c_int()
        check(lib().tn_mesh_interface_info(self.h, C.byref(g), C.byref(m)))
        return g.value, m.value

    # ---- level-1 operations on device tensors
    def interface_sum(self, v, ncomp, stream=None):
        check(lib().tn_interface_sum(self.h, dtype_code(v.dtype), ncomp, _ptr(v), _stream_ptr(stream)))

    def dot(self, a, b, ncomp, stream=None):
        out = C.c_double()
        check(lib().tn_dot(self.h, dtype_code(a.dtype), ncomp, _ptr(a), _ptr(b), C.byref(out), _stream_ptr(stream)))
        return out.value

    def lincomb(self, y, coeffs, xs, c0=0.0, ncomp=3, stream=None):
        nx = len(xs)
        c = (C.c_double * max(nx, 1))(*coeffs)
        x = (C.c_void_p * max(nx, 1))(*[t.data_ptr() for t in xs])
        check(lib().tn_lincomb(self.h, dtype_code(y.dtype), ncomp, _ptr(y), nx, c, x, c0, _stream_ptr(stream)))

    def __del__(self):
        try:
            if self.h:
                lib().tn_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Eps:
    """epsilon-div-div operator (EpsilonDivDivKerngen drop-in)."""

    def __init__(self, mesh, eta, bc_cmb=TN_NEUMANN, bc_surf=TN_NEUMANN):
        self.mesh = mesh
        self.eta = eta  # keep alive: the operator holds a view
        self.dtype = dtype_code(eta.dtype)
        h = C.c_void_p()
        check(lib().tn_eps_create(mesh.h, self.dtype, _ptr(eta), bc_cmb, bc_surf, C.byref(h)))
        self.h = h

    def set_modes(self, apply_mode=TN_REPLACE, comm_mode=TN_COMM_ADDITIVE):
        check(lib().tn_eps_set_modes(self.h, apply_mode, comm_mode))

    def set_diagonal(self, diagonal):
        check(lib().tn_eps_set_diagonal(self.h, int(bool(diagonal))))

    def set_boundary_conditions(self, bc_cmb, bc_surf):
        check(lib().tn_eps_set_boundary_conditions(self.h, bc_cmb, bc_surf))

    def set_eta(self, eta):
        self.eta = eta
        check(lib().tn_eps_set_eta(self.h, _ptr(eta)))

    def apply(self, src, dst, stream=None):
        check(lib().tn_eps_apply(self.h, _ptr(src), _ptr(dst), _stream_ptr(stream)))
        return dst

    def __del__(self):
        try:
            if self.h:
                lib().tn_eps_destroy(self.h)
                self.h = None
        except Exception:
            pass
