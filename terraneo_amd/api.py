"""Thin Python handles over the C ABI (plumbing for tests and benchmarks).

Device arrays are torch tensors on the HIP device; their data pointers are
passed straight to the C ABI.  Layouts: scalar fields [10, n, n, nr], vector
fields [3, 10, n, n, nr] (contiguous).
"""
import ctypes as C

import numpy as np

from ._capi import (TN_ADD, TN_COMM_ADDITIVE, TN_F32, TN_F64, TN_NEUMANN, TN_REPLACE, GmgOptions, KrylovOptions,
                    SolveInfo, check, lib)


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

    def invert_entries(self, v, ncomp=3, stream=None):
        check(lib().tn_invert_entries(self.h, dtype_code(v.dtype), ncomp, _ptr(v), _stream_ptr(stream)))

    def scale_with_vector(self, y, d, ncomp=3, stream=None):
        check(lib().tn_scale_with_vector(self.h, dtype_code(y.dtype), ncomp, _ptr(y), _ptr(d), _stream_ptr(stream)))

    def randomize(self, y, seed, ncomp=3, stream=None):
        check(lib().tn_randomize(self.h, dtype_code(y.dtype), ncomp, _ptr(y), seed, _stream_ptr(stream)))

    def max_abs_entry(self, v, ncomp=3, stream=None):
        out = C.c_double()
        check(lib().tn_max_abs_entry(self.h, dtype_code(v.dtype), ncomp, _ptr(v), C.byref(out), _stream_ptr(stream)))
        return out.value

    def has_nan_or_inf(self, v, ncomp=3, stream=None):
        out = C.c_int()
        check(lib().tn_has_nan_or_inf(self.h, dtype_code(v.dtype), ncomp, _ptr(v), C.byref(out), _stream_ptr(stream)))
        return bool(out.value)

    def __del__(self):
        try:
            if self.h:
                lib().tn_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Eps:
    """epsilon-div-div operator (EpsilonDivDivKerngen drop-in)."""

    def __init__(self, mesh, eta, bc_cmb=TN_NEUMANN, bc_surf=TN_NEUMANN, eta_cell=None):
        """eta: nodal viscosity [10][n][n][nr]; eta_cell (optional): per hex cell
        viscosity [10][n-1][n-1][nr-1], which then replaces the nodal one."""
        self.mesh = mesh
        self.eta = eta  # keep alive: the operator holds a view
        self.eta_cell = eta_cell
        ref = eta if eta is not None else eta_cell
        self.dtype = dtype_code(ref.dtype)
        h = C.c_void_p()
        check(lib().tn_eps_create(mesh.h, self.dtype, _ptr(eta) if eta is not None else None, bc_cmb, bc_surf,
                                  C.byref(h)))
        self.h = h
        if eta_cell is not None:
            check(lib().tn_eps_set_cell_eta(self.h, _ptr(eta_cell)))

    def set_cell_eta(self, eta_cell):
        self.eta_cell = eta_cell
        check(lib().tn_eps_set_cell_eta(self.h, _ptr(eta_cell) if eta_cell is not None else None))

    def grad_log_cell_eta(self, u, lam, grad, stream=None):
        """grad[cell] = -eta_cell (T lam)^T A_cell(1) u (adjoint viscosity gradient)."""
        check(lib().tn_eps_grad_log_cell_eta(self.h, _ptr(u), _ptr(lam), _ptr(grad), _stream_ptr(stream)))
        return grad

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

    # ---- Galerkin coarse operators
    def gca_from(self, fine, store_dtype=TN_F64, stream=None, bracket="reference"):
        """Make this operator's stored matrices the Galerkin coarsening of `fine`
        (bracket "reference": the reference's local prolongation, "geometric":
        the exact Galerkin product)."""
        self.fine_ref = fine
        b = {"reference": 0, "geometric": 1}[bracket]
        check(lib().tn_eps_gca_from_bracket(self.h, fine.h, store_dtype, b, _stream_ptr(stream)))

    def gca_select(self, coarsest, flags, stream=None):
        """Selective GCA: rediscretized matrices on cells whose coarsest ancestor is not flagged."""
        check(lib().tn_eps_gca_select(self.h, coarsest.h, _ptr(flags), _stream_ptr(stream)))

    def local_matrix(self, sd, x, y, r, w):
        out = np.zeros(324)
        check(lib().tn_eps_get_local_matrix(self.h, sd, x, y, r, w, out.ctypes.data))
        return out.reshape(18, 18)

    def set_local_matrix(self, sd, x, y, r, w, mat):
        a = np.ascontiguousarray(mat, dtype=np.float64).reshape(324)
        check(lib().tn_eps_set_local_matrix(self.h, sd, x, y, r, w, a.ctypes.data))

    def set_stored_matrices(self, store_dtype=TN_F64, stream=None):
        """stored-matrix path with every cell's own matrices (set_stored_matrix_mode Full)"""
        check(lib().tn_eps_set_stored_matrices(self.h, store_dtype, _stream_ptr(stream)))

    def set_penalty_epsilon(self, eps):
        check(lib().tn_eps_set_penalty_epsilon(self.h, eps))

    def set_penalty_once(self, once=True):
        """fs/fs rotation penalty at the owned copy only (symmetric operator) instead of
        at every node copy (the reference's rule)"""
        check(lib().tn_eps_set_penalty_once(self.h, int(bool(once))))


# ---------------------------------------------------------------- mesh-level helpers
def vector_mass(mesh, x, y, stream=None):
    check(lib().tn_vector_mass_apply(mesh.h, dtype_code(x.dtype), _ptr(x), _ptr(y), _stream_ptr(stream)))
    return y


def kmass_lumped(mesh, k, d, stream=None):
    check(lib().tn_kmass_lumped(mesh.h, dtype_code(k.dtype), _ptr(k), _ptr(d), _stream_ptr(stream)))
    return d


def prolongate_add(coarse, fine, xc, xf, ncomp=3, stream=None):
    check(lib().tn_prolongate_add(coarse.h, fine.h, dtype_code(xc.dtype), ncomp, _ptr(xc), _ptr(xf),
                                  _stream_ptr(stream)))
    return xf


def restrict(fine, coarse, xf, xc, ncomp=3, stream=None):
    check(lib().tn_restrict(fine.h, coarse.h, dtype_code(xf.dtype), ncomp, _ptr(xf), _ptr(xc), _stream_ptr(stream)))
    return xc


def gca_collect_elements(fine, coarsest, k, flags, stream=None):
    """GCAElementsCollector: flags (scalar field on `coarsest`) = 1 where a descendant fine cell has
    |grad k| > 10 (reference coordinates), else 0."""
    check(lib().tn_gca_collect_elements(fine.h, coarsest.h, dtype_code(k.dtype), _ptr(k), _ptr(flags),
                                        _stream_ptr(stream)))
    return flags


def prolongate_linear(coarse, fine, xc, xf, ncomp=3, add=False, vec_early_return=False, stream=None):
    """Geometry-aware linear prolongation (ProlongationLinear / ProlongationVecLinear)."""
    check(lib().tn_prolongate_linear(coarse.h, fine.h, dtype_code(xc.dtype), ncomp, _ptr(xc), _ptr(xf),
                                     TN_ADD if add else TN_REPLACE, int(bool(vec_early_return)), _stream_ptr(stream)))
    return xf


def restrict_linear(fine, coarse, xf, xc, ncomp=3, vec_early_return=False, stream=None):
    """Transpose of the linear prolongation over owned fine nodes (RestrictionLinear / RestrictionVecLinear)."""
    check(lib().tn_restrict_linear(fine.h, coarse.h, dtype_code(xf.dtype), ncomp, _ptr(xf), _ptr(xc),
                                   int(bool(vec_early_return)), _stream_ptr(stream)))
    return xc


def boundary_assign(mesh, dst, src, which, ncomp=3, stream=None):
    check(lib().tn_boundary_assign(mesh.h, dtype_code(dst.dtype), ncomp, _ptr(dst), _ptr(src), which,
                                   _stream_ptr(stream)))


def freeslip_project(mesh, v, which, stream=None):
    check(lib().tn_freeslip_project(mesh.h, dtype_code(v.dtype), _ptr(v), which, _stream_ptr(stream)))


def stokes_rhs_from_temperature(K, temperature, rayleigh, fu, fp, stream=None):
    """Stokes right-hand side of one mantle-circulation time step (mantlecirculation.cpp:876-898):
    fu = M (Ra T n) with K's velocity BCs enforced, fp = 0."""
    check(lib().tn_stokes_rhs_from_temperature(K.h, _ptr(temperature), float(rayleigh), _ptr(fu), _ptr(fp),
                                               _stream_ptr(stream)))
    return fu, fp


def pressure_normalize(mesh, p, stream=None):
    """p -= owned mean of p (mantlecirculation.cpp:926-932)."""
    check(lib().tn_pressure_normalize(mesh.h, dtype_code(p.dtype), _ptr(p), _stream_ptr(stream)))
    return p


class Stokes:
    """EpsDivDivStokes drop-in: [A B^T; B 0] on (velocity mesh, pressure mesh)."""

    def __init__(self, vel, pre, A):
        self.vel, self.pre, self.A = vel, pre, A
        h = C.c_void_p()
        check(lib().tn_stokes_create(vel.h, pre.h, A.h, C.byref(h)))
        self.h = h

    def apply(self, u, p, yu, yp, stream=None):
        check(lib().tn_stokes_apply(self.h, _ptr(u), _ptr(p), _ptr(yu), _ptr(yp), _stream_ptr(stream)))

    def div(self, u, p, stream=None):
        check(lib().tn_div_apply(self.h, _ptr(u), _ptr(p), _stream_ptr(stream)))
        return p

    def grad(self, p, u, apply_mode=TN_REPLACE, comm_mode=TN_COMM_ADDITIVE, stream=None):
        check(lib().tn_grad_apply(self.h, _ptr(p), _ptr(u), apply_mode, comm_mode, _stream_ptr(stream)))
        return u

    def __del__(self):
        try:
            if self.h:
                lib().tn_stokes_destroy(self.h)
                self.h = None
        except Exception:
            pass


def gmg_options(cheby_order=2, cheby_iterations=3, power_iterations=10, num_cycles=1, coarse_max_iterations=1000,
                coarse_rtol=1e-6, coarse_atol=1e-16, seed=12345, coarse_solver="pcg",
                freeslip_projected_transfer=False, freeslip_block_jacobi=False):
    cs = {"pcg": 0, "dense": 1}[coarse_solver]
    return GmgOptions(cheby_order, cheby_iterations, power_iterations, num_cycles, coarse_max_iterations, coarse_rtol,
                      coarse_atol, seed, cs, int(bool(freeslip_projected_transfer)), int(bool(freeslip_block_jacobi)))


class Gmg:
    """Multigrid V-cycle over epsilon operators ordered coarse -> fine."""

    def __init__(self, ops, options=None, stream=None):
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[o.h.value for o in self.ops])
        self.opts = options or gmg_options()
        h = C.c_void_p()
        check(lib().tn_gmg_create(len(self.ops), arr, C.byref(self.opts), _stream_ptr(stream), C.byref(h)))
        self.h = h

    def solve(self, x, b, stream=None):
        check(lib().tn_gmg_solve(self.h, _ptr(x), _ptr(b), _stream_ptr(stream)))
        return x

    def lambda_max(self, level):
        v = C.c_double()
        check(lib().tn_gmg_get_lambda(self.h, level, C.byref(v)))
        return v.value

    def set_lambda_max(self, level, value):
        check(lib().tn_gmg_set_lambda(self.h, level, value))

    def __del__(self):
        try:
            if self.h:
                lib().tn_gmg_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _info_dict(info, hist):
    n = min(info.history_len, len(hist))
    return {"iterations": info.iterations, "initial_residual": info.initial_residual,
            "final_residual": info.final_residual, "seconds": info.seconds, "history": hist[:n].copy(),
            "restarts": info.restarts, "prec_skipped": info.prec_skipped}


def pcg_solve(A, x, b, prec=None, max_iterations=100, rtol=1e-8, atol=1e-30, stream=None):
    o = KrylovOptions(max_iterations, rtol, atol, 30, 0)
    info = SolveInfo()
    hist = np.zeros(max_iterations + 2)
    check(lib().tn_pcg_solve(A.h, prec.h if prec is not None else None, _ptr(x), _ptr(b), C.byref(o), C.byref(info),
                             hist.ctypes.data, len(hist), _stream_ptr(stream)))
    return _info_dict(info, hist)


def stokes_solve(K, method, prec_u, schur_diag, u, p, bu, bp, max_iterations=100, rtol=1e-8, atol=1e-30,
                 restart=30, pressure_mean_free=False, stream=None):
    o = KrylovOptions(max_iterations, rtol, atol, restart, int(bool(pressure_mean_free)))
    info = SolveInfo()
    hist = np.zeros(max_iterations + 2)
    check(lib().tn_stokes_solve(K.h, method, prec_u.h, _ptr(schur_diag), _ptr(u), _ptr(p), _ptr(bu), _ptr(bp),
                                C.byref(o), C.byref(info), hist.ctypes.data, len(hist), _stream_ptr(stream)))
    return _info_dict(info, hist)


# ---------------------------------------------------------------- energy step (hex finite volumes)
class FV:
    """FV state on a velocity mesh (cell centres, ghost map, FCT work arrays):
    the reference's FVFCTBuffers + initialize_cell_centers.  FV scalars are
    FP64 torch tensors of shape fv_shape = (10, n+1, n+1, nr+1) with one
    ghost layer; the velocity is the FP64 nodal field [3, 10, n, n, nr]."""

    def __init__(self, mesh):
        self.mesh = mesh
        h = C.c_void_p()
        check(lib().tn_fv_create(mesh.h, C.byref(h)))
        self.h = h
        n, nr = mesh.n, mesh.nr
        self.fv_shape = (10, n + 1, n + 1, nr + 1)

    def empty(self, fill=0.0):
        import torch

        return torch.full(self.fv_shape, fill, dtype=torch.float64, device="cuda")

    def cell_centers(self, stream=None):
        import torch

        c = torch.empty(self.fv_shape + (3,), dtype=torch.float64, device="cuda")
        check(lib().tn_fv_cell_centers(self.h, _ptr(c), _stream_ptr(stream)))
        return c

    def ghost_update(self, f, ncomp=1, stream=None):
        check(lib().tn_fv_ghost_update(self.h, _ptr(f), ncomp, _stream_ptr(stream)))
        return f

    def dt_stable(self, vel, kappa=0.0, stream=None):
        out = C.c_double()
        check(lib().tn_fv_dt_stable(self.h, _ptr(vel), float(kappa), C.byref(out), _stream_ptr(stream)))
        return out.value

    def apply_dirichlet(self, f, cmb=None, surf=None, stream=None):
        check(lib().tn_fv_apply_dirichlet(self.h, _ptr(f), int(cmb is not None), float(cmb or 0.0),
                                          int(surf is not None), float(surf or 0.0), _stream_ptr(stream)))
        return f

    def predictor(self, T, vel, dt, kappa=0.0, source=None, subtract_divergence=True, stream=None):
        check(lib().tn_fct_predictor(self.h, _ptr(T), _ptr(vel), float(dt), float(kappa),
                                     _ptr(source) if source is not None else None, int(bool(subtract_divergence)),
                                     _stream_ptr(stream)))

    def limiter(self, stream=None):
        check(lib().tn_fct_limiter(self.h, _stream_ptr(stream)))

    def correction(self, T, stream=None):
        check(lib().tn_fct_correction(self.h, _ptr(T), _stream_ptr(stream)))
        return T

    def buffers(self):
        """host copies (numpy) of the state's T_L, antidiffusive fluxes [6, ...], R+, R-"""
        ptrs = [C.c_void_p() for _ in range(4)]
        check(lib().tn_fct_buffers(self.h, *[C.byref(p) for p in ptrs]))
        check(lib().tn_sync())
        out = []
        for p, k in zip(ptrs, (1, 6, 1, 1)):
            a = np.empty(((k,) if k > 1 else ()) + self.fv_shape)
            check(lib().tn_memcpy_d2h(a.ctypes.data, p, a.nbytes))
            out.append(a)
        return out

    def explicit_step(self, T, vel, dt, kappa=0.0, source=None, subtract_divergence=True, cmb=None, surf=None,
                      stream=None):
        check(lib().tn_fct_explicit_step(self.h, _ptr(T), _ptr(vel), float(dt), float(kappa),
                                         _ptr(source) if source is not None else None,
                                         int(bool(subtract_divergence)), int(cmb is not None), float(cmb or 0.0),
                                         int(surf is not None), float(surf or 0.0), _stream_ptr(stream)))
        return T

    def upwind_step(self, T, vel, dt, kappa=0.0, source=None, subtract_divergence=True, stream=None):
        check(lib().tn_fct_upwind_step(self.h, _ptr(T), _ptr(vel), float(dt), float(kappa),
                                       _ptr(source) if source is not None else None, int(bool(subtract_divergence)),
                                       _stream_ptr(stream)))
        return T

    def to_fe_rhs(self, Tfv, b, stream=None):
        check(lib().tn_fv_to_fe_rhs(self.h, _ptr(Tfv), _ptr(b), _stream_ptr(stream)))
        return b

    def to_fe_lumped(self, Tfv, Tfe, tmp, stream=None):
        """bound-preserving FV -> nodal transfer b / D through the lumped mass (tmp receives D)"""
        check(lib().tn_fv_to_fe_lumped(self.h, _ptr(Tfv), _ptr(Tfe), _ptr(tmp), _stream_ptr(stream)))
        return Tfe

    def from_fe(self, Tfe, Tfv, stream=None):
        """FE -> FV cell averages, lateral ghost layers updated"""
        check(lib().tn_fe_to_fv(self.h, _ptr(Tfe), _ptr(Tfv), _stream_ptr(stream)))
        return Tfv

    def to_fe(self, Tfv, Tfe, max_iterations=1000, rtol=1e-12, atol=1e-12, stream=None):
        """L2 projection FV -> nodal (CG on the mass matrix from Tfe's content)"""
        o = KrylovOptions(max_iterations, rtol, atol, 30, 0)
        info = SolveInfo()
        hist = np.zeros(max_iterations + 2)
        check(lib().tn_fv_to_fe(self.h, _ptr(Tfv), _ptr(Tfe), C.byref(o), C.byref(info), hist.ctypes.data,
                                len(hist), _stream_ptr(stream)))
        return _info_dict(info, hist)

    def __del__(self):
        try:
            if self.h:
                lib().tn_fv_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------- XDMF output (C ABI writer)
class XDMFWriter:
    """tn_xdmf_*: the reference's XDMF / checkpoint writer over device fields
    (torch tensors in this library's layout, kept alive while registered)."""

    def __init__(self, mesh, directory, points_bytes=4, connectivity_bytes=8):
        self.mesh = mesh
        self._keep = []
        h = C.c_void_p()
        check(lib().tn_xdmf_create(mesh.h, str(directory).encode(), points_bytes, connectivity_bytes, C.byref(h)))
        self.h = h

    def add(self, name, field, vec_dim=1, out_bytes=4):
        self._keep.append(field)
        check(lib().tn_xdmf_add(self.h, name.encode(), dtype_code(field.dtype), vec_dim, out_bytes, _ptr(field)))

    def set_write_counter(self, step):
        check(lib().tn_xdmf_set_write_counter(self.h, int(step)))

    def write(self, stream=None):
        check(lib().tn_xdmf_write(self.h, _stream_ptr(stream)))

    def __del__(self):
        try:
            if self.h:
                lib().tn_xdmf_destroy(self.h)
                self.h = None
        except Exception:
            pass


def read_checkpoint_grid(mesh, directory, name, step, vec_dim=1, dtype=None, stream=None):
    """tn_xdmf_read_checkpoint_grid: a checkpointed field as a device tensor"""
    import torch

    dtype = dtype or torch.float64
    out = torch.empty(vec_dim * mesh.N, dtype=dtype, device="cuda")
    check(lib().tn_xdmf_read_checkpoint_grid(mesh.h, str(directory).encode(), name.encode(), int(step),
                                             dtype_code(dtype), vec_dim, _ptr(out), _stream_ptr(stream)))
    return out
