"""ctypes declarations for include/terraneo_amd.h.

Loading fails loudly when the shared library has not been built: there is no
CPU fallback for any compute entry point.
"""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libterraneo_amd.so")

TN_F64, TN_F32 = 0, 1
TN_NEUMANN, TN_DIRICHLET, TN_FREESLIP = 0, 1, 2
TN_REPLACE, TN_ADD = 0, 1
TN_COMM_SKIP, TN_COMM_ADDITIVE = 0, 1
TN_MINRES_BLOCKDIAG, TN_FGMRES_BLOCKTRI, TN_FGMRES_BLOCKDIAG = 0, 1, 2
TN_COARSE_PCG, TN_COARSE_DENSE = 0, 1
TN_GCA_REFERENCE, TN_GCA_GEOMETRIC = 0, 1


class GmgOptions(C.Structure):
    _fields_ = [("cheby_order", C.c_int), ("cheby_iterations", C.c_int), ("power_iterations", C.c_int),
                ("num_cycles", C.c_int), ("coarse_max_iterations", C.c_int), ("coarse_rtol", C.c_double),
                ("coarse_atol", C.c_double), ("seed", C.c_uint64), ("coarse_solver", C.c_int), ("freeslip_projected_transfer", C.c_int), ("freeslip_block_jacobi", C.c_int)]


class KrylovOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("rtol", C.c_double), ("atol", C.c_double), ("restart", C.c_int),
                ("pressure_mean_free", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("seconds", C.c_double), ("history_len", C.c_int), ("restarts", C.c_int),
                ("prec_skipped", C.c_int)]


class TNError(RuntimeError):
    pass


_lib = None

vp = C.c_void_p
i32, i64, f64 = C.c_int, C.c_int64, C.c_double
pvp = C.POINTER(C.c_void_p)

# name: (restype, [argtypes])
SIGNATURES = {
    "tn_last_error": (C.c_char_p, []),
    "tn_version": (i32, []),
    "tn_malloc": (i32, [pvp, C.c_size_t]),
    "tn_free": (i32, [vp]),
    "tn_memcpy_h2d": (i32, [vp, vp, C.c_size_t]),
    "tn_memcpy_d2h": (i32, [vp, vp, C.c_size_t]),
    "tn_sync": (i32, []),
    "tn_set_wait_limit": (i32, [i32]),
    "tn_mesh_create": (i32, [i32, i32, f64, f64, pvp]),
    "tn_mesh_create_subdomains": (i32, [i32, i32, f64, f64, i32, i32, pvp]),
    "tn_mesh_create_radii": (i32, [i32, i32, C.POINTER(f64), pvp]),
    "tn_mesh_destroy": (i32, [vp]),
    "tn_mesh_dims": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)]),
    "tn_mesh_owned_nodes": (i64, [vp]),
    "tn_mesh_lateral_coords": (i32, [vp, vp]),
    "tn_mesh_radii": (i32, [vp, vp]),
    "tn_mesh_owned_mask": (i32, [vp, vp]),
    "tn_mesh_interface_info": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "tn_interface_sum": (i32, [vp, i32, i32, vp, vp]),
    "tn_lincomb": (i32, [vp, i32, i32, vp, i32, C.POINTER(f64), pvp, f64, vp]),
    "tn_dot": (i32, [vp, i32, i32, vp, vp, C.POINTER(f64), vp]),
    "tn_dot_noalloc": (i32, [vp, i32, i32, vp, vp, C.POINTER(f64), vp]),
    "tn_invert_entries": (i32, [vp, i32, i32, vp, vp]),
    "tn_scale_with_vector": (i32, [vp, i32, i32, vp, vp, vp]),
    "tn_randomize": (i32, [vp, i32, i32, vp, C.c_uint64, vp]),
    "tn_max_abs_entry": (i32, [vp, i32, i32, vp, C.POINTER(f64), vp]),
    "tn_has_nan_or_inf": (i32, [vp, i32, i32, vp, C.POINTER(C.c_int), vp]),
    "tn_eps_create": (i32, [vp, i32, vp, i32, i32, pvp]),
    "tn_eps_destroy": (i32, [vp]),
    "tn_eps_set_modes": (i32, [vp, i32, i32]),
    "tn_eps_set_diagonal": (i32, [vp, i32]),
    "tn_eps_set_boundary_conditions": (i32, [vp, i32, i32]),
    "tn_eps_set_eta": (i32, [vp, vp]),
    "tn_eps_apply": (i32, [vp, vp, vp, vp]),
    "tn_eps_set_penalty_epsilon": (i32, [vp, f64]),
    "tn_eps_set_penalty_once": (i32, [vp, i32]),
    "tn_eps_gca_from": (i32, [vp, vp, i32, vp]),
    "tn_eps_get_local_matrix": (i32, [vp, i32, i32, i32, i32, i32, vp]),
    "tn_eps_set_local_matrix": (i32, [vp, i32, i32, i32, i32, i32, vp]),
    "tn_eps_set_stored_matrices": (i32, [vp, i32, vp]),
    "tn_eps_set_cell_eta": (i32, [vp, vp]),
    "tn_eps_gca_from_bracket": (i32, [vp, vp, i32, i32, vp]),
    "tn_eps_grad_log_cell_eta": (i32, [vp, vp, vp, vp, vp]),
    "tn_stokes_create": (i32, [vp, vp, vp, pvp]),
    "tn_stokes_destroy": (i32, [vp]),
    "tn_stokes_apply": (i32, [vp, vp, vp, vp, vp, vp]),
    "tn_div_apply": (i32, [vp, vp, vp, vp]),
    "tn_grad_apply": (i32, [vp, vp, vp, i32, i32, vp]),
    "tn_vector_mass_apply": (i32, [vp, i32, vp, vp, vp]),
    "tn_kmass_lumped": (i32, [vp, i32, vp, vp, vp]),
    "tn_prolongate_add": (i32, [vp, vp, i32, i32, vp, vp, vp]),
    "tn_restrict": (i32, [vp, vp, i32, i32, vp, vp, vp]),
    "tn_gca_collect_elements": (i32, [vp, vp, i32, vp, vp, vp]),
    "tn_eps_gca_select": (i32, [vp, vp, vp, vp]),
    "tn_prolongate_linear": (i32, [vp, vp, i32, i32, vp, vp, i32, i32, vp]),
    "tn_restrict_linear": (i32, [vp, vp, i32, i32, vp, vp, i32, vp]),
    "tn_boundary_assign": (i32, [vp, i32, i32, vp, vp, i32, vp]),
    "tn_freeslip_project": (i32, [vp, i32, vp, i32, vp]),
    "tn_stokes_rhs_from_temperature": (i32, [vp, vp, f64, vp, vp, vp]),
    "tn_pressure_normalize": (i32, [vp, i32, vp, vp]),
    "tn_gmg_create": (i32, [i32, pvp, vp, vp, pvp]),
    "tn_gmg_destroy": (i32, [vp]),
    "tn_gmg_solve": (i32, [vp, vp, vp, vp]),
    "tn_gmg_get_lambda": (i32, [vp, i32, C.POINTER(f64)]),
    "tn_gmg_set_lambda": (i32, [vp, i32, f64]),
    "tn_pcg_solve": (i32, [vp, vp, vp, vp, vp, vp, vp, i32, vp]),
    "tn_stokes_solve": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp]),
    "tn_fv_create": (i32, [vp, pvp]),
    "tn_fv_destroy": (i32, [vp]),
    "tn_fv_size": (i32, [vp, C.POINTER(i64)]),
    "tn_fv_cell_centers": (i32, [vp, vp, vp]),
    "tn_fv_ghost_update": (i32, [vp, vp, i32, vp]),
    "tn_fv_dt_stable": (i32, [vp, vp, f64, C.POINTER(f64), vp]),
    "tn_fv_apply_dirichlet": (i32, [vp, vp, i32, f64, i32, f64, vp]),
    "tn_fct_predictor": (i32, [vp, vp, vp, f64, f64, vp, i32, vp]),
    "tn_fct_limiter": (i32, [vp, vp]),
    "tn_fct_correction": (i32, [vp, vp, vp]),
    "tn_fct_buffers": (i32, [vp, pvp, pvp, pvp, pvp]),
    "tn_fct_explicit_step": (i32, [vp, vp, vp, f64, f64, vp, i32, i32, f64, i32, f64, vp]),
    "tn_fct_upwind_step": (i32, [vp, vp, vp, f64, f64, vp, i32, vp]),
    "tn_fv_to_fe_rhs": (i32, [vp, vp, vp, vp]),
    "tn_fv_to_fe_lumped": (i32, [vp, vp, vp, vp, vp]),
    "tn_fe_to_fv": (i32, [vp, vp, vp, vp]),
    "tn_fv_to_fe": (i32, [vp, vp, vp, vp, vp, vp, i32, vp]),
    "tn_xdmf_create": (i32, [vp, C.c_char_p, i32, i32, pvp]),
    "tn_xdmf_destroy": (i32, [vp]),
    "tn_xdmf_add": (i32, [vp, C.c_char_p, i32, i32, i32, vp]),
    "tn_xdmf_set_write_counter": (i32, [vp, i32]),
    "tn_xdmf_write": (i32, [vp, vp]),
    "tn_xdmf_read_checkpoint_grid": (i32, [vp, C.c_char_p, C.c_char_p, i32, i32, i32, vp, vp]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TNError(
                f"{LIB_PATH} is missing: build it with `python -m terraneo_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().tn_last_error()
        raise TNError(msg.decode() if msg else f"terraneo_amd error {rc}")
    return rc
