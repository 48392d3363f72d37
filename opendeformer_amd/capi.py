"""ctypes binding of libodhip.so (the C-ABI declared in include/od.h).

There is no CPU fallback: if the HIP library is missing or fails to load,
`lib()` raises, and every handle constructor goes through it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libodhip.so")
_LIB = None

OD_MAT_COROTATIONAL, OD_MAT_NEOHOOKEAN, OD_MAT_COROT_PLASTIC = 0, 1, 2
OD_PRE_NONE, OD_PRE_JACOBI, OD_PRE_ICT, OD_PRE_IC0, OD_PRE_LDLT = 0, 1, 2, 3, 4
OD_FP64, OD_MIXED = 0, 1
OD_OK, OD_ERR_ARG, OD_ERR_HIP, OD_ERR_STATE, OD_ERR_UNSUPPORTED = 0, -1, -2, -3, -4

# every symbol include/od.h declares
EXPORTS = [
    "od_last_error", "od_version", "od_ctx_create", "od_ctx_destroy", "od_ctx_sync", "od_dev_alloc", "od_dev_free",
    "od_copy_h2d", "od_copy_d2h", "od_memset_d", "od_mesh_create", "od_mesh_destroy", "od_mesh_info",
    "od_pattern_export", "od_mesh_rows", "od_mesh_set_material", "od_body_force", "od_matrix_create",
    "od_matrix_destroy", "od_matrix_values", "od_matrix_upload", "od_matrix_download", "od_matrix_zero",
    "od_matrix_axpby", "od_mass_matrix", "od_assemble", "od_spmv", "od_matrix_sync_f32", "od_spmv_f32",
    "od_solver_create", "od_solver_destroy", "od_pcg_solve", "od_precond_apply", "od_ic0_factor",
    "od_newmark_create", "od_backward_euler_create", "od_integrator_destroy", "od_integrator_set_external",
    "od_integrator_set_impulse", "od_integrator_set_active", "od_integrator_step", "od_integrator_get_state",
    "od_integrator_set_state", "od_integrator_state_dev", "od_integrator_stats", "od_backward_euler_update_mesh",
    "od_integrator_positions", "od_spmv_timed", "od_integrator_matrix", "od_ctx_profile", "od_ctx_profile_read",
    "od_solver_set_refinement", "od_integrator_set_refinement", "od_mesh_autotune", "od_mesh_spmv_variant", "od_mesh_sweep_form",
    "od_solver_set_graph", "od_integrator_set_graph", "od_integrator_graph_info", "od_mesh_bind_ctx", "od_integrator_device_bytes",
    "od_server_create", "od_server_destroy", "od_server_add_object", "od_server_run", "od_server_object_state",
    "od_server_object_info", "od_mesh_set_plastic", "od_mesh_plastic_strain", "od_plane_damping",
    "od_integrator_add_plane_damping", "od_integrator_clear_damping", "od_integrator_damping", "od_relabel_gps",
    "od_solver_info", "od_integrator_solver", "od_solver_set_precond_params", "od_pcg_solve_noalloc",
    "od_integrator_damping_matrix", "od_precond_preprocess", "od_precond_solve",
]


class ODError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"od error {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ODError(-4, f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C opendeformer_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, i, d, ll = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int, C.c_double, C.c_longlong
        pvp = C.POINTER(C.c_void_p)
        L.od_last_error.restype = C.c_char_p
        sig = {
            "od_ctx_create": [i, pvp], "od_ctx_destroy": [vp], "od_ctx_sync": [vp],
            "od_dev_alloc": [vp, ll, pvp], "od_dev_free": [vp, vp], "od_copy_h2d": [vp, vp, vp, ll],
            "od_copy_d2h": [vp, vp, vp, ll], "od_memset_d": [vp, vp, i, ll],
            "od_mesh_create": [vp, i, dp, i, ip, i, ip, i, ip, ip, pvp], "od_mesh_destroy": [vp],
            "od_mesh_info": [vp, C.POINTER(ll)], "od_pattern_export": [vp, ip, ip, ip, ip], "od_mesh_rows": [vp, ip, ip],
            "od_mesh_set_material": [vp, ip, dp, dp, dp, d, d], "od_body_force": [vp, dp, vp],
            "od_matrix_create": [vp, pvp], "od_matrix_destroy": [vp], "od_matrix_values": [vp, pvp],
            "od_matrix_upload": [vp, dp], "od_matrix_download": [vp, dp], "od_matrix_zero": [vp],
            "od_matrix_axpby": [vp, d, d, vp], "od_mass_matrix": [vp, vp], "od_assemble": [vp, vp, vp, vp],
            "od_spmv": [vp, vp, vp, i], "od_matrix_sync_f32": [vp], "od_spmv_f32": [vp, vp, vp, i],
            "od_solver_create": [vp, i, d, i, pvp], "od_solver_destroy": [vp],
            "od_solver_set_refinement": [vp, d, i], "od_integrator_set_refinement": [vp, d, i],
            "od_mesh_autotune": [vp, dp, ip], "od_mesh_spmv_variant": [vp, i, i, i, i], "od_mesh_sweep_form": [vp, i, i],
            "od_solver_set_graph": [vp, i], "od_integrator_set_graph": [vp, i],
            "od_integrator_graph_info": [vp, C.POINTER(ll)],
            "od_mesh_bind_ctx": [vp, vp], "od_integrator_device_bytes": [vp, C.POINTER(ll)],
            "od_server_create": [i, ll, i, i, i, d, pvp], "od_server_destroy": [vp],
            "od_server_add_object": [vp, i, dp, i, ip, i, ip, ip, dp, dp, dp, dp, d, ip],
            "od_server_run": [vp, i, ip, ip, dp, ip, ip, dp, dp, dp, ip, C.POINTER(ll)],
            "od_server_object_state": [vp, i, dp, dp, dp], "od_server_object_info": [vp, i, C.POINTER(ll)],
            "od_mesh_set_plastic": [vp, dp, dp, dp], "od_mesh_plastic_strain": [vp, dp, dp],
            "od_pcg_solve": [vp, vp, vp, vp, d, i, ip, dp], "od_precond_apply": [vp, vp, vp, vp],
            "od_ic0_factor": [vp, vp, dp],
            "od_newmark_create": [vp, d, d, d, d, d, i, d, d, i, i, i, pvp],
            "od_backward_euler_create": [vp, d, d, d, i, d, d, i, pvp], "od_integrator_destroy": [vp],
            "od_integrator_set_external": [vp, dp], "od_integrator_set_impulse": [vp, dp],
            "od_integrator_set_active": [vp, C.POINTER(C.c_ubyte)], "od_integrator_step": [vp, i],
            "od_integrator_get_state": [vp, dp, dp, dp], "od_integrator_set_state": [vp, dp, dp, dp],
            "od_integrator_state_dev": [vp, pvp, pvp, pvp], "od_integrator_stats": [vp, ip, dp, ip, ip],
            "od_backward_euler_update_mesh": [vp], "od_integrator_positions": [vp, C.POINTER(C.c_float)],
            "od_plane_damping": [vp, i, dp, ip, d, vp, vp], "od_integrator_add_plane_damping": [vp, i, dp, ip, d],
            "od_integrator_clear_damping": [vp], "od_integrator_damping": [vp, dp],
            "od_relabel_gps": [i, i, i, ip, ip],
            "od_spmv_timed": [vp, vp, vp, i, i, dp], "od_integrator_matrix": [vp, pvp],
            "od_ctx_profile": [vp, i], "od_ctx_profile_read": [vp, dp, C.POINTER(ll)],
            "od_solver_info": [vp, C.POINTER(ll)], "od_integrator_solver": [vp, pvp],
            "od_solver_set_precond_params": [vp, d, d],
            "od_pcg_solve_noalloc": [vp, i, vp, vp, d, i, ip, dp], "od_integrator_damping_matrix": [vp, pvp],
            "od_precond_preprocess": [vp, vp], "od_precond_solve": [vp, vp, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _LIB = L
    return _LIB


def check(code):
    if code != 0:
        raise ODError(code, lib().od_last_error().decode())
    return code


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
