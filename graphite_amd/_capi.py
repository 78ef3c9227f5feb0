"""ctypes declarations of libgraphite_hip.so (include/graphite_hip.h)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# GRAPHITE_HIP_LIB selects another build of the same library (the sanitizer build of the host code)
LIB_PATH = os.environ.get("GRAPHITE_HIP_LIB") or os.path.join(HERE, "libgraphite_hip.so")

GR_OK, GR_ERR_INVALID, GR_ERR_TOPOLOGY, GR_ERR_IO, GR_ERR_NOT_FOUND, GR_ERR_DEVICE, GR_ERR_DIVERGED = range(7)
GR_F32, GR_F64 = 4, 8
GR_TRI, GR_QUAD, GR_TET, GR_HEX, GR_WEDGE, GR_PYRAMID = 5, 9, 10, 12, 13, 14
GR_LIM_MICHALAK, GR_LIM_BARTH_JESPERSEN, GR_LIM_VENKATAKRISHNAN = 0, 1, 2
GR_BC_SLIP, GR_BC_FARFIELD = 0, 1
(GR_OP_LAPLACIAN_C, GR_OP_SCAL_GRAD_C, GR_OP_SCAL_GRAD_C_CORRECT, GR_OP_VEC_GRAD_C, GR_OP_BOUND_SCAL_GRAD_C,
 GR_OP_BOUND_VEC_GRAD_C, GR_OP_DIV_C, GR_OP_CONV_CONST_SCAL_U, GR_OP_CONV_CONST_SCAL_U2, GR_OP_CONV_SCAL_U,
 GR_OP_CONV_VEC_U) = range(11)
OP_OUTPUTS = (1, 3, 3, 9, 3, 9, 1, 1, 1, 1, 3)

# Every function the header declares: name -> (restype, argtypes).
_vp = C.c_void_p
_u32 = C.c_uint32
_u32p = C.POINTER(C.c_uint32)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i = C.c_int
_d = C.c_double
_s = C.c_char_p


class ElementsInfo(C.Structure):
    _fields_ = [("n_nodes", _u32), ("n_elems", _u32), ("n_elem_nodes", _u32), ("n_markers", _u32)]


class MeshSizes(C.Structure):
    _fields_ = [("n_nodes", _u32), ("n_faces", _u32), ("n_cells", _u32), ("n_ghosts", _u32),
                ("n_regions", _u32), ("n_face_nodes", _u32), ("n_cell_slots", _u32)]


class SolverConfig(C.Structure):
    _fields_ = [("dtype", _i), ("cfl", _d), ("limiter", _i), ("corrected", _i), ("use_graph", _i)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", _i), ("converged", _i), ("residual", _d), ("ratio", _d), ("reference", _d)]


class WingParams(C.Structure):
    _fields_ = [("n_around", _u32), ("n_prism", _u32), ("n_tet", _u32), ("n_span", _u32), ("n_cap", _u32),
                ("thickness", _d), ("sweep_deg", _d), ("taper", _d), ("semispan", _d),
                ("first_height", _d), ("growth", _d), ("radius", _d), ("scale", _d)]


class NsiConfig(C.Structure):
    _fields_ = [("dtype", _i), ("nu", _d), ("dt", _d), ("cg_max", _i), ("cg_tol", _d)]


class WarmNorm(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("in_mean", C.c_float * 6), ("in_std", C.c_float * 6),
                ("out_mean", C.c_float * 5), ("out_std", C.c_float * 5), ("trust", C.c_float * 2)]


SIGNATURES = {
    "gr_last_error": (C.c_char_p, []),
    "gr_abi_version": (_i, []),
    "gr_device_count": (_i, []),
    "gr_elements_create": (_i, [_u32, _dp, _u32, _u32p, _u32p, _u32p, C.POINTER(_vp)]),
    "gr_elements_add_marker": (_i, [_vp, _s, _u32, _u32p, _u32p, _u32p]),
    "gr_elements_read_su2": (_i, [_s, C.POINTER(_vp)]),
    "gr_elements_write_su2": (_i, [_vp, _s]),
    "gr_elements_info_get": (_i, [_vp, C.POINTER(ElementsInfo)]),
    "gr_elements_get": (_i, [_vp, _dp, _u32p, _u32p, _u32p]),
    "gr_elements_get_marker": (_i, [_vp, _u32, C.c_char_p, _u32p, _u32p, _u32p, _u32p, _u32p]),
    "gr_elements_free": (None, [_vp]),
    "gr_gen_box": (_i, [_u32, _u32, _u32, _d, _d, _d, _i, _d, C.c_uint64, C.POINTER(_vp)]),
    "gr_gen_ogrid": (_i, [_u32, _u32, _u32, _d, _d, _d, _d, C.POINTER(_vp)]),
    "gr_gen_cgrid_naca": (_i, [_u32, _u32, _u32, _d, _d, _d, _d, C.POINTER(_vp)]),
    "gr_gen_wing": (_i, [C.POINTER(WingParams), C.POINTER(_vp)]),
    "gr_mesh_build": (_i, [_vp, C.POINTER(_vp)]),
    "gr_mesh_from_su2": (_i, [_s, _s, _i]),
    "gr_mesh_read_fvm": (_i, [_s, _i, C.POINTER(_vp)]),
    "gr_mesh_write_fvm": (_i, [_vp, _s, _i]),
    "gr_mesh_make_1d": (_i, [C.c_size_t, _d, _d, C.POINTER(_vp)]),
    "gr_mesh_free": (None, [_vp]),
    "gr_mesh_sizes_get": (_i, [_vp, C.POINTER(MeshSizes)]),
    "gr_mesh_get_connectivity": (_i, [_vp, _dp, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]),
    "gr_mesh_region_id": (_i, [_vp, _s, _u32p]),
    "gr_mesh_region_name": (_i, [_vp, _u32, C.c_char_p]),
    "gr_mesh_region_size": (_i, [_vp, _u32, _u32p]),
    "gr_mesh_renumber_rcm": (_i, [_vp, _u32p, _u32p]),
    "gr_mesh_renumber_hilbert": (_i, [_vp, _u32p, _u32p]),
    "gr_mesh_face_colors": (_i, [_vp, _u32p, _u32p]),
    "gr_mesh_bandwidth": (_i, [_vp, _u32p]),
    "gr_mesh_compute_metrics": (_i, [_vp]),
    "gr_mesh_get_metrics": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "gr_mesh_get_view": (_i, [_vp, _vp]),
    "gr_field_create": (_i, [_vp, _u32, _i, _s, _i, C.POINTER(_vp)]),
    "gr_field_free": (None, [_vp]),
    "gr_field_upload": (_i, [_vp, _dp]),
    "gr_field_download": (_i, [_vp, _dp, _dp, _dp]),
    "gr_field_init": (_i, [_vp, _dp]),
    "gr_field_operator": (_i, [_vp, _i, _dp, _vp, _dp]),
    "gr_field_get_view": (_i, [_vp, _vp]),
    "gr_field_reduce_sum": (_i, [_vp, _dp]),
    "gr_field_compute_gradient": (_i, [_vp]),
    "gr_field_compute_limiter": (_i, [_vp, _i]),
    "gr_field_gradlim_bc": (_i, [_vp]),
    "gr_field_bc_fixed": (_i, [_vp, _u32, _dp]),
    "gr_field_bc_zero_grad": (_i, [_vp, _u32]),
    "gr_euler_init": (_i, [_vp, _vp, _vp, _vp, _dp, _d, _d]),
    "gr_euler_bc_slip": (_i, [_vp, _vp, _vp, _vp, _s]),
    "gr_euler_bc_farfield": (_i, [_vp, _vp, _vp, _vp, _dp, _d, _d, _s]),
    "gr_euler_step_prepare": (_i, [_vp, _vp, _vp, _vp, _i]),
    "gr_euler_step": (_i, [_vp, _vp, _vp, _vp, _d, _s, _vp, _vp]),
    "gr_euler_residual": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "gr_euler_residual_noalloc": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "gr_euler_set_corrected_reconstruction": (_i, [_vp, _i]),
    "gr_solver_create": (_i, [_vp, C.POINTER(SolverConfig), C.POINTER(_vp)]),
    "gr_solver_free": (None, [_vp]),
    "gr_solver_set_bc": (_i, [_vp, _s, _i, _dp, _d, _d]),
    "gr_solver_init_freestream": (_i, [_vp, _dp, _d, _d]),
    "gr_solver_set_state": (_i, [_vp, _dp, _dp, _dp]),
    "gr_solver_get_state": (_i, [_vp, _dp, _dp, _dp]),
    "gr_solver_iterate": (_i, [_vp, _i, _d, _d, C.POINTER(SolveResult)]),
    "gr_solver_history": (_i, [_vp, _dp, _i]),
    "gr_solver_residual_norm": (_i, [_vp, _dp]),
    "gr_solver_profile": (_i, [_vp, _i, _dp]),
    "gr_solver_autotune": (_i, [_vp, _i, C.POINTER(C.c_uint32)]),
    "gr_solver_kernel_bytes": (_i, [_vp, _dp]),
    "gr_solver_layout_bytes": (_i, [_vp, _dp]),
    "gr_stream_copy_probe": (_i, [C.c_size_t, _i, _dp]),
    "gr_batch_create": (_i, [_vp, _u32, C.POINTER(SolverConfig), C.POINTER(_vp)]),
    "gr_batch_free": (None, [_vp]),
    "gr_batch_set_bc": (_i, [_vp, _s, _i]),
    "gr_batch_set_freestream": (_i, [_vp, _dp, _d, _d]),
    "gr_batch_set_state": (_i, [_vp, _u32, _dp, _dp, _dp]),
    "gr_batch_get_state": (_i, [_vp, _u32, _dp, _dp, _dp]),
    "gr_batch_iterate": (_i, [_vp, _i, _d, _dp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "gr_batch_iterate_masked": (_i, [_vp, _i, _d, _dp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                                     C.POINTER(_i)]),
    "gr_batch_reset_freestream": (_i, [_vp, C.POINTER(_i)]),
    "gr_batch_history": (_i, [_vp, _u32, _dp, _i]),
    "gr_batch_residual_norms": (_i, [_vp, _dp]),
    "gr_batch_profile": (_i, [_vp, _i, _dp]),
    "gr_batch_kernel_bytes": (_i, [_vp, _dp]),
    "gr_batch_set_graph_iters": (_i, [_vp, _i]),
    "gr_batch_get_schedule": (_i, [_vp, C.POINTER(C.c_int)]),
    "gr_batch_set_schedule": (_i, [_vp, C.POINTER(C.c_int)]),
    "gr_batch_autotune": (_i, [_vp, _i, C.POINTER(C.c_int)]),
    "gr_solver_force": (_i, [_vp, _s, _dp]),
    "gr_batch_force": (_i, [_vp, _s, _dp]),
    "gr_mesh_wall_distance": (_i, [_vp, C.POINTER(C.c_char_p), _u32, _dp]),
    "gr_write_vtu": (_i, [_vp, _s, _u32, C.POINTER(C.c_char_p), _u32p, C.POINTER(_dp)]),
    "gr_case_check": (_i, [_s]),
    "gr_case_run": (_i, [_s, _i, C.POINTER(SolveResult)]),
    "gr_nsi_create": (_i, [_vp, C.POINTER(NsiConfig), C.POINTER(_vp)]),
    "gr_nsi_free": (None, [_vp]),
    "gr_nsi_set_wall": (_i, [_vp, _s, _dp]),
    "gr_nsi_set_state": (_i, [_vp, _dp, _dp, _dp]),
    "gr_nsi_get_state": (_i, [_vp, _dp, _dp, _dp]),
    "gr_nsi_step": (_i, [_vp, _i, _dp, C.POINTER(_i)]),
    "gr_nsi_iterate": (_i, [_vp, _i, _d, C.POINTER(_i), _dp]),
    "gr_nsi_history": (_i, [_vp, _dp, C.POINTER(_i), _i]),
    "gr_nsi_profile": (_i, [_vp, _i, _dp]),
    "gr_mlp_create": (_i, [_u32, _u32, C.c_uint64, C.POINTER(_vp)]),
    "gr_mlp_free": (None, [_vp]),
    "gr_mlp_n_params": (_i, [_vp, C.POINTER(C.c_size_t)]),
    "gr_mlp_get_params": (_i, [_vp, _fp]),
    "gr_mlp_set_params": (_i, [_vp, _fp]),
    "gr_mlp_forward": (_i, [_vp, _fp, _i, _fp]),
    "gr_mlp_loss_grad": (_i, [_vp, _fp, _fp, _i, _dp, _fp]),
    "gr_mlp_train_step": (_i, [_vp, _fp, _fp, _i, C.c_float, _dp]),
    "gr_mlp_set_data": (_i, [_vp, _fp, _fp, C.c_uint64]),
    "gr_mlp_train": (_i, [_vp, _i, _i, C.c_float, C.c_float, C.c_uint64, _dp]),
    "gr_batch_set_states": (_i, [_vp, _dp, _dp, _dp]),
    "gr_batch_get_states": (_i, [_vp, _dp, _dp, _dp]),
    "gr_batch_warm_rows": (_i, [_vp, _dp, _fp, _fp, _fp]),
    "gr_batch_warm_start": (_i, [_vp, _vp, _dp, C.POINTER(WarmNorm), _d, _dp, C.POINTER(_i)]),
}

_lib = None


def load() -> C.CDLL:
    """Load the in-tree library; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m graphite_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
