"""ctypes binding of the C ABI declared in include/xrtailor_amd.h.

The shared library is built in-tree by ``python -m xrtailor_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libxrtailor_amd.so"

XRT_OK = 0
ERRORS = {
    -1: "XRT_ERR_INVALID_ARG",
    -2: "XRT_ERR_HIP",
    -3: "XRT_ERR_STATE",
    -4: "XRT_ERR_CAPACITY",
    -5: "XRT_ERR_NOT_FINITE",
    -6: "XRT_ERR_UNSUPPORTED",
}
XRT_ERR_CAPACITY = -4


class XrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class XrtParams(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 3),
        ("damping", C.c_float),
        ("relaxation_factor", C.c_float),
        ("long_range_stretchiness", C.c_float),
        ("bvh_tolerance", C.c_float),
        ("ground_height", C.c_float),
        ("ground_friction", C.c_float),
        ("pbd_thickness", C.c_float),
        ("repulsion_relaxation", C.c_float),
        ("frame_time", C.c_float),
        ("num_substeps", C.c_int32),
        ("num_iterations", C.c_int32),
        ("num_collision_passes", C.c_int32),
        ("solve_bending", C.c_int32),
        ("enable_ground", C.c_int32),
        ("enable_obstacle_contact", C.c_int32),
        ("enable_self_contact", C.c_int32),
        ("pair_update_interval", C.c_int32),
        ("bend_model", C.c_int32),
        ("coloring", C.c_int32),
        ("pair_capacity_per_face", C.c_int32),
        ("repulsion_vf_items_per_face", C.c_int32),
        ("repulsion_ee_items_per_face", C.c_int32),
        ("enable_particle_collision", C.c_int32),
        ("max_num_neighbors", C.c_int32),
        ("inter_leaved_hash", C.c_int32),
        ("particle_diameter", C.c_float),
        ("hash_cell_size", C.c_float),
        ("constraint_model", C.c_int32),
        ("fem_xx_stiffness", C.c_float),
        ("fem_yy_stiffness", C.c_float),
        ("fem_xy_stiffness", C.c_float),
        ("fem_xy_poisson_ratio", C.c_float),
        ("fem_yx_poisson_ratio", C.c_float),
        ("fem_bending_stiffness", C.c_float),
        ("contact_friction", C.c_float),
    ]


class XrtTuning(C.Structure):
    _fields_ = [
        ("persistent_block", C.c_int32),
        ("persistent_unroll", C.c_int32),
        ("wide_per_thread", C.c_int32),
        ("path", C.c_int32),
        ("best_ms", C.c_float),
        ("default_ms", C.c_float),
        ("candidates", C.c_int32),
    ]


_P = C.c_void_p
_I = C.c_int32
_F = C.c_float
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)

# name -> argtypes (all functions return int status unless listed in _RESTYPES)
SIGNATURES = {
    "xrt_abi_version": [],
    "xrt_last_error": [],
    "xrt_default_params": [C.POINTER(XrtParams)],
    "xrt_create": [C.POINTER(_P), C.c_int],
    "xrt_destroy": [_P],
    "xrt_set_params": [_P, C.POINTER(XrtParams)],
    "xrt_get_params": [_P, C.POINTER(XrtParams)],
    "xrt_sync": [_P],
    "xrt_device_info": [_P, C.POINTER(C.c_int64)],
    "xrt_build_edges": [_IP, _I, _IP, _I, _IP],
    "xrt_build_bends": [_IP, _I, _IP, _I, _IP, _I, _IP],
    "xrt_color_constraints": [_IP, _I, _I, _I, C.c_uint64, _I, _IP, _IP],
    "xrt_add_cloth": [_P, _FP, _I, _IP, _I, _F, _F, _IP],
    "xrt_pin_vertices": [_P, _I, _IP, _I],
    "xrt_set_stretch_compliance": [_P, _I, _FP],
    "xrt_add_lra_slots": [_P, _I, _IP, _I],
    "xrt_set_obstacle": [_P, _I, _FP, _I, _IP, _I],
    "xrt_start": [_P, C.c_uint64],
    "xrt_update_obstacle": [_P, _I, _FP],
    "xrt_get_contacts": [_P, _I, _IP, _FP, _IP, _FP, _FP, _IP],
    "xrt_num_instances": [_P, _IP],
    "xrt_get_counts": [_P, _I, _IP],
    "xrt_get_edges": [_P, _I, _IP, _FP, _IP],
    "xrt_get_bends": [_P, _I, _IP, _FP, _IP],
    "xrt_get_lra": [_P, _I, _IP, _IP, _IP, _FP],
    "xrt_get_inv_mass": [_P, _I, _FP],
    "xrt_set_state": [_P, _I, _FP, _FP],
    "xrt_get_state": [_P, _I, _FP, _FP],
    "xrt_get_predicted": [_P, _I, _FP],
    "xrt_predict_positions": [_P, _I, _F],
    "xrt_stretch_solve": [_P, _I, _I, _F, _I],
    "xrt_bend_solve": [_P, _I, _I, _F, _I],
    "xrt_strain_solve": [_P, _I, _I],
    "xrt_isometric_bending_solve": [_P, _I, _I, _I, _F],
    "xrt_get_fem": [_P, _I, _IP, _IP, _FP, _FP, _FP, _FP],
    "xrt_time_fem": [_P, _I, _FP],
    "xrt_solve_lra": [_P, _I],
    "xrt_solve_ground": [_P, _I],
    "xrt_apply_deltas": [_P, _I],
    "xrt_obstacle_bvh_construct": [_P, _I],
    "xrt_cloth_bvh_construct": [_P, _I],
    "xrt_obstacle_interpolate": [_P, _I, _I, _I],
    "xrt_obstacle_update_midstep_normals": [_P, _I],
    "xrt_obstacle_bvh_refit_midstep": [_P, _I, _F],
    "xrt_obstacle_predicted_to_positions": [_P, _I],
    "xrt_vf_dcd_generate": [_P, _I],
    "xrt_vf_dcd_solve": [_P, _I],
    "xrt_vf_ccd_generate": [_P, _I],
    "xrt_vf_ccd_solve": [_P, _I],
    "xrt_ee_dcd_generate": [_P, _I],
    "xrt_ee_dcd_solve": [_P, _I],
    "xrt_repulsion_update_pairs": [_P, _I, _F],
    "xrt_repulsion_solve_no_proximity": [_P, _I, _F, _F],
    "xrt_finalize": [_P, _I, _F],
    "xrt_add_bindings": [_P, _I, _IP, _FP, _FP, _I],
    "xrt_set_obstacle_skinning": [_P, _I, _IP, _IP, _FP, _I],
    "xrt_skin_obstacles": [_P, _FP],
    "xrt_skin_obstacle": [_P, _I, _FP],
    "xrt_query_nearest_barycentric": [_P, _I, _IP, _FP],
    "xrt_get_bindings": [_P, _I, _IP, _FP],
    "xrt_solve_bindings": [_P, _I],
    "xrt_get_obstacle": [_P, _I, _FP, _FP],
    "xrt_time_binding": [_P, _I, _FP],
    "xrt_particle_hash": [_P, _I],
    "xrt_collide_particles": [_P, _I],
    "xrt_get_particle_hash": [_P, _I, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _IP,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _FP],
    "xrt_substep_plain": [_P, _I],
    "xrt_frames_plain": [_P, _I],
    "xrt_substep_batched": [_P],
    "xrt_frames_batched": [_P, _I],
    "xrt_time_batched_substep": [_P, _I, _FP],
    "xrt_record_frames": [_P, _I, _FP, _FP],
    "xrt_write_obj": [C.c_char_p, _FP, _FP, _I, _IP, _I],
    "xrt_write_obj_sequence": [C.c_char_p, _I, _I, _FP, _FP, _I, _IP, _I],
    "xrt_time_frames": [_P, _I, _I, _FP],
    "xrt_adjoint_begin": [_P, _I, _I],
    "xrt_adjoint_seed": [_P, _I, _FP, _FP],
    "xrt_adjoint_backward": [_P],
    "xrt_adjoint_get_grad_compliance": [_P, _I, _FP, _FP],
    "xrt_adjoint_get_grad_state": [_P, _I, _FP, _FP],
    "xrt_adjoint_substep": [_P],
    "xrt_time_adjoint": [_P, _I, _FP],
    "xrt_time_repulsion": [_P, _I, _FP, C.POINTER(C.c_int64)],
    "xrt_time_particles": [_P, _I, _FP, C.POINTER(C.c_int64)],
    "xrt_bvh_create": [C.POINTER(_P), C.c_int],
    "xrt_bvh_destroy": [_P],
    "xrt_bvh_build": [_P, _FP, _FP, _FP, _I, _IP, _I, _F, _I, _F],
    "xrt_bvh_refit_midstep": [_P, _FP, _FP, _F],
    "xrt_bvh_update": [_P, _FP, _FP, _I],
    "xrt_bvh_num_objects": [_P, _IP],
    "xrt_bvh_get": [_P, C.POINTER(C.c_uint32), _IP, _IP, _FP, _IP, _IP, _FP],
    "xrt_bvh_query": [_P, _FP, _FP, _I, _IP, _IP],
    "xrt_bvh_traverse": [_P, _P, _F, _IP, C.c_int64, C.POINTER(C.c_int64)],
    "xrt_bvh_traverse_noalloc": [_P, _P, _F, _IP, C.c_int64, C.POINTER(C.c_int64)],
    "xrt_vertex_normals": [_P, _FP, _I, _IP, _I, _FP],
    "xrt_autotune": [_P, _I, C.POINTER(XrtTuning)],
    "xrt_get_tuning": [_P, C.POINTER(XrtTuning)],
    "xrt_set_tuning": [_P, C.POINTER(XrtTuning)],
}
_RESTYPES = {"xrt_last_error": C.c_char_p, "xrt_default_params": None}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m xrtailor_amd.build` "
            "(the solver has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != XRT_OK:
        raise XrtError(rc, load().xrt_last_error().decode(errors="replace"))


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)
