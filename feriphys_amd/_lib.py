"""Loader for libferiphys_gpu.so (the gfx950 springy step behind the C ABI in include/feriphys_gpu.h).

There is no CPU fallback: if the library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libferiphys_gpu.so")

FP_OK, FP_EINVAL, FP_ENONMANIFOLD, FP_EHIP, FP_ESTATE, FP_ENOMEM, FP_ENODEV = range(7)
FP_EULER, FP_RK4 = 0, 1
FP_PLANE_XY, FP_PLANE_XZ = 0, 1
FP_KERNEL_AUTO, FP_KERNEL_CSR, FP_KERNEL_GRID, FP_KERNEL_GRID_SIMPLE = 0, 1, 2, 3
FP_MATH_EXACT, FP_MATH_FAST = 0, 1
FP_TILE_15x15, FP_TILE_7x31, FP_TILE_16x16 = 0, 1, 2
FP_STEP_TILED, FP_STEP_GRID_SIMPLE, FP_STEP_PATCH, FP_STEP_CSR = 0, 1, 2, 3

# Every symbol declared in include/feriphys_gpu.h, with its ctypes signature.
_U64 = C.c_uint64
_U64P = C.POINTER(C.c_uint64)
_I64P = C.POINTER(C.c_int64)
_FP = C.POINTER(C.c_float)
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)


class FpConfig(C.Structure):
    _fields_ = [("integration", C.c_int32), ("dt", C.c_float), ("gravity", C.c_float * 3), ("wind", C.c_float * 3),
                ("lift_coefficient", C.c_float), ("drag_coefficient", C.c_float),
                ("coefficient_of_restitution", C.c_float), ("coefficient_of_friction", C.c_float)]


class FpStrutOverride(C.Structure):
    _fields_ = [("a", C.c_uint64), ("b", C.c_uint64), ("constant", C.c_float), ("damping", C.c_float)]


class FpTorsionConfig(C.Structure):
    _fields_ = [("spring_constant", C.c_float), ("spring_damping", C.c_float)]


class FpSurrogateConfig(C.Structure):
    _fields_ = [("dt", C.c_float), ("ratio", C.c_int32), ("strain_scale", C.c_float), ("out_scale", C.c_float),
                ("base_struts", C.c_int32), ("strain_limit", C.c_float), ("energy_limit", C.c_float)]


class FpSurrogateStats(C.Structure):
    _fields_ = [("big_steps", C.c_uint64), ("fallbacks", C.c_uint64), ("mesh_fallbacks", C.c_uint64),
                ("plain_only", C.c_uint64), ("max_strain_residual", C.c_float), ("max_energy_residual", C.c_float)]


class FpRigidConfig(C.Structure):
    _fields_ = [("integration", C.c_int32), ("dt", C.c_float), ("coefficient_of_restitution", C.c_float),
                ("gravity", C.c_float * 3), ("torque", C.c_float * 3)]


FP_RIGID_RECORD = 29
FP_SURROGATE_PARAMS = 42755
FP_SURROGATE_FEATURES = 72


SIGNATURES = {
    "fp_last_error": (C.c_char_p, []),
    "fp_version": (C.c_char_p, []),
    "fp_selftest_exact_math": (C.c_int, [C.POINTER(C.c_uint64)]),
    "fp_sim_count_nonfinite": (C.c_int, [_P, _U64P]),
    "fp_sim_count_nonfinite_noalloc": (C.c_int, [_P, _U64P]),
    "fp_rigid_config_default": (None, [C.POINTER(FpRigidConfig)]),
    "fp_rigid_new": (C.c_int, [_FP, C.c_float, _PP]),
    "fp_rigid_get_state": (C.c_int, [_P, _FP]),
    "fp_rigid_apply_impulse": (C.c_int, [_P, _FP, _FP]),
    "fp_rigid_free": (None, [_P]),
    "fp_rigid_sim_new": (C.c_int, [_PP, _U64, _PP, _U64, _PP]),
    "fp_rigid_sim_free": (None, [_P]),
    "fp_rigid_sim_count": (C.c_int, [_P, _U64P]),
    "fp_rigid_sim_set_config": (C.c_int, [_P, C.POINTER(FpRigidConfig)]),
    "fp_rigid_sim_get_config": (C.c_int, [_P, C.POINTER(FpRigidConfig)]),
    "fp_rigid_sim_step": (C.c_int, [_P, _U64, _FP]),
    "fp_rigid_sim_time_steps": (C.c_int, [_P, _U64, _FP]),
    "fp_rigid_sim_get_states": (C.c_int, [_P, _FP]),
    "fp_rigid_sim_set_states": (C.c_int, [_P, _FP, _U64]),
    "fp_rigid_sim_apply_impulse": (C.c_int, [_P, _U64, _FP, _FP]),
    "fp_config_default": (None, [C.POINTER(FpConfig)]),
    "fp_mesh_new": (C.c_int, [_FP, _U64, _U64P, _U64, C.c_float, C.c_float, C.c_float, C.POINTER(FpTorsionConfig),
                              C.POINTER(FpStrutOverride), _U64, _PP]),
    "fp_cloth_new": (C.c_int, [_U64, _U64, C.c_float, _FP, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_float, C.c_float, _U64P, _U64, _PP]),
    "fp_cloth_new_in_plane": (C.c_int, [_U64, _U64, C.c_float, _FP, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_float, _U64P, _U64, C.c_int32, _PP]),
    "fp_mesh_add_strut": (C.c_int, [_P, _U64, _U64, C.c_float, C.c_float]),
    "fp_mesh_add_struts": (C.c_int, [_P, _U64P, _U64, C.c_float, C.c_float]),
    "fp_mesh_add_pin": (C.c_int, [_P, _U64]),
    "fp_mesh_counts": (C.c_int, [_P, _U64P, _U64P, _U64P, _U64P]),
    "fp_mesh_get_struts": (C.c_int, [_P, _U64P, _FP, _I64P]),
    "fp_mesh_get_torsion": (C.c_int, [_P, C.POINTER(C.c_uint8), _FP]),
    "fp_mesh_get_faces": (C.c_int, [_P, _U64P, _U64P]),
    "fp_mesh_get_pins": (C.c_int, [_P, _U64P]),
    "fp_mesh_get_points": (C.c_int, [_P, _FP]),
    "fp_mesh_set_points": (C.c_int, [_P, _FP, _U64]),
    "fp_mesh_get_vertices": (C.c_int, [_P, _FP, _U64P]),
    "fp_mesh_free": (None, [_P]),
    "fp_obstacle_new": (C.c_int, [_FP, _U64, _U64P, _U64, _PP]),
    "fp_obstacle_counts": (C.c_int, [_P, _U64P, _U64P, _U64P]),
    "fp_obstacle_free": (None, [_P]),
    "fp_sim_new": (C.c_int, [_PP, _U64, _PP, _U64, _PP]),
    "fp_sim_set_config": (C.c_int, [_P, C.POINTER(FpConfig)]),
    "fp_sim_get_config": (C.c_int, [_P, C.POINTER(FpConfig)]),
    "fp_sim_set_mesh_config": (C.c_int, [_P, _U64, C.POINTER(FpConfig)]),
    "fp_sim_set_mesh_configs": (C.c_int, [_P, C.POINTER(FpConfig), _U64]),
    "fp_sim_step": (C.c_int, [_P, _U64, _FP]),
    "fp_sim_step_async": (C.c_int, [_P, _U64]),
    "fp_sim_synchronize": (C.c_int, [_P]),
    "fp_sim_get_timestep": (C.c_int, [_P, _FP]),
    "fp_sim_mesh_count": (C.c_int, [_P, _U64P]),
    "fp_sim_mesh_counts": (C.c_int, [_P, _U64, _U64P, _U64P, _U64P, _U64P]),
    "fp_sim_get_points": (C.c_int, [_P, _U64, _FP]),
    "fp_sim_set_points": (C.c_int, [_P, _U64, _FP, _U64]),
    "fp_sim_get_vertices": (C.c_int, [_P, _U64, _FP, _U64P]),
    "fp_host_alloc": (C.c_int, [_U64, _PP]),
    "fp_host_free": (C.c_int, [_P]),
    "fp_sim_accumulate_forces": (C.c_int, [_P, _U64]),
    "fp_sim_update_points": (C.c_int, [_P, _U64, _FP, _U64]),
    "fp_sim_clear_forces": (C.c_int, [_P, _U64]),
    "fp_sim_set_kernel": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "fp_sim_set_tile_shape": (C.c_int, [_P, _U64, C.c_int32]),
    "fp_sim_get_tile_shape": (C.c_int, [_P, _U64, C.POINTER(C.c_int32)]),
    "fp_sim_get_mesh_kernel": (C.c_int, [_P, _U64, C.POINTER(C.c_int32)]),
    "fp_sim_autotune": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int32)]),
    "fp_sim_step_graph": (C.c_int, [_P, _U64, C.c_uint32]),
    "fp_sim_time_steps": (C.c_int, [_P, _U64, _FP]),
    "fp_sim_time_steps_graph": (C.c_int, [_P, _U64, C.c_uint32, _FP]),
    "fp_sim_free": (None, [_P]),
    "fp_surrogate_config_default": (C.c_int, [C.POINTER(FpSurrogateConfig)]),
    "fp_surrogate_new": (C.c_int, [C.POINTER(FpSurrogateConfig), _U64, _FP, _PP]),
    "fp_surrogate_get_config": (C.c_int, [_P, C.POINTER(FpSurrogateConfig)]),
    "fp_surrogate_get_params": (C.c_int, [_P, _FP]),
    "fp_surrogate_set_params": (C.c_int, [_P, _FP]),
    "fp_surrogate_grad": (C.c_int, [_P, _FP, _FP, _U64, _FP, _FP, _FP]),
    "fp_surrogate_reserve": (C.c_int, [_P, _U64]),
    "fp_surrogate_sample_count": (C.c_int, [_P, _U64P]),
    "fp_surrogate_clear_samples": (C.c_int, [_P]),
    "fp_surrogate_add_samples": (C.c_int, [_P, _FP, _FP, _U64]),
    "fp_surrogate_train": (C.c_int, [_P, _U64, C.c_uint32, C.c_float, _U64, _FP]),
    "fp_surrogate_free": (None, [_P]),
    "fp_sim_collect_samples": (C.c_int, [_P, _P, _U64, C.c_float, _U64]),
    "fp_sim_apply_torsional_forces": (C.c_int, [_P, _U64]),
    "fp_sim_set_torsion": (C.c_int, [_P, C.c_int32]),
    "fp_sim_set_surrogate": (C.c_int, [_P, _P]),
    "fp_sim_set_surrogate_shape": (C.c_int, [_P, C.c_int32]),
    "fp_sim_step_surrogate": (C.c_int, [_P, _U64, C.POINTER(FpSurrogateStats)]),
    "fp_sim_time_surrogate": (C.c_int, [_P, _U64, _FP, _FP, C.POINTER(FpSurrogateStats)]),
    "fp_sim_surrogate_features": (C.c_int, [_P, _U64, _FP]),
    "fp_sim_surrogate_accel": (C.c_int, [_P, _U64, _FP]),
    "fp_sim_surrogate_slots": (C.c_int, [_P, _U64, _I64P, _FP]),
}

_LIB = None


class FpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class FpIndexError(FpError, IndexError):
    pass


class FpValueError(FpError, ValueError):
    pass


def lib() -> C.CDLL:
    """The loaded C ABI library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C feriphys_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(code: int):
    if code == FP_OK:
        return
    msg = lib().fp_last_error().decode(errors="replace")
    if code == FP_EINVAL and ("out of range" in msg or "index" in msg):
        raise FpIndexError(code, msg)
    if code in (FP_EINVAL, FP_ENONMANIFOLD):
        raise FpValueError(code, msg)
    raise FpError(code, msg)
