"""feriphys_amd — the spring-mass-damper mesh step of jalberse/feriphys on MI355X (gfx950).

The compute path is libferiphys_gpu.so (hand-written HIP kernels behind the C ABI of
include/feriphys_gpu.h); this package is a thin ctypes mirror of the reference's Rust API.
"""
from ._lib import (FP_EULER, FP_KERNEL_AUTO, FP_KERNEL_CSR, FP_KERNEL_GRID, FP_KERNEL_GRID_SIMPLE, FP_MATH_EXACT,
                   FP_MATH_FAST,
                   FP_PLANE_XY, FP_PLANE_XZ, FP_RK4, FP_STEP_CSR, FP_STEP_GRID_SIMPLE, FP_STEP_PATCH, FP_STEP_TILED,
                   FP_TILE_7x31, FP_TILE_15x15, FP_TILE_16x16, FpError, lib)
from .springy import (Cloth, CollidableMesh, Config, Integration, Simulation, SpringyMesh, config_default,
                      pinned_empty, selftest_exact_math)
from .rigid import RigidBody, RigidConfig, RigidSimulation, rigid_config_default
from .surrogate import Surrogate, SurrogateConfig, SurrogateStats

__all__ = [
    "Cloth", "CollidableMesh", "Config", "Integration", "Simulation", "SpringyMesh", "config_default", "FpError",
    "selftest_exact_math", "pinned_empty",
    "Surrogate", "SurrogateConfig", "SurrogateStats", "RigidBody", "RigidConfig", "RigidSimulation",
    "rigid_config_default",
    "lib", "FP_EULER", "FP_RK4", "FP_PLANE_XY", "FP_PLANE_XZ", "FP_KERNEL_AUTO", "FP_KERNEL_CSR", "FP_KERNEL_GRID",
    "FP_KERNEL_GRID_SIMPLE",
    "FP_MATH_EXACT", "FP_MATH_FAST", "FP_STEP_TILED", "FP_STEP_GRID_SIMPLE", "FP_STEP_PATCH", "FP_STEP_CSR",
    "FP_TILE_15x15", "FP_TILE_7x31", "FP_TILE_16x16",
]
