"""Python mirror of the reference's springy-mesh API over the gfx950 C ABI.

Names, argument meaning and error behaviour follow the reference crate:
  Integration, Config            springy/config.rs:9-33, state.rs:5-8
  SpringyMesh                    springy_mesh.rs:258-661 (new, add_strut, add_pin, get_points, get_vertices)
  Cloth                          springy/cloth.rs:9-113
  CollidableMesh                 collidable_mesh.rs:50-197
  Simulation                     springy/simulation.rs:10-78 (new, step, get_timestep, get_meshes)
Where the reference panics (bad index, wrong record width, non-manifold edge) these raise
IndexError / ValueError.  Point records are numpy (N, 10) float32 arrays in the Stateful layout
[mass, x, y, z, vx, vy, vz, Fx, Fy, Fz] (springy_mesh.rs:242-255).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib

_F = C.POINTER(C.c_float)
_U64P = C.POINTER(C.c_uint64)


def _fptr(a):
    return a.ctypes.data_as(_F)


def _uptr(a):
    return a.ctypes.data_as(_U64P)


class Integration(IntEnum):
    """state.rs:5-8"""
    Euler = _lib.FP_EULER
    Rk4 = _lib.FP_RK4


@dataclass
class Config:
    """springy/config.rs:9-18; defaults are Config::default() (:20-33).  `gravity` is a force
    added to every vertex (springy_mesh.rs:194-196)."""
    integration: Integration = Integration.Rk4
    dt: float = 0.001
    gravity: tuple = (0.0, -10.0, 0.0)
    wind: tuple = (0.0, 0.0, 0.0)
    lift_coefficient: float = 1.0
    drag_coefficient: float = 1.0
    coefficient_of_restitution: float = 0.95
    coefficient_of_friction: float = 0.3

    def to_c(self) -> _lib.FpConfig:
        c = _lib.FpConfig()
        c.integration = int(self.integration)
        c.dt = self.dt
        for i in range(3):
            c.gravity[i] = self.gravity[i]
            c.wind[i] = self.wind[i]
        c.lift_coefficient = self.lift_coefficient
        c.drag_coefficient = self.drag_coefficient
        c.coefficient_of_restitution = self.coefficient_of_restitution
        c.coefficient_of_friction = self.coefficient_of_friction
        return c

    @staticmethod
    def from_c(c: _lib.FpConfig) -> "Config":
        return Config(Integration(c.integration), c.dt, tuple(c.gravity), tuple(c.wind), c.lift_coefficient,
                      c.drag_coefficient, c.coefficient_of_restitution, c.coefficient_of_friction)


class _Handle:
    _free = None

    def __init__(self, h):
        self._h = h

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("handle was moved into a Simulation")
        return self._h

    def _release(self):
        h, self._h = self._h, None
        return h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and self._free:
            getattr(lib(), self._free)(h)
            self._h = None


class SpringyMesh(_Handle):
    """springy_mesh.rs:258-661.  `mass` is the total mass (mass / N per vertex, :277)."""
    _free = "fp_mesh_free"

    def __init__(self, vertex_positions, vertex_indices, mass: float, default_stiffness: float,
                 default_damping: float, torsional_spring_config: Optional[tuple] = None,
                 strut_overrides: Optional[dict] = None, _handle=None):
        if _handle is not None:
            super().__init__(_handle)
            return
        pos = np.ascontiguousarray(np.asarray(vertex_positions, np.float32).reshape(-1, 3))
        idx = np.ascontiguousarray(np.asarray(vertex_indices, np.uint64).reshape(-1))
        tor = None
        if torsional_spring_config is not None:
            tor = _lib.FpTorsionConfig(*torsional_spring_config)
        ovr = None
        n_ovr = 0
        if strut_overrides:
            items = list(strut_overrides.items())
            n_ovr = len(items)
            ovr = (_lib.FpStrutOverride * n_ovr)()
            for k, ((a, b), (kk, cc)) in enumerate(items):
                ovr[k] = _lib.FpStrutOverride(a, b, kk, cc)
        h = C.c_void_p()
        check(lib().fp_mesh_new(_fptr(pos), pos.shape[0], _uptr(idx), idx.size, mass, default_stiffness,
                                default_damping, C.byref(tor) if tor is not None else None, ovr, n_ovr, C.byref(h)))
        super().__init__(h.value)

    def add_strut(self, vertex_indices: Sequence[int], stiffness: float, damping: float):
        check(lib().fp_mesh_add_strut(self.handle, int(vertex_indices[0]), int(vertex_indices[1]), stiffness, damping))

    def add_struts(self, pairs, stiffness: float, damping: float):
        """add_strut for each (v0, v1) row of `pairs`, in order, in one call (no reference
        counterpart; the same struts as the loop)."""
        v01 = np.ascontiguousarray(np.asarray(pairs, np.uint64).reshape(-1, 2))
        check(lib().fp_mesh_add_struts(self.handle, _uptr(v01), v01.shape[0], stiffness, damping))

    def add_pin(self, pin_index: int):
        check(lib().fp_mesh_add_pin(self.handle, int(pin_index)))

    def counts(self):
        v = [C.c_uint64() for _ in range(4)]
        check(lib().fp_mesh_counts(self.handle, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def struts(self):
        """(v01 (S,2) uint64, kcl (S,3) float32 [stiffness, damping, length], faces (S,2) int64, -1=None)"""
        S = self.counts()[1]
        v01 = np.zeros((max(S, 1), 2), np.uint64)
        kcl = np.zeros((max(S, 1), 3), np.float32)
        fc = np.zeros((max(S, 1), 2), np.int64)
        check(lib().fp_mesh_get_struts(self.handle, _uptr(v01), _fptr(kcl), fc.ctypes.data_as(C.POINTER(C.c_int64))))
        return v01[:S], kcl[:S], fc[:S]

    def torsion(self):
        S = self.counts()[1]
        has = np.zeros(max(S, 1), np.uint8)
        kcr = np.zeros((max(S, 1), 3), np.float32)
        check(lib().fp_mesh_get_torsion(self.handle, has.ctypes.data_as(C.POINTER(C.c_uint8)), _fptr(kcr)))
        return has[:S].astype(bool), kcr[:S]

    def faces(self):
        F = self.counts()[2]
        s3 = np.zeros((max(F, 1), 3), np.uint64)
        v3 = np.zeros((max(F, 1), 3), np.uint64)
        check(lib().fp_mesh_get_faces(self.handle, _uptr(s3), _uptr(v3)))
        return s3[:F], v3[:F]

    def pins(self):
        P = self.counts()[3]
        out = np.zeros(max(P, 1), np.uint64)
        check(lib().fp_mesh_get_pins(self.handle, _uptr(out)))
        return out[:P]

    def get_points(self) -> np.ndarray:
        N = self.counts()[0]
        out = np.zeros((N, 10), np.float32)
        check(lib().fp_mesh_get_points(self.handle, _fptr(out)))
        return out

    def set_points(self, records):
        rec = np.ascontiguousarray(np.asarray(records, np.float32).reshape(-1, 10))
        check(lib().fp_mesh_set_points(self.handle, _fptr(rec), rec.shape[0]))

    def get_vertices(self):
        N, _, F, _ = self.counts()
        pos = np.zeros((N, 3), np.float32)
        idx = np.zeros(max(3 * F, 1), np.uint64)
        check(lib().fp_mesh_get_vertices(self.handle, _fptr(pos), _uptr(idx)))
        return pos, idx[:3 * F]


class Cloth:
    """springy/cloth.rs:9-113.  `plane` 0 = xy (the reference), 1 = xz (rows along +z)."""

    def __init__(self, rows: int, cols: int, spacing: float, position, point_mass: float, tensile_stiffness: float,
                 tensile_damping: float, shear_stiffness: float, shear_damping: float,
                 binding_spring_stiffness: float, binding_spring_damping: float, pinned_vertices: Iterable[int],
                 plane: int = _lib.FP_PLANE_XY):
        pos = np.asarray(position, np.float32).reshape(3)
        pins = np.asarray(list(pinned_vertices), np.uint64)
        pins_buf = pins if pins.size else np.zeros(1, np.uint64)
        h = C.c_void_p()
        check(lib().fp_cloth_new_in_plane(rows, cols, spacing, _fptr(pos), point_mass, tensile_stiffness,
                                          tensile_damping, shear_stiffness, shear_damping, binding_spring_stiffness,
                                          binding_spring_damping, _uptr(pins_buf), pins.size, plane, C.byref(h)))
        self.mesh = SpringyMesh(None, None, 0, 0, 0, _handle=h.value)


class CollidableMesh(_Handle):
    """collidable_mesh.rs:50-197."""
    _free = "fp_obstacle_free"

    def __init__(self, vertex_positions, vertex_indices):
        pos = np.ascontiguousarray(np.asarray(vertex_positions, np.float32).reshape(-1, 3))
        idx = np.ascontiguousarray(np.asarray(vertex_indices, np.uint64).reshape(-1))
        h = C.c_void_p()
        check(lib().fp_obstacle_new(_fptr(pos), pos.shape[0], _uptr(idx), idx.size, C.byref(h)))
        super().__init__(h.value)

    def counts(self):
        v = [C.c_uint64() for _ in range(3)]
        check(lib().fp_obstacle_counts(self.handle, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)


class Simulation(_Handle):
    """springy/simulation.rs:10-78, device resident.  Takes ownership of the meshes and obstacles
    (their Python handles become unusable), as Simulation::new takes them by value."""
    _free = "fp_sim_free"

    def __init__(self, meshes: Sequence[SpringyMesh], obstacles: Sequence[CollidableMesh]):
        mh = [m._release() for m in meshes]
        oh = [o._release() for o in obstacles]
        marr = (C.c_void_p * max(1, len(mh)))(*mh)
        oarr = (C.c_void_p * max(1, len(oh)))(*oh)
        h = C.c_void_p()
        check(lib().fp_sim_new(marr, len(mh), oarr, len(oh), C.byref(h)))
        super().__init__(h.value)
        self.n_meshes = len(mh)

    # --- config (the reference sets it from its GUI, springy/simulation.rs:64-77) ---
    @property
    def config(self) -> Config:
        c = _lib.FpConfig()
        check(lib().fp_sim_get_config(self.handle, C.byref(c)))
        return Config.from_c(c)

    @config.setter
    def config(self, cfg: Config):
        c = cfg.to_c()
        check(lib().fp_sim_set_config(self.handle, C.byref(c)))

    def set_mesh_config(self, mesh: int, cfg: Config):
        c = cfg.to_c()
        check(lib().fp_sim_set_mesh_config(self.handle, mesh, C.byref(c)))

    def set_mesh_configs(self, cfgs):
        arr = (_lib.FpConfig * len(cfgs))(*[c.to_c() for c in cfgs])
        check(lib().fp_sim_set_mesh_configs(self.handle, arr, len(cfgs)))

    def autotune(self, steps_per_trial: int = 20):
        out = (C.c_int32 * 4)()
        check(lib().fp_sim_autotune(self.handle, steps_per_trial, out))
        return list(out)

    def step_graph(self, n_steps: int, chunk: int = 100):
        check(lib().fp_sim_step_graph(self.handle, n_steps, chunk))

    def time_steps_graph(self, n_steps: int, chunk: int = 100) -> float:
        ms = C.c_float()
        check(lib().fp_sim_time_steps_graph(self.handle, n_steps, chunk, C.byref(ms)))
        return ms.value

    def step(self, n_steps: int = 1) -> float:
        """Simulation::step (springy/simulation.rs:27-45); returns Duration(dt) in seconds."""
        ts = C.c_float()
        check(lib().fp_sim_step(self.handle, n_steps, C.byref(ts)))
        return ts.value

    def step_async(self, n_steps: int = 1):
        check(lib().fp_sim_step_async(self.handle, n_steps))

    def synchronize(self):
        check(lib().fp_sim_synchronize(self.handle))

    def get_timestep(self) -> float:
        ts = C.c_float()
        check(lib().fp_sim_get_timestep(self.handle, C.byref(ts)))
        return ts.value

    def mesh_counts(self, mesh: int = 0):
        v = [C.c_uint64() for _ in range(4)]
        check(lib().fp_sim_mesh_counts(self.handle, mesh, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def get_points(self, mesh: int = 0, out: np.ndarray = None) -> np.ndarray:
        """The mesh's 10-float records; `out`, when given (a C-contiguous float32 array of shape
        (points, 10), e.g. a pinned buffer a render loop reuses), receives them instead of a new array."""
        N = self.mesh_counts(mesh)[0]
        if out is None:
            out = np.zeros((N, 10), np.float32)
        elif out.shape != (N, 10) or out.dtype != np.float32 or not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError(f"out must be a writeable C-contiguous float32 array of shape ({N}, 10)")
        check(lib().fp_sim_get_points(self.handle, mesh, _fptr(out)))
        return out

    def set_points(self, mesh: int, records):
        rec = np.ascontiguousarray(np.asarray(records, np.float32).reshape(-1, 10))
        check(lib().fp_sim_set_points(self.handle, mesh, _fptr(rec), rec.shape[0]))

    def get_vertices(self, mesh: int = 0):
        N, _, F, _ = self.mesh_counts(mesh)
        pos = np.zeros((N, 3), np.float32)
        idx = np.zeros(max(3 * F, 1), np.uint64)
        check(lib().fp_sim_get_vertices(self.handle, mesh, _fptr(pos), _uptr(idx)))
        return pos, idx[:3 * F]

    def accumulate_forces(self, mesh: int = 0):
        check(lib().fp_sim_accumulate_forces(self.handle, mesh))

    def update_points(self, mesh: int, new_records):
        rec = np.ascontiguousarray(np.asarray(new_records, np.float32).reshape(-1, 10))
        check(lib().fp_sim_update_points(self.handle, mesh, _fptr(rec), rec.shape[0]))

    def apply_torsional_forces(self, mesh: int = 0):
        """SpringyMesh::apply_torsional_forces (springy_mesh.rs:553-627) on one mesh's force records."""
        check(lib().fp_sim_apply_torsional_forces(self.handle, mesh))

    def set_torsion(self, enabled: bool):
        """Restores the torsional forces the reference comments out of accumulate_forces
        (springy_mesh.rs:518-519), for every step; off by default, as in the reference."""
        check(lib().fp_sim_set_torsion(self.handle, 1 if enabled else 0))

    def clear_forces(self, mesh: int = 0):
        check(lib().fp_sim_clear_forces(self.handle, mesh))

    def set_kernel(self, kernel: int = _lib.FP_KERNEL_AUTO, math: int = _lib.FP_MATH_EXACT):
        check(lib().fp_sim_set_kernel(self.handle, kernel, math))

    def set_tile_shape(self, mesh: int, shape: int):
        check(lib().fp_sim_set_tile_shape(self.handle, mesh, shape))

    def get_tile_shape(self, mesh: int = 0) -> int:
        v = C.c_int32()
        check(lib().fp_sim_get_tile_shape(self.handle, mesh, C.byref(v)))
        return v.value

    def get_mesh_kernel(self, mesh: int = 0) -> int:
        """The kernel stepping `mesh` now: FP_STEP_TILED, FP_STEP_GRID_SIMPLE, FP_STEP_PATCH or
        FP_STEP_CSR."""
        v = C.c_int32()
        check(lib().fp_sim_get_mesh_kernel(self.handle, mesh, C.byref(v)))
        return v.value

    def count_nonfinite(self) -> np.ndarray:
        """Per mesh, the vertices whose position or velocity is not finite (a diagnostic the
        reference does not have; its steps, like these, let NaN propagate)."""
        out = np.zeros(max(1, self.n_meshes), np.uint64)
        check(lib().fp_sim_count_nonfinite(self.handle, _uptr(out)))
        return out[:self.n_meshes]

    def count_nonfinite_noalloc(self) -> np.ndarray:
        """count_nonfinite through the entry point that allocates no device memory (its counters are
        a fixed array of the library; at most 65,536 meshes)."""
        out = np.zeros(max(1, self.n_meshes), np.uint64)
        check(lib().fp_sim_count_nonfinite_noalloc(self.handle, _uptr(out)))
        return out[:self.n_meshes]

    def time_steps(self, n_steps: int) -> float:
        """Runs n_steps bracketed by HIP events on the simulation's stream; returns milliseconds."""
        ms = C.c_float()
        check(lib().fp_sim_time_steps(self.handle, n_steps, C.byref(ms)))
        return ms.value


    # --- learned surrogate (feriphys_amd.surrogate) ---
    def set_surrogate(self, net):
        """Copies the network into the simulation (None detaches)."""
        check(lib().fp_sim_set_surrogate(self.handle, None if net is None else net.handle))

    def set_surrogate_shape(self, vertices_per_wave: int):
        check(lib().fp_sim_set_surrogate_shape(self.handle, vertices_per_wave))

    def step_surrogate(self, n_big: int = 1):
        from .surrogate import SurrogateStats

        st = _lib.FpSurrogateStats()
        check(lib().fp_sim_step_surrogate(self.handle, n_big, C.byref(st)))
        return SurrogateStats.from_c(st)

    def time_surrogate(self, n_big: int):
        """(total wall ms, network-kernel ms from HIP events, stats) of n_big surrogate steps."""
        from .surrogate import SurrogateStats

        st = _lib.FpSurrogateStats()
        tot, mlp = C.c_float(), C.c_float()
        check(lib().fp_sim_time_surrogate(self.handle, n_big, C.byref(tot), C.byref(mlp), C.byref(st)))
        return tot.value, mlp.value, SurrogateStats.from_c(st)

    def surrogate_features(self, mesh: int = 0) -> np.ndarray:
        n = self.mesh_counts(mesh)[0]
        out = np.zeros((n, _lib.FP_SURROGATE_FEATURES), np.float32)
        check(lib().fp_sim_surrogate_features(self.handle, mesh, _fptr(out)))
        return out

    def surrogate_accel(self, mesh: int = 0) -> np.ndarray:
        n = self.mesh_counts(mesh)[0]
        out = np.zeros((n, 3), np.float32)
        check(lib().fp_sim_surrogate_accel(self.handle, mesh, _fptr(out)))
        return out

    def surrogate_slots(self, mesh: int = 0):
        n = self.mesh_counts(mesh)[0]
        nbr = np.zeros((n, 12), np.int64)
        rest = np.zeros((n, 12), np.float32)
        check(lib().fp_sim_surrogate_slots(self.handle, mesh, nbr.ctypes.data_as(C.POINTER(C.c_int64)), _fptr(rest)))
        return nbr, rest


def config_default() -> Config:
    c = _lib.FpConfig()
    lib().fp_config_default(C.byref(c))
    return Config.from_c(c)


def selftest_exact_math() -> dict:
    """Runs fp_selftest_exact_math on the current GPU: over every finite f32, how often the exact
    kernels' division-free forms differ from the operations they replace (all 0 expected)."""
    out = (C.c_uint64 * 5)()
    check(lib().fp_selftest_exact_math(out))
    return {"rcp": int(out[0]), "div_pi": int(out[1]), "div3": int(out[2]), "atan2": int(out[3]),
            "sqrt": int(out[4])}


class _PinnedBlock:
    """Page-locked host memory from fp_host_alloc, exposed as a NumPy array interface and released
    with fp_host_free when the last array viewing it is collected."""

    def __init__(self, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        self.ptr = C.c_void_p()
        check(lib().fp_host_alloc(n, C.byref(self.ptr)))
        self.__array_interface__ = {"shape": tuple(shape), "typestr": dt.str, "data": (self.ptr.value, False),
                                    "version": 3}

    def __del__(self):
        ptr, self.ptr = getattr(self, "ptr", None), None
        if ptr is not None and ptr.value:
            try:
                lib().fp_host_free(ptr)
            except Exception:  # interpreter shutdown: the library may already be unloaded
                pass


def pinned_empty(shape, dtype=np.float32) -> np.ndarray:
    """An uninitialised array in page-locked host memory (fp_host_alloc): get_points, set_points and
    get_vertices copy to and from it at the bus's full rate, as a render loop's reused staging
    buffer.  No counterpart in the reference."""
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    if int(np.prod(shape)) == 0:
        return np.empty(shape, dtype)
    return np.asarray(_PinnedBlock(shape, dtype))
