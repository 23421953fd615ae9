"""CPU ORACLE — test infrastructure only.

ctypes wrapper over oracle/liboracle.so, the C++ restatement of the reference springy step
(springy_oracle.hpp).  Imported by the test-suite, by ``__graft_entry__.smoke()`` as the checker
and by bench.py's cpu_baseline leg; the product package never imports it.

Every class takes ``dtype`` = np.float32 (the reference's arithmetic) or np.float64 (the f64 run).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_P = C.c_void_p
_U64P = C.POINTER(C.c_uint64)
_I64P = C.POINTER(C.c_int64)
_U8P = C.POINTER(C.c_uint8)


def _cfg_type(ct):
    class CConfig(C.Structure):
        _fields_ = [("integration", C.c_int), ("dt", ct), ("gravity", ct * 3), ("wind", ct * 3),
                    ("lift_coefficient", ct), ("drag_coefficient", ct),
                    ("coefficient_of_restitution", ct), ("coefficient_of_friction", ct)]
    return CConfig


_CFG = {np.float32: _cfg_type(C.c_float), np.float64: _cfg_type(C.c_double)}
_CT = {np.float32: C.c_float, np.float64: C.c_double}
_SFX = {np.float32: "_f32", np.float64: "_f64"}


def _declare(L):
    for dt, sfx in _SFX.items():
        ct = _CT[dt]
        RP = C.POINTER(ct)
        CP = C.POINTER(_CFG[dt])
        sig = {
            "orc_mesh_new": (_P, [RP, C.c_size_t, _U64P, C.c_size_t, ct, ct, ct, C.c_int, ct, ct, _U64P, RP,
                                  C.c_size_t, C.c_int, C.c_char_p, C.c_size_t]),
            "orc_cloth_new": (_P, [C.c_size_t, C.c_size_t, ct, RP, ct, ct, ct, ct, ct, ct, ct, _U64P, C.c_size_t,
                                   C.c_int, C.c_char_p, C.c_size_t]),
            "orc_mesh_free": (None, [_P]),
            "orc_mesh_add_strut": (C.c_int, [_P, C.c_uint64, C.c_uint64, ct, ct]),
            "orc_mesh_add_pin": (None, [_P, C.c_uint64]),
            "orc_mesh_counts": (None, [_P, _U64P]),
            "orc_mesh_struts": (None, [_P, _U64P, RP, _I64P, _U8P, RP]),
            "orc_mesh_set_torsion_rest": (None, [_P, C.c_uint64, ct]),
            "orc_mesh_faces": (None, [_P, _U64P, _U64P]),
            "orc_mesh_pins": (None, [_P, _U64P]),
            "orc_mesh_points": (None, [_P, RP]),
            "orc_mesh_set_points": (C.c_int, [_P, RP, C.c_size_t]),
            "orc_mesh_accumulate_forces": (C.c_int, [_P, CP]),
            "orc_mesh_apply_strut_forces": (None, [_P]),
            "orc_mesh_apply_face_forces": (None, [_P, CP]),
            "orc_mesh_apply_torsional_forces": (C.c_int, [_P]),
            "orc_mesh_set_torsion_in_step": (None, [_P, C.c_int]),
            "orc_mesh_clear_forces": (None, [_P]),
            "orc_mesh_update_points": (C.c_int, [_P, RP, C.c_size_t, C.POINTER(_P), C.c_size_t, CP]),
            "orc_mesh_integrate": (C.c_int, [_P, CP, RP]),
            "orc_mesh_face_angle": (ct, [_P, C.c_uint64, C.c_int]),
            "orc_mesh_face_normal": (None, [_P, C.c_uint64, RP]),
            "orc_mesh_get_vertices": (None, [_P, RP, _U64P]),
            "orc_obstacle_new": (_P, [RP, C.c_size_t, _U64P, C.c_size_t, C.c_char_p, C.c_size_t]),
            "orc_obstacle_free": (None, [_P]),
            "orc_obstacle_counts": (None, [_P, _U64P]),
            "orc_obstacle_edges": (None, [_P, RP]),
            "orc_obstacle_faces": (None, [_P, RP]),
            "orc_collided_face": (C.c_int64, [C.POINTER(_P), C.c_size_t, RP, RP, ct]),
            "orc_collided_faces": (None, [C.POINTER(_P), C.c_size_t, RP, RP, C.c_size_t, ct, _I64P]),
            "orc_duration_roundtrip": (ct, [ct]),
            "orc_sim_new": (_P, [C.POINTER(_P), C.c_size_t, C.POINTER(_P), C.c_size_t]),
            "orc_sim_free": (None, [_P]),
            "orc_sim_set_config": (None, [_P, CP]),
            "orc_sim_step": (C.c_int, [_P, C.c_uint64]),
            "orc_sim_mesh": (_P, [_P, C.c_uint64]),
            "orc_rigid_sim_new": (_P, [RP, RP, C.c_size_t, C.POINTER(_P), C.c_size_t]),
            "orc_rigid_sim_free": (None, [_P]),
            "orc_rigid_sim_set_config": (None, [_P, C.c_int, ct, ct, RP, RP]),
            "orc_rigid_sim_step": (C.c_int, [_P, C.c_uint64]),
            "orc_rigid_sim_get_states": (None, [_P, RP]),
            "orc_rigid_sim_set_states": (None, [_P, RP]),
            "orc_rigid_sim_apply_impulse": (None, [_P, C.c_uint64, RP, RP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name + sfx)
            f.restype = res
            f.argtypes = args
    L.orc_state_test_euler.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_state_test_euler.restype = None
    L.orc_state_test_rk4.argtypes = [C.POINTER(C.c_float)]
    L.orc_state_test_rk4.restype = None
    L.orc_tuple_difference.argtypes = [C.c_int64] * 5
    L.orc_tuple_difference.restype = C.c_int64


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _fn(name, dtype):
    return getattr(lib(), name + _SFX[dtype])


@dataclass
class Config:
    """springy/config.rs:9-33 (defaults from :20-33)."""
    integration: int = 1  # 0 Euler, 1 Rk4
    dt: float = 0.001
    gravity: tuple = (0.0, -10.0, 0.0)
    wind: tuple = (0.0, 0.0, 0.0)
    lift_coefficient: float = 1.0
    drag_coefficient: float = 1.0
    coefficient_of_restitution: float = 0.95
    coefficient_of_friction: float = 0.3

    def c(self, dtype):
        npt = np.dtype(dtype).type
        cfg = _CFG[npt]()
        cfg.integration = int(self.integration)
        # Values are first rounded to f32, as the reference stores f32; the f64 run widens them.
        f = lambda v: float(np.float32(v)) if npt is np.float32 else float(v)
        cfg.dt = f(self.dt)
        for i in range(3):
            cfg.gravity[i] = f(self.gravity[i])
            cfg.wind[i] = f(self.wind[i])
        cfg.lift_coefficient = f(self.lift_coefficient)
        cfg.drag_coefficient = f(self.drag_coefficient)
        cfg.coefficient_of_restitution = f(self.coefficient_of_restitution)
        cfg.coefficient_of_friction = f(self.coefficient_of_friction)
        return cfg


class Mesh:
    """SpringyMesh (springy_mesh.rs:258-661)."""

    def __init__(self, handle, dtype, owned=True):
        self.h = handle
        self.dtype = np.dtype(dtype).type
        self.owned = owned

    def __del__(self):
        if getattr(self, "owned", False) and self.h:
            _fn("orc_mesh_free", self.dtype)(self.h)
            self.h = None

    @classmethod
    def new(cls, positions, indices, mass, k, c, torsion=None, overrides=None, dtype=np.float32):
        dt = np.dtype(dtype).type
        ct = _CT[dt]
        pos = np.ascontiguousarray(np.asarray(positions, dtype=dt).reshape(-1, 3))
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64).reshape(-1))
        ovr_pairs = np.zeros((max(1, len(overrides or [])), 2), dtype=np.uint64)
        ovr_kc = np.zeros((max(1, len(overrides or [])), 2), dtype=dt)
        for i, (a, b, kk, cc) in enumerate(overrides or []):
            ovr_pairs[i] = (a, b)
            ovr_kc[i] = (kk, cc)
        err = C.create_string_buffer(256)
        tk, tc = torsion if torsion is not None else (0.0, 0.0)
        h = _fn("orc_mesh_new", dt)(_ptr(pos, ct), pos.shape[0], _ptr(idx, C.c_uint64), idx.size, mass, k, c,
                                     1 if torsion is not None else 0, tk, tc, _ptr(ovr_pairs, C.c_uint64),
                                     _ptr(ovr_kc, ct), len(overrides or []), 1 if overrides is not None else 0,
                                     err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h, dt)

    @classmethod
    def cloth(cls, rows, cols, spacing, position, point_mass, tk, tc, sk, sc, bk, bc, pins, plane=0,
              dtype=np.float32):
        dt = np.dtype(dtype).type
        ct = _CT[dt]
        p = np.asarray(position, dtype=dt)
        pins_a = np.asarray(list(pins) or [0], dtype=np.uint64)
        err = C.create_string_buffer(256)
        h = _fn("orc_cloth_new", dt)(rows, cols, spacing, _ptr(p, ct), point_mass, tk, tc, sk, sc, bk, bc,
                                      _ptr(pins_a, C.c_uint64), len(list(pins)), plane, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h, dt)

    def add_strut(self, a, b, k, c):
        if _fn("orc_mesh_add_strut", self.dtype)(self.h, a, b, k, c):
            raise IndexError("add_strut index out of range")

    def add_pin(self, i):
        _fn("orc_mesh_add_pin", self.dtype)(self.h, i)

    def counts(self):
        out = np.zeros(4, dtype=np.uint64)
        _fn("orc_mesh_counts", self.dtype)(self.h, _ptr(out, C.c_uint64))
        return tuple(int(x) for x in out)

    def struts(self):
        """Returns dict of arrays: v01 (S,2), kcl (S,3: stiffness, damping, length), faces (S,2; -1=None),
        has_torsion (S,), torsion (S,3: k, c, rest_angle)."""
        ct = _CT[self.dtype]
        S = self.counts()[1]
        n = max(S, 1)
        v01 = np.zeros((n, 2), np.uint64)
        kcl = np.zeros((n, 3), self.dtype)
        fc = np.zeros((n, 2), np.int64)
        ht = np.zeros(n, np.uint8)
        tor = np.zeros((n, 3), self.dtype)
        _fn("orc_mesh_struts", self.dtype)(self.h, _ptr(v01, C.c_uint64), _ptr(kcl, ct), _ptr(fc, C.c_int64),
                                            _ptr(ht, C.c_uint8), _ptr(tor, ct))
        return dict(v01=v01[:S], kcl=kcl[:S], faces=fc[:S], has_torsion=ht[:S].astype(bool), torsion=tor[:S])

    def set_torsion_rest(self, strut, rest):
        _fn("orc_mesh_set_torsion_rest", self.dtype)(self.h, strut, rest)

    def faces(self):
        F = self.counts()[2]
        n = max(F, 1)
        s3 = np.zeros((n, 3), np.uint64)
        v3 = np.zeros((n, 3), np.uint64)
        _fn("orc_mesh_faces", self.dtype)(self.h, _ptr(s3, C.c_uint64), _ptr(v3, C.c_uint64))
        return s3[:F], v3[:F]

    def pins(self):
        P = self.counts()[3]
        out = np.zeros(max(P, 1), np.uint64)
        _fn("orc_mesh_pins", self.dtype)(self.h, _ptr(out, C.c_uint64))
        return out[:P]

    def points(self):
        N = self.counts()[0]
        out = np.zeros((N, 10), self.dtype)
        _fn("orc_mesh_points", self.dtype)(self.h, _ptr(out, _CT[self.dtype]))
        return out

    def set_points(self, rec):
        rec = np.ascontiguousarray(np.asarray(rec, dtype=self.dtype).reshape(-1, 10))
        if _fn("orc_mesh_set_points", self.dtype)(self.h, _ptr(rec, _CT[self.dtype]), rec.shape[0]):
            raise ValueError("point count mismatch")

    def accumulate_forces(self, cfg: Config):
        c = cfg.c(self.dtype)
        if _fn("orc_mesh_accumulate_forces", self.dtype)(self.h, C.byref(c)):
            raise IndexError("pin index out of range")

    def apply_strut_forces(self):
        _fn("orc_mesh_apply_strut_forces", self.dtype)(self.h)

    def apply_face_forces(self, cfg: Config):
        c = cfg.c(self.dtype)
        _fn("orc_mesh_apply_face_forces", self.dtype)(self.h, C.byref(c))

    def set_torsion_in_step(self, on: bool):
        """Restores the apply_torsional_forces call the reference comments out of accumulate_forces
        (springy_mesh.rs:518-519), between the strut and the face forces."""
        _fn("orc_mesh_set_torsion_in_step", self.dtype)(self.h, 1 if on else 0)

    def apply_torsional_forces(self):
        if _fn("orc_mesh_apply_torsional_forces", self.dtype)(self.h):
            raise RuntimeError("torsion failed")

    def clear_forces(self):
        _fn("orc_mesh_clear_forces", self.dtype)(self.h)

    def integrate(self, cfg: Config):
        """State::new(points).{rk4,euler}_step(dt).get_elements() as (N,10) records."""
        N = self.counts()[0]
        out = np.zeros((N, 10), self.dtype)
        c = cfg.c(self.dtype)
        _fn("orc_mesh_integrate", self.dtype)(self.h, C.byref(c), _ptr(out, _CT[self.dtype]))
        return out

    def update_points(self, new_rec, obstacles, cfg: Config):
        new_rec = np.ascontiguousarray(np.asarray(new_rec, dtype=self.dtype).reshape(-1, 10))
        arr = (_P * max(1, len(obstacles)))(*[o.h for o in obstacles])
        c = cfg.c(self.dtype)
        if _fn("orc_mesh_update_points", self.dtype)(self.h, _ptr(new_rec, _CT[self.dtype]), new_rec.shape[0], arr,
                                                      len(obstacles), C.byref(c)):
            raise IndexError("pin index out of range")

    def face_angle(self, face, corner):
        return _fn("orc_mesh_face_angle", self.dtype)(self.h, face, corner)

    def face_normal(self, face):
        out = np.zeros(3, self.dtype)
        _fn("orc_mesh_face_normal", self.dtype)(self.h, face, _ptr(out, _CT[self.dtype]))
        return out

    def get_vertices(self):
        N, _, F, _ = self.counts()
        pos = np.zeros((N, 3), self.dtype)
        idx = np.zeros(max(3 * F, 1), np.uint64)
        _fn("orc_mesh_get_vertices", self.dtype)(self.h, _ptr(pos, _CT[self.dtype]), _ptr(idx, C.c_uint64))
        return pos, idx[:3 * F]


class Obstacle:
    """CollidableMesh (collidable_mesh.rs:50-197)."""

    def __init__(self, positions, indices, dtype=np.float32):
        self.dtype = np.dtype(dtype).type
        ct = _CT[self.dtype]
        pos = np.ascontiguousarray(np.asarray(positions, dtype=self.dtype).reshape(-1, 3))
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64).reshape(-1))
        err = C.create_string_buffer(256)
        self.h = _fn("orc_obstacle_new", self.dtype)(_ptr(pos, ct), pos.shape[0], _ptr(idx, C.c_uint64), idx.size,
                                                     err, 256)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            _fn("orc_obstacle_free", self.dtype)(self.h)
            self.h = None

    def counts(self):
        out = np.zeros(3, np.uint64)
        _fn("orc_obstacle_counts", self.dtype)(self.h, _ptr(out, C.c_uint64))
        return tuple(int(x) for x in out)

    def edges(self):
        E = self.counts()[1]
        out = np.zeros((max(E, 1), 2, 3), self.dtype)
        _fn("orc_obstacle_edges", self.dtype)(self.h, _ptr(out, _CT[self.dtype]))
        return out[:E]

    def faces(self):
        F = self.counts()[2]
        out = np.zeros((max(F, 1), 3, 3), self.dtype)
        _fn("orc_obstacle_faces", self.dtype)(self.h, _ptr(out, _CT[self.dtype]))
        return out[:F]


def collided_faces(obstacles, old, new, dt_d, dtype=np.float32):
    """First-hit face index per segment over the concatenated face list (-1 = none)."""
    dt = np.dtype(dtype).type
    ct = _CT[dt]
    old = np.ascontiguousarray(np.asarray(old, dtype=dt).reshape(-1, 3))
    new = np.ascontiguousarray(np.asarray(new, dtype=dt).reshape(-1, 3))
    out = np.zeros(old.shape[0], np.int64)
    arr = (_P * max(1, len(obstacles)))(*[o.h for o in obstacles])
    _fn("orc_collided_faces", dt)(arr, len(obstacles), _ptr(old, ct), _ptr(new, ct), old.shape[0], dt_d,
                                  _ptr(out, C.c_int64))
    return out


def duration_roundtrip(dt, dtype=np.float32):
    return _fn("orc_duration_roundtrip", dtype)(dt)


class Simulation:
    """springy::Simulation (springy/simulation.rs:10-78); meshes and obstacles are copied in."""

    def __init__(self, meshes, obstacles, dtype=np.float32):
        self.dtype = np.dtype(dtype).type
        ma = (_P * max(1, len(meshes)))(*[m.h for m in meshes])
        oa = (_P * max(1, len(obstacles)))(*[o.h for o in obstacles])
        self.h = _fn("orc_sim_new", self.dtype)(ma, len(meshes), oa, len(obstacles))
        self.n_meshes = len(meshes)

    def __del__(self):
        if getattr(self, "h", None):
            _fn("orc_sim_free", self.dtype)(self.h)
            self.h = None

    def set_config(self, cfg: Config):
        c = cfg.c(self.dtype)
        _fn("orc_sim_set_config", self.dtype)(self.h, C.byref(c))

    def step(self, n=1):
        if _fn("orc_sim_step", self.dtype)(self.h, n):
            raise IndexError("step failed (index out of range)")

    def mesh(self, i) -> Mesh:
        return Mesh(_fn("orc_sim_mesh", self.dtype)(self.h, i), self.dtype, owned=False)


def build_scene_per_mesh(scene, dtype=np.float32):
    """One oracle Simulation per cloth (a reference Simulation holds one Config; per-cloth winds
    need one each), with the scene's jitter, obstacles and per-mesh configs."""
    from feriphys_amd.workloads import jitter_records, mesh_configs

    dt = np.dtype(dtype).type
    sims = []
    for k, (c, cfg) in enumerate(zip(scene.cloths, mesh_configs(scene))):
        m = Mesh.cloth(*c.args(), plane=c.plane, dtype=np.float32)
        rec = m.points()
        if scene.jitter:
            rec = jitter_records(rec, c.pins, scene.jitter, scene.seed + k)
        if dt is not np.float32:
            m = Mesh.cloth(*c.args(), plane=c.plane, dtype=dt)
        m.set_points(rec.astype(dt))
        obs = [Obstacle(np.asarray(p, np.float32).astype(dt), i, dtype=dt) for p, i in scene.obstacles]
        sim = Simulation([m], obs, dtype=dt)
        sim.set_config(Config(cfg.integration, cfg.dt, cfg.gravity, cfg.wind, cfg.lift_coefficient,
                              cfg.drag_coefficient, cfg.coefficient_of_restitution, cfg.coefficient_of_friction))
        sims.append(sim)
    return sims


def build_surface_scene(scene, dtype=np.float32) -> Simulation:
    """The oracle Simulation of a feriphys_amd.workloads.SurfaceScene (SpringyMesh::new from the
    triangle list, then add_strut across every shared edge in strut order)."""
    from feriphys_amd.workloads import bend_pairs

    dt = np.dtype(dtype).type
    meshes = []
    for spec in scene.meshes:
        m = Mesh.new(spec.positions.astype(dt), spec.indices, spec.mass, spec.k, spec.c, dtype=dt)
        st = m.struts()
        _, fv = m.faces()
        for a, b in bend_pairs(st["v01"], st["faces"], fv):
            m.add_strut(int(a), int(b), spec.bk, spec.bc)
        meshes.append(m)
    obs = [Obstacle(np.asarray(p, np.float32).astype(dt), i, dtype=dt) for p, i in scene.obstacles]
    sim = Simulation(meshes, obs, dtype=dt)
    c = scene.config
    sim.set_config(Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                          c.coefficient_of_restitution, c.coefficient_of_friction))
    return sim


def build_scene(scene, dtype=np.float32) -> Simulation:
    """The oracle Simulation of a feriphys_amd.workloads.Scene: the same Cloth::new arguments, the
    same seeded jitter (computed on the f32 records; the f64 run starts from them widened), the
    same obstacles and Config."""
    from feriphys_amd.workloads import jitter_records

    dt = np.dtype(dtype).type
    meshes = []
    for k, c in enumerate(scene.cloths):
        m32 = Mesh.cloth(*c.args(), plane=c.plane, dtype=np.float32)
        rec = m32.points()
        if scene.jitter:
            rec = jitter_records(rec, c.pins, scene.jitter, scene.seed + k)
        m = m32 if dt is np.float32 else Mesh.cloth(*c.args(), plane=c.plane, dtype=dt)
        m.set_points(rec.astype(dt))
        meshes.append(m)
    obs = [Obstacle(np.asarray(p, np.float32).astype(dt), i, dtype=dt) for p, i in scene.obstacles]
    sim = Simulation(meshes, obs, dtype=dt)
    c = scene.config
    sim.set_config(Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                          c.coefficient_of_restitution, c.coefficient_of_friction))
    return sim


def state_test_euler():
    init = (C.c_float * 6)()
    out = (C.c_float * 6)()
    lib().orc_state_test_euler(init, out)
    return np.array(init[:], np.float32), np.array(out[:], np.float32)


def state_test_rk4():
    out = (C.c_float * 12)()
    lib().orc_state_test_rk4(out)
    return np.array(out[:], np.float32).reshape(4, 3)


def tuple_difference(trio, duo):
    r = lib().orc_tuple_difference(*trio, *duo)
    if r < 0:
        raise AssertionError("tuple_difference: duo is not a subset of trio")
    return r


class RigidSimulation:
    """rigidbody::Simulation (rigidbody/simulation.rs:10-52) of unit cubes (RigidBody::new,
    rigidbody.rs:213-251) among static obstacles; states are the 29-float Stateful records
    (rigidbody.rs:62-93)."""

    def __init__(self, positions, masses, obstacles=(), dtype=np.float32):
        self.dtype = np.dtype(dtype).type
        ct = _CT[self.dtype]
        pos = np.ascontiguousarray(np.asarray(positions, dtype=self.dtype).reshape(-1, 3))
        m = np.ascontiguousarray(np.asarray(masses, dtype=self.dtype).reshape(-1))
        self.n = pos.shape[0]
        self._obstacles = list(obstacles)
        oa = (_P * max(1, len(obstacles)))(*[o.h for o in obstacles])
        self.h = _fn("orc_rigid_sim_new", self.dtype)(_ptr(pos, ct), _ptr(m, ct), self.n, oa, len(obstacles))
        if not self.h:
            raise ValueError("Uninvertable moment of inertia!")

    def __del__(self):
        if getattr(self, "h", None):
            _fn("orc_rigid_sim_free", self.dtype)(self.h)
            self.h = None

    def set_config(self, integration=1, dt=0.001, restitution=0.7, gravity=(0, 0, 0), torque=(0, 0, 0)):
        ct = _CT[self.dtype]
        g = np.asarray(gravity, self.dtype)
        t = np.asarray(torque, self.dtype)
        _fn("orc_rigid_sim_set_config", self.dtype)(self.h, int(integration), self.dtype(dt), self.dtype(restitution),
                                                    _ptr(g, ct), _ptr(t, ct))

    def step(self, n=1):
        if _fn("orc_rigid_sim_step", self.dtype)(self.h, n):
            raise RuntimeError("rigid-body step failed")

    def states(self):
        out = np.zeros((self.n, 29), self.dtype)
        _fn("orc_rigid_sim_get_states", self.dtype)(self.h, _ptr(out, _CT[self.dtype]))
        return out

    def set_states(self, states):
        st = np.ascontiguousarray(np.asarray(states, self.dtype).reshape(self.n, 29))
        _fn("orc_rigid_sim_set_states", self.dtype)(self.h, _ptr(st, _CT[self.dtype]))

    def apply_impulse(self, body, impulse, at):
        ct = _CT[self.dtype]
        i = np.asarray(impulse, self.dtype)
        a = np.asarray(at, self.dtype)
        _fn("orc_rigid_sim_apply_impulse", self.dtype)(self.h, int(body), _ptr(i, ct), _ptr(a, ct))
