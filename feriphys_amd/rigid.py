"""Rigid bodies on the GPU: a ctypes mirror of the reference's rigidbody module
(src/simulation/rigidbody/{rigidbody,simulation,config}.rs) over the C ABI of
include/feriphys_gpu.h.  A body is the reference's unit cube; its state is the 29-float Stateful
record (rigidbody.rs:62-93)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check, lib
from .springy import CollidableMesh, Integration, _Handle, _fptr

RECORD = _lib.FP_RIGID_RECORD


@dataclass
class RigidConfig:
    """rigidbody/config.rs:7-25 (Default: RK4, dt 1 ms, restitution 0.7, no gravity or torque)."""
    integration: Integration = Integration.Rk4
    dt: float = 0.001
    coefficient_of_restitution: float = 0.7
    gravity: tuple = (0.0, 0.0, 0.0)
    torque: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> _lib.FpRigidConfig:
        return _lib.FpRigidConfig(int(self.integration), self.dt, self.coefficient_of_restitution,
                                  (C.c_float * 3)(*self.gravity), (C.c_float * 3)(*self.torque))

    @staticmethod
    def from_c(c: _lib.FpRigidConfig) -> "RigidConfig":
        return RigidConfig(Integration(c.integration), c.dt, c.coefficient_of_restitution, tuple(c.gravity),
                           tuple(c.torque))


def rigid_config_default() -> RigidConfig:
    c = _lib.FpRigidConfig()
    lib().fp_rigid_config_default(C.byref(c))
    return RigidConfig.from_c(c)


class RigidBody(_Handle):
    """RigidBody::new (rigidbody.rs:213-251): a unit cube of `mass` at `position`, at rest."""
    _free = "fp_rigid_free"

    def __init__(self, position, mass: float):
        p = np.ascontiguousarray(np.asarray(position, np.float32).reshape(3))
        h = C.c_void_p()
        check(lib().fp_rigid_new(_fptr(p), mass, C.byref(h)))
        super().__init__(h.value)

    def state(self) -> np.ndarray:
        out = np.zeros(RECORD, np.float32)
        check(lib().fp_rigid_get_state(self.handle, _fptr(out)))
        return out

    def apply_impulse(self, impulse, position):
        i = np.ascontiguousarray(np.asarray(impulse, np.float32).reshape(3))
        a = np.ascontiguousarray(np.asarray(position, np.float32).reshape(3))
        check(lib().fp_rigid_apply_impulse(self.handle, _fptr(i), _fptr(a)))


class RigidSimulation(_Handle):
    """rigidbody::Simulation (rigidbody/simulation.rs:10-78), device resident.  Takes ownership of
    the bodies and obstacles (their Python handles become unusable)."""
    _free = "fp_rigid_sim_free"

    def __init__(self, bodies: Sequence[RigidBody], obstacles: Sequence[CollidableMesh] = ()):
        bh = [b._release() for b in bodies]
        oh = [o._release() for o in obstacles]
        barr = (C.c_void_p * max(1, len(bh)))(*bh)
        oarr = (C.c_void_p * max(1, len(oh)))(*oh)
        h = C.c_void_p()
        check(lib().fp_rigid_sim_new(barr, len(bh), oarr, len(oh), C.byref(h)))
        super().__init__(h.value)
        self.n = len(bh)

    @property
    def config(self) -> RigidConfig:
        c = _lib.FpRigidConfig()
        check(lib().fp_rigid_sim_get_config(self.handle, C.byref(c)))
        return RigidConfig.from_c(c)

    @config.setter
    def config(self, cfg: RigidConfig):
        c = cfg.to_c()
        check(lib().fp_rigid_sim_set_config(self.handle, C.byref(c)))

    def step(self, n_steps: int = 1) -> float:
        """Simulation::step n times; returns Duration(dt) in seconds."""
        ts = C.c_float()
        check(lib().fp_rigid_sim_step(self.handle, n_steps, C.byref(ts)))
        return ts.value

    def time_steps(self, n_steps: int) -> float:
        ms = C.c_float()
        check(lib().fp_rigid_sim_time_steps(self.handle, n_steps, C.byref(ms)))
        return ms.value

    def states(self) -> np.ndarray:
        out = np.zeros((self.n, RECORD), np.float32)
        check(lib().fp_rigid_sim_get_states(self.handle, _fptr(out)))
        return out

    def set_states(self, records):
        r = np.ascontiguousarray(np.asarray(records, np.float32).reshape(-1, RECORD))
        check(lib().fp_rigid_sim_set_states(self.handle, _fptr(r), r.shape[0]))

    def apply_impulse(self, body: int, impulse, position):
        i = np.ascontiguousarray(np.asarray(impulse, np.float32).reshape(3))
        a = np.ascontiguousarray(np.asarray(position, np.float32).reshape(3))
        check(lib().fp_rigid_sim_apply_impulse(self.handle, body, _fptr(i), _fptr(a)))
