"""The learned large-step surrogate (the fp_surrogate_* and fp_sim_*surrogate* entry points of
include/feriphys_gpu.h): a per-vertex bf16 MLP trained on the GPU from plain-solver trajectories,
stepping springy meshes with H = ratio * dt and falling back to plain steps per mesh when a step's
strain or energy residual is too large.  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib

N_PARAMS = _lib.FP_SURROGATE_PARAMS
N_FEATURES = _lib.FP_SURROGATE_FEATURES


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class SurrogateConfig:
    dt: float = 1.0 / 600.0
    ratio: int = 8
    strain_scale: float = 10.0
    out_scale: float = 0.0  # 0: the library default for dt and ratio
    base_struts: bool = False
    strain_limit: float = 0.02
    energy_limit: float = 0.002

    def to_c(self) -> _lib.FpSurrogateConfig:
        d = _lib.FpSurrogateConfig()
        lib().fp_surrogate_config_default(C.byref(d))
        H = np.float32(np.float32(self.dt) * np.float32(self.ratio))
        out = self.out_scale if self.out_scale > 0 else float(np.float32(0.005) / (H * H))
        return _lib.FpSurrogateConfig(self.dt, self.ratio, self.strain_scale, out, int(self.base_struts),
                                      self.strain_limit, self.energy_limit)

    @staticmethod
    def from_c(c: _lib.FpSurrogateConfig) -> "SurrogateConfig":
        return SurrogateConfig(c.dt, c.ratio, c.strain_scale, c.out_scale, bool(c.base_struts), c.strain_limit,
                               c.energy_limit)


@dataclass
class SurrogateStats:
    big_steps: int
    fallbacks: int
    mesh_fallbacks: int
    plain_only: int
    max_strain_residual: float
    max_energy_residual: float

    @staticmethod
    def from_c(s: _lib.FpSurrogateStats) -> "SurrogateStats":
        return SurrogateStats(s.big_steps, s.fallbacks, s.mesh_fallbacks, s.plain_only, s.max_strain_residual,
                              s.max_energy_residual)


class Surrogate:
    """Network weights, optimiser state and training set, all device resident."""

    def __init__(self, config: SurrogateConfig | None = None, seed: int = 0, params=None):
        c = (config or SurrogateConfig()).to_c()
        h = C.c_void_p()
        p = None
        if params is not None:
            self._init = np.ascontiguousarray(params, dtype=np.float32).reshape(N_PARAMS)
            p = _fptr(self._init)
        check(lib().fp_surrogate_new(C.byref(c), seed, p, C.byref(h)))
        self.handle = h.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                lib().fp_surrogate_free(h)
            except Exception:
                pass
            self.handle = None

    @property
    def config(self) -> SurrogateConfig:
        c = _lib.FpSurrogateConfig()
        check(lib().fp_surrogate_get_config(self.handle, C.byref(c)))
        return SurrogateConfig.from_c(c)

    def get_params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        check(lib().fp_surrogate_get_params(self.handle, _fptr(out)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(N_PARAMS)
        check(lib().fp_surrogate_set_params(self.handle, _fptr(p)))

    def grad(self, X, T):
        """GPU forward + backward on (X [rows, 72], T [rows, 3]): (loss, grad [42755], Y [rows, 3])."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        rows = X.shape[0]
        loss = np.zeros(1, np.float32)
        g = np.zeros(N_PARAMS, np.float32)
        Y = np.zeros((rows, 3), np.float32)
        check(lib().fp_surrogate_grad(self.handle, _fptr(X), _fptr(T), rows, _fptr(loss), _fptr(g), _fptr(Y)))
        return float(loss[0]), g, Y

    def reserve(self, capacity: int):
        check(lib().fp_surrogate_reserve(self.handle, capacity))

    def sample_count(self) -> int:
        v = C.c_uint64()
        check(lib().fp_surrogate_sample_count(self.handle, C.byref(v)))
        return v.value

    def clear_samples(self):
        check(lib().fp_surrogate_clear_samples(self.handle))

    def add_samples(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        check(lib().fp_surrogate_add_samples(self.handle, _fptr(X), _fptr(T), X.shape[0]))

    def train(self, steps: int, batch: int = 8192, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(1, steps), np.float32)
        check(lib().fp_surrogate_train(self.handle, steps, batch, lr, seed, _fptr(losses)))
        return losses[:steps]

    def collect(self, sim, n_big: int, noise: float = 1e-3, seed: int = 0):
        """Adds training samples from n_big * ratio plain steps of `sim` (a Simulation)."""
        check(lib().fp_sim_collect_samples(sim.handle, self.handle, n_big, noise, seed))
