"""Helpers shared by the GPU parity tests: build a scene on both sides, compare records."""
import numpy as np

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def gpu_scene(scene, kernel=fa.FP_KERNEL_AUTO, math=fa.FP_MATH_EXACT):
    sim = wl.build_gpu(scene)
    sim.set_kernel(kernel, math)
    return sim


def oracle_scene(scene, dtype=np.float32):
    return orc.build_scene(scene, dtype)


def force_rel_err(f_gpu, f_ref):
    """max |F_gpu - F_ref| over all components, relative to the largest force component."""
    scale = max(float(np.abs(f_ref).max()), 1e-30)
    return float(np.abs(f_gpu.astype(np.float64) - f_ref.astype(np.float64)).max()) / scale


def announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
