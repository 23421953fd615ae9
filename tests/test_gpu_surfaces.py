"""GPU parity on irregular meshes (the patch kernel, and the CSR kernel with its face kernel):
seeded, radially perturbed icospheres with tensile struts on edges and bending struts across
shared edges, dropped onto a ground plane; a fan whose hub has more faces than the patch kernel
takes."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR])
@pytest.mark.parametrize("math", [fa.FP_MATH_EXACT, fa.FP_MATH_FAST])
def test_icospheres_on_ground_match_oracle(math, kernel):
    sc = wl.config4(n_meshes=5, seed=3, levels=[2, 3, 3, 4, 1])
    for m in sc.meshes:  # bring them close to the ground so they land within the run
        m.positions[:, 1] -= m.positions[:, 1].min() - 0.01
    n = sum(m.positions.shape[0] for m in sc.meshes)
    announce(f"icospheres: {len(sc.meshes)} meshes, {n} vertices, math {math}, kernel {kernel}")
    g = wl.build_surface_gpu(sc)
    g.set_kernel(kernel, math)
    o = orc.build_surface_scene(sc)
    steps = 240
    g.step(steps)
    o.step(steps)
    worst = 0.0
    for k in range(len(sc.meshes)):
        a, b = g.get_points(k), o.mesh(k).points()
        assert np.isfinite(a).all()
        if math == fa.FP_MATH_EXACT:
            assert (bits(a) == bits(b)).all(), k
        worst = max(worst, float(np.abs(a[:, 1:4].astype(np.float64) - b[:, 1:4]).max()))
        assert b[:, 2].min() > -1e-3  # resting on the ground plane, not through it
    announce(f"  max position difference {worst:.2e} m")
    assert worst <= 5e-3


def test_icosphere_one_step_forces_match_oracle():
    sc = wl.config4(n_meshes=2, seed=5, levels=[3, 2])
    g = wl.build_surface_gpu(sc)
    o = orc.build_surface_scene(sc)
    rng = np.random.default_rng(0)
    for k in range(2):
        rec = o.mesh(k).points()
        rec[:, 4:7] = rng.normal(scale=0.3, size=(rec.shape[0], 3)).astype(np.float32)
        o.mesh(k).set_points(rec)
        g.set_points(k, rec)
        c = sc.config
        o.mesh(k).accumulate_forces(orc.Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient,
                                               c.drag_coefficient, c.coefficient_of_restitution,
                                               c.coefficient_of_friction))
        g.accumulate_forces(k)
        assert (bits(g.get_points(k)) == bits(o.mesh(k).points())).all()


def test_fan_with_many_faces_per_vertex_matches_oracle():
    """A hub with 12 faces (more than the patch kernel holds per vertex) keeps the mesh on the CSR
    kernel; an icosphere in the same simulation is on the patch kernel.  Both bit-identical to the
    oracle over 120 steps with wind, drag and lift."""
    k = 12
    ang = np.arange(k) * (2 * np.pi / k)
    fan_pos = np.concatenate([[[0.0, 0.3, 0.0]], np.stack([0.2 * np.cos(ang), 0.3 + 0.01 * np.sin(3 * ang),
                                                          0.2 * np.sin(ang)], 1)]).astype(np.float32)
    fan_idx = np.array([[0, 1 + j, 1 + (j + 1) % k] for j in range(k)], np.uint64).reshape(-1)
    sc = wl.config4(n_meshes=1, seed=7, levels=[3])
    sc.meshes.append(wl.SurfaceSpec(fan_pos, fan_idx, 0.05, 5.0, 0.02, 0.5, 0.002))
    announce(f"fan of {k} faces + level-3 icosphere")
    g = wl.build_surface_gpu(sc)
    o = orc.build_surface_scene(sc)
    g.step(120)
    o.step(120)
    for m in range(len(sc.meshes)):
        a, b = g.get_points(m), o.mesh(m).points()
        assert np.isfinite(a).all()
        assert (bits(a) == bits(b)).all(), m
