"""GPU parity on irregular meshes (the CSR kernel with its face kernel): seeded, radially perturbed
icospheres with tensile struts on edges and bending struts across shared edges, dropped onto a
ground plane."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("math", [fa.FP_MATH_EXACT, fa.FP_MATH_FAST])
def test_icospheres_on_ground_match_oracle(math):
    sc = wl.config4(n_meshes=5, seed=3, levels=[2, 3, 3, 4, 1])
    for m in sc.meshes:  # bring them close to the ground so they land within the run
        m.positions[:, 1] -= m.positions[:, 1].min() - 0.01
    n = sum(m.positions.shape[0] for m in sc.meshes)
    announce(f"icospheres: {len(sc.meshes)} meshes, {n} vertices, math {math}")
    g = wl.build_surface_gpu(sc)
    g.set_kernel(fa.FP_KERNEL_AUTO, math)
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
