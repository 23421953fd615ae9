"""GPU parity of the fast-math tiled grid kernel (grid_fast.hip): each strut and face evaluated once
per tile, one quad per thread.  Bars as for every FAST kernel: one-step forces within 1e-5 of the
largest force component and one-step states within 1e-5 relative of the oracle; runs are
deterministic (fixed summation order, no atomics), so repeated runs and graph replays are
bit-identical."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits, force_rel_err, gpu_scene, oracle_scene

pytestmark = pytest.mark.gpu


def orc_cfg(scene):
    c = scene.config
    return orc.Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                      c.coefficient_of_restitution, c.coefficient_of_friction)


def check_one_step(sc, rec, shape=None, label=""):
    g = gpu_scene(sc, fa.FP_KERNEL_GRID, fa.FP_MATH_FAST)
    if shape is not None:
        g.set_tile_shape(0, shape)
    assert g.get_mesh_kernel(0) == fa.FP_STEP_TILED
    g.set_points(0, rec)
    g.accumulate_forces(0)
    fg = g.get_points(0)[:, 7:10]
    o = oracle_scene(sc)
    o.mesh(0).set_points(rec)
    o.mesh(0).accumulate_forces(orc_cfg(sc))
    fr = o.mesh(0).points()[:, 7:10]
    ferr = force_rel_err(fg, fr)
    g.clear_forces(0)
    g.set_points(0, rec)
    g.step(1)
    sg = g.get_points(0)
    o = oracle_scene(sc)
    o.mesh(0).set_points(rec)
    o.step(1)
    sr = o.mesh(0).points()
    dx = float(np.abs(sg[:, 1:4].astype(np.float64) - sr[:, 1:4]).max())
    dv = float(np.abs(sg[:, 4:7].astype(np.float64) - sr[:, 4:7]).max())
    announce(f"fast kernel {label}: force rel err {ferr:.2e}, max |dx| {dx:.2e}, max |dv| {dv:.2e}")
    assert np.isfinite(fg).all() and np.isfinite(sg).all()
    assert ferr <= 1e-5
    assert dx <= 1e-5 * max(1.0, float(np.abs(sr[:, 1:4]).max()))
    assert dv <= 1e-5 * max(float(np.abs(sr[:, 4:7]).max()), 1e-3)
    pins = sc.cloths[0].pins
    assert (bits(sg[pins, :7]) == bits(rec[pins, :7])).all()


@pytest.mark.parametrize("shape", [fa.FP_TILE_16x16, fa.FP_TILE_7x31])
@pytest.mark.parametrize("rows,cols,pins", [(32, 32, "corners"), (5, 3, "edge"), (2, 2, "edge"), (19, 129, "edge"),
                                            (37, 70, "corners"), (31, 31, "edge"), (8, 250, "edge")])
def test_fast_kernel_one_step(rows, cols, pins, shape):
    sc = wl.config0(rows=rows, cols=cols, pins=pins)
    o = oracle_scene(sc)
    o.step(25)
    check_one_step(sc, o.mesh(0).points(), shape, f"{rows}x{cols} shape {shape}")


def test_fast_kernel_one_step_full_size():
    """The 1024 x 1024 benchmark cloth with its sphere: one-step forces and state against the oracle."""
    sc = wl.config1(steps=1)
    announce("fast kernel, full size: building scenes")
    rec = wl.build_gpu(sc).get_points(0)
    check_one_step(sc, rec, None, "1024x1024")


def test_fast_kernel_deterministic_and_graph_replay():
    """Two runs of 200 steps give the same bits; replaying 50-step graphs gives the same bits as
    stepping eagerly."""
    sc = wl.config0(rows=64, cols=96, pins="edge")
    outs = []
    for mode in ("eager", "eager", "graph"):
        g = gpu_scene(sc, fa.FP_KERNEL_GRID, fa.FP_MATH_FAST)
        if mode == "graph":
            g.step_graph(200, 50)
        else:
            g.step(200)
        outs.append(g.get_points(0))
    assert np.isfinite(outs[0]).all()
    assert (bits(outs[0]) == bits(outs[1])).all()
    assert (bits(outs[0]) == bits(outs[2])).all()


def test_fast_kernel_batched_mixed_shapes_track_exact():
    """Forty cloths of uneven sizes with per-cloth configs (tile shapes chosen per size): 100 fast steps
    stay within 1e-4 m of the exact kernel's on every cloth."""
    sc = wl.config2(n_cloths=40, smin=3, smax=60, steps=100)
    ge = wl.build_gpu(sc)
    ge.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT)
    gf = wl.build_gpu(sc)
    gf.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_FAST)
    ge.step(sc.steps)
    gf.step(sc.steps)
    worst = 0.0
    for k in range(len(sc.cloths)):
        xe = ge.get_points(k)[:, 1:4].astype(np.float64)
        xf = gf.get_points(k)[:, 1:4].astype(np.float64)
        if not (np.isfinite(xe).all() and np.abs(xe).max() < 1e3):
            continue  # a corner-pinned cloth that the method itself blows up
        worst = max(worst, float(np.abs(xf - xe).max()))
    announce(f"batched fast vs exact, 100 steps: worst cloth {worst:.2e} m")
    assert worst <= 1e-4
