"""GPU parity of the batched path: many cloths of uneven sizes, each with its own Config, packed
into one simulation and stepped by the tiled kernel (one launch per tile shape); HIP-graph replays
of captured step chunks; the tile-shape autotuner."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu


def small_batch(n=40, seed=1, steps=60):
    return wl.config2(n_cloths=n, seed=seed, smin=3, smax=40, steps=steps)


def test_batched_cloths_match_per_cloth_oracle():
    sc = small_batch()
    announce(f"batched: {len(sc.cloths)} cloths, {sum(c.rows * c.cols for c in sc.cloths)} vertices")
    g = wl.build_gpu(sc)
    g.step(sc.steps)
    sims = orc.build_scene_per_mesh(sc)
    for k, o in enumerate(sims):
        o.step(sc.steps)
        a, b = g.get_points(k), o.mesh(0).points()
        assert (bits(a) == bits(b)).all(), f"cloth {k} ({sc.cloths[k].rows}x{sc.cloths[k].cols}) differs"


def test_all_tile_shapes_in_one_step_match_oracle():
    """Cloths given the three tile shapes in turn (three launches per step) stay bit-identical to
    the per-cloth oracle, also after the shapes are switched mid-run."""
    sc = small_batch(n=18, seed=9, steps=50)
    g = wl.build_gpu(sc)
    shapes = [fa.FP_TILE_15x15, fa.FP_TILE_7x31, fa.FP_TILE_16x16]
    for k in range(len(sc.cloths)):
        g.set_tile_shape(k, shapes[k % 3])
    assert all(g.get_mesh_kernel(k) == fa.FP_STEP_TILED for k in range(len(sc.cloths)))
    sims = orc.build_scene_per_mesh(sc)
    g.step(25)
    for k in range(len(sc.cloths)):
        g.set_tile_shape(k, shapes[(k + 1) % 3])
    g.step(sc.steps - 25)
    for k, o in enumerate(sims):
        o.step(sc.steps)
        assert (bits(g.get_points(k)) == bits(o.mesh(0).points())).all(), k


def test_batched_mixed_kernels_match_oracle():
    """Cloths extended with add_strut leave the grid kernel for the CSR kernel; both kinds step in
    the same simulation and stay bit-identical to the oracle."""
    sc = small_batch(n=12, seed=4, steps=40)
    meshes = []
    omeshes = []
    for k, c in enumerate(sc.cloths):
        cl = fa.Cloth(*c.args(), plane=c.plane)
        oc = orc.Mesh.cloth(*c.args(), plane=c.plane)
        rec = wl.jitter_records(cl.mesh.get_points(), c.pins, sc.jitter, sc.seed + k)
        cl.mesh.set_points(rec)
        oc.set_points(rec)
        if k % 3 == 0:
            cl.mesh.add_strut((0, c.rows * c.cols - 1), 3.0, 0.5)
            oc.add_strut(0, c.rows * c.cols - 1, 3.0, 0.5)
        meshes.append(cl.mesh)
        omeshes.append(oc)
    g = fa.Simulation(meshes, [])
    cfgs = wl.mesh_configs(sc)
    to = lambda c: fa.Config(fa.Integration(c.integration), c.dt, c.gravity, c.wind, c.lift_coefficient,
                             c.drag_coefficient, c.coefficient_of_restitution, c.coefficient_of_friction)
    g.set_mesh_configs([to(c) for c in cfgs])
    g.step(sc.steps)
    for k, (m, c) in enumerate(zip(omeshes, cfgs)):
        o = orc.Simulation([m], [])
        o.set_config(orc.Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                                c.coefficient_of_restitution, c.coefficient_of_friction))
        o.step(sc.steps)
        assert (bits(g.get_points(k)) == bits(o.mesh(0).points())).all(), k


@pytest.mark.parametrize("chunk", [20, 7])
def test_graph_replays_match_eager_steps(chunk):
    """Every replay of the captured chunk is checked against eager stepping, with the program's own
    eager work (a readback and a synchronisation) between replays."""
    sc = small_batch(n=30, seed=2)
    eager = wl.build_gpu(sc)
    graph = wl.build_gpu(sc)
    for rnd in range(4):
        graph.step_graph(chunk, chunk)
        eager.step(chunk)
        got = [graph.get_points(k) for k in range(len(sc.cloths))]
        ref = [eager.get_points(k) for k in range(len(sc.cloths))]
        same = all((bits(a) == bits(b)).all() for a, b in zip(got, ref))
        announce(f"graph chunk {chunk}, replay {rnd}: identical={same}")
        assert same
    # a config change invalidates the captured graph; replays then follow the new config
    cfgs = wl.mesh_configs(sc)
    to = lambda c, w: fa.Config(fa.Integration(c.integration), c.dt, c.gravity, w, c.lift_coefficient,
                                c.drag_coefficient, c.coefficient_of_restitution, c.coefficient_of_friction)
    new = [to(c, (2.0, 0.5, -1.0)) for c in cfgs]
    graph.set_mesh_configs(new)
    eager.set_mesh_configs(new)
    graph.step_graph(2 * chunk + 3, chunk)
    eager.step(2 * chunk + 3)
    for k in range(len(sc.cloths)):
        assert (bits(graph.get_points(k)) == bits(eager.get_points(k))).all()


def test_autotune_keeps_results_and_state():
    sc = small_batch(n=30, seed=3)
    a = wl.build_gpu(sc)
    b = wl.build_gpu(sc)
    before = [b.get_points(k) for k in range(len(sc.cloths))]
    chosen = b.autotune(5)
    announce(f"autotune chose tile shapes per size class: {chosen}")
    after = [b.get_points(k) for k in range(len(sc.cloths))]
    assert all((bits(x) == bits(y)).all() for x, y in zip(before, after))  # state restored
    a.step(30)
    b.step(30)
    for k in range(len(sc.cloths)):
        assert (bits(a.get_points(k)) == bits(b.get_points(k))).all()
