"""GPU parity on irregular meshes (the patch kernel, and the CSR kernel with its face kernel):
seeded, radially perturbed icospheres with tensile struts on edges and bending struts across
shared edges, dropped onto a ground plane; a fan whose hub has twice an icosphere vertex's faces;
the per-mesh choice between the two kernels."""
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


@pytest.mark.parametrize("kernel", [fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR])
def test_fan_with_many_faces_per_vertex_matches_oracle(kernel):
    """A hub vertex with 12 faces (twice an icosphere vertex's) and an icosphere in one simulation,
    on the patch kernel and on the CSR kernel: bit-identical to the oracle over 120 steps with wind,
    drag and lift."""
    k = 12
    ang = np.arange(k) * (2 * np.pi / k)
    fan_pos = np.concatenate([[[0.0, 0.3, 0.0]], np.stack([0.2 * np.cos(ang), 0.3 + 0.01 * np.sin(3 * ang),
                                                          0.2 * np.sin(ang)], 1)]).astype(np.float32)
    fan_idx = np.array([[0, 1 + j, 1 + (j + 1) % k] for j in range(k)], np.uint64).reshape(-1)
    sc = wl.config4(n_meshes=1, seed=7, levels=[3])
    sc.meshes.append(wl.SurfaceSpec(fan_pos, fan_idx, 0.05, 5.0, 0.02, 0.5, 0.002))
    announce(f"fan of {k} faces + level-3 icosphere, kernel {kernel}")
    g = wl.build_surface_gpu(sc)
    g.set_kernel(kernel, fa.FP_MATH_EXACT)
    expect = fa.FP_STEP_PATCH if kernel == fa.FP_KERNEL_AUTO else fa.FP_STEP_CSR
    assert [g.get_mesh_kernel(m) for m in range(2)] == [expect, expect]
    o = orc.build_surface_scene(sc)
    g.step(120)
    o.step(120)
    for m in range(len(sc.meshes)):
        a, b = g.get_points(m), o.mesh(m).points()
        assert np.isfinite(a).all()
        assert (bits(a) == bits(b)).all(), m


def _orc_config(sc):
    c = sc.config
    return orc.Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                      c.coefficient_of_restitution, c.coefficient_of_friction)


def test_icosphere_graph_replays_match_eager_steps():
    """Captured step chunks on the patch kernel: every replay equals eager stepping, with readbacks
    and synchronisations between the replays."""
    sc = wl.config4(n_meshes=3, seed=11, levels=[3, 2, 4])
    for m in sc.meshes:
        m.positions[:, 1] -= m.positions[:, 1].min() - 0.005
    eager, graph = wl.build_surface_gpu(sc), wl.build_surface_gpu(sc)
    for rnd in range(3):
        graph.step_graph(9, 9)
        eager.step(9)
        same = all((bits(graph.get_points(k)) == bits(eager.get_points(k))).all() for k in range(3))
        announce(f"icosphere graph replay {rnd}: identical={same}")
        assert same


def test_icosphere_records_round_trip_and_update_points_match_oracle():
    """A general mesh is stored in the library's own vertex order: records written and read back
    come out unchanged, and SpringyMesh::update_points (pins kept, ground-plane collision) equals the
    oracle bit for bit in the reference's order."""
    sc = wl.config4(n_meshes=2, seed=13, levels=[3, 2])
    g, o = wl.build_surface_gpu(sc), orc.build_surface_scene(sc)
    rng = np.random.default_rng(5)
    obs = [orc.Obstacle(np.asarray(p, np.float32), i) for p, i in sc.obstacles]
    for k in range(2):
        rec = o.mesh(k).points()
        rec[:, 4:10] = rng.normal(scale=0.5, size=(rec.shape[0], 6)).astype(np.float32)
        g.set_points(k, rec)
        assert (bits(g.get_points(k)) == bits(rec)).all()
        o.mesh(k).set_points(rec)
        new = rec.copy()
        new[:, 1:4] -= np.float32(1.2) * rec[:, 1:4].min(axis=0)[None, :] * np.array([0, 1, 0], np.float32)
        new[:, 1:4] += rng.normal(scale=0.01, size=(rec.shape[0], 3)).astype(np.float32)
        new[:, 4:7] = rng.normal(size=(rec.shape[0], 3)).astype(np.float32)
        g.update_points(k, new)
        o.mesh(k).update_points(new, obs, _orc_config(sc))
        hits = int((bits(o.mesh(k).points()[:, 1:4]) != bits(new[:, 1:4])).any(axis=1).sum())
        announce(f"mesh {k}: {rec.shape[0]} vertices, {hits} crossed the ground plane")
        assert hits > 0
        assert (bits(g.get_points(k)) == bits(o.mesh(k).points())).all()


def test_icosphere_surrogate_slots_features_and_outputs():
    """On a general mesh (stored in the library's own order) the surrogate's slots are the vertex's
    struts in strut order, in the reference's vertex indices; its features and the network's outputs
    match the network oracle's."""
    import mlp_oracle as mo

    sc = wl.config4(n_meshes=2, seed=17, levels=[2, 3])
    sim, o = wl.build_surface_gpu(sc), orc.build_surface_scene(sc)
    net = fa.Surrogate(seed=3)
    sim.set_surrogate(net)
    H = float(np.float32(np.float32(sc.config.dt) * np.float32(net.config.ratio)))
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(9)
    for m in range(2):
        rec = sim.get_points(m)
        rec[:, 4:7] = rng.normal(scale=0.3, size=(rec.shape[0], 3)).astype(np.float32)
        sim.set_points(m, rec)
        st = o.mesh(m).struts()
        n = rec.shape[0]
        exp_nbr = -np.ones((n, 12), np.int64)
        exp_rest = np.zeros((n, 12), np.float32)
        used = np.zeros(n, np.int64)
        for (a, b), L in zip(st["v01"].astype(np.int64), st["kcl"][:, 2]):
            if not L > 0:
                continue
            for s, t in ((a, b), (b, a)):
                if used[s] < 12:
                    exp_nbr[s, used[s]], exp_rest[s, used[s]] = t, L
                    used[s] += 1
        nbr, rest = sim.surrogate_slots(m)
        assert (nbr == exp_nbr).all()
        assert (bits(rest[exp_nbr >= 0]) == bits(exp_rest[exp_nbr >= 0])).all()
        f = sim.surrogate_features(m)
        fo = mo.features(rec[:, 1:4], rec[:, 4:7], nbr, rest, H, net.config.strain_scale)
        np.testing.assert_allclose(f, fo, rtol=2e-5, atol=2e-6)
        yo, _ = mo.forward(p, fo, emulate=True)
        got = sim.surrogate_accel(m)
        err = np.abs(got - yo * net.config.out_scale).max() / np.abs(yo * net.config.out_scale).max()
        announce(f"surrogate on icosphere {m}: {n} vertices, output error {err:.2e} of the largest")
        assert err < 1e-2


def test_kernel_switch_mid_run_matches_oracle():
    """Patch kernel for 60 steps, then the CSR kernel for 60, then back: the block lists and face
    shares are rebuilt on each switch and the trajectory stays the oracle's."""
    sc = wl.config4(n_meshes=3, seed=21, levels=[3, 4, 2])
    for m in sc.meshes:
        m.positions[:, 1] -= m.positions[:, 1].min() - 0.01
    g, o = wl.build_surface_gpu(sc), orc.build_surface_scene(sc)
    for kernel in (fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR, fa.FP_KERNEL_AUTO):
        g.set_kernel(kernel, fa.FP_MATH_EXACT)
        g.step(60)
        o.step(60)
        for k in range(3):
            assert (bits(g.get_points(k)) == bits(o.mesh(k).points())).all(), (kernel, k)


def test_config4_one_step_full_size():
    """The 64-icosphere workload at full size (2,145,408 vertices): one step on the patch kernel
    equals the oracle bit for bit."""
    sc = wl.config4()
    n = sum(m.positions.shape[0] for m in sc.meshes)
    announce(f"config4 full size: {len(sc.meshes)} meshes, {n} vertices: building the library's meshes")
    g = wl.build_surface_gpu(sc)
    announce("building the oracle's meshes (about half a minute)")
    o = orc.build_surface_scene(sc)
    announce("one step on each side")
    g.step(1)
    o.step(1)
    for k in range(len(sc.meshes)):
        assert (bits(g.get_points(k)) == bits(o.mesh(k).points())).all(), k


@pytest.mark.parametrize("n_over", [14, 15])
def test_strut_classes_choose_the_kernel_and_match_oracle(n_over):
    """The patch kernel's 8-byte strut entries rebuild each strut's stiffness and damping as
    (1.0 / len) * nominal from its mesh's table of nominal pairs, at most 16.  An icosphere whose
    edge overrides make 16 pairs (the default, the bending struts' and 14 overrides) runs on it; one
    more pair moves the mesh to the CSR kernel.  Both equal the oracle bit for bit over 120 steps."""
    sc = wl.config4(n_meshes=1, seed=23, levels=[3])
    spec = sc.meshes[0]
    spec.positions[:, 1] -= spec.positions[:, 1].min() - 0.01
    edges, seen = [], set()
    for t in spec.indices.reshape(-1, 3).astype(np.int64):
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            key = (int(min(a, b)), int(max(a, b)))
            if key not in seen:
                seen.add(key)
                edges.append(key)
    ovr = {e: (spec.k * (1.0 + 0.05 * (j + 1)), spec.c * (1.0 + 0.03 * (j + 1))) for j, e in enumerate(edges[:n_over])}
    m = fa.SpringyMesh(spec.positions, spec.indices, spec.mass, spec.k, spec.c, strut_overrides=ovr)
    om = orc.Mesh.new(spec.positions, spec.indices, spec.mass, spec.k, spec.c,
                      overrides=[(a, b, kk, cc) for (a, b), (kk, cc) in ovr.items()])
    v01, _, faces = m.struts()
    _, fv = m.faces()
    for a, b in wl.bend_pairs(v01, faces, fv):
        m.add_strut((int(a), int(b)), spec.bk, spec.bc)
        om.add_strut(int(a), int(b), spec.bk, spec.bc)
    g = fa.Simulation([m], [fa.CollidableMesh(p, i) for p, i in sc.obstacles])
    o = orc.Simulation([om], [orc.Obstacle(np.asarray(p, np.float32), i) for p, i in sc.obstacles])
    c = sc.config
    g.config = fa.Config(fa.Integration(c.integration), c.dt, c.gravity, c.wind, c.lift_coefficient,
                         c.drag_coefficient, c.coefficient_of_restitution, c.coefficient_of_friction)
    o.set_config(_orc_config(sc))
    kernel = g.get_mesh_kernel(0)
    announce(f"icosphere with {n_over} edge overrides: kernel {kernel}")
    assert kernel == (fa.FP_STEP_PATCH if n_over == 14 else fa.FP_STEP_CSR)
    g.step(120)
    o.step(120)
    assert (bits(g.get_points(0)) == bits(o.mesh(0).points())).all()


def test_get_vertices_positions_and_faces():
    """SpringyMesh::get_vertices through the simulation: the positions equal the records' (in the
    reference's vertex order, packed on the device) and the face indices the oracle's; a NULL index
    buffer returns the positions alone."""
    import ctypes as C

    sc = wl.config4(n_meshes=2, seed=29, levels=[3, 2])
    g, o = wl.build_surface_gpu(sc), orc.build_surface_scene(sc)
    g.step(30)
    o.step(30)
    for k in range(2):
        pos, idx = g.get_vertices(k)
        rec = g.get_points(k)
        assert (bits(pos) == bits(rec[:, 1:4])).all()
        assert (bits(pos) == bits(o.mesh(k).points()[:, 1:4])).all()
        assert (idx == o.mesh(k).faces()[1].reshape(-1)).all()
        only = np.zeros_like(pos)
        rc = fa.lib().fp_sim_get_vertices(g.handle, k, only.ctypes.data_as(C.POINTER(C.c_float)), None)
        assert rc == 0 and (bits(only) == bits(pos)).all()


@pytest.mark.parametrize("kernel", [fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR])
def test_icospheres_euler_and_per_vertex_masses_match_oracle(kernel):
    """Explicit Euler (state.rs:75-83) on general meshes, with every vertex given its own mass
    (the records' mass field, so the kernels take the per-vertex mass path): bit-identical to the
    oracle over 120 steps on the ground plane."""
    import dataclasses

    sc = wl.config4(n_meshes=2, seed=31, levels=[3, 2])
    sc.config = dataclasses.replace(sc.config, integration=0)
    for m in sc.meshes:
        m.positions[:, 1] -= m.positions[:, 1].min() - 0.01
    g, o = wl.build_surface_gpu(sc), orc.build_surface_scene(sc)
    g.set_kernel(kernel, fa.FP_MATH_EXACT)
    rng = np.random.default_rng(17)
    for k in range(2):
        rec = o.mesh(k).points()
        rec[:, 0] = (rec[:, 0] * rng.uniform(0.5, 2.0, rec.shape[0])).astype(np.float32)
        o.mesh(k).set_points(rec)
        g.set_points(k, rec)
    assert g.get_mesh_kernel(0) == (fa.FP_STEP_PATCH if kernel == fa.FP_KERNEL_AUTO else fa.FP_STEP_CSR)
    g.step(120)
    o.step(120)
    for k in range(2):
        a, b = g.get_points(k), o.mesh(k).points()
        assert np.isfinite(a).all()
        assert (bits(a) == bits(b)).all(), k


@pytest.mark.parametrize("kernel", [fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR])
def test_icosphere_onto_sphere_obstacle_matches_oracle(kernel):
    """A general mesh dropped onto a curved obstacle (a 320-face icosphere, the broad-phase grids'
    path) beside the ground plane: first-hit faces and responses bit-identical to the oracle over
    240 steps, the mesh draped over the sphere's cap."""
    sc = wl.config4(n_meshes=2, seed=37, levels=[3, 3])
    for m in sc.meshes:
        m.positions[:, 1] -= m.positions[:, 1].min() - 0.02
    c0 = sc.meshes[0].positions.astype(np.float64).mean(axis=0)
    sc.obstacles = list(sc.obstacles) + [wl.sphere_obstacle(0.2, (c0[0] + 0.01, -0.19, c0[2] + 0.005), 2)]
    g, o = wl.build_surface_gpu(sc), orc.build_surface_scene(sc)
    g.set_kernel(kernel, fa.FP_MATH_EXACT)
    g.step(240)
    o.step(240)
    for k in range(2):
        a, b = g.get_points(k), o.mesh(k).points()
        assert np.isfinite(a).all()
        assert (bits(a) == bits(b)).all(), k
    # the first mesh (soft) drapes over the sphere's cap, whose top is ~9.5 mm above the ground;
    # the second lies flat on the ground
    assert o.mesh(0).points()[:, 2].max() > 0.009 > 0.008 > o.mesh(1).points()[:, 2].max()
