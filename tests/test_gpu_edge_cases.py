"""Degenerate and boundary inputs on the GPU, each against the oracle: a simulation with no meshes,
a general mesh with vertices that no triangle uses (patches of one vertex), pinned vertices on the
patch kernel, a single triangle, a mesh of struts without faces, a zero timestep, per-mesh counts of
non-finite vertices, and a rigid-body simulation with no bodies; and records read into a reused
buffer, against a fresh read."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu

FLOOR = (np.array([[-5, 0, 5], [5, 0, 5], [5, 0, -5], [-5, 0, -5]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint64))


def cfg_pair(dt=1.0 / 600.0):
    g = fa.Config(fa.Integration.Rk4, dt, (0.0, -9.81 * 0.01, 0.0), (1.0, 0.0, 0.5), 0.2, 0.5, 0.3, 0.2)
    o = orc.Config(1, dt, (0.0, -9.81 * 0.01, 0.0), (1.0, 0.0, 0.5), 0.2, 0.5, 0.3, 0.2)
    return g, o


def test_simulation_without_meshes_steps():
    sim = fa.Simulation([], [fa.CollidableMesh(*FLOOR)])
    assert sim.step(5) == pytest.approx(0.001, rel=1e-6)


def test_general_mesh_with_unused_vertices_and_pins():
    """Two triangles sharing an edge plus three vertices no triangle uses; two pinned vertices."""
    pos = np.array([[0, 0.3, 0], [0.1, 0.3, 0], [0, 0.3, 0.1], [0.1, 0.31, 0.1], [0.5, 0.5, 0.5], [0.6, 0.2, 0.0],
                    [-0.4, 0.1, 0.3]], np.float32)
    idx = np.array([0, 1, 2, 1, 3, 2], np.uint64)
    gc, oc = cfg_pair()
    gm = fa.SpringyMesh(pos, idx, 0.07, 20.0, 0.1)
    om = orc.Mesh.new(pos, idx, 0.07, 20.0, 0.1)
    for p in (1, 5):
        gm.add_pin(p)
        om.add_pin(p)
    g = fa.Simulation([gm], [fa.CollidableMesh(*FLOOR)])
    g.config = gc
    o = orc.Simulation([om], [orc.Obstacle(*FLOOR)])
    o.set_config(oc)
    g.step(300)
    o.step(300)
    a, b = g.get_points(0), o.mesh(0).points()
    announce(f"unused vertices: lowest y {a[:, 2].min():.4f}")
    assert (bits(a) == bits(b)).all()
    assert (a[[1, 5], 1:4] == pos[[1, 5]]).all()


@pytest.mark.parametrize("kernel", [fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR])
def test_single_triangle(kernel):
    pos = np.array([[0, 0.2, 0], [0.3, 0.25, 0], [0, 0.2, 0.3]], np.float32)
    idx = np.array([0, 1, 2], np.uint64)
    gc, oc = cfg_pair()
    g = fa.Simulation([fa.SpringyMesh(pos, idx, 0.03, 10.0, 0.05)], [fa.CollidableMesh(*FLOOR)])
    g.config = gc
    g.set_kernel(kernel, fa.FP_MATH_EXACT)
    o = orc.Simulation([orc.Mesh.new(pos, idx, 0.03, 10.0, 0.05)], [orc.Obstacle(*FLOOR)])
    o.set_config(oc)
    g.step(200)
    o.step(200)
    assert (bits(g.get_points(0)) == bits(o.mesh(0).points())).all()


def test_rigid_simulation_without_bodies():
    sim = fa.RigidSimulation([], [fa.CollidableMesh(*FLOOR)])
    assert sim.step(3) == np.float32(0.001)
    assert sim.states().shape == (0, 29)


def test_nonfinite_counts_per_mesh():
    """A NaN planted in one vertex of the second cloth: counted there only, and (as in the
    reference, which does not trap it) spreading through that cloth's struts as it steps."""
    from feriphys_amd import workloads as wl

    a = wl.config0(steps=1, seed=1, rows=8, cols=9, pins="edge")
    b = wl.config0(steps=1, seed=2, rows=7, cols=6, pins="edge")
    sc = wl.Scene(a.cloths + b.cloths, a.config, [], 1e-3, 1, 1, "two cloths")
    sim = wl.build_gpu(sc)
    assert list(sim.count_nonfinite()) == [0, 0]
    rec = sim.get_points(1)
    rec[20, 2] = np.nan
    sim.set_points(1, rec)
    assert list(sim.count_nonfinite()) == [0, 1]
    sim.step(3)
    c = sim.count_nonfinite()
    announce(f"non-finite vertices after 3 steps: {list(c)}")
    assert c[0] == 0 and c[1] > 1


@pytest.mark.parametrize("kernel", [fa.FP_KERNEL_AUTO, fa.FP_KERNEL_CSR])
def test_strut_chain_without_faces(kernel):
    """A mesh with no triangles, only add_strut struts: a chain of 40 vertices hanging from a pin
    and falling onto the floor (no face forces, so the patch kernel runs with no face LDS)."""
    n = 40
    pos = np.stack([np.linspace(0.0, 0.39, n), np.full(n, 0.05), np.zeros(n)], 1).astype(np.float32)
    gm = fa.SpringyMesh(pos, np.zeros(0, np.uint64), 0.4, 1.0, 1.0)
    om = orc.Mesh.new(pos, np.zeros(0, np.uint64), 0.4, 1.0, 1.0)
    for i in range(n - 1):
        gm.add_strut((i, i + 1), 5.0, 0.01)
        om.add_strut(i, i + 1, 5.0, 0.01)
    gm.add_pin(0)
    om.add_pin(0)
    gc, oc = cfg_pair()
    g = fa.Simulation([gm], [fa.CollidableMesh(*FLOOR)])
    g.config = gc
    g.set_kernel(kernel, fa.FP_MATH_EXACT)
    assert g.get_mesh_kernel(0) == (fa.FP_STEP_PATCH if kernel == fa.FP_KERNEL_AUTO else fa.FP_STEP_CSR)
    o = orc.Simulation([om], [orc.Obstacle(*FLOOR)])
    o.set_config(oc)
    g.step(300)
    o.step(300)
    a, b = g.get_points(0), o.mesh(0).points()
    assert np.isfinite(a).all() and b[:, 2].min() > -1e-3
    assert (bits(a) == bits(b)).all()


def test_zero_timestep_leaves_the_state():
    """dt = 0 (a valid Duration): every step returns Duration 0 and the records stay as the oracle's
    (x + 0 v, v + 0 a), bit for bit."""
    pos = np.array([[0, 0.3, 0], [0.1, 0.3, 0], [0, 0.3, 0.1]], np.float32)
    idx = np.array([0, 1, 2], np.uint64)
    gc, oc = cfg_pair(dt=0.0)
    g = fa.Simulation([fa.SpringyMesh(pos, idx, 0.03, 20.0, 0.1)], [fa.CollidableMesh(*FLOOR)])
    g.config = gc
    o = orc.Simulation([orc.Mesh.new(pos, idx, 0.03, 20.0, 0.1)], [orc.Obstacle(*FLOOR)])
    o.set_config(oc)
    assert g.step(3) == 0.0
    o.step(3)
    assert (bits(g.get_points(0)) == bits(o.mesh(0).points())).all()


def test_get_points_into_a_reused_buffer():
    """get_points(out=...) fills the caller's buffer (as a render loop reuses its staging buffer, here
    also one in page-locked memory from fp_host_alloc) with the records a fresh get_points returns,
    bit for bit, and refuses a buffer of the wrong shape."""
    announce("get_points into a reused buffer: 24 x 17 cloth, 7 steps")
    cloth = fa.Cloth(24, 17, 0.05, (0.0, 1.0, 0.0), 0.01, 20.0, 0.1, 15.0, 0.1, 5.0, 0.05, [0, 16])
    g = fa.Simulation([cloth.mesh], [fa.CollidableMesh(*FLOOR)])
    g.config = cfg_pair()[0]
    buf = np.full((24 * 17, 10), np.nan, np.float32)
    for _ in range(7):
        g.step(1)
        got = g.get_points(0, out=buf)
        assert got is buf
        assert (bits(buf) == bits(g.get_points(0))).all()
    pinned = fa.pinned_empty((24 * 17, 10))
    assert pinned.shape == (24 * 17, 10) and pinned.dtype == np.float32
    assert (bits(g.get_points(0, out=pinned)) == bits(g.get_points(0))).all()
    with pytest.raises(ValueError):
        g.get_points(0, out=np.zeros((24 * 17, 9), np.float32))
