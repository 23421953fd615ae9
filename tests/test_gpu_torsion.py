"""GPU parity of the torsional hinge springs (SpringyMesh::apply_torsional_forces,
springy_mesh.rs:553-627): the reference's own hinge case, perturbed hinges, and steps with the
torsional forces restored in accumulate_forces (the reference comments the call out at :518-519),
bit for bit against the oracle."""
import math

import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu
F32 = np.float32


def hinge_case():
    r2 = float(F32(math.sqrt(2.0)) / F32(2.0))
    verts = np.array([[0, 0, 0], [0, 0, -1], [1, 0, 0], [r2, r2, 0]], F32)
    return verts, [1, 2, 0, 3, 1, 0]


def both(verts, idx, mass, k, c, torsion):
    g = fa.Simulation([fa.SpringyMesh(verts, idx, mass, k, c, torsion)], [])
    o = orc.Mesh.new(verts, idx, mass, k, c, torsion=torsion)
    return g, o


def test_reference_hinge_case_at_rest_gives_zero():
    """springy_mesh.rs:887-936: at the rest angle no torsional force."""
    verts, idx = hinge_case()
    g, o = both(verts, idx, 2.0, 1.0, 1.0, (1.0, 1.0))
    g.apply_torsional_forces(0)
    o.apply_torsional_forces()
    got = g.get_points(0)
    assert (bits(got) == bits(o.points())).all()
    assert (got[:, 7:10] == 0.0).all()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_perturbed_hinge_forces_match_oracle(seed):
    verts, idx = hinge_case()
    g, o = both(verts, idx, 2.0, 1.0, 1.0, (1.0, 1.0))
    rng = np.random.default_rng(seed)
    rec = o.points()
    rec[:, 1:4] += rng.uniform(-0.2, 0.2, size=(4, 3)).astype(F32)
    rec[:, 4:7] = rng.normal(scale=0.5, size=(4, 3)).astype(F32)
    o.set_points(rec)
    g.set_points(0, rec)
    g.apply_torsional_forces(0)
    o.apply_torsional_forces()
    got, ref = g.get_points(0), o.points()
    announce(f"hinge seed {seed}: forces {got[:, 7:10].tolist()}")
    assert np.abs(ref[:, 7:10]).max() > 1e-3  # the hinge is away from rest
    assert (bits(got) == bits(ref)).all()
    # Newton 3 for the hinge: the four forces cancel (up to rounding)
    assert np.abs(got[:, 7:10].astype(np.float64).sum(axis=0)).max() < 1e-5


def sphere_mesh(level, seed):
    v, f = wl.icosphere_fast(level)
    rng = np.random.default_rng(seed)
    v = v * (1.0 + rng.uniform(-0.02, 0.02, size=(v.shape[0], 1)))
    return v.astype(F32), f.reshape(-1).astype(np.int64)


@pytest.mark.parametrize("math_", [fa.FP_MATH_EXACT])
def test_steps_with_torsion_match_oracle(math_):
    v, idx = sphere_mesh(2, 3)
    torsion = (0.02, 0.001)
    cfg = wl.ConfigSpec(1, 1.0 / 600.0, (0.0, -0.0049, 0.0), (1.0, 0.0, 0.5), 0.2, 0.5, 0.95, 0.3).f32()
    ccfg = fa.Config(fa.Integration(cfg.integration), cfg.dt, cfg.gravity, cfg.wind, cfg.lift_coefficient,
                     cfg.drag_coefficient, cfg.coefficient_of_restitution, cfg.coefficient_of_friction)
    ocfg = orc.Config(cfg.integration, cfg.dt, cfg.gravity, cfg.wind, cfg.lift_coefficient, cfg.drag_coefficient,
                      cfg.coefficient_of_restitution, cfg.coefficient_of_friction)
    rng = np.random.default_rng(5)
    meshes = []
    for on in (True, False):
        g = fa.Simulation([fa.SpringyMesh(v, idx, 0.8, 30.0, 0.2, torsion)], [])
        g.config = ccfg
        g.set_kernel(fa.FP_KERNEL_AUTO, math_)
        g.set_torsion(on)
        om = orc.Mesh.new(v, idx, 0.8, 30.0, 0.2, torsion=torsion)
        om.set_torsion_in_step(on)
        rec = om.points()
        if not meshes:
            dv = rng.normal(scale=0.05, size=(rec.shape[0], 3)).astype(F32)
            dx = rng.normal(scale=0.01, size=(rec.shape[0], 3)).astype(F32)
        rec[:, 4:7] = dv
        rec[:, 1:4] = (rec[:, 1:4] + dx).astype(F32)
        om.set_points(rec)
        g.set_points(0, rec)
        o = orc.Simulation([om], [])
        o.set_config(ocfg)
        meshes.append((g, o))
    for (g, o), on in zip(meshes, (True, False)):
        g.step(20)
        o.step(20)
        a, b = g.get_points(0), o.mesh(0).points()
        announce(f"torsion {'on' if on else 'off'}: 20 steps, max |x| {np.abs(a[:, 1:4]).max():.3f}")
        assert np.isfinite(a).all()
        assert (bits(a) == bits(b)).all()
    on_x, off_x = meshes[0][0].get_points(0)[:, 1:4], meshes[1][0].get_points(0)[:, 1:4]
    assert np.abs(on_x - off_x).max() > 1e-7  # the torsional forces act


def test_accumulate_forces_with_torsion_matches_oracle():
    v, idx = sphere_mesh(1, 7)
    torsion = (0.05, 0.01)
    g = fa.Simulation([fa.SpringyMesh(v, idx, 0.5, 20.0, 0.1, torsion)], [])
    om = orc.Mesh.new(v, idx, 0.5, 20.0, 0.1, torsion=torsion)
    rng = np.random.default_rng(9)
    rec = om.points()
    rec[:, 1:4] = (rec[:, 1:4] + rng.normal(scale=0.02, size=(rec.shape[0], 3))).astype(F32)
    rec[:, 4:7] = rng.normal(scale=0.1, size=(rec.shape[0], 3)).astype(F32)
    om.set_points(rec)
    g.set_points(0, rec)
    c = g.config
    ocfg = orc.Config(int(c.integration), c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                      c.coefficient_of_restitution, c.coefficient_of_friction)
    g.set_torsion(True)
    om.set_torsion_in_step(True)
    g.accumulate_forces(0)
    om.accumulate_forces(ocfg)
    assert (bits(g.get_points(0)) == bits(om.points())).all()
