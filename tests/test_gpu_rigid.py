"""GPU parity of the rigid-body step (rigidbody/simulation.rs:26-52) with the CPU oracle: unit
cubes with random impulses tumbling onto a ground plane and an icosphere (the broad-phase grid
path), RK4 and Euler, with gravity and torque; every 29-float record bit for bit.

The reference's response (one impulse per penetrating vertex, each from the pre-step velocity)
adds energy in resting contact when the restitution is above 0: in the oracle itself a cube of
the 96-cube pile below reaches NaN at step 371 at restitution 0.6 (after which the sign of a NaN,
which x86 and gfx950 produce differently, steers the collision tests).  The long runs therefore
use restitution 0; the run at 0.6 stops at step 360, with momenta already growing."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu

FLOOR = (np.array([[-60, 0, 60], [60, 0, 60], [60, 0, -60], [-60, 0, -60]], np.float32),
         np.array([0, 1, 2, 0, 2, 3], np.uint64))


def scene(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.concatenate([rng.uniform(-6, 6, (n, 1)), rng.uniform(1.0, 6.0, (n, 1)), rng.uniform(-6, 6, (n, 1))], 1)
    mass = rng.uniform(0.5, 4.0, n)
    imp = rng.normal(scale=2.0, size=(n, 3))
    at = rng.uniform(-0.5, 0.5, size=(n, 3))
    return pos.astype(np.float32), mass.astype(np.float32), imp.astype(np.float32), at.astype(np.float32)


def obstacles():
    return [FLOOR, wl.sphere_obstacle(2.0, (0.0, 1.0, 0.0), 2)]


def build_both(n, seed, integration, gravity, torque, restitution=0.0):
    pos, mass, imp, at = scene(n, seed)
    obs = obstacles()
    bodies = [fa.RigidBody(pos[i], float(mass[i])) for i in range(n)]
    for i in range(n // 2):
        bodies[i].apply_impulse(imp[i], at[i])
    g = fa.RigidSimulation(bodies, [fa.CollidableMesh(p, i) for p, i in obs])
    for i in range(n // 2, n):
        g.apply_impulse(i, imp[i], at[i])
    cfg = fa.RigidConfig(fa.Integration(integration), 1.0 / 600.0, restitution, gravity, torque)
    g.config = cfg
    o = orc.RigidSimulation(pos, mass, [orc.Obstacle(p, i) for p, i in obs])
    for i in range(n):
        o.apply_impulse(i, imp[i], at[i])
    o.set_config(integration=integration, dt=np.float32(1.0 / 600.0), restitution=restitution, gravity=gravity,
                 torque=torque)
    return g, o


def same_bits(a, b):
    return ((bits(a) == bits(b)) | (np.isnan(a) & np.isnan(b))).all()


@pytest.mark.parametrize("integration,steps,restitution", [(1, (100, 300, 600), 0.0), (0, (50, 200), 0.0),
                                                           (1, (200, 360), 0.6)])
def test_tumbling_cubes_bit_exact(integration, steps, restitution):
    g, o = build_both(96, 3 + integration, integration, (0.0, -9.81, 0.0), (0.0, 0.0, 0.0), restitution)
    assert (bits(g.states()) == bits(o.states())).all()
    done = 0
    for s in steps:
        g.step(s - done)
        o.step(s - done)
        done = s
        a, b = g.states(), o.states()
        moved = int((np.abs(a[:, 8]) > 0).sum())
        announce(f"rigid {'RK4' if integration else 'Euler'} step {s}: {moved} bodies moving vertically, "
                 f"lowest centre {a[:, 1].min():.3f}")
        assert same_bits(a, b)
    if restitution == 0.0:
        assert np.isfinite(a).all()


def test_torque_and_checkpoint_restore():
    g, o = build_both(32, 11, 1, (0.0, -3.0, 0.0), (0.2, 0.0, -0.1))
    g.step(120)
    o.step(120)
    snap = g.states()
    assert (bits(snap) == bits(o.states())).all()
    g.step(60)
    g.set_states(snap)
    assert (bits(g.states()) == bits(snap)).all()
    g.step(60)
    o.step(60)
    assert (bits(g.states()) == bits(o.states())).all()


def test_step_returns_the_duration_of_dt():
    g = fa.RigidSimulation([fa.RigidBody([0, 1, 0], 1.0)])
    assert g.step(1) == np.float32(0.001)  # Duration::from_millis(1) round trip
