"""The rigid-body oracle (oracle/rigid_oracle.hpp), pinned where the reference has no test: the
constructor's record, closed forms of free flight under a constant force and of a torque-free
spin, the flat landing on a floor (whose four simultaneous vertex impulses add up to the
restitution law), and f32 against f64 runs."""
import numpy as np
import pytest

import pyoracle as orc

FLOOR = (np.array([[-50, 0, 50], [50, 0, 50], [50, 0, -50], [-50, 0, -50]], np.float32),
         np.array([0, 1, 2, 0, 2, 3], np.uint64))


def test_constructor_record():
    s = orc.RigidSimulation([[1.0, 2.0, 3.0]], [3.0])
    r = s.states()[0]
    assert list(r[0:3]) == [1.0, 2.0, 3.0]
    assert list(r[3:7]) == [0.0, 0.0, 0.0, 1.0]          # Quaternion::one(): v = 0, s = 1
    assert not np.any(r[7:13]) and not np.signbit(r[7:13]).any()
    assert r[13] == np.float32(3.0)
    inv = r[14:23].reshape(3, 3)                            # columns
    assert not np.any(inv - np.diag(np.diag(inv)))
    assert np.allclose(np.diag(inv), 6.0 / 3.0, rtol=2e-7)  # (mass / 6)^-1
    assert not np.any(r[23:29])


@pytest.mark.parametrize("integration", [0, 1])
def test_free_flight_under_constant_force(integration):
    m, g, h, n = 2.0, np.array([0.3, -9.81, 0.0]), 1.0 / 600.0, 600
    s = orc.RigidSimulation([[0.0, 10.0, 0.0]], [m], dtype=np.float64)
    s.set_config(integration=integration, dt=h, gravity=g)
    s.step(n)
    r = s.states()[0]
    t = n * np.float64(np.float32(h)) if False else n * h
    np.testing.assert_allclose(r[7:10], g * t, rtol=1e-12, atol=1e-12)  # p = F t
    x = np.array([0.0, 10.0, 0.0]) + g / m * t * t / 2
    if integration == 1:  # RK4 integrates the quadratic exactly
        np.testing.assert_allclose(r[0:3], x, rtol=1e-12, atol=1e-11)
    else:                 # Euler: x_n = x0 + h^2 F/m n(n-1)/2
        np.testing.assert_allclose(r[0:3], np.array([0.0, 10.0, 0.0]) + g / m * h * h * n * (n - 1) / 2, atol=1e-11)
    assert list(r[3:7]) == [0.0, 0.0, 0.0, 1.0]


def test_torque_free_spin_about_a_principal_axis():
    m, h, n, w = 6.0, 1.0 / 600.0, 900, 3.0
    s = orc.RigidSimulation([[0.0, 0.0, 0.0]], [m])
    st = s.states()
    st[0, 12] = np.float32(w * m / 6.0)  # L_z = I w, I = m / 6
    s.set_states(st)
    s.set_config(dt=h)
    s.step(n)
    r = s.states()[0]
    th = w * n * np.float32(h)
    np.testing.assert_allclose(r[3:7], [0.0, 0.0, np.sin(th / 2), np.cos(th / 2)], atol=2e-5)
    assert r[12] == st[0, 12] and r[10] == 0 and r[11] == 0  # angular momentum untouched
    assert abs(np.linalg.norm(r[3:7].astype(np.float64)) - 1.0) < 1e-6


def test_flat_landing_follows_restitution():
    e, m, h = 0.6, 1.0, 1.0 / 600.0
    s = orc.RigidSimulation([[0.0, 0.52, 0.0]], [m], [orc.Obstacle(*FLOOR)])
    st = s.states()
    st[0, 8] = np.float32(-2.0 * m)  # falling at 2 m/s
    s.set_states(st)
    s.set_config(dt=h, restitution=e)
    vy, y = [], []
    for _ in range(40):
        s.step(1)
        vy.append(float(s.states()[0, 8]) / m)
        y.append(float(s.states()[0, 1]))
    k = next(i for i, v in enumerate(vy) if v > 0)
    assert vy[k] == pytest.approx(e * 2.0, rel=1e-6)             # the 4 corner impulses sum to -(1+e) m v_n
    assert y[k] == pytest.approx(0.5 + 2e-4, abs=1e-6)           # contact point + half the cube + 2 EPSILON
    r = s.states()[0]
    assert np.abs(r[10:13]).max() < 1e-6                          # symmetric corners: no spin
    assert all(v > 0 for v in vy[k:])


def test_tumbling_cube_f32_tracks_f64():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-3, 3, size=(16, 3)) + np.array([0, 5, 0])
    mass = rng.uniform(0.5, 4.0, size=16)
    imp, at = rng.normal(size=(16, 3)), rng.normal(scale=0.3, size=(16, 3))
    runs = []
    for dt in (np.float32, np.float64):
        s = orc.RigidSimulation(pos, mass, [orc.Obstacle(*FLOOR, dtype=dt)], dtype=dt)
        s.set_config(dt=1.0 / 600.0, restitution=0.5, gravity=(0.0, -9.81, 0.0), torque=(0.0, 0.0, 0.0))
        for b in range(16):
            s.apply_impulse(b, imp[b], at[b])
        s.step(300)
        runs.append(s.states().astype(np.float64))
    assert np.abs(runs[0][:, 0:3] - runs[1][:, 0:3]).max() < 1e-3
    assert np.abs(np.linalg.norm(runs[0][:, 3:7], axis=1) - 1).max() < 1e-6
