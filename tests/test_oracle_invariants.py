"""Pins the oracle where the reference's own tests do not: closed forms, invariants, textbook
collision response, the Duration round trip.  Each check is computed independently (in Python
double precision from the mathematics), so a dropped term, a wrong sign or a swapped operand in
the restatement fails one of them."""
import math
import struct

import numpy as np
import pytest

import pyoracle as orc
from feriphys_amd import workloads as wl

f32, f64 = np.float32, np.float64


def cfg(**kw):
    c = orc.Config(integration=1, dt=0.01, gravity=(0.0, 0.0, 0.0), wind=(0.0, 0.0, 0.0), lift_coefficient=0.0,
                   drag_coefficient=0.0, coefficient_of_restitution=0.5, coefficient_of_friction=0.25)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def lone_points(pos, mass, dtype):
    """A SpringyMesh with points and no triangles (so no struts or faces)."""
    return orc.Mesh.new(np.asarray(pos, dtype), [], mass, 0.0, 0.0, dtype=dtype)


@pytest.mark.parametrize("integration", [0, 1])
def test_gravity_only_closed_form(integration):
    """No struts, no faces: the force is the gravity force g, a = g/m.  Frozen-force RK4 gives
    x_n = x0 + n h v0 + (n h)^2 a / 2 exactly; Euler gives x0 + n h v0 + h^2 a n(n-1)/2."""
    h, n, m_tot = 0.01, 50, 3.0
    g = (0.5, -9.0, 0.25)
    x0 = np.array([[0.0, 1.0, 2.0], [1.0, -1.0, 0.5], [-2.0, 0.0, 1.0]])
    v0 = np.array([[0.1, 0.2, -0.3], [0.0, 0.0, 0.0], [1.0, -1.0, 2.0]])
    for dtype, tol in ((f64, 1e-12), (f32, 2e-5)):
        mesh = lone_points(x0, m_tot, dtype)
        rec = mesh.points()
        rec[:, 4:7] = v0
        mesh.set_points(rec)
        sim = orc.Simulation([mesh], [], dtype=dtype)
        sim.set_config(cfg(integration=integration, dt=h, gravity=g))
        sim.step(n)
        p = sim.mesh(0).points().astype(f64)
        hh = float(dtype(h))
        a = np.array([float(dtype(x)) for x in g]) / float(dtype(m_tot) / dtype(3))
        t = n * hh
        if integration == 1:
            xe = x0 + t * v0 + 0.5 * t * t * a
        else:
            xe = x0 + t * v0 + hh * hh * a * n * (n - 1) / 2
        ve = v0 + t * a
        assert np.abs(p[:, 1:4] - xe).max() <= tol * max(1.0, np.abs(xe).max())
        assert np.abs(p[:, 4:7] - ve).max() <= tol * max(1.0, np.abs(ve).max())


def test_single_strut_closed_form():
    """One strut: spring k/L (|d| - L) u plus damping c/L ((v1 - v0) . u) u on vertex 0, the
    negative on vertex 1 (springy_mesh.rs:533-551)."""
    L = 0.5
    k_nom, c_nom = 3.0, 0.7
    for dtype in (f64, f32):
        mesh = lone_points([[0, 0, 0], [L, 0, 0], [5, 5, 5]], 3.0, dtype)
        mesh.add_strut(0, 1, k_nom, c_nom)
        rec = mesh.points()
        rec[1, 1:4] = [0.6, 0.1, -0.2]       # stretched and rotated
        rec[0, 4:7] = [0.3, -0.1, 0.2]
        rec[1, 4:7] = [-0.2, 0.4, 0.1]
        mesh.set_points(rec)
        mesh.accumulate_forces(cfg())
        F = mesh.points()[:, 7:10].astype(f64)
        d = np.array([0.6, 0.1, -0.2])
        ln = np.linalg.norm(d)
        u = d / ln
        k, c = k_nom / L, c_nom / L
        fs = k * (ln - L) * u
        fd = c * np.dot(np.array([-0.2, 0.4, 0.1]) - np.array([0.3, -0.1, 0.2]), u) * u
        tol = 1e-12 if dtype is f64 else 2e-6
        assert np.abs(F[0] - (fs + fd)).max() <= tol
        assert np.abs(F[1] + (fs + fd)).max() <= tol
        assert (F[2] == 0).all()


def test_face_force_closed_form():
    """One right isosceles triangle at rest in the xz plane with wind w: relative velocity -w,
    effective area A |n . (-w)|, drag -Cd A_eff (-w), lift -Cl A_eff (-w) |n x (-w)| (parallel to the
    relative velocity, as the reference computes it), shared over the corners by angle / pi:
    1/2 at the right angle, 1/4 at the others (springy_mesh.rs:629-654)."""
    s = 0.4
    pos = [[0, 0, 0], [s, 0, 0], [0, 0, s]]
    Cd, Cl = 0.7, 0.3
    for wind in ([0.0, -2.0, 0.0], [1.0, -0.5, 2.0], [3.0, 0.0, 1.0]):
        mesh = orc.Mesh.new(np.array(pos, f64), [0, 1, 2], 3.0, 0.0, 0.0, dtype=f64)
        mesh.accumulate_forces(cfg(wind=tuple(wind), drag_coefficient=Cd, lift_coefficient=Cl))
        F = mesh.points()[:, 7:10]
        rv = -np.array(wind)
        n = np.array([0.0, -1.0, 0.0])  # (p1 - p0) x (p2 - p0) = x x z = -y
        A = s * s / 2
        eff = A * abs(np.dot(n, rv))
        tot = -Cd * eff * rv - Cl * eff * rv * np.linalg.norm(np.cross(n, rv))
        assert np.abs(F[0] - 0.5 * tot).max() <= 1e-12
        assert np.abs(F[1] - 0.25 * tot).max() <= 1e-12
        assert np.abs(F[2] - 0.25 * tot).max() <= 1e-12


def test_undeformed_cloth_has_exactly_zero_force():
    """At rest, undeformed, with no wind and no gravity: every strut's |d| - L is exactly 0 (rest
    lengths come from the same subtraction) and every face's relative velocity is 0."""
    c = wl.config0(rows=12, cols=9).cloths[0]
    mesh = orc.Mesh.cloth(*c.args(), plane=c.plane)
    mesh.accumulate_forces(cfg(drag_coefficient=0.5, lift_coefficient=0.2))
    assert (mesh.points()[:, 7:10] == 0).all()


def test_strut_forces_obey_newton_third_law():
    sc = wl.config0(rows=10, cols=10)
    mesh = orc.Mesh.cloth(*sc.cloths[0].args(), plane=1, dtype=f64)
    rec = wl.jitter_records(mesh.points().astype(f32), [], 1e-2, 5).astype(f64)
    rec[:, 4:7] = np.random.default_rng(1).normal(size=(100, 3))
    mesh.set_points(rec)
    mesh.apply_strut_forces()
    F = mesh.points()[:, 7:10]
    assert np.abs(F.sum(axis=0)).max() <= 1e-12 * np.abs(F).sum()
    assert np.abs(F).max() > 1e-3


def test_zero_relative_velocity_gives_no_face_force():
    sc = wl.config0(rows=8, cols=8)
    w = (0.3, -0.2, 0.5)
    a = orc.Mesh.cloth(*sc.cloths[0].args(), plane=1)
    rec = wl.jitter_records(a.points(), [], 1e-3, 2)
    rec[:, 4:7] = w
    a.set_points(rec)
    b = orc.Mesh.cloth(*sc.cloths[0].args(), plane=1)
    b.set_points(rec)
    a.accumulate_forces(cfg(wind=w, drag_coefficient=0.5, lift_coefficient=0.2))
    b.accumulate_forces(cfg(wind=w))
    assert (a.points()[:, 7:10] == b.points()[:, 7:10]).all()


def test_pinned_points_keep_their_record():
    sc = wl.config0(steps=200, pins="corners")
    o = orc.build_scene(sc)
    before = o.mesh(0).points()
    o.step(200)
    after = o.mesh(0).points()
    pins = sc.cloths[0].pins
    assert (after[pins, :7] == before[pins, :7]).all()
    assert not (after[:, 1:4] == before[:, 1:4]).all()


def test_ccd_floor_response_textbook():
    """A point crossing the demo floor (y = -2, normal +y) between two states: placed on the plane
    plus 1e-4 along the normal, normal velocity reversed and scaled by the restitution, tangential
    velocity reduced by min(mu |v_n|, |v_t|) (springy_mesh.rs:444-480)."""
    dt, e, mu, m_tot = 0.01, 0.5, 0.25, 3.0
    floor = orc.Obstacle(*wl.floor_obstacle(), dtype=f64)
    mesh = lone_points([[0.1, -1.995, 0.2], [5, 5, 5], [6, 6, 6]], m_tot, f64)
    old = mesh.points()
    v_old = np.array([0.3, -1.0, 0.2])
    F = np.array([0.0, -3.0, 0.5])
    old[0, 4:7] = v_old
    old[0, 7:10] = F
    mesh.set_points(old)
    new = old.copy()
    new[0, 1:4] = old[0, 1:4] + dt * v_old + np.array([0.0, -0.01, 0.0])
    new[0, 4:7] = [0.31, -1.2, 0.21]
    mesh.update_points(new, [floor], cfg(dt=dt, coefficient_of_restitution=e, coefficient_of_friction=mu))
    p = mesh.points()[0]
    m = m_tot / 3
    d_old = old[0, 2] + 2.0
    d_new = new[0, 2] + 2.0
    frac = d_old / (d_old - d_new)
    cp = old[0, 1:4] + dt * frac * v_old
    vc = v_old + dt * frac * F / m
    n = np.array([0.0, 1.0, 0.0])
    vn = np.dot(vc, n) * n
    vt = vc - vn
    vt_out = vt - vt / np.linalg.norm(vt) * min(mu * np.linalg.norm(vn), np.linalg.norm(vt))
    assert np.abs(p[1:4] - (cp + 1e-4 * n)).max() <= 1e-9
    assert np.abs(p[4:7] - (-e * vn + vt_out)).max() <= 1e-9
    assert abs(p[2] - (-2.0 + 1e-4)) < 1e-12 + abs(cp[1] + 2.0)


def test_ccd_first_face_in_list_order_wins():
    """Two floors crossed by one segment: the first face in obstacle-then-face order is taken even
    though the segment meets the other plane first (collidable_mesh.rs:113, `find`), which shows
    in the height the point is placed at."""
    lo = orc.Obstacle(*wl.floor_obstacle(y=-1.0, half=2.0))
    hi = orc.Obstacle(*wl.floor_obstacle(y=-0.5, half=2.0))
    for obstacles, plane in (([lo, hi], -1.0), ([hi, lo], -0.5)):
        mesh = lone_points([[0.1, 0.0, 0.1], [5, 5, 5], [6, 6, 6]], 3.0, f32)
        old = mesh.points()
        old[0, 4:7] = [0.0, -150.0, 0.0]
        mesh.set_points(old)
        new = old.copy()
        new[0, 1:4] = [0.1, -1.5, 0.1]
        mesh.update_points(new, obstacles, cfg(dt=0.01))
        assert abs(mesh.points()[0, 2] - (plane + 1e-4)) < 1e-5
    old = np.array([[0.1, 0.0, 0.1]], f32)
    dt_d = orc.duration_roundtrip(np.float32(0.01))
    # no crossing, no hit; a crossing outside the triangles, no hit
    assert orc.collided_faces([lo], old, np.array([[0.1, -0.2, 0.1]], f32), dt_d)[0] == -1
    assert orc.collided_faces([lo], np.array([[9.0, 0.0, 9.0]], f32), np.array([[9.0, -1.5, 9.0]], f32), dt_d)[0] == -1
    # the diagonal x + z = 0 splits the square: face 0 holds x + z > 0, face 1 x + z < 0
    assert orc.collided_faces([lo], old, np.array([[0.1, -1.5, 0.1]], f32), dt_d)[0] == 0
    assert orc.collided_faces([lo], -old * [1, 0, 1], np.array([[-0.1, -1.5, -0.1]], f32), dt_d)[0] == 1


def _duration_rt_py(dt32):
    """Duration::from_secs_f32 (exact, round half to even in ns) then as_secs_f32, in integers."""
    from fractions import Fraction
    ns = Fraction(float(dt32)) * 1_000_000_000
    q, r = divmod(ns.numerator, ns.denominator)
    if 2 * r > ns.denominator or (2 * r == ns.denominator and q % 2 == 1):
        q += 1
    secs, nanos = divmod(q, 1_000_000_000)
    return np.float32(np.float32(secs) + np.float32(nanos) / np.float32(1e9))


def test_duration_roundtrip():
    h = np.float32(1.0 / 600.0)
    got = orc.duration_roundtrip(h)
    assert struct.unpack("<I", struct.pack("<f", got))[0] == 0x3ADA7411
    assert orc.duration_roundtrip(np.float32(0.001)) == np.float32(0.001)
    rng = np.random.default_rng(0)
    for dt in rng.uniform(1e-5, 3.0, 200).astype(f32):
        assert orc.duration_roundtrip(dt) == _duration_rt_py(dt)


def test_f32_oracle_tracks_f64_over_600_steps():
    sc = wl.config0(pins="edge")
    a = orc.build_scene(sc, f32)
    b = orc.build_scene(sc, f64)
    a.step(600)
    b.step(600)
    e = np.abs(a.mesh(0).points()[:, 1:4].astype(f64) - b.mesh(0).points()[:, 1:4]).max()
    assert e < 1e-3
