// CPU ORACLE — test infrastructure only (imported by tests/, __graft_entry__.smoke() and bench.py's
// CPU baseline; never by the product library).
//
// Restatement of the reference rigid-body step in its own operation order:
//   src/simulation/rigidbody/rigidbody.rs:14-199 (State, Stateful: as_state, derivative,
//     from_state_vector), :201-251 (RigidBody::new: the unit cube of graphics/forms.rs:91-146 as a
//     CollidableMesh, inertia mass/6 on the diagonal, inverted), :257-319 (update_state: vertex-face
//     collisions with impulses), :322-338 (forces, torques);
//   src/simulation/rigidbody/simulation.rs:26-52 (Simulation::step), config.rs:7-25;
//   src/simulation/state.rs:75-106 (the generic Euler / RK4 over the state vector).
// cgmath 0.18 operations are restated from the crate's source semantics (not vendored):
//   Quaternion * Quaternion: (s s' - x x' - y y' - z z', s x' + x s' + y z' - z y', ...), left to right;
//   scalar * Quaternion, Quaternion * scalar: per component;
//   Quaternion::normalize: q * (1 / sqrt(s s + ((x x + y y) + z z)));
//   Matrix3::from(Quaternion): x2 = x + x, ..., columns (1 - yy2 - zz2, xy2 + sz2, xz2 - sy2), ...;
//   Matrix3 * Vector3: row(i) . v with dot = (a.x b.x + a.y b.y) + a.z b.z; Matrix3 * Matrix3 column
//   by column; Matrix3::invert: columns (c1 x c2, c2 x c0, c0 x c1) / det, transposed, det by the
//   first-column cofactor expansion.
// Parity unpinned by the reference: it has no rigid-body test.  Pinned here by closed forms and
// invariants (tests/test_rigid_oracle.py).
#pragma once

#include "springy_oracle.hpp"

namespace oracle {

template <class R> struct Quat {
  R s;
  V3<R> v;
};
template <class R> Quat<R> qmul(Quat<R> a, Quat<R> b) {
  return {a.s * b.s - a.v.x * b.v.x - a.v.y * b.v.y - a.v.z * b.v.z,
          {a.s * b.v.x + a.v.x * b.s + a.v.y * b.v.z - a.v.z * b.v.y,
           a.s * b.v.y + a.v.y * b.s + a.v.z * b.v.x - a.v.x * b.v.z,
           a.s * b.v.z + a.v.z * b.s + a.v.x * b.v.y - a.v.y * b.v.x}};
}
template <class R> Quat<R> qscale(R k, Quat<R> q) { return {k * q.s, k * q.v}; }
template <class R> Quat<R> qnormalize(Quat<R> q) {
  const R inv = R(1) / std::sqrt(q.s * q.s + dot(q.v, q.v));
  return {q.s * inv, q.v * inv};
}

// Matrix3 as three columns.
template <class R> struct M3 {
  V3<R> c[3];
};
template <class R> V3<R> row(const M3<R>& m, int i) {
  auto at = [&](const V3<R>& c) { return i == 0 ? c.x : i == 1 ? c.y : c.z; };
  return {at(m.c[0]), at(m.c[1]), at(m.c[2])};
}
template <class R> V3<R> mul(const M3<R>& m, V3<R> v) { return {dot(row(m, 0), v), dot(row(m, 1), v), dot(row(m, 2), v)}; }
template <class R> M3<R> mul(const M3<R>& a, const M3<R>& b) { return {{mul(a, b.c[0]), mul(a, b.c[1]), mul(a, b.c[2])}}; }
template <class R> M3<R> transpose(const M3<R>& m) { return {{row(m, 0), row(m, 1), row(m, 2)}}; }
template <class R> M3<R> from_quat(Quat<R> q) {
  const R x2 = q.v.x + q.v.x, y2 = q.v.y + q.v.y, z2 = q.v.z + q.v.z;
  const R xx2 = x2 * q.v.x, xy2 = x2 * q.v.y, xz2 = x2 * q.v.z;
  const R yy2 = y2 * q.v.y, yz2 = y2 * q.v.z, zz2 = z2 * q.v.z;
  const R sy2 = y2 * q.s, sz2 = z2 * q.s, sx2 = x2 * q.s;
  return {{{R(1) - yy2 - zz2, xy2 + sz2, xz2 - sy2},
           {xy2 - sz2, R(1) - xx2 - zz2, yz2 + sx2},
           {xz2 + sy2, yz2 - sx2, R(1) - xx2 - yy2}}};
}
template <class R> R determinant(const M3<R>& m) {
  const V3<R>&a = m.c[0], &b = m.c[1], &c = m.c[2];
  return a.x * (b.y * c.z - c.y * b.z) - b.x * (a.y * c.z - c.y * a.z) + c.x * (a.y * b.z - b.y * a.z);
}
template <class R> M3<R> invert(const M3<R>& m) {
  const R det = determinant(m);
  if (det == R(0)) throw std::invalid_argument("Uninvertable moment of inertia!");  // rigidbody.rs:225-227
  return transpose(M3<R>{{cross(m.c[1], m.c[2]) / det, cross(m.c[2], m.c[0]) / det, cross(m.c[0], m.c[1]) / det}});
}

// rigidbody/config.rs:7-25
template <class R> struct RigidConfig {
  Integration integration = Integration::Rk4;
  R dt = R(0.001);
  R coefficient_of_restitution = R(0.7);
  V3<R> gravity = {R(0), R(0), R(0)};
  V3<R> torque = {R(0), R(0), R(0)};
};

// rigidbody.rs:14-199
template <class R> struct RigidState {
  V3<R> position;
  Quat<R> rotation;
  V3<R> linear_momentum, angular_momentum;
  R mass;
  M3<R> inertia_inv0;  // initial_moment_of_intertia_inverted
  V3<R> force, torque;

  static size_t num_state_elements() { return 29; }
  void normalize_rotation() { rotation = qnormalize(rotation); }
  M3<R> inertia_inv() const {
    const M3<R> r = from_quat(rotation);
    return mul(mul(r, inertia_inv0), transpose(r));
  }
  void apply_impulse(V3<R> impulse, V3<R> at) {
    linear_momentum = linear_momentum + impulse;
    angular_momentum = angular_momentum + cross(at, impulse);
  }
  V3<R> velocity() const { return linear_momentum / mass; }
  V3<R> angular_velocity() const { return mul(inertia_inv(), angular_momentum); }

  std::vector<R> as_state() const {
    const M3<R>& m = inertia_inv0;
    return {position.x, position.y, position.z, rotation.v.x, rotation.v.y, rotation.v.z, rotation.s,
            linear_momentum.x, linear_momentum.y, linear_momentum.z, angular_momentum.x, angular_momentum.y,
            angular_momentum.z, mass, m.c[0].x, m.c[0].y, m.c[0].z, m.c[1].x, m.c[1].y, m.c[1].z, m.c[2].x,
            m.c[2].y, m.c[2].z, force.x, force.y, force.z, torque.x, torque.y, torque.z};
  }
  std::vector<R> derivative() const {
    const V3<R> v = velocity();
    const Quat<R> qd = qmul(qscale(R(0.5), Quat<R>{R(0), angular_velocity()}), rotation);
    std::vector<R> d = {v.x, v.y, v.z, qd.v.x, qd.v.y, qd.v.z, qd.s, force.x, force.y, force.z,
                        torque.x, torque.y, torque.z};
    d.resize(29, R(0));
    return d;
  }
  static RigidState from_state_vector(const std::vector<R>& s) {
    if (s.size() != 29) throw std::invalid_argument("Incorrect size of state vector!");
    RigidState r;
    r.position = {s[0], s[1], s[2]};
    r.rotation = {s[6], {s[3], s[4], s[5]}};
    r.linear_momentum = {s[7], s[8], s[9]};
    r.angular_momentum = {s[10], s[11], s[12]};
    r.mass = s[13];
    r.inertia_inv0 = {{{s[14], s[15], s[16]}, {s[17], s[18], s[19]}, {s[20], s[21], s[22]}}};
    r.force = {s[23], s[24], s[25]};
    r.torque = {s[26], s[27], s[28]};
    return r;
  }
};

// graphics/forms.rs:91-146
template <class R> CollidableMesh<R> unit_cube() {
  const R h = R(0.5);
  std::vector<V3<R>> p = {{-h, -h, h}, {h, -h, h},  {h, h, h},  {-h, h, h},
                          {-h, -h, -h}, {h, -h, -h}, {h, h, -h}, {-h, h, -h}};
  std::vector<size_t> idx = {0, 1, 2, 2, 3, 0, 1, 5, 6, 6, 2, 1, 7, 6, 5, 5, 4, 7,
                             4, 0, 3, 3, 7, 4, 4, 5, 1, 1, 0, 4, 3, 2, 6, 6, 7, 3};
  return CollidableMesh<R>(p, idx);
}

template <class R> struct RigidBody {
  RigidState<R> state;
  CollidableMesh<R> mesh = unit_cube<R>();

  // rigidbody.rs:213-251
  RigidBody(V3<R> position, R mass) {
    const R i = mass / R(6);
    const M3<R> inertia = {{{i, R(0), R(0)}, {R(0), i, R(0)}, {R(0), R(0), i}}};
    state.position = position;
    state.rotation = {R(1), {R(0), R(0), R(0)}};
    state.linear_momentum = mass * zero3<R>();
    state.angular_momentum = mul(inertia, zero3<R>());
    state.mass = mass;
    state.inertia_inv0 = invert(inertia);
    state.force = zero3<R>();
    state.torque = zero3<R>();
  }

  // rigidbody.rs:257-319
  void update_state(RigidState<R> ns, const std::vector<const ObstacleFace<R>*>& fl, const RigidConfig<R>& cfg) {
    const R dt_d = duration_roundtrip<R>(cfg.dt);
    const M3<R> r_old = from_quat(state.rotation), r_new = from_quat(ns.rotation);
    std::vector<V3<R>> olds, news;
    for (const V3<R>& v : mesh.vertices) olds.push_back(mul(r_old, v) + state.position);
    for (const V3<R>& v : mesh.vertices) news.push_back(mul(r_new, v) + ns.position);
    for (size_t k = 0; k < olds.size(); ++k) {
      const V3<R> np = news[k], op = olds[k];
      const long hit = CollidableMesh<R>::get_collided_face_from_list(fl, op, np, dt_d);
      if (hit < 0) continue;
      const ObstacleFace<R>& f = *fl[(size_t)hit];
      const R od = f.distance_from_plane(op), nd = f.distance_from_plane(np);
      const V3<R> r = op - state.position;
      const R frac = od / (od - nd);
      const V3<R> cv = state.velocity() + cross(state.angular_velocity(), r);
      const V3<R> cp = op + (cfg.dt * frac) * cv;
      const R vn = dot(cv, f.normal());
      const R j = (-(R(1) + cfg.coefficient_of_restitution) * vn) /
                  (R(1) / state.mass + dot(f.normal(), mul(state.inertia_inv(), cross(cross(r, f.normal()), r))));
      const V3<R> impulse = j * f.normal();
      ns.position = (cp - r) + (epsilon<R>() * R(2)) * f.normal();
      ns.apply_impulse(impulse, r);
    }
    state = ns;
  }
};

// rigidbody/simulation.rs:10-52
template <class R> struct RigidSimulation {
  RigidConfig<R> config;
  std::vector<RigidBody<R>> bodies;
  std::vector<CollidableMesh<R>> obstacles;

  R step() {
    std::vector<const ObstacleFace<R>*> fl;
    for (const auto& o : obstacles)
      for (const auto& f : o.faces) fl.push_back(&f);
    for (auto& b : bodies) {
      b.state.force = b.state.force + config.gravity;     // accumulate_forces (:322-324)
      b.state.torque = b.state.torque + config.torque;    // accumulate_torques (:326-328)
      State<R, RigidState<R>> st({b.state});
      State<R, RigidState<R>> next = config.integration == Integration::Rk4 ? st.rk4_step(config.dt)
                                                                              : st.euler_step(config.dt);
      RigidState<R> ns = next.get_elements()[0];
      ns.normalize_rotation();
      b.update_state(ns, fl, config);
      b.state.force = zero3<R>();   // clear_forces (:330-332)
      b.state.torque = zero3<R>();  // clear_torques (:334-336)
    }
    return duration_roundtrip<R>(config.dt);
  }
};

}  // namespace oracle
