// CPU ORACLE — test infrastructure only (see springy_oracle.hpp).
//
// extern "C" entry points over the oracle restatement, so the Python test-suite and bench.py's
// cpu_baseline leg can drive it through ctypes.  Every entry point has an f32 and an f64 form
// (suffix _f32 / _f64); the f64 form takes double arrays.  Errors that the reference reports by
// panicking (bad index, bad state width, non-manifold edge) are returned as non-zero codes with a
// message in the caller's buffer.
#include "rigid_oracle.hpp"
#include "springy_oracle.hpp"

#include <cstdio>
#include <memory>
#include <string>

using namespace oracle;

namespace {

template <class R> struct CConfig {
  int integration;  // 0 = Euler, 1 = Rk4
  R dt;
  R gravity[3];
  R wind[3];
  R lift_coefficient, drag_coefficient, coefficient_of_restitution, coefficient_of_friction;
};

template <class R> Config<R> to_cfg(const CConfig<R>* c) {
  Config<R> o;
  if (!c) return o;
  o.integration = c->integration == 0 ? Integration::Euler : Integration::Rk4;
  o.dt = c->dt;
  o.gravity = {c->gravity[0], c->gravity[1], c->gravity[2]};
  o.wind = {c->wind[0], c->wind[1], c->wind[2]};
  o.lift_coefficient = c->lift_coefficient;
  o.drag_coefficient = c->drag_coefficient;
  o.coefficient_of_restitution = c->coefficient_of_restitution;
  o.coefficient_of_friction = c->coefficient_of_friction;
  return o;
}

void set_err(char* err, size_t n, const char* msg) {
  if (err && n) { std::snprintf(err, n, "%s", msg); }
}

template <class R> std::vector<V3<R>> to_v3(const R* p, size_t n) {
  std::vector<V3<R>> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
  return out;
}

template <class R> std::vector<Point<R>> to_points(const R* rec, size_t n) {
  std::vector<Point<R>> out;
  out.reserve(n);
  for (size_t i = 0; i < n; ++i) out.push_back(Point<R>::from_state_vector(std::vector<R>(rec + 10 * i, rec + 10 * i + 10)));
  return out;
}

template <class R> void* mesh_new(const R* pos, size_t n, const uint64_t* idx, size_t nidx, R mass, R k, R c,
                                  int has_torsion, R tk, R tc, const uint64_t* ovr_pairs, const R* ovr_kc,
                                  size_t novr, int has_overrides, char* err, size_t errlen) {
  try {
    std::vector<size_t> ix(idx, idx + nidx);
    std::vector<SpringOverride<R>> ovr;
    for (size_t i = 0; i < novr; ++i) ovr.push_back({ovr_pairs[2 * i], ovr_pairs[2 * i + 1], ovr_kc[2 * i], ovr_kc[2 * i + 1]});
    TorsionalSpring<R> t{tk, tc, R(0)};
    return new SpringyMesh<R>(to_v3(pos, n), ix, mass, k, c, has_torsion ? &t : nullptr, has_overrides ? &ovr : nullptr);
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

template <class R> void* cloth_new(size_t rows, size_t cols, R spacing, const R* position, R pm, R tk, R tc, R sk,
                                   R sc, R bk, R bc, const uint64_t* pins, size_t npins, int plane, char* err,
                                   size_t errlen) {
  try {
    std::vector<size_t> p(pins, pins + npins);
    return new SpringyMesh<R>(oracle::cloth_new<R>(rows, cols, spacing, V3<R>{position[0], position[1], position[2]}, pm, tk, tc,
                                           sk, sc, bk, bc, p, plane));
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

template <class R> struct Sim {
  Simulation<R> sim;
};

}  // namespace

#define ORC_INSTANTIATE(R, SFX)                                                                                      \
  extern "C" {                                                                                                       \
  void* orc_mesh_new##SFX(const R* pos, size_t n, const uint64_t* idx, size_t nidx, R mass, R k, R c,               \
                          int has_torsion, R tk, R tc, const uint64_t* ovr_pairs, const R* ovr_kc, size_t novr,     \
                          int has_overrides, char* err, size_t errlen) {                                            \
    return mesh_new<R>(pos, n, idx, nidx, mass, k, c, has_torsion, tk, tc, ovr_pairs, ovr_kc, novr, has_overrides, \
                       err, errlen);                                                                                 \
  }                                                                                                                  \
  void* orc_cloth_new##SFX(size_t rows, size_t cols, R spacing, const R* position, R pm, R tk, R tc, R sk, R sc,   \
                           R bk, R bc, const uint64_t* pins, size_t npins, int plane, char* err, size_t errlen) {   \
    return cloth_new<R>(rows, cols, spacing, position, pm, tk, tc, sk, sc, bk, bc, pins, npins, plane, err, errlen); \
  }                                                                                                                  \
  void orc_mesh_free##SFX(void* h) { delete static_cast<SpringyMesh<R>*>(h); }                                      \
  int orc_mesh_add_strut##SFX(void* h, uint64_t a, uint64_t b, R k, R c) {                                           \
    try { static_cast<SpringyMesh<R>*>(h)->add_strut(a, b, k, c); return 0; } catch (...) { return 1; }             \
  }                                                                                                                  \
  void orc_mesh_add_pin##SFX(void* h, uint64_t i) { static_cast<SpringyMesh<R>*>(h)->add_pin(i); }                  \
  void orc_mesh_counts##SFX(void* h, uint64_t* out4) {                                                               \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    out4[0] = m->points.size(); out4[1] = m->struts.size(); out4[2] = m->faces.size();                              \
    out4[3] = m->pinned_points.size();                                                                               \
  }                                                                                                                  \
  void orc_mesh_struts##SFX(void* h, uint64_t* v01, R* kcl, int64_t* faces, uint8_t* has_tor, R* tor3) {            \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    for (size_t i = 0; i < m->struts.size(); ++i) {                                                                 \
      const auto& s = m->struts[i];                                                                                  \
      v01[2 * i] = s.v0; v01[2 * i + 1] = s.v1;                                                                      \
      kcl[3 * i] = s.stiffness; kcl[3 * i + 1] = s.damping; kcl[3 * i + 2] = s.length;                              \
      faces[2 * i] = s.face0; faces[2 * i + 1] = s.face1;                                                            \
      has_tor[i] = s.has_torsion ? 1 : 0;                                                                            \
      tor3[3 * i] = s.torsion.spring_constant; tor3[3 * i + 1] = s.torsion.damping;                                 \
      tor3[3 * i + 2] = s.torsion.rest_angle;                                                                        \
    }                                                                                                                \
  }                                                                                                                  \
  void orc_mesh_set_torsion_rest##SFX(void* h, uint64_t strut, R rest) {                                             \
    static_cast<SpringyMesh<R>*>(h)->struts[strut].torsion.rest_angle = rest;                                        \
  }                                                                                                                  \
  void orc_mesh_faces##SFX(void* h, uint64_t* s3, uint64_t* v3) {                                                    \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    for (size_t i = 0; i < m->faces.size(); ++i) {                                                                   \
      const auto& f = m->faces[i];                                                                                   \
      s3[3 * i] = f.s0; s3[3 * i + 1] = f.s1; s3[3 * i + 2] = f.s2;                                                  \
      v3[3 * i] = f.v0; v3[3 * i + 1] = f.v1; v3[3 * i + 2] = f.v2;                                                  \
    }                                                                                                                \
  }                                                                                                                  \
  void orc_mesh_pins##SFX(void* h, uint64_t* out) {                                                                  \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    for (size_t i = 0; i < m->pinned_points.size(); ++i) out[i] = m->pinned_points[i];                               \
  }                                                                                                                  \
  void orc_mesh_points##SFX(void* h, R* out10) {                                                                     \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    for (size_t i = 0; i < m->points.size(); ++i) {                                                                  \
      auto s = m->points[i].as_state();                                                                              \
      for (int j = 0; j < 10; ++j) out10[10 * i + j] = s[j];                                                         \
    }                                                                                                                \
  }                                                                                                                  \
  int orc_mesh_set_points##SFX(void* h, const R* in10, size_t n) {                                                   \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    if (n != m->points.size()) return 1;                                                                             \
    m->points = to_points(in10, n);                                                                                  \
    return 0;                                                                                                        \
  }                                                                                                                  \
  int orc_mesh_accumulate_forces##SFX(void* h, const CConfig<R>* cfg) {                                              \
    try { static_cast<SpringyMesh<R>*>(h)->accumulate_forces(to_cfg(cfg)); return 0; } catch (...) { return 1; }    \
  }                                                                                                                  \
  void orc_mesh_apply_strut_forces##SFX(void* h) { static_cast<SpringyMesh<R>*>(h)->apply_strut_forces(); }          \
  void orc_mesh_apply_face_forces##SFX(void* h, const CConfig<R>* cfg) {                                             \
    static_cast<SpringyMesh<R>*>(h)->apply_face_forces(to_cfg(cfg));                                                 \
  }                                                                                                                  \
  int orc_mesh_apply_torsional_forces##SFX(void* h) {                                                                \
    try { static_cast<SpringyMesh<R>*>(h)->apply_torsional_forces(); return 0; } catch (...) { return 1; }          \
  }                                                                                                                  \
  void orc_mesh_set_torsion_in_step##SFX(void* h, int on) { static_cast<SpringyMesh<R>*>(h)->torsion_in_step = on != 0; } \
  void orc_mesh_clear_forces##SFX(void* h) { static_cast<SpringyMesh<R>*>(h)->clear_forces(); }                      \
  int orc_mesh_update_points##SFX(void* h, const R* new10, size_t n, void** obs, size_t nobs,                        \
                                  const CConfig<R>* cfg) {                                                           \
    try {                                                                                                            \
      std::vector<const CollidableMesh<R>*> o;                                                                       \
      for (size_t i = 0; i < nobs; ++i) o.push_back(static_cast<CollidableMesh<R>*>(obs[i]));                         \
      static_cast<SpringyMesh<R>*>(h)->update_points(to_points(new10, n), o, to_cfg(cfg));                           \
      return 0;                                                                                                      \
    } catch (...) { return 1; }                                                                                      \
  }                                                                                                                  \
  int orc_mesh_integrate##SFX(void* h, const CConfig<R>* cfg, R* out10) {                                            \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    Config<R> c = to_cfg(cfg);                                                                                       \
    State<R, Point<R>> st(m->points);                                                                                \
    auto ns = c.integration == Integration::Rk4 ? st.rk4_step(c.dt) : st.euler_step(c.dt);                           \
    auto els = ns.get_elements();                                                                                    \
    for (size_t i = 0; i < els.size(); ++i) {                                                                        \
      auto s = els[i].as_state();                                                                                    \
      for (int j = 0; j < 10; ++j) out10[10 * i + j] = s[j];                                                         \
    }                                                                                                                \
    return 0;                                                                                                        \
  }                                                                                                                  \
  R orc_mesh_face_angle##SFX(void* h, uint64_t face, int corner) {                                                   \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    const auto& f = m->faces[face];                                                                                  \
    return corner == 0 ? m->vertex_angle_0(f) : corner == 1 ? m->vertex_angle_1(f) : m->vertex_angle_2(f);          \
  }                                                                                                                  \
  void orc_mesh_face_normal##SFX(void* h, uint64_t face, R* out3) {                                                  \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    auto n = m->face_normal(m->faces[face]);                                                                         \
    out3[0] = n.x; out3[1] = n.y; out3[2] = n.z;                                                                     \
  }                                                                                                                  \
  void orc_mesh_get_vertices##SFX(void* h, R* pos3, uint64_t* idx) {                                                 \
    auto* m = static_cast<SpringyMesh<R>*>(h);                                                                       \
    for (size_t i = 0; i < m->points.size(); ++i) {                                                                  \
      pos3[3 * i] = m->points[i].position.x; pos3[3 * i + 1] = m->points[i].position.y;                              \
      pos3[3 * i + 2] = m->points[i].position.z;                                                                     \
    }                                                                                                                \
    for (size_t i = 0; i < m->faces.size(); ++i) {                                                                   \
      idx[3 * i] = m->faces[i].v0; idx[3 * i + 1] = m->faces[i].v1; idx[3 * i + 2] = m->faces[i].v2;                 \
    }                                                                                                                \
  }                                                                                                                  \
  void* orc_obstacle_new##SFX(const R* pos, size_t n, const uint64_t* idx, size_t nidx, char* err, size_t errlen) {  \
    try {                                                                                                            \
      return new CollidableMesh<R>(to_v3(pos, n), std::vector<size_t>(idx, idx + nidx));                             \
    } catch (const std::exception& e) { set_err(err, errlen, e.what()); return nullptr; }                            \
  }                                                                                                                  \
  void orc_obstacle_free##SFX(void* h) { delete static_cast<CollidableMesh<R>*>(h); }                                \
  void orc_obstacle_counts##SFX(void* h, uint64_t* out3) {                                                           \
    auto* o = static_cast<CollidableMesh<R>*>(h);                                                                    \
    out3[0] = o->vertices.size(); out3[1] = o->edges.size(); out3[2] = o->faces.size();                              \
  }                                                                                                                  \
  void orc_obstacle_edges##SFX(void* h, R* out6) {                                                                   \
    auto* o = static_cast<CollidableMesh<R>*>(h);                                                                    \
    for (size_t i = 0; i < o->edges.size(); ++i) {                                                                   \
      const auto& e = o->edges[i];                                                                                   \
      R v[6] = {e.first.x, e.first.y, e.first.z, e.second.x, e.second.y, e.second.z};                                \
      for (int j = 0; j < 6; ++j) out6[6 * i + j] = v[j];                                                            \
    }                                                                                                                \
  }                                                                                                                  \
  void orc_obstacle_faces##SFX(void* h, R* out9) {                                                                   \
    auto* o = static_cast<CollidableMesh<R>*>(h);                                                                    \
    for (size_t i = 0; i < o->faces.size(); ++i) {                                                                   \
      const auto& f = o->faces[i];                                                                                   \
      R v[9] = {f.v0.x, f.v0.y, f.v0.z, f.v1.x, f.v1.y, f.v1.z, f.v2.x, f.v2.y, f.v2.z};                             \
      for (int j = 0; j < 9; ++j) out9[9 * i + j] = v[j];                                                            \
    }                                                                                                                \
  }                                                                                                                  \
  int64_t orc_collided_face##SFX(void** obs, size_t nobs, const R* old3, const R* new3, R dt_d) {                    \
    std::vector<const ObstacleFace<R>*> fl;                                                                          \
    for (size_t i = 0; i < nobs; ++i)                                                                                \
      for (auto& f : static_cast<CollidableMesh<R>*>(obs[i])->faces) fl.push_back(&f);                               \
    return CollidableMesh<R>::get_collided_face_from_list(fl, {old3[0], old3[1], old3[2]},                           \
                                                          {new3[0], new3[1], new3[2]}, dt_d);                        \
  }                                                                                                                  \
  /* Batch form of the scan above: one first-hit face index per segment (-1 for none). */                           \
  void orc_collided_faces##SFX(void** obs, size_t nobs, const R* old3, const R* new3, size_t n, R dt_d,             \
                               int64_t* out) {                                                                       \
    std::vector<const ObstacleFace<R>*> fl;                                                                          \
    for (size_t i = 0; i < nobs; ++i)                                                                                \
      for (auto& f : static_cast<CollidableMesh<R>*>(obs[i])->faces) fl.push_back(&f);                               \
    for (size_t k = 0; k < n; ++k)                                                                                   \
      out[k] = CollidableMesh<R>::get_collided_face_from_list(                                                       \
          fl, {old3[3 * k], old3[3 * k + 1], old3[3 * k + 2]}, {new3[3 * k], new3[3 * k + 1], new3[3 * k + 2]}, dt_d); \
  }                                                                                                                  \
  R orc_duration_roundtrip##SFX(R dt) { return duration_roundtrip<R>(dt); }                                          \
  void* orc_sim_new##SFX(void** meshes, size_t nm, void** obs, size_t nobs) {                                        \
    auto* s = new Sim<R>();                                                                                          \
    for (size_t i = 0; i < nm; ++i) s->sim.meshes.push_back(*static_cast<SpringyMesh<R>*>(meshes[i]));               \
    for (size_t i = 0; i < nobs; ++i) s->sim.obstacles.push_back(*static_cast<CollidableMesh<R>*>(obs[i]));          \
    return s;                                                                                                        \
  }                                                                                                                  \
  void orc_sim_free##SFX(void* h) { delete static_cast<Sim<R>*>(h); }                                                \
  void orc_sim_set_config##SFX(void* h, const CConfig<R>* cfg) { static_cast<Sim<R>*>(h)->sim.config = to_cfg(cfg); } \
  int orc_sim_step##SFX(void* h, uint64_t n) {                                                                       \
    try { for (uint64_t i = 0; i < n; ++i) static_cast<Sim<R>*>(h)->sim.step(); return 0; } catch (...) { return 1; } \
  }                                                                                                                  \
  void* orc_sim_mesh##SFX(void* h, uint64_t i) { return &static_cast<Sim<R>*>(h)->sim.meshes[i]; }                   \
  /* rigid bodies (rigid_oracle.hpp): RigidBody::new per body, the obstacles copied */                               \
  void* orc_rigid_sim_new##SFX(const R* pos3, const R* mass, size_t n, void** obs, size_t nobs) {                    \
    try {                                                                                                            \
      auto* s = new RigidSimulation<R>();                                                                            \
      for (size_t i = 0; i < n; ++i) s->bodies.emplace_back(V3<R>{pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]}, mass[i]); \
      for (size_t i = 0; i < nobs; ++i) s->obstacles.push_back(*static_cast<CollidableMesh<R>*>(obs[i]));            \
      return s;                                                                                                      \
    } catch (...) { return nullptr; }                                                                                \
  }                                                                                                                  \
  void orc_rigid_sim_free##SFX(void* h) { delete static_cast<RigidSimulation<R>*>(h); }                             \
  void orc_rigid_sim_set_config##SFX(void* h, int integration, R dt, R restitution, const R* g3, const R* t3) {      \
    auto& c = static_cast<RigidSimulation<R>*>(h)->config;                                                           \
    c.integration = integration == 0 ? Integration::Euler : Integration::Rk4;                                        \
    c.dt = dt; c.coefficient_of_restitution = restitution;                                                           \
    c.gravity = {g3[0], g3[1], g3[2]}; c.torque = {t3[0], t3[1], t3[2]};                                             \
  }                                                                                                                  \
  int orc_rigid_sim_step##SFX(void* h, uint64_t n) {                                                                 \
    try { for (uint64_t i = 0; i < n; ++i) static_cast<RigidSimulation<R>*>(h)->step(); return 0; } catch (...) { return 1; } \
  }                                                                                                                  \
  void orc_rigid_sim_get_states##SFX(void* h, R* out) {                                                              \
    auto* s = static_cast<RigidSimulation<R>*>(h);                                                                   \
    for (size_t i = 0; i < s->bodies.size(); ++i) {                                                                  \
      auto v = s->bodies[i].state.as_state();                                                                        \
      for (size_t k = 0; k < 29; ++k) out[29 * i + k] = v[k];                                                        \
    }                                                                                                                \
  }                                                                                                                  \
  void orc_rigid_sim_set_states##SFX(void* h, const R* in) {                                                         \
    auto* s = static_cast<RigidSimulation<R>*>(h);                                                                   \
    for (size_t i = 0; i < s->bodies.size(); ++i)                                                                    \
      s->bodies[i].state = RigidState<R>::from_state_vector(std::vector<R>(in + 29 * i, in + 29 * i + 29));          \
  }                                                                                                                  \
  void orc_rigid_sim_apply_impulse##SFX(void* h, uint64_t body, const R* impulse3, const R* at3) {                    \
    static_cast<RigidSimulation<R>*>(h)->bodies[body].state.apply_impulse({impulse3[0], impulse3[1], impulse3[2]},   \
                                                                          {at3[0], at3[1], at3[2]});                 \
  }                                                                                                                  \
  }

ORC_INSTANTIATE(float, _f32)
ORC_INSTANTIATE(double, _f64)

// ---------------------------------------------------------------------------------------------
// The reference's own integrator tests, state.rs:115-280, restated with their Stateful test types.
namespace {
struct TestPoint {  // state.rs:120-164
  float px, py, pz, vx, vy, vz;
  static size_t num_state_elements() { return 6; }
  static TestPoint from_state_vector(const std::vector<float>& s) {
    if (s.size() != 6) throw std::invalid_argument("State Vector incorrect size!");
    return TestPoint{s[0], s[1], s[2], s[3], s[4], s[5]};
  }
  std::vector<float> derivative() const { return {vx, vy, vz, 1.0f, -1.0f, 0.0f}; }
  std::vector<float> as_state() const { return {px, py, pz, vx, vy, vz}; }
};
struct ExampleFn {  // state.rs:187-216
  float y, t, timestep;
  static size_t num_state_elements() { return 3; }
  static ExampleFn from_state_vector(const std::vector<float>& s) {
    if (s.size() != 3) throw std::invalid_argument("State Vector incorrect size!");
    return ExampleFn{s[0], s[1], s[2]};
  }
  std::vector<float> derivative() const { return {y - t * t + 1.0f, 1.0f, 0.0f}; }  // powi(t, 2) = t*t
  std::vector<float> as_state() const { return {y, t, timestep}; }
};
}  // namespace

extern "C" {
// state.rs:166-185: one Euler step (h = 0.5) of the 6-element test point.  out6 = x, v after.
void orc_state_test_euler(float* initial6, float* out6) {
  State<float, TestPoint> st({TestPoint{0, 0, 0, 0, 0, 1}});
  auto iv = st.as_vector();
  for (int i = 0; i < 6; ++i) initial6[i] = iv[i];
  auto ns = st.euler_step(0.5f).get_elements()[0];
  float v[6] = {ns.px, ns.py, ns.pz, ns.vx, ns.vy, ns.vz};
  for (int i = 0; i < 6; ++i) out6[i] = v[i];
}
// state.rs:218-280: four RK4 steps (h = 0.5) of y' = y - t^2 + 1, y(0) = 0.5; out = (y,t,h) x 4.
void orc_state_test_rk4(float* out12) {
  std::vector<ExampleFn> els{ExampleFn{0.5f, 0.0f, 0.5f}};
  for (int k = 0; k < 4; ++k) {
    State<float, ExampleFn> st(els);
    els = st.rk4_step(0.5f).get_elements();
    out12[3 * k] = els[0].y; out12[3 * k + 1] = els[0].t; out12[3 * k + 2] = els[0].timestep;
  }
}
// utils.rs:26-40: returns the missing element, or -1 when the reference would panic.
int64_t orc_tuple_difference(int64_t t0, int64_t t1, int64_t t2, int64_t d0, int64_t d1) {
  try { return tuple_difference<int64_t>(t0, t1, t2, d0, d1); } catch (...) { return -1; }
}
}
