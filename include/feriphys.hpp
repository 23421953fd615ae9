// C++ mirror of the reference crate's springy-mesh API over the C ABI of feriphys_gpu.h: the same
// type and method names, argument meanings and order as springy_mesh.rs, springy/cloth.rs,
// collidable_mesh.rs, springy/config.rs and springy/simulation.rs, so that C++ callers (and tests)
// read like the reference.  Header-only; link with -lferiphys_gpu.
//
// Where Rust and C++ differ:
//   * SpringyMesh::new / Cloth::new / CollidableMesh::new / Simulation::new are constructors;
//   * the reference panics, the mirror throws feriphys::Panic (with the library's message and
//     status code);
//   * Simulation::new takes its meshes and obstacles by value (std::move), as the Rust one does, and
//     get_meshes()[i].method(...) is Simulation::method(i, ...);
//   * the construction tables that the reference's unit tests read through private fields are
//     exposed read-only (struts(), faces()).
#pragma once

#include <array>
#include <chrono>
#include <cstdint>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "feriphys_gpu.h"

namespace feriphys {

using Vector3 = std::array<float, 3>;

struct Panic : std::runtime_error {
  int code;
  Panic(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
  if (rc != FP_OK) throw Panic(rc, fp_last_error());
}

// Point (springy_mesh.rs:177-192): the 10-float Stateful record (:242-255).
struct Point {
  float mass;
  Vector3 position, velocity, accumulated_force;
};
static_assert(sizeof(Point) == 10 * sizeof(float), "Point is the 10-float record");

// TorsionalSpringConfig (springy_mesh.rs:33-36), SpringConfig and StrutKey (:38-42, 95-107).
struct TorsionalSpringConfig {
  float spring_constant, spring_damping;
};
struct SpringConfig {
  float constant, damping;
};
using StrutKey = std::pair<uint64_t, uint64_t>;

// state.rs:5-8 and springy/config.rs:9-33 (Config::default() from the library).
enum class Integration : int32_t { Euler = FP_EULER, Rk4 = FP_RK4 };
struct Config {
  Integration integration;
  float dt;
  Vector3 gravity, wind;
  float lift_coefficient, drag_coefficient, coefficient_of_restitution, coefficient_of_friction;

  static Config default_config() {
    fp_config c;
    fp_config_default(&c);
    return from_c(c);
  }
  static Config from_c(const fp_config& c) {
    return Config{static_cast<Integration>(c.integration), c.dt, {c.gravity[0], c.gravity[1], c.gravity[2]},
                  {c.wind[0], c.wind[1], c.wind[2]}, c.lift_coefficient, c.drag_coefficient,
                  c.coefficient_of_restitution, c.coefficient_of_friction};
  }
  fp_config to_c() const {
    return fp_config{static_cast<int32_t>(integration), dt, {gravity[0], gravity[1], gravity[2]},
                     {wind[0], wind[1], wind[2]}, lift_coefficient, drag_coefficient, coefficient_of_restitution,
                     coefficient_of_friction};
  }
};

// Strut (springy_mesh.rs:62-71) and Face (:110-116) as the constructor built them.
struct TorsionalSpring {
  float spring_constant, damping, rest_angle;
};
struct Strut {
  std::pair<uint64_t, uint64_t> vertex_indices;
  float stiffness, damping, length;
  std::pair<std::optional<uint64_t>, std::optional<uint64_t>> face_indices;
  std::optional<TorsionalSpring> torsional_spring;
};
struct Face {
  std::array<uint64_t, 3> strut_indices, vertex_indices;
};

class SpringyMesh {
 public:
  // SpringyMesh::new (springy_mesh.rs:267-401)
  SpringyMesh(const std::vector<Vector3>& vertex_positions, const std::vector<uint64_t>& vertex_indices, float mass,
              float default_stiffness, float default_damping,
              std::optional<TorsionalSpringConfig> torsional_spring_config = std::nullopt,
              const std::map<StrutKey, SpringConfig>* strut_overrides = nullptr) {
    std::vector<fp_strut_override> ovr;
    if (strut_overrides)
      for (const auto& [k, v] : *strut_overrides) ovr.push_back(fp_strut_override{k.first, k.second, v.constant, v.damping});
    fp_torsion_config tc{};
    if (torsional_spring_config)
      tc = fp_torsion_config{torsional_spring_config->spring_constant, torsional_spring_config->spring_damping};
    check(fp_mesh_new(vertex_positions.empty() ? nullptr : vertex_positions[0].data(), vertex_positions.size(),
                      vertex_indices.empty() ? nullptr : vertex_indices.data(), vertex_indices.size(), mass,
                      default_stiffness, default_damping, torsional_spring_config ? &tc : nullptr,
                      ovr.empty() ? nullptr : ovr.data(), ovr.size(), &h_));
  }
  SpringyMesh(SpringyMesh&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  SpringyMesh& operator=(SpringyMesh&& o) noexcept {
    if (this != &o) {
      reset();
      h_ = o.h_;
      o.h_ = nullptr;
    }
    return *this;
  }
  SpringyMesh(const SpringyMesh&) = delete;
  SpringyMesh& operator=(const SpringyMesh&) = delete;
  ~SpringyMesh() { reset(); }

  // add_strut, add_pin (springy_mesh.rs:403-417)
  void add_strut(std::pair<uint64_t, uint64_t> vertex_indices, float stiffness, float damping) {
    check(fp_mesh_add_strut(h_, vertex_indices.first, vertex_indices.second, stiffness, damping));
  }
  void add_pin(uint64_t pin_index) { check(fp_mesh_add_pin(h_, pin_index)); }

  // get_points (:419-421), get_vertices (:500-513)
  std::vector<Point> get_points() const {
    std::vector<Point> p(counts()[0]);
    check(fp_mesh_get_points(h_, p.empty() ? nullptr : &p[0].mass));
    return p;
  }
  std::pair<std::vector<Vector3>, std::vector<uint64_t>> get_vertices() const {
    const auto n = counts();
    std::pair<std::vector<Vector3>, std::vector<uint64_t>> out{std::vector<Vector3>(n[0]),
                                                               std::vector<uint64_t>(3 * n[2])};
    check(fp_mesh_get_vertices(h_, out.first.empty() ? nullptr : out.first[0].data(),
                               out.second.empty() ? nullptr : out.second.data()));
    return out;
  }

  // the construction tables
  std::vector<Strut> struts() const {
    const uint64_t ns = counts()[1];
    std::vector<uint64_t> v01(2 * ns + 1);
    std::vector<float> kcl(3 * ns + 1), kcr(3 * ns + 1);
    std::vector<int64_t> f(2 * ns + 1);
    std::vector<uint8_t> has(ns + 1);
    check(fp_mesh_get_struts(h_, v01.data(), kcl.data(), f.data()));
    check(fp_mesh_get_torsion(h_, has.data(), kcr.data()));
    std::vector<Strut> s(ns);
    for (uint64_t k = 0; k < ns; ++k) {
      auto face = [&](int64_t x) { return x < 0 ? std::optional<uint64_t>{} : std::optional<uint64_t>{(uint64_t)x}; };
      s[k] = Strut{{v01[2 * k], v01[2 * k + 1]}, kcl[3 * k], kcl[3 * k + 1], kcl[3 * k + 2],
                   {face(f[2 * k]), face(f[2 * k + 1])},
                   has[k] ? std::optional<TorsionalSpring>{TorsionalSpring{kcr[3 * k], kcr[3 * k + 1], kcr[3 * k + 2]}}
                          : std::nullopt};
    }
    return s;
  }
  std::vector<Face> faces() const {
    const uint64_t nf = counts()[2];
    std::vector<uint64_t> s3(3 * nf + 1), v3(3 * nf + 1);
    check(fp_mesh_get_faces(h_, s3.data(), v3.data()));
    std::vector<Face> f(nf);
    for (uint64_t k = 0; k < nf; ++k)
      f[k] = Face{{s3[3 * k], s3[3 * k + 1], s3[3 * k + 2]}, {v3[3 * k], v3[3 * k + 1], v3[3 * k + 2]}};
    return f;
  }
  // points, struts, faces, pins
  std::array<uint64_t, 4> counts() const {
    std::array<uint64_t, 4> c{};
    check(fp_mesh_counts(h_, &c[0], &c[1], &c[2], &c[3]));
    return c;
  }

 private:
  friend class Simulation;
  friend struct Cloth;
  explicit SpringyMesh(fp_mesh* h) : h_(h) {}
  fp_mesh* release() {
    fp_mesh* h = h_;
    h_ = nullptr;
    return h;
  }
  void reset() {
    if (h_) fp_mesh_free(h_);
    h_ = nullptr;
  }
  fp_mesh* h_ = nullptr;
};

// Cloth::new (springy/cloth.rs:20-113); the cloth is its `mesh`, as in the reference.
struct Cloth {
  SpringyMesh mesh;
  Cloth(uint64_t rows, uint64_t cols, float spacing, Vector3 position, float point_mass, float tensile_stiffness,
        float tensile_damping, float shear_stiffness, float shear_damping, float binding_stiffness,
        float binding_damping, const std::vector<uint64_t>& pinned_vertices)
      : mesh(make(rows, cols, spacing, position, point_mass, tensile_stiffness, tensile_damping, shear_stiffness,
                  shear_damping, binding_stiffness, binding_damping, pinned_vertices)) {}

 private:
  static SpringyMesh make(uint64_t rows, uint64_t cols, float spacing, Vector3 position, float point_mass, float tk,
                          float tc, float sk, float sc, float bk, float bc, const std::vector<uint64_t>& pins) {
    fp_mesh* h = nullptr;
    check(fp_cloth_new(rows, cols, spacing, position.data(), point_mass, tk, tc, sk, sc, bk, bc,
                       pins.empty() ? nullptr : pins.data(), pins.size(), &h));
    return SpringyMesh(h);
  }
};

// CollidableMesh::new (collidable_mesh.rs:57-105)
class CollidableMesh {
 public:
  CollidableMesh(const std::vector<Vector3>& vertex_positions, const std::vector<uint64_t>& vertex_indices) {
    check(fp_obstacle_new(vertex_positions.empty() ? nullptr : vertex_positions[0].data(), vertex_positions.size(),
                          vertex_indices.empty() ? nullptr : vertex_indices.data(), vertex_indices.size(), &h_));
  }
  CollidableMesh(CollidableMesh&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  CollidableMesh(const CollidableMesh&) = delete;
  CollidableMesh& operator=(const CollidableMesh&) = delete;
  ~CollidableMesh() {
    if (h_) fp_obstacle_free(h_);
  }

 private:
  friend class Simulation;
  fp_obstacle* h_ = nullptr;
};

// springy::Simulation (springy/simulation.rs:10-57)
class Simulation {
 public:
  // Simulation::new: the meshes and obstacles move in (on every path, as Rust's move does).
  Simulation(std::vector<SpringyMesh> meshes, std::vector<CollidableMesh> obstacles) {
    std::vector<fp_mesh*> m;
    std::vector<fp_obstacle*> o;
    for (auto& x : meshes) m.push_back(x.release());
    for (auto& x : obstacles) {
      o.push_back(x.h_);
      x.h_ = nullptr;
    }
    check(fp_sim_new(m.empty() ? nullptr : m.data(), m.size(), o.empty() ? nullptr : o.data(), o.size(), &h_));
  }
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;
  ~Simulation() {
    if (h_) fp_sim_free(h_);
  }

  // step (:27-45), get_timestep: Duration(dt) as the reference returns it
  std::chrono::duration<float> step() {
    float ts = 0.0f;
    check(fp_sim_step(h_, 1, &ts));
    return std::chrono::duration<float>(ts);
  }
  std::chrono::duration<float> get_timestep() const {
    float ts = 0.0f;
    check(fp_sim_get_timestep(h_, &ts));
    return std::chrono::duration<float>(ts);
  }
  // the Config the reference syncs from its UI (:64-77)
  Config get_config() const {
    fp_config c;
    check(fp_sim_get_config(h_, &c));
    return Config::from_c(c);
  }
  void set_config(const Config& c) {
    const fp_config cc = c.to_c();
    check(fp_sim_set_config(h_, &cc));
  }
  uint64_t mesh_count() const {
    uint64_t n = 0;
    check(fp_sim_mesh_count(h_, &n));
    return n;
  }
  // get_meshes()[mesh].<method> (springy_mesh.rs)
  std::vector<Point> get_points(uint64_t mesh) {
    std::vector<Point> p(points_of(mesh));
    check(fp_sim_get_points(h_, mesh, p.empty() ? nullptr : &p[0].mass));
    return p;
  }
  std::pair<std::vector<Vector3>, std::vector<uint64_t>> get_vertices(uint64_t mesh) {
    uint64_t np = 0, nf = 0;
    check(fp_sim_mesh_counts(h_, mesh, &np, nullptr, &nf, nullptr));
    std::pair<std::vector<Vector3>, std::vector<uint64_t>> out{std::vector<Vector3>(np), std::vector<uint64_t>(3 * nf)};
    check(fp_sim_get_vertices(h_, mesh, out.first.empty() ? nullptr : out.first[0].data(),
                              out.second.empty() ? nullptr : out.second.data()));
    return out;
  }
  void accumulate_forces(uint64_t mesh) { check(fp_sim_accumulate_forces(h_, mesh)); }
  void update_points(uint64_t mesh, const std::vector<Point>& new_points) {
    check(fp_sim_update_points(h_, mesh, new_points.empty() ? nullptr : &new_points[0].mass, new_points.size()));
  }
  void clear_forces(uint64_t mesh) { check(fp_sim_clear_forces(h_, mesh)); }

 private:
  uint64_t points_of(uint64_t mesh) const {
    uint64_t np = 0;
    check(fp_sim_mesh_counts(h_, mesh, &np, nullptr, nullptr, nullptr));
    return np;
  }
  fp_sim* h_ = nullptr;
};

}  // namespace feriphys
