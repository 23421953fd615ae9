// The reference's SpringyMesh unit tests (springy_mesh.rs:666-883, the constructor, strip and
// get_vertices cases) restated through the C++ mirror (include/feriphys.hpp) of the library's host
// builder, plus Cloth::new's counts and the reference's panics as exceptions.  Without a GPU the
// simulation reports FP_ENODEV; with one it steps.  Prints "ok" or the first failed check.
#include <cmath>
#include <cstdio>

#include "feriphys.hpp"

using namespace feriphys;

#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      std::printf("failed: %s (line %d)\n", #c, __LINE__);    \
      return 1;                                               \
    }                                                         \
  } while (0)

namespace {

const Vector3 kZero{0, 0, 0}, kX{1, 0, 0}, kY{0, 1, 0}, kZ{0, 0, 1};
Vector3 add(Vector3 a, Vector3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vector3 sub(Vector3 a, Vector3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vector3 mul(Vector3 a, float s) { return {a[0] * s, a[1] * s, a[2] * s}; }

// get_triangle (:666-679): (0,0,0), 2x, 2y; mass 1, k 2, c 3, torsion (4, 5)
SpringyMesh get_triangle() {
  return SpringyMesh({kZero, mul(kX, 2.0f), mul(kY, 2.0f)}, {0, 1, 2}, 1.0f, 2.0f, 3.0f,
                     TorsionalSpringConfig{4.0f, 5.0f});
}
// get_strip (:694-716): four triangles, the last bent 90 degrees
SpringyMesh get_strip() {
  return SpringyMesh({kZero, kX, add(kX, kZ), sub(kY, kX), kY, add(kY, kX)}, {0, 4, 3, 0, 1, 4, 1, 5, 4, 1, 2, 5}, 1.0f,
                     2.0f, 3.0f, TorsionalSpringConfig{4.0f, 5.0f});
}

}  // namespace

int main() {
  {  // ctor_triangle (:720-757)
    const SpringyMesh m = get_triangle();
    const auto c = m.counts();
    CHECK(c[0] == 3 && c[1] == 3 && c[2] == 1);
    const auto f = m.faces();
    CHECK((f[0].strut_indices == std::array<uint64_t, 3>{0, 1, 2}));
    CHECK((f[0].vertex_indices == std::array<uint64_t, 3>{0, 1, 2}));
    const auto p = m.get_points();
    CHECK(p[0].mass == 1.0f / 3.0f);
    CHECK(p[0].position == kZero && p[0].velocity == kZero && p[0].accumulated_force == kZero);
    CHECK(p[1].position == mul(kX, 2.0f) && p[1].velocity == kZero);
    const auto s = m.struts();
    CHECK(s[0].stiffness == 2.0f * (1.0f / 2.0f));  // NOMINAL_STRUT_LENGTH = 1
    CHECK(s[0].damping == 3.0f * (1.0f / 2.0f));
    CHECK(s[0].length == 2.0f);
    CHECK(!s[0].torsional_spring.has_value());  // one face: no hinge
  }
  {  // strip_ctor (:762-866), the tables
    const SpringyMesh m = get_strip();
    const auto c = m.counts();
    CHECK(c[0] == 6 && c[1] == 9 && c[2] == 4);
    const auto f = m.faces();
    const auto s = m.struts();
    CHECK((f[0].vertex_indices == std::array<uint64_t, 3>{0, 4, 3}));
    CHECK((f[1].vertex_indices == std::array<uint64_t, 3>{0, 1, 4}));
    CHECK((f[2].vertex_indices == std::array<uint64_t, 3>{1, 5, 4}));
    CHECK((f[3].vertex_indices == std::array<uint64_t, 3>{1, 2, 5}));
    CHECK((f[0].strut_indices == std::array<uint64_t, 3>{0, 1, 2}));
    CHECK((f[1].strut_indices == std::array<uint64_t, 3>{3, 4, 0}));
    CHECK((f[2].strut_indices == std::array<uint64_t, 3>{5, 6, 4}));
    CHECK((f[3].strut_indices == std::array<uint64_t, 3>{7, 8, 5}));
    const std::pair<uint64_t, uint64_t> v[9] = {{0, 4}, {3, 4}, {0, 3}, {0, 1}, {1, 4}, {1, 5}, {4, 5}, {1, 2}, {2, 5}};
    for (int k = 0; k < 9; ++k) CHECK(s[k].vertex_indices == v[k]);
    using OF = std::optional<uint64_t>;
    const std::pair<OF, OF> fi[9] = {{0, 1}, {0, {}}, {0, {}}, {1, {}}, {1, 2}, {2, 3}, {2, {}}, {3, {}}, {3, {}}};
    for (int k = 0; k < 9; ++k) CHECK(s[k].face_indices == fi[k]);
    const bool hinge[9] = {true, false, false, false, true, true, false, false, false};
    for (int k = 0; k < 9; ++k) CHECK(s[k].torsional_spring.has_value() == hinge[k]);
    for (int k : {0, 4, 5}) CHECK(s[k].torsional_spring->spring_constant == 4.0f && s[k].torsional_spring->damping == 5.0f);
    CHECK(std::fabs(s[0].torsional_spring->rest_angle) <= 1e-6f);
    CHECK(std::fabs(s[4].torsional_spring->rest_angle) <= 1e-6f);
    CHECK(std::fabs(s[5].torsional_spring->rest_angle + 3.14159265f / 2.0f) <= 1e-6f);
  }
  {  // get_vertices (:868-883)
    const auto [pos, idx] = get_strip().get_vertices();
    const std::vector<Vector3> ep{kZero, kX, add(kX, kZ), sub(kY, kX), kY, add(kY, kX)};
    CHECK(pos == ep);
    CHECK((idx == std::vector<uint64_t>{0, 4, 3, 0, 1, 4, 1, 5, 4, 1, 2, 5}));
  }
  // Cloth::new: 32 x 32 gives 4,745 struts (1,984 tensile + 961 shear + 1,800 bending), 1,922 faces
  Cloth cloth(32, 32, 0.05f, kZero, 0.01f, 25.0f, 0.1f, 21.2f, 0.0707f, 5.0f, 0.05f, {992, 1023});
  const auto cc = cloth.mesh.counts();
  CHECK(cc[0] == 1024 && cc[1] == 4745 && cc[2] == 1922 && cc[3] == 2);
  // the reference's panics: an index out of range, a record count that does not match
  bool threw = false;
  try {
    cloth.mesh.add_strut({0, 4096}, 1.0f, 1.0f);
  } catch (const Panic& e) {
    threw = e.code == FP_EINVAL;
  }
  CHECK(threw);
  CHECK(Config::default_config().integration == Integration::Rk4 && Config::default_config().dt == 0.001f);

  // Simulation::new moves the meshes in; without a device it panics with FP_ENODEV
  std::vector<SpringyMesh> meshes;
  meshes.push_back(std::move(cloth.mesh));
  std::vector<CollidableMesh> obstacles;
  obstacles.emplace_back(std::vector<Vector3>{{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}},
                         std::vector<uint64_t>{0, 1, 2, 0, 2, 3});
  try {
    Simulation sim(std::move(meshes), std::move(obstacles));
    Config cfg = Config::default_config();
    cfg.dt = 1.0f / 600.0f;
    sim.set_config(cfg);
    CHECK(sim.step().count() > 0.0f);
    CHECK(sim.get_points(0).size() == 1024);
  } catch (const Panic& e) {
    CHECK(e.code == FP_ENODEV);
  }
  std::printf("ok\n");
  return 0;
}
