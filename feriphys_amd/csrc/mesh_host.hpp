// Host-side SpringyMesh / Cloth / CollidableMesh construction for the gfx950 step.
//
// Builds the same strut, face, pin and point tables as the reference constructors
// (springy_mesh.rs:267-417, springy/cloth.rs:20-113, collidable_mesh.rs:57-105), so strut and
// face indices match bit for bit, then hands them to the device packer (sim.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/feriphys_gpu.h"
#include "vec.hpp"

namespace fp {

// What the grid kernel needs to rebuild a Cloth::new topology on chip.
struct ClothDesc {
  bool valid = false;  // built by Cloth::new and not extended by add_strut since
  uint32_t rows = 0, cols = 0;
  float spacing = 0.f;
  float position[3] = {0.f, 0.f, 0.f};
  int32_t plane = FP_PLANE_XY;
  float tk = 0.f, tc = 0.f, sk = 0.f, sc = 0.f, bk = 0.f, bc = 0.f;  // nominal constants
};

struct HostMesh {
  uint64_t n = 0;              // points
  std::vector<float> rec;      // n * 10 Stateful records (springy_mesh.rs:242-255)
  // struts in reference index order
  std::vector<uint32_t> s_v0, s_v1;
  std::vector<float> s_k, s_c, s_len;
  std::vector<float> s_kn, s_cn;  // the nominal constants that s_k, s_c were scaled from
  std::vector<int64_t> s_f0, s_f1;  // -1 = None
  std::vector<uint8_t> t_has;       // torsional spring present (springy_mesh.rs:348-393)
  std::vector<float> t_k, t_c, t_rest;
  // faces: 3 vertex and 3 strut indices each
  std::vector<uint32_t> f_v, f_s;
  std::vector<uint64_t> pins;
  ClothDesc cloth;
  bool moved = false;  // true once consumed by fp_sim_new

  uint64_t n_struts() const { return s_v0.size(); }
  uint64_t n_faces() const { return f_v.size() / 3; }
};

struct HostObstacle {
  std::vector<float> vertices;  // n * 3
  uint64_t n_edges = 0;         // size of the BTreeSet of edges (collidable_mesh.rs:63-88)
  std::vector<float> faces;     // per face v0, v1, v2 (9 floats), index order
  bool moved = false;
  uint64_t n_faces() const { return faces.size() / 9; }
};

// Status plus message; the C ABI turns these into fp_last_error().
struct Status {
  int code = FP_OK;
  std::string msg;
  static Status ok() { return Status{}; }
  static Status err(int c, std::string m) { return Status{c, std::move(m)}; }
};

Status build_springy_mesh(const float* positions, uint64_t n, const uint64_t* indices, uint64_t n_idx, float mass,
                          float k_default, float c_default, const fp_torsion_config* torsion,
                          const fp_strut_override* overrides, uint64_t n_overrides, HostMesh& out);

Status build_cloth(uint64_t rows, uint64_t cols, float spacing, const float position[3], float point_mass,
                   float tk, float tc, float sk, float sc, float bk, float bc, const uint64_t* pins,
                   uint64_t n_pins, int32_t plane, HostMesh& out);

Status mesh_add_strut(HostMesh& m, uint64_t a, uint64_t b, float k, float c);

Status build_obstacle(const float* positions, uint64_t n, const uint64_t* indices, uint64_t n_idx, HostObstacle& out);

}  // namespace fp
