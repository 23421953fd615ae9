// Host-side topology construction.  See mesh_host.hpp.
//
// Strut de-duplication: the reference walks the triangle list and inserts each edge key
// (min,max) into a hash map the first time it appears (springy_mesh.rs:284-314), so strut ids are
// in order of first appearance.  Here the 3F edge occurrences are sorted by (key, occurrence) and
// each group's first occurrence ranks the strut: the same order, without a hash map, and the
// per-occurrence strut id falls out for the face tables (:316-330).
#include "mesh_host.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <unordered_map>
#include <utility>

namespace fp {

namespace {

inline uint64_t edge_key(uint64_t a, uint64_t b) {
  uint64_t lo = a <= b ? a : b, hi = a <= b ? b : a;
  return (lo << 32) | hi;
}

inline f3 rec_pos(const std::vector<float>& rec, uint64_t i) {
  return f3{rec[10 * i + 1], rec[10 * i + 2], rec[10 * i + 3]};
}

// The vertex of face (t0,t1,t2) that is not on strut (d0,d1) (utils.rs:26-40).
bool third_vertex(uint32_t t0, uint32_t t1, uint32_t t2, uint32_t d0, uint32_t d1, uint32_t& out) {
  if (t0 != d0 && t0 != d1) {
    out = t0;
    return (d0 == t1 && d1 == t2) || (d0 == t2 && d1 == t1);
  }
  if (t1 != d0 && t1 != d1) {
    out = t1;
    return (d0 == t0 && d1 == t2) || (d0 == t2 && d1 == t0);
  }
  out = t2;
  return (d0 == t0 && d1 == t1) || (d0 == t1 && d1 == t0);
}

// Strut::new (springy_mesh.rs:74-90): constants scaled by NOMINAL_STRUT_LENGTH (1.0) / length.
inline void push_strut(HostMesh& m, uint32_t a, uint32_t b, float k_nom, float c_nom, float len) {
  m.s_v0.push_back(a);
  m.s_v1.push_back(b);
  m.s_k.push_back(1.0f / len * k_nom);
  m.s_c.push_back(1.0f / len * c_nom);
  m.s_kn.push_back(k_nom);
  m.s_cn.push_back(c_nom);
  m.s_len.push_back(len);
  m.s_f0.push_back(-1);
  m.s_f1.push_back(-1);
  m.t_has.push_back(0);
  m.t_k.push_back(0.f);
  m.t_c.push_back(0.f);
  m.t_rest.push_back(0.f);
}

}  // namespace

Status build_springy_mesh(const float* positions, uint64_t n, const uint64_t* indices, uint64_t n_idx, float mass,
                          float k_default, float c_default, const fp_torsion_config* torsion,
                          const fp_strut_override* overrides, uint64_t n_overrides, HostMesh& m) {
  if (n > 0 && !positions) return Status::err(FP_EINVAL, "positions is NULL");
  if (n_idx > 0 && !indices) return Status::err(FP_EINVAL, "indices is NULL");
  if (n_overrides > 0 && !overrides) return Status::err(FP_EINVAL, "overrides is NULL");
  if (n_idx % 3 != 0) return Status::err(FP_EINVAL, "index count is not a multiple of 3");
  if (n >= (1ull << 32)) return Status::err(FP_EINVAL, "more than 2^32-1 vertices in one mesh");
  for (uint64_t i = 0; i < n_idx; ++i)
    if (indices[i] >= n) return Status::err(FP_EINVAL, "vertex index out of range");
  m = HostMesh{};
  m.n = n;
  const float mass_per_vert = mass / (float)n;  // springy_mesh.rs:277
  m.rec.assign(n * 10, 0.0f);
  for (uint64_t i = 0; i < n; ++i) {
    m.rec[10 * i] = mass_per_vert;
    m.rec[10 * i + 1] = positions[3 * i];
    m.rec[10 * i + 2] = positions[3 * i + 1];
    m.rec[10 * i + 3] = positions[3 * i + 2];
  }

  // Overrides: HashMap<StrutKey, SpringConfig>; a repeated key keeps the last value (insert).
  std::unordered_map<uint64_t, std::pair<float, float>> ovr;
  ovr.reserve(n_overrides * 2 + 1);
  for (uint64_t i = 0; i < n_overrides; ++i)
    ovr[edge_key(overrides[i].a, overrides[i].b)] = {overrides[i].constant, overrides[i].damping};

  const uint64_t n_occ = n_idx;  // one edge occurrence per corner: (v0v1), (v1v2), (v2v0)
  std::vector<uint64_t> key(n_occ);
  for (uint64_t t = 0; t < n_idx; t += 3) {
    key[t] = edge_key(indices[t], indices[t + 1]);
    key[t + 1] = edge_key(indices[t + 1], indices[t + 2]);
    key[t + 2] = edge_key(indices[t + 2], indices[t]);
  }
  std::vector<uint32_t> occ(n_occ);
  std::iota(occ.begin(), occ.end(), 0u);
  std::sort(occ.begin(), occ.end(), [&](uint32_t a, uint32_t b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
  // group heads (first occurrence of each key)
  std::vector<uint32_t> head_of(n_occ);
  std::vector<uint32_t> heads;
  for (uint64_t i = 0; i < n_occ; ++i) {
    if (i == 0 || key[occ[i]] != key[occ[i - 1]]) heads.push_back(occ[i]);
    head_of[occ[i]] = heads.back();
  }
  std::sort(heads.begin(), heads.end());  // strut id = rank of the first occurrence
  std::vector<uint32_t> strut_of_head_pos;  // map first occurrence -> strut id via binary search
  const uint64_t S = heads.size();
  m.s_v0.reserve(S);
  for (uint64_t s = 0; s < S; ++s) {
    uint64_t kk = key[heads[s]];
    uint32_t lo = (uint32_t)(kk >> 32), hi = (uint32_t)(kk & 0xffffffffu);
    float ks = k_default, cs = c_default;
    auto it = ovr.find(kk);
    if (it != ovr.end()) {
      ks = it->second.first;
      cs = it->second.second;
    }
    // length = |p[key.0] - p[key.1]| (springy_mesh.rs:303-304)
    push_strut(m, lo, hi, ks, cs, h_magnitude(sub(rec_pos(m.rec, lo), rec_pos(m.rec, hi))));
  }
  auto strut_id = [&](uint32_t o) -> uint32_t {
    uint32_t h = head_of[o];
    return (uint32_t)(std::lower_bound(heads.begin(), heads.end(), h) - heads.begin());
  };

  // Faces and strut -> face adjacency, in face order (springy_mesh.rs:316-345).
  const uint64_t F = n_idx / 3;
  m.f_v.resize(3 * F);
  m.f_s.resize(3 * F);
  for (uint64_t t = 0; t < F; ++t) {
    for (int j = 0; j < 3; ++j) {
      m.f_v[3 * t + j] = (uint32_t)indices[3 * t + j];
      m.f_s[3 * t + j] = strut_id((uint32_t)(3 * t + j));
    }
    for (int j = 0; j < 3; ++j) {
      uint32_t s = m.f_s[3 * t + j];
      if (m.s_f0[s] < 0) m.s_f0[s] = (int64_t)t;
      else if (m.s_f1[s] < 0) m.s_f1[s] = (int64_t)t;
      else return Status::err(FP_ENONMANIFOLD, "A strut shouldn't have more than two adjacent faces!");
    }
  }

  // Torsional springs across interior edges (springy_mesh.rs:348-393).
  if (torsion) {
    for (uint64_t s = 0; s < S; ++s) {
      if (m.s_f0[s] < 0 || m.s_f1[s] < 0) continue;
      uint32_t a = m.s_v0[s], b = m.s_v1[s];
      f3 x0 = rec_pos(m.rec, a), x1 = rec_pos(m.rec, b);
      f3 h = h_normalize(sub(x1, x0));
      const uint32_t* f1 = &m.f_v[3 * m.s_f0[s]];
      const uint32_t* f2 = &m.f_v[3 * m.s_f1[s]];
      uint32_t i2, i3;
      if (!third_vertex(f1[0], f1[1], f1[2], a, b, i2) || !third_vertex(f2[0], f2[1], f2[2], a, b, i3))
        return Status::err(FP_EINVAL, "strut is not an edge of its adjacent face");
      f3 x2 = rec_pos(m.rec, i2), x3 = rec_pos(m.rec, i3);
      f3 nl = h_normalize(cross(sub(x2, x0), sub(x1, x0)));
      f3 nr = h_normalize(cross(sub(x1, x0), sub(x3, x0)));
      m.t_has[s] = 1;
      m.t_k[s] = torsion->spring_constant;
      m.t_c[s] = torsion->spring_damping;
      m.t_rest[s] = atan2f(dot(cross(nl, nr), h), dot(nl, nr));
    }
  }
  return Status::ok();
}

Status mesh_add_strut(HostMesh& m, uint64_t a, uint64_t b, float k, float c) {
  if (a >= m.n || b >= m.n) return Status::err(FP_EINVAL, "add_strut: vertex index out of range");
  // length = |p[v0] - p[v1]| (springy_mesh.rs:403-413)
  push_strut(m, (uint32_t)a, (uint32_t)b, k, c, h_magnitude(sub(rec_pos(m.rec, a), rec_pos(m.rec, b))));
  m.cloth.valid = false;
  return Status::ok();
}

Status build_cloth(uint64_t rows, uint64_t cols, float spacing, const float position[3], float point_mass,
                   float tk, float tc, float sk, float sc, float bk, float bc, const uint64_t* pins,
                   uint64_t n_pins, int32_t plane, HostMesh& m) {
  if (rows < 2 || cols < 2) return Status::err(FP_EINVAL, "cloth needs at least 2 x 2 vertices");
  if (plane != FP_PLANE_XY && plane != FP_PLANE_XZ) return Status::err(FP_EINVAL, "unknown plane");
  if (n_pins > 0 && !pins) return Status::err(FP_EINVAL, "pins is NULL");
  if (!position) return Status::err(FP_EINVAL, "position is NULL");
  if (rows * cols >= (1ull << 32)) return Status::err(FP_EINVAL, "cloth too large");
  // Grid in the reference frame (springy/cloth.rs:34-55): x = col*s, y = row*s, z = 0, minus the
  // centre (cols/2*s, rows/2*s, 0), plus position.  For the xz plane the same construction runs on
  // (p.x, p.z, p.y) and each vertex is permuted back as (x, z, y).
  const float p0 = position[0];
  const float p1 = plane == FP_PLANE_XZ ? position[2] : position[1];
  const float p2 = plane == FP_PLANE_XZ ? position[1] : position[2];
  const float cx = (float)cols / 2.0f * spacing;
  const float cy = (float)rows / 2.0f * spacing;
  std::vector<float> pos(rows * cols * 3);
  for (uint64_t r = 0; r < rows; ++r)
    for (uint64_t c = 0; c < cols; ++c) {
      float x = ((float)c * spacing - cx) + p0;
      float y = ((float)r * spacing - cy) + p1;
      float z = (0.0f - 0.0f) + p2;
      float* o = &pos[3 * (r * cols + c)];
      o[0] = x;
      o[1] = plane == FP_PLANE_XZ ? z : y;
      o[2] = plane == FP_PLANE_XZ ? y : z;
    }
  std::vector<uint64_t> idx;
  idx.reserve(6 * (rows - 1) * (cols - 1));
  std::vector<fp_strut_override> ovr;
  ovr.reserve((rows - 1) * (cols - 1));
  for (uint64_t r = 0; r + 1 < rows; ++r)  // top-left triangles (cloth.rs:63-74)
    for (uint64_t c = 0; c + 1 < cols; ++c) {
      uint64_t i = r * cols + c;
      idx.push_back(i + 1);
      idx.push_back(i + cols);
      idx.push_back(i);
      ovr.push_back(fp_strut_override{i + cols, i + 1, sk, sc});
    }
  for (uint64_t r = 1; r < rows; ++r)  // bottom-right triangles (cloth.rs:75-83)
    for (uint64_t c = 0; c + 1 < cols; ++c) {
      uint64_t i = r * cols + c;
      idx.push_back(i - cols + 1);
      idx.push_back(i + 1);
      idx.push_back(i);
    }
  Status st = build_springy_mesh(pos.data(), rows * cols, idx.data(), idx.size(),
                                 point_mass * (float)rows * (float)cols, tk, tc, nullptr, ovr.data(), ovr.size(), m);
  if (st.code != FP_OK) return st;
  for (uint64_t i = 0; i < n_pins; ++i) m.pins.push_back(pins[i]);  // cloth.rs:95-97
  if (rows >= 3 && cols >= 3)                                       // cloth.rs:99-110
    for (uint64_t r = 0; r + 2 < rows; ++r)
      for (uint64_t c = 0; c + 2 < cols; ++c) {
        uint64_t i = r * cols + c;
        mesh_add_strut(m, i, i + 2, bk, bc);
        mesh_add_strut(m, i, i + 2 * cols, bk, bc);
      }
  ClothDesc& d = m.cloth;
  d.valid = true;
  d.rows = (uint32_t)rows;
  d.cols = (uint32_t)cols;
  d.spacing = spacing;
  d.position[0] = position[0];
  d.position[1] = position[1];
  d.position[2] = position[2];
  d.plane = plane;
  d.tk = tk; d.tc = tc; d.sk = sk; d.sc = sc; d.bk = bk; d.bc = bc;
  return Status::ok();
}

Status build_obstacle(const float* positions, uint64_t n, const uint64_t* indices, uint64_t n_idx, HostObstacle& o) {
  if (n > 0 && !positions) return Status::err(FP_EINVAL, "positions is NULL");
  if (n_idx > 0 && !indices) return Status::err(FP_EINVAL, "indices is NULL");
  for (uint64_t i = 0; i < n_idx; ++i)
    if (indices[i] >= n) return Status::err(FP_EINVAL, "obstacle vertex index out of range");
  o = HostObstacle{};
  o.vertices.assign(positions, positions + 3 * n);
  std::vector<uint64_t> edges;
  const uint64_t F = n_idx / 3;  // trailing indices that do not make a triangle are ignored (tuples())
  for (uint64_t t = 0; t < F; ++t) {
    uint64_t a = indices[3 * t], b = indices[3 * t + 1], c = indices[3 * t + 2];
    // A repeated corner makes a one-element edge set, which the reference indexes out of bounds
    // (collidable_mesh.rs:83-86).
    if (a == b || b == c || c == a) return Status::err(FP_EINVAL, "obstacle triangle repeats a vertex");
    edges.push_back(edge_key(a, b));
    edges.push_back(edge_key(b, c));
    edges.push_back(edge_key(c, a));
    for (uint64_t v : {a, b, c})
      for (int k = 0; k < 3; ++k) o.faces.push_back(positions[3 * v + k]);
  }
  std::sort(edges.begin(), edges.end());
  o.n_edges = (uint64_t)(std::unique(edges.begin(), edges.end()) - edges.begin());
  return Status::ok();
}

}  // namespace fp
