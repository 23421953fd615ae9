// CPU ORACLE — test infrastructure only.
//
// A restatement, in C++, of the reference crate's spring-mass-damper ("springy") step, used by
// the test-suite and by bench.py's cpu_baseline leg as the checker. Nothing in the product library
// (feriphys_amd/) includes, links or calls this file.
//
// Each function follows the reference file:line it cites, in the reference's own operation order:
// Rust evaluates left to right and never contracts a*b+c into an FMA, so this file is compiled with
// -ffp-contract=off and without fast-math, and every literal is typed as the scalar R so that the
// float instantiation never widens to double.  cgmath 0.18 vector semantics (not vendored in the
// reference) are restated as:
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)   = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   magnitude(v) = sqrt(dot(v,v));   normalize(v) = v * (1 / magnitude(v))
//   angle(a,b)   = atan2(magnitude(cross(a,b)), dot(a,b))
//
// Instantiated for R = float (the reference's arithmetic) and R = double (the "f64 run" that
// reports the rounding error of both the oracle and the GPU path).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <map>
#include <set>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------------------------
// cgmath Vector3 semantics (cgmath 0.18.0, Cargo.lock:209-212).
template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> v) { return {s * v.x, s * v.y, s * v.z}; }
template <class R> inline V3<R> operator*(V3<R> v, R s) { return {v.x * s, v.y * s, v.z * s}; }
template <class R> inline V3<R> operator/(V3<R> v, R s) { return {v.x / s, v.y / s, v.z / s}; }
template <class R> inline void operator+=(V3<R>& a, V3<R> b) { a.x = a.x + b.x; a.y = a.y + b.y; a.z = a.z + b.z; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
template <class R> inline R magnitude(V3<R> v) { return std::sqrt(dot(v, v)); }
template <class R> inline V3<R> normalize(V3<R> v) { return v * (R(1) / magnitude(v)); }
template <class R> inline R angle(V3<R> a, V3<R> b) { return std::atan2(magnitude(cross(a, b)), dot(a, b)); }
template <class R> inline bool is_zero(V3<R> v) { return v.x == R(0) && v.y == R(0) && v.z == R(0); }
template <class R> inline V3<R> zero3() { return {R(0), R(0), R(0)}; }
// Rust f32::is_sign_positive: the sign bit is clear (also for NaN).
template <class R> inline bool is_sign_positive(R v) { return !std::signbit(v); }
// Rust f32::min: returns the non-NaN argument when one is NaN.
template <class R> inline R rust_min(R a, R b) { return std::fmin(a, b); }

// std::f32::consts::PI / std::f64::consts::PI
template <class R> inline R pi() { return R(3.14159265358979323846264338327950288); }

// ---------------------------------------------------------------------------------------------
// Duration::from_secs_f32(dt).as_secs_f32() (springy_mesh.rs:442, collidable_mesh.rs:114,129).
// Rust's float->Duration conversion is exact and rounds half-to-even to whole nanoseconds;
// as_secs_f32 is (secs as f32) + (nanos as f32) / 1e9.  The f64 instantiation uses
// from_secs_f64/as_secs_f64, the same rule at double width.
template <class R> R duration_roundtrip(R dt);
template <> inline float duration_roundtrip<float>(float dt) {
  if (!(dt >= 0.0f) || !std::isfinite(dt)) throw std::invalid_argument("Duration::from_secs_f32: negative or non-finite");
  // dt = mant * 2^e exactly.
  int e = 0;
  float fr = std::frexp(dt, &e);  // dt = fr * 2^e, fr in [0.5,1)
  uint64_t mant = (uint64_t)std::ldexp(fr, 24);  // exact, 24-bit
  int shift = 24 - e;                              // dt = mant / 2^shift
  unsigned __int128 ns_num = (unsigned __int128)mant * 1000000000ull;
  unsigned __int128 q, r;
  if (shift <= 0) { q = ns_num << (-shift); r = 0; }
  else if (shift >= 120) { q = 0; r = ns_num; }
  else { q = ns_num >> shift; r = ns_num & ((((unsigned __int128)1) << shift) - 1); }
  if (shift > 0 && shift < 120) {
    unsigned __int128 half = ((unsigned __int128)1) << (shift - 1);
    if (r > half || (r == half && (q & 1))) q += 1;
  }
  uint64_t total_ns = (uint64_t)q;
  uint64_t secs = total_ns / 1000000000ull;
  uint32_t nanos = (uint32_t)(total_ns % 1000000000ull);
  return (float)secs + (float)nanos / 1000000000.0f;
}
template <> inline double duration_roundtrip<double>(double dt) {
  if (!(dt >= 0.0) || !std::isfinite(dt)) throw std::invalid_argument("Duration::from_secs_f64: negative or non-finite");
  int e = 0;
  double fr = std::frexp(dt, &e);
  uint64_t mant = (uint64_t)std::ldexp(fr, 53);
  int shift = 53 - e;
  unsigned __int128 ns_num = (unsigned __int128)mant * 1000000000ull;
  unsigned __int128 q, r;
  if (shift <= 0) { q = ns_num << (-shift); r = 0; }
  else if (shift >= 120) { q = 0; r = ns_num; }
  else { q = ns_num >> shift; r = ns_num & ((((unsigned __int128)1) << shift) - 1); }
  if (shift > 0 && shift < 120) {
    unsigned __int128 half = ((unsigned __int128)1) << (shift - 1);
    if (r > half || (r == half && (q & 1))) q += 1;
  }
  uint64_t total_ns = (uint64_t)q;
  uint64_t secs = total_ns / 1000000000ull;
  uint32_t nanos = (uint32_t)(total_ns % 1000000000ull);
  return (double)secs + (double)nanos / 1000000000.0;
}

// ---------------------------------------------------------------------------------------------
// utils.rs:5-21 — vec_add and scale over flat state vectors.
template <class R> std::vector<R> vec_add(const std::vector<R>& a, const std::vector<R>& b) {
  if (a.size() != b.size()) throw std::invalid_argument("vec_add: length mismatch");  // utils.rs:9-11
  std::vector<R> out(a.size());
  for (size_t i = 0; i < a.size(); ++i) out[i] = a[i] + b[i];
  return out;
}
template <class R> std::vector<R> scale(const std::vector<R>& v, R s) {
  std::vector<R> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i] * s;
  return out;
}
// utils.rs:26-40 — the member of `trio` missing from `duo` (torsion only).
template <class T> T tuple_difference(T t0, T t1, T t2, T d0, T d1) {
  if (t0 != d0 && t0 != d1) {
    if (!((d0 == t1 && d1 == t2) || (d0 == t2 && d1 == t1))) throw std::logic_error("tuple_difference");
    return t0;
  } else if (t1 != d0 && t1 != d1) {
    if (!((d0 == t0 && d1 == t2) || (d0 == t2 && d1 == t0))) throw std::logic_error("tuple_difference");
    return t1;
  } else {
    if (!((d0 == t0 && d1 == t1) || (d0 == t1 && d1 == t0))) throw std::logic_error("tuple_difference");
    return t2;
  }
}

// ---------------------------------------------------------------------------------------------
// state.rs:10-112 — the Stateful contract and the generic explicit integrators.
// E provides: static size_t num_state_elements(); static E from_state_vector(const std::vector<R>&);
//             std::vector<R> derivative() const; std::vector<R> as_state() const.
template <class R, class E> struct State {
  std::vector<E> elements;

  explicit State(std::vector<E> e) : elements(std::move(e)) {}  // state.rs:42-44

  static State from_state_vector(const std::vector<R>& sv) {  // state.rs:46-54
    const size_t w = E::num_state_elements();
    std::vector<E> els;
    els.reserve(sv.size() / w + 1);
    for (size_t i = 0; i < sv.size(); i += w) {
      size_t n = std::min(w, sv.size() - i);
      els.push_back(E::from_state_vector(std::vector<R>(sv.begin() + i, sv.begin() + i + n)));
    }
    return State(std::move(els));
  }
  std::vector<R> derivative() const {  // state.rs:56-62
    std::vector<R> out;
    for (const auto& e : elements) { auto d = e.derivative(); out.insert(out.end(), d.begin(), d.end()); }
    return out;
  }
  std::vector<R> as_vector() const {  // state.rs:64-70
    std::vector<R> out;
    for (const auto& e : elements) { auto d = e.as_state(); out.insert(out.end(), d.begin(), d.end()); }
    return out;
  }
  State euler_step(R h) const {  // state.rs:75-83
    auto d = derivative();
    for (auto& x : d) x = x * h;
    return from_state_vector(vec_add(as_vector(), d));
  }
  State rk4_step(R h) const {  // state.rs:86-106
    auto k1 = derivative();
    auto half_k1 = scale(k1, h * R(0.5));
    auto k2 = from_state_vector(vec_add(as_vector(), half_k1)).derivative();
    auto half_k2 = scale(k2, h * R(0.5));
    auto k3 = from_state_vector(vec_add(as_vector(), half_k2)).derivative();
    auto k3d = scale(k3, h);
    auto k4 = from_state_vector(vec_add(as_vector(), k3d)).derivative();
    std::vector<R> delta(k1.size());
    for (size_t i = 0; i < k1.size(); ++i)
      delta[i] = h / R(6) * k1[i] + h / R(3) * k2[i] + h / R(3) * k3[i] + h / R(6) * k4[i];
    return from_state_vector(vec_add(as_vector(), delta));
  }
  std::vector<E> get_elements() const { return elements; }  // state.rs:110-112
};

enum class Integration { Euler = 0, Rk4 = 1 };  // state.rs:5-8

// springy/config.rs:9-33
template <class R> struct Config {
  Integration integration = Integration::Rk4;
  R dt = R(0.001);
  V3<R> gravity = {R(0), R(-10), R(0)};
  V3<R> wind = {R(0), R(0), R(0)};
  R lift_coefficient = R(1);
  R drag_coefficient = R(1);
  R coefficient_of_restitution = R(0.95);
  R coefficient_of_friction = R(0.3);
};

// consts.rs:2
template <class R> inline R epsilon() { return R(0.0001f); }

// ---------------------------------------------------------------------------------------------
// springy_mesh.rs:177-256 — Point and its Stateful record [m, x,y,z, vx,vy,vz, Fx,Fy,Fz].
template <class R> struct Point {
  R mass;
  V3<R> position, velocity, accumulated_force;

  static size_t num_state_elements() { return 10; }
  static Point from_state_vector(const std::vector<R>& s) {  // :204-221
    if (s.size() != 10) throw std::invalid_argument("State Vector incorrect size!");
    return Point{s[0], {s[1], s[2], s[3]}, {s[4], s[5], s[6]}, {s[7], s[8], s[9]}};
  }
  std::vector<R> derivative() const {  // :223-240
    return {R(0), velocity.x, velocity.y, velocity.z,
            accumulated_force.x / mass, accumulated_force.y / mass, accumulated_force.z / mass,
            R(0), R(0), R(0)};
  }
  std::vector<R> as_state() const {  // :242-255
    return {mass, position.x, position.y, position.z, velocity.x, velocity.y, velocity.z,
            accumulated_force.x, accumulated_force.y, accumulated_force.z};
  }
};

// springy_mesh.rs:26-31, 57-91
template <class R> struct TorsionalSpring {
  R spring_constant, damping, rest_angle;
};
template <class R> struct Strut {
  R stiffness, damping, length;
  size_t v0, v1;
  long face0 = -1, face1 = -1;  // Option<usize>
  bool has_torsion = false;
  TorsionalSpring<R> torsion{};
};
// springy_mesh.rs:74-90: k = (NOMINAL_STRUT_LENGTH / length) * nominal, NOMINAL_STRUT_LENGTH = 1.
template <class R> Strut<R> make_strut(R k_nom, R c_nom, R length, size_t v0, size_t v1) {
  Strut<R> s;
  s.stiffness = R(1) / length * k_nom;
  s.damping = R(1) / length * c_nom;
  s.length = length;
  s.v0 = v0;
  s.v1 = v1;
  return s;
}

template <class R> struct Face {
  size_t s0, s1, s2;  // strut indices
  size_t v0, v1, v2;  // vertex indices
};

template <class R> struct SpringOverride {
  size_t a, b;  // StrutKey vertex pair (any order; keyed on (min,max))
  R constant, damping;
};

// ---------------------------------------------------------------------------------------------
// collidable_mesh.rs:40-47, 57-169
template <class R> struct ObstacleFace {
  V3<R> v0, v1, v2;
  V3<R> normal() const { return normalize(cross(v1 - v0, v2 - v0)); }           // :41-43
  R distance_from_plane(V3<R> p) const { return dot(p - v0, normal()); }      // :45-47
};

template <class R> struct CollidableMesh {
  std::vector<V3<R>> vertices;
  std::vector<std::pair<V3<R>, V3<R>>> edges;
  std::vector<ObstacleFace<R>> faces;

  // :57-105 — edges are the BTreeSet of vertex-index pairs (ordered), faces keep index order.
  CollidableMesh(const std::vector<V3<R>>& pos, const std::vector<size_t>& idx) {
    vertices = pos;
    std::set<std::pair<size_t, size_t>> es;
    for (size_t t = 0; t + 2 < idx.size(); t += 3) {
      size_t a = idx[t], b = idx[t + 1], c = idx[t + 2];
      if (a >= pos.size() || b >= pos.size() || c >= pos.size()) throw std::out_of_range("obstacle index");
      es.insert({std::min(a, b), std::max(a, b)});
      es.insert({std::min(b, c), std::max(b, c)});
      es.insert({std::min(c, a), std::max(c, a)});
    }
    for (auto& e : es) edges.push_back({pos[e.first], pos[e.second]});
    for (size_t t = 0; t + 2 < idx.size(); t += 3) faces.push_back({pos[idx[t]], pos[idx[t + 1]], pos[idx[t + 2]]});
  }

  // :107-169 — first face (in list order) whose plane the segment crosses and whose flattened
  // triangle contains the crossing point.  dt_d is Duration(dt).as_secs_f32().
  static long get_collided_face_from_list(const std::vector<const ObstacleFace<R>*>& fl, V3<R> old_p,
                                          V3<R> new_p, R dt_d) {
    for (size_t fi = 0; fi < fl.size(); ++fi) {
      const ObstacleFace<R>& f = *fl[fi];
      V3<R> old_velocity = (new_p - old_p) / dt_d;
      R od = f.distance_from_plane(old_p);
      R nd = f.distance_from_plane(new_p);
      bool crossed = is_sign_positive(od) != is_sign_positive(nd);
      if (!crossed) continue;
      R frac = od / (od - nd);
      V3<R> cp = old_p + dt_d * frac * old_velocity;
      V3<R> n = f.normal();
      V3<R> a, b, c, p;
      if (n.x >= n.y && n.x >= n.z) {
        a = {R(0), f.v0.y, f.v0.z}; b = {R(0), f.v1.y, f.v1.z}; c = {R(0), f.v2.y, f.v2.z}; p = {R(0), cp.y, cp.z};
      } else if (n.y >= n.x && n.y >= n.z) {
        a = {f.v0.x, R(0), f.v0.z}; b = {f.v1.x, R(0), f.v1.z}; c = {f.v2.x, R(0), f.v2.z}; p = {cp.x, R(0), cp.z};
      } else {
        a = {f.v0.x, f.v0.y, R(0)}; b = {f.v1.x, f.v1.y, R(0)}; c = {f.v2.x, f.v2.y, R(0)}; p = {cp.x, cp.y, R(0)};
      }
      V3<R> c0 = cross(b - a, p - a);
      V3<R> c1 = cross(c - b, p - b);
      V3<R> c2 = cross(a - c, p - c);
      bool o0 = is_sign_positive(dot(c0, f.normal()));
      bool o1 = is_sign_positive(dot(c1, f.normal()));
      bool o2 = is_sign_positive(dot(c2, f.normal()));
      if (o0 == o1 && o1 == o2) return (long)fi;
    }
    return -1;
  }
};

// ---------------------------------------------------------------------------------------------
// springy_mesh.rs:258-661 — SpringyMesh.
template <class R> struct SpringyMesh {
  std::vector<Strut<R>> struts;
  std::vector<Face<R>> faces;
  std::vector<Point<R>> points;
  std::vector<size_t> pinned_points;

  // :267-401.  `overrides` restates the FxHashMap<StrutKey, SpringConfig>: a later entry for the
  // same key replaces an earlier one, as HashMap::insert does.
  SpringyMesh(const std::vector<V3<R>>& positions, const std::vector<size_t>& indices, R mass, R k_default,
              R c_default, const TorsionalSpring<R>* torsion_cfg /*spring_constant, damping; rest unused*/,
              const std::vector<SpringOverride<R>>* overrides) {
    const R mass_per_vert = mass / (R)positions.size();  // :277
    for (auto& p : positions) points.push_back(Point<R>{mass_per_vert, p, zero3<R>(), zero3<R>()});
    std::unordered_map<uint64_t, std::pair<R, R>> ovr;
    auto key = [](size_t a, size_t b) -> uint64_t {
      size_t lo = a <= b ? a : b, hi = a <= b ? b : a;  // StrutKey::new, :102-106
      return ((uint64_t)lo << 32) | (uint64_t)hi;
    };
    if (overrides)
      for (auto& o : *overrides) ovr[key(o.a, o.b)] = {o.constant, o.damping};
    if (indices.size() % 3 != 0) throw std::invalid_argument("index count not a multiple of 3");
    for (size_t i : indices)
      if (i >= positions.size()) throw std::out_of_range("vertex index");
    std::unordered_map<uint64_t, size_t> strut_indices;
    for (size_t t = 0; t < indices.size(); t += 3) {  // :286-314
      auto add = [&](size_t i1, size_t i2) {
        uint64_t k = key(i1, i2);
        if (strut_indices.find(k) == strut_indices.end()) {
          strut_indices[k] = struts.size();
          R ks = k_default, cs = c_default;
          if (overrides) {
            auto it = ovr.find(k);
            if (it != ovr.end()) { ks = it->second.first; cs = it->second.second; }
          }
          size_t lo = (size_t)(k >> 32), hi = (size_t)(k & 0xffffffffu);
          struts.push_back(make_strut<R>(ks, cs, magnitude(points[lo].position - points[hi].position), lo, hi));
        }
      };
      add(indices[t], indices[t + 1]);
      add(indices[t + 1], indices[t + 2]);
      add(indices[t + 2], indices[t]);
    }
    for (size_t t = 0; t < indices.size(); t += 3) {  // :316-345
      size_t a = indices[t], b = indices[t + 1], c = indices[t + 2];
      size_t s0 = strut_indices.at(key(a, b)), s1 = strut_indices.at(key(b, c)), s2 = strut_indices.at(key(c, a));
      size_t fi = faces.size();
      faces.push_back(Face<R>{s0, s1, s2, a, b, c});
      for (size_t si : {s0, s1, s2}) {
        Strut<R>& s = struts[si];
        if (s.face0 < 0) s.face0 = (long)fi;
        else if (s.face1 < 0) s.face1 = (long)fi;
        else throw std::domain_error("A strut shouldn't have more than two adjacent faces!");  // :339
      }
    }
    if (torsion_cfg) {  // :348-393
      for (size_t si = 0; si < struts.size(); ++si) {
        Strut<R>& s = struts[si];
        if (s.face0 < 0 || s.face1 < 0) continue;
        const V3<R> x0 = points[s.v0].position, x1 = points[s.v1].position;
        V3<R> h = normalize(x1 - x0);
        const Face<R>& f1 = faces[(size_t)s.face0];
        const Face<R>& f2 = faces[(size_t)s.face1];
        size_t i2 = tuple_difference(f1.v0, f1.v1, f1.v2, s.v0, s.v1);
        size_t i3 = tuple_difference(f2.v0, f2.v1, f2.v2, s.v0, s.v1);
        V3<R> x2 = points[i2].position, x3 = points[i3].position;
        V3<R> nl = normalize(cross(x2 - x0, x1 - x0));
        V3<R> nr = normalize(cross(x1 - x0, x3 - x0));
        R theta = std::atan2(dot(cross(nl, nr), h), dot(nl, nr));
        s.has_torsion = true;
        s.torsion = TorsionalSpring<R>{torsion_cfg->spring_constant, torsion_cfg->damping, theta};
      }
    }
  }

  void add_strut(size_t a, size_t b, R k, R c) {  // :403-413
    if (a >= points.size() || b >= points.size()) throw std::out_of_range("add_strut index");
    struts.push_back(make_strut<R>(k, c, magnitude(points[a].position - points[b].position), a, b));
  }
  void add_pin(size_t i) {  // :415-417 (the index is checked when used, :495/:523)
    pinned_points.push_back(i);
  }

  // Face helpers :118-172
  V3<R> face_normal(const Face<R>& f) const {
    return normalize(cross(points[f.v1].position - points[f.v0].position, points[f.v2].position - points[f.v0].position));
  }
  R face_area(const Face<R>& f) const {
    return magnitude(cross(points[f.v1].position - points[f.v0].position, points[f.v2].position - points[f.v0].position)) / R(2);
  }
  R vertex_angle(size_t i0, size_t i1, size_t i2) const {
    return angle(points[i2].position - points[i0].position, points[i1].position - points[i0].position);
  }
  R vertex_angle_0(const Face<R>& f) const { return vertex_angle(f.v0, f.v1, f.v2); }
  R vertex_angle_1(const Face<R>& f) const { return vertex_angle(f.v1, f.v2, f.v0); }
  R vertex_angle_2(const Face<R>& f) const { return vertex_angle(f.v2, f.v0, f.v1); }

  // :423-497
  void update_points(std::vector<Point<R>> new_points, const std::vector<const CollidableMesh<R>*>& obstacles,
                     const Config<R>& cfg) {
    std::vector<const ObstacleFace<R>*> fl;
    for (auto* o : obstacles)
      for (auto& f : o->faces) fl.push_back(&f);
    const R dt_d = duration_roundtrip<R>(cfg.dt);
    const size_t n = std::min(new_points.size(), points.size());  // zip
    for (size_t i = 0; i < n; ++i) {
      Point<R>& np = new_points[i];
      const Point<R>& op = points[i];
      long hit = CollidableMesh<R>::get_collided_face_from_list(fl, op.position, np.position, dt_d);
      if (hit < 0) continue;
      const ObstacleFace<R>& f = *fl[(size_t)hit];
      R od = f.distance_from_plane(op.position);
      R nd = f.distance_from_plane(np.position);
      R frac = od / (od - nd);
      V3<R> cp = op.position + cfg.dt * frac * op.velocity;
      V3<R> vc = op.velocity + cfg.dt * frac * op.accumulated_force / op.mass;
      V3<R> newp = cp + f.normal() * epsilon<R>();
      V3<R> vcn = dot(vc, f.normal()) * f.normal();
      V3<R> vct = vc - vcn;
      V3<R> vrn = R(-1) * vcn * cfg.coefficient_of_restitution;
      V3<R> vrt;
      if (is_zero(vct) || std::isnan(magnitude(vct))) {
        vrt = zero3<R>();
      } else {
        vrt = vct - normalize(vct) * rust_min(cfg.coefficient_of_friction * magnitude(vcn), magnitude(vct));
      }
      np.position = newp;
      np.velocity = vrn + vrt;
    }
    std::vector<Point<R>> original = points;  // :492-496
    points = std::move(new_points);
    for (size_t pi : pinned_points) {
      if (pi >= points.size() || pi >= original.size()) throw std::out_of_range("pin index");
      points[pi] = original[pi];
    }
  }

  // :515-525
  // The reference comments out apply_torsional_forces at :518-519; torsion_in_step restores that
  // call in its place (between the strut and the face forces).
  bool torsion_in_step = false;

  void accumulate_forces(const Config<R>& cfg) {
    for (auto& p : points) p.accumulated_force += cfg.gravity;  // :527-531, :194-196
    apply_strut_forces();
    if (torsion_in_step) apply_torsional_forces();
    apply_face_forces(cfg);
    for (size_t pi : pinned_points) {
      if (pi >= points.size()) throw std::out_of_range("pin index");
      points[pi].accumulated_force = zero3<R>();
    }
  }

  // :533-551
  void apply_strut_forces() {
    for (const Strut<R>& s : struts) {
      V3<R> p0 = points[s.v0].position, p1 = points[s.v1].position;
      V3<R> u = normalize(p1 - p0);
      V3<R> fs0 = s.stiffness * (magnitude(p1 - p0) - s.length) * u;
      points[s.v0].accumulated_force += fs0;
      V3<R> fs1 = R(-1) * fs0;
      points[s.v1].accumulated_force += fs1;
      V3<R> v0 = points[s.v0].velocity, v1 = points[s.v1].velocity;
      V3<R> fd0 = s.damping * dot(v1 - v0, u) * u;
      points[s.v0].accumulated_force += fd0;
      V3<R> fd1 = R(-1) * fd0;
      points[s.v1].accumulated_force += fd1;
    }
  }

  // :553-627 (disabled in accumulate_forces at :518-519; kept for the hinge test and as the
  // next component on the path).
  void apply_torsional_forces() {
    // FxHashMap<usize, Vector3>: each vertex's entry sums in strut order; entries are then added to
    // distinct points, so the map's iteration order does not affect any result.
    std::map<size_t, V3<R>> vf;
    for (const Strut<R>& s : struts) {
      if (s.face0 < 0 || s.face1 < 0) continue;
      const Point<R>& x0 = points[s.v0];
      const Point<R>& x1 = points[s.v1];
      const Face<R>& f1 = faces[(size_t)s.face0];
      const Face<R>& f2 = faces[(size_t)s.face1];
      size_t i2 = tuple_difference(f1.v0, f1.v1, f1.v2, s.v0, s.v1);
      size_t i3 = tuple_difference(f2.v0, f2.v1, f2.v2, s.v0, s.v1);
      const Point<R>& x2 = points[i2];
      const Point<R>& x3 = points[i3];
      R l01 = magnitude(x1.position - x0.position);
      V3<R> h = normalize(x1.position - x0.position);
      R d02 = dot(x2.position - x0.position, h);
      R d03 = dot(x3.position - x0.position, h);
      V3<R> rl = (x2.position - x0.position) - d02 * h;
      V3<R> rr = (x3.position - x0.position) - d03 * h;
      V3<R> nl = normalize(cross(x1.position - x0.position, x2.position - x0.position));
      V3<R> nr = normalize(cross(x3.position - x0.position, x1.position - x0.position));
      R theta = std::atan2(dot(cross(nl, nr), h), dot(nl, nr));
      R tl = dot(x2.velocity, nl) / magnitude(rl);
      R tr = dot(x3.velocity, nr) / magnitude(rr);
      V3<R> spring_torque = s.torsion.spring_constant * (theta - s.torsion.rest_angle) * h;
      V3<R> damping_torque = R(-1) * s.torsion.damping * (tl + tr) * h;
      V3<R> torque = spring_torque + damping_torque;
      V3<R> f3 = dot(torque, h) / magnitude(rr) * nr;
      V3<R> f2v = dot(torque, h) / magnitude(rl) * nl;
      V3<R> f1v = (d02 * f2v + d03 * f3) / l01;
      V3<R> f0v = R(-1) * (f1v + f2v + f3);
      auto acc = [&](size_t i, V3<R> f) {
        auto it = vf.find(i);
        if (it == vf.end()) vf[i] = zero3<R>();
        vf[i] += f;
      };
      acc(s.v0, f0v);
      acc(s.v1, f1v);
      acc(i2, f2v);
      acc(i3, f3);
    }
    for (auto& kv : vf) points[kv.first].accumulated_force += kv.second;
  }

  // :629-654
  void apply_face_forces(const Config<R>& cfg) {
    for (const Face<R>& f : faces) {
      const Point<R> p0 = points[f.v0], p1 = points[f.v1], p2 = points[f.v2];
      V3<R> avg = (p0.velocity + p1.velocity + p2.velocity) / R(3);
      V3<R> rv = avg - cfg.wind;
      R eff = face_area(f) * std::fabs(dot(face_normal(f), rv));
      V3<R> drag = R(-1) * cfg.drag_coefficient * eff * rv;
      V3<R> lift = R(-1) * cfg.lift_coefficient * eff * (rv * magnitude(cross(face_normal(f), rv)));
      V3<R> f0 = vertex_angle_0(f) / pi<R>() * (drag + lift);
      V3<R> f1 = vertex_angle_1(f) / pi<R>() * (drag + lift);
      V3<R> f2 = vertex_angle_2(f) / pi<R>() * (drag + lift);
      points[f.v0].accumulated_force += f0;
      points[f.v1].accumulated_force += f1;
      points[f.v2].accumulated_force += f2;
    }
  }

  void clear_forces() {  // :656-660
    for (auto& p : points) p.accumulated_force = zero3<R>();
  }
};

// springy/cloth.rs:20-113.  `plane` selects the grid plane: 0 = the reference's xy plane; 1 = xz
// (the grid's row axis along +z, the reference's z offset along y), which is the same constructor
// sequence (SpringyMesh::new, add_pin, add_strut) applied to component-permuted positions.
template <class R> SpringyMesh<R> cloth_new(size_t rows, size_t cols, R spacing, V3<R> position, R point_mass,
                                            R tk, R tc, R sk, R sc, R bk, R bc,
                                            const std::vector<size_t>& pins, int plane = 0) {
  if (rows < 2 || cols < 2) throw std::invalid_argument("cloth needs at least 2x2 vertices");
  V3<R> pos_ref = plane == 1 ? V3<R>{position.x, position.z, position.y} : position;
  std::vector<V3<R>> vp;
  vp.reserve(rows * cols);
  for (size_t r = 0; r < rows; ++r)
    for (size_t c = 0; c < cols; ++c) vp.push_back(V3<R>{(R)c * spacing, (R)r * spacing, R(0)});
  const V3<R> centre{(R)cols / R(2) * spacing, (R)rows / R(2) * spacing, R(0)};
  for (auto& v : vp) {
    v = (v - centre) + pos_ref;
    if (plane == 1) v = V3<R>{v.x, v.z, v.y};
  }
  std::vector<SpringOverride<R>> ovr;
  std::vector<size_t> idx;
  for (size_t r = 0; r + 1 < rows; ++r)
    for (size_t c = 0; c + 1 < cols; ++c) {
      size_t i = r * cols + c;
      idx.push_back(i + 1); idx.push_back(i + cols); idx.push_back(i);
      ovr.push_back(SpringOverride<R>{i + cols, i + 1, sk, sc});
    }
  for (size_t r = 1; r < rows; ++r)
    for (size_t c = 0; c + 1 < cols; ++c) {
      size_t i = r * cols + c;
      idx.push_back(i - cols + 1); idx.push_back(i + 1); idx.push_back(i);
    }
  SpringyMesh<R> m(vp, idx, point_mass * (R)rows * (R)cols, tk, tc, nullptr, &ovr);
  for (size_t p : pins) m.add_pin(p);
  if (rows >= 3 && cols >= 3)
    for (size_t r = 0; r + 2 < rows; ++r)
      for (size_t c = 0; c + 2 < cols; ++c) {
        size_t i = r * cols + c;
        m.add_strut(i, i + 2, bk, bc);
        m.add_strut(i, i + 2 * cols, bk, bc);
      }
  return m;
}

// springy/simulation.rs:18-45 — one Config for all meshes, meshes stepped serially.
template <class R> struct Simulation {
  Config<R> config;
  std::vector<SpringyMesh<R>> meshes;
  std::vector<CollidableMesh<R>> obstacles;

  R step() {
    std::vector<const CollidableMesh<R>*> obs;
    for (auto& o : obstacles) obs.push_back(&o);
    for (auto& mesh : meshes) {
      mesh.accumulate_forces(config);
      State<R, Point<R>> st(mesh.points);
      State<R, Point<R>> ns = config.integration == Integration::Rk4 ? st.rk4_step(config.dt) : st.euler_step(config.dt);
      mesh.update_points(ns.get_elements(), obs, config);
      mesh.clear_forces();
    }
    return duration_roundtrip<R>(config.dt);
  }
};

}  // namespace oracle
