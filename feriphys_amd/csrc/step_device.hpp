// Device-side pieces of the springy step shared by the CSR and grid kernels: per-strut and
// per-face force terms, the integrators, the obstacle first-hit scan and the collision response.
//
// Every vertex GATHERS its own force: struts in strut-index order, then faces in face-index order,
// exactly the per-vertex summation order of the reference's scatter loops
// (springy_mesh.rs:533-551, 629-654).  No atomics, so every run is bit-reproducible.
#pragma once

#include "../../include/feriphys_gpu.h"
#include "vec.hpp"

namespace fp {

// Per-mesh step constants (springy/config.rs:9-18 plus values derived on the host).
struct StepParams {
  int32_t integration;  // FP_EULER / FP_RK4
  float dt;
  float dt_half;  // dt * 0.5           (state.rs:88)
  float dt6;      // dt / 6.0           (state.rs:99)
  float dt3;      // dt / 3.0           (state.rs:100)
  float dt_d;     // Duration::from_secs_f32(dt).as_secs_f32()  (springy_mesh.rs:442)
  f3 gravity;
  f3 wind;
  float neg_drag;  // -1.0 * drag_coefficient  (springy_mesh.rs:638)
  float neg_lift;  // -1.0 * lift_coefficient  (:639-640)
  float restitution;
  float friction;
  float mass;        // the mesh's per-vertex mass when uniform
  int32_t uniform_mass;
};

// Obstacle face for the first-hit scan (collidable_mesh.rs:30-47): corners, the unit normal
// (computed once on the host with the reference's operations; Face::normal recomputes the same
// value on every call) and the axis the containment test drops (:133-154).
struct alignas(16) ObsFace {
  float v0[3], v1[3], v2[3], n[3];
  int32_t axis;  // 0: drop x, 1: drop y, 2: drop z
  int32_t pad[3];
};
static_assert(sizeof(ObsFace) == 64, "one 64-byte record per face");

// Face i's corners and normal as three 16-byte loads issued together (the plane test needs the
// normal and the first corner, the containment test the rest: one round trip serves both), and
// its dropped axis recomputed from the normal by the comparisons that set it on the host
// (collidable_mesh.rs:133-154) rather than loaded.
__device__ __forceinline__ ObsFace load_face(const ObsFace* __restrict__ faces, uint32_t i) {
  const float4* q = reinterpret_cast<const float4*>(faces + i);
  const float4 a = q[0], b = q[1], c = q[2];
  ObsFace f;
  f.v0[0] = a.x; f.v0[1] = a.y; f.v0[2] = a.z;
  f.v1[0] = a.w; f.v1[1] = b.x; f.v1[2] = b.y;
  f.v2[0] = b.z; f.v2[1] = b.w; f.v2[2] = c.x;
  f.n[0] = c.y; f.n[1] = c.z; f.n[2] = c.w;
  f.axis = (c.y >= c.z && c.y >= c.w) ? 0 : ((c.z >= c.y && c.z >= c.w) ? 1 : 2);
  f.pad[0] = f.pad[1] = f.pad[2] = 0;
  return f;
}

// Broad phase over the obstacle faces: uniform grids of face ids, each cell listing (in ascending
// id order) every face whose bounding box, grown by `margin`, touches the cell.
//   grid 0   : 3D grid of faces whose dropped normal component is at least 0.02.  A hit needs the
//              crossing point in the triangle, and the point lies on the segment, so only faces
//              whose grown box meets the segment's grown box can be hit (the margin covers the
//              crossing point's rounding divided by the dropped component).
//   grid 1+a : 2D grid (axis a collapsed) of faces that drop axis a with a component in
//              [1e-3, 0.02).  Their flattened containment test is thin but still confines the
//              crossing point to the triangle's box in the two kept axes; along the dropped axis
//              the plane may be met anywhere, so that axis is not binned.
//   always   : faces with a smaller (or non-finite) dropped component, tested for every vertex.
struct GridView {
  float lo[3];
  float inv_cell[3];  // 0 on a collapsed axis: every coordinate maps to cell 0
  int32_t dims[3];
  const uint32_t* cell_start;  // prod(dims)+1 (nullptr: empty grid)
  const uint32_t* cell_faces;
};

struct ObsView {
  const ObsFace* faces;
  uint32_t n_faces;
  uint32_t use_grid;  // 0: scan every face in order
  float margin;
  GridView grid[4];
  const uint32_t* always;  // ascending
  uint32_t n_always;
};

__device__ __forceinline__ f3 ld3(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                                  uint32_t i) {
  return f3{x[i], y[i], z[i]};
}

// One strut's spring + damping force on the vertex at p_self (springy_mesh.rs:535-549).
// The reference orients the strut (v0, v1) and gives v1 exactly -1 * the v0 force; computing with
// the gathering vertex as v0 yields that same value bit for bit, because p_a - p_b = -(p_b - p_a)
// and every later product carries the sign through exactly.
template <class M>
__device__ __forceinline__ void strut_accumulate(f3& F, f3 p_self, f3 v_self, f3 p_nbr, f3 v_nbr, float k, float c,
                                                 float rest) {
  f3 d = sub(p_nbr, p_self);
  // |d| - rest cancels to the strain, so |d| is always formed exactly (unfused dot, IEEE sqrt):
  // one ulp of |d| is k * ulp(L), already ~5e-5 of a typical strut force.
  float len = sqrt_rn(dot(d, d));
  float inv = M::rcp(len);
  f3 u = f3{d.x * inv, d.y * inv, d.z * inv};
  float s = k * (len - rest);
  F = add(F, f3{s * u.x, s * u.y, s * u.z});
  float dd = c * M::dot(sub(v_nbr, v_self), u);
  F = add(F, f3{dd * u.x, dd * u.y, dd * u.z});
}

// Drag + lift on one face (springy_mesh.rs:631-646); the weighted share for corner `corner` is
// (angle / PI) * this (:647-652).
template <class M>
__device__ __forceinline__ f3 face_force(f3 p0, f3 p1, f3 p2, f3 u0, f3 u1, f3 u2, const StepParams& P) {
  // (v0 + v1 + v2) / 3.0 (springy_mesh.rs:634); div3_rn is that division, correctly rounded
  f3 avg;
  if constexpr (M::kExact)
    avg = f3{div3_rn((u0.x + u1.x) + u2.x), div3_rn((u0.y + u1.y) + u2.y), div3_rn((u0.z + u1.z) + u2.z)};
  else
    avg = f3{((u0.x + u1.x) + u2.x) * (1.0f / 3.0f), ((u0.y + u1.y) + u2.y) * (1.0f / 3.0f),
             ((u0.z + u1.z) + u2.z) * (1.0f / 3.0f)};
  f3 rv = sub(avg, P.wind);
  f3 cr = cross(sub(p1, p0), sub(p2, p0));
  float cm = M::sqrt_(M::dot(cr, cr));
  float area = cm / 2.0f;
  float inv = M::rcp(cm);
  f3 n = f3{cr.x * inv, cr.y * inv, cr.z * inv};
  float eff = area * fabsf(M::dot(n, rv));
  float sd = P.neg_drag * eff;
  f3 drag = f3{sd * rv.x, sd * rv.y, sd * rv.z};
  f3 nxr = cross(n, rv);
  float lm = M::sqrt_(M::dot(nxr, nxr));
  float sl = P.neg_lift * eff;
  f3 lift = f3{sl * (rv.x * lm), sl * (rv.y * lm), sl * (rv.z * lm)};
  return add(drag, lift);
}

// Face::vertex_angle (springy_mesh.rs:163-172): angle between (p_b - p_a) and (p_c - p_a) at p_a,
// called as (p2 - p0).angle(p1 - p0) for corner 0 etc.; returns angle / PI.
template <class M>
__device__ __forceinline__ float corner_weight(f3 pa, f3 pnext, f3 pprev) {
  // vertex_angle(v0, v1, v2) = (p[v2] - p[v0]).angle(p[v1] - p[v0]); angle = atan2(|a x b|, a.b)
  f3 a = sub(pprev, pa);
  f3 b = sub(pnext, pa);
  f3 axb = cross(a, b);
  float th = M::atan2_(M::sqrt_(M::dot(axb, axb)), M::dot(a, b));
  return M::div_pi(th);  // Rad / Rad(PI), PI = std::f32::consts::PI
}

// Forces of one torsional hinge (springy_mesh.rs:555-610, the reference's operation order): hinge
// strut x0 -> x1, x2 the other vertex of the strut's first face, x3 of its second; velocities of
// x2 and x3; out[0..3] are the forces on x0, x1, x2, x3.
template <class M>
__device__ __forceinline__ void hinge_forces(f3 x0, f3 x1, f3 x2, f3 x3, f3 v2, f3 v3, float k, float c, float rest,
                                             f3 out[4]) {
  const f3 d01 = sub(x1, x0), d02 = sub(x2, x0), d03 = sub(x3, x0);
  const float l01 = M::sqrt_(M::dot(d01, d01));
  const f3 h = smul(M::rcp(l01), d01);  // normalize: v * (1 / |v|)
  const float d02h = M::dot(d02, h), d03h = M::dot(d03, h);
  const f3 rl = sub(d02, smul(d02h, h)), rr = sub(d03, smul(d03h, h));
  const f3 cl = cross(d01, d02), cr = cross(d03, d01);
  const f3 nl = smul(M::rcp(M::sqrt_(M::dot(cl, cl))), cl);
  const f3 nr = smul(M::rcp(M::sqrt_(M::dot(cr, cr))), cr);
  const float theta = M::atan2_(M::dot(cross(nl, nr), h), M::dot(nl, nr));
  const float ml = M::sqrt_(M::dot(rl, rl)), mr = M::sqrt_(M::dot(rr, rr));
  const float tl = M::div(M::dot(v2, nl), ml), tr = M::div(M::dot(v3, nr), mr);
  const f3 spring = smul(k * (theta - rest), h);
  const f3 damp = smul((-1.0f * c) * (tl + tr), h);
  const f3 torque = add(spring, damp);
  const float th = M::dot(torque, h);
  out[3] = smul(M::div(th, mr), nr);
  out[2] = smul(M::div(th, ml), nl);
  out[1] = sdiv(add(smul(d02h, out[2]), smul(d03h, out[3])), l01);  // Vector3 / f32: per component
  out[0] = smul(-1.0f, add(add(out[1], out[2]), out[3]));
}

// The explicit integrators of state.rs:75-106 applied to one Point record (springy_mesh.rs:223-240)
// with the force frozen over the step (its derivative entries are 0).  RK4 expands to the stage
// values below; each line is the reference's vec_add/scale of that record component.
template <class M>
__device__ __forceinline__ void integrate(const StepParams& P, float m, f3 x, f3 v, f3 F, f3& xn, f3& vn) {
  f3 a1 = f3{M::div(F.x, m), M::div(F.y, m), M::div(F.z, m)};
  if (P.integration == FP_EULER) {  // S + S' * h  (state.rs:75-83)
    xn = f3{x.x + v.x * P.dt, x.y + v.y * P.dt, x.z + v.z * P.dt};
    vn = f3{v.x + a1.x * P.dt, v.y + a1.y * P.dt, v.z + a1.z * P.dt};
    return;
  }
  // Stages 2-4 see the force record F + 0 (the force slot's derivative is 0), which differs from F
  // only for F = -0: then (F + 0) / m = +0 / m, otherwise it is a1.
  const float z0 = M::div(0.0f, m);
  f3 a2 = f3{F.x == 0.0f ? z0 : a1.x, F.y == 0.0f ? z0 : a1.y, F.z == 0.0f ? z0 : a1.z};
  const float hh = P.dt_half, h = P.dt, h6 = P.dt6, h3 = P.dt3;
  f3 v2 = f3{v.x + a1.x * hh, v.y + a1.y * hh, v.z + a1.z * hh};
  f3 v3 = f3{v.x + a2.x * hh, v.y + a2.y * hh, v.z + a2.z * hh};
  f3 v4 = f3{v.x + a2.x * h, v.y + a2.y * h, v.z + a2.z * h};
  // position slot: k1 = v, k2 = v2, k3 = v3, k4 = v4 ; velocity slot: k1 = a1, k2..k4 = a2
  xn.x = x.x + (((h6 * v.x + h3 * v2.x) + h3 * v3.x) + h6 * v4.x);
  xn.y = x.y + (((h6 * v.y + h3 * v2.y) + h3 * v3.y) + h6 * v4.y);
  xn.z = x.z + (((h6 * v.z + h3 * v2.z) + h3 * v3.z) + h6 * v4.z);
  vn.x = v.x + (((h6 * a1.x + h3 * a2.x) + h3 * a2.x) + h6 * a2.x);
  vn.y = v.y + (((h6 * a1.y + h3 * a2.y) + h3 * a2.y) + h6 * a2.y);
  vn.z = v.z + (((h6 * a1.z + h3 * a2.z) + h3 * a2.z) + h6 * a2.z);
}

// CollidableMesh::get_collided_face_from_list's per-face predicate (collidable_mesh.rs:113-163).
// Always exact: it decides a branch, and its inputs are the same bits in every math policy.
__device__ __forceinline__ bool face_hit(const ObsFace& f, f3 old_p, f3 new_p, float dt_d) {
  f3 n = f3{f.n[0], f.n[1], f.n[2]};
  f3 v0 = f3{f.v0[0], f.v0[1], f.v0[2]};
  float od = dot(sub(old_p, v0), n);
  float nd = dot(sub(new_p, v0), n);
  if (sign_positive(od) == sign_positive(nd)) return false;
  // old_velocity = (new - old) / dt (collidable_mesh.rs:114), needed only once a plane is crossed
  const f3 ov = f3{(new_p.x - old_p.x) / dt_d, (new_p.y - old_p.y) / dt_d, (new_p.z - old_p.z) / dt_d};
  float frac = od / (od - nd);
  float s = dt_d * frac;
  f3 cp = f3{old_p.x + s * ov.x, old_p.y + s * ov.y, old_p.z + s * ov.z};
  f3 a = v0, b = f3{f.v1[0], f.v1[1], f.v1[2]}, c = f3{f.v2[0], f.v2[1], f.v2[2]}, p = cp;
  if (f.axis == 0) { a.x = 0.0f; b.x = 0.0f; c.x = 0.0f; p.x = 0.0f; }
  else if (f.axis == 1) { a.y = 0.0f; b.y = 0.0f; c.y = 0.0f; p.y = 0.0f; }
  else { a.z = 0.0f; b.z = 0.0f; c.z = 0.0f; p.z = 0.0f; }
  bool o0 = sign_positive(dot(cross(sub(b, a), sub(p, a)), n));
  bool o1 = sign_positive(dot(cross(sub(c, b), sub(p, b)), n));
  bool o2 = sign_positive(dot(cross(sub(a, c), sub(p, c)), n));
  return o0 == o1 && o1 == o2;
}

// Lowest-index face hit in one grid by the segment whose grown box is [lo, hi]; `best` is the
// lowest hit so far (faces at or above it need no test).
__device__ __forceinline__ void grid_hits(const GridView& G, const ObsFace* faces, const float lo[3], const float hi[3],
                                          f3 old_p, f3 new_p, float dt_d, uint32_t& best) {
  if (G.cell_start == nullptr) return;
  int32_t c0[3], c1[3];
  for (int a = 0; a < 3; ++a) {
    const float l = (lo[a] - G.lo[a]) * G.inv_cell[a];
    const float h = (hi[a] - G.lo[a]) * G.inv_cell[a];
    if (!(h >= 0.0f) || !(l < (float)G.dims[a])) return;  // outside the grid (or NaN)
    c0[a] = l < 0.0f ? 0 : (int32_t)l;
    c1[a] = h >= (float)G.dims[a] ? G.dims[a] - 1 : (int32_t)h;
  }
  for (int32_t cz = c0[2]; cz <= c1[2]; ++cz)
    for (int32_t cy = c0[1]; cy <= c1[1]; ++cy)
      for (int32_t cx = c0[0]; cx <= c1[0]; ++cx) {
        const uint32_t cell = ((uint32_t)cz * (uint32_t)G.dims[1] + (uint32_t)cy) * (uint32_t)G.dims[0] + (uint32_t)cx;
        // the cell's faces in ascending order, each index loaded while the one before is tested
        uint32_t k = G.cell_start[cell];
        const uint32_t e = G.cell_start[cell + 1];
        uint32_t f = k < e ? G.cell_faces[k] : 0xffffffffu;
        while (f < best) {
          ++k;
          const uint32_t next = k < e ? G.cell_faces[k] : 0xffffffffu;
          if (face_hit(load_face(faces, f), old_p, new_p, dt_d)) {
            best = f;
            break;
          }
          f = next;
        }
      }
}

// First face (lowest index over all obstacles, in order) hit by the segment old_p -> new_p, or -1
// (CollidableMesh::get_collided_face_from_list, collidable_mesh.rs:107-169).  With the grids the
// candidates are fewer but the answer is the same: the lowest index whose predicate holds.
__device__ __forceinline__ int32_t first_hit(const ObsView& O, f3 old_p, f3 new_p, float dt_d) {
  if (O.n_faces == 0) return -1;
  if (!O.use_grid) {
    for (uint32_t f = 0; f < O.n_faces; ++f)
      if (face_hit(load_face(O.faces, f), old_p, new_p, dt_d)) return (int32_t)f;
    return -1;
  }
  uint32_t best = 0xffffffffu;
  for (uint32_t k = 0; k < O.n_always; ++k) {
    const uint32_t f = O.always[k];
    if (f >= best) break;
    if (face_hit(load_face(O.faces, f), old_p, new_p, dt_d)) { best = f; break; }
  }
  const float lo[3] = {fminf(old_p.x, new_p.x) - O.margin, fminf(old_p.y, new_p.y) - O.margin,
                       fminf(old_p.z, new_p.z) - O.margin};
  const float hi[3] = {fmaxf(old_p.x, new_p.x) + O.margin, fmaxf(old_p.y, new_p.y) + O.margin,
                       fmaxf(old_p.z, new_p.z) + O.margin};
  for (int g = 0; g < 4; ++g) grid_hits(O.grid[g], O.faces, lo, hi, old_p, new_p, dt_d, best);
  return best == 0xffffffffu ? -1 : (int32_t)best;
}

// Collision response of SpringyMesh::update_points for a hit (springy_mesh.rs:444-480).
// Always exact (it runs for few vertices and decides positions next to a surface).
__device__ __forceinline__ void collide_response(const ObsFace& f, const StepParams& P, float m, f3 x, f3 v, f3 F,
                                                 f3& xn, f3& vn) {
  f3 n = f3{f.n[0], f.n[1], f.n[2]};
  f3 v0 = f3{f.v0[0], f.v0[1], f.v0[2]};
  float od = dot(sub(x, v0), n);
  float nd = dot(sub(xn, v0), n);
  float frac = od / (od - nd);
  float s = P.dt * frac;
  f3 cp = f3{x.x + s * v.x, x.y + s * v.y, x.z + s * v.z};
  f3 vc = f3{v.x + (s * F.x) / m, v.y + (s * F.y) / m, v.z + (s * F.z) / m};
  const float eps = 0.0001f;  // consts::EPSILON (consts.rs:2)
  f3 np = f3{cp.x + n.x * eps, cp.y + n.y * eps, cp.z + n.z * eps};
  float vdn = dot(vc, n);
  f3 vcn = f3{vdn * n.x, vdn * n.y, vdn * n.z};
  f3 vct = sub(vc, vcn);
  f3 vrn = f3{(-1.0f * vcn.x) * P.restitution, (-1.0f * vcn.y) * P.restitution, (-1.0f * vcn.z) * P.restitution};
  f3 vrt = f3{0.0f, 0.0f, 0.0f};
  float mt = sqrtf(dot(vct, vct));
  bool zero = vct.x == 0.0f && vct.y == 0.0f && vct.z == 0.0f;
  if (!zero && !__builtin_isnan(mt)) {
    float inv = 1.0f / mt;
    float lim = fminf(P.friction * sqrtf(dot(vcn, vcn)), mt);
    vrt = f3{vct.x - (vct.x * inv) * lim, vct.y - (vct.y * inv) * lim, vct.z - (vct.z * inv) * lim};
  }
  xn = np;
  vn = add(vrn, vrt);
}

// The learned surrogate's big step for one free vertex (every kernel that takes it calls this, with
// explicit fused multiply-adds, so that its results do not depend on a file's contraction setting):
// a = ((gravity + Ff) + Fx) / m + da, v' = v + H a, x' = x + H v', then the collision response over
// H with the force m a; returns the face-force work Ff . (x' - x).  Fx: explicit strut forces when
// the big step keeps them (their energy is counted as spring energy, not as work).
__device__ __forceinline__ float sur_update(const StepParams& P, const ObsView& O, float m, f3 p, f3 v, f3 Ff,
                                            f3 Fx, f3 da, f3& xn, f3& vn) {
  const f3 F = add(add(P.gravity, Ff), Fx);
  const float im = 1.0f / m;
  const f3 a = f3{__builtin_fmaf(F.x, im, da.x), __builtin_fmaf(F.y, im, da.y), __builtin_fmaf(F.z, im, da.z)};
  const float H = P.dt;
  vn = f3{__builtin_fmaf(H, a.x, v.x), __builtin_fmaf(H, a.y, v.y), __builtin_fmaf(H, a.z, v.z)};
  xn = f3{__builtin_fmaf(H, vn.x, p.x), __builtin_fmaf(H, vn.y, p.y), __builtin_fmaf(H, vn.z, p.z)};
  const int32_t hit = first_hit(O, p, xn, P.dt_d);
  if (hit >= 0) collide_response(O.faces[hit], P, m, p, v, f3{a.x * m, a.y * m, a.z * m}, xn, vn);
  return __builtin_fmaf(Ff.z, xn.z - p.z, __builtin_fmaf(Ff.y, xn.y - p.y, Ff.x * (xn.x - p.x)));
}

__device__ __forceinline__ bool pinned(const uint32_t* __restrict__ bits, uint32_t i) {
  return bits != nullptr && ((bits[i >> 5] >> (i & 31u)) & 1u);
}

}  // namespace fp
