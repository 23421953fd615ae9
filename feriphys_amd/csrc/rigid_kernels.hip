// The rigid-body step (rigidbody/simulation.rs:26-52) for a batch of unit cubes, one thread per
// body, in the reference's operation order (no FMA contraction; correctly rounded division and
// square root):
//   accumulate_forces / accumulate_torques (rigidbody.rs:322-328);
//   State::rk4_step or euler_step over the body's 29-float record (state.rs:75-106), whose
//     derivative (rigidbody.rs:94-157) is (p / m, 0.5 (0, w) q, F, tau, 0...) with
//     w = (R I0^-1 R^T) L and R = Matrix3::from(q);
//   normalize_rotation (:29-31);
//   update_state (:257-319): each of the cube's 8 vertices, old and new world positions, the first
//     obstacle face its segment crosses (CollidableMesh::get_collided_face_from_list), and the
//     reference's impulse response on the new state;
//   clear_forces / clear_torques (:330-336).
// cgmath's quaternion and matrix operations are restated below with their component order.
#include "rigid.hpp"

namespace fp {

namespace {

struct Q {
  float s;
  f3 v;
};
// Quaternion * Quaternion (Hamilton product), each component summed left to right.
__device__ __forceinline__ Q qmul(Q a, Q b) {
  return Q{a.s * b.s - a.v.x * b.v.x - a.v.y * b.v.y - a.v.z * b.v.z,
           f3{a.s * b.v.x + a.v.x * b.s + a.v.y * b.v.z - a.v.z * b.v.y,
              a.s * b.v.y + a.v.y * b.s + a.v.z * b.v.x - a.v.x * b.v.z,
              a.s * b.v.z + a.v.z * b.s + a.v.x * b.v.y - a.v.y * b.v.x}};
}
struct M3 {
  f3 c[3];  // columns
};
__device__ __forceinline__ f3 row(const M3& m, int i) {
  return i == 0 ? f3{m.c[0].x, m.c[1].x, m.c[2].x} : i == 1 ? f3{m.c[0].y, m.c[1].y, m.c[2].y}
                                                             : f3{m.c[0].z, m.c[1].z, m.c[2].z};
}
// Matrix3 * Vector3: row(i) . v
__device__ __forceinline__ f3 mv(const M3& m, f3 v) { return f3{dot(row(m, 0), v), dot(row(m, 1), v), dot(row(m, 2), v)}; }
__device__ __forceinline__ M3 mm(const M3& a, const M3& b) { return M3{{mv(a, b.c[0]), mv(a, b.c[1]), mv(a, b.c[2])}}; }
__device__ __forceinline__ M3 tr(const M3& m) { return M3{{row(m, 0), row(m, 1), row(m, 2)}}; }
// Matrix3::from(Quaternion)
__device__ __forceinline__ M3 from_quat(Q q) {
  const float x2 = q.v.x + q.v.x, y2 = q.v.y + q.v.y, z2 = q.v.z + q.v.z;
  const float xx2 = x2 * q.v.x, xy2 = x2 * q.v.y, xz2 = x2 * q.v.z;
  const float yy2 = y2 * q.v.y, yz2 = y2 * q.v.z, zz2 = z2 * q.v.z;
  const float sy2 = y2 * q.s, sz2 = z2 * q.s, sx2 = x2 * q.s;
  return M3{{f3{1.0f - yy2 - zz2, xy2 + sz2, xz2 - sy2}, f3{xy2 - sz2, 1.0f - xx2 - zz2, yz2 + sx2},
             f3{xz2 + sy2, yz2 - sx2, 1.0f - xx2 - yy2}}};
}
__device__ __forceinline__ Q rot(const float* S) { return Q{S[6], f3{S[3], S[4], S[5]}}; }
__device__ __forceinline__ M3 inertia0(const float* S) {
  return M3{{f3{S[14], S[15], S[16]}, f3{S[17], S[18], S[19]}, f3{S[20], S[21], S[22]}}};
}
// State::get_moment_of_inertia_inverted (rigidbody.rs:33-36): (R I0^-1) R^T
__device__ __forceinline__ M3 inertia_inv(const float* S) {
  const M3 R = from_quat(rot(S));
  return mm(mm(R, inertia0(S)), tr(R));
}

// Stateful::derivative (rigidbody.rs:94-157)
__device__ __forceinline__ void derivative(const float* S, float* D) {
  const float m = S[13];
  const f3 v = f3{S[7] / m, S[8] / m, S[9] / m};  // linear_momentum / mass
  const f3 w = mv(inertia_inv(S), f3{S[10], S[11], S[12]});
  const Q qd = qmul(Q{0.5f * 0.0f, f3{0.5f * w.x, 0.5f * w.y, 0.5f * w.z}}, rot(S));
  D[0] = v.x; D[1] = v.y; D[2] = v.z;
  D[3] = qd.v.x; D[4] = qd.v.y; D[5] = qd.v.z; D[6] = qd.s;
  for (int k = 7; k < 13; ++k) D[k] = S[k + 16];  // force, torque
#pragma unroll
  for (int k = 13; k < kRigidRecord; ++k) D[k] = 0.0f;
}

__global__ void __launch_bounds__(kBlock) rigid_step_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                             uint32_t n, uint32_t pitch, RigidParams P, ObsView O) {
  const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= n) return;
  float S[kRigidRecord], T[kRigidRecord], K[kRigidRecord], A[kRigidRecord];
  // field-major records: each field is one coalesced wavefront load
#pragma unroll
  for (int k = 0; k < kRigidRecord; ++k) S[k] = in[(size_t)k * pitch + b];
  // accumulate_forces, accumulate_torques
  S[23] = S[23] + P.gravity.x; S[24] = S[24] + P.gravity.y; S[25] = S[25] + P.gravity.z;
  S[26] = S[26] + P.torque.x; S[27] = S[27] + P.torque.y; S[28] = S[28] + P.torque.z;
  const float h = P.dt;
  if (P.integration == FP_EULER) {  // S + S' h (state.rs:75-83)
    derivative(S, K);
#pragma unroll
    for (int k = 0; k < kRigidRecord; ++k) T[k] = S[k] + K[k] * h;
  } else {  // state.rs:86-106, delta accumulated left to right
    const float hh = h * 0.5f, h6 = h / 6.0f, h3 = h / 3.0f;
    derivative(S, K);
#pragma unroll
    for (int k = 0; k < kRigidRecord; ++k) { A[k] = h6 * K[k]; T[k] = S[k] + K[k] * hh; }
    derivative(T, K);
#pragma unroll
    for (int k = 0; k < kRigidRecord; ++k) { A[k] = A[k] + h3 * K[k]; T[k] = S[k] + K[k] * hh; }
    derivative(T, K);
#pragma unroll
    for (int k = 0; k < kRigidRecord; ++k) { A[k] = A[k] + h3 * K[k]; T[k] = S[k] + K[k] * h; }
    derivative(T, K);
#pragma unroll
    for (int k = 0; k < kRigidRecord; ++k) { A[k] = A[k] + h6 * K[k]; T[k] = S[k] + A[k]; }
  }
  // normalize_rotation: q * (1 / sqrt(s s + v . v))
  {
    const float inv = 1.0f / __builtin_sqrtf(T[6] * T[6] + ((T[3] * T[3] + T[4] * T[4]) + T[5] * T[5]));
    T[3] = T[3] * inv; T[4] = T[4] * inv; T[5] = T[5] * inv; T[6] = T[6] * inv;
  }
  // update_state: the cube's vertices against the obstacles
  if (O.n_faces != 0) {
    const M3 R_old = from_quat(rot(S)), R_new = from_quat(rot(T));
    const f3 x_old = f3{S[0], S[1], S[2]}, x_new = f3{T[0], T[1], T[2]};
    const M3 I_old = inertia_inv(S);
    const float m = S[13];
    const f3 vel = f3{S[7] / m, S[8] / m, S[9] / m};
    const f3 w = mv(I_old, f3{S[10], S[11], S[12]});
    for (int k = 0; k < 8; ++k) {
      const f3 c = f3{kCubeCorner[k][0], kCubeCorner[k][1], kCubeCorner[k][2]};
      const f3 op = add(mv(R_old, c), x_old), np = add(mv(R_new, c), x_new);
      const int32_t hit = first_hit(O, op, np, P.dt_d);
      if (hit < 0) continue;
      const ObsFace& f = O.faces[hit];
      const f3 nrm = f3{f.n[0], f.n[1], f.n[2]}, v0 = f3{f.v0[0], f.v0[1], f.v0[2]};
      const float od = dot(sub(op, v0), nrm), nd = dot(sub(np, v0), nrm);
      const f3 r = sub(op, x_old);
      const float frac = od / (od - nd);
      const f3 cv = add(vel, cross(w, r));
      const f3 cp = add(op, smul(P.dt * frac, cv));
      const float vn = dot(cv, nrm);
      const float j = (-(1.0f + P.restitution) * vn) / (1.0f / m + dot(nrm, mv(I_old, cross(cross(r, nrm), r))));
      const f3 imp = smul(j, nrm);
      const f3 pos = add(sub(cp, r), smul(0.0001f * 2.0f, nrm));  // consts::EPSILON * 2.0
      T[0] = pos.x; T[1] = pos.y; T[2] = pos.z;
      T[7] = T[7] + imp.x; T[8] = T[8] + imp.y; T[9] = T[9] + imp.z;  // apply_impulse (:38-41)
      const f3 dl = cross(r, imp);
      T[10] = T[10] + dl.x; T[11] = T[11] + dl.y; T[12] = T[12] + dl.z;
    }
  }
  // clear_forces, clear_torques
#pragma unroll
  for (int k = 23; k < kRigidRecord; ++k) T[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < kRigidRecord; ++k) out[(size_t)k * pitch + b] = T[k];
}

// Field-major device records <-> the caller's 29-float records, body after body.
__global__ void __launch_bounds__(kBlock) rigid_pack_kernel(const float* __restrict__ in, uint32_t n, uint32_t pitch,
                                                             float* __restrict__ out) {
  const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= n) return;
#pragma unroll
  for (int k = 0; k < kRigidRecord; ++k) out[(size_t)kRigidRecord * b + k] = in[(size_t)k * pitch + b];
}
__global__ void __launch_bounds__(kBlock) rigid_unpack_kernel(const float* __restrict__ in, uint32_t n, uint32_t pitch,
                                                               float* __restrict__ out) {
  const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= n) return;
#pragma unroll
  for (int k = 0; k < kRigidRecord; ++k) out[(size_t)k * pitch + b] = in[(size_t)kRigidRecord * b + k];
}

}  // namespace

hipError_t launch_rigid_records(const float* in, float* out, uint32_t n, uint32_t pitch, bool pack, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (pitch < n) return hipErrorInvalidValue;
  const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
  if (pack) rigid_pack_kernel<<<g, b, 0, s>>>(in, n, pitch, out);
  else rigid_unpack_kernel<<<g, b, 0, s>>>(in, n, pitch, out);
  return hipGetLastError();
}

hipError_t launch_rigid_step(const float* in, float* out, uint32_t n, uint32_t pitch, const RigidParams& p,
                             const ObsView& obs, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (pitch < n) return hipErrorInvalidValue;
  rigid_step_kernel<<<dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s>>>(in, out, n, pitch, p, obs);
  return hipGetLastError();
}

}  // namespace fp
