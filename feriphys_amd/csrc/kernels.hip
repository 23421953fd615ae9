// gfx950 kernels of the springy step: one thread per vertex, each vertex gathers its own force
// (struts in strut order, faces in face order), integrates, resolves obstacle collision, restores
// pins and writes its new state.  Reads come from `in`, writes go to `out` (ping-pong), so one
// launch is one Simulation::step for every vertex of every mesh it covers.
//
//   csr_kernel  : any SpringyMesh (SpringyMesh::new + add_strut), topology from a vertex-centric
//                 CSR built on the host.
//   grid_kernel : meshes built by Cloth::new and left unchanged; struts, rest lengths and faces are
//                 rebuilt from (row, col) on chip, so the only HBM traffic is the state itself.
#include "kernels.hpp"

namespace fp {

namespace {

__device__ __forceinline__ void store_state(const SoA& o, uint32_t i, f3 x, f3 v) {
  o.x[i] = x.x; o.y[i] = x.y; o.z[i] = x.z;
  o.vx[i] = v.x; o.vy[i] = v.y; o.vz[i] = v.z;
}

// Integration, collision and pins: the tail of Simulation::step for one vertex
// (springy/simulation.rs:145-155 -> state.rs, springy_mesh.rs:423-497).
template <class M>
__device__ __forceinline__ void finish_step(const StepArgs& A, const StepParams& P, uint32_t i, float m, f3 p, f3 v,
                                            f3 F, bool pin) {
  f3 xn, vn;
  integrate<M>(P, m, p, v, F, xn, vn);
  int32_t hit = first_hit(A.obs, p, xn, P.dt_d);
  if (hit >= 0) collide_response(A.obs.faces[hit], P, m, p, v, F, xn, vn);
  if (pin) { xn = p; vn = v; }  // pinned points keep their pre-step record (:492-496)
  store_state(A.out, i, xn, vn);
  if (A.force_init) { A.fx[i] = 0.0f; A.fy[i] = 0.0f; A.fz[i] = 0.0f; }  // clear_forces (:656-660)
}

template <class M, int MODE>
__global__ void __launch_bounds__(kBlock) csr_kernel(StepArgs A, CsrView T, const BlockRange* __restrict__ blocks) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  if (i >= br.end) return;
  const StepParams P = A.params[br.mesh];
  const SoA& S = A.in;
  const f3 p = ld3(S.x, S.y, S.z, i);
  const f3 v = ld3(S.vx, S.vy, S.vz, i);
  const float m = P.uniform_mass ? P.mass : A.mass[i];
  f3 F = A.force_init ? f3{A.fx[i], A.fy[i], A.fz[i]} : f3{0.0f, 0.0f, 0.0f};
  F = add(F, P.gravity);  // apply_external_point_forces (:527-531)
  // apply_strut_forces (:533-551), in strut order.  Four entries at a time: their loads (entries,
  // then neighbour states) are all issued before the first force is added, so one row costs a few
  // memory latencies rather than two per strut; the additions stay in strut order.
  {
    const StrutEnt* ent = T.s_ent + T.s_first[i];
    const uint32_t ns = T.s_cnt[i];
    for (uint32_t e = 0; e < ns; e += 4) {
      StrutEnt se[4];
      f3 q[4], w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) se[j] = ent[kSlice * (e + j < ns ? e + j : e)];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        q[j] = ld3(S.x, S.y, S.z, se[j].nbr);
        w[j] = ld3(S.vx, S.vy, S.vz, se[j].nbr);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (e + j < ns) strut_accumulate<M>(F, p, v, q[j], w[j], se[j].k, se[j].c, se[j].len);
    }
  }
  // apply_torsional_forces (:553-627), when the step includes it (the reference comments the call
  // out at :518-519): the vertex's hinge forces summed in strut order from zero, then added once
  if (T.t_row != nullptr) {
    uint32_t e = T.t_row[i];
    const uint32_t ee = T.t_row[i + 1];
    if (e < ee) {
      f3 ts = f3{0.0f, 0.0f, 0.0f};
      for (; e < ee; ++e) {
        const float4 hf = T.hinge_f[T.t_ent[e]];
        ts = add(ts, f3{hf.x, hf.y, hf.z});
      }
      F = add(F, ts);
    }
  }
  // apply_face_forces (:629-654): the face kernel has stored each corner's weight * (drag + lift)
  // in this vertex's face entries; they are added in face order
  {
    const float4* fe = T.face_c + T.f_first[i];
    const uint32_t nf = T.f_cnt[i];
    for (uint32_t e = 0; e < nf; e += 4) {
      float4 fc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) fc[j] = fe[kSlice * (e + j < nf ? e + j : e)];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (e + j < nf) F = add(F, f3{fc[j].x, fc[j].y, fc[j].z});
    }
  }
  const bool pin = pinned(A.pins, i);
  if (pin) F = f3{0.0f, 0.0f, 0.0f};  // (:522-524)
  if (MODE == MODE_FORCES) {
    A.fx[i] = F.x; A.fy[i] = F.y; A.fz[i] = F.z;
    return;
  }
  finish_step<M>(A, P, i, m, p, v, F, pin);
}

// The step of one patch of a general mesh per workgroup (the operations of face_kernel and
// csr_kernel, in the same order per vertex):
//   phase 0  stage the states of the patch and its halo in LDS with one pass of loads, and the
//            mesh's table of nominal strut constants;
//   phase 1  drag + lift of every face with a corner in the patch into LDS (a face on the patch
//            border is computed by each patch that holds one of its corners);
//   phase 2  every patch vertex: struts in strut order (8-byte entries, neighbour states from LDS),
//            then its faces in face order, each its own corner's weight (angle / PI) times the
//            face's drag + lift; integrate, collide, pin, store.
// LDS: the staged states as (x, y, z, vx) and (vy, vz) at fixed strides (one address per
// neighbour, two wide reads), and the face forces sized per launch, nf_cap float4.
template <class M>
__global__ void __launch_bounds__(kBlock) patch_kernel(StepArgs A, CsrView T, const PatchRange* __restrict__ patches) {
  __shared__ float4 s_p[kPatchLocal];  // x, y, z, vx
  __shared__ float2 s_w[kPatchLocal];  // vy, vz
  __shared__ float2 s_cls[kPatchClasses];
  extern __shared__ float4 s_ff[];     // per patch face: drag + lift
  const PatchRange R = patches[blockIdx.x];
  const uint32_t np = R.end - R.begin, nl = np + R.n_halo;
  const SoA& S = A.in;
  const StepParams P = A.params[R.mesh];
  if (threadIdx.x < kPatchClasses) s_cls[threadIdx.x] = T.p_cls[R.mesh * kPatchClasses + threadIdx.x];
  for (uint32_t e = threadIdx.x; e < nl; e += kBlock) {
    const uint32_t g = e < np ? R.begin + e : T.halo[R.halo_begin + (e - np)];
    s_p[e] = make_float4(S.x[g], S.y[g], S.z[g], S.vx[g]);
    s_w[e] = make_float2(S.vy[g], S.vz[g]);
  }
  __syncthreads();
  auto Pl = [&](uint32_t l) { const float4 a = s_p[l]; return f3{a.x, a.y, a.z}; };
  auto Vl = [&](uint32_t l) { const float4 a = s_p[l]; const float2 b = s_w[l]; return f3{a.w, b.x, b.y}; };
  for (uint32_t q = threadIdx.x; q < R.n_faces; q += kBlock) {
    const PatchFace pf = T.patch_faces[R.face_begin + q];
    const f3 sum = face_force<M>(Pl(pf.v[0]), Pl(pf.v[1]), Pl(pf.v[2]), Vl(pf.v[0]), Vl(pf.v[1]), Vl(pf.v[2]), P);
    s_ff[q] = make_float4(sum.x, sum.y, sum.z, 0.0f);
  }
  __syncthreads();
  if (threadIdx.x >= np) return;
  const uint32_t t = threadIdx.x, i = R.begin + t;
  const f3 p = Pl(t), v = Vl(t);
  const float m = P.uniform_mass ? P.mass : A.mass[i];
  f3 F = A.force_init ? f3{A.fx[i], A.fy[i], A.fz[i]} : f3{0.0f, 0.0f, 0.0f};
  F = add(F, P.gravity);  // apply_external_point_forces (:527-531)
  // apply_strut_forces (:533-551), in strut order; each strut's stiffness and damping rebuilt as
  // Strut::new makes them, (1.0 / len) * nominal (:83-84)
  {
    const PatchStrut* ent = T.p_ent + T.s_first[i];
    const uint32_t ns = T.s_cnt[i];
    for (uint32_t e = 0; e < ns; e += 4) {
      PatchStrut se[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) se[j] = ent[kSlice * (e + j < ns ? e + j : e)];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (e + j < ns) {
          const float2 kc = s_cls[se[j].cls];
          const float il = rcp_rn(se[j].len);
          strut_accumulate<M>(F, p, v, Pl(se[j].loc), Vl(se[j].loc), il * kc.x, il * kc.y, se[j].len);
        }
    }
  }
  // apply_torsional_forces, when the step includes it (see csr_kernel)
  if (T.t_row != nullptr) {
    uint32_t e = T.t_row[i];
    const uint32_t ee = T.t_row[i + 1];
    if (e < ee) {
      f3 ts = f3{0.0f, 0.0f, 0.0f};
      for (; e < ee; ++e) {
        const float4 hf = T.hinge_f[T.t_ent[e]];
        ts = add(ts, f3{hf.x, hf.y, hf.z});
      }
      F = add(F, ts);
    }
  }
  // apply_face_forces (:629-654): in face order, this vertex's corner weight (vertex_angle / PI,
  // :647-649) times the face's drag + lift (:650-652)
  {
    const PatchCorner* cor = T.p_corner + T.f_first[i];
    const uint32_t nf = T.f_cnt[i];
    for (uint32_t j = 0; j < nf; j += 4) {
      PatchCorner ce[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) ce[k] = cor[kSlice * (j + k < nf ? j + k : j)];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (j + k < nf) {
          const float w = corner_weight<M>(p, Pl(ce[k].next), Pl(ce[k].prev));
          const float4 ff = s_ff[ce[k].q];
          F = add(F, f3{w * ff.x, w * ff.y, w * ff.z});
        }
    }
  }
  const bool pin = pinned(A.pins, i);
  if (pin) F = f3{0.0f, 0.0f, 0.0f};  // (:522-524)
  finish_step<M>(A, P, i, m, p, v, F, pin);
}

// Drag + lift (springy_mesh.rs:631-646) and the three corner weights angle / PI (:647-649) of every
// face of the CSR meshes, once per face; corner k's share weight_k * (drag + lift) (the product of
// :650-652) goes to its vertex's face entry.
template <class M>
__global__ void __launch_bounds__(kBlock) face_kernel(StepArgs A, CsrView T, const BlockRange* __restrict__ blocks) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t f = br.begin + threadIdx.x;
  if (f >= br.end) return;
  const StepParams P = A.params[br.mesh];
  const SoA& S = A.in;
  const uint32_t a = T.f_v[3 * f], b = T.f_v[3 * f + 1], c = T.f_v[3 * f + 2];
  const f3 p0 = ld3(S.x, S.y, S.z, a), p1 = ld3(S.x, S.y, S.z, b), p2 = ld3(S.x, S.y, S.z, c);
  const f3 sum = face_force<M>(p0, p1, p2, ld3(S.vx, S.vy, S.vz, a), ld3(S.vx, S.vy, S.vz, b),
                               ld3(S.vx, S.vy, S.vz, c), P);
  const float w0 = corner_weight<M>(p0, p1, p2), w1 = corner_weight<M>(p1, p2, p0), w2 = corner_weight<M>(p2, p0, p1);
  T.face_c[T.f_slot[3 * f]] = make_float4(w0 * sum.x, w0 * sum.y, w0 * sum.z, 0.0f);
  T.face_c[T.f_slot[3 * f + 1]] = make_float4(w1 * sum.x, w1 * sum.y, w1 * sum.z, 0.0f);
  T.face_c[T.f_slot[3 * f + 2]] = make_float4(w2 * sum.x, w2 * sum.y, w2 * sum.z, 0.0f);
}

// The four forces of each torsional hinge, once per hinge.
template <class M>
__global__ void __launch_bounds__(kBlock) hinge_kernel(StepArgs A, CsrView T, const HingeEnt* __restrict__ hinges,
                                                        uint32_t first, uint32_t count) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  const uint32_t hgi = first + k;
  const HingeEnt H = hinges[hgi];
  const SoA& S = A.in;
  f3 out[4];
  hinge_forces<M>(ld3(S.x, S.y, S.z, H.v[0]), ld3(S.x, S.y, S.z, H.v[1]), ld3(S.x, S.y, S.z, H.v[2]),
                  ld3(S.x, S.y, S.z, H.v[3]), ld3(S.vx, S.vy, S.vz, H.v[2]), ld3(S.vx, S.vy, S.vz, H.v[3]), H.k, H.c,
                  H.rest, out);
  for (int r = 0; r < 4; ++r) T.hinge_f[4 * hgi + r] = make_float4(out[r].x, out[r].y, out[r].z, 0.0f);
}

// Split API: F += the vertex's hinge forces (vertices with no hinge are left untouched).
__global__ void __launch_bounds__(kBlock) torsion_gather_kernel(StepArgs A, CsrView T,
                                                                 const BlockRange* __restrict__ blocks) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  if (i >= br.end) return;
  uint32_t e = T.t_row[i];
  const uint32_t ee = T.t_row[i + 1];
  if (e == ee) return;
  f3 ts = f3{0.0f, 0.0f, 0.0f};
  for (; e < ee; ++e) {
    const float4 hf = T.hinge_f[T.t_ent[e]];
    ts = add(ts, f3{hf.x, hf.y, hf.z});
  }
  A.fx[i] = A.fx[i] + ts.x;
  A.fy[i] = A.fy[i] + ts.y;
  A.fz[i] = A.fz[i] + ts.z;
}

// Grid cloth built by Cloth::new (springy/cloth.rs:20-113).  Vertex (r, c) of an R x C grid.
// Strut ids follow first appearance over the triangle list (top-left triangles of every quad, then
// bottom-right ones), then the bending struts; so the incident struts of a vertex, in id order, are:
//   top-left-created  : shear to (r-1,c+1), tensile to (r-1,c)   [quad (r-1,c)]
//                       shear to (r+1,c-1), tensile to (r,c-1)   [quad (r,c-1)]
//                       tensile to (r+1,c), tensile to (r,c+1)   [quad (r,c)]
//   bottom-right-created (last column verticals, last row horizontals), ids after all of the above:
//                       (r-1,c) if c=C-1, (r,c-1) if r=R-1, (r+1,c) if c=C-1, (r,c+1) if r=R-1
//   bending (add_strut, cloth.rs:99-110): (r-2,c), (r,c-2), (r,c+2), (r+2,c)
// Faces: A(q) = [(r',c'+1), (r'+1,c'), (r',c')] for every quad q in row-major order, then
// B(q) = [(r',c'+1), (r'+1,c'+1), (r'+1,c')]; a vertex sees A(r-1,c), A(r,c-1), A(r,c),
// B(r-1,c-1), B(r-1,c), B(r,c-1) in that order, as corners 1, 0, 2, 1, 2, 0.
// Rest lengths are |p_a - p_b| of the construction grid: one of the three components is exactly 0
// in either plane, so |d| = sqrt(dX*dX + dY*dY) with dX, dY from the grid coordinates below.
template <class M, int MODE>
__global__ void __launch_bounds__(kBlock) grid_kernel(StepArgs A, const ClothView* __restrict__ cloths,
                                                       const BlockRange* __restrict__ blocks) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  if (i >= br.end) return;
  const StepParams P = A.params[br.mesh];
  const ClothView G = cloths[br.mesh];
  const SoA& S = A.in;
  const int R = (int)G.rows, C = (int)G.cols;
  const int li = (int)(i - G.base);
  const int r = li / C, c = li - r * C;
  auto id = [&](int rr, int cc) -> uint32_t { return G.base + (uint32_t)(rr * C + cc); };
  auto X = [&](int cc) -> float { return ((float)cc * G.spacing - G.cx) + G.p0; };
  auto Y = [&](int rr) -> float { return ((float)rr * G.spacing - G.cy) + G.p1; };
  const f3 p = ld3(S.x, S.y, S.z, i);
  const f3 v = ld3(S.vx, S.vy, S.vz, i);
  const float m = P.uniform_mass ? P.mass : A.mass[i];
  f3 F = A.force_init ? f3{A.fx[i], A.fy[i], A.fz[i]} : f3{0.0f, 0.0f, 0.0f};
  F = add(F, P.gravity);
  const float xc = X(c), yr = Y(r);
  auto strut = [&](int rr, int cc, float kn, float cn) {
    const float dX = xc - X(cc), dY = yr - Y(rr);
    const float L = sqrt_rn(dX * dX + dY * dY);
    const float inv = rcp_rn(L);  // Strut::new: NOMINAL_STRUT_LENGTH / length (springy_mesh.rs:83-84)
    const uint32_t j = id(rr, cc);
    strut_accumulate<M>(F, p, v, ld3(S.x, S.y, S.z, j), ld3(S.vx, S.vy, S.vz, j), inv * kn, inv * cn, L);
  };
  if (r >= 1 && c <= C - 2) { strut(r - 1, c + 1, G.sk, G.sc); strut(r - 1, c, G.tk, G.tc); }
  if (r <= R - 2 && c >= 1) { strut(r + 1, c - 1, G.sk, G.sc); strut(r, c - 1, G.tk, G.tc); }
  if (r <= R - 2 && c <= C - 2) { strut(r + 1, c, G.tk, G.tc); strut(r, c + 1, G.tk, G.tc); }
  if (c == C - 1 && r >= 1) strut(r - 1, c, G.tk, G.tc);
  if (r == R - 1 && c >= 1) strut(r, c - 1, G.tk, G.tc);
  if (c == C - 1 && r <= R - 2) strut(r + 1, c, G.tk, G.tc);
  if (r == R - 1 && c <= C - 2) strut(r, c + 1, G.tk, G.tc);
  if (r >= 2 && c <= C - 3) strut(r - 2, c, G.bk, G.bc);
  if (c >= 2 && r <= R - 3) strut(r, c - 2, G.bk, G.bc);
  if (r <= R - 3 && c <= C - 3) { strut(r, c + 2, G.bk, G.bc); strut(r + 2, c, G.bk, G.bc); }

  auto P3 = [&](int rr, int cc) { return ld3(S.x, S.y, S.z, id(rr, cc)); };
  auto V3 = [&](int rr, int cc) { return ld3(S.vx, S.vy, S.vz, id(rr, cc)); };
  // corner 0/1/2 of face (q0, q1, q2): weight = angle at that corner / PI
  auto face = [&](int r0, int c0, int r1, int c1, int r2, int c2, int corner) {
    const f3 p0 = P3(r0, c0), p1 = P3(r1, c1), p2 = P3(r2, c2);
    const f3 sum = face_force<M>(p0, p1, p2, V3(r0, c0), V3(r1, c1), V3(r2, c2), P);
    const float w = corner == 0 ? corner_weight<M>(p0, p1, p2)
                                : (corner == 1 ? corner_weight<M>(p1, p2, p0) : corner_weight<M>(p2, p0, p1));
    F = add(F, f3{w * sum.x, w * sum.y, w * sum.z});
  };
  if (r >= 1 && c <= C - 2) face(r - 1, c + 1, r, c, r - 1, c, 1);          // A(r-1, c)
  if (r <= R - 2 && c >= 1) face(r, c, r + 1, c - 1, r, c - 1, 0);          // A(r, c-1)
  if (r <= R - 2 && c <= C - 2) face(r, c + 1, r + 1, c, r, c, 2);          // A(r, c)
  if (r >= 1 && c >= 1) face(r - 1, c, r, c, r, c - 1, 1);                  // B(r-1, c-1)
  if (r >= 1 && c <= C - 2) face(r - 1, c + 1, r, c + 1, r, c, 2);          // B(r-1, c)
  if (r <= R - 2 && c >= 1) face(r, c, r + 1, c, r + 1, c - 1, 0);          // B(r, c-1)

  const bool pin = pinned(A.pins, i);
  if (pin) F = f3{0.0f, 0.0f, 0.0f};
  if (MODE == MODE_FORCES) {
    A.fx[i] = F.x; A.fy[i] = F.y; A.fz[i] = F.z;
    return;
  }
  finish_step<M>(A, P, i, m, p, v, F, pin);
}

// SpringyMesh::update_points (springy_mesh.rs:423-497) for records supplied by the caller.
__global__ void __launch_bounds__(kBlock) update_points_kernel(StepArgs A, const float* __restrict__ rec,
                                                                uint32_t rec_base, float* __restrict__ mass_out,
                                                                const BlockRange* __restrict__ blocks) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  if (i >= br.end) return;
  const StepParams P = A.params[br.mesh];
  const float* q = rec + 10ull * (i - rec_base);
  const float m_old = P.uniform_mass ? P.mass : A.mass[i];
  const f3 p = ld3(A.in.x, A.in.y, A.in.z, i), v = ld3(A.in.vx, A.in.vy, A.in.vz, i);
  const f3 F = f3{A.fx[i], A.fy[i], A.fz[i]};
  f3 xn = f3{q[1], q[2], q[3]}, vn = f3{q[4], q[5], q[6]};
  float mn = q[0];
  f3 Fn = f3{q[7], q[8], q[9]};
  const int32_t hit = first_hit(A.obs, p, xn, P.dt_d);
  if (hit >= 0) collide_response(A.obs.faces[hit], P, m_old, p, v, F, xn, vn);
  if (pinned(A.pins, i)) { xn = p; vn = v; mn = m_old; Fn = F; }
  store_state(A.in, i, xn, vn);
  mass_out[i] = mn;
  A.fx[i] = Fn.x; A.fy[i] = Fn.y; A.fz[i] = Fn.z;
}

// The Stateful records [m, x, y, z, vx, vy, vz, Fx, Fy, Fz] (springy_mesh.rs:242-255) of device
// vertices [base, base + n) in the reference's vertex order: record i from device slot
// base + pos[i] (base + i without a permutation).
__global__ void __launch_bounds__(kBlock) pack_records_kernel(SoA S, const float* __restrict__ mass,
                                                               const float* __restrict__ fx, const float* __restrict__ fy,
                                                               const float* __restrict__ fz,
                                                               const uint32_t* __restrict__ pos, uint32_t base,
                                                               uint32_t n, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = base + (pos != nullptr ? pos[i] : i);
  float* o = out + 10ull * i;
  o[0] = mass[d];
  o[1] = S.x[d]; o[2] = S.y[d]; o[3] = S.z[d];
  o[4] = S.vx[d]; o[5] = S.vy[d]; o[6] = S.vz[d];
  o[7] = fx[d]; o[8] = fy[d]; o[9] = fz[d];
}

// The positions of device vertices [base, base + n) in the reference's vertex order, 3 floats each.
__global__ void __launch_bounds__(kBlock) pack_positions_kernel(SoA S, const uint32_t* __restrict__ pos, uint32_t base,
                                                                 uint32_t n, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = base + (pos != nullptr ? pos[i] : i);
  out[3ull * i] = S.x[d]; out[3ull * i + 1] = S.y[d]; out[3ull * i + 2] = S.z[d];
}

// The inverse of pack_records_kernel: records in the reference's vertex order into the state,
// masses and force records of device vertices [base, base + n).
__global__ void __launch_bounds__(kBlock) unpack_records_kernel(const float* __restrict__ in, const uint32_t* __restrict__ pos,
                                                                 uint32_t base, uint32_t n, SoA S, float* __restrict__ mass,
                                                                 float* __restrict__ fx, float* __restrict__ fy,
                                                                 float* __restrict__ fz) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = base + (pos != nullptr ? pos[i] : i);
  const float* q = in + 10ull * i;
  mass[d] = q[0];
  S.x[d] = q[1]; S.y[d] = q[2]; S.z[d] = q[3];
  S.vx[d] = q[4]; S.vy[d] = q[5]; S.vz[d] = q[6];
  fx[d] = q[7]; fy[d] = q[8]; fz[d] = q[9];
}

// Per mesh, the number of vertices whose position or velocity is not finite.
__global__ void __launch_bounds__(kBlock) nonfinite_kernel(SoA S, const BlockRange* __restrict__ blocks,
                                                            unsigned long long* __restrict__ counts) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  bool bad = false;
  if (i < br.end)
    bad = !(__builtin_isfinite(S.x[i]) && __builtin_isfinite(S.y[i]) && __builtin_isfinite(S.z[i]) &&
            __builtin_isfinite(S.vx[i]) && __builtin_isfinite(S.vy[i]) && __builtin_isfinite(S.vz[i]));
  const int c = __syncthreads_count(bad);
  if (threadIdx.x == 0 && c > 0) atomicAdd(&counts[br.mesh], (unsigned long long)c);
}

__global__ void fill_kernel(float* p, uint64_t n, float v) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
    p[k] = v;
}

}  // namespace

hipError_t launch_csr(const StepArgs& a, const CsrView& t, const BlockRange* blocks, uint32_t n_blocks, int mode,
                      bool exact, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  dim3 g(n_blocks), b(kBlock);
  if (exact) {
    if (mode == MODE_STEP) csr_kernel<ExactMath, MODE_STEP><<<g, b, 0, s>>>(a, t, blocks);
    else csr_kernel<ExactMath, MODE_FORCES><<<g, b, 0, s>>>(a, t, blocks);
  } else {
    if (mode == MODE_STEP) csr_kernel<FastMath, MODE_STEP><<<g, b, 0, s>>>(a, t, blocks);
    else csr_kernel<FastMath, MODE_FORCES><<<g, b, 0, s>>>(a, t, blocks);
  }
  return hipGetLastError();
}

// Every finite f32 x (grid-stride over all bit patterns): the exact kernels' shortened forms
// against the operations they stand for, bad[0] rcp_rn(x) vs 1 / x, bad[1] div_pi_rn(x) vs x / PI,
// bad[2] div3_rn(x) vs x / 3, bad[3] the branch-free atan2 vs fdlibm's control flow at y = 1 and
// y = -1.7, bad[4] sqrt_rn(x) vs the compiler's correctly rounded sqrt (NaN payloads aside).
__global__ void __launch_bounds__(kBlock) selftest_kernel(unsigned long long* __restrict__ bad) {
  unsigned long long b[5] = {0, 0, 0, 0, 0};
  auto bits = [](float f) { return __float_as_uint(f); };
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < (1ull << 32); k += (uint64_t)gridDim.x * kBlock) {
    const float x = __uint_as_float((uint32_t)k);
    if (!__builtin_isfinite(x)) continue;
    b[0] += bits(rcp_rn(x)) != bits(1.0f / x);
    b[1] += bits(div_pi_rn(x)) != bits(x / 3.14159274101257324f);
    b[2] += bits(div3_rn(x)) != bits(x / 3.0f);
    const float ys[2] = {1.0f, -1.7f};
    for (int j = 0; j < 2; ++j) {
      const float a = fdlibm_atan2f(ys[j], x), r = fdlibm_atan2f_ref(ys[j], x);
      b[3] += bits(a) != bits(r) && !(__builtin_isnan(a) && __builtin_isnan(r));
    }
    const float sq = sqrt_rn(x), sr = __builtin_sqrtf(x);
    b[4] += bits(sq) != bits(sr) && !(__builtin_isnan(sq) && __builtin_isnan(sr));
  }
  for (int q = 0; q < 5; ++q)
    if (b[q]) atomicAdd(&bad[q], b[q]);
}

hipError_t launch_selftest(unsigned long long* bad, hipStream_t s) {
  selftest_kernel<<<dim3(4096), dim3(kBlock), 0, s>>>(bad);
  return hipGetLastError();
}

hipError_t launch_patches(const StepArgs& a, const CsrView& t, const PatchRange* patches, uint32_t n_patches,
                          uint32_t nf_cap, bool exact, hipStream_t s) {
  if (n_patches == 0) return hipSuccess;
  if (nf_cap > kPatchFaceCap) return hipErrorInvalidValue;
  const size_t lds = sizeof(float4) * (size_t)nf_cap;
  if (exact) patch_kernel<ExactMath><<<dim3(n_patches), dim3(kBlock), lds, s>>>(a, t, patches);
  else patch_kernel<FastMath><<<dim3(n_patches), dim3(kBlock), lds, s>>>(a, t, patches);
  return hipGetLastError();
}

hipError_t launch_faces(const StepArgs& a, const CsrView& t, const BlockRange* face_blocks, uint32_t n_blocks,
                        bool exact, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  if (exact) face_kernel<ExactMath><<<dim3(n_blocks), dim3(kBlock), 0, s>>>(a, t, face_blocks);
  else face_kernel<FastMath><<<dim3(n_blocks), dim3(kBlock), 0, s>>>(a, t, face_blocks);
  return hipGetLastError();
}

hipError_t launch_grid(const StepArgs& a, const ClothView* cloths, const BlockRange* blocks, uint32_t n_blocks,
                       int mode, bool exact, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  dim3 g(n_blocks), b(kBlock);
  if (exact) {
    if (mode == MODE_STEP) grid_kernel<ExactMath, MODE_STEP><<<g, b, 0, s>>>(a, cloths, blocks);
    else grid_kernel<ExactMath, MODE_FORCES><<<g, b, 0, s>>>(a, cloths, blocks);
  } else {
    if (mode == MODE_STEP) grid_kernel<FastMath, MODE_STEP><<<g, b, 0, s>>>(a, cloths, blocks);
    else grid_kernel<FastMath, MODE_FORCES><<<g, b, 0, s>>>(a, cloths, blocks);
  }
  return hipGetLastError();
}

hipError_t launch_update_points(const StepArgs& a, const float* nw_rec, uint32_t rec_base, float* mass_out,
                                const BlockRange* blocks, uint32_t n_blocks, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  update_points_kernel<<<dim3(n_blocks), dim3(kBlock), 0, s>>>(a, nw_rec, rec_base, mass_out, blocks);
  return hipGetLastError();
}

hipError_t launch_pack_records(const SoA& st, const float* mass, const float* fx, const float* fy, const float* fz,
                               const uint32_t* pos, uint32_t base, uint32_t n, float* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  pack_records_kernel<<<dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s>>>(st, mass, fx, fy, fz, pos, base, n,
                                                                               out);
  return hipGetLastError();
}

hipError_t launch_pack_positions(const SoA& st, const uint32_t* pos, uint32_t base, uint32_t n, float* out,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  pack_positions_kernel<<<dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s>>>(st, pos, base, n, out);
  return hipGetLastError();
}

hipError_t launch_unpack_records(const float* in, const uint32_t* pos, uint32_t base, uint32_t n, const SoA& st,
                                 float* mass, float* fx, float* fy, float* fz, hipStream_t s) {
  if (n == 0) return hipSuccess;
  unpack_records_kernel<<<dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s>>>(in, pos, base, n, st, mass, fx, fy,
                                                                                 fz);
  return hipGetLastError();
}

hipError_t launch_nonfinite(const SoA& s_in, const BlockRange* blocks, uint32_t n_blocks, unsigned long long* counts,
                            hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  nonfinite_kernel<<<dim3(n_blocks), dim3(kBlock), 0, s>>>(s_in, blocks, counts);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, uint64_t n, float v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  fill_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_hinges(const StepArgs& a, const CsrView& t, const HingeEnt* hinges, uint32_t first, uint32_t count,
                         bool exact, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const dim3 g((count + kBlock - 1) / kBlock);
  if (exact)
    hipLaunchKernelGGL(hinge_kernel<ExactMath>, g, dim3(kBlock), 0, s, a, t, hinges, first, count);
  else
    hipLaunchKernelGGL(hinge_kernel<FastMath>, g, dim3(kBlock), 0, s, a, t, hinges, first, count);
  return hipGetLastError();
}

hipError_t launch_torsion_gather(const StepArgs& a, const CsrView& t, const BlockRange* blocks, uint32_t n_blocks,
                                 hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(torsion_gather_kernel, dim3(n_blocks), dim3(kBlock), 0, s, a, t, blocks);
  return hipGetLastError();
}

}  // namespace fp
