// LDS-tiled step kernel for meshes built by Cloth::new (springy/cloth.rs:20-113).
//
// One 256-thread workgroup owns a TR x TC tile of the grid:
//   phase 0  stage x, y, z, vx, vy, vz of the tile plus a 2-vertex halo (bending struts reach two
//            vertices) from HBM into LDS, coalesced along rows;
//   phase 1  compute, once per tile, every face's drag + lift (each face is shared by three
//            vertices) and the strut constants (rest length, stiffness, damping) per column pair,
//            row pair and quad; a Cloth::new grid's rest lengths depend only on those;
//   phase 2  every vertex gathers its struts (from LDS state and the constant tables) and its faces
//            (phase-1 force times its own corner's angle / PI) in the reference's order,
//            integrates, resolves collision and pins, and writes its new state to HBM.
// The per-vertex operations and their order are those of grid_kernel, so the EXACT form returns
// the same bits as the reference.  HBM traffic: the state read once (plus halo rows that the L2
// serves) and written once.
#include "kernels.hpp"

namespace fp {

namespace {

// Strut constants (Strut::new, springy_mesh.rs:74-90) of a strut whose ends differ by dX, dY in the
// construction grid: rest length |d| = sqrt(dX^2 + dY^2) (the third component is exactly 0),
// stiffness and damping = nominal / length.
struct StrutK {
  float k, c, L;
};
__device__ __forceinline__ StrutK strut_k(float dX, float dY, float kn, float cn) {
  const float L = sqrt_rn(dX * dX + dY * dY);
  const float inv = rcp_rn(L);
  return StrutK{inv * kn, inv * cn, L};
}

// The LDS of one TR x TC tile: state tile + halo (rows r0-2 .. r0+TR+1, cols c0-2 .. c0+TC+1), the
// drag + lift of the faces of quads (r', c') with r' in [r0-1, r0+TR), c' in [c0-1, c0+TC), and the
// strut constant tables.  A staged row holds its six fields one after another (field q of column c
// at q * (TC + 4) + c) and rows are kStRow words apart, kStRow = 16 (mod 64): the four rows of
// sixteen columns a wavefront reads in the per-vertex pass, shifted by any neighbour offset, then
// fall on distinct LDS banks.
template <int TR, int TC>
struct TileLds {
  static constexpr int kStRow = (6 * (TC + 4) - 16 + 63) / 64 * 64 + 16;
  float st[TR + 4][kStRow];
  float face[2][TR + 1][TC + 1][3];
  StrutK shear[TR + 1][TC + 1];
  StrutK h[TC + 1], v[TR + 1], bh[TC + 2], bv[TR + 2];
};

// One tile's step by one workgroup (all of its threads call it).
template <class M, int TR, int TC, int MODE>
__device__ __forceinline__ void tile_step(TileLds<TR, TC>& L, const StepArgs& A, const ClothView* __restrict__ cloths,
                                          const TileRange T) {
  constexpr int TH = TR + 4, TW = TC + 4;
  constexpr int QR = TR + 1, QC = TC + 1;
  auto& s_st = L.st;
  auto& s_face = L.face;
  auto& s_shear = L.shear;
  auto& s_h = L.h;
  auto& s_v = L.v;
  auto& s_bh = L.bh;
  auto& s_bv = L.bv;

  const StepParams P = A.params[T.mesh];
  const ClothView G = cloths[T.mesh];
  const int R = (int)G.rows, C = (int)G.cols;
  const int r0 = (int)T.r0, c0 = (int)T.c0;
  const int tid = (int)threadIdx.x;
  auto X = [&](int cc) -> float { return ((float)cc * G.spacing - G.cx) + G.p0; };
  auto Y = [&](int rr) -> float { return ((float)rr * G.spacing - G.cy) + G.p1; };

  // ---- phase 0: state tile + halo -> LDS
  const float* src[6] = {A.in.x, A.in.y, A.in.z, A.in.vx, A.in.vy, A.in.vz};
  for (int e = tid; e < TH * TW; e += kBlock) {
    const int lr = e / TW, lc = e - lr * TW;
    const int r = r0 - 2 + lr, c = c0 - 2 + lc;
    if (r < 0 || r >= R || c < 0 || c >= C) continue;
    const uint32_t g = G.base + (uint32_t)(r * C + c);
#pragma unroll
    for (int q = 0; q < 6; ++q) s_st[lr][q * TW + lc] = src[q][g];
  }
  // strut constant tables (construction grid only: no state needed)
  for (int j = tid; j < TC + 1; j += kBlock) {  // (c0-1+j, c0+j)
    const int c = c0 - 1 + j;
    s_h[j] = strut_k(X(c) - X(c + 1), 0.0f, G.tk, G.tc);
  }
  for (int j = tid; j < TR + 1; j += kBlock) {  // rows (r0-1+j, r0+j)
    const int r = r0 - 1 + j;
    s_v[j] = strut_k(0.0f, Y(r) - Y(r + 1), G.tk, G.tc);
  }
  for (int j = tid; j < TC + 2; j += kBlock) {  // (c0-2+j, c0+j)
    const int c = c0 - 2 + j;
    s_bh[j] = strut_k(X(c) - X(c + 2), 0.0f, G.bk, G.bc);
  }
  for (int j = tid; j < TR + 2; j += kBlock) {  // rows (r0-2+j, r0+j)
    const int r = r0 - 2 + j;
    s_bv[j] = strut_k(0.0f, Y(r) - Y(r + 2), G.bk, G.bc);
  }
  for (int e = tid; e < QR * QC; e += kBlock) {  // anti-diagonal (r', c'+1)-(r'+1, c') of quad (r', c')
    const int i = e / QC, j = e - i * QC;
    const int r = r0 - 1 + i, c = c0 - 1 + j;
    s_shear[i][j] = strut_k(X(c + 1) - X(c), Y(r) - Y(r + 1), G.sk, G.sc);
  }
  __syncthreads();

  auto Pl = [&](int r, int c) {
    const float* row = s_st[r - r0 + 2] + (c - c0 + 2);
    return f3{row[0], row[TW], row[2 * TW]};
  };
  auto Vl = [&](int r, int c) {
    const float* row = s_st[r - r0 + 2] + (c - c0 + 2);
    return f3{row[3 * TW], row[4 * TW], row[5 * TW]};
  };

  // ---- phase 1: drag + lift of every face that touches the tile, once per face
  // A(q) = [(r',c'+1), (r'+1,c'), (r',c')], B(q) = [(r',c'+1), (r'+1,c'+1), (r'+1,c')]
  for (int e = tid; e < QR * QC; e += kBlock) {
    const int i = e / QC, j = e - i * QC;
    const int r = r0 - 1 + i, c = c0 - 1 + j;
    if (r < 0 || r > R - 2 || c < 0 || c > C - 2) continue;
    const f3 pa = Pl(r, c + 1), pb = Pl(r + 1, c), pc = Pl(r, c), pd = Pl(r + 1, c + 1);
    const f3 va = Vl(r, c + 1), vb = Vl(r + 1, c), vc = Vl(r, c), vd = Vl(r + 1, c + 1);
    const f3 fa = face_force<M>(pa, pb, pc, va, vb, vc, P);
    const f3 fb = face_force<M>(pa, pd, pb, va, vd, vb, P);
    s_face[0][i][j][0] = fa.x; s_face[0][i][j][1] = fa.y; s_face[0][i][j][2] = fa.z;
    s_face[1][i][j][0] = fb.x; s_face[1][i][j][1] = fb.y; s_face[1][i][j][2] = fb.z;
  }
  __syncthreads();

  // ---- phase 2: per vertex
  for (int e = tid; e < TR * TC; e += kBlock) {
    const int lr = e / TC, lc = e - lr * TC;
    const int r = r0 + lr, c = c0 + lc;
    if (r >= R || c >= C) continue;
    const uint32_t i = G.base + (uint32_t)(r * C + c);
    const f3 p = Pl(r, c), v = Vl(r, c);
    const float m = P.uniform_mass ? P.mass : A.mass[i];
    // MODE_SURROGATE: F collects only the face forces (from zero); gravity and the network's
    // acceleration enter in sur_update
    f3 F = (MODE != MODE_SURROGATE && A.force_init) ? f3{A.fx[i], A.fy[i], A.fz[i]} : f3{0.0f, 0.0f, 0.0f};
    if (MODE != MODE_SURROGATE) F = add(F, P.gravity);
    auto strut = [&](int rr, int cc, const StrutK& K) {
      if (MODE != MODE_SURROGATE) strut_accumulate<M>(F, p, v, Pl(rr, cc), Vl(rr, cc), K.k, K.c, K.L);
    };
    const int qi = r - r0, qj = c - c0;  // quad (r-1, c-1) is s_*[qi][qj]
    // struts in strut-index order (see grid_kernel in kernels.hip)
    if (r >= 1 && c <= C - 2) { strut(r - 1, c + 1, s_shear[qi][qj + 1]); strut(r - 1, c, s_v[qi]); }
    if (r <= R - 2 && c >= 1) { strut(r + 1, c - 1, s_shear[qi + 1][qj]); strut(r, c - 1, s_h[qj]); }
    if (r <= R - 2 && c <= C - 2) { strut(r + 1, c, s_v[qi + 1]); strut(r, c + 1, s_h[qj + 1]); }
    if (c == C - 1 && r >= 1) strut(r - 1, c, s_v[qi]);
    if (r == R - 1 && c >= 1) strut(r, c - 1, s_h[qj]);
    if (c == C - 1 && r <= R - 2) strut(r + 1, c, s_v[qi + 1]);
    if (r == R - 1 && c <= C - 2) strut(r, c + 1, s_h[qj + 1]);
    if (r >= 2 && c <= C - 3) strut(r - 2, c, s_bv[qi]);
    if (c >= 2 && r <= R - 3) strut(r, c - 2, s_bh[qj]);
    if (r <= R - 3 && c <= C - 3) { strut(r, c + 2, s_bh[qj + 2]); strut(r + 2, c, s_bv[qi + 2]); }
    // faces in face-index order: A(r-1,c) A(r,c-1) A(r,c) B(r-1,c-1) B(r-1,c) B(r,c-1)
    // this vertex's corner weight (angle / PI) times the face's drag + lift
    auto face = [&](int fs, int i2, int j2, float w) {
      F = add(F, f3{w * s_face[fs][i2][j2][0], w * s_face[fs][i2][j2][1], w * s_face[fs][i2][j2][2]});
    };
    if (r >= 1 && c <= C - 2) face(0, qi, qj + 1, corner_weight<M>(p, Pl(r - 1, c), Pl(r - 1, c + 1)));
    if (r <= R - 2 && c >= 1) face(0, qi + 1, qj, corner_weight<M>(p, Pl(r + 1, c - 1), Pl(r, c - 1)));
    if (r <= R - 2 && c <= C - 2) face(0, qi + 1, qj + 1, corner_weight<M>(p, Pl(r, c + 1), Pl(r + 1, c)));
    if (r >= 1 && c >= 1) face(1, qi, qj, corner_weight<M>(p, Pl(r, c - 1), Pl(r - 1, c)));
    if (r >= 1 && c <= C - 2) face(1, qi, qj + 1, corner_weight<M>(p, Pl(r - 1, c + 1), Pl(r, c + 1)));
    if (r <= R - 2 && c >= 1) face(1, qi + 1, qj, corner_weight<M>(p, Pl(r + 1, c), Pl(r + 1, c - 1)));
    const bool pin = pinned(A.pins, i);
    if (MODE == MODE_SURROGATE) {
      f3 xn = p, vn = v;
      float w = 0.0f;
      if (!pin)
        w = sur_update(P, A.obs, m, p, v, F, f3{0.0f, 0.0f, 0.0f},
                       f3{A.da[3 * (size_t)i], A.da[3 * (size_t)i + 1], A.da[3 * (size_t)i + 2]}, xn, vn);
      A.out.x[i] = xn.x; A.out.y[i] = xn.y; A.out.z[i] = xn.z;
      A.out.vx[i] = vn.x; A.out.vy[i] = vn.y; A.out.vz[i] = vn.z;
      A.work[i] = w;
      continue;
    }
    if (pin) F = f3{0.0f, 0.0f, 0.0f};
    if (MODE == MODE_FORCES) {
      A.fx[i] = F.x; A.fy[i] = F.y; A.fz[i] = F.z;
      continue;
    }
    f3 xn, vn;
    integrate<M>(P, m, p, v, F, xn, vn);
    const int32_t hit = first_hit(A.obs, p, xn, P.dt_d);
    if (hit >= 0) collide_response(A.obs.faces[hit], P, m, p, v, F, xn, vn);
    if (pin) { xn = p; vn = v; }
    A.out.x[i] = xn.x; A.out.y[i] = xn.y; A.out.z[i] = xn.z;
    A.out.vx[i] = vn.x; A.out.vy[i] = vn.y; A.out.vz[i] = vn.z;
    if (A.force_init) { A.fx[i] = 0.0f; A.fy[i] = 0.0f; A.fz[i] = 0.0f; }
  }
}

template <class M, int TR, int TC, int MODE>
__global__ void __launch_bounds__(kBlock) grid_tiled_kernel(StepArgs A, const ClothView* __restrict__ cloths,
                                                             const TileRange* __restrict__ tiles) {
  __shared__ TileLds<TR, TC> L;
  tile_step<M, TR, TC, MODE>(L, A, cloths, tiles[blockIdx.x]);
}

template <class M, int TR, int TC>
hipError_t launch_shape(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n, int mode,
                        hipStream_t s) {
  if (mode == MODE_STEP) grid_tiled_kernel<M, TR, TC, MODE_STEP><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles);
  else if (mode == MODE_FORCES) grid_tiled_kernel<M, TR, TC, MODE_FORCES><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles);
  else grid_tiled_kernel<M, TR, TC, MODE_SURROGATE><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles);
  return hipGetLastError();
}

template <class M>
hipError_t launch_math(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n, int shape,
                       int mode, hipStream_t s) {
  switch (shape) {
    case TILE_15x15: return launch_shape<M, 15, 15>(a, cloths, tiles, n, mode, s);
    case TILE_7x31: return launch_shape<M, 7, 31>(a, cloths, tiles, n, mode, s);
    case TILE_16x16: return launch_shape<M, 16, 16>(a, cloths, tiles, n, mode, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_grid_tiled(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n_tiles,
                             int shape, int mode, bool exact, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  return exact ? launch_math<ExactMath>(a, cloths, tiles, n_tiles, shape, mode, s)
               : launch_math<FastMath>(a, cloths, tiles, n_tiles, shape, mode, s);
}

}  // namespace fp
