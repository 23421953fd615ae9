// The fast-math tiled step for meshes built by Cloth::new (springy/cloth.rs:20-113): the work of
// one TR x TC tile, split into the three phases one workgroup runs between its barriers.
//
// Built around its own instruction budget rather than the reference's loop structure: every strut
// and every face touching the tile is evaluated ONCE (the reference's scatter loops,
// springy_mesh.rs:533-551 and 629-654, visit each once too; the exact kernel instead has each
// vertex gather its struts, evaluating each strut from both ends), and each face's corner weights
// use two angles, the third being what is left of PI.
//
// Work decomposition: the tile's (TR+1) x (TC+1) = 256 quads, one per thread.  Thread (i, j) owns
// quad (R', C') = (r0-1+i, c0-1+j) with corners A = (R', C'), B = (R', C'+1), Cc = (R'+1, C'),
// D = (R'+1, C'+1), and evaluates
//   top     the row strut A -> B              (i >= 1; for i == 0 the bend strut (r0-2, C') -> Cc)
//   left    the column strut A -> Cc          (j >= 1; for j == 0 the bend strut (R', c0-2) -> B)
//   shear   the anti-diagonal B -> Cc
//   bh, bv  the bend struts A -> (R', C'+2) and A -> (R'+2, C')
//   faces   (B, Cc, A) and (B, D, Cc), the reference's two triangles of the quad
// which are exactly the struts and faces with an end or corner in the tile (the two substituted
// bend struts are the row r0-2 and column c0-2 ones no quad of the tile starts).  Each strut's
// force on its first end is f, on its second -f.  The thread sums what it owes each corner into
// four vectors (to_A .. to_D) plus the far ends of its two bend struts, and phase 2 gives every
// tile vertex the six vectors addressed to it.  Summation order is fixed, so runs are
// deterministic; it is not the reference's order (the exact kernel keeps that order).
//
// Tiles whose staged border lies inside the grid (all but the cloth's outer ring of tiles) take an
// INTERIOR instantiation with no bounds checks and no validity selects: every strut and face their
// threads evaluate exists, except thread (0, 0)'s two substituted bend struts, whose results land
// only in slots that no vertex of the tile reads.
//
// Arithmetic: a strut's length |d| is the reference's (unfused dot, correctly rounded square root:
// |d| - L cancels to the strain, and one ulp of |d| is already k ulp(L) of force); the rest uses
// fused multiply-adds and the hardware reciprocal, reciprocal square root and square root; corner
// angles come from atan2(|cross|, dot) by a degree-11 odd polynomial for atan(a)/PI on [0, 1]
// (least-squares fit refined towards minimax, max error 5.6e-7 in the weight).
#pragma once

#include "kernels.hpp"

#ifndef FP_LDS_CHECK
#define FP_LDS_CHECK(cond) ((void)0)
#endif

namespace fp {
namespace gfast {

FP_HD float rcp_f(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
FP_HD float rsq_f(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}
FP_HD float sqrt_f(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
// The correctly rounded square root for every input but subnormals (where it is within one ulp):
// the hardware square root and the choice among it and its two neighbours by the exact residuals
// (sqrt_rn of exact_ops.hpp without its range check, which only subnormal and huge inputs need;
// zero, infinity and NaN come out right through the same selects).
FP_HD float sqrt_cr(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u), up = __uint_as_float(__float_as_uint(s) + 1u);
  const float rdn = __builtin_fmaf(-dn, s, x), rup = __builtin_fmaf(-up, s, x);
  const float t = rdn <= 0.0f ? dn : s;
  return rup > 0.0f ? up : t;
#else
  return sqrtf(x);
#endif
}
FP_HD float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
FP_HD f3 fsub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FP_HD float fdot(f3 a, f3 b) { return fma_(a.z, b.z, fma_(a.y, b.y, a.x * b.x)); }
FP_HD f3 fcross(f3 a, f3 b) {
  return f3{fma_(a.y, b.z, -(a.z * b.y)), fma_(a.z, b.x, -(a.x * b.z)), fma_(a.x, b.y, -(a.y * b.x))};
}
FP_HD f3 fmad3(float s, f3 v, f3 acc) { return f3{fma_(s, v.x, acc.x), fma_(s, v.y, acc.y), fma_(s, v.z, acc.z)}; }
FP_HD f3 sel3(bool c, f3 a, f3 b) { return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }

// Weight angle / PI of a triangle corner with |cross| = cm > 0 and dot = d: atan2(cm, d) / PI.
FP_HD float corner_w(float cm, float d) {
  const float ad = fabsf(d);
  const float mx = fmaxf(ad, cm), mn = fminf(ad, cm);
  const float a = mn * rcp_f(mx);
  const float t = a * a;
  float p = -0.0037303001154214144f;
  p = fma_(p, t, 0.016758132725954056f);
  p = fma_(p, t, -0.037059664726257324f);
  p = fma_(p, t, 0.06160580366849899f);
  p = fma_(p, t, -0.10587713122367859f);
  p = fma_(p, t, 0.31830263137817383f);
  p = p * a;                                     // atan(a) / PI, in [0, 1/4]
  const float w0 = ad > cm ? p : 0.5f - p;       // the angle for d >= 0
  return d < 0.0f ? 1.0f - w0 : w0;
}

// Force of one strut on its first end p0 (springy_mesh.rs:535-549): the spring k (|d| - L) and
// damping c ((v1 - v0) . u) along u = d / |d|, d = p1 - p0.
FP_HD f3 strut_f(f3 p0, f3 v0, f3 p1, f3 v1, float k, float c, float L) {
  const f3 d = fsub(p1, p0);
  const float len = sqrt_cr((d.x * d.x + d.y * d.y) + d.z * d.z);
  const float inv = rcp_f(len);
  const float s = k * (len - L);
  const float dvd = fdot(fsub(v1, v0), d);
  const float g = fma_(c * inv, dvd, s) * inv;
  return f3{g * d.x, g * d.y, g * d.z};
}

// Drag + lift of a face from its cross product cr = (p1 - p0) x (p2 - p0), 1 / |cr| and its
// relative velocity vr = (u0 + u1 + u2) / 3 - wind (springy_mesh.rs:631-646):
// area |n . vr| = |cr . vr| / 2 and |n x vr| = |cr x vr| / |cr|.
FP_HD f3 face_f(f3 cr, float rs, f3 vr, float neg_drag, float neg_lift) {
  const float crv = fdot(cr, vr);
  const f3 cx = fcross(cr, vr);
  const float lm = sqrt_f(fdot(cx, cx)) * rs;
  const float s = (0.5f * fabsf(crv)) * fma_(neg_lift, lm, neg_drag);
  return f3{s * vr.x, s * vr.y, s * vr.z};
}

// Strut constants of one table entry (Strut::new, springy_mesh.rs:74-90, for a Cloth::new grid:
// the rest length depends only on the two construction-grid coordinates' difference).
struct KTab {
  float d, L, k, c;  // d: the coordinate difference (for the shear struts' lengths)
};

template <int TR, int TC>
struct Lds {
  static constexpr int SH = TR + 4, SW = TC + 4;  // staged rows [r0-2, r0+TR+2), cols [c0-2, c0+TC+2)
  static constexpr int QR = TR + 1, QC = TC + 1;  // quads = threads
  static constexpr int NQ = QR * QC;
  static constexpr int NV = TR * TC;              // tile vertices
  // strut tables, one array: row struts by first column c0-1+j (at H0), column struts by first row
  // r0-1+i (V0), row bends by first column c0-2+j (BH0), column bends by first row r0-2+i (BV0)
  static constexpr int H0 = 0, V0 = QC, BH0 = QC + QR, BV0 = 2 * QC + QR + 1, NTAB = 2 * (QC + QR) + 2;
  float4 pv[SH * SW];  // x, y, z, vx
  float2 vw[SH * SW];  // vy, vz
  KTab tab[NTAB];
  // what the quads owe the tile vertices, by recipient (tile-local vertex (a, b) at a * TC + b):
  // from the quad with the vertex as corner B, as corner Cc, as corner D, and the far ends of the
  // row and column bends reaching it (to_A stays in its thread's registers)
  float out[5][3][NV];
};

// Whether a tile's staged region (the tile and a 2-vertex border) lies inside its grid.
template <int TR, int TC>
FP_HD bool interior_tile(const ClothView& G, int r0, int c0) {
  return r0 >= 2 && c0 >= 2 && r0 + TR + 2 <= (int)G.rows && c0 + TC + 2 <= (int)G.cols;
}

FP_HD float grid_x(const ClothView& G, int c) { return ((float)c * G.spacing - G.cx) + G.p0; }
FP_HD float grid_y(const ClothView& G, int r) { return ((float)r * G.spacing - G.cy) + G.p1; }

// A strut along one grid axis: the reference's rest length sqrt(d * d + 0 * 0) is |d| exactly (a
// correctly rounded square root of a correctly rounded square returns |d| in binary floating point).
FP_HD KTab ktab(float d, float kn, float cn) {
  const float L = fabsf(d);
  const float inv = rcp_f(L);
  return KTab{d, L, inv * kn, inv * cn};
}

// Phase 0: stage the state of the tile and its 2-vertex border; the strut constant tables.
template <int TR, int TC, bool INTERIOR>
FP_HD void phase0(Lds<TR, TC>& L, const StepArgs& A, const ClothView& G, int r0, int c0, int tid, int nthreads) {
  using S = Lds<TR, TC>;
  const int R = (int)G.rows, C = (int)G.cols;
  // The staged region has more entries than the workgroup has threads (361 or 385 for 256).  Every
  // round's loads are issued, unconditionally, before the first LDS store, so that no thread waits
  // for two HBM round trips in a row: a lane past the region's end loads its last entry again, and
  // on a tile at the grid's edge a lane outside the grid the nearest grid vertex; only the stores
  // are conditional.  (The rounds assume nthreads == kBlock, the kernel's only block size.)
  constexpr int NS = S::SH * S::SW, RND = (NS + (int)kBlock - 1) / (int)kBlock;
  FP_LDS_CHECK(nthreads == (int)kBlock);
  float4 q[RND];
  float2 w[RND];
  if (INTERIOR) {
    // the staged region's first vertex (r0-2, c0-2) lies in the grid: every array is addressed
    // from its (workgroup-uniform) pointer to that vertex plus a 32-bit byte offset per lane
    const uint32_t g0 = G.base + (uint32_t)((r0 - 2) * C + (c0 - 2));
    const char *bx = (const char*)(A.in.x + g0), *by = (const char*)(A.in.y + g0), *bz = (const char*)(A.in.z + g0);
    const char *bu = (const char*)(A.in.vx + g0), *bv = (const char*)(A.in.vy + g0), *bw = (const char*)(A.in.vz + g0);
#pragma unroll
    for (int k = 0; k < RND; ++k) {
      const int e = tid + k * nthreads < NS ? tid + k * nthreads : NS - 1, lr = e / S::SW, lc = e - lr * S::SW;
      const uint32_t ob = (uint32_t)(lr * C + lc) * 4u;
      q[k] = float4{*(const float*)(bx + ob), *(const float*)(by + ob), *(const float*)(bz + ob),
                    *(const float*)(bu + ob)};
      w[k] = float2{*(const float*)(bv + ob), *(const float*)(bw + ob)};
    }
#pragma unroll
    for (int k = 0; k < RND; ++k) {
      const int e = tid + k * nthreads;
      if (e < NS) {
        FP_LDS_CHECK(e >= 0 && e < NS);
        L.pv[e] = q[k];
        L.vw[e] = w[k];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < RND; ++k) {
      const int e = tid + k * nthreads < NS ? tid + k * nthreads : NS - 1, lr = e / S::SW, lc = e - lr * S::SW;
      const int rr = r0 - 2 + lr, cc = c0 - 2 + lc;
      const int r = rr < 0 ? 0 : rr >= R ? R - 1 : rr, c = cc < 0 ? 0 : cc >= C ? C - 1 : cc;
      const uint32_t g = G.base + (uint32_t)(r * C + c);
      q[k] = float4{A.in.x[g], A.in.y[g], A.in.z[g], A.in.vx[g]};
      w[k] = float2{A.in.vy[g], A.in.vz[g]};
    }
#pragma unroll
    for (int k = 0; k < RND; ++k) {
      const int e = tid + k * nthreads;
      const int lr = e / S::SW, lc = e - lr * S::SW;
      const int r = r0 - 2 + lr, c = c0 - 2 + lc;
      if (e < NS && r >= 0 && r < R && c >= 0 && c < C) {
        FP_LDS_CHECK(e >= 0 && e < NS);
        L.pv[e] = q[k];
        L.vw[e] = w[k];
      }
    }
  }
  // the tables are computed by the upper half of the workgroup (waves 2 and 3 of 4), which store one
  // round of the staging above where the lower half stores two
  const int t0 = nthreads >= 2 * S::NTAB ? nthreads / 2 : 0;
  for (int e = tid - t0; e >= 0 && e < S::NTAB; e += nthreads) {
    // which table: coordinate axis (columns or rows), first index, reach (1 or 2), nominal constants
    const bool bend = e >= S::BH0;
    const int q = e - (e >= S::BV0 ? S::BV0 : e >= S::BH0 ? S::BH0 : e >= S::V0 ? S::V0 : S::H0);
    const bool cols = e < S::V0 || (e >= S::BH0 && e < S::BV0);
    const int first = (cols ? c0 : r0) - (bend ? 2 : 1) + q, reach = bend ? 2 : 1;
    const float d = cols ? grid_x(G, first) - grid_x(G, first + reach) : grid_y(G, first) - grid_y(G, first + reach);
    FP_LDS_CHECK(e >= 0 && e < S::NTAB);
    L.tab[e] = ktab(d, bend ? G.bk : G.tk, bend ? G.bc : G.tc);
  }
}

// Phase 1: thread t = i * QC + j evaluates quad (r0-1+i, c0-1+j)'s struts and faces; it returns what
// the quad owes its corner A, the vertex the same thread owns in phase 2 (for i, j >= 1).
template <int TR, int TC, bool INTERIOR>
FP_HD f3 phase1(Lds<TR, TC>& L, const StepParams& P, const ClothView& G, int r0, int c0, int t) {
  using S = Lds<TR, TC>;
  if (t >= S::NQ) return f3{0.0f, 0.0f, 0.0f};
  const int R = (int)G.rows, C = (int)G.cols;
  const int i = t / S::QC, j = t - i * S::QC;
  const int Rq = r0 - 1 + i, Cq = c0 - 1 + j;  // corner A
  const int la = (i + 1) * S::SW + (j + 1);     // A's staged index
  auto P3 = [&](int e) {
    FP_LDS_CHECK(e >= 0 && e < S::SH * S::SW);
    const float4 q = L.pv[e];
    return f3{q.x, q.y, q.z};
  };
  auto V3 = [&](int e) {
    FP_LDS_CHECK(e >= 0 && e < S::SH * S::SW);
    const float4 q = L.pv[e];
    const float2 w = L.vw[e];
    return f3{q.w, w.x, w.y};
  };
  const int lb = la + 1, lc = la + S::SW, ld = lc + 1;
  const f3 pA = P3(la), pB = P3(lb), pC = P3(lc), pD = P3(ld);
  const f3 uA = V3(la), uB = V3(lb), uC = V3(lc), uD = V3(ld);
  const bool rowA = INTERIOR || (Rq >= 0 && Rq < R), colA = INTERIOR || (Cq >= 0 && Cq < C);
  const bool quad = INTERIOR || (Rq >= 0 && Rq < R - 1 && Cq >= 0 && Cq < C - 1);
  const bool bend = INTERIOR || (Rq >= 0 && Rq < R - 2 && Cq >= 0 && Cq < C - 2);
  const f3 z{0.0f, 0.0f, 0.0f};

  // top: row strut A -> B, or for i == 0 the bend strut (r0-2, C') -> Cc
  f3 f_top;
  {
    const bool fr = i == 0;
    const int e0 = fr ? la - S::SW : la, e1 = fr ? lc : lb;
    const KTab K = L.tab[fr ? S::BV0 : S::H0 + j];
    const f3 f = strut_f(P3(e0), V3(e0), P3(e1), V3(e1), K.k, K.c, K.L);
    if (INTERIOR) {
      f_top = f;
    } else {
      const bool ok = fr ? (j >= 1 && r0 >= 2 && Cq >= 0 && Cq < C - 2) : (rowA && Cq >= 0 && Cq < C - 1);
      f_top = sel3(ok, f, z);
    }
  }
  // left: column strut A -> Cc, or for j == 0 the bend strut (R', c0-2) -> B
  f3 f_left;
  {
    const bool fr = j == 0;
    const int e0 = fr ? la - 1 : la, e1 = fr ? lb : lc;
    const KTab K = L.tab[fr ? S::BH0 : S::V0 + i];
    const f3 f = strut_f(P3(e0), V3(e0), P3(e1), V3(e1), K.k, K.c, K.L);
    if (INTERIOR) {
      f_left = f;
    } else {
      const bool ok = fr ? (i >= 1 && c0 >= 2 && Rq >= 0 && Rq < R - 2) : (colA && Rq >= 0 && Rq < R - 1);
      f_left = sel3(ok, f, z);
    }
  }
  // shear: B -> Cc, rest length from the quad's column and row differences
  f3 f_sh;
  {
    const float dx = L.tab[S::H0 + j].d, dy = L.tab[S::V0 + i].d;
    const float Ls = sqrt_cr(dx * dx + dy * dy);
    const float inv = rcp_f(Ls);
    const f3 f = strut_f(pB, uB, pC, uC, inv * G.sk, inv * G.sc, Ls);
    f_sh = INTERIOR ? f : sel3(quad, f, z);
  }
  // bends from A
  f3 f_bh, f_bv;
  {
    const int e1 = la + 2;
    const KTab K = L.tab[S::BH0 + j + 1];
    const f3 f = strut_f(pA, uA, P3(e1), V3(e1), K.k, K.c, K.L);
    f_bh = INTERIOR ? f : sel3(bend, f, z);
  }
  {
    const int e1 = la + 2 * S::SW;
    const KTab K = L.tab[S::BV0 + i + 1];
    const f3 f = strut_f(pA, uA, P3(e1), V3(e1), K.k, K.c, K.L);
    f_bv = INTERIOR ? f : sel3(bend, f, z);
  }

  // faces (B, Cc, A) and (B, D, Cc): drag + lift, corner weights from two angles each
  f3 sA, sB, sC, sD;
  {
    const f3 eBC = fsub(pC, pB), eBA = fsub(pA, pB), eBD = fsub(pD, pB);
    const float third = 1.0f / 3.0f;
    const f3 uBC = f3{uB.x + uC.x, uB.y + uC.y, uB.z + uC.z};
    const f3 vrA = f3{fma_(uBC.x + uA.x, third, -P.wind.x), fma_(uBC.y + uA.y, third, -P.wind.y),
                      fma_(uBC.z + uA.z, third, -P.wind.z)};
    const f3 vrB = f3{fma_(uBC.x + uD.x, third, -P.wind.x), fma_(uBC.y + uD.y, third, -P.wind.y),
                      fma_(uBC.z + uD.z, third, -P.wind.z)};
    const float nBC = fdot(eBC, eBC);
    // face (B, Cc, A): cr = (Cc - B) x (A - B); corners B: (A-B).(Cc-B), Cc: |Cc-B|^2 - that
    const f3 crA = fcross(eBC, eBA);
    const float cmA2 = fdot(crA, crA);
    const float rsA = rsq_f(cmA2), cmA = cmA2 * rsA;
    const float dA0 = fdot(eBA, eBC);
    const float wA0 = corner_w(cmA, dA0), wA1 = corner_w(cmA, nBC - dA0);
    const float wA2 = (1.0f - wA0) - wA1;
    const f3 FA = face_f(crA, rsA, vrA, P.neg_drag, P.neg_lift);
    // face (B, D, Cc): cr = (D - B) x (Cc - B); corners B: (Cc-B).(D-B), D: |D-B|^2 - that
    const f3 crB = fcross(eBD, eBC);
    const float cmB2 = fdot(crB, crB);
    const float rsB = rsq_f(cmB2), cmB = cmB2 * rsB;
    const float dB0 = fdot(eBC, eBD);
    const float wB0 = corner_w(cmB, dB0), wB1 = corner_w(cmB, fdot(eBD, eBD) - dB0);
    const float wB2 = (1.0f - wB0) - wB1;
    const f3 FB = face_f(crB, rsB, vrB, P.neg_drag, P.neg_lift);
    sA = f3{wA2 * FA.x, wA2 * FA.y, wA2 * FA.z};
    sB = fmad3(wB0, FB, f3{wA0 * FA.x, wA0 * FA.y, wA0 * FA.z});
    sC = fmad3(wB2, FB, f3{wA1 * FA.x, wA1 * FA.y, wA1 * FA.z});
    sD = f3{wB1 * FB.x, wB1 * FB.y, wB1 * FB.z};
    if (!INTERIOR) {
      sA = sel3(quad, sA, z);
      sB = sel3(quad, sB, z);
      sC = sel3(quad, sC, z);
      sD = sel3(quad, sD, z);
    }
  }

  // what this quad owes each corner, and the bend struts' far ends
  const f3 toA = add(add(add(add(sA, f_top), f_left), f_bh), f_bv);
  const f3 toB = sub(sub(add(sB, f_sh), f_top), sel3(j == 0, f_left, z));
  const f3 toC = sub(sub(sub(sC, f_sh), f_left), sel3(i == 0, f_top, z));
  // stored by recipient: B = (i-1, j), Cc = (i, j-1), D = (i, j), the row bend's far end (i-1, j+1)
  // and the column bend's (i+1, j-1) in tile-local coordinates, skipped outside the tile (the bends'
  // far ends take -f: phase 2 subtracts)
  auto put = [&](int o, bool in_tile, int slot, f3 f) {
    if (in_tile) {
      FP_LDS_CHECK(slot >= 0 && slot < S::NV);
      L.out[o][0][slot] = f.x;
      L.out[o][1][slot] = f.y;
      L.out[o][2][slot] = f.z;
    }
  };
  put(0, i >= 1 && j < TC, (i - 1) * TC + j, toB);
  put(1, i < TR && j >= 1, i * TC + j - 1, toC);
  put(2, i < TR && j < TC, i * TC + j, sD);
  put(3, i >= 1 && j + 1 < TC, (i - 1) * TC + j + 1, f_bh);
  put(4, i + 1 < TR && j >= 1, (i + 1) * TC + j - 1, f_bv);
  return toA;
}

FP_HD bool pin_bit(const uint32_t* __restrict__ bits, uint32_t i) {
  return bits != nullptr && ((bits[i >> 5] >> (i & 31u)) & 1u);
}

// Phase 2: thread t = i * QC + j with i, j >= 1 owns tile vertex (r0+i-1, c0+j-1), corner A of its
// own phase-1 quad (toA, passed in): it sums the vectors addressed to the vertex, then gravity,
// pins, the integrator, collision (`collide(p, v, F, xn, vn)`: the obstacle response of
// update_points, springy_mesh.rs:423-497) and the store.
template <int TR, int TC, int MODE, bool INTERIOR, class Collide>
FP_HD void phase2(const Lds<TR, TC>& L, const StepArgs& A, const StepParams& P, const ClothView& G, int r0, int c0,
                  int t, f3 toA, Collide collide) {
  using S = Lds<TR, TC>;
  if (t >= S::NQ) return;
  const int i = t / S::QC, j = t - i * S::QC;
  if (i == 0 || j == 0) return;
  const int a = i - 1, b = j - 1;
  const int r = r0 + a, c = c0 + b;
  if (!INTERIOR && (r >= (int)G.rows || c >= (int)G.cols)) return;
  const uint32_t gi = G.base + (uint32_t)(r * (int)G.cols + c);
  // the vertex's records: the tile's first vertex (r0, c0), which lies in the grid, as a
  // workgroup-uniform base and a 32-bit byte offset per lane (no per-lane 64-bit addresses)
  const uint32_t g0 = G.base + (uint32_t)(r0 * (int)G.cols + c0), ob = (uint32_t)(a * (int)G.cols + b) * 4u;
  auto rec = [&](float* base) -> float& { return *(float*)((char*)(base + g0) + ob); };
  const int slot = a * TC + b;  // the vertex's slot in L.out
  auto get = [&](int o) {
    FP_LDS_CHECK(slot >= 0 && slot < S::NV);
    return f3{L.out[o][0][slot], L.out[o][1][slot], L.out[o][2][slot]};
  };
  // then what the quads with this vertex as corner B, Cc and D (threads t - 1, t - QC, t - QC - 1)
  // and the bends ending at it left in its slot
  f3 F = A.force_init ? f3{rec(A.fx), rec(A.fy), rec(A.fz)} : f3{0.0f, 0.0f, 0.0f};
  F = add(F, P.gravity);
  F = add(F, toA);
  F = add(F, get(0));
  F = add(F, get(1));
  F = add(F, get(2));
  const f3 z{0.0f, 0.0f, 0.0f};
  F = sub(F, b >= 1 ? get(3) : z);  // bend (r, c-2) -> (r, c) from thread (i, j-2); for b == 0 a left strut
  F = sub(F, a >= 1 ? get(4) : z);  // bend (r-2, c) -> (r, c) from thread (i-2, j); for a == 0 a top strut
  const bool pin = pin_bit(A.pins, gi);
  if (pin) F = z;
  if (MODE == MODE_FORCES) {
    rec(A.fx) = F.x;
    rec(A.fy) = F.y;
    rec(A.fz) = F.z;
    return;
  }
  const int ls = (a + 2) * S::SW + (b + 2);
  const float4 q = L.pv[ls];
  const float2 w = L.vw[ls];
  const f3 p{q.x, q.y, q.z}, v{q.w, w.x, w.y};
  const float im = rcp_f(P.mass);
  const f3 acc{F.x * im, F.y * im, F.z * im};
  f3 xn, vn;
  if (P.integration == FP_EULER) {
    xn = f3{fma_(v.x, P.dt, p.x), fma_(v.y, P.dt, p.y), fma_(v.z, P.dt, p.z)};
  } else {  // frozen-force RK4 (state.rs:86-106): x + (h v + h^2 a / 2)
    const float h2 = 0.5f * P.dt * P.dt;
    xn = f3{p.x + fma_(h2, acc.x, P.dt * v.x), p.y + fma_(h2, acc.y, P.dt * v.y), p.z + fma_(h2, acc.z, P.dt * v.z)};
  }
  vn = f3{fma_(acc.x, P.dt, v.x), fma_(acc.y, P.dt, v.y), fma_(acc.z, P.dt, v.z)};
  collide(p, v, F, xn, vn);
  if (pin) {
    xn = p;
    vn = v;
  }
  rec(A.out.x) = xn.x;
  rec(A.out.y) = xn.y;
  rec(A.out.z) = xn.z;
  rec(A.out.vx) = vn.x;
  rec(A.out.vy) = vn.y;
  rec(A.out.vz) = vn.z;
  if (A.force_init) {
    rec(A.fx) = 0.0f;
    rec(A.fy) = 0.0f;
    rec(A.fz) = 0.0f;
  }
}

}  // namespace gfast
}  // namespace fp
