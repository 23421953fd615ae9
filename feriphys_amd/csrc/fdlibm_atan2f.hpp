// Single-precision atan2 that rounds exactly as the C library the reference links (glibc's
// single-precision atan2f, which is fdlibm's algorithm converted to float: Sun Microsystems,
// fdlibm 5.3, e_atan2.c / s_atan.c).  Rust's f32::atan2 calls it, so the reference's face corner
// angles (cgmath Vector3::angle, springy_mesh.rs:163-172) are these values bit for bit.
//
// Method (fdlibm): reduce atan(|y/x|) to one of five intervals, atan(t) = atan(c) + atan((t-c)/(1+ct))
// with c in {0, 0.5, 1, 1.5, inf}, evaluate an 11-term odd polynomial split into even and odd
// halves, and add atan(c) as a (hi, lo) pair; then place the result in its quadrant.  The
// polynomial coefficients are fdlibm's double-precision ones rounded to f32; atan(c) hi parts are
// truncated to f32 and lo = round(atan(c) - hi).  Must be compiled without FMA contraction.
//
// fdlibm_atan2f_ref follows fdlibm's control flow; fdlibm_atan2f computes the same bits without its
// data-dependent branches (one reduction division with selected constants; see below), which is
// what a wavefront wants when its lanes fall in different reduction intervals.  Both are checked
// bit for bit against the host C library's atan2f (tests).
#pragma once

#include <stdint.h>
#include <string.h>

#ifndef FP_HD
#define FP_HD inline
#endif

namespace fp {

FP_HD float fdl_from_bits(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}
FP_HD uint32_t fdl_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  return b;
}
FP_HD float fdl_abs(float f) { return fdl_from_bits(fdl_bits(f) & 0x7fffffffu); }

FP_HD float fdlibm_atanf(float x) {
  const float atanhi0 = fdl_from_bits(0x3eed6338u), atanhi1 = fdl_from_bits(0x3f490fdau),
              atanhi2 = fdl_from_bits(0x3f7b985eu), atanhi3 = fdl_from_bits(0x3fc90fdau);
  const float atanlo0 = fdl_from_bits(0x31ac376au), atanlo1 = fdl_from_bits(0x33222169u),
              atanlo2 = fdl_from_bits(0x33140fb5u), atanlo3 = fdl_from_bits(0x33a22169u);
  const float a0 = fdl_from_bits(0x3eaaaaabu), a1 = fdl_from_bits(0xbe4ccccdu), a2 = fdl_from_bits(0x3e124925u),
              a3 = fdl_from_bits(0xbde38e38u), a4 = fdl_from_bits(0x3dba2e6eu), a5 = fdl_from_bits(0xbd9d8795u),
              a6 = fdl_from_bits(0x3d886b35u), a7 = fdl_from_bits(0xbd6ef16bu), a8 = fdl_from_bits(0x3d4bda59u),
              a9 = fdl_from_bits(0xbd15a221u), a10 = fdl_from_bits(0x3c8569d7u);
  const int32_t hx = (int32_t)fdl_bits(x);
  const int32_t ix = hx & 0x7fffffff;
  int id;
  if (ix >= 0x4c000000) {  // |x| >= 2^25
    if (ix > 0x7f800000) return x + x;  // NaN
    return hx > 0 ? atanhi3 + atanlo3 : -atanhi3 - atanlo3;
  }
  if (ix < 0x3ee00000) {  // |x| < 0.4375
    if (ix < 0x31000000) return x;  // |x| < 2^-29
    id = -1;
  } else {
    x = fdl_abs(x);
    if (ix < 0x3f980000) {    // |x| < 1.1875
      if (ix < 0x3f300000) {  // 7/16 <= |x| < 11/16
        id = 0;
        x = (2.0f * x - 1.0f) / (2.0f + x);
      } else {  // 11/16 <= |x| < 19/16
        id = 1;
        x = (x - 1.0f) / (x + 1.0f);
      }
    } else if (ix < 0x401c0000) {  // |x| < 2.4375
      id = 2;
      x = (x - 1.5f) / (1.0f + 1.5f * x);
    } else {  // 2.4375 <= |x| < 2^25
      id = 3;
      x = -1.0f / x;
    }
  }
  const float z = x * x;
  const float w = z * z;
  const float s1 = z * (a0 + w * (a2 + w * (a4 + w * (a6 + w * (a8 + w * a10)))));
  const float s2 = w * (a1 + w * (a3 + w * (a5 + w * (a7 + w * a9))));
  if (id < 0) return x - x * (s1 + s2);
  const float hi = id == 0 ? atanhi0 : id == 1 ? atanhi1 : id == 2 ? atanhi2 : atanhi3;
  const float lo = id == 0 ? atanlo0 : id == 1 ? atanlo1 : id == 2 ? atanlo2 : atanlo3;
  const float r = hi - ((x * (s1 + s2) - lo) - x);
  return hx < 0 ? -r : r;
}

FP_HD float fdlibm_atan2f_ref(float y, float x) {
  const float tiny = 1.0e-30f;
  const float pi_o_4 = fdl_from_bits(0x3f490fdbu), pi_o_2 = fdl_from_bits(0x3fc90fdbu),
              pi = fdl_from_bits(0x40490fdbu), pi_lo = fdl_from_bits(0xb3bbbd2eu);
  const int32_t hx = (int32_t)fdl_bits(x), hy = (int32_t)fdl_bits(y);
  const int32_t ix = hx & 0x7fffffff, iy = hy & 0x7fffffff;
  if (ix > 0x7f800000 || iy > 0x7f800000) return x + y;  // NaN
  if (hx == 0x3f800000) return fdlibm_atanf(y);          // x == 1
  const int m = ((hy >> 31) & 1) | ((hx >> 30) & 2);     // 2*sign(x) + sign(y)
  if (iy == 0) {
    if (m <= 1) return y;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if (ix == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (ix == 0x7f800000) {
    if (iy == 0x7f800000) {
      if (m == 0) return pi_o_4 + tiny;
      if (m == 1) return -pi_o_4 - tiny;
      if (m == 2) return 3.0f * pi_o_4 + tiny;
      return -3.0f * pi_o_4 - tiny;
    }
    if (m == 0) return 0.0f;
    if (m == 1) return -0.0f;
    if (m == 2) return pi + tiny;
    return -pi - tiny;
  }
  if (iy == 0x7f800000) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  const int32_t k = (iy - ix) >> 23;
  float z;
  if (k > 60) z = pi_o_2 + 0.5f * pi_lo;  // |y/x| > 2^60
  else if (hx < 0 && k < -60) z = 0.0f;   // |y|/x < -2^60
  else {
    z = fdlibm_atanf(fdl_abs(y / x));
  }
  if (m == 0) return z;
  if (m == 1) return -z;
  if (m == 2) return pi - (z - pi_lo);
  return (z - pi_lo) - pi;
}

// fdlibm_atan2f_ref without data-dependent branches, for finite non-zero y and x (other inputs take
// the reference path).  Why each fdlibm branch can go:
//   x == 1 calls atanf(y) directly: the general path computes atanf(|y / 1|) and the same sign;
//   |y/x| > 2^60 and (x < 0, |y/x| < 2^-60) return constants that the general path rounds to as
//   well (atan(t >= 2^25) is pi/2 in f32; a tiny z vanishes against pi_lo);
//   the four reduction intervals share one form, atan(t) = hi + atan((t A - B) / (A + t B)) with
//   (A, B) = (1, 0), (2, 1), (1, 1), (1, 1.5), (0, 1) for t < 7/16, < 11/16, < 19/16, < 39/16,
//   above: each numerator and denominator then rounds exactly like fdlibm's (t*1 = t, t*0 = 0,
//   0 + t = t), and for the first interval hi = lo = 0 gives 0 - ((u S - 0) - u) = u - u S.
FP_HD float fdlibm_atan2f(float y, float x) {
  const uint32_t ix = fdl_bits(x) & 0x7fffffffu, iy = fdl_bits(y) & 0x7fffffffu;
  if (ix - 1u >= 0x7f7fffffu || iy - 1u >= 0x7f7fffffu) return fdlibm_atan2f_ref(y, x);  // 0, inf, NaN
  const float pi = fdl_from_bits(0x40490fdbu), pi_lo = fdl_from_bits(0xb3bbbd2eu);
  const float t = fdl_abs(y / x);
  const uint32_t it = fdl_bits(t);
  const bool c0 = it >= 0x3ee00000u, c1 = it >= 0x3f300000u, c2 = it >= 0x3f980000u, c3 = it >= 0x401c0000u;
  const float A = c3 ? 0.0f : (c0 && !c1) ? 2.0f : 1.0f;
  const float B = !c0 ? 0.0f : (c2 && !c3) ? 1.5f : 1.0f;
  const float hi = !c0   ? 0.0f
                   : !c1 ? fdl_from_bits(0x3eed6338u)
                   : !c2 ? fdl_from_bits(0x3f490fdau)
                   : !c3 ? fdl_from_bits(0x3f7b985eu)
                         : fdl_from_bits(0x3fc90fdau);
  const float lo = !c0   ? 0.0f
                   : !c1 ? fdl_from_bits(0x31ac376au)
                   : !c2 ? fdl_from_bits(0x33222169u)
                   : !c3 ? fdl_from_bits(0x33140fb5u)
                         : fdl_from_bits(0x33a22169u);
  const float u = (t * A - B) / (A + t * B);
  const float a0 = fdl_from_bits(0x3eaaaaabu), a1 = fdl_from_bits(0xbe4ccccdu), a2 = fdl_from_bits(0x3e124925u),
              a3 = fdl_from_bits(0xbde38e38u), a4 = fdl_from_bits(0x3dba2e6eu), a5 = fdl_from_bits(0xbd9d8795u),
              a6 = fdl_from_bits(0x3d886b35u), a7 = fdl_from_bits(0xbd6ef16bu), a8 = fdl_from_bits(0x3d4bda59u),
              a9 = fdl_from_bits(0xbd15a221u), a10 = fdl_from_bits(0x3c8569d7u);
  const float z = u * u;
  const float w = z * z;
  const float s1 = z * (a0 + w * (a2 + w * (a4 + w * (a6 + w * (a8 + w * a10)))));
  const float s2 = w * (a1 + w * (a3 + w * (a5 + w * (a7 + w * a9))));
  float r = hi - ((u * (s1 + s2) - lo) - u);
  r = it < 0x31000000u ? t : r;                                // |t| < 2^-29: atan(t) = t
  r = it >= 0x4c000000u ? fdl_from_bits(0x3fc90fdbu) : r;      // t >= 2^25: atanhi[3] + atanlo[3]
  const bool yneg = fdl_bits(y) >> 31, xneg = fdl_bits(x) >> 31;
  if (!xneg) return yneg ? -r : r;
  const float q = yneg ? (r - pi_lo) - pi : pi - (r - pi_lo);
  return q;
}

}  // namespace fp
