// Correctly rounded f32 operations computed without the division instruction sequence.  Each is
// checked bit for bit against the plain operator on every finite f32 input (tests/native).
#pragma once

#ifndef FP_HD
#define FP_HD inline
#endif

namespace fp {

// x / 3.0f, correctly rounded: q0 = x * RN(1/3), then one FMA correction with the exact remainder
// fma(-q0, 3, x).  The zero guard keeps the sign of a zero quotient.
FP_HD float div3_rn(float x) {
  const float third = 0.3333333432674408f;  // RN(1/3) = 0x3eaaaaab
  const float q0 = x * third;
  const float r = __builtin_fmaf(-q0, 3.0f, x);
  return q0 == 0.0f ? q0 : __builtin_fmaf(r, third, q0);
}

// 1.0f / x, correctly rounded.  On gfx950 for |x| in [2^-126, 2^126]: the hardware reciprocal
// (within 1 ulp) refined once by Newton's step with the exact residual fma(-x, r, 1); checked on
// every such input against the division (fp_selftest_exact_math).  Other inputs (zero, subnormal,
// huge, inf, NaN) and host code take the division.
FP_HD float rcp_rn(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  const float ax = __builtin_fabsf(x);
  if (ax >= 0x1p-126f && ax <= 0x1p126f) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
  }
#endif
  return 1.0f / x;
}

// sqrt(x), correctly rounded.  On gfx950 for x in [2^-96, FLT_MAX]: the hardware square root
// (within 1 ulp) and the choice among it and its two neighbours by the sign of the exact
// residuals x - s' s (the compiler's own correctly rounded sequence without its rescaling of tiny
// inputs and its zero/infinity fix-up, which that range does not need); checked on every such
// input against the full sequence (fp_selftest_exact_math).  Other inputs take the full sequence.
FP_HD float sqrt_rn(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  if (x >= 0x1p-96f && x <= 3.40282347e38f) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u), up = __uint_as_float(__float_as_uint(s) + 1u);
    const float rdn = __builtin_fmaf(-dn, s, x), rup = __builtin_fmaf(-up, s, x);
    const float t = rdn <= 0.0f ? dn : s;
    return rup > 0.0f ? up : t;
  }
#endif
  return __builtin_sqrtf(x);
}

// x / PI (PI = std::f32::consts::PI = 0x40490fdb), correctly rounded: for x >= 2^-100 the product
// with RN(1/PI) corrected once by the exact remainder fma(-q0, PI, x) (the same method as div3_rn;
// below 2^-104 the remainder is no longer exact), otherwise (tiny, zero, negative, NaN) the
// division itself.
FP_HD float div_pi_rn(float x) {
  const float pi = 3.14159274101257324f, inv_pi = 0.31830987334251404f;  // 0x40490fdb, RN(1/PI) = 0x3ea2f983
  if (!(x >= 0x1p-100f)) return x / pi;
  const float q0 = x * inv_pi;
  return __builtin_fmaf(__builtin_fmaf(-q0, pi, x), inv_pi, q0);
}

}  // namespace fp
