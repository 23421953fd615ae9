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

}  // namespace fp
