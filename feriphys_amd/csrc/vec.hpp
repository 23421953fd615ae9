// Vector arithmetic shared by the host builders and the gfx950 kernels.
//
// The exact policy reproduces the reference's f32 operation order (Rust never fuses a*b+c; cgmath
// 0.18 defines dot as (x*x' + y*y') + z*z', normalize as v * (1/|v|), angle as atan2(|a x b|, a.b)).
// All sources are compiled with -ffp-contract=off, so every a*b+c below stays two roundings unless
// a policy asks for a fused multiply-add explicitly (the fast policy does).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define FP_HD __host__ __device__ __forceinline__

#include "exact_ops.hpp"
#include "fdlibm_atan2f.hpp"

namespace fp {

struct f3 {
  float x, y, z;
};

FP_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
FP_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
FP_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FP_HD f3 smul(float s, f3 v) { return f3{s * v.x, s * v.y, s * v.z}; }
FP_HD f3 sdiv(f3 v, float s) { return f3{v.x / s, v.y / s, v.z / s}; }
FP_HD f3 neg(f3 v) { return f3{-v.x, -v.y, -v.z}; }
FP_HD f3 cross(f3 a, f3 b) { return f3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)}; }
FP_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
FP_HD bool sign_positive(float v) {
#ifdef __HIP_DEVICE_COMPILE__
  return !__builtin_signbit(v);
#else
  return !signbit(v);
#endif
}

// Host-side exact helpers (used by the topology builder for rest lengths and obstacle normals).
inline float h_magnitude(f3 v) { return sqrtf(dot(v, v)); }
inline f3 h_normalize(f3 v) { return smul(1.0f / h_magnitude(v), v); }

// ---- math policies for the kernels ------------------------------------------------------------
// ExactMath: correctly rounded sqrt, division and reciprocal (the HIP defaults for f32, or the
// correctly rounded forms of exact_ops.hpp), the C library's atan2f rounding (fdlibm_atan2f.hpp),
// no fused multiply-add in the reference's arithmetic: the reference's f32 results bit for bit.
struct ExactMath {
  static constexpr bool kExact = true;
  __device__ static float sqrt_(float x) { return sqrt_rn(x); }
  __device__ static float rcp(float x) { return rcp_rn(x); }
  __device__ static float div(float a, float b) { return a / b; }
  __device__ static float atan2_(float y, float x) { return fdlibm_atan2f(y, x); }
  __device__ static float div_pi(float x) { return div_pi_rn(x); }
  __device__ static float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
  __device__ static float mad(float a, float b, float c) { return a * b + c; }
};

// atan2 for the fast kernels: atan of min/max(|x|,|y|) in [0, 1] by an odd degree-13 polynomial
// (least-squares fit on Chebyshev nodes, refined towards minimax; max error 3.2e-7 rad), then
// placed in its octant.
__device__ __forceinline__ float fast_atan2(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float a = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
  const float t = a * a;
  float p = 0.006811788771301508f;
  p = __builtin_fmaf(p, t, -0.0336042158305645f);
  p = __builtin_fmaf(p, t, 0.07962367683649063f);
  p = __builtin_fmaf(p, t, -0.132333442568779f);
  p = __builtin_fmaf(p, t, 0.19807817041873932f);
  p = __builtin_fmaf(p, t, -0.3331736922264099f);
  p = __builtin_fmaf(p, t, 0.9999961256980896f);
  float r = p * a;
  if (ay > ax) r = 1.5707963267948966f - r;
  if (x < 0.0f) r = 3.141592653589793f - r;
  return __builtin_copysignf(r, y);
}

// FastMath: hardware sqrt/reciprocal (1 ulp), fused multiply-adds, polynomial atan2.  Not
// bit-identical; the parity tests hold it to the 1e-5 relative one-step bound.
struct FastMath {
  static constexpr bool kExact = false;
  __device__ static float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  __device__ static float atan2_(float y, float x) { return fast_atan2(y, x); }
  __device__ static float div_pi(float x) { return div(x, 3.14159274101257324f); }
  __device__ static float dot(f3 a, f3 b) { return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)); }
  __device__ static float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};

}  // namespace fp
