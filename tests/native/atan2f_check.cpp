// Compares fp::fdlibm_atan2f (the device kernels' branch-free atan2) and fp::fdlibm_atan2f_ref
// (fdlibm's control flow) with the host C library's atan2f, which is what the reference's
// f32::atan2 calls.  Prints the number of bit mismatches (NaN payloads aside).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../feriphys_amd/csrc/fdlibm_atan2f.hpp"

static uint64_t s = 88172645463325252ull;
static uint64_t xr() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float fb(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
static uint32_t bf(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }

int main(int argc, char** argv) {
  long n = argc > 1 ? std::atol(argv[1]) : 20000000;
  long mism = 0;
  const float special[] = {0.0f, -0.0f, 1.0f, -1.0f, 0.5f, 1.5f, 2.4375f, 0.4375f, 1.1875f, INFINITY, -INFINITY,
                           1e-30f, -1e-30f, 1e30f, -1e30f, 3.0f, -3.0f, 1e-45f, 16777216.0f};
  for (float y : special)
    for (float x : special) {
      float a = std::atan2(y, x), b = fp::fdlibm_atan2f(y, x), c = fp::fdlibm_atan2f_ref(y, x);
      if (bf(a) != bf(b)) { ++mism; std::printf("special y=%a x=%a libm=%a fdlibm=%a\n", y, x, a, b); }
      if (bf(a) != bf(c)) { ++mism; std::printf("special y=%a x=%a libm=%a fdlibm_ref=%a\n", y, x, a, c); }
    }
  for (long i = 0; i < n; ++i) {
    float y, x;
    uint64_t r = xr();
    if (i % 3 == 0) { y = fb((uint32_t)r); x = fb((uint32_t)(r >> 32)); }  // any bit pattern
    else {  // the face-angle regime: y = |a x b| >= 0, x = a . b of either sign
      y = (float)((r >> 11) * (1.0 / 9007199254740992.0)) * 4.0f;
      x = (float)((xr() >> 11) * (1.0 / 9007199254740992.0)) * 8.0f - 4.0f;
    }
    float a = std::atan2(y, x), b = fp::fdlibm_atan2f(y, x), c = fp::fdlibm_atan2f_ref(y, x);
    if (bf(a) != bf(b) && !(std::isnan(a) && std::isnan(b))) ++mism;
    if (bf(a) != bf(c) && !(std::isnan(a) && std::isnan(c))) ++mism;
  }
  std::printf("%ld\n", mism);
  return 0;
}
