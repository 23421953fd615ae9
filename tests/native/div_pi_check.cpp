// Compares fp::div_pi_rn with x / PI (PI = 0x40490fdb) on every finite f32; prints the number of bit
// mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../feriphys_amd/csrc/exact_ops.hpp"

int main() {
  const float pi = 3.14159274101257324f;
  unsigned long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long i = 0; i < (1ll << 32); ++i) {
    const uint32_t b = (uint32_t)i;
    float x;
    std::memcpy(&x, &b, 4);
    if (!std::isfinite(x)) continue;
    const float ref = x / pi, got = fp::div_pi_rn(x);
    uint32_t rb, gb;
    std::memcpy(&rb, &ref, 4);
    std::memcpy(&gb, &got, 4);
    bad += rb != gb;
  }
  std::printf("%llu\n", bad);
  return 0;
}
