// Host-side helpers shared by the C-ABI translation units: the thread-local error message behind
// fp_last_error, status plumbing and a minimal owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/feriphys_gpu.h"
#include "mesh_host.hpp"

namespace fph {

inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
inline int fail(const fp::Status& s) { return fail(s.code, s.msg); }

// Equal as bit patterns (so -0 differs from +0 and a NaN equals its own copy).
inline bool same_bits(float a, float b) {
  uint32_t x, y;
  std::memcpy(&x, &a, 4);
  std::memcpy(&y, &b, 4);
  return x == y;
}

// Duration::from_secs_f32(dt).as_secs_f32() (springy_mesh.rs:442; springy/simulation.rs:158).
// The float -> Duration conversion is exact with round-half-to-even to whole nanoseconds: dt has
// a 24-bit significand and 1e9 a 21-bit odd part, so dt * 1e9 is exact in x87 long double and
// nearbyintl rounds it half-to-even.  Then secs as f32 + (nanos as f32) / 1e9.
inline bool duration_roundtrip(float dt, float& out) {
  if (!(dt >= 0.0f) || !std::isfinite(dt)) return false;
  long double ns = (long double)dt * 1000000000.0L;
  long double r = nearbyintl(ns);
  if (r > 1.8e19L) return false;
  uint64_t total = (uint64_t)r;
  uint64_t secs = total / 1000000000ull;
  uint32_t nanos = (uint32_t)(total % 1000000000ull);
  out = (float)secs + (float)nanos / 1000000000.0f;
  return true;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fph::fail(FP_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)&p, count * sizeof(T));
  }
  // Allocates only when the buffer holds fewer than `count` elements.
  hipError_t reserve(size_t count) {
    if (p != nullptr && n >= count) return hipSuccess;
    return alloc(count);
  }
  hipError_t upload(const T* h, size_t count) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice);
  }
  // Like upload, but keeps the allocation (and so the device address) when the size is unchanged.
  hipError_t assign(const T* h, size_t count) {
    if (count != n || p == nullptr) return upload(h, count);
    return hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice);
  }
};

}  // namespace fph
