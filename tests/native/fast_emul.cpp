// CPU emulation of the fast-math tiled grid kernel (feriphys_amd/csrc/grid_fast_core.hpp): every
// tile's three phases run on the host in barrier order, with each LDS index bounds-checked and the
// LDS filled with NaN beforehand, so a read of a slot that no phase wrote shows in the results.
// Test infrastructure only: it lets the tests compare the kernel's tile logic with the oracle
// without a GPU.  No obstacles (the collision functions are the exact kernels' own).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#define FP_LDS_CHECK(cond)                                                      \
  do {                                                                          \
    if (!(cond)) {                                                              \
      std::fprintf(stderr, "fast_emul: index check failed: %s\n", #cond);       \
      std::abort();                                                             \
    }                                                                           \
  } while (0)
#include "../../feriphys_amd/csrc/grid_fast_core.hpp"

namespace {

template <int TR, int TC, int MODE, bool INTERIOR>
void tile(fp::gfast::Lds<TR, TC>& L, const fp::StepArgs& A, const fp::ClothView& G, const fp::StepParams& P, int r0,
          int c0) {
  const int nthreads = (int)fp::kBlock;
  for (int t = 0; t < nthreads; ++t) fp::gfast::phase0<TR, TC, INTERIOR>(L, A, G, r0, c0, t, nthreads);
  std::vector<fp::f3> toA(nthreads);  // each thread's registers across the barrier
  for (int t = 0; t < nthreads; ++t) toA[t] = fp::gfast::phase1<TR, TC, INTERIOR>(L, P, G, r0, c0, t);
  for (int t = 0; t < nthreads; ++t)
    fp::gfast::phase2<TR, TC, MODE, INTERIOR>(L, A, P, G, r0, c0, t, toA[t],
                                              [](fp::f3, fp::f3, fp::f3, fp::f3&, fp::f3&) {});
}

template <int TR, int TC, int MODE>
void run(const fp::StepArgs& A, const fp::ClothView& G, const fp::StepParams& P) {
  using L_t = fp::gfast::Lds<TR, TC>;
  std::unique_ptr<L_t> L(new L_t);
  for (uint32_t r0 = 0; r0 < G.rows; r0 += TR)
    for (uint32_t c0 = 0; c0 < G.cols; c0 += TC) {
      std::memset((void*)L.get(), 0xff, sizeof(L_t));  // NaN everywhere
      if (fp::gfast::interior_tile<TR, TC>(G, (int)r0, (int)c0))
        tile<TR, TC, MODE, true>(*L, A, G, P, (int)r0, (int)c0);
      else
        tile<TR, TC, MODE, false>(*L, A, G, P, (int)r0, (int)c0);
    }
}

}  // namespace

extern "C" {

// state: 6 planes of n floats (x, y, z, vx, vy, vz); out: the same after the step (mode 0); forces:
// 3 planes, read when force_init and written in mode 1 (the forces) or cleared (mode 0 with
// force_init); pins: bit mask over the vertices (may be NULL); cfg: integration, dt, dt_d, gravity[3],
// wind[3], drag, lift, restitution, friction, mass; shape: 0 = 15x15 tiles, 1 = 7x31.
int emul_fast(uint32_t rows, uint32_t cols, float spacing, const float* position, const float* kc6, const float* state,
              float* out, float* forces, int32_t force_init, const uint32_t* pins, int32_t integration,
              const float* cfg, int32_t mode, int32_t shape) {
  const uint32_t n = rows * cols;
  fp::ClothView G{};
  G.base = 0;
  G.rows = rows;
  G.cols = cols;
  G.spacing = spacing;
  G.cx = (float)cols / 2.0f * spacing;
  G.cy = (float)rows / 2.0f * spacing;
  G.p0 = position[0];
  G.p1 = position[1];
  G.tk = kc6[0]; G.tc = kc6[1]; G.sk = kc6[2]; G.sc = kc6[3]; G.bk = kc6[4]; G.bc = kc6[5];
  fp::StepParams P{};
  P.integration = integration;
  P.dt = cfg[0];
  P.dt_half = cfg[0] * 0.5f;
  P.dt6 = cfg[0] / 6.0f;
  P.dt3 = cfg[0] / 3.0f;
  P.dt_d = cfg[1];
  P.gravity = fp::f3{cfg[2], cfg[3], cfg[4]};
  P.wind = fp::f3{cfg[5], cfg[6], cfg[7]};
  P.neg_drag = -1.0f * cfg[8];
  P.neg_lift = -1.0f * cfg[9];
  P.restitution = cfg[10];
  P.friction = cfg[11];
  P.mass = cfg[12];
  P.uniform_mass = 1;
  // every plane in an allocation of its own, so that an address sanitizer reports any access past
  // a plane's end (one contiguous buffer would hide such a read in the next plane)
  std::vector<std::vector<float>> in(6), o(6, std::vector<float>(n)), f(3);
  for (int k = 0; k < 6; ++k) in[k].assign(state + (size_t)k * n, state + (size_t)(k + 1) * n);
  for (int k = 0; k < 3; ++k) f[k].assign(forces + (size_t)k * n, forces + (size_t)(k + 1) * n);
  std::vector<uint32_t> pin_words;
  if (pins) pin_words.assign(pins, pins + (n + 31) / 32);
  fp::StepArgs A{};
  A.in = fp::SoA{in[0].data(), in[1].data(), in[2].data(), in[3].data(), in[4].data(), in[5].data()};
  A.out = fp::SoA{o[0].data(), o[1].data(), o[2].data(), o[3].data(), o[4].data(), o[5].data()};
  A.fx = f[0].data();
  A.fy = f[1].data();
  A.fz = f[2].data();
  A.pins = pins ? pin_words.data() : nullptr;
  A.force_init = force_init;
  A.obs = fp::ObsView{};
  if (shape == 0) {
    if (mode == 0) run<15, 15, fp::MODE_STEP>(A, G, P);
    else run<15, 15, fp::MODE_FORCES>(A, G, P);
  } else {
    if (mode == 0) run<7, 31, fp::MODE_STEP>(A, G, P);
    else run<7, 31, fp::MODE_FORCES>(A, G, P);
  }
  for (int k = 0; k < 6; ++k) std::copy(o[k].begin(), o[k].end(), out + (size_t)k * n);
  for (int k = 0; k < 3; ++k) std::copy(f[k].begin(), f[k].end(), forces + (size_t)k * n);
  return 0;
}

}  // extern "C"
