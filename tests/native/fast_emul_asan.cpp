// Drives the CPU emulation of the fast grid kernel (fast_emul.cpp) over small and ragged grids,
// both tile shapes and both modes, built with an address sanitizer: every staging load, the clamped
// ones of lanes past the staged region or outside the grid included, stays inside its state plane,
// and every pin-word read inside the bit mask.  Test infrastructure only; prints "ok".
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" int emul_fast(uint32_t rows, uint32_t cols, float spacing, const float* position, const float* kc6,
                         const float* state, float* out, float* forces, int32_t force_init, const uint32_t* pins,
                         int32_t integration, const float* cfg, int32_t mode, int32_t shape);

int main() {
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> jitter(-0.01f, 0.01f);
  const uint32_t sides[] = {2, 3, 5, 8, 16, 17, 19, 20, 33, 34};
  const float spacing = 0.05f, dt = 1.0f / 600.0f;
  const float position[3] = {0.0f, 0.0f, 0.0f};
  const float kc6[6] = {20.0f, 0.1f, 15.0f, 0.1f, 5.0f, 0.05f};
  const float cfg[13] = {dt, dt, 0.0f, -9.81f, 0.0f, 1.0f, 0.0f, 0.5f, 0.2f, 0.5f, 0.3f, 0.2f, 0.01f};
  for (uint32_t rows : sides)
    for (uint32_t cols : sides)
      for (int32_t shape = 0; shape < 2; ++shape)
        for (int32_t mode = 0; mode < 2; ++mode) {
          const uint32_t n = rows * cols;
          std::vector<float> state(6 * (size_t)n), out(6 * (size_t)n), forces(3 * (size_t)n);
          for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t c = 0; c < cols; ++c) {
              const size_t i = (size_t)r * cols + c;
              state[i] = (float)c * spacing - (float)cols * spacing / 2.0f + jitter(rng);
              state[n + i] = (float)r * spacing - (float)rows * spacing / 2.0f + jitter(rng);
              state[2 * (size_t)n + i] = jitter(rng);
              for (int k = 3; k < 6; ++k) state[k * (size_t)n + i] = jitter(rng);
              for (int k = 0; k < 3; ++k) forces[k * (size_t)n + i] = jitter(rng);
            }
          std::vector<uint32_t> pins((n + 31) / 32, 0u);
          pins[0] |= 1u;
          pins[(n - 1) >> 5] |= 1u << ((n - 1) & 31u);
          if (emul_fast(rows, cols, spacing, position, kc6, state.data(), out.data(), forces.data(), mode,
                        pins.data(), 1, cfg, mode, shape) != 0) {
            std::printf("emul_fast failed: %u x %u, shape %d, mode %d\n", rows, cols, shape, mode);
            return 1;
          }
        }
  std::puts("ok");
  return 0;
}
