// Learned per-vertex surrogate for large springy steps: shared constants, device-visible structs
// and the launchers of surrogate_kernels.hip.
//
// The network maps 72 features of a vertex (its up-to-12 strut neighbours, 6 values each) through
// three hidden layers of 128 GELU units (a piecewise-quadratic GELU, see surrogate_kernels.hip) to a
// 3-vector, the correction acceleration for a step of H = ratio * dt.  Weights and activations
// are bf16, products accumulate in f32 (MFMA).
#pragma once

#include "kernels.hpp"

namespace fp {

constexpr int kSlots = 12;
constexpr int kFeat = 6 * kSlots;  // 72
constexpr int kHid = 128;
constexpr int kOut = 3;
constexpr uint32_t kNoSlot = 0xffffffffu;

// Parameter vector (f32), each matrix row-major [out][in] (z = W x + b)
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kFeat;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kOut * kHid;
constexpr int kParams = kOffB4 + kOut;  // 42,755

// LDS image of the inference kernel: bf16 weight matrices with the K index permuted to the MFMA
// operand layout and rows padded (104 / 136 elements) so that the 16 rows one MFMA operand read
// touches fall in distinct LDS banks; then the f32 biases.
constexpr int kLd1 = 104, kLd = 136;
constexpr int kImgW1 = 0;                              // [128][104]
constexpr int kImgW2 = kImgW1 + kHid * kLd1 * 2;       // [128][136]
constexpr int kImgW3 = kImgW2 + kHid * kLd * 2;        // [128][136]
constexpr int kImgW4 = kImgW3 + kHid * kLd * 2;        // [16][136], rows 3..15 zero
constexpr int kImgB = kImgW4 + 16 * kLd * 2;           // f32 b1[128] b2[128] b3[128] b4[16]
constexpr int kImgBytes = kImgB + (3 * kHid + 16) * 4;  // 102,208
static_assert(kImgBytes % 16 == 0, "image is copied in 16-byte units");

// One feature slot of a vertex: the strut neighbour and the strut's rest length.
struct SlotEnt {
  uint32_t nbr;  // global vertex id, kNoSlot when the slot is empty
  float rest;
};

struct FeatParams {
  float H;             // big step (s)
  float strain_scale;  // position features: strain_scale * (|d| - L) / L * d / |d|
};

// Per-block diagnostics of a big step: the largest strut-strain change; the energy (kinetic +
// spring + gravity potential) before, and after minus the work of the face forces; and the
// mechanical energy (kinetic + spring) before plus after, the scale of the energy residual.
struct DiagOut {
  float max_dstrain, e_old, e_new, emech_old;
};

// variant: vertices per wave / 16 (2: 1024-thread workgroups, the default; 4: 512 threads)
hipError_t launch_sur_mlp(const SoA& st, const SlotEnt* slots, uint32_t n, const void* image, FeatParams fp,
                          float out_scale, float* da, int grid, int variant, hipStream_t s);
// Big-step update of the vertices in `blocks` from `a.in` to `a.out` with acceleration
// F_ext / m + da (F_ext: gravity and face forces, plus strut forces when `struts`); `work` gets
// each vertex's face-force work over the step.
hipError_t launch_sur_integrate(const StepArgs& a, const CsrView& t, const BlockRange* blocks, uint32_t n_blocks,
                                const float* da, float* work, bool struts, hipStream_t s);
hipError_t launch_sur_diag(const SoA& in, const SoA& out, const SlotEnt* slots, const float* slot_k,
                           const StepParams* params, const float* mass, const float* work, const BlockRange* blocks,
                           uint32_t n_blocks, DiagOut* diag, hipStream_t s);
// Per block of `blocks`: max over its vertices of |v| / vmax and |strain| / smax (+inf when any
// value is not finite).
hipError_t launch_sur_health(const SoA& st, const SlotEnt* slots, const BlockRange* blocks, uint32_t n_blocks,
                             float vmax, float smax, float* out, hipStream_t s);
// Features (natural order: slot * 6 + component, f32) of the vertices `ids`.
hipError_t launch_sur_features(const SoA& st, const SlotEnt* slots, const uint32_t* ids, uint32_t n, FeatParams fp,
                               float* out, hipStream_t s);

// Training samples at big step b from positions x_b (cur) and x_{b+1} (next) of the plain solver.
// noise: `noisy` = positions x_b + e1 and velocities ((x_b + e1) - (x_{b-1} + e2)) / H with e1, e2
// normal(0, sigma) per component (zero for pinned vertices).
struct Pos3 {
  const float *x, *y, *z;
};
hipError_t launch_sur_noise(Pos3 prev, Pos3 cur, SoA noisy, const uint32_t* pins, uint32_t n, float H, float sigma,
                            uint64_t seed, hipStream_t s);
// For the vertices `ids` (mesh ids in `mesh_of`): features of the noisy state a.in (bf16 row of
// 72) and the normalised target ((x_{b+1} - x~_b) / H^2 - v~ / H - F_ext / m) / out_scale, which is
// ((x_{b+1} - 2 x~_b + x~_{b-1}) / H^2 - F_ext / m) / out_scale.
hipError_t launch_sur_samples(const StepArgs& a, const CsrView& t, const uint32_t* ids, const uint32_t* mesh_of,
                              uint32_t n, Pos3 next, const SlotEnt* slots, FeatParams fp, float out_scale, bool struts,
                              __bf16* X, float* T, hipStream_t s);

// Training kernels
// C(m, n) = sum_k A(m, k) B(k, n) (+ C when accumulate), bf16 operands at arbitrary strides, f32 C.
hipError_t launch_gemm_bf16(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbk,
                            long sbn, float* C, long scm, long scn, bool accumulate, hipStream_t s);
// Z (f32 [rows][n]) += bias; H = bf16(gelu(Z)) (act) or bf16(Z)
hipError_t launch_bias_act(float* Z, const float* bias, __bf16* H, int rows, int n, bool act, hipStream_t s);
// dZ = bf16(dH * gelu'(Z))
hipError_t launch_gelu_bwd(const float* dH, const float* Z, __bf16* dZ, long count, int n, hipStream_t s);
// db[j] = sum_r dZ[r][j]
hipError_t launch_colsum(const __bf16* dZ, int rows, int n, float* db, hipStream_t s);
// Y (f32 [rows][3], pre-bias) += b4; dY = bf16(2 (Y - T) / (3 rows)); per-block squared-error sums
hipError_t launch_mse(float* Y, const float* b4, const float* T, int rows, __bf16* dY, float* partial, hipStream_t s);
hipError_t launch_sum(const float* partial, int n, float scale, float* out, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s);
hipError_t launch_to_bf16(const float* in, __bf16* out, long n, hipStream_t s);
// rows idx[j] of X (bf16 [.][72]) and T (f32 [.][3]) -> Xb, Tb; idx from a hash of (seed, j) mod count
hipError_t launch_gather_batch(const __bf16* X, const float* T, uint64_t count, uint64_t seed, int batch, __bf16* Xb,
                               float* Tb, hipStream_t s);

}  // namespace fp
