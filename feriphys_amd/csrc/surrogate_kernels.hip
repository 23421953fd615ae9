// gfx950 kernels of the learned large-step surrogate: per-vertex features, the fused bf16 MFMA
// inference kernel, the big-step update, step diagnostics, training-sample generation and the
// training kernels (bf16 MFMA GEMM, GELU forward/backward, MSE, Adam).
//
// MFMA used throughout: v_mfma_f32_16x16x32_bf16.  Operand layout (lane l of a wave, s = 0..7):
// A[16 x 32]: row l%16, k = (l/16)*8 + s;  B[32 x 16]: column l%16, k = (l/16)*8 + s;
// C[16 x 16] (4 f32 per lane): rows (l/16)*4 + r, column l%16.
#include "surrogate.hpp"

#pragma clang fp contract(fast)

namespace fp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The network's activation, a GELU without transcendentals: x s(x) with the normal CDF replaced by
// a C1 piecewise quadratic,
//   s(x) = clamp(0.5 + x (a - b min(|x|, x*)), 0, 1),  a = 1 / x*,  b = 1 / (2 x*^2),
// which reaches 1 (and 0) with zero slope at |x| = x*, so act(x) is exactly x for x >= x* and 0 for
// x <= -x*; x* = 2.46298 minimises max |act(x) - x Phi(x)| (0.0197).  Four VALU operations per
// element (the two clamps ride on a multiply and an FMA): the activation is what the inference
// kernel spends most of its vector instructions on.
constexpr float kXs = 2.4629791631638027f, kA = 0.4060123670374283f, kB = 0.08242302109366773f;
constexpr float kIXs = (float)(1.0 / 2.4629791631638027), kBX = (float)(0.08242302109366773 * 2.4629791631638027);
__device__ __forceinline__ float gelu(float x) {
  const float t = __builtin_amdgcn_fmed3f(fabsf(x) * kIXs, 0.0f, 1.0f);  // min(|x|, x*) / x*
  const float s = __builtin_amdgcn_fmed3f(fmaf(x, fmaf(-kBX, t, kA), 0.5f), 0.0f, 1.0f);
  return x * s;
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float t = __builtin_amdgcn_fmed3f(fabsf(x) * kIXs, 0.0f, 1.0f);
  const float s = __builtin_amdgcn_fmed3f(fmaf(x, fmaf(-kBX, t, kA), 0.5f), 0.0f, 1.0f);
  const float ds = fabsf(x) < kXs ? fmaf(-2.0f * kB, fabsf(x), kA) : 0.0f;  // s'(x) = a - 2 b |x|
  return fmaf(x, ds, s);
}

__device__ __forceinline__ void store_state(const SoA& o, uint32_t i, f3 x, f3 v) {
  o.x[i] = x.x; o.y[i] = x.y; o.z[i] = x.z;
  o.vx[i] = v.x; o.vy[i] = v.y; o.vz[i] = v.z;
}

// The 6 features of one slot: strain_scale * strain * d/|d| and (v_nbr - v) * H / L, from the
// neighbour's position q and velocity w; zeros when `has` is false.  Branch-free, so that a caller
// can issue every load of a pass before the first feature is computed.
__device__ __forceinline__ void slot_feat_from(f3 p, f3 v, f3 q, f3 w, float rest, bool has, FeatParams fp, float* f) {
  const f3 d = sub(q, p);
  const float len = __builtin_sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
  const float il = __builtin_amdgcn_rcpf(rest);
  const float st = (has && len > 0.0f) ? fp.strain_scale * ((len - rest) * il) * __builtin_amdgcn_rcpf(len) : 0.0f;
  const float hv = has ? fp.H * il : 0.0f;
  f[0] = st * d.x; f[1] = st * d.y; f[2] = st * d.z;
  f[3] = (w.x - v.x) * hv; f[4] = (w.y - v.y) * hv; f[5] = (w.z - v.z) * hv;
}

__device__ __forceinline__ void slot_feat(const SoA& S, f3 p, f3 v, uint32_t self, SlotEnt e, FeatParams fp,
                                          float* f) {
  const bool has = e.nbr != kNoSlot;
  const uint32_t nb = has ? e.nbr : self;
  slot_feat_from(p, v, ld3(S.x, S.y, S.z, nb), ld3(S.vx, S.vy, S.vz, nb), has ? e.rest : 1.0f, has, fp, f);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float unit_open(uint64_t h) { return ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------------------------
// Inference: a workgroup takes 512 vertices per pass, NT MFMA column tiles of 16 vertices per wave
// (NT = 2: 16 waves of 32 vertices, 4 waves per SIMD within 128 registers; NT = 4: 8 waves of 64).
// Neurons are the MFMA rows, vertices the columns, so a layer's f32 output tile already holds, per
// lane, 4 consecutive neurons of one vertex; two such tiles form the next layer's 8-element bf16
// operand once the weights' K index is permuted the same way (done when the image is packed).  The
// activations never leave registers; the weights stay in LDS for the whole (persistent) launch.

constexpr int kMlpTile = 512;

// One layer, acc = W B + bias then the activation, pipelined by output block: the 16 neurons of
// block mt accumulate (KB MFMAs per column tile, the bias as the accumulators' initial value)
// while the activation of block mt - 1 is evaluated, so that each wave keeps the matrix cores and
// the vector ALUs busy together; two accumulator blocks are live instead of eight.  `out` must not
// alias `B`, which is read until the last block.
template <int KB, int LDW, int NT>
__device__ __forceinline__ void mfma_act_layer(const __bf16* __restrict__ W, const float* __restrict__ bias,
                                               const bf16x8 (&B)[KB][NT], bf16x8 (&out)[4][NT], int col, int g) {
  f32x4 acc[2][NT];
#pragma unroll
  for (int mt = 0; mt <= 8; ++mt) {
    if (mt < 8) {
      const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 16 * mt + 4 * g);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt & 1][nt] = b;
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(W + (16 * mt + col) * LDW + kb * 32 + g * 8);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt & 1][nt] = mfma(a, B[kb][nt], acc[mt & 1][nt]);
      }
    }
    if (mt > 0) {
      const int p = mt - 1;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) out[p >> 1][nt][(p & 1) * 4 + r] = (__bf16)gelu(acc[p & 1][nt][r]);
    }
  }
}

template <int NT>
__global__ void __launch_bounds__(2048 / NT) sur_mlp_kernel(SoA S, const SlotEnt* __restrict__ slots, uint32_t n,
                                                            const uint4* __restrict__ image, FeatParams fp,
                                                            float out_scale, float* __restrict__ da) {
  constexpr int kThreads = 2048 / NT;
  __shared__ uint4 s_img[kImgBytes / 16];
  for (int i = threadIdx.x; i < kImgBytes / 16; i += kThreads) s_img[i] = image[i];
  __syncthreads();
  const char* img = reinterpret_cast<const char*>(s_img);
  const __bf16* W1 = reinterpret_cast<const __bf16*>(img + kImgW1);
  const __bf16* W2 = reinterpret_cast<const __bf16*>(img + kImgW2);
  const __bf16* W3 = reinterpret_cast<const __bf16*>(img + kImgW3);
  const __bf16* W4 = reinterpret_cast<const __bf16*>(img + kImgW4);
  const float* Bias = reinterpret_cast<const float*>(img + kImgB);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 15, g = lane >> 4;
  for (uint32_t t0 = blockIdx.x * kMlpTile; t0 < n; t0 += gridDim.x * kMlpTile) {
    const uint32_t v0 = t0 + wave * 16 * NT;
    if (v0 >= n) continue;
    // features: lane group g computes slots 3g..3g+2 of its vertex (18 values, then 6 zeros), which
    // are its 8-element K slices of the three 32-wide K blocks of layer 1.  Every load is
    // unconditional (past-the-end columns read vertex n-1, empty slots the vertex itself; both are
    // zeroed), so a pass's loads can all be in flight together.
    bf16x8 B1[3][NT];
    SlotEnt se[NT][3];
    f3 p[NT], v[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const uint32_t iv = v0 + nt * 16 + col;
      const uint32_t i = iv < n ? iv : n - 1;
      p[nt] = ld3(S.x, S.y, S.z, i);
      v[nt] = ld3(S.vx, S.vy, S.vz, i);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        se[nt][j] = slots[(size_t)i * kSlots + 3 * g + j];
        if (iv >= n) se[nt][j].nbr = kNoSlot;
      }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const uint32_t iv = v0 + nt * 16 + col;
      const uint32_t i = iv < n ? iv : n - 1;
      float f[24];
#pragma unroll
      for (int j = 0; j < 3; ++j) slot_feat(S, p[nt], v[nt], i, se[nt][j], fp, f + 6 * j);
#pragma unroll
      for (int k = 18; k < 24; ++k) f[k] = 0.0f;
#pragma unroll
      for (int kb = 0; kb < 3; ++kb)
#pragma unroll
        for (int s = 0; s < 8; ++s) B1[kb][nt][s] = (__bf16)f[kb * 8 + s];
    }
    bf16x8 Bh[4][NT], Bg[4][NT];
    mfma_act_layer<3, kLd1, NT>(W1, Bias, B1, Bh, col, g);
    mfma_act_layer<4, kLd, NT>(W2, Bias + kHid, Bh, Bg, col, g);
    mfma_act_layer<4, kLd, NT>(W3, Bias + 2 * kHid, Bg, Bh, col, g);
    f32x4 o[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) o[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(W4 + col * kLd + kb * 32 + g * 8);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) o[nt] = mfma(a, Bh[kb][nt], o[nt]);
    }
    if (g == 0) {
      const float b0 = Bias[3 * kHid], b1 = Bias[3 * kHid + 1], b2 = Bias[3 * kHid + 2];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const uint32_t i = v0 + nt * 16 + col;
        if (i < n) {
          da[3 * (size_t)i] = out_scale * (o[nt][0] + b0);
          da[3 * (size_t)i + 1] = out_scale * (o[nt][1] + b1);
          da[3 * (size_t)i + 2] = out_scale * (o[nt][2] + b2);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Big-step update: a = (gravity + face forces [+ strut forces]) / m + da; v' = v + H a;
// x' = x + H v'; then the obstacle first-hit test and the reference's collision response over the
// step H, and pinned vertices keep their record.  `work` receives each vertex's face-force work
// F_face . (x' - x), for the energy balance of the diagnostics.
template <bool STRUTS>
__global__ void __launch_bounds__(kBlock) sur_integrate_kernel(StepArgs A, CsrView T, const BlockRange* __restrict__ blocks,
                                                                const float* __restrict__ da, float* __restrict__ work) {
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  if (i >= br.end) return;
  const StepParams P = A.params[br.mesh];
  const SoA& S = A.in;
  const f3 p = ld3(S.x, S.y, S.z, i);
  const f3 v = ld3(S.vx, S.vy, S.vz, i);
  if (pinned(A.pins, i)) {
    store_state(A.out, i, p, v);
    work[i] = 0.0f;
    return;
  }
  const float m = P.uniform_mass ? P.mass : A.mass[i];
  f3 Ff = f3{0.0f, 0.0f, 0.0f};
  for (uint32_t e = 0, nf = T.f_cnt[i]; e < nf; ++e) {
    const float4 fc = T.face_c[T.f_first[i] + kSlice * e];
    Ff = add(Ff, f3{fc.x, fc.y, fc.z});
  }
  f3 Fs = f3{0.0f, 0.0f, 0.0f};
  if (STRUTS)
    for (uint32_t e = 0, ns = T.s_cnt[i]; e < ns; ++e) {
      const StrutEnt se = T.s_ent[T.s_first[i] + kSlice * e];
      strut_accumulate<FastMath>(Fs, p, v, ld3(S.x, S.y, S.z, se.nbr), ld3(S.vx, S.vy, S.vz, se.nbr), se.k, se.c,
                                 se.len);
    }
  const f3 d = f3{da[3 * (size_t)i], da[3 * (size_t)i + 1], da[3 * (size_t)i + 2]};
  f3 xn, vn;
  const float w = sur_update(P, A.obs, m, p, v, Ff, Fs, d, xn, vn);
  store_state(A.out, i, xn, vn);
  work[i] = w;
}

// ---------------------------------------------------------------------------------------------
// Diagnostics of a big step over the vertices of `blocks`: the largest change of any strut strain
// (each strut seen from both ends; NaN propagates), energies before and after (kinetic, spring
// with each strut counted half from each end, gravity potential -F_g . x), the face-force work
// subtracted from the energy after.
__global__ void __launch_bounds__(kBlock) sur_diag_kernel(SoA I, SoA O, const SlotEnt* __restrict__ slots,
                                                           const float* __restrict__ slot_k,
                                                           const StepParams* __restrict__ params,
                                                           const float* __restrict__ mass, const float* __restrict__ work,
                                                           const BlockRange* __restrict__ blocks, DiagOut* __restrict__ out) {
  __shared__ float s_ds[kBlock], s_eo[kBlock], s_en[kBlock], s_em[kBlock];
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t iv = br.begin + threadIdx.x;
  const bool valid = iv < br.end;
  const uint32_t i = valid ? iv : br.begin;
  float ds = 0.0f, eo = 0.0f, en = 0.0f, em = 0.0f;
  {
    const StepParams P = params[br.mesh];
    const float m = P.uniform_mass ? P.mass : mass[i];
    const f3 po = ld3(I.x, I.y, I.z, i), pn = ld3(O.x, O.y, O.z, i);
    const f3 vo = ld3(I.vx, I.vy, I.vz, i), vn = ld3(O.vx, O.vy, O.vz, i);
    // all slot entries first, then all neighbour positions (empty slots read the vertex itself), so
    // the loads are in flight together
    SlotEnt e[kSlots];
    float k[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      e[j] = slots[(size_t)i * kSlots + j];
      k[j] = slot_k[(size_t)i * kSlots + j];
    }
    f3 qo[kSlots], qn[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const uint32_t nb = e[j].nbr != kNoSlot ? e[j].nbr : i;
      qo[j] = ld3(I.x, I.y, I.z, nb);
      qn[j] = ld3(O.x, O.y, O.z, nb);
    }
    float so = 0.0f, sn = 0.0f;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool has = e[j].nbr != kNoSlot;
      const f3 d_o = sub(qo[j], po), d_n = sub(qn[j], pn);
      const float r = has ? e[j].rest : 1.0f;
      const float lo = __builtin_sqrtf(dot(d_o, d_o)) - r, ln = __builtin_sqrtf(dot(d_n, d_n)) - r;
      const float d = has ? __builtin_fabsf(ln - lo) * __builtin_amdgcn_rcpf(r) : 0.0f;
      ds = (d > ds || d != d) ? d : ds;
      const float kk = has ? 0.25f * k[j] : 0.0f;
      so += kk * lo * lo;
      sn += kk * ln * ln;
    }
    const float ko = 0.5f * m * dot(vo, vo), kn = 0.5f * m * dot(vn, vn);
    em = ko + so + kn + sn;
    eo = ko + so - dot(P.gravity, po);
    en = kn + sn - dot(P.gravity, pn) - work[i];
    if (!valid) ds = eo = en = em = 0.0f;
  }
  s_ds[threadIdx.x] = ds; s_eo[threadIdx.x] = eo; s_en[threadIdx.x] = en; s_em[threadIdx.x] = em;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const float a = s_ds[threadIdx.x], b = s_ds[threadIdx.x + w];
      s_ds[threadIdx.x] = (b > a || b != b) ? b : a;
      s_eo[threadIdx.x] += s_eo[threadIdx.x + w];
      s_en[threadIdx.x] += s_en[threadIdx.x + w];
      s_em[threadIdx.x] += s_em[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = DiagOut{s_ds[0], s_eo[0], s_en[0], s_em[0]};
}

// Health of the plain trajectories that feed training: per block, the largest of |v| / vmax and
// strain / smax over its vertices (NaN and infinities count as unhealthy: the result is then +inf).
__global__ void __launch_bounds__(kBlock) sur_health_kernel(SoA S, const SlotEnt* __restrict__ slots,
                                                             const BlockRange* __restrict__ blocks, float vmax,
                                                             float smax, float* __restrict__ out) {
  __shared__ float s_h[kBlock];
  const BlockRange br = blocks[blockIdx.x];
  const uint32_t i = br.begin + threadIdx.x;
  float h = 0.0f;
  if (i < br.end) {
    const f3 p = ld3(S.x, S.y, S.z, i), v = ld3(S.vx, S.vy, S.vz, i);
    h = __builtin_sqrtf(dot(v, v)) / vmax;
    for (int j = 0; j < kSlots; ++j) {
      const SlotEnt e = slots[(size_t)i * kSlots + j];
      if (e.nbr == kNoSlot) continue;
      const f3 d = sub(ld3(S.x, S.y, S.z, e.nbr), p);
      h = fmaxf(h, __builtin_fabsf(__builtin_sqrtf(dot(d, d)) - e.rest) / (e.rest * smax));
    }
    if (!(h <= 3.0e38f)) h = __builtin_inff();
  }
  s_h[threadIdx.x] = h;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s_h[threadIdx.x] = fmaxf(s_h[threadIdx.x], s_h[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s_h[0];
}

__global__ void __launch_bounds__(kBlock) sur_features_kernel(SoA S, const SlotEnt* __restrict__ slots,
                                                               const uint32_t* __restrict__ ids, uint32_t n,
                                                               FeatParams fp, float* __restrict__ out) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = ids[j];
  const f3 p = ld3(S.x, S.y, S.z, i), v = ld3(S.vx, S.vy, S.vz, i);
  for (int s = 0; s < kSlots; ++s) slot_feat(S, p, v, i, slots[(size_t)i * kSlots + s], fp, out + (size_t)j * kFeat + 6 * s);
}

__global__ void __launch_bounds__(kBlock) sur_noise_kernel(Pos3 prev, Pos3 cur, SoA o, const uint32_t* __restrict__ pins,
                                                            uint32_t n, float H, float sigma, uint64_t seed) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float e[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (sigma > 0.0f && !pinned(pins, i)) {
    for (int k = 0; k < 3; ++k) {
      const uint64_t h = mix64(seed ^ mix64(((uint64_t)i << 3) | (uint64_t)k));
      const float r = sigma * sqrtf(-2.0f * logf(unit_open(h))), th = 6.283185307179586f * unit_open(mix64(h));
      e[2 * k] = r * cosf(th);
      e[2 * k + 1] = r * sinf(th);
    }
  }
  const float xb = cur.x[i] + e[0], yb = cur.y[i] + e[1], zb = cur.z[i] + e[2];
  const float xp = prev.x[i] + e[3], yp = prev.y[i] + e[4], zp = prev.z[i] + e[5];
  const float iH = 1.0f / H;
  o.x[i] = xb; o.y[i] = yb; o.z[i] = zb;
  o.vx[i] = (xb - xp) * iH; o.vy[i] = (yb - yp) * iH; o.vz[i] = (zb - zp) * iH;
}

template <bool STRUTS>
__global__ void __launch_bounds__(kBlock) sur_samples_kernel(StepArgs A, CsrView T, const uint32_t* __restrict__ ids,
                                                              const uint32_t* __restrict__ mesh_of, uint32_t n, Pos3 nx,
                                                              const SlotEnt* __restrict__ slots, FeatParams fp,
                                                              float out_scale, __bf16* __restrict__ X,
                                                              float* __restrict__ Tg) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = ids[j];
  const StepParams P = A.params[mesh_of[i]];
  const SoA& S = A.in;
  const f3 p = ld3(S.x, S.y, S.z, i), v = ld3(S.vx, S.vy, S.vz, i);
  float f[kFeat];
  bool ok = true;
  for (int s = 0; s < kSlots; ++s) slot_feat(S, p, v, i, slots[(size_t)i * kSlots + s], fp, f + 6 * s);
  for (int c = 0; c < kFeat; ++c) ok = ok && __builtin_isfinite(f[c]);
  const float m = P.uniform_mass ? P.mass : A.mass[i];
  f3 F = P.gravity;
  for (uint32_t e = 0, nf = T.f_cnt[i]; e < nf; ++e) {
    const float4 fc = T.face_c[T.f_first[i] + kSlice * e];
    F = add(F, f3{fc.x, fc.y, fc.z});
  }
  if (STRUTS)
    for (uint32_t e = 0, ns = T.s_cnt[i]; e < ns; ++e) {
      const StrutEnt se = T.s_ent[T.s_first[i] + kSlice * e];
      strut_accumulate<FastMath>(F, p, v, ld3(S.x, S.y, S.z, se.nbr), ld3(S.vx, S.vy, S.vz, se.nbr), se.k, se.c, se.len);
    }
  const float H = fp.H, iH = 1.0f / H, iH2 = iH * iH, im = 1.0f / m, is = 1.0f / out_scale;
  float t[3] = {((nx.x[i] - p.x) * iH2 - v.x * iH - F.x * im) * is, ((nx.y[i] - p.y) * iH2 - v.y * iH - F.y * im) * is,
                ((nx.z[i] - p.z) * iH2 - v.z * iH - F.z * im) * is};
  for (int c = 0; c < 3; ++c) ok = ok && __builtin_isfinite(t[c]);
  // a non-finite sample (from a trajectory the health check has not yet dropped) becomes a zero row
  for (int c = 0; c < kFeat; ++c) X[(size_t)j * kFeat + c] = (__bf16)(ok ? f[c] : 0.0f);
  for (int c = 0; c < 3; ++c) Tg[3 * (size_t)j + c] = ok ? t[c] : 0.0f;
}

// ---------------------------------------------------------------------------------------------
// Training GEMM: 64 x 64 C tile per 256-thread workgroup (4 waves of 32 x 32 = 2 x 2 MFMA tiles),
// K in steps of 32 through LDS.  Operands are read at arbitrary strides (the transposes of the
// backward pass), staged as As[m][k] and Bs[n][k] so every MFMA operand is one 16-byte LDS read.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const __bf16* __restrict__ A, long sam,
                                                        long sak, const __bf16* __restrict__ B, long sbk, long sbn,
                                                        float* __restrict__ C, long scm, long scn, int accumulate) {
  __shared__ __attribute__((aligned(16))) __bf16 As[64][40];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[64][40];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, col = lane & 15, g = lane >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const __bf16 zero = (__bf16)0.0f;
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + e * 256;
      int mm, kk;
      if (sak == 1) { mm = idx >> 5; kk = idx & 31; } else { kk = idx >> 6; mm = idx & 63; }
      const int gm = m0 + mm, gk = k0 + kk;
      As[mm][kk] = (gm < M && gk < K) ? A[(long)gm * sam + (long)gk * sak] : zero;
      int nn;
      if (sbk == 1) { nn = idx >> 5; kk = idx & 31; } else { kk = idx >> 6; nn = idx & 63; }
      const int gn = n0 + nn, gk2 = k0 + kk;
      Bs[nn][kk] = (gn < N && gk2 < K) ? B[(long)gk2 * sbk + (long)gn * sbn] : zero;
    }
    __syncthreads();
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + col][g * 8]);
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + col][g * 8]);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * g + r, nn = n0 + wn + 16 * j + col;
        if (m < M && nn < N) {
          float* c = C + (long)m * scm + (long)nn * scn;
          *c = accumulate ? *c + acc[i][j][r] : acc[i][j][r];
        }
      }
}

__global__ void __launch_bounds__(kBlock) bias_act_kernel(float* __restrict__ Z, const float* __restrict__ bias,
                                                           __bf16* __restrict__ H, long count, int n, int act) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const float z = Z[i] + bias[i % n];
  Z[i] = z;
  H[i] = (__bf16)(act ? gelu(z) : z);
}

__global__ void __launch_bounds__(kBlock) gelu_bwd_kernel(const float* __restrict__ dH, const float* __restrict__ Z,
                                                           __bf16* __restrict__ dZ, long count) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  dZ[i] = (__bf16)(dH[i] * gelu_grad(Z[i]));
}

__global__ void __launch_bounds__(kBlock) colsum_kernel(const __bf16* __restrict__ dZ, int rows, int n,
                                                         float* __restrict__ db) {
  __shared__ float s[kBlock];
  const int j = blockIdx.x;
  float acc = 0.0f;
  for (int r = threadIdx.x; r < rows; r += kBlock) acc += (float)dZ[(long)r * n + j];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = s[0];
}

__global__ void __launch_bounds__(kBlock) mse_kernel(float* __restrict__ Y, const float* __restrict__ b4,
                                                      const float* __restrict__ T, int rows, __bf16* __restrict__ dY,
                                                      float* __restrict__ partial) {
  __shared__ float s[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float e2 = 0.0f;
  if (i < 3L * rows) {
    const float y = Y[i] + b4[i % 3];
    Y[i] = y;
    const float d = y - T[i];
    e2 = d * d;
    dY[i] = (__bf16)(2.0f * d / (3.0f * (float)rows));
  }
  s[threadIdx.x] = e2;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kBlock) sum_kernel(const float* __restrict__ partial, int n, float scale,
                                                      float* __restrict__ out) {
  __shared__ float s[kBlock];
  float acc = 0.0f;
  for (int i = threadIdx.x; i < n; i += kBlock) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0] * scale;
}

__global__ void __launch_bounds__(kBlock) adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                       float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                       float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(kBlock) to_bf16_kernel(const float* __restrict__ in, __bf16* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (__bf16)in[i];
}

__global__ void __launch_bounds__(kBlock) gather_batch_kernel(const __bf16* __restrict__ X, const float* __restrict__ T,
                                                               uint64_t count, uint64_t seed, int batch,
                                                               __bf16* __restrict__ Xb, float* __restrict__ Tb) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= batch) return;
  const uint64_t r = mix64(seed ^ mix64((uint64_t)j)) % count;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(X + r * kFeat);
  uint32_t* dst = reinterpret_cast<uint32_t*>(Xb + (uint64_t)j * kFeat);
  for (int k = 0; k < kFeat / 2; ++k) dst[k] = src[k];
  for (int c = 0; c < 3; ++c) Tb[3 * (uint64_t)j + c] = T[3 * r + c];
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_sur_mlp(const SoA& st, const SlotEnt* slots, uint32_t n, const void* image, FeatParams fp,
                          float out_scale, float* da, int grid, int variant, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int tiles = (int)((n + kMlpTile - 1) / kMlpTile);
  const dim3 g(grid < tiles ? grid : tiles);
  if (variant == 4)
    hipLaunchKernelGGL(sur_mlp_kernel<4>, g, dim3(512), 0, s, st, slots, n, reinterpret_cast<const uint4*>(image), fp,
                       out_scale, da);
  else
    hipLaunchKernelGGL(sur_mlp_kernel<2>, g, dim3(1024), 0, s, st, slots, n, reinterpret_cast<const uint4*>(image), fp,
                       out_scale, da);
  return hipGetLastError();
}

hipError_t launch_sur_integrate(const StepArgs& a, const CsrView& t, const BlockRange* blocks, uint32_t n_blocks,
                                const float* da, float* work, bool struts, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  if (struts)
    hipLaunchKernelGGL(sur_integrate_kernel<true>, dim3(n_blocks), dim3(kBlock), 0, s, a, t, blocks, da, work);
  else
    hipLaunchKernelGGL(sur_integrate_kernel<false>, dim3(n_blocks), dim3(kBlock), 0, s, a, t, blocks, da, work);
  return hipGetLastError();
}

hipError_t launch_sur_diag(const SoA& in, const SoA& out, const SlotEnt* slots, const float* slot_k,
                           const StepParams* params, const float* mass, const float* work, const BlockRange* blocks,
                           uint32_t n_blocks, DiagOut* diag, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sur_diag_kernel, dim3(n_blocks), dim3(kBlock), 0, s, in, out, slots, slot_k, params, mass, work,
                     blocks, diag);
  return hipGetLastError();
}

hipError_t launch_sur_health(const SoA& st, const SlotEnt* slots, const BlockRange* blocks, uint32_t n_blocks,
                             float vmax, float smax, float* out, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sur_health_kernel, dim3(n_blocks), dim3(kBlock), 0, s, st, slots, blocks, vmax, smax, out);
  return hipGetLastError();
}

hipError_t launch_sur_features(const SoA& st, const SlotEnt* slots, const uint32_t* ids, uint32_t n, FeatParams fp,
                               float* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(sur_features_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, st, slots, ids, n, fp, out);
  return hipGetLastError();
}

hipError_t launch_sur_noise(Pos3 prev, Pos3 cur, SoA noisy, const uint32_t* pins, uint32_t n, float H, float sigma,
                            uint64_t seed, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(sur_noise_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, prev, cur, noisy, pins, n, H, sigma,
                     seed);
  return hipGetLastError();
}

hipError_t launch_sur_samples(const StepArgs& a, const CsrView& t, const uint32_t* ids, const uint32_t* mesh_of,
                              uint32_t n, Pos3 next, const SlotEnt* slots, FeatParams fp, float out_scale, bool struts,
                              __bf16* X, float* T, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (struts)
    hipLaunchKernelGGL(sur_samples_kernel<true>, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, t, ids, mesh_of, n, next,
                       slots, fp, out_scale, X, T);
  else
    hipLaunchKernelGGL(sur_samples_kernel<false>, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, t, ids, mesh_of, n, next,
                       slots, fp, out_scale, X, T);
  return hipGetLastError();
}

hipError_t launch_gemm_bf16(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbk,
                            long sbn, float* C, long scm, long scn, bool accumulate, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, M, N, K, A, sam, sak, B,
                     sbk, sbn, C, scm, scn, accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_bias_act(float* Z, const float* bias, __bf16* H, int rows, int n, bool act, hipStream_t s) {
  const long count = (long)rows * n;
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, s, Z, bias, H, count, n, act ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_gelu_bwd(const float* dH, const float* Z, __bf16* dZ, long count, int n, hipStream_t s) {
  (void)n;
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, s, dH, Z, dZ, count);
  return hipGetLastError();
}

hipError_t launch_colsum(const __bf16* dZ, int rows, int n, float* db, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(colsum_kernel, dim3(n), dim3(kBlock), 0, s, dZ, rows, n, db);
  return hipGetLastError();
}

hipError_t launch_mse(float* Y, const float* b4, const float* T, int rows, __bf16* dY, float* partial, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(mse_kernel, dim3(blocks_for(3L * rows)), dim3(kBlock), 0, s, Y, b4, T, rows, dY, partial);
  return hipGetLastError();
}

hipError_t launch_sum(const float* partial, int n, float scale, float* out, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kBlock), 0, s, partial, n, scale, out);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* in, __bf16* out, long n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_gather_batch(const __bf16* X, const float* T, uint64_t count, uint64_t seed, int batch, __bf16* Xb,
                               float* Tb, hipStream_t s) {
  if (batch == 0 || count == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_batch_kernel, dim3(blocks_for(batch)), dim3(kBlock), 0, s, X, T, count, seed, batch, Xb,
                     Tb);
  return hipGetLastError();
}

}  // namespace fp
