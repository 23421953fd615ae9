// Fast-math LDS-tiled step kernel for Cloth::new meshes: one workgroup per tile, the three phases
// of grid_fast_core.hpp separated by barriers.  The host deals the tiles to the XCDs in runs of
// consecutive tiles (as for the exact tiled kernel), so vertically adjacent tiles mostly share an L2.
#include "grid_fast.hpp"
#include "grid_fast_core.hpp"

namespace fp {

namespace {

// At most 72 VGPRs per lane (room for seven waves per SIMD); the LDS, 23,232 B per workgroup with
// 15x15 tiles, allocated in 512-byte granules, holds six workgroups per CU.  (Entries of 12 bytes
// in the strut table bring it to 22,960 B and seven workgroups per CU: measured no faster.)
template <int TR, int TC, int MODE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 8)))
grid_fast_kernel(StepArgs A, const ClothView* __restrict__ cloths, const TileRange* __restrict__ tiles, HitBox hb) {
  static_assert((TR + 1) * (TC + 1) == (int)kBlock, "one quad per thread");
  __shared__ gfast::Lds<TR, TC> L;
  const TileRange T = tiles[blockIdx.x];
  const ClothView G = cloths[T.mesh];
  const StepParams P = A.params[T.mesh];
  const int r0 = (int)T.r0, c0 = (int)T.c0, tid = (int)threadIdx.x;
  // the obstacle response of update_points for segments that can reach an obstacle face
  auto collide = [&](f3 p, f3 v, f3 F, f3& xn, f3& vn) {
    const float m = A.obs.margin;
    const bool near = fminf(p.x, xn.x) - m <= hb.hi[0] && fmaxf(p.x, xn.x) + m >= hb.lo[0] &&
                      fminf(p.y, xn.y) - m <= hb.hi[1] && fmaxf(p.y, xn.y) + m >= hb.lo[1] &&
                      fminf(p.z, xn.z) - m <= hb.hi[2] && fmaxf(p.z, xn.z) + m >= hb.lo[2];
    if (near) {
      const int32_t hit = first_hit(A.obs, p, xn, P.dt_d);
      if (hit >= 0) collide_response(A.obs.faces[hit], P, P.mass, p, v, F, xn, vn);
    }
  };
  // Phase 0 runs at a raised issue priority: a workgroup that has just started gets its loads out
  // and its strut table computed ahead of the phase-1 arithmetic of the workgroups already resident
  // on the CU (measured 24.1 -> 23.8 us per step on a 1024 x 1024 cloth).
  if (gfast::interior_tile<TR, TC>(G, r0, c0)) {
    __builtin_amdgcn_s_setprio(2);
    gfast::phase0<TR, TC, true>(L, A, G, r0, c0, tid, (int)kBlock);
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    const f3 toA = gfast::phase1<TR, TC, true>(L, P, G, r0, c0, tid);
    __syncthreads();
    gfast::phase2<TR, TC, MODE, true>(L, A, P, G, r0, c0, tid, toA, collide);
  } else {
    __builtin_amdgcn_s_setprio(2);
    gfast::phase0<TR, TC, false>(L, A, G, r0, c0, tid, (int)kBlock);
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    const f3 toA = gfast::phase1<TR, TC, false>(L, P, G, r0, c0, tid);
    __syncthreads();
    gfast::phase2<TR, TC, MODE, false>(L, A, P, G, r0, c0, tid, toA, collide);
  }
}

template <int TR, int TC>
hipError_t launch_shape(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n, int mode,
                        const HitBox& hb, hipStream_t s) {
  if (mode == MODE_STEP) grid_fast_kernel<TR, TC, MODE_STEP><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles, hb);
  else if (mode == MODE_FORCES)
    grid_fast_kernel<TR, TC, MODE_FORCES><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles, hb);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace

hipError_t launch_grid_fast(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n_tiles,
                            int shape, int mode, const HitBox& box, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  switch (shape) {
    case TILE_15x15: return launch_shape<15, 15>(a, cloths, tiles, n_tiles, mode, box, s);
    case TILE_7x31: return launch_shape<7, 31>(a, cloths, tiles, n_tiles, mode, box, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fp
