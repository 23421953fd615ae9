// Fast-math LDS-tiled step kernel for Cloth::new meshes: one workgroup per tile, the three phases
// of grid_fast_core.hpp separated by barriers.  Tiles are dealt to the XCDs in contiguous bands by
// the host (as for the exact tiled kernel), so vertically adjacent tiles share an L2.
#include "grid_fast.hpp"
#include "grid_fast_core.hpp"

namespace fp {

namespace {

template <int TR, int TC, int MODE>
__global__ void __launch_bounds__(kBlock) grid_fast_kernel(StepArgs A, const ClothView* __restrict__ cloths,
                                                           const TileRange* __restrict__ tiles) {
  static_assert((TR + 1) * (TC + 1) == (int)kBlock, "one quad per thread");
  __shared__ gfast::Lds<TR, TC> L;
  const TileRange T = tiles[blockIdx.x];
  const ClothView G = cloths[T.mesh];
  const StepParams P = A.params[T.mesh];
  const int r0 = (int)T.r0, c0 = (int)T.c0, tid = (int)threadIdx.x;
  gfast::phase0<TR, TC>(L, A, G, r0, c0, tid, (int)kBlock);
  __syncthreads();
  gfast::phase1<TR, TC>(L, P, G, r0, c0, tid);
  __syncthreads();
  gfast::phase2<TR, TC, MODE>(L, A, P, G, r0, c0, tid, [&](f3 p, f3 v, f3 F, f3& xn, f3& vn) {
    const int32_t hit = first_hit(A.obs, p, xn, P.dt_d);
    if (hit >= 0) collide_response(A.obs.faces[hit], P, P.mass, p, v, F, xn, vn);
  });
}

template <int TR, int TC>
hipError_t launch_shape(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n, int mode,
                        hipStream_t s) {
  if (mode == MODE_STEP) grid_fast_kernel<TR, TC, MODE_STEP><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles);
  else if (mode == MODE_FORCES) grid_fast_kernel<TR, TC, MODE_FORCES><<<dim3(n), dim3(kBlock), 0, s>>>(a, cloths, tiles);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace

hipError_t launch_grid_fast(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n_tiles,
                            int shape, int mode, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  switch (shape) {
    case TILE_15x15: return launch_shape<15, 15>(a, cloths, tiles, n_tiles, mode, s);
    case TILE_7x31: return launch_shape<7, 31>(a, cloths, tiles, n_tiles, mode, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fp
