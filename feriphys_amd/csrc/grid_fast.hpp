// Launch interface of the fast-math tiled grid kernel (grid_fast.hip).
#pragma once

#include "kernels.hpp"

namespace fp {

// The step (MODE_STEP) or the forces (MODE_FORCES) of Cloth::new meshes with the fast math
// policy, one 256-thread workgroup per tile of shape TILE_15x15 or TILE_7x31 ((TR+1) x (TC+1) =
// 256 quads, one per thread).
hipError_t launch_grid_fast(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n_tiles,
                            int shape, int mode, hipStream_t s);

}  // namespace fp
