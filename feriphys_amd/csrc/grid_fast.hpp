// Launch interface of the fast-math tiled grid kernel (grid_fast.hip).
#pragma once

#include "kernels.hpp"

namespace fp {

// A box outside which no segment can hit an obstacle face: the union of the faces' bounding boxes
// grown by the broad phase's margin (unbounded along the dropped axis of a thin face, everything
// when some face is tested for every vertex or the faces are few enough to be scanned in order, and
// empty without obstacles).  A segment whose grown bounding box misses it meets no grid cell that
// the first-hit search would look in.
struct HitBox {
  float lo[3], hi[3];
};

// The step (MODE_STEP) or the forces (MODE_FORCES) of Cloth::new meshes with the fast math
// policy, one 256-thread workgroup per tile of shape TILE_15x15 or TILE_7x31 ((TR+1) x (TC+1) =
// 256 quads, one per thread).
hipError_t launch_grid_fast(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n_tiles,
                            int shape, int mode, const HitBox& box, hipStream_t s);

}  // namespace fp
