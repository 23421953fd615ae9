// Static obstacles on the device: the faces of every CollidableMesh of a simulation (in list order,
// corners, unit normal and dropped axis) and the broad-phase grids over them, shared by the springy
// and rigid-body steps (both call CollidableMesh::get_collided_face_from_list, collidable_mesh.rs:107-169).
#pragma once

#include <vector>

#include "grid_fast.hpp"
#include "host_util.hpp"
#include "mesh_host.hpp"
#include "step_device.hpp"

namespace fph {

struct ObstacleSet {
  DBuf<fp::ObsFace> faces;
  DBuf<uint32_t> cell_start[4], cell_faces[4], always;
  uint64_t n_always = 0, n_binned[4] = {0, 0, 0, 0};
  fp::ObsView view{};
  fp::HitBox hit_box{};  // for the fast kernel's early-out (see fp::HitBox)
  uint64_t n_vertices = 0, n_faces = 0;
};

// Uploads the faces of `obstacles` (in order) and builds the grids; FP_OK or a status with its message.
int upload_obstacles(ObstacleSet& s, const std::vector<fp::HostObstacle>& obstacles);

}  // namespace fph
