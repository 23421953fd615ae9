// The C ABI's opaque host-side handles that more than one translation unit opens: meshes and
// obstacles built on the host, moved into a simulation by fp_sim_new / fp_rigid_sim_new.
#pragma once

#include "mesh_host.hpp"

struct fp_mesh {
  fp::HostMesh m;
};
struct fp_obstacle {
  fp::HostObstacle o;
};
