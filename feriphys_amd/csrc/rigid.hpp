// Batched rigid bodies (rigidbody/rigidbody.rs, rigidbody/simulation.rs): device record layout,
// step constants and the step launcher.
#pragma once

#include "kernels.hpp"

namespace fp {

// A body's record: the reference's 29-float Stateful vector (rigidbody.rs:62-93: position,
// rotation v then s, linear momentum, angular momentum, mass, initial inverse inertia by columns,
// accumulated force, accumulated torque).  On the device the records are stored field-major:
// field k of body b at k * pitch + b (pitch = the body count rounded up to kRigidAlign), so that
// a wavefront's load of one field reads consecutive floats.
constexpr int kRigidRecord = 29;
constexpr uint64_t kRigidAlign = 64;

// The vertices of the unit cube every RigidBody::new builds (graphics/forms.rs:91-146), in its order.
constexpr float kCubeCorner[8][3] = {{-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f},   {0.5f, 0.5f, 0.5f},
                                     {-0.5f, 0.5f, 0.5f},  {-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f},
                                     {0.5f, 0.5f, -0.5f},  {-0.5f, 0.5f, -0.5f}};

// rigidbody/config.rs:7-13 plus Duration(dt) (the collision test's timestep).
struct RigidParams {
  int32_t integration;  // FP_EULER / FP_RK4
  float dt;
  float dt_d;  // Duration::from_secs_f32(dt).as_secs_f32()
  float restitution;
  f3 gravity;  // added to the accumulated force each step
  f3 torque;   // added to the accumulated torque each step
};

// pack: field-major records `in` -> 29 floats per body, body after body, in `out`; otherwise the
// inverse.
hipError_t launch_rigid_records(const float* in, float* out, uint32_t n, uint32_t pitch, bool pack, hipStream_t s);

// One Simulation::step of n bodies: records in -> out (field-major, `pitch` floats per field).
hipError_t launch_rigid_step(const float* in, float* out, uint32_t n, uint32_t pitch, const RigidParams& p,
                             const ObsView& obs, hipStream_t s);

}  // namespace fp
