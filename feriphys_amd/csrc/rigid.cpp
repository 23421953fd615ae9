// C ABI of the rigid-body step (include/feriphys_gpu.h, "rigid bodies"): RigidBody::new on the
// host, rigidbody::Simulation on the device (rigid_kernels.hip), one 29-float record per body.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <vector>

#include "../../include/feriphys_gpu.h"
#include "handles.hpp"
#include "host_util.hpp"
#include "obstacles.hpp"
#include "rigid.hpp"

using fph::DBuf;
using fph::fail;

struct fp_rigid {
  float rec[fp::kRigidRecord];
  bool moved = false;
};

struct fp_rigid_sim {
  uint64_t n = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t pitch = 0;  // floats per field of the device records (n rounded up to kRigidAlign)
  DBuf<float> rec[2];  // ping-pong, field-major (rigid.hpp)
  DBuf<float> stage;   // the caller's layout (29 floats per body) for one copy each way
  int cur = 0;
  fph::ObstacleSet obst;
  fp_rigid_config cfg{};
  ~fp_rigid_sim() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// cgmath's Matrix3 (columns) operations that RigidBody::new uses, in their component order.
struct HM3 {
  fp::f3 c[3];
};
float h_det(const HM3& m) {  // first-column cofactor expansion
  const fp::f3 &a = m.c[0], &b = m.c[1], &c = m.c[2];
  return a.x * (b.y * c.z - c.y * b.z) - b.x * (a.y * c.z - c.y * a.z) + c.x * (a.y * b.z - b.y * a.z);
}
// Matrix3::invert: columns (c1 x c2, c2 x c0, c0 x c1) / det, transposed; false when det == 0.
bool h_invert(const HM3& m, HM3& out) {
  const float det = h_det(m);
  if (det == 0.0f) return false;
  const fp::f3 k[3] = {fp::sdiv(fp::cross(m.c[1], m.c[2]), det), fp::sdiv(fp::cross(m.c[2], m.c[0]), det),
                       fp::sdiv(fp::cross(m.c[0], m.c[1]), det)};
  for (int i = 0; i < 3; ++i)  // transpose
    out.c[i] = fp::f3{i == 0 ? k[0].x : i == 1 ? k[0].y : k[0].z, i == 0 ? k[1].x : i == 1 ? k[1].y : k[1].z,
                      i == 0 ? k[2].x : i == 1 ? k[2].y : k[2].z};
  return true;
}
fp::f3 h_mv(const HM3& m, fp::f3 v) {  // row(i) . v
  return fp::f3{fp::dot(fp::f3{m.c[0].x, m.c[1].x, m.c[2].x}, v), fp::dot(fp::f3{m.c[0].y, m.c[1].y, m.c[2].y}, v),
                fp::dot(fp::f3{m.c[0].z, m.c[1].z, m.c[2].z}, v)};
}

int check_config(const fp_rigid_config* c) {
  if (!c) return fail(FP_EINVAL, "config is NULL");
  if (c->integration != FP_EULER && c->integration != FP_RK4) return fail(FP_EINVAL, "unknown integration");
  float dd = 0.0f;
  if (!fph::duration_roundtrip(c->dt, dd)) return fail(FP_EINVAL, "dt must be finite and non-negative (Duration)");
  return FP_OK;
}

fp::RigidParams params_of(const fp_rigid_config& c) {
  fp::RigidParams p{};
  p.integration = c.integration;
  p.dt = c.dt;
  fph::duration_roundtrip(c.dt, p.dt_d);
  p.restitution = c.coefficient_of_restitution;
  p.gravity = fp::f3{c.gravity[0], c.gravity[1], c.gravity[2]};
  p.torque = fp::f3{c.torque[0], c.torque[1], c.torque[2]};
  return p;
}

}  // namespace

extern "C" {

void fp_rigid_config_default(fp_rigid_config* out) {
  if (!out) return;
  fp_rigid_config c{};
  c.integration = FP_RK4;
  c.dt = 0.001f;  // Duration::from_millis(1).as_secs_f32()
  c.coefficient_of_restitution = 0.7f;
  *out = c;  // gravity and torque zero
}

int fp_rigid_new(const float position[3], float mass, fp_rigid** out) {
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  if (!position) return fail(FP_EINVAL, "position is NULL");
  std::unique_ptr<fp_rigid> b(new (std::nothrow) fp_rigid());
  if (!b) return fail(FP_ENOMEM, "out of host memory");
  const float i = mass / 6.0f;  // the unit cube's moment of inertia (rigidbody.rs:217-227)
  const HM3 inertia{{fp::f3{i, 0.0f, 0.0f}, fp::f3{0.0f, i, 0.0f}, fp::f3{0.0f, 0.0f, i}}};
  HM3 inv;
  if (!h_invert(inertia, inv)) return fail(FP_EINVAL, "Uninvertable moment of inertia!");
  const fp::f3 p = fp::smul(mass, fp::f3{0.0f, 0.0f, 0.0f});  // mass * Vector3::zero()
  const fp::f3 l = h_mv(inertia, fp::f3{0.0f, 0.0f, 0.0f});     // moment_of_inertia * Vector3::zero()
  float* r = b->rec;
  r[0] = position[0]; r[1] = position[1]; r[2] = position[2];
  r[3] = 0.0f; r[4] = 0.0f; r[5] = 0.0f; r[6] = 1.0f;  // Quaternion::one()
  r[7] = p.x; r[8] = p.y; r[9] = p.z;
  r[10] = l.x; r[11] = l.y; r[12] = l.z;
  r[13] = mass;
  for (int c = 0; c < 3; ++c) { r[14 + 3 * c] = inv.c[c].x; r[15 + 3 * c] = inv.c[c].y; r[16 + 3 * c] = inv.c[c].z; }
  for (int k = 23; k < fp::kRigidRecord; ++k) r[k] = 0.0f;
  *out = b.release();
  return FP_OK;
}

int fp_rigid_get_state(const fp_rigid* b, float* record) {
  if (!b || !record) return fail(FP_EINVAL, "NULL argument");
  if (b->moved) return fail(FP_ESTATE, "body was moved into a simulation");
  std::memcpy(record, b->rec, sizeof(b->rec));
  return FP_OK;
}

// State::apply_impulse (rigidbody.rs:38-41): linear += impulse; angular += position x impulse
static void apply_impulse(float* r, const float* imp, const float* at) {
  const fp::f3 i{imp[0], imp[1], imp[2]}, a{at[0], at[1], at[2]};
  r[7] = r[7] + i.x; r[8] = r[8] + i.y; r[9] = r[9] + i.z;
  const fp::f3 d = fp::cross(a, i);
  r[10] = r[10] + d.x; r[11] = r[11] + d.y; r[12] = r[12] + d.z;
}

int fp_rigid_apply_impulse(fp_rigid* b, const float impulse[3], const float position[3]) {
  if (!b || !impulse || !position) return fail(FP_EINVAL, "NULL argument");
  if (b->moved) return fail(FP_ESTATE, "body was moved into a simulation");
  apply_impulse(b->rec, impulse, position);
  return FP_OK;
}

void fp_rigid_free(fp_rigid* b) { delete b; }

int fp_rigid_sim_new(fp_rigid* const* bodies, uint64_t n_bodies, fp_obstacle* const* obstacles, uint64_t n_obstacles,
                     fp_rigid_sim** out) {
  // ownership moves in, as Simulation::new takes the Vecs by value: free the handles on every path
  std::vector<std::unique_ptr<fp_rigid>> bs;
  std::vector<std::unique_ptr<fp_obstacle>> os;
  for (uint64_t i = 0; i < n_bodies; ++i)
    if (bodies && bodies[i]) bs.emplace_back(bodies[i]);
  for (uint64_t i = 0; i < n_obstacles; ++i)
    if (obstacles && obstacles[i]) os.emplace_back(obstacles[i]);
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  if (bs.size() != n_bodies || os.size() != n_obstacles) return fail(FP_EINVAL, "NULL body or obstacle handle");
  for (auto& b : bs)
    if (b->moved) return fail(FP_ESTATE, "body was moved into a simulation");
  if (n_bodies >= (1ull << 31)) return fail(FP_EINVAL, "too many bodies");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return fail(FP_ENODEV, "no HIP device");
  std::unique_ptr<fp_rigid_sim> s(new (std::nothrow) fp_rigid_sim());
  if (!s) return fail(FP_ENOMEM, "out of host memory");
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&s->ev0));
  HIP_TRY(hipEventCreate(&s->ev1));
  fp_rigid_config_default(&s->cfg);
  s->n = n_bodies;
  s->pitch = (n_bodies + fp::kRigidAlign - 1) / fp::kRigidAlign * fp::kRigidAlign;
  std::vector<float> h(fp::kRigidRecord * s->pitch + 1, 0.0f);
  for (uint64_t i = 0; i < n_bodies; ++i)
    for (int k = 0; k < fp::kRigidRecord; ++k) h[k * s->pitch + i] = bs[i]->rec[k];
  HIP_TRY(s->rec[0].upload(h.data(), h.size()));
  HIP_TRY(s->rec[1].upload(h.data(), h.size()));
  std::vector<fp::HostObstacle> obs;
  for (auto& o : os) obs.push_back(std::move(o->o));
  if (int e = fph::upload_obstacles(s->obst, obs)) return e;
  HIP_TRY(hipDeviceSynchronize());
  *out = s.release();
  return FP_OK;
}

void fp_rigid_sim_free(fp_rigid_sim* s) { delete s; }

int fp_rigid_sim_count(const fp_rigid_sim* s, uint64_t* n_bodies) {
  if (!s || !n_bodies) return fail(FP_EINVAL, "NULL argument");
  *n_bodies = s->n;
  return FP_OK;
}

int fp_rigid_sim_set_config(fp_rigid_sim* s, const fp_rigid_config* c) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (int e = check_config(c)) return e;
  s->cfg = *c;
  return FP_OK;
}

int fp_rigid_sim_get_config(const fp_rigid_sim* s, fp_rigid_config* c) {
  if (!s || !c) return fail(FP_EINVAL, "NULL argument");
  *c = s->cfg;
  return FP_OK;
}

static int run_rigid(fp_rigid_sim* s, uint64_t n) {
  const fp::RigidParams p = params_of(s->cfg);
  for (uint64_t k = 0; k < n; ++k) {
    HIP_TRY(fp::launch_rigid_step(s->rec[s->cur].p, s->rec[s->cur ^ 1].p, (uint32_t)s->n, (uint32_t)s->pitch, p,
                                  s->obst.view, s->stream));
    s->cur ^= 1;
  }
  return FP_OK;
}

int fp_rigid_sim_step(fp_rigid_sim* s, uint64_t n_steps, float* timestep) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (int e = run_rigid(s, n_steps)) return e;
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (timestep) fph::duration_roundtrip(s->cfg.dt, *timestep);  // Duration::from_secs_f32(dt)
  return FP_OK;
}

int fp_rigid_sim_time_steps(fp_rigid_sim* s, uint64_t n_steps, float* ms) {
  if (!s || !ms) return fail(FP_EINVAL, "NULL argument");
  HIP_TRY(hipEventRecord(s->ev0, s->stream));
  if (int e = run_rigid(s, n_steps)) return e;
  HIP_TRY(hipEventRecord(s->ev1, s->stream));
  HIP_TRY(hipEventSynchronize(s->ev1));
  HIP_TRY(hipEventElapsedTime(ms, s->ev0, s->ev1));
  return FP_OK;
}

int fp_rigid_sim_get_states(fp_rigid_sim* s, float* records) {
  if (!s || (!records && s->n)) return fail(FP_EINVAL, "NULL argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->n == 0) return FP_OK;
  HIP_TRY(s->stage.reserve(fp::kRigidRecord * s->n));
  HIP_TRY(fp::launch_rigid_records(s->rec[s->cur].p, s->stage.p, (uint32_t)s->n, (uint32_t)s->pitch, true, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(records, s->stage.p, fp::kRigidRecord * s->n * sizeof(float), hipMemcpyDeviceToHost));
  return FP_OK;
}

int fp_rigid_sim_set_states(fp_rigid_sim* s, const float* records, uint64_t n) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (n != s->n) return fail(FP_EINVAL, "Incorrect size of state vector! (record count != body count)");
  if (n && !records) return fail(FP_EINVAL, "records is NULL");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (n == 0) return FP_OK;
  HIP_TRY(s->stage.reserve(fp::kRigidRecord * n));
  HIP_TRY(hipMemcpy(s->stage.p, records, fp::kRigidRecord * n * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(fp::launch_rigid_records(s->stage.p, s->rec[s->cur].p, (uint32_t)n, (uint32_t)s->pitch, false, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_rigid_sim_apply_impulse(fp_rigid_sim* s, uint64_t body, const float impulse[3], const float position[3]) {
  if (!s || !impulse || !position) return fail(FP_EINVAL, "NULL argument");
  if (body >= s->n) return fail(FP_EINVAL, "body index out of range");
  HIP_TRY(hipStreamSynchronize(s->stream));
  float r[fp::kRigidRecord];
  float* d = s->rec[s->cur].p + body;  // field k at d + k * pitch
  const size_t fp_bytes = s->pitch * sizeof(float);
  HIP_TRY(hipMemcpy2D(r, sizeof(float), d, fp_bytes, sizeof(float), fp::kRigidRecord, hipMemcpyDeviceToHost));
  apply_impulse(r, impulse, position);
  HIP_TRY(hipMemcpy2D(d, fp_bytes, r, sizeof(float), sizeof(float), fp::kRigidRecord, hipMemcpyHostToDevice));
  return FP_OK;
}

}  // extern "C"
