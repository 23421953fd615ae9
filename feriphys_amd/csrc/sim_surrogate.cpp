// C ABI of the learned surrogate on a simulation (include/feriphys_gpu.h, "learned large-step
// surrogate"): feature slots, the big step with its per-mesh fallback, training-sample collection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sim_internal.hpp"

using namespace fpsim;
using fph::DBuf;
using fph::fail;

// ---------------------------------------------------------------------------------------------
// Learned surrogate on a simulation
namespace {

// Feature slots of every vertex (global ids): meshes built by Cloth::new by grid offset, other
// meshes their first 12 struts in strut order; with each slot's rest length and stiffness.
int build_slots(fp_sim* s) {
  auto& S = s->sur;
  const uint64_t N = s->N;
  S.h_slots.assign(N * fp::kSlots, fp::SlotEnt{fp::kNoSlot, 1.0f});
  std::vector<float> sk(N * fp::kSlots, 0.0f);
  std::vector<uint32_t> mesh_of(N), ids;
  std::vector<uint8_t> used(N, 0);
  static const int offs[10][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}, {-1, 1}, {1, -1}, {0, 2}, {0, -2}, {2, 0}, {-2, 0}};
  for (uint32_t mi = 0; mi < s->meshes.size(); ++mi) {
    const MeshSlot& ms = s->meshes[mi];
    const fp::HostMesh& h = ms.host;
    std::vector<uint8_t> pin(h.n, 0);
    for (uint64_t p : h.pins) pin[p] = 1;
    for (uint64_t i = 0; i < h.n; ++i) {
      mesh_of[ms.vbase + i] = mi;
      if (!pin[ms.ref(i)]) ids.push_back((uint32_t)(ms.vbase + i));
    }
    const bool grid = h.cloth.valid;
    const int64_t C = h.cloth.cols;
    for (uint64_t k = 0; k < h.n_struts(); ++k) {
      if (!(h.s_len[k] > 0.0f)) continue;
      for (int side = 0; side < 2; ++side) {
        const uint32_t self = side ? h.s_v1[k] : h.s_v0[k], other = side ? h.s_v0[k] : h.s_v1[k];
        const uint64_t gi = ms.dev(self);
        int slot = -1;
        if (grid) {
          const int64_t dr = (int64_t)(other / C) - (int64_t)(self / C), dc = (int64_t)(other % C) - (int64_t)(self % C);
          for (int j = 0; j < 10; ++j)
            if (offs[j][0] == dr && offs[j][1] == dc) slot = j;
        } else if (used[gi] < fp::kSlots) {
          slot = used[gi]++;
        }
        if (slot < 0) continue;
        S.h_slots[gi * fp::kSlots + slot] = fp::SlotEnt{(uint32_t)ms.dev(other), h.s_len[k]};
        sk[gi * fp::kSlots + slot] = h.s_k[k];
      }
    }
  }
  HIP_TRY(S.slots.upload(S.h_slots.data(), S.h_slots.size()));
  HIP_TRY(S.slot_k.upload(sk.data(), sk.size()));
  HIP_TRY(S.mesh_of.upload(mesh_of.data(), mesh_of.size()));
  HIP_TRY(S.ids.upload(ids.data(), ids.size()));
  return FP_OK;
}

int check_dt(const fp_sim* s, const fp_surrogate_config& c) {
  for (const auto& ms : s->meshes)
    if (!(ms.cfg.dt == c.dt)) return fail(FP_EINVAL, "a mesh's dt differs from the surrogate network's dt");
  return FP_OK;
}

int upload_big_params(fp_sim* s, const fp_surrogate_config& c) {
  std::vector<fp::StepParams> bp;
  for (const auto& ms : s->meshes) {
    fp_config k = ms.cfg;
    k.dt = c.dt * (float)c.ratio;
    bp.push_back(make_params(k, ms.params.mass, ms.uniform_mass));
  }
  HIP_TRY(s->sur.big_params.assign(bp.data(), bp.size()));
  return FP_OK;
}

int prepare_surrogate(fp_sim* s, const fp_surrogate_config& c) {
  if (int e = check_dt(s, c)) return e;
  if (!s->sur.slots.p && s->N) {
    if (int e = build_slots(s)) return e;
  }
  if (int e = upload_big_params(s, c)) return e;
  HIP_TRY(s->sur.da.reserve(3 * s->N + 1));
  HIP_TRY(s->sur.work.reserve(s->N + 1));
  HIP_TRY(s->sur.diag.reserve(s->h_blocks.size() + 1));
  return FP_OK;
}

fp::FeatParams feat_params(const fp_surrogate_config& c) {
  return fp::FeatParams{c.dt * (float)c.ratio, c.strain_scale};
}

// One big step of every mesh, with the per-mesh fallback.
int sur_big_step(fp_sim* s, fp_surrogate_stats& st, float* mlp_ms) {
  auto& S = s->sur;
  const uint32_t nblk = (uint32_t)s->h_blocks.size();
  const size_t M = s->meshes.size();
  S.skip.resize(M, 0);
  S.backoff.resize(M, 0);
  bool all_skip = true;
  for (size_t m = 0; m < M; ++m) all_skip = all_skip && S.skip[m] > 0;
  if (all_skip) {  // every mesh is backing off: plain steps, no network attempt
    for (size_t m = 0; m < M; ++m) --S.skip[m];
    ++st.big_steps;
    ++st.fallbacks;
    ++st.plain_only;
    st.mesh_fallbacks += M;
    return run_steps(s, (uint64_t)S.cfg.ratio);
  }
  fp::StepArgs a = s->args(false);
  a.params = S.big_params.p;
  a.da = S.da.p;
  a.work = S.work.p;
  // meshes off the tiled kernel: face kernel now, per-vertex update after the network; cloths on
  // the tiled kernel take their faces and the update in one LDS-tiled pass (same arithmetic)
  // (with base_struts every mesh takes the generic path: the tiled pass has no strut forces)
  const bool fused = S.cfg.base_struts == 0;
  const fp::BlockRange* fb = fused ? s->sur_face_blocks.p : s->all_face_blocks.p;
  const uint32_t nfb = fused ? s->n_sur_face_blocks : (uint32_t)s->h_face_blocks.size();
  const fp::BlockRange* vb = fused ? s->sur_blocks.p : s->all_blocks.p;
  const uint32_t nvb = fused ? s->n_sur_blocks : nblk;
  HIP_TRY(fp::launch_faces(a, s->csr(), fb, nfb, false, s->stream));
  if (mlp_ms) HIP_TRY(hipEventRecord(s->ev0, s->stream));
  HIP_TRY(fp::launch_sur_mlp(a.in, S.slots.p, (uint32_t)s->N, S.image.p, feat_params(S.cfg), S.cfg.out_scale, S.da.p,
                             S.grid, S.variant, s->stream));
  if (mlp_ms) HIP_TRY(hipEventRecord(s->ev1, s->stream));
  HIP_TRY(fp::launch_sur_integrate(a, s->csr(), vb, nvb, S.da.p, S.work.p, !fused, s->stream));
  if (fused)
    for (int sh = 0; sh < fp::N_TILE_SHAPES; ++sh)
      HIP_TRY(fp::launch_grid_tiled(a, s->cloths.p, s->tiles[sh].p, s->n_tiles[sh], sh, fp::MODE_SURROGATE, false,
                                    s->stream));
  HIP_TRY(fp::launch_sur_diag(a.in, a.out, S.slots.p, S.slot_k.p, S.big_params.p, s->mass.p, S.work.p, s->all_blocks.p,
                              nblk, S.diag.p, s->stream));
  std::vector<fp::DiagOut> d(nblk);
  HIP_TRY(hipMemcpyAsync(d.data(), S.diag.p, nblk * sizeof(fp::DiagOut), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (mlp_ms) {
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *mlp_ms += t;
  }
  std::vector<double> eo(M, 0.0), en(M, 0.0), em(M, 0.0);
  std::vector<float> ds(M, 0.0f);
  for (uint32_t b = 0; b < nblk; ++b) {
    const uint32_t m = s->h_blocks[b].mesh;
    const float x = d[b].max_dstrain;
    ds[m] = (x > ds[m] || x != x) ? x : ds[m];
    eo[m] += d[b].e_old;
    en[m] += d[b].e_new;
    em[m] += d[b].emech_old;
  }
  std::vector<uint8_t> bad(M, 0);
  size_t nbad = 0;
  for (size_t m = 0; m < M; ++m) {
    if (S.skip[m] > 0) {  // backing off: plain, whatever its residuals
      --S.skip[m];
      bad[m] = 1;
      ++nbad;
      continue;
    }
    const float er = (float)((en[m] - eo[m]) / std::max(em[m], 1e-30));
    bad[m] = !(ds[m] <= S.cfg.strain_limit) || !(er <= S.cfg.energy_limit);
    nbad += bad[m];
    S.backoff[m] = bad[m] ? std::min<uint32_t>(S.backoff[m] ? 2 * S.backoff[m] : 1, 16) : 0;
    S.skip[m] = bad[m] ? S.backoff[m] : 0;
    st.max_strain_residual = (ds[m] > st.max_strain_residual || ds[m] != ds[m]) ? ds[m] : st.max_strain_residual;
    st.max_energy_residual = (er > st.max_energy_residual || er != er) ? er : st.max_energy_residual;
  }
  ++st.big_steps;
  if (nbad == 0) {
    s->cur ^= 1;
    return FP_OK;
  }
  ++st.fallbacks;
  st.mesh_fallbacks += nbad;
  const int out = s->cur ^ 1;
  const bool keep = nbad < M;
  if (keep)
    for (int k = 0; k < 6; ++k) {
      HIP_TRY(S.keep[k].reserve(s->N));
      HIP_TRY(hipMemcpyAsync(S.keep[k].p, s->state[out][k].p, s->N * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    }
  if (int e = run_steps(s, (uint64_t)S.cfg.ratio)) return e;
  if (keep)
    for (size_t m = 0; m < M; ++m) {
      if (bad[m]) continue;
      const MeshSlot& ms = s->meshes[m];
      for (int k = 0; k < 6; ++k)
        HIP_TRY(hipMemcpyAsync(s->state[s->cur][k].p + ms.vbase, S.keep[k].p + ms.vbase, ms.host.n * sizeof(float),
                               hipMemcpyDeviceToDevice, s->stream));
    }
  return FP_OK;
}

int run_surrogate(fp_sim* s, uint64_t n_big, fp_surrogate_stats* stats, float* mlp_ms) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (!s->sur.on) return fail(FP_ESTATE, "no surrogate attached (fp_sim_set_surrogate)");
  if (int e = prepare_surrogate(s, s->sur.cfg)) return e;
  fp_surrogate_stats st{};
  for (uint64_t b = 0; b < n_big; ++b)
    if (int e = sur_big_step(s, st, mlp_ms)) return e;
  if (stats) *stats = st;
  return FP_OK;
}

}  // namespace

extern "C" {

int fp_sim_set_surrogate(fp_sim* s, const fp_surrogate* n) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (!n) {
    s->sur.on = false;
    return FP_OK;
  }
  if (int e = prepare_surrogate(s, n->cfg)) return e;
  std::vector<float> P(fp::kParams);
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(P.data(), n->p.p, sizeof(float) * fp::kParams, hipMemcpyDeviceToHost));
  const std::vector<uint8_t> img = fph::pack_image(P);
  HIP_TRY(s->sur.image.assign(img.data(), img.size()));
  int dev = 0, cus = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  s->sur.grid = cus > 0 ? cus : 256;
  s->sur.cfg = n->cfg;
  s->sur.skip.assign(s->meshes.size(), 0);
  s->sur.backoff.assign(s->meshes.size(), 0);
  s->sur.on = true;
  return FP_OK;
}

int fp_sim_set_surrogate_shape(fp_sim* s, int32_t vertices_per_wave) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (vertices_per_wave != 32 && vertices_per_wave != 64) return fail(FP_EINVAL, "vertices_per_wave must be 32 or 64");
  s->sur.variant = vertices_per_wave / 16;
  return FP_OK;
}

int fp_sim_step_surrogate(fp_sim* s, uint64_t n_big, fp_surrogate_stats* stats) {
  return run_surrogate(s, n_big, stats, nullptr);
}

int fp_sim_time_surrogate(fp_sim* s, uint64_t n_big, float* total_ms, float* mlp_ms, fp_surrogate_stats* stats) {
  if (!total_ms || !mlp_ms) return fail(FP_EINVAL, "NULL argument");
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  HIP_TRY(hipStreamSynchronize(s->stream));
  *mlp_ms = 0.0f;
  const auto t0 = std::chrono::steady_clock::now();
  if (int e = run_surrogate(s, n_big, stats, mlp_ms)) return e;
  HIP_TRY(hipStreamSynchronize(s->stream));
  *total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return FP_OK;
}

int fp_sim_collect_samples(fp_sim* s, fp_surrogate* n, uint64_t n_big, float noise, uint64_t seed) {
  if (!s || !n) return fail(FP_EINVAL, "NULL argument");
  if (!(noise >= 0.0f)) return fail(FP_EINVAL, "noise must be >= 0");
  if (int e = prepare_surrogate(s, n->cfg)) return e;
  auto& S = s->sur;
  const uint64_t N = s->N, nfree = S.ids.n;
  if (N == 0 || nfree == 0 || n_big == 0) return FP_OK;
  if (int e = fp_surrogate_reserve(n, n->count + n_big * nfree)) return e;
  HIP_TRY(hipStreamSynchronize(n->stream));
  DBuf<float> pos[3][3], nz[6];
  for (int t = 0; t < 3; ++t)
    for (int k = 0; k < 3; ++k) HIP_TRY(pos[t][k].alloc(N));
  for (int k = 0; k < 6; ++k) HIP_TRY(nz[k].alloc(N));
  auto snap = [&](int slot) -> int {
    for (int k = 0; k < 3; ++k)
      HIP_TRY(hipMemcpyAsync(pos[slot][k].p, s->state[s->cur][k].p, N * sizeof(float), hipMemcpyDeviceToDevice,
                             s->stream));
    return FP_OK;
  };
  // time t lives in ring slot (t + 3) % 3; the state before the first big step is taken to be at
  // rest, so time -1 has the same positions as time 0
  if (int e = snap(0)) return e;
  if (int e = snap(2)) return e;
  const fp::FeatParams fpp = feat_params(n->cfg);
  const uint32_t nfb = (uint32_t)s->h_face_blocks.size();
  // Trajectory health: a mesh whose plain state stops being finite, or reaches a speed of 30 m/s or
  // a strain of 100 %, contributes no samples from that big step on (the reference method itself
  // diverges for some pinnings, see the cloth workloads).
  const size_t M = s->meshes.size();
  const uint32_t nblk = (uint32_t)s->h_blocks.size();
  std::vector<uint8_t> sick(M, 0);
  std::vector<float> health(nblk);
  DBuf<float> d_health;
  DBuf<uint32_t> ids_now;
  HIP_TRY(d_health.alloc(nblk + 1));
  uint64_t n_ids = 0;
  auto rebuild_ids = [&]() -> int {
    std::vector<uint32_t> ids;
    for (size_t m = 0; m < M; ++m) {
      if (sick[m]) continue;
      const MeshSlot& ms = s->meshes[m];
      std::vector<uint8_t> pin(ms.host.n, 0);
      for (uint64_t q : ms.host.pins) pin[q] = 1;
      for (uint64_t i = 0; i < ms.host.n; ++i)
        if (!pin[ms.ref(i)]) ids.push_back((uint32_t)(ms.vbase + i));
    }
    n_ids = ids.size();
    HIP_TRY(ids_now.upload(ids.data(), ids.size()));
    return FP_OK;
  };
  if (int e = rebuild_ids()) return e;
  for (uint64_t b = 0; b < n_big; ++b) {
    if (int e = run_steps(s, (uint64_t)n->cfg.ratio)) return e;
    const int ic = (int)(b % 3), in = (int)((b + 1) % 3), ip = (int)((b + 2) % 3);
    if (int e = snap(in)) return e;
    fp::StepArgs st_now = s->args(false);
    HIP_TRY(fp::launch_sur_health(st_now.in, S.slots.p, s->all_blocks.p, nblk, 30.0f, 1.0f, d_health.p, s->stream));
    HIP_TRY(hipMemcpyAsync(health.data(), d_health.p, nblk * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    bool changed = false;
    for (uint32_t k = 0; k < nblk; ++k) {
      const uint32_t m = s->h_blocks[k].mesh;
      if (!sick[m] && !(health[k] <= 1.0f)) { sick[m] = 1; changed = true; }
    }
    if (changed)
      if (int e = rebuild_ids()) return e;
    if (n_ids == 0) break;
    const fp::Pos3 prev{pos[ip][0].p, pos[ip][1].p, pos[ip][2].p}, cur{pos[ic][0].p, pos[ic][1].p, pos[ic][2].p},
        next{pos[in][0].p, pos[in][1].p, pos[in][2].p};
    fp::SoA noisy{nz[0].p, nz[1].p, nz[2].p, nz[3].p, nz[4].p, nz[5].p};
    uint64_t h = seed ^ (0x9e3779b97f4a7c15ull * (b + 1));
    HIP_TRY(fp::launch_sur_noise(prev, cur, noisy, s->pins.p, (uint32_t)N, fpp.H, noise, h, s->stream));
    fp::StepArgs a = s->args(false);
    a.in = noisy;
    a.params = S.big_params.p;
    HIP_TRY(fp::launch_faces(a, s->csr(), s->all_face_blocks.p, nfb, false, s->stream));
    HIP_TRY(fp::launch_sur_samples(a, s->csr(), ids_now.p, S.mesh_of.p, (uint32_t)n_ids, next, S.slots.p, fpp,
                                   n->cfg.out_scale, n->cfg.base_struts != 0, n->X.p + n->count * fp::kFeat,
                                   n->T.p + n->count * 3, s->stream));
    n->count += n_ids;
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (S.on) return upload_big_params(s, S.cfg);
  return FP_OK;
}

int fp_sim_surrogate_features(fp_sim* s, uint64_t mesh, float* features) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (!features) return fail(FP_EINVAL, "features is NULL");
  if (!s->sur.on) return fail(FP_ESTATE, "no surrogate attached (fp_sim_set_surrogate)");
  if (int e = prepare_surrogate(s, s->sur.cfg)) return e;
  const MeshSlot& m = s->meshes[mesh];
  std::vector<uint32_t> ids(m.host.n);
  for (uint64_t i = 0; i < m.host.n; ++i) ids[i] = (uint32_t)m.dev(i);  // rows in reference order
  DBuf<uint32_t> d_ids;
  DBuf<float> out;
  HIP_TRY(d_ids.upload(ids.data(), ids.size()));
  HIP_TRY(out.alloc(m.host.n * fp::kFeat + 1));
  fp::StepArgs a = s->args(false);
  HIP_TRY(fp::launch_sur_features(a.in, s->sur.slots.p, d_ids.p, (uint32_t)m.host.n, feat_params(s->sur.cfg), out.p,
                                  s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(features, out.p, m.host.n * fp::kFeat * sizeof(float), hipMemcpyDeviceToHost));
  return FP_OK;
}

int fp_sim_surrogate_accel(fp_sim* s, uint64_t mesh, float* accel) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (!accel) return fail(FP_EINVAL, "accel is NULL");
  if (!s->sur.on) return fail(FP_ESTATE, "no surrogate attached (fp_sim_set_surrogate)");
  if (int e = prepare_surrogate(s, s->sur.cfg)) return e;
  auto& S = s->sur;
  fp::StepArgs a = s->args(false);
  HIP_TRY(fp::launch_sur_mlp(a.in, S.slots.p, (uint32_t)s->N, S.image.p, feat_params(S.cfg), S.cfg.out_scale, S.da.p,
                             S.grid, S.variant, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const MeshSlot& m = s->meshes[mesh];
  std::vector<float> da(3 * m.host.n);
  HIP_TRY(hipMemcpy(da.data(), S.da.p + 3 * m.vbase, da.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < m.host.n; ++i)
    for (int c = 0; c < 3; ++c) accel[3 * i + c] = da[3 * (m.dev(i) - m.vbase) + c];
  return FP_OK;
}

int fp_sim_surrogate_slots(fp_sim* s, uint64_t mesh, int64_t* nbr, float* rest) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (!nbr || !rest) return fail(FP_EINVAL, "NULL argument");
  if (s->sur.h_slots.empty() && s->N)
    if (int e = build_slots(s)) return e;
  const MeshSlot& m = s->meshes[mesh];
  for (uint64_t i = 0; i < m.host.n; ++i)
    for (int j = 0; j < fp::kSlots; ++j) {
      const fp::SlotEnt e = s->sur.h_slots[m.dev(i) * fp::kSlots + j];
      nbr[i * fp::kSlots + j] = e.nbr == fp::kNoSlot ? -1 : (int64_t)m.ref(e.nbr - m.vbase);
      rest[i * fp::kSlots + j] = e.nbr == fp::kNoSlot ? 0.0f : e.rest;
    }
  return FP_OK;
}

}  // extern "C"
