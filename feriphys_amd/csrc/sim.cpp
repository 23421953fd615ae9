// C ABI of the gfx950 springy step: host handles, device packing, kernel dispatch.
//
// Device layout of a simulation (all meshes packed, global vertex ids = mesh base + local id):
//   state[2][6]   SoA f32 x, y, z, vx, vy, vz, ping-ponged every step (24 B read + 24 B written
//                 per vertex-step: the whole HBM traffic of the grid kernel)
//   force[3]      f32 Fx, Fy, Fz records, touched only by the split API (accumulate_forces /
//                 update_points) or when a step starts from non-zero forces
//   mass          f32 per vertex (read only for meshes whose masses are not uniform)
//   pins          1 bit per vertex
//   incidence     per vertex: incident struts (neighbour, rest length, k, c) in strut order and
//                 one float4 per incident face (that corner's share of the face's drag + lift,
//                 written by the face kernel) in face order, sliced 64 vertices at a time (entry j
//                 of a vertex at first + 64 j); per face 3 vertex ids and the positions of its
//                 corners' entries
//   obstacles     faces (corners, unit normal, dropped axis) + a uniform grid over them
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "grid_fast.hpp"
#include "sim_internal.hpp"


using namespace fpsim;

namespace {

using fph::DBuf;
using fph::fail;
using fph::g_err;

using fph::duration_roundtrip;
using fph::same_bits;

fp_config default_config() {
  fp_config c;
  c.integration = FP_RK4;
  c.dt = 0.001f;  // Duration::from_millis(1).as_secs_f32()
  c.gravity[0] = 0.0f; c.gravity[1] = -10.0f; c.gravity[2] = 0.0f;
  c.wind[0] = c.wind[1] = c.wind[2] = 0.0f;
  c.lift_coefficient = 1.0f;
  c.drag_coefficient = 1.0f;
  c.coefficient_of_restitution = 0.95f;
  c.coefficient_of_friction = 0.3f;
  return c;
}

// Strut adjacency of a mesh (CSR over reference vertex ids).
struct Adjacency {
  std::vector<uint32_t> row, adj;
  uint32_t deg(uint32_t v) const { return row[v + 1] - row[v]; }
};
Adjacency strut_adjacency(const fp::HostMesh& h) {
  Adjacency A;
  A.row.assign(h.n + 1, 0);
  A.adj.resize(2 * h.n_struts());
  for (uint64_t k = 0; k < h.n_struts(); ++k) { A.row[h.s_v0[k] + 1]++; A.row[h.s_v1[k] + 1]++; }
  for (uint64_t i = 0; i < h.n; ++i) A.row[i + 1] += A.row[i];
  std::vector<uint32_t> fill(A.row.begin(), A.row.end() - 1);
  for (uint64_t k = 0; k < h.n_struts(); ++k) {
    A.adj[fill[h.s_v0[k]]++] = h.s_v1[k];
    A.adj[fill[h.s_v1[k]]++] = h.s_v0[k];
  }
  return A;
}

// Reverse Cuthill-McKee order of a strut graph: breadth-first from a lowest-degree vertex,
// neighbours in ascending degree, each connected component in turn, the whole order reversed.
std::vector<uint32_t> rcm_order(const Adjacency& A, uint64_t n) {
  std::vector<uint32_t> by_deg(n), order;
  for (uint64_t i = 0; i < n; ++i) by_deg[i] = (uint32_t)i;
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](uint32_t a, uint32_t b) { return A.deg(a) < A.deg(b); });
  std::vector<uint8_t> seen(n, 0);
  std::vector<uint32_t> nb;
  order.reserve(n);
  for (uint32_t start : by_deg) {
    if (seen[start]) continue;
    seen[start] = 1;
    const size_t head = order.size();
    order.push_back(start);
    for (size_t q = head; q < order.size(); ++q) {
      const uint32_t v = order[q];
      nb.clear();
      for (uint32_t e = A.row[v]; e < A.row[v + 1]; ++e)
        if (!seen[A.adj[e]]) { seen[A.adj[e]] = 1; nb.push_back(A.adj[e]); }
      std::stable_sort(nb.begin(), nb.end(), [&](uint32_t a, uint32_t b) { return A.deg(a) < A.deg(b); });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Device order of a general mesh: patches of at most kBlock vertices grown breadth-first over the
// strut graph from seeds taken in reverse Cuthill-McKee order, each patch's vertices in the order
// they were reached.  A patch is a compact piece of surface, so its halo (the neighbours outside
// it that its struts reach) is small and one workgroup can stage patch and halo in LDS; the
// sweep order keeps successive patches adjacent.  perm[d] = the reference vertex at device slot d,
// bstart = the patches' first slots, ending with n.
void patch_order(const fp::HostMesh& h, std::vector<uint32_t>& perm, std::vector<uint32_t>& bstart) {
  const Adjacency A = strut_adjacency(h);
  const std::vector<uint32_t> seeds = rcm_order(A, h.n);
  std::vector<uint8_t> taken(h.n, 0);
  perm.clear();
  perm.reserve(h.n);
  bstart.clear();
  for (uint32_t seed : seeds) {
    if (taken[seed]) continue;
    const size_t head = perm.size();
    bstart.push_back((uint32_t)head);
    taken[seed] = 1;
    perm.push_back(seed);
    for (size_t q = head; q < perm.size() && perm.size() - head < fp::kBlock; ++q) {
      const uint32_t v = perm[q];
      for (uint32_t e = A.row[v]; e < A.row[v + 1] && perm.size() - head < fp::kBlock; ++e)
        if (!taken[A.adj[e]]) { taken[A.adj[e]] = 1; perm.push_back(A.adj[e]); }
    }
  }
  bstart.push_back((uint32_t)h.n);
}

// Default tile shape for a cloth of this size: 16x16 (fastest measured on the 1024 x 1024 cloth,
// one vertex per thread); 7x31 for short wide cloths.
int default_tile_shape(uint32_t rows, uint32_t cols) {
  if (rows < 15 && cols >= 31) return fp::TILE_7x31;
  return fp::TILE_16x16;
}

}  // namespace


namespace fpsim {

fp::StepParams make_params(const fp_config& c, float mass, bool uniform) {
  fp::StepParams P{};
  P.integration = c.integration;
  P.dt = c.dt;
  P.dt_half = c.dt * 0.5f;
  P.dt6 = c.dt / 6.0f;
  P.dt3 = c.dt / 3.0f;
  float dd = 0.0f;
  duration_roundtrip(c.dt, dd);
  P.dt_d = dd;
  P.gravity = fp::f3{c.gravity[0], c.gravity[1], c.gravity[2]};
  P.wind = fp::f3{c.wind[0], c.wind[1], c.wind[2]};
  P.neg_drag = -1.0f * c.drag_coefficient;
  P.neg_lift = -1.0f * c.lift_coefficient;
  P.restitution = c.coefficient_of_restitution;
  P.friction = c.coefficient_of_friction;
  P.mass = mass;
  P.uniform_mass = uniform ? 1 : 0;
  return P;
}

int check_config(const fp_config* c) {
  if (!c) return fail(FP_EINVAL, "config is NULL");
  if (c->integration != FP_EULER && c->integration != FP_RK4) return fail(FP_EINVAL, "unknown integration");
  float dd;
  if (!duration_roundtrip(c->dt, dd)) return fail(FP_EINVAL, "dt must be finite and non-negative (Duration)");
  return FP_OK;
}

// Re-derives whether the mesh's masses are all equal (the grid kernel and the scalar-mass path
// need that) from its records.
void refresh_mass(MeshSlot& ms, const float* rec, uint64_t n) {
  bool uni = true;
  float m0 = n ? rec[0] : 0.0f;
  for (uint64_t i = 1; i < n && uni; ++i) uni = (rec[10 * i] == m0) && !(std::signbit(rec[10 * i]) != std::signbit(m0));
  ms.uniform_mass = uni;
  ms.params = make_params(ms.cfg, m0, uni);
}

int upload_params(fp_sim* s) {
  std::vector<fp::StepParams> p;
  for (auto& m : s->meshes) p.push_back(m.params);
  if (s->params.p != nullptr && p.size() == s->h_params_last.size() &&
      std::memcmp(p.data(), s->h_params_last.data(), p.size() * sizeof(fp::StepParams)) == 0)
    return FP_OK;  // the device copy already holds these values
  HIP_TRY(s->params.assign(p.data(), p.size()));
  s->h_params_last = p;
  ++s->epoch;
  return FP_OK;
}

// The tiles of one cloth in row-major tile order.
void cloth_tiles(uint32_t mesh, const fp::ClothDesc& d, int shape, std::vector<fp::TileRange>& out) {
  const uint32_t tr = (uint32_t)fp::kTileRows[shape], tc = (uint32_t)fp::kTileCols[shape];
  for (uint32_t r = 0; r < d.rows; r += tr)
    for (uint32_t c = 0; c < d.cols; c += tc) out.push_back(fp::TileRange{mesh, r, c, (uint32_t)shape});
}

// Runs of the tile list per XCD (see xcd_order).
constexpr size_t kTilePieces = 8;

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8), each XCD with its
// own L2.  Launch order is permuted so that each XCD receives `pieces` contiguous runs of the work
// list: the list is cut into 8 * pieces runs, XCD x taking runs x, x + 8, x + 16, ... (its total
// q or q + 1 items for n = 8q + r, the first r XCDs one more).  Within a run, vertically adjacent
// tiles read each other's halo rows from one L2; several runs per XCD spread work that is not
// uniform over the list (the tiles over an obstacle) across all eight.
template <class T>
std::vector<T> xcd_order(const std::vector<T>& t, size_t pieces = 1) {
  const size_t n = t.size(), q = n / 8, r = n % 8;
  std::vector<std::vector<T>> per(8);
  size_t at = 0;
  for (size_t j = 0; j < pieces; ++j)
    for (size_t x = 0; x < 8; ++x) {
      const size_t quota = q + (x < r ? 1 : 0);
      const size_t len = (j + 1) * quota / pieces - j * quota / pieces;
      per[x].insert(per[x].end(), t.begin() + at, t.begin() + at + len);
      at += len;
    }
  std::vector<T> out(n);
  for (size_t b = 0; b < n; ++b) out[b] = per[b % 8][b / 8];  // b / 8 < q + (b % 8 < r)
  return out;
}

int build_block_lists(fp_sim* s) {
  // the lists depend on the kernel selection and, per mesh, on whether its masses are uniform and on
  // its tile shape (topology, blocks and patches are fixed when the simulation is built)
  std::vector<int32_t> sig{s->kernel, s->math};
  for (const auto& m : s->meshes) {
    sig.push_back(m.uniform_mass ? 1 : 0);
    sig.push_back(m.tile_shape);
  }
  if (sig == s->lists_sig) return FP_OK;
  std::vector<fp::BlockRange> g, c, cf, sb, sf;
  std::vector<fp::PatchRange> pl;
  std::vector<fp::TileRange> by_shape[fp::N_TILE_SHAPES], all_tiles;
  for (uint32_t mi = 0; mi < s->meshes.size(); ++mi) {
    MeshSlot& ms = s->meshes[mi];
    const bool grid_ok = ms.host.cloth.valid && ms.uniform_mass;
    ms.kernel = !grid_ok || s->kernel == FP_KERNEL_CSR ? MK_CSR
                : s->kernel == FP_KERNEL_GRID_SIMPLE   ? MK_GRID_SIMPLE
                                                       : MK_GRID_TILED;
    ms.tile_begin = ms.tile_end = (uint32_t)all_tiles.size();
    ms.on_patch = false;
    if (ms.kernel == MK_GRID_TILED) {
      // the fast kernel has one quad per thread: 15x15 and 7x31 tiles (16x16 becomes 15x15)
      ms.launch_shape = s->math == FP_MATH_FAST && ms.tile_shape == fp::TILE_16x16 ? fp::TILE_15x15 : ms.tile_shape;
      std::vector<fp::TileRange> mt;
      cloth_tiles(mi, ms.host.cloth, ms.launch_shape, mt);
      by_shape[ms.launch_shape].insert(by_shape[ms.launch_shape].end(), mt.begin(), mt.end());
      all_tiles.insert(all_tiles.end(), mt.begin(), mt.end());
      ms.tile_end = (uint32_t)all_tiles.size();
      continue;
    }
    // general meshes: the patch kernel where every patch fits it, unless FP_KERNEL_CSR asks for the
    // one-thread-per-vertex CSR kernel (with the face kernel before it)
    const bool on_patch = ms.kernel == MK_CSR && ms.patch_ok && s->kernel != FP_KERNEL_CSR;
    ms.on_patch = on_patch;
    for (uint32_t b = ms.blk_begin; b < ms.blk_end; ++b) {
      if (on_patch) pl.push_back(s->h_patches[b]);
      else (ms.kernel == MK_GRID_SIMPLE ? g : c).push_back(s->h_blocks[b]);
      sb.push_back(s->h_blocks[b]);
    }
    for (uint32_t b = ms.fblk_begin; b < ms.fblk_end; ++b) {
      if (ms.kernel == MK_CSR && !on_patch) cf.push_back(s->h_face_blocks[b]);
      sf.push_back(s->h_face_blocks[b]);
    }
  }
  // the surrogate's generic path (face kernel + per-vertex update) covers every mesh that is not
  // on the tiled kernel
  c = xcd_order(c);
  pl = xcd_order(pl);
  s->n_patch_list = (uint32_t)pl.size();
  s->patch_nf_cap = 0;
  for (const auto& r : pl) s->patch_nf_cap = std::max(s->patch_nf_cap, r.n_faces);
  HIP_TRY(s->patch_list.upload(pl.data(), pl.size()));
  cf = xcd_order(cf);
  sb = xcd_order(sb);
  sf = xcd_order(sf);
  s->n_sur_blocks = (uint32_t)sb.size();
  s->n_sur_face_blocks = (uint32_t)sf.size();
  HIP_TRY(s->sur_blocks.upload(sb.data(), sb.size()));
  HIP_TRY(s->sur_face_blocks.upload(sf.data(), sf.size()));
  s->n_csr_face_blocks = (uint32_t)cf.size();
  HIP_TRY(s->csr_face_blocks.upload(cf.data(), cf.size()));
  s->n_grid_blocks = (uint32_t)g.size();
  s->n_csr_blocks = (uint32_t)c.size();
  HIP_TRY(s->grid_blocks.upload(g.data(), g.size()));
  HIP_TRY(s->csr_blocks.upload(c.data(), c.size()));
  for (int sh = 0; sh < fp::N_TILE_SHAPES; ++sh) {
    std::vector<fp::TileRange> o = xcd_order(by_shape[sh], kTilePieces);
    s->n_tiles[sh] = (uint32_t)o.size();
    HIP_TRY(s->tiles[sh].upload(o.data(), o.size()));
  }
  HIP_TRY(s->mesh_tiles.upload(all_tiles.data(), all_tiles.size()));
  s->lists_sig = sig;
  ++s->epoch;
  return FP_OK;
}

int check_mesh(const fp_mesh* m) {
  if (!m) return fail(FP_EINVAL, "mesh is NULL");
  if (m->m.moved) return fail(FP_ESTATE, "mesh was moved into a simulation");
  return FP_OK;
}

int check_sim_mesh(const fp_sim* s, uint64_t mesh) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (mesh >= s->meshes.size()) return fail(FP_EINVAL, "mesh index out of range");
  return FP_OK;
}

// The plain step (or the forces) of cloths on the tiled kernels: the exact kernel, or for the fast
// math form the fast kernel (which has no 16x16 shape: its tiles are built as 15x15).
hipError_t tiled_launch(fp_sim* s, const fp::StepArgs& a, const fp::TileRange* tiles, uint32_t n, int shape, int mode) {
  if (s->math == FP_MATH_FAST && shape != fp::TILE_16x16)
    return fp::launch_grid_fast(a, s->cloths.p, tiles, n, shape, mode, s->obst.hit_box, s->stream);
  return fp::launch_grid_tiled(a, s->cloths.p, tiles, n, shape, mode, s->math == FP_MATH_EXACT, s->stream);
}

int run_steps(fp_sim* s, uint64_t n) {
  for (uint64_t k = 0; k < n; ++k) {
    fp::StepArgs a = s->args(s->forces_dirty);
    bool exact = s->math == FP_MATH_EXACT;
    for (int sh = 0; sh < fp::N_TILE_SHAPES; ++sh)
      HIP_TRY(tiled_launch(s, a, s->tiles[sh].p, s->n_tiles[sh], sh, fp::MODE_STEP));
    HIP_TRY(fp::launch_grid(a, s->cloths.p, s->grid_blocks.p, s->n_grid_blocks, fp::MODE_STEP, exact, s->stream));
    HIP_TRY(fp::launch_faces(a, s->csr(), s->csr_face_blocks.p, s->n_csr_face_blocks, exact, s->stream));
    if (s->torsion_on)
      HIP_TRY(fp::launch_hinges(a, s->csr(), s->hinges.p, 0, s->n_hinges, exact, s->stream));
    HIP_TRY(fp::launch_csr(a, s->csr(), s->csr_blocks.p, s->n_csr_blocks, fp::MODE_STEP, exact, s->stream));
    HIP_TRY(fp::launch_patches(a, s->csr(), s->patch_list.p, s->n_patch_list, s->patch_nf_cap, exact, s->stream));
    s->cur ^= 1;
    s->forces_dirty = false;
  }
  return FP_OK;
}

// Enqueues n steps, replaying captured `chunk`-step graphs where possible.  A step that starts from
// non-zero force records runs eagerly first (the captured launches assume clear forces).
int run_steps_graph(fp_sim* s, uint64_t n, uint32_t chunk) {
  if (chunk == 0) return run_steps(s, n);
  if (s->forces_dirty && n > 0) {
    if (int e = run_steps(s, 1)) return e;
    --n;
  }
  while (n >= chunk) {
    const int par = s->cur;
    if (!s->graph[par] || s->graph_epoch[par] != s->epoch || s->graph_chunk[par] != chunk) {
      if (s->graph[par]) { HIP_TRY(hipGraphExecDestroy(s->graph[par])); s->graph[par] = nullptr; }
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      const int e = run_steps(s, chunk);
      const hipError_t ee = hipStreamEndCapture(s->stream, &g);
      s->cur = par;  // capture records the launches without running them
      if (e) return e;
      if (ee != hipSuccess) return fail(FP_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
      const hipError_t ie = hipGraphInstantiate(&s->graph[par], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(FP_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      s->graph_epoch[par] = s->epoch;
      s->graph_chunk[par] = chunk;
    }
    HIP_TRY(hipGraphLaunch(s->graph[par], s->stream));
    s->cur = par ^ (int)(chunk & 1u);
    n -= chunk;
  }
  return run_steps(s, n);
}

int size_class(const fp::ClothDesc& d) {
  const uint32_t m = std::max(d.rows, d.cols);
  return m < 16 ? 0 : m < 32 ? 1 : m < 64 ? 2 : 3;
}

// Uploads the incidence lists of the packed meshes (vertex-centric: incident struts in strut
// order, incident faces in face order) in the sliced layout of fp::CsrView, the strut classes of
// the patch kernel, and the patches of the general meshes.
int build_incidence(fp_sim* s) {
  const uint64_t N = s->N, F = s->F;
  std::vector<uint32_t> scnt(N, 0), fcnt(N, 0), sfirst(N), ffirst(N), fv(3 * F);
  for (auto& ms_ : s->meshes) {
    const auto& h = ms_.host;
    for (uint64_t k = 0; k < h.n_struts(); ++k) {
      scnt[ms_.dev(h.s_v0[k])]++;
      scnt[ms_.dev(h.s_v1[k])]++;
    }
    for (uint64_t k = 0; k < h.f_v.size(); ++k) fcnt[ms_.dev(h.f_v[k])]++;
  }
  // slices: kSlice consecutive device ids from each block start (the wavefronts of the block's
  // workgroup); a slice takes kSlice x (its largest count) entries, lane l's entry j at the slice
  // base + kSlice j + l
  auto slices = [&](const std::vector<uint32_t>& cnt, std::vector<uint32_t>& first) -> uint64_t {
    uint64_t total = 0;
    for (auto& ms_ : s->meshes)
      for (size_t b = 0; b + 1 < ms_.bstart.size(); ++b)
        for (uint64_t g0 = ms_.bstart[b]; g0 < ms_.bstart[b + 1]; g0 += fp::kSlice) {
          const uint64_t g1 = std::min<uint64_t>(ms_.bstart[b + 1], g0 + fp::kSlice);
          uint32_t mx = 0;
          for (uint64_t d = g0; d < g1; ++d) mx = std::max(mx, cnt[ms_.vbase + d]);
          for (uint64_t d = g0; d < g1; ++d) first[ms_.vbase + d] = (uint32_t)(total + (d - g0));
          total += (uint64_t)fp::kSlice * mx;
        }
    return total;
  };
  const uint64_t n_sent = slices(scnt, sfirst), n_fent = slices(fcnt, ffirst);
  if (n_sent >= (1ull << 32) || n_fent >= (1ull << 32)) return fail(FP_EINVAL, "simulation too large for 32-bit ids");
  std::vector<fp::StrutEnt> sent(n_sent + 1, fp::StrutEnt{0, 1.0f, 0.0f, 0.0f});
  std::vector<uint32_t> fslot(3 * F), sfill(N, 0), ffill(N, 0);
  std::vector<uint32_t> frank(3 * F);  // per face corner (reference face order): the face's rank among its vertex's faces
  // the patch kernel's strut classes: per mesh the distinct nominal (stiffness, damping) pairs in
  // first-appearance order, each strut's constants checked to be (1.0 / len) * its pair's
  std::vector<uint16_t> scls(n_sent + 1, 0);
  std::vector<float2> cls_tab(s->meshes.size() * fp::kPatchClasses, float2{0.0f, 0.0f});
  std::vector<uint8_t> cls_ok(s->meshes.size(), 1);
  for (size_t mi = 0; mi < s->meshes.size(); ++mi) {
    auto& ms_ = s->meshes[mi];
    const auto& h = ms_.host;
    uint32_t n_cls = 0;
    for (uint64_t k = 0; k < h.n_struts(); ++k) {
      uint32_t a = (uint32_t)ms_.dev(h.s_v0[k]), b = (uint32_t)ms_.dev(h.s_v1[k]);
      const uint32_t ea = sfirst[a] + fp::kSlice * sfill[a]++, eb = sfirst[b] + fp::kSlice * sfill[b]++;
      sent[ea] = fp::StrutEnt{b, h.s_len[k], h.s_k[k], h.s_c[k]};
      sent[eb] = fp::StrutEnt{a, h.s_len[k], h.s_k[k], h.s_c[k]};
      if (!cls_ok[mi]) continue;
      const float kn = h.s_kn[k], cn = h.s_cn[k], inv = 1.0f / h.s_len[k];
      if (!same_bits(inv * kn, h.s_k[k]) || !same_bits(inv * cn, h.s_c[k])) { cls_ok[mi] = 0; continue; }
      float2* tab = cls_tab.data() + mi * fp::kPatchClasses;
      uint32_t c = 0;
      while (c < n_cls && !(same_bits(tab[c].x, kn) && same_bits(tab[c].y, cn))) ++c;
      if (c == n_cls) {
        if (n_cls == fp::kPatchClasses) { cls_ok[mi] = 0; continue; }
        tab[n_cls++] = float2{kn, cn};
      }
      scls[ea] = scls[eb] = (uint16_t)c;
    }
    // faces are stored (for the face kernel) in ascending order of their lowest device vertex id,
    // so that neighbouring threads gather neighbouring vertices and write neighbouring entries;
    // each vertex's entries list its faces in face-index order
    std::vector<uint32_t> slot3(3 * h.n_faces()), lo(h.n_faces()), at(h.n_faces());
    for (uint64_t t = 0; t < h.n_faces(); ++t) {
      uint32_t m = 0xffffffffu;
      for (uint32_t j = 0; j < 3; ++j) {
        uint32_t v = (uint32_t)ms_.dev(h.f_v[3 * t + j]);
        frank[ms_.fbase * 3 + 3 * t + j] = ffill[v];
        slot3[3 * t + j] = ffirst[v] + fp::kSlice * ffill[v]++;
        m = std::min(m, v);
      }
      lo[t] = m;
      at[t] = (uint32_t)t;
    }
    std::stable_sort(at.begin(), at.end(), [&](uint32_t a, uint32_t b) { return lo[a] < lo[b]; });
    for (uint64_t q = 0; q < h.n_faces(); ++q)
      for (uint32_t j = 0; j < 3; ++j) {
        fv[3 * (ms_.fbase + q) + j] = (uint32_t)ms_.dev(h.f_v[3 * (uint64_t)at[q] + j]);
        fslot[3 * (ms_.fbase + q) + j] = slot3[3 * (uint64_t)at[q] + j];
      }
  }
  // patches of the general meshes (block b of every mesh is h_patches[b]): halo, faces, LDS
  // indices, each patch vertex's corners.  A mesh with a patch whose vertices and halo exceed
  // kPatchLocal, or whose faces exceed kPatchFaceCap, stays on the CSR kernel.
  std::vector<uint16_t> sloc(n_sent + 1, 0);
  std::vector<uint32_t> halo;
  std::vector<fp::PatchFace> pfaces;
  std::vector<fp::PatchCorner> pcorner(n_fent + 1, fp::PatchCorner{0, 0, 0, 0});
  std::vector<int32_t> mark(N, -1);
  for (uint32_t mi = 0; mi < s->meshes.size(); ++mi) {
    auto& ms_ = s->meshes[mi];
    const auto& h = ms_.host;
    ms_.patch_ok = !ms_.perm.empty() && cls_ok[mi];
    std::vector<uint32_t> vf_row, vf;  // reference faces of each device-local vertex
    if (ms_.patch_ok) {
      vf_row.assign(h.n + 1, 0);
      for (uint64_t k = 0; k < h.f_v.size(); ++k) vf_row[ms_.pos[h.f_v[k]] + 1]++;
      for (uint64_t d = 0; d < h.n; ++d) vf_row[d + 1] += vf_row[d];
      vf.resize(h.f_v.size());
      std::vector<uint32_t> fill(vf_row.begin(), vf_row.end() - 1);
      for (uint64_t k = 0; k < h.f_v.size(); ++k) vf[fill[ms_.pos[h.f_v[k]]]++] = (uint32_t)(k / 3);
    }
    std::vector<uint8_t> fseen(ms_.patch_ok ? h.n_faces() : 0, 0);
    for (size_t b = 0; b + 1 < ms_.bstart.size(); ++b) {
      const uint32_t b0 = (uint32_t)ms_.vbase + ms_.bstart[b], b1 = (uint32_t)ms_.vbase + ms_.bstart[b + 1];
      fp::PatchRange pr{mi, b0, b1, (uint32_t)halo.size(), 0u, (uint32_t)pfaces.size(), 0u, 0u};
      if (ms_.patch_ok) {
        const size_t h0 = halo.size(), q0 = pfaces.size();
        auto local = [&](uint32_t g) -> uint32_t {  // LDS index of device vertex g (adds it to the halo)
          if (g >= b0 && g < b1) return g - b0;
          if (mark[g] < 0) { mark[g] = (int32_t)(halo.size() - h0); halo.push_back(g); }
          return (b1 - b0) + (uint32_t)mark[g];
        };
        for (uint32_t i = b0; i < b1; ++i)
          for (uint32_t j = 0; j < scnt[i]; ++j) {
            const uint32_t e = sfirst[i] + fp::kSlice * j;
            sloc[e] = (uint16_t)std::min<uint32_t>(local(sent[e].nbr), fp::kPatchLocal - 1);
          }
        std::vector<uint32_t> fl;
        for (uint32_t i = b0; i < b1; ++i)
          for (uint32_t k = vf_row[i - ms_.vbase]; k < vf_row[i - ms_.vbase + 1]; ++k)
            if (!fseen[vf[k]]) { fseen[vf[k]] = 1; fl.push_back(vf[k]); }
        for (uint32_t qi = 0; qi < (uint32_t)fl.size(); ++qi) {
          const uint32_t f = fl[qi];
          fp::PatchFace pf{};
          uint32_t g3[3];
          for (int c = 0; c < 3; ++c) {
            g3[c] = (uint32_t)ms_.dev(h.f_v[3 * (uint64_t)f + c]);
            pf.v[c] = (uint16_t)std::min<uint32_t>(local(g3[c]), fp::kPatchLocal - 1);
          }
          // corner c of a patch vertex: Face::vertex_angle's operands are the next and the
          // previous corner (springy_mesh.rs:163-172), in the entry of the face's rank among
          // that vertex's faces
          for (int c = 0; c < 3; ++c)
            if (g3[c] >= b0 && g3[c] < b1)
              pcorner[ffirst[g3[c]] + fp::kSlice * frank[3 * (ms_.fbase + f) + c]] =
                  fp::PatchCorner{(uint16_t)std::min<uint32_t>(qi, fp::kPatchFaceCap - 1), pf.v[(c + 1) % 3],
                                  pf.v[(c + 2) % 3], 0};
          pfaces.push_back(pf);
        }
        for (uint32_t f : fl) fseen[f] = 0;
        pr.n_halo = (uint32_t)(halo.size() - h0);
        pr.n_faces = (uint32_t)(pfaces.size() - q0);
        if ((b1 - b0) + pr.n_halo > fp::kPatchLocal || pr.n_faces > fp::kPatchFaceCap) ms_.patch_ok = false;
        for (size_t k = h0; k < halo.size(); ++k) mark[halo[k]] = -1;
      }
      s->h_patches.push_back(pr);
    }
  }
  pfaces.push_back(fp::PatchFace{});
  HIP_TRY(s->patch_faces.upload(pfaces.data(), pfaces.size()));
  HIP_TRY(s->p_corner.upload(pcorner.data(), pcorner.size()));
  {
    std::vector<fp::PatchStrut> pent(n_sent + 1);
    for (uint64_t e = 0; e <= n_sent; ++e) pent[e] = fp::PatchStrut{sent[e].len, sloc[e], scls[e]};
    HIP_TRY(s->p_ent.upload(pent.data(), pent.size()));
    HIP_TRY(s->p_cls.upload(cls_tab.data(), cls_tab.size()));
  }
  halo.push_back(0);
  HIP_TRY(s->halo.upload(halo.data(), halo.size()));
  HIP_TRY(s->s_first.upload(sfirst.data(), sfirst.size()));
  HIP_TRY(s->s_cnt.upload(scnt.data(), scnt.size()));
  HIP_TRY(s->s_ent.upload(sent.data(), sent.size()));
  HIP_TRY(s->face_c.alloc(n_fent + 1));
  HIP_TRY(s->f_first.upload(ffirst.data(), ffirst.size()));
  HIP_TRY(s->f_cnt.upload(fcnt.data(), fcnt.size()));
  HIP_TRY(s->f_slot.upload(fslot.data(), fslot.size()));
  HIP_TRY(s->f_v.upload(fv.data(), fv.size()));
  return FP_OK;
}

// Uploads the torsional hinges (springy_mesh.rs:348-393 builds them; :555-567 names the four
// vertices) and each vertex's hinge entries.
int build_hinges(fp_sim* s) {
  const uint64_t N = s->N;
  std::vector<fp::HingeEnt> hs;
  for (auto& ms_ : s->meshes) {
    const auto& h = ms_.host;
    ms_.hinge_begin = (uint32_t)hs.size();
    for (uint64_t k = 0; k < h.n_struts(); ++k) {
      if (!h.t_has[k] || h.s_f0[k] < 0 || h.s_f1[k] < 0) continue;
      const uint32_t a = h.s_v0[k], b = h.s_v1[k];
      auto other = [&](int64_t f) -> uint32_t {  // utils::tuple_difference(face vertices, strut)
        for (int j = 0; j < 3; ++j) {
          const uint32_t v = (uint32_t)h.f_v[3 * (size_t)f + j];
          if (v != a && v != b) return v;
        }
        return a;
      };
      fp::HingeEnt e{};
      e.v[0] = (uint32_t)ms_.dev(a);
      e.v[1] = (uint32_t)ms_.dev(b);
      e.v[2] = (uint32_t)ms_.dev(other(h.s_f0[k]));
      e.v[3] = (uint32_t)ms_.dev(other(h.s_f1[k]));
      e.k = h.t_k[k];
      e.c = h.t_c[k];
      e.rest = h.t_rest[k];
      hs.push_back(e);
    }
    ms_.hinge_end = (uint32_t)hs.size();
  }
  s->n_hinges = (uint32_t)hs.size();
  if (!hs.empty()) {
    std::vector<uint32_t> trow(N + 1, 0);
    for (const auto& e : hs)
      for (int r = 0; r < 4; ++r) trow[e.v[r] + 1]++;
    for (uint64_t i = 0; i < N; ++i) trow[i + 1] += trow[i];
    std::vector<uint32_t> tent(trow[N]), tfill(trow.begin(), trow.end() - 1);
    for (uint32_t k = 0; k < hs.size(); ++k)
      for (uint32_t r = 0; r < 4; ++r) tent[tfill[hs[k].v[r]]++] = (k << 2) | r;
    HIP_TRY(s->hinges.upload(hs.data(), hs.size()));
    HIP_TRY(s->t_row.upload(trow.data(), trow.size()));
    HIP_TRY(s->t_ent.upload(tent.data(), tent.size()));
    HIP_TRY(s->hinge_f.alloc(4 * hs.size()));
  }
  return FP_OK;
}

// Uploads the cloth descriptors of the grid kernels and the vertex and face block tables.
int build_cloth_views(fp_sim* s) {
  std::vector<fp::ClothView> cv;
  for (uint32_t mi = 0; mi < s->meshes.size(); ++mi) {
    MeshSlot& ms_ = s->meshes[mi];
    const fp::ClothDesc& d = ms_.host.cloth;
    fp::ClothView c{};
    c.base = (uint32_t)ms_.vbase;
    c.rows = d.rows;
    c.cols = d.cols;
    c.spacing = d.spacing;
    c.cx = (float)d.cols / 2.0f * d.spacing;
    c.cy = (float)d.rows / 2.0f * d.spacing;
    c.p0 = d.position[0];
    c.p1 = d.plane == FP_PLANE_XZ ? d.position[2] : d.position[1];
    c.tk = d.tk; c.tc = d.tc; c.sk = d.sk; c.sc = d.sc; c.bk = d.bk; c.bc = d.bc;
    cv.push_back(c);
    ms_.blk_begin = (uint32_t)s->h_blocks.size();
    for (size_t b = 0; b + 1 < ms_.bstart.size(); ++b)
      s->h_blocks.push_back(fp::BlockRange{mi, (uint32_t)(ms_.vbase + ms_.bstart[b]),
                                           (uint32_t)(ms_.vbase + ms_.bstart[b + 1]), 0u});
    ms_.blk_end = (uint32_t)s->h_blocks.size();
    ms_.fblk_begin = (uint32_t)s->h_face_blocks.size();
    const uint64_t nf = ms_.host.n_faces();
    for (uint64_t b = 0; b < nf; b += fp::kBlock)
      s->h_face_blocks.push_back(fp::BlockRange{mi, (uint32_t)(ms_.fbase + b),
                                                (uint32_t)(ms_.fbase + std::min<uint64_t>(nf, b + fp::kBlock)), 0u});
    ms_.fblk_end = (uint32_t)s->h_face_blocks.size();
  }
  HIP_TRY(s->cloths.upload(cv.data(), cv.size()));
  HIP_TRY(s->all_blocks.upload(s->h_blocks.data(), s->h_blocks.size()));
  HIP_TRY(s->all_face_blocks.upload(s->h_face_blocks.data(), s->h_face_blocks.size()));
  return FP_OK;
}

}  // namespace fpsim

extern "C" {

const char* fp_last_error(void) { return g_err.c_str(); }
const char* fp_version(void) { return "feriphys_gpu 0.1 gfx950"; }

int fp_selftest_exact_math(uint64_t mismatches[5]) {
  if (!mismatches) return fail(FP_EINVAL, "mismatches is NULL");
  DBuf<unsigned long long> d;
  HIP_TRY(d.alloc(5));
  HIP_TRY(hipMemset(d.p, 0, 5 * sizeof(unsigned long long)));
  HIP_TRY(fp::launch_selftest(d.p, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  unsigned long long h[5];
  HIP_TRY(hipMemcpy(h, d.p, sizeof(h), hipMemcpyDeviceToHost));
  for (int q = 0; q < 5; ++q) mismatches[q] = h[q];
  return FP_OK;
}
void fp_config_default(fp_config* out) {
  if (out) *out = default_config();
}

int fp_mesh_new(const float* positions, uint64_t n_vertices, const uint64_t* indices, uint64_t n_indices, float mass,
                float k, float c, const fp_torsion_config* torsion, const fp_strut_override* overrides,
                uint64_t n_overrides, fp_mesh** out) {
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  std::unique_ptr<fp_mesh> m(new (std::nothrow) fp_mesh());
  if (!m) return fail(FP_ENOMEM, "out of host memory");
  try {
    fp::Status st = fp::build_springy_mesh(positions, n_vertices, indices, n_indices, mass, k, c, torsion, overrides,
                                           n_overrides, m->m);
    if (st.code != FP_OK) return fail(st);
  } catch (const std::bad_alloc&) {
    return fail(FP_ENOMEM, "out of host memory");
  }
  *out = m.release();
  return FP_OK;
}

int fp_cloth_new_in_plane(uint64_t rows, uint64_t cols, float spacing, const float position[3], float point_mass,
                          float tk, float tc, float sk, float sc, float bk, float bc, const uint64_t* pins,
                          uint64_t n_pins, int32_t plane, fp_mesh** out) {
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  std::unique_ptr<fp_mesh> m(new (std::nothrow) fp_mesh());
  if (!m) return fail(FP_ENOMEM, "out of host memory");
  try {
    fp::Status st = fp::build_cloth(rows, cols, spacing, position, point_mass, tk, tc, sk, sc, bk, bc, pins, n_pins,
                                    plane, m->m);
    if (st.code != FP_OK) return fail(st);
  } catch (const std::bad_alloc&) {
    return fail(FP_ENOMEM, "out of host memory");
  }
  *out = m.release();
  return FP_OK;
}

int fp_cloth_new(uint64_t rows, uint64_t cols, float spacing, const float position[3], float point_mass, float tk,
                 float tc, float sk, float sc, float bk, float bc, const uint64_t* pins, uint64_t n_pins,
                 fp_mesh** out) {
  return fp_cloth_new_in_plane(rows, cols, spacing, position, point_mass, tk, tc, sk, sc, bk, bc, pins, n_pins,
                               FP_PLANE_XY, out);
}

int fp_mesh_add_strut(fp_mesh* mesh, uint64_t v0, uint64_t v1, float k, float c) {
  if (int e = check_mesh(mesh)) return e;
  fp::Status st = fp::mesh_add_strut(mesh->m, v0, v1, k, c);
  return st.code == FP_OK ? FP_OK : fail(st);
}

int fp_mesh_add_struts(fp_mesh* mesh, const uint64_t* v01, uint64_t n, float k, float c) {
  if (int e = check_mesh(mesh)) return e;
  if (n && !v01) return fail(FP_EINVAL, "strut pairs are NULL");
  for (uint64_t s = 0; s < 2 * n; ++s)  // all or nothing: check every index before appending
    if (v01[s] >= mesh->m.n) return fail(FP_EINVAL, "add_strut: vertex index out of range");
  for (uint64_t s = 0; s < n; ++s) {
    fp::Status st = fp::mesh_add_strut(mesh->m, v01[2 * s], v01[2 * s + 1], k, c);
    if (st.code != FP_OK) return fail(st);
  }
  return FP_OK;
}

int fp_mesh_add_pin(fp_mesh* mesh, uint64_t index) {
  if (int e = check_mesh(mesh)) return e;
  mesh->m.pins.push_back(index);
  return FP_OK;
}

int fp_mesh_counts(const fp_mesh* mesh, uint64_t* np, uint64_t* ns, uint64_t* nf, uint64_t* npin) {
  if (int e = check_mesh(mesh)) return e;
  if (np) *np = mesh->m.n;
  if (ns) *ns = mesh->m.n_struts();
  if (nf) *nf = mesh->m.n_faces();
  if (npin) *npin = mesh->m.pins.size();
  return FP_OK;
}

static void copy_struts(const fp::HostMesh& m, uint64_t* v01, float* kcl, int64_t* faces) {
  for (uint64_t s = 0; s < m.n_struts(); ++s) {
    if (v01) { v01[2 * s] = m.s_v0[s]; v01[2 * s + 1] = m.s_v1[s]; }
    if (kcl) { kcl[3 * s] = m.s_k[s]; kcl[3 * s + 1] = m.s_c[s]; kcl[3 * s + 2] = m.s_len[s]; }
    if (faces) { faces[2 * s] = m.s_f0[s]; faces[2 * s + 1] = m.s_f1[s]; }
  }
}

int fp_mesh_get_struts(const fp_mesh* mesh, uint64_t* v01, float* kcl, int64_t* faces) {
  if (int e = check_mesh(mesh)) return e;
  copy_struts(mesh->m, v01, kcl, faces);
  return FP_OK;
}

int fp_mesh_get_torsion(const fp_mesh* mesh, uint8_t* has, float* kcr) {
  if (int e = check_mesh(mesh)) return e;
  const fp::HostMesh& m = mesh->m;
  for (uint64_t s = 0; s < m.n_struts(); ++s) {
    if (has) has[s] = m.t_has[s];
    if (kcr) { kcr[3 * s] = m.t_k[s]; kcr[3 * s + 1] = m.t_c[s]; kcr[3 * s + 2] = m.t_rest[s]; }
  }
  return FP_OK;
}

int fp_mesh_get_faces(const fp_mesh* mesh, uint64_t* s3, uint64_t* v3) {
  if (int e = check_mesh(mesh)) return e;
  const fp::HostMesh& m = mesh->m;
  for (uint64_t k = 0; k < m.f_v.size(); ++k) {
    if (s3) s3[k] = m.f_s[k];
    if (v3) v3[k] = m.f_v[k];
  }
  return FP_OK;
}

int fp_mesh_get_pins(const fp_mesh* mesh, uint64_t* pins) {
  if (int e = check_mesh(mesh)) return e;
  if (pins) std::copy(mesh->m.pins.begin(), mesh->m.pins.end(), pins);
  return FP_OK;
}

int fp_mesh_get_points(const fp_mesh* mesh, float* rec) {
  if (int e = check_mesh(mesh)) return e;
  if (!rec) return fail(FP_EINVAL, "records is NULL");
  std::copy(mesh->m.rec.begin(), mesh->m.rec.end(), rec);
  return FP_OK;
}

int fp_mesh_set_points(fp_mesh* mesh, const float* rec, uint64_t n) {
  if (int e = check_mesh(mesh)) return e;
  if (n != mesh->m.n) return fail(FP_EINVAL, "State Vector incorrect size! (record count != point count)");
  if (n && !rec) return fail(FP_EINVAL, "records is NULL");
  std::copy(rec, rec + 10 * n, mesh->m.rec.begin());
  return FP_OK;
}

static void copy_vertices(const fp::HostMesh& m, const float* rec, float* pos, uint64_t* idx) {
  if (pos)
    for (uint64_t i = 0; i < m.n; ++i)
      for (int k = 0; k < 3; ++k) pos[3 * i + k] = rec[10 * i + 1 + k];
  if (idx)
    for (uint64_t k = 0; k < m.f_v.size(); ++k) idx[k] = m.f_v[k];
}

int fp_mesh_get_vertices(const fp_mesh* mesh, float* pos, uint64_t* idx) {
  if (int e = check_mesh(mesh)) return e;
  copy_vertices(mesh->m, mesh->m.rec.data(), pos, idx);
  return FP_OK;
}

void fp_mesh_free(fp_mesh* mesh) { delete mesh; }

int fp_obstacle_new(const float* positions, uint64_t n, const uint64_t* indices, uint64_t n_idx, fp_obstacle** out) {
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  std::unique_ptr<fp_obstacle> o(new (std::nothrow) fp_obstacle());
  if (!o) return fail(FP_ENOMEM, "out of host memory");
  fp::Status st = fp::build_obstacle(positions, n, indices, n_idx, o->o);
  if (st.code != FP_OK) return fail(st);
  *out = o.release();
  return FP_OK;
}

int fp_obstacle_counts(const fp_obstacle* o, uint64_t* nv, uint64_t* ne, uint64_t* nf) {
  if (!o) return fail(FP_EINVAL, "obstacle is NULL");
  if (o->o.moved) return fail(FP_ESTATE, "obstacle was moved into a simulation");
  if (nv) *nv = o->o.vertices.size() / 3;
  if (ne) *ne = o->o.n_edges;
  if (nf) *nf = o->o.n_faces();
  return FP_OK;
}

void fp_obstacle_free(fp_obstacle* o) { delete o; }

int fp_sim_new(fp_mesh* const* meshes, uint64_t n_meshes, fp_obstacle* const* obstacles, uint64_t n_obstacles,
               fp_sim** out) {
  // Ownership moves in, as Simulation::new takes the Vecs by value: free the handles on every path.
  std::vector<std::unique_ptr<fp_mesh>> ms;
  std::vector<std::unique_ptr<fp_obstacle>> os;
  for (uint64_t i = 0; i < n_meshes; ++i)
    if (meshes && meshes[i]) ms.emplace_back(meshes[i]);
  for (uint64_t i = 0; i < n_obstacles; ++i)
    if (obstacles && obstacles[i]) os.emplace_back(obstacles[i]);
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  if (ms.size() != n_meshes || os.size() != n_obstacles) return fail(FP_EINVAL, "NULL mesh or obstacle handle");
  for (auto& m : ms)
    if (m->m.moved) return fail(FP_ESTATE, "mesh was moved into a simulation");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return fail(FP_ENODEV, "no HIP device");

  std::unique_ptr<fp_sim> s(new (std::nothrow) fp_sim());
  if (!s) return fail(FP_ENOMEM, "out of host memory");
  HIP_TRY(hipGetDevice(&s->device));
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&s->ev0));
  HIP_TRY(hipEventCreate(&s->ev1));
  const fp_config cfg0 = default_config();
  uint64_t N = 0, F = 0, Sdir = 0, Fent = 0;
  for (auto& m : ms) {
    for (uint64_t p : m->m.pins)
      if (p >= m->m.n) return fail(FP_EINVAL, "pin index out of range");  // springy_mesh.rs:523 panics
    MeshSlot slot;
    slot.host = std::move(m->m);
    slot.vbase = N;
    slot.fbase = F;
    slot.cfg = cfg0;
    N += slot.host.n;
    F += slot.host.n_faces();
    Sdir += 2 * slot.host.n_struts();
    Fent += 3 * slot.host.n_faces();
    refresh_mass(slot, slot.host.rec.data(), slot.host.n);
    if (slot.host.cloth.valid) slot.tile_shape = default_tile_shape(slot.host.cloth.rows, slot.host.cloth.cols);
    if (!slot.host.cloth.valid && slot.host.n > 1) {
      patch_order(slot.host, slot.perm, slot.bstart);
      slot.pos.assign(slot.host.n, 0);
      for (uint64_t d = 0; d < slot.host.n; ++d) slot.pos[slot.perm[d]] = (uint32_t)d;
    } else {
      for (uint64_t b = 0; b < slot.host.n; b += fp::kBlock) slot.bstart.push_back((uint32_t)b);
      slot.bstart.push_back((uint32_t)slot.host.n);
    }
    s->meshes.push_back(std::move(slot));
  }
  if (N >= (1ull << 32) || F >= (1ull << 30) || Sdir >= (1ull << 32) || Fent >= (1ull << 32))
    return fail(FP_EINVAL, "simulation too large for 32-bit ids");
  s->N = N;
  s->F = F;

  // state, masses, forces, pins
  {
    std::vector<float> comp(N), m(N);
    for (int k = 0; k < 6; ++k) {
      for (auto& ms_ : s->meshes)
        for (uint64_t i = 0; i < ms_.host.n; ++i) comp[ms_.dev(i)] = ms_.host.rec[10 * i + 1 + k];
      HIP_TRY(s->state[0][k].upload(comp.data(), N));
      HIP_TRY(s->state[1][k].alloc(N));
    }
    for (auto& ms_ : s->meshes)
      for (uint64_t i = 0; i < ms_.host.n; ++i) m[ms_.dev(i)] = ms_.host.rec[10 * i];
    HIP_TRY(s->mass.upload(m.data(), N));
    std::vector<uint32_t> pos(N, 0);
    for (auto& ms_ : s->meshes)
      for (uint64_t i = 0; i < ms_.pos.size(); ++i) pos[ms_.vbase + i] = ms_.pos[i];
    HIP_TRY(s->d_pos.upload(pos.data(), N));
    for (int k = 0; k < 3; ++k) {
      for (auto& ms_ : s->meshes)
        for (uint64_t i = 0; i < ms_.host.n; ++i) comp[ms_.dev(i)] = ms_.host.rec[10 * i + 7 + k];
      HIP_TRY(s->force[k].upload(comp.data(), N));
    }
    bool nonzero_f = false;
    for (auto& ms_ : s->meshes)
      for (uint64_t i = 0; i < ms_.host.n && !nonzero_f; ++i)
        nonzero_f = ms_.host.rec[10 * i + 7] != 0.0f || ms_.host.rec[10 * i + 8] != 0.0f || ms_.host.rec[10 * i + 9] != 0.0f ||
                    std::signbit(ms_.host.rec[10 * i + 7]) || std::signbit(ms_.host.rec[10 * i + 8]) ||
                    std::signbit(ms_.host.rec[10 * i + 9]);
    s->forces_dirty = nonzero_f;
    std::vector<uint32_t> bits((N + 31) / 32 + 1, 0u);
    for (auto& ms_ : s->meshes)
      for (uint64_t p : ms_.host.pins) {
        uint64_t g = ms_.dev(p);
        bits[g >> 5] |= 1u << (g & 31);
      }
    HIP_TRY(s->pins.upload(bits.data(), bits.size()));
  }

  if (int e = build_incidence(s.get())) return e;
  if (int e = build_hinges(s.get())) return e;
  if (int e = build_cloth_views(s.get())) return e;
  std::vector<fp::HostObstacle> obs;
  for (auto& o : os) obs.push_back(std::move(o->o));
  if (int e = fph::upload_obstacles(s->obst, obs)) return e;
  if (int e = upload_params(s.get())) return e;
  if (int e = build_block_lists(s.get())) return e;
  HIP_TRY(hipDeviceSynchronize());
  *out = s.release();
  return FP_OK;
}

int fp_sim_set_config(fp_sim* s, const fp_config* c) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (int e = check_config(c)) return e;
  for (auto& m : s->meshes) {
    m.cfg = *c;
    m.params = make_params(m.cfg, m.params.mass, m.uniform_mass);
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return upload_params(s);
}

int fp_sim_get_config(const fp_sim* s, fp_config* c) {
  if (!s || !c) return fail(FP_EINVAL, "NULL argument");
  *c = s->meshes.empty() ? default_config() : s->meshes[0].cfg;
  return FP_OK;
}

int fp_sim_set_mesh_config(fp_sim* s, uint64_t mesh, const fp_config* c) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (int e = check_config(c)) return e;
  MeshSlot& m = s->meshes[mesh];
  m.cfg = *c;
  m.params = make_params(m.cfg, m.params.mass, m.uniform_mass);
  HIP_TRY(hipStreamSynchronize(s->stream));
  return upload_params(s);
}

int fp_sim_set_mesh_configs(fp_sim* s, const fp_config* c, uint64_t n) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (n != s->meshes.size()) return fail(FP_EINVAL, "one config per mesh expected");
  for (uint64_t i = 0; i < n; ++i)
    if (int e = check_config(&c[i])) return e;
  for (uint64_t i = 0; i < n; ++i) {
    MeshSlot& m = s->meshes[i];
    m.cfg = c[i];
    m.params = make_params(m.cfg, m.params.mass, m.uniform_mass);
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return upload_params(s);
}

int fp_sim_step_graph(fp_sim* s, uint64_t n, uint32_t chunk) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (int e = run_steps_graph(s, n, chunk)) return e;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_sim_time_steps_graph(fp_sim* s, uint64_t n, uint32_t chunk, float* ms) {
  if (!s || !ms) return fail(FP_EINVAL, "NULL argument");
  HIP_TRY(hipEventRecord(s->ev0, s->stream));
  if (int e = run_steps_graph(s, n, chunk)) return e;
  HIP_TRY(hipEventRecord(s->ev1, s->stream));
  HIP_TRY(hipEventSynchronize(s->ev1));
  HIP_TRY(hipEventElapsedTime(ms, s->ev0, s->ev1));
  return FP_OK;
}

int fp_sim_autotune(fp_sim* s, uint32_t steps, int32_t* chosen_out) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (steps == 0) steps = 1;
  HIP_TRY(hipStreamSynchronize(s->stream));
  // snapshot the state (and force records) so every trial starts from it and it is restored
  DBuf<float> snap[9];
  const int c0 = s->cur;
  for (int k = 0; k < 9; ++k) {
    const float* src = k < 6 ? s->state[c0][k].p : s->force[k - 6].p;
    HIP_TRY(snap[k].alloc(s->N));
    if (s->N) HIP_TRY(hipMemcpy(snap[k].p, src, s->N * sizeof(float), hipMemcpyDeviceToDevice));
  }
  const bool dirty0 = s->forces_dirty;
  auto restore = [&]() -> int {
    s->cur = c0;
    for (int k = 0; k < 9; ++k) {
      float* dst = k < 6 ? s->state[c0][k].p : s->force[k - 6].p;
      if (s->N) HIP_TRY(hipMemcpy(dst, snap[k].p, s->N * sizeof(float), hipMemcpyDeviceToDevice));
    }
    s->forces_dirty = dirty0;
    return FP_OK;
  };
  int32_t chosen[4] = {-1, -1, -1, -1};
  for (int cls = 0; cls < 4; ++cls) {
    std::vector<MeshSlot*> members;
    for (auto& m : s->meshes)
      if (m.kernel == MK_GRID_TILED && size_class(m.host.cloth) == cls) members.push_back(&m);
    if (members.empty()) continue;
    float best_ms = INFINITY;
    int best = members[0]->tile_shape;
    for (int shape = 0; shape < fp::N_TILE_SHAPES; ++shape) {
      for (auto* m : members) m->tile_shape = shape;
      if (int e = build_block_lists(s)) return e;
      if (int e = restore()) return e;
      if (int e = run_steps(s, 2)) return e;  // warm-up
      float ms = 0.0f;
      HIP_TRY(hipEventRecord(s->ev0, s->stream));
      if (int e = run_steps(s, steps)) return e;
      HIP_TRY(hipEventRecord(s->ev1, s->stream));
      HIP_TRY(hipEventSynchronize(s->ev1));
      HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
      if (ms < best_ms) { best_ms = ms; best = shape; }
    }
    for (auto* m : members) m->tile_shape = best;
    chosen[cls] = best;
  }
  if (int e = build_block_lists(s)) return e;
  if (int e = restore()) return e;
  HIP_TRY(hipDeviceSynchronize());
  if (chosen_out)
    for (int k = 0; k < 4; ++k) chosen_out[k] = chosen[k];
  return FP_OK;
}

int fp_sim_step_async(fp_sim* s, uint64_t n) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  return run_steps(s, n);
}

int fp_sim_synchronize(fp_sim* s) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_sim_get_timestep(const fp_sim* s, float* out) {
  if (!s || !out) return fail(FP_EINVAL, "NULL argument");
  fp_config c = s->meshes.empty() ? default_config() : s->meshes[0].cfg;
  if (!duration_roundtrip(c.dt, *out)) return fail(FP_EINVAL, "bad dt");
  return FP_OK;
}

int fp_sim_step(fp_sim* s, uint64_t n, float* timestep_out) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (int e = run_steps(s, n)) return e;
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (timestep_out) return fp_sim_get_timestep(s, timestep_out);
  return FP_OK;
}

int fp_sim_mesh_count(const fp_sim* s, uint64_t* out) {
  if (!s || !out) return fail(FP_EINVAL, "NULL argument");
  *out = s->meshes.size();
  return FP_OK;
}

int fp_sim_mesh_counts(const fp_sim* s, uint64_t mesh, uint64_t* np, uint64_t* ns, uint64_t* nf, uint64_t* npin) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  const fp::HostMesh& h = s->meshes[mesh].host;
  if (np) *np = h.n;
  if (ns) *ns = h.n_struts();
  if (nf) *nf = h.n_faces();
  if (npin) *npin = h.pins.size();
  return FP_OK;
}

int fp_sim_count_nonfinite(fp_sim* s, uint64_t* counts) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (!counts && !s->meshes.empty()) return fail(FP_EINVAL, "counts is NULL");
  const size_t M = s->meshes.size();
  DBuf<unsigned long long> d;
  HIP_TRY(d.alloc(M + 1));
  HIP_TRY(hipMemsetAsync(d.p, 0, (M + 1) * sizeof(unsigned long long), s->stream));
  fp::StepArgs a = s->args(false);
  HIP_TRY(fp::launch_nonfinite(a.in, s->all_blocks.p, (uint32_t)s->h_blocks.size(), d.p, s->stream));
  std::vector<unsigned long long> h(M + 1);
  HIP_TRY(hipMemcpyAsync(h.data(), d.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (size_t m = 0; m < M; ++m) counts[m] = h[m];
  return FP_OK;
}

int fp_sim_get_points(fp_sim* s, uint64_t mesh, float* rec) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (!rec) return fail(FP_EINVAL, "records is NULL");
  const MeshSlot& m = s->meshes[mesh];
  const uint64_t n = m.host.n, b = m.vbase;
  if (n == 0) return FP_OK;
  // the records packed in the reference's order on the device, then one copy
  HIP_TRY(s->rec_stage.reserve(10 * n));
  const auto& st = s->state[s->cur];
  const fp::SoA cur{st[0].p, st[1].p, st[2].p, st[3].p, st[4].p, st[5].p};
  HIP_TRY(fp::launch_pack_records(cur, s->mass.p, s->force[0].p, s->force[1].p, s->force[2].p,
                                  m.pos.empty() ? nullptr : s->d_pos.p + b, (uint32_t)b, (uint32_t)n,
                                  s->rec_stage.p, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(rec, s->rec_stage.p, 10 * n * sizeof(float), hipMemcpyDeviceToHost));
  return FP_OK;
}

int fp_sim_set_points(fp_sim* s, uint64_t mesh, const float* rec, uint64_t n) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  MeshSlot& m = s->meshes[mesh];
  if (n != m.host.n) return fail(FP_EINVAL, "State Vector incorrect size! (record count != point count)");
  if (n && !rec) return fail(FP_EINVAL, "records is NULL");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (n) {  // one copy, then the records scattered into the device order on the device
    HIP_TRY(s->rec_stage.reserve(10 * n));
    HIP_TRY(hipMemcpy(s->rec_stage.p, rec, 10 * n * sizeof(float), hipMemcpyHostToDevice));
    const auto& st = s->state[s->cur];
    const fp::SoA cur{st[0].p, st[1].p, st[2].p, st[3].p, st[4].p, st[5].p};
    HIP_TRY(fp::launch_unpack_records(s->rec_stage.p, m.pos.empty() ? nullptr : s->d_pos.p + m.vbase,
                                      (uint32_t)m.vbase, (uint32_t)n, cur, s->mass.p, s->force[0].p, s->force[1].p,
                                      s->force[2].p, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  bool nonzero_f = false;
  for (uint64_t i = 0; i < n && !nonzero_f; ++i)
    for (int k = 7; k < 10; ++k) nonzero_f = nonzero_f || rec[10 * i + k] != 0.0f || std::signbit(rec[10 * i + k]);
  s->forces_dirty = s->forces_dirty || nonzero_f;
  refresh_mass(m, rec, n);
  if (int e = upload_params(s)) return e;
  return build_block_lists(s);
}

int fp_sim_get_vertices(fp_sim* s, uint64_t mesh, float* pos, uint64_t* idx) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  const MeshSlot& m = s->meshes[mesh];
  const uint64_t n = m.host.n;
  if (pos && n) {  // positions packed in the reference's order on the device, one copy
    HIP_TRY(s->rec_stage.reserve(3 * n));
    const auto& st = s->state[s->cur];
    const fp::SoA cur{st[0].p, st[1].p, st[2].p, st[3].p, st[4].p, st[5].p};
    HIP_TRY(fp::launch_pack_positions(cur, m.pos.empty() ? nullptr : s->d_pos.p + m.vbase, (uint32_t)m.vbase,
                                      (uint32_t)n, s->rec_stage.p, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(pos, s->rec_stage.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
  }
  if (idx) std::copy(m.host.f_v.begin(), m.host.f_v.end(), idx);
  return FP_OK;
}

int fp_host_alloc(uint64_t bytes, void** out) {
  if (!out) return fail(FP_EINVAL, "out is NULL");
  *out = nullptr;
  if (bytes == 0) return FP_OK;
  HIP_TRY(hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
  return FP_OK;
}

int fp_host_free(void* ptr) {
  if (ptr) HIP_TRY(hipHostFree(ptr));
  return FP_OK;
}

int fp_sim_accumulate_forces(fp_sim* s, uint64_t mesh) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  const MeshSlot& m = s->meshes[mesh];
  fp::StepArgs a = s->args(true);
  const fp::BlockRange* blk = s->all_blocks.p + m.blk_begin;
  uint32_t nb = m.blk_end - m.blk_begin;
  bool exact = s->math == FP_MATH_EXACT;
  if (m.kernel == MK_GRID_TILED)
    HIP_TRY(tiled_launch(s, a, s->mesh_tiles.p + m.tile_begin, m.tile_end - m.tile_begin, m.launch_shape,
                         fp::MODE_FORCES));
  else if (m.kernel == MK_GRID_SIMPLE) HIP_TRY(fp::launch_grid(a, s->cloths.p, blk, nb, fp::MODE_FORCES, exact, s->stream));
  else {
    HIP_TRY(fp::launch_faces(a, s->csr(), s->all_face_blocks.p + m.fblk_begin, m.fblk_end - m.fblk_begin, exact,
                             s->stream));
    if (s->torsion_on)
      HIP_TRY(fp::launch_hinges(a, s->csr(), s->hinges.p, m.hinge_begin, m.hinge_end - m.hinge_begin, exact,
                                s->stream));
    HIP_TRY(fp::launch_csr(a, s->csr(), blk, nb, fp::MODE_FORCES, exact, s->stream));
  }
  s->forces_dirty = true;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_sim_apply_torsional_forces(fp_sim* s, uint64_t mesh) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  const MeshSlot& m = s->meshes[mesh];
  if (m.hinge_end == m.hinge_begin) return FP_OK;
  fp::StepArgs a = s->args(true);
  const bool exact = s->math == FP_MATH_EXACT;
  HIP_TRY(fp::launch_hinges(a, s->csr(true), s->hinges.p, m.hinge_begin, m.hinge_end - m.hinge_begin, exact,
                            s->stream));
  HIP_TRY(fp::launch_torsion_gather(a, s->csr(true), s->all_blocks.p + m.blk_begin, m.blk_end - m.blk_begin,
                                    s->stream));
  s->forces_dirty = true;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_sim_set_torsion(fp_sim* s, int32_t enabled) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  s->torsion_on = enabled != 0;
  ++s->epoch;  // captured step graphs no longer match
  return FP_OK;
}

int fp_sim_update_points(fp_sim* s, uint64_t mesh, const float* rec, uint64_t n) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  MeshSlot& m = s->meshes[mesh];
  if (n != m.host.n) return fail(FP_EINVAL, "record count != point count");
  if (n && !rec) return fail(FP_EINVAL, "records is NULL");
  if (n == 0) return FP_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (m.perm.empty()) {
    HIP_TRY(s->tmp_rec.upload(rec, 10 * n));
  } else {  // records in device order
    std::vector<float> r(10 * n);
    for (uint64_t d = 0; d < n; ++d) std::memcpy(&r[10 * d], rec + 10 * (uint64_t)m.perm[d], 10 * sizeof(float));
    HIP_TRY(s->tmp_rec.upload(r.data(), r.size()));
  }
  fp::StepArgs a = s->args(true);
  HIP_TRY(fp::launch_update_points(a, s->tmp_rec.p, (uint32_t)m.vbase, s->mass.p, s->all_blocks.p + m.blk_begin,
                                   m.blk_end - m.blk_begin, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->forces_dirty = true;
  // masses may have changed: re-derive the uniform-mass flag from the device copy
  std::vector<float> mm(n), r10(10 * n, 0.0f);
  HIP_TRY(hipMemcpy(mm.data(), s->mass.p + m.vbase, n * sizeof(float), hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < n; ++i) r10[10 * i] = mm[i];  // device order: only the uniformity matters
  refresh_mass(m, r10.data(), n);
  if (int e = upload_params(s)) return e;
  return build_block_lists(s);
}

int fp_sim_clear_forces(fp_sim* s, uint64_t mesh) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  const MeshSlot& m = s->meshes[mesh];
  for (int k = 0; k < 3; ++k) HIP_TRY(fp::launch_fill(s->force[k].p + m.vbase, m.host.n, 0.0f, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_sim_set_kernel(fp_sim* s, int32_t kernel, int32_t math) {
  if (!s) return fail(FP_EINVAL, "sim is NULL");
  if (kernel < FP_KERNEL_AUTO || kernel > FP_KERNEL_GRID_SIMPLE) return fail(FP_EINVAL, "unknown kernel");
  if (math != FP_MATH_EXACT && math != FP_MATH_FAST) return fail(FP_EINVAL, "unknown math mode");
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->kernel = kernel;
  s->math = math;
  return build_block_lists(s);
}

int fp_sim_set_tile_shape(fp_sim* s, uint64_t mesh, int32_t shape) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (shape < 0 || shape >= fp::N_TILE_SHAPES) return fail(FP_EINVAL, "unknown tile shape");
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->meshes[mesh].tile_shape = shape;
  return build_block_lists(s);
}

int fp_sim_get_tile_shape(const fp_sim* s, uint64_t mesh, int32_t* shape) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (!shape) return fail(FP_EINVAL, "shape is NULL");
  *shape = s->meshes[mesh].tile_shape;
  return FP_OK;
}

int fp_sim_get_mesh_kernel(const fp_sim* s, uint64_t mesh, int32_t* kernel) {
  if (int e = check_sim_mesh(s, mesh)) return e;
  if (!kernel) return fail(FP_EINVAL, "kernel is NULL");
  const MeshSlot& m = s->meshes[mesh];
  *kernel = m.kernel == MK_GRID_TILED    ? FP_STEP_TILED
            : m.kernel == MK_GRID_SIMPLE ? FP_STEP_GRID_SIMPLE
            : m.on_patch                 ? FP_STEP_PATCH
                                         : FP_STEP_CSR;
  return FP_OK;
}

int fp_sim_time_steps(fp_sim* s, uint64_t n, float* ms) {
  if (!s || !ms) return fail(FP_EINVAL, "NULL argument");
  HIP_TRY(hipEventRecord(s->ev0, s->stream));
  if (int e = run_steps(s, n)) return e;
  HIP_TRY(hipEventRecord(s->ev1, s->stream));
  HIP_TRY(hipEventSynchronize(s->ev1));
  HIP_TRY(hipEventElapsedTime(ms, s->ev0, s->ev1));
  return FP_OK;
}

void fp_sim_free(fp_sim* s) { delete s; }

}  // extern "C"
