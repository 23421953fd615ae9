// Internal state of a simulation handle (fp_sim), shared by the translation units of the C ABI:
// sim.cpp (construction, device packing, stepping, readback) and sim_surrogate.cpp (the learned
// surrogate on a simulation).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/feriphys_gpu.h"
#include "handles.hpp"
#include "host_util.hpp"
#include "kernels.hpp"
#include "mesh_host.hpp"
#include "obstacles.hpp"
#include "surrogate_host.hpp"

namespace fpsim {

using fph::DBuf;

enum MeshKernel { MK_CSR = 0, MK_GRID_SIMPLE = 1, MK_GRID_TILED = 2 };

struct MeshSlot {
  fp::HostMesh host;
  uint64_t vbase = 0, fbase = 0;
  uint32_t hinge_begin = 0, hinge_end = 0;  // torsional hinges (range in fp_sim::hinges)
  fp_config cfg;
  fp::StepParams params;
  bool uniform_mass = true;
  int kernel = MK_CSR;
  int tile_shape = fp::TILE_16x16;
  int launch_shape = fp::TILE_16x16;  // the shape its tiles are built with under the current math form
  uint32_t blk_begin = 0, blk_end = 0;
  uint32_t fblk_begin = 0, fblk_end = 0;  // face blocks (range in fp_sim::h_face_blocks)
  uint32_t tile_begin = 0, tile_end = 0;  // range in fp_sim::mesh_tiles
  // Device storage order of a general mesh's vertices (empty: the reference order, as for cloths):
  // perm[d] = the reference index stored at device slot vbase + d, pos = its inverse.
  std::vector<uint32_t> perm, pos;
  // device-local starts of the mesh's vertex blocks (one workgroup each), ending with n: patches of
  // at most kBlock connected vertices for general meshes, kBlock strides otherwise
  std::vector<uint32_t> bstart;
  // every patch's vertices and halo fit the patch kernel's LDS (at most kPatchLocal, faces at most
  // kPatchFaceCap), and the struts' nominal constants form at most kPatchClasses pairs
  bool patch_ok = false;
  bool on_patch = false;  // stepped by the patch kernel under the current selection (block lists)
  uint64_t dev(uint64_t i) const { return vbase + (pos.empty() ? i : pos[i]); }
  uint64_t ref(uint64_t d) const { return perm.empty() ? d : perm[d]; }
};

}  // namespace fpsim

struct fp_sim {
  std::vector<fpsim::MeshSlot> meshes;
  uint64_t N = 0, F = 0;
  hipStream_t stream = nullptr;
  int device = 0;  // the device current when the simulation was built (its buffers and stream live there)
  int cur = 0;
  fph::DBuf<float> state[2][6];
  fph::DBuf<float> force[3];
  bool forces_dirty = false;
  fph::DBuf<float> mass;
  // per general mesh, for every reference vertex its device-local slot (MeshSlot::pos), at vbase
  fph::DBuf<uint32_t> d_pos;
  fph::DBuf<float> rec_stage;  // 10-float records packed on the device for one readback
  fph::DBuf<uint32_t> pins;
  fph::DBuf<uint32_t> s_first, s_cnt, f_first, f_cnt, f_v, f_slot;
  fph::DBuf<fp::StrutEnt> s_ent;
  fph::DBuf<float4> face_c;
  // patch kernel: per block of a general mesh its patch and halo (h_patches[b] for block b), the
  // halo ids, per strut entry its compact entry, the meshes' class tables; the launch list of the
  // meshes on it
  std::vector<fp::PatchRange> h_patches;
  fph::DBuf<uint32_t> halo;
  fph::DBuf<fp::PatchStrut> p_ent;
  fph::DBuf<float2> p_cls;
  fph::DBuf<fp::PatchFace> patch_faces;
  fph::DBuf<fp::PatchCorner> p_corner;
  fph::DBuf<fp::PatchRange> patch_list;
  uint32_t n_patch_list = 0;
  uint32_t patch_nf_cap = 0;  // the most faces of a listed patch
  // face blocks of the meshes on the CSR kernel (BlockRange over global face ids), all meshes'
  // face blocks (for the split API, indexed per mesh by fblk_begin/fblk_end)
  fph::DBuf<fp::BlockRange> csr_face_blocks, all_face_blocks;
  std::vector<fp::BlockRange> h_face_blocks;
  uint32_t n_csr_face_blocks = 0;
  fph::ObstacleSet obst;  // obstacle faces and their broad-phase grids
  fph::DBuf<fp::BlockRange> all_blocks, grid_blocks, csr_blocks;
  fph::DBuf<fp::BlockRange> sur_blocks, sur_face_blocks;  // meshes off the tiled kernel (surrogate step)
  uint32_t n_sur_blocks = 0, n_sur_face_blocks = 0;
  std::vector<fp::BlockRange> h_blocks;
  uint32_t n_grid_blocks = 0, n_csr_blocks = 0;
  // LDS-tiled grid kernel: per mesh its tiles (all of the mesh's tile shape), and per shape the
  // tiles of every tiled mesh in XCD-aware launch order
  fph::DBuf<fp::TileRange> tiles[fp::N_TILE_SHAPES], mesh_tiles;
  uint32_t n_tiles[fp::N_TILE_SHAPES] = {0, 0, 0};
  fph::DBuf<fp::ClothView> cloths;
  fph::DBuf<fp::StepParams> params;
  int32_t kernel = FP_KERNEL_AUTO, math = FP_MATH_EXACT;
  fph::DBuf<float> tmp_rec;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // captured step chunks, one per state-buffer parity; valid while `epoch` is unchanged
  uint64_t epoch = 1;
  // what the last upload_params / build_block_lists derived their device lists from: a call that
  // would upload the same lists again does nothing (so a checkpoint restore keeps the graphs)
  std::vector<fp::StepParams> h_params_last;
  std::vector<int32_t> lists_sig;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  uint64_t graph_epoch[2] = {0, 0};
  uint32_t graph_chunk[2] = {0, 0};
  // learned surrogate (fp_sim_set_surrogate): the network's inference image and config, the
  // per-vertex feature slots, and the buffers of a big step
  struct Surrogate {
    bool on = false;
    fp_surrogate_config cfg{};
    fph::DBuf<uint8_t> image;
    fph::DBuf<fp::SlotEnt> slots;
    fph::DBuf<float> slot_k, da, work;
    fph::DBuf<fp::StepParams> big_params;
    fph::DBuf<fp::DiagOut> diag;
    fph::DBuf<float> keep[6];  // surrogate results of the meshes that keep them during a fallback
    fph::DBuf<uint32_t> mesh_of, ids;
    std::vector<fp::SlotEnt> h_slots;
    int grid = 256;
    int variant = 2;  // inference kernel: vertices per wave / 16
    // per mesh: big steps left to run plain before the network is tried again, and the current
    // backoff length (doubles on each rejection up to 16, cleared by an accepted step)
    std::vector<uint32_t> skip, backoff;
  } sur;

  void drop_graphs() {
    for (int k = 0; k < 2; ++k)
      if (graph[k]) { (void)hipGraphExecDestroy(graph[k]); graph[k] = nullptr; }
  }

  ~fp_sim() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  fp::StepArgs args(bool force_init) {
    fp::StepArgs a{};
    a.in = fp::SoA{state[cur][0].p, state[cur][1].p, state[cur][2].p, state[cur][3].p, state[cur][4].p, state[cur][5].p};
    int o = cur ^ 1;
    a.out = fp::SoA{state[o][0].p, state[o][1].p, state[o][2].p, state[o][3].p, state[o][4].p, state[o][5].p};
    a.fx = force[0].p;
    a.fy = force[1].p;
    a.fz = force[2].p;
    a.mass = mass.p;
    a.pins = pins.p;
    a.params = params.p;
    a.obs = obst.view;
    a.force_init = force_init ? 1 : 0;
    return a;
  }
  // torsional hinges of every mesh built with a torsion config (strut order), per vertex their
  // (hinge << 2 | role) entries, and the per-hinge forces; in the step only when torsion_on
  fph::DBuf<fp::HingeEnt> hinges;
  fph::DBuf<uint32_t> t_row, t_ent;
  fph::DBuf<float4> hinge_f;
  uint32_t n_hinges = 0;
  bool torsion_on = false;

  fp::CsrView csr(bool with_torsion = false) const {
    const bool t = (with_torsion || torsion_on) && n_hinges > 0;
    return fp::CsrView{s_first.p, s_cnt.p, s_ent.p, f_first.p, f_cnt.p, f_v.p, f_slot.p, face_c.p, p_ent.p, p_cls.p, halo.p,
                       patch_faces.p, p_corner.p,
                       t ? t_row.p : nullptr, t ? t_ent.p : nullptr, hinge_f.p};
  }
};

namespace fpsim {

fp::StepParams make_params(const fp_config& c, float mass, bool uniform);
int check_sim_mesh(const fp_sim* s, uint64_t mesh);
int run_steps(fp_sim* s, uint64_t n);

}  // namespace fpsim
