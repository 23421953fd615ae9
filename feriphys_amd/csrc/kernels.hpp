// Launch interface of the gfx950 springy-step kernels (kernels.hip).
#pragma once

#include "step_device.hpp"

namespace fp {

struct SoA {
  float *x, *y, *z, *vx, *vy, *vz;
};

// One workgroup's share of one mesh: vertices [begin, end) of mesh `mesh`.
struct BlockRange {
  uint32_t mesh, begin, end, pad;
};

// One incident strut of a vertex: the other end and the strut's constants (16 bytes, one load).
struct StrutEnt {
  uint32_t nbr;
  float len, k, c;
};

// One torsional hinge: the strut (x0, x1), the vertex x2 of its first face and x3 of its second
// that are not on the strut (utils::tuple_difference), its spring constant, damping and rest angle.
struct HingeEnt {
  uint32_t v[4];
  float k, c, rest, pad;
};

// One workgroup of the patch kernel: device vertices [begin, end) of mesh `mesh` (a connected patch
// of at most kBlock vertices), its halo, halo[halo_begin .. halo_begin + n_halo) (every vertex
// outside the patch that a patch vertex's struts or faces reach; patch plus halo is at most
// kPatchLocal), and its faces, patch_faces[face_begin .. face_begin + n_faces) (every face with a
// corner in the patch, at most kPatchFaceCap).
struct PatchRange {
  uint32_t mesh, begin, end, halo_begin, n_halo, face_begin, n_faces, pad;
};
constexpr uint32_t kPatchLocal = 512;
constexpr uint32_t kPatchFaceCap = 2048;
// A face of a patch: its corners' indices in the patch's LDS copy.
struct alignas(8) PatchFace {
  uint16_t v[3], pad;
};
// A face corner of a patch vertex (one per face entry of the vertex, in face-index order): the
// face's index in its patch's face list and the LDS indices of the face's next and previous
// corners, the operands of Face::vertex_angle for this corner.
struct alignas(8) PatchCorner {
  uint16_t q, next, prev, pad;
};
// A strut entry of a patch vertex (8 bytes): the rest length, the neighbour's index in the patch's
// LDS copy and the strut's class, an index into its mesh's table of nominal (stiffness, damping)
// pairs.  Strut::new scales both by NOMINAL_STRUT_LENGTH / length = 1.0 / len
// (springy_mesh.rs:74-90), so the kernel recomputes them bit for bit as (1.0 / len) * nominal.
struct alignas(8) PatchStrut {
  float len;
  uint16_t loc, cls;
};
constexpr uint32_t kPatchClasses = 16;  // classes per mesh; a mesh with more stays on the CSR kernel

// Vertex-centric incidence lists of a packed set of meshes (global device vertex and face ids), in
// a sliced layout: the vertices of one mesh are cut into slices of kSlice consecutive ids (one
// wavefront), and entry j of a vertex is at first[v] + kSlice * j, so that one load instruction of
// a wavefront reads kSlice consecutive entries.
constexpr uint32_t kSlice = 64;
struct CsrView {
  const uint32_t* s_first;  // per vertex: its strut entries, in strut-index order
  const uint32_t* s_cnt;
  const StrutEnt* s_ent;
  const uint32_t* f_first;  // per vertex: its incident faces, in face-index order
  const uint32_t* f_cnt;
  const uint32_t* f_v;      // 3 global vertex ids per face
  const uint32_t* f_slot;   // 3 per face: the position of (face, corner) among its vertex's entries
  // per face entry: that corner's share of its face's drag + lift, (angle / PI) * (drag + lift),
  // written by the face kernel at f_slot
  float4* face_c;
  // patch kernel: per strut entry (the same slots as s_ent) the compact entry, whose neighbour index
  // is into its patch's LDS copy (the patch's own vertices first, then its halo); the class tables
  // (mesh m's at m * kPatchClasses); the halo vertex ids of all patches
  const PatchStrut* p_ent;
  const float2* p_cls;
  const uint32_t* halo;
  const PatchFace* patch_faces;
  const PatchCorner* p_corner;  // per face entry (the slots of face_c)
  // torsional hinges in the step (nullptr: none): per vertex its (hinge << 2 | role) entries in
  // strut order (plain CSR, N+1 row pointers), and the 4 forces per hinge written by the hinge kernel
  const uint32_t* t_row;
  const uint32_t* t_ent;
  float4* hinge_f;
};

// A Cloth::new grid: enough to rebuild its struts, rest lengths and faces on chip.
struct ClothView {
  uint32_t base, rows, cols, pad;
  float spacing, cx, cy, p0, p1;  // construction grid: X(c) = (c*s - cx) + p0, Y(r) = (r*s - cy) + p1
  float tk, tc, sk, sc, bk, bc;
};

struct StepArgs {
  SoA in, out;
  float *fx, *fy, *fz;  // force records (split API / non-zero initial force)
  const float* mass;    // per-vertex mass (used when a mesh's mass is not uniform)
  const uint32_t* pins; // pin bitmask over global vertex ids
  const StepParams* params;
  ObsView obs;
  int32_t force_init;   // 1: start from the force records and clear them after the step
  const float* da;      // surrogate mode: the network's acceleration, 3 per vertex
  float* work;          // surrogate mode: each vertex's face-force work over the step
};

// MODE_SURROGATE: the learned big step (gravity + face forces, the network's acceleration,
// position-difference update; see sur_update in step_device.hpp).
enum StepMode { MODE_STEP = 0, MODE_FORCES = 1, MODE_SURROGATE = 2 };

// One workgroup's tile of a Cloth::new grid for the LDS-tiled kernel: rows [r0, r0+TR) and columns
// [c0, c0+TC) of mesh `mesh` (clipped to the grid).
struct TileRange {
  uint32_t mesh, r0, c0, shape;
};

// Tile shapes compiled for the tiled grid kernel (rows x columns per 256-thread workgroup).  A
// TR x TC tile touches (TR+1) x (TC+1) quads; 15x15 and 7x31 make that exactly 256, one quad (two
// faces) per thread in the face phase.
enum TileShape { TILE_15x15 = 0, TILE_7x31 = 1, TILE_16x16 = 2, N_TILE_SHAPES = 3 };
constexpr int kTileRows[N_TILE_SHAPES] = {15, 7, 16};
constexpr int kTileCols[N_TILE_SHAPES] = {15, 31, 16};

hipError_t launch_csr(const StepArgs& a, const CsrView& t, const BlockRange* blocks, uint32_t n_blocks, int mode,
                      bool exact, hipStream_t s);
// Adds, per mesh (counts[mesh]), the vertices of `blocks` whose state has a NaN or infinity.
hipError_t launch_nonfinite(const SoA& state, const BlockRange* blocks, uint32_t n_blocks, unsigned long long* counts,
                            hipStream_t s);
// Counts, over every finite f32, where the exact kernels' shortened forms differ from the
// operations they replace (bad[0..4], added to; see selftest_kernel).
hipError_t launch_selftest(unsigned long long* bad, hipStream_t s);
// The step (MODE_STEP) of general meshes one patch per workgroup, the neighbour states read from
// the patch's LDS copy; the same results as launch_csr.  nf_cap bounds every patch's faces
// (<= kPatchFaceCap) and sizes the face part of its LDS.
hipError_t launch_patches(const StepArgs& a, const CsrView& t, const PatchRange* patches, uint32_t n_patches,
                          uint32_t nf_cap, bool exact, hipStream_t s);
// Drag + lift and corner weights of the faces in `face_blocks` (BlockRange over global face ids),
// for the CSR kernel that runs after it in the same stream.
hipError_t launch_faces(const StepArgs& a, const CsrView& t, const BlockRange* face_blocks, uint32_t n_blocks,
                        bool exact, hipStream_t s);
// Forces of hinges [first, first + count) (SpringyMesh::apply_torsional_forces, springy_mesh.rs:553-627)
// into t.hinge_f.
hipError_t launch_hinges(const StepArgs& a, const CsrView& t, const HingeEnt* hinges, uint32_t first, uint32_t count,
                         bool exact, hipStream_t s);
// Adds each vertex's hinge forces (summed in strut order from zero) to its force record.
hipError_t launch_torsion_gather(const StepArgs& a, const CsrView& t, const BlockRange* blocks, uint32_t n_blocks,
                                 hipStream_t s);
hipError_t launch_grid(const StepArgs& a, const ClothView* cloths, const BlockRange* blocks, uint32_t n_blocks,
                       int mode, bool exact, hipStream_t s);
// LDS-tiled grid kernel over `n_tiles` tiles that all have tile shape `shape`.
hipError_t launch_grid_tiled(const StepArgs& a, const ClothView* cloths, const TileRange* tiles, uint32_t n_tiles,
                             int shape, int mode, bool exact, hipStream_t s);
// SpringyMesh::update_points on device: old state = a.in (+ force records, masses), new = `nw_rec`
// (10-float Stateful records, indexed by global vertex id); the result is written in place into
// a.in, the force records and the masses.
hipError_t launch_update_points(const StepArgs& a, const float* nw_rec, uint32_t rec_base, float* mass_out,
                                const BlockRange* blocks,
                                uint32_t n_blocks, hipStream_t s);
hipError_t launch_fill(float* p, uint64_t n, float v, hipStream_t s);
// The 10-float Stateful records of device vertices [base, base + n) into out (10 n floats), record
// i from device slot base + pos[i] (pos == nullptr: base + i).
hipError_t launch_pack_records(const SoA& st, const float* mass, const float* fx, const float* fy, const float* fz,
                               const uint32_t* pos, uint32_t base, uint32_t n, float* out, hipStream_t s);
// The positions of device vertices [base, base + n), 3 floats each, in the reference's order.
hipError_t launch_pack_positions(const SoA& st, const uint32_t* pos, uint32_t base, uint32_t n, float* out,
                                 hipStream_t s);
// The inverse: 10 n floats of records (reference order) into the state, masses and forces.
hipError_t launch_unpack_records(const float* in, const uint32_t* pos, uint32_t base, uint32_t n, const SoA& st,
                                 float* mass, float* fx, float* fy, float* fz, hipStream_t s);

constexpr uint32_t kBlock = 256;

}  // namespace fp
