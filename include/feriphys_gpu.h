/* feriphys_gpu.h — C ABI of the MI355X (gfx950) spring-mass-damper mesh step.
 *
 * This is the drop-in boundary for the springy-mesh path of the reference crate
 * (src/simulation/springy/, src/simulation/state.rs, src/simulation/collidable_mesh.rs).  The
 * reference is a Rust binary crate with no FFI; each entry point below replaces one Rust method
 * and cites it (file:line).  A Rust caller binds these with `extern "C"` declarations (see
 * INTEGRATION.md); plain pointers and sizes only.
 *
 * Ownership mirrors the Rust API: meshes and obstacles are built on the host, then MOVED into a
 * simulation (fp_sim_new consumes the handles, as Simulation::new takes Vec<SpringyMesh> by value).
 * The simulation owns all device memory.  Caller buffers are copied in at create time and copied
 * out by the get_* calls; nothing aliases caller memory.
 *
 * Errors: the reference panics (bad index, bad record width, non-manifold edge); every entry point
 * here returns an int status instead, with a message available from fp_last_error().
 *
 * Arithmetic: f32 throughout, as the reference.  Vertex records use the reference's Stateful
 * layout for Point (springy_mesh.rs:199-256): 10 floats [mass, x, y, z, vx, vy, vz, Fx, Fy, Fz].
 */
#ifndef FERIPHYS_GPU_H
#define FERIPHYS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------- */
#define FP_OK 0
#define FP_EINVAL 1        /* bad argument: index out of range, wrong record count, NULL      */
#define FP_ENONMANIFOLD 2  /* an edge with more than two faces (springy_mesh.rs:339)           */
#define FP_EHIP 3          /* HIP runtime error (message from hipGetErrorString)               */
#define FP_ESTATE 4        /* handle used after it was moved into a simulation                 */
#define FP_ENOMEM 5        /* host or device allocation failed                                 */
#define FP_ENODEV 6        /* no gfx950 device visible                                         */

#define FP_STATE_WIDTH 10  /* Point::num_state_elements (springy_mesh.rs:200-202)              */

typedef struct fp_mesh fp_mesh;         /* SpringyMesh (springy_mesh.rs:258-264)               */
typedef struct fp_obstacle fp_obstacle; /* CollidableMesh (collidable_mesh.rs:50-55)           */
typedef struct fp_sim fp_sim;           /* springy::Simulation (springy/simulation.rs:10-15)   */

/* state.rs:5-8 */
#define FP_EULER 0
#define FP_RK4 1

/* springy/config.rs:9-18.  gravity is applied as a FORCE per vertex (springy_mesh.rs:194-196). */
typedef struct fp_config {
  int32_t integration; /* FP_EULER or FP_RK4 */
  float dt;            /* seconds */
  float gravity[3];
  float wind[3];
  float lift_coefficient;
  float drag_coefficient;
  float coefficient_of_restitution;
  float coefficient_of_friction;
} fp_config;

/* SpringConfig keyed by StrutKey (springy_mesh.rs:38-42, 95-107): (a,b) in any order. */
typedef struct fp_strut_override {
  uint64_t a, b;
  float constant;
  float damping;
} fp_strut_override;

/* TorsionalSpringConfig (springy_mesh.rs:33-36). */
typedef struct fp_torsion_config {
  float spring_constant;
  float spring_damping;
} fp_torsion_config;

/* ---- library ------------------------------------------------------------------------------ */
/* Message for the last non-FP_OK status returned on the calling thread. */
const char* fp_last_error(void);
/* Library version string. */
const char* fp_version(void);
/* Config::default() (springy/config.rs:20-33): RK4, dt = Duration::from_millis(1), gravity
 * (0,-10,0), no wind, lift 1, drag 1, restitution 0.95, friction 0.3. */
void fp_config_default(fp_config* out);

/* ---- SpringyMesh construction (host side; no device work) ---------------------------------- */
/* SpringyMesh::new (springy_mesh.rs:267-401).  positions: n_vertices x 3 floats; indices:
 * n_indices (multiple of 3) triangle corners.  `mass` is the TOTAL mesh mass (mass / N per
 * vertex, :277).  torsion may be NULL (None).  overrides may be NULL with n_overrides 0 (None). */
int fp_mesh_new(const float* positions, uint64_t n_vertices, const uint64_t* indices, uint64_t n_indices,
                float mass, float default_stiffness, float default_damping, const fp_torsion_config* torsion,
                const fp_strut_override* overrides, uint64_t n_overrides, fp_mesh** out);

/* Cloth::new (springy/cloth.rs:20-113): rows x cols grid in the xy plane, centred on `position`;
 * tensile/shear/bending ("binding") nominal spring constants; pinned vertex indices. */
int fp_cloth_new(uint64_t rows, uint64_t cols, float spacing, const float position[3], float point_mass,
                 float tensile_stiffness, float tensile_damping, float shear_stiffness, float shear_damping,
                 float binding_stiffness, float binding_damping, const uint64_t* pinned_vertices,
                 uint64_t n_pinned, fp_mesh** out);

/* Extension of Cloth::new with the grid plane chosen: plane 0 = xy (identical to fp_cloth_new),
 * plane 1 = xz (rows along +z, position.y is the cloth's height).  The xz cloth is the same
 * SpringyMesh::new / add_pin / add_strut sequence applied to component-permuted positions. */
#define FP_PLANE_XY 0
#define FP_PLANE_XZ 1
int fp_cloth_new_in_plane(uint64_t rows, uint64_t cols, float spacing, const float position[3], float point_mass,
                          float tensile_stiffness, float tensile_damping, float shear_stiffness,
                          float shear_damping, float binding_stiffness, float binding_damping,
                          const uint64_t* pinned_vertices, uint64_t n_pinned, int32_t plane, fp_mesh** out);

/* SpringyMesh::add_strut (springy_mesh.rs:403-413): appended, no de-duplication. */
int fp_mesh_add_strut(fp_mesh* mesh, uint64_t v0, uint64_t v1, float stiffness, float damping);
/* SpringyMesh::add_pin (springy_mesh.rs:415-417): appended, duplicates kept.  The index is
 * checked when the mesh is moved into a simulation (the reference panics at first use). */
int fp_mesh_add_pin(fp_mesh* mesh, uint64_t index);

/* Introspection of the built topology (the reference's private fields, read by its tests). */
int fp_mesh_counts(const fp_mesh* mesh, uint64_t* n_points, uint64_t* n_struts, uint64_t* n_faces,
                   uint64_t* n_pins);
/* struts: v01 (2 per strut, (min,max) for triangle edges, as given for add_strut), kcl (stiffness,
 * damping, rest length), faces (2 per strut, -1 = None).  Any output pointer may be NULL. */
int fp_mesh_get_struts(const fp_mesh* mesh, uint64_t* v01, float* kcl, int64_t* faces);
/* Torsional springs built by SpringyMesh::new (:348-393): has (0/1), spring, damping, rest angle. */
int fp_mesh_get_torsion(const fp_mesh* mesh, uint8_t* has, float* kcr);
int fp_mesh_get_faces(const fp_mesh* mesh, uint64_t* strut_indices3, uint64_t* vertex_indices3);
int fp_mesh_get_pins(const fp_mesh* mesh, uint64_t* pins);
/* get_points (springy_mesh.rs:419-421) as 10-float Stateful records. */
int fp_mesh_get_points(const fp_mesh* mesh, float* records);
/* Replace every point with the given 10-float records (a Stateful checkpoint restore; the record
 * format of State::as_vector, state.rs:64-70).  n must equal the point count. */
int fp_mesh_set_points(fp_mesh* mesh, const float* records, uint64_t n);
/* get_vertices (springy_mesh.rs:500-513): positions (3 per point) and the flat face index list. */
int fp_mesh_get_vertices(const fp_mesh* mesh, float* positions, uint64_t* indices);
void fp_mesh_free(fp_mesh* mesh);

/* ---- CollidableMesh ----------------------------------------------------------------------- */
/* CollidableMesh::new (collidable_mesh.rs:57-105). */
int fp_obstacle_new(const float* positions, uint64_t n_vertices, const uint64_t* indices, uint64_t n_indices,
                    fp_obstacle** out);
int fp_obstacle_counts(const fp_obstacle* obstacle, uint64_t* n_vertices, uint64_t* n_edges, uint64_t* n_faces);
void fp_obstacle_free(fp_obstacle* obstacle);

/* ---- Simulation (device resident) ---------------------------------------------------------- */
/* Simulation::new (springy/simulation.rs:18-25): takes ownership of the meshes and obstacles
 * (their handles are freed by this call, on success and on failure) and uploads them to the GPU.
 * The config starts as Config::default(). */
int fp_sim_new(fp_mesh* const* meshes, uint64_t n_meshes, fp_obstacle* const* obstacles, uint64_t n_obstacles,
               fp_sim** out);
/* The reference sets Config only through its GUI (springy/simulation.rs:64-77, every frame);
 * these set and read it directly.  fp_sim_set_mesh_config gives one mesh its own Config, which is
 * what one reference Simulation per mesh computes (meshes do not interact). */
int fp_sim_set_config(fp_sim* sim, const fp_config* config);
int fp_sim_get_config(const fp_sim* sim, fp_config* config);
int fp_sim_set_mesh_config(fp_sim* sim, uint64_t mesh, const fp_config* config);
/* One Config per mesh for all meshes at once (configs[n_meshes]); one upload. */
int fp_sim_set_mesh_configs(fp_sim* sim, const fp_config* configs, uint64_t n_meshes);
/* Simulation::step (springy/simulation.rs:27-45), n_steps times; blocks until done.
 * timestep_out (may be NULL) receives Duration::from_secs_f32(dt).as_secs_f32(). */
int fp_sim_step(fp_sim* sim, uint64_t n_steps, float* timestep_out);
/* The same, enqueued on the simulation's HIP stream without waiting. */
int fp_sim_step_async(fp_sim* sim, uint64_t n_steps);
int fp_sim_synchronize(fp_sim* sim);
/* Simulation::get_timestep (springy/simulation.rs:47-49). */
int fp_sim_get_timestep(const fp_sim* sim, float* out);
int fp_sim_mesh_count(const fp_sim* sim, uint64_t* out);
int fp_sim_mesh_counts(const fp_sim* sim, uint64_t mesh, uint64_t* n_points, uint64_t* n_struts,
                       uint64_t* n_faces, uint64_t* n_pins);
/* get_meshes()[mesh].get_points() (springy/simulation.rs:51-53, springy_mesh.rs:419-421). */
int fp_sim_get_points(fp_sim* sim, uint64_t mesh, float* records);
/* Checkpoint restore of one mesh's records (same layout; n must equal its point count). */
int fp_sim_set_points(fp_sim* sim, uint64_t mesh, const float* records, uint64_t n);
/* get_meshes()[mesh].get_vertices() (springy_mesh.rs:500-513). */
int fp_sim_get_vertices(fp_sim* sim, uint64_t mesh, float* positions, uint64_t* indices);
/* SpringyMesh::accumulate_forces (springy_mesh.rs:515-525) with the mesh's config: gravity,
 * struts, faces, pinned F = 0.  The forces appear in fp_sim_get_points records until cleared. */
int fp_sim_accumulate_forces(fp_sim* sim, uint64_t mesh);
/* SpringyMesh::update_points (springy_mesh.rs:423-497) with the simulation's obstacles: the
 * records are the integrated points; collisions are resolved against the pre-call points,
 * then pinned points are restored from them. */
int fp_sim_update_points(fp_sim* sim, uint64_t mesh, const float* new_records, uint64_t n);
/* SpringyMesh::clear_forces (springy_mesh.rs:656-660). */
int fp_sim_clear_forces(fp_sim* sim, uint64_t mesh);

/* Kernel selection.  FP_KERNEL_AUTO uses the LDS-tiled grid kernel for meshes built by Cloth::new
 * and left unchanged, the CSR kernel otherwise; FP_KERNEL_GRID_SIMPLE is the one-thread-per-vertex
 * grid kernel without LDS (kept as a cross-check).  EXACT kernels follow the reference's operation
 * order with no FMA contraction, correctly rounded division and square root and the C library's
 * atan2f rounding (bit-identical results); FAST kernels use fused multiply-adds and the hardware
 * reciprocal and square root (one-step forces within 1e-5 relative). */
#define FP_KERNEL_AUTO 0
#define FP_KERNEL_CSR 1
#define FP_KERNEL_GRID 2
#define FP_KERNEL_GRID_SIMPLE 3
#define FP_MATH_EXACT 0
#define FP_MATH_FAST 1
int fp_sim_set_kernel(fp_sim* sim, int32_t kernel, int32_t math);
/* Tile shape of the LDS-tiled grid kernel for one cloth mesh (rows x columns per workgroup):
 * FP_TILE_15x15, FP_TILE_7x31 or FP_TILE_16x16.  The default is chosen from the cloth's size. */
#define FP_TILE_15x15 0
#define FP_TILE_7x31 1
#define FP_TILE_16x16 2
int fp_sim_set_tile_shape(fp_sim* sim, uint64_t mesh, int32_t shape);
int fp_sim_get_tile_shape(const fp_sim* sim, uint64_t mesh, int32_t* shape);
/* Autotuner: meshes on the tiled grid kernel are grouped by size class (longer side < 16, < 32,
 * < 64, >= 64 vertices); for each class every tile shape is timed over `steps_per_trial` steps
 * (HIP events on the simulation's stream, state restored afterwards) and the fastest is kept.
 * chosen_out[4] (may be NULL) receives the shape per class (-1: no mesh in the class). */
int fp_sim_autotune(fp_sim* sim, uint32_t steps_per_trial, int32_t* chosen_out);
/* n_steps of Simulation::step replayed from HIP graphs that capture `chunk` steps each (the
 * launch sequence of one chunk, captured once per state-buffer parity and reused until the
 * simulation's kernels, configs or tiles change); any remainder runs eagerly.  Blocks. */
int fp_sim_step_graph(fp_sim* sim, uint64_t n_steps, uint32_t chunk);
/* Runs n_steps on the simulation's stream bracketed by HIP events recorded on that stream;
 * kernel_ms receives the event-measured time of the whole run (device time, inputs resident). */
int fp_sim_time_steps(fp_sim* sim, uint64_t n_steps, float* kernel_ms);
/* The same through fp_sim_step_graph's captured chunks. */
int fp_sim_time_steps_graph(fp_sim* sim, uint64_t n_steps, uint32_t chunk, float* kernel_ms);
void fp_sim_free(fp_sim* sim);

#ifdef __cplusplus
}
#endif
#endif /* FERIPHYS_GPU_H */
