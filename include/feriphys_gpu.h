/* feriphys_gpu.h — C ABI of the MI355X (gfx950) spring-mass-damper mesh step.
 *
 * This is the drop-in boundary for the springy-mesh path of the reference crate
 * (src/simulation/springy/, src/simulation/state.rs, src/simulation/collidable_mesh.rs).  The
 * reference is a Rust binary crate with no FFI; each entry point below replaces one Rust method
 * and cites it (file:line).  A Rust caller binds these with `extern "C"` declarations; plain
 * pointers and sizes only.
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
/* No counterpart in the reference.  Checks on the current GPU, over every finite f32 x, that the
 * exact kernels' shortened forms equal the operations they stand for: mismatches[0] the
 * reciprocal vs 1 / x, [1] x / PI, [2] x / 3, [3] the branch-free atan2 vs fdlibm's control flow
 * at y = 1 and y = -1.7, [4] the square root vs the compiler's correctly rounded sequence.  All
 * five are 0 on a GPU whose results the exact kernels reproduce. */
int fp_selftest_exact_math(uint64_t mismatches[5]);
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
 * tensile/shear/bending ("binding") nominal spring constants; pinned vertex indices.  rows or cols
 * below 2 returns FP_EINVAL: the reference's bend loops run over 0..(rows - 2) and 0..(cols - 2)
 * (cloth.rs:101-102), which underflow and panic there. */
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
/* No counterpart in the reference: n calls of SpringyMesh::add_strut (springy_mesh.rs:403-413) in
 * order, pairs v01[2 s], v01[2 s + 1], all with the same stiffness and damping.  Every index is
 * checked first; on FP_EINVAL nothing is appended. */
int fp_mesh_add_struts(fp_mesh* mesh, const uint64_t* v01, uint64_t n, float stiffness, float damping);
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
/* One Config per mesh for all meshes at once (an array of n_meshes Configs); one upload. */
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
/* No counterpart in the reference (it lets NaN propagate silently, as this library's steps do):
 * per mesh, the number of vertices whose position or velocity holds a NaN or an infinity;
 * counts has one entry per mesh. */
int fp_sim_count_nonfinite(fp_sim* sim, uint64_t* counts);
/* No counterpart in the reference.  fp_sim_count_nonfinite without any device allocation, for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation): the per-mesh counters live in a fixed device array of the library, one
 * per GPU, of at most 65,536 meshes (a larger simulation gets FP_EINVAL).  Calls that overlap on
 * one GPU share that array and race. */
int fp_sim_count_nonfinite_noalloc(fp_sim* sim, uint64_t* counts);
/* get_meshes()[mesh].get_points() (springy/simulation.rs:51-53, springy_mesh.rs:419-421). */
int fp_sim_get_points(fp_sim* sim, uint64_t mesh, float* records);
/* Checkpoint restore of one mesh's records (same layout; n must equal its point count). */
int fp_sim_set_points(fp_sim* sim, uint64_t mesh, const float* records, uint64_t n);
/* get_meshes()[mesh].get_vertices() (springy_mesh.rs:500-513): 3 floats per vertex and the 3 vertex
 * indices of every face.  Either output may be NULL; a renderer that keeps the (static) index buffer
 * passes NULL for it and gets the positions in one 12-byte-per-vertex copy. */
int fp_sim_get_vertices(fp_sim* sim, uint64_t mesh, float* positions, uint64_t* indices);
/* No counterpart in the reference, whose records live in ordinary Vec<Point> memory.  Page-locked
 * host memory for the buffers of the readback and restore calls above, for callers that read the
 * state back every frame (render loops): a copy to or from it runs at the bus's full rate, where a
 * pageable buffer is staged through the runtime's own pinned chunks.  `*out` is NULL for bytes == 0;
 * fp_host_free releases what fp_host_alloc returned (NULL is a no-op). */
int fp_host_alloc(uint64_t bytes, void** out);
int fp_host_free(void* ptr);
/* SpringyMesh::accumulate_forces (springy_mesh.rs:515-525) with the mesh's config: gravity,
 * struts, faces, pinned F = 0.  The forces appear in fp_sim_get_points records until cleared. */
int fp_sim_accumulate_forces(fp_sim* sim, uint64_t mesh);
/* SpringyMesh::update_points (springy_mesh.rs:423-497) with the simulation's obstacles: the
 * records are the integrated points; collisions are resolved against the pre-call points,
 * then pinned points are restored from them. */
int fp_sim_update_points(fp_sim* sim, uint64_t mesh, const float* new_records, uint64_t n);
/* SpringyMesh::clear_forces (springy_mesh.rs:656-660). */
int fp_sim_clear_forces(fp_sim* sim, uint64_t mesh);
/* SpringyMesh::apply_torsional_forces (springy_mesh.rs:553-627): every strut with a torsional
 * spring (meshes built with a torsion config) adds its hinge forces to its four vertices; each
 * vertex's sum, taken in strut order from zero, is added once to its force record. */
int fp_sim_apply_torsional_forces(fp_sim* sim, uint64_t mesh);
/* The reference comments apply_torsional_forces out of accumulate_forces (springy_mesh.rs:518-519,
 * "torsional forces are broken").  enabled = 1 restores that call in its place, between the strut
 * and the face forces, for every step and fp_sim_accumulate_forces; 0 (the default) is the
 * reference's behaviour.  Meshes built by Cloth::new have no torsional springs. */
int fp_sim_set_torsion(fp_sim* sim, int32_t enabled);

/* Kernel selection.  FP_KERNEL_AUTO uses the LDS-tiled grid kernel for meshes built by Cloth::new
 * and left unchanged, and for other meshes the patch kernel (one workgroup per connected patch of
 * up to 256 vertices, patch and halo staged in LDS, faces computed in the same pass) when every
 * patch and halo fits it (at most 512 vertices, at most 2,048 faces with a corner in the patch,
 * at most 16 strut classes; see fp_sim_get_mesh_kernel), the CSR kernel (one thread per vertex,
 * after a face kernel) otherwise.  FP_KERNEL_CSR puts every mesh on the CSR
 * kernel; FP_KERNEL_GRID_SIMPLE is the one-thread-per-vertex grid kernel without LDS (kept as a
 * cross-check).  All kernels return the same results.  EXACT kernels follow the reference's operation
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
/* The kernel that steps one mesh under the current selection (no counterpart in the reference,
 * whose step has one code path): FP_STEP_TILED (LDS-tiled grid kernel), FP_STEP_GRID_SIMPLE,
 * FP_STEP_PATCH or FP_STEP_CSR.  The patch kernel also needs the mesh's struts to use at most 16
 * distinct nominal (stiffness, damping) pairs (SpringyMesh::new's default, each override, each
 * add_strut call's values). */
#define FP_STEP_TILED 0
#define FP_STEP_GRID_SIMPLE 1
#define FP_STEP_PATCH 2
#define FP_STEP_CSR 3
int fp_sim_get_mesh_kernel(const fp_sim* sim, uint64_t mesh, int32_t* kernel);
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

/* ---------------------------------------------------------------------------------------------
 * Learned large-step surrogate.  No counterpart in the reference: the reference only steps with
 * the explicit integrators of state.rs.  A per-vertex network (72 features -> 3 x 128 GELU -> 3;
 * bf16 weights and activations, f32 accumulation on the matrix cores) stands in for the strut
 * forces over a step of H = ratio * dt.  Its GELU is x * clamp(0.5 + x (a - b min(|x|, x*)), 0, 1),
 * a C1 piecewise quadratic for the normal CDF (surrogate_kernels.hip; within 0.0197 of x Phi(x)).
 * The step:
 *     a  = (gravity + face forces [+ strut forces when base_struts]) / m + out_scale * net(f)
 *     v' = v + H a,   x' = x + H v'
 * then the reference's obstacle collision response over H and pinned records kept.  Features of
 * a vertex, per strut neighbour slot (up to 12; Cloth::new meshes use fixed grid offsets (0,+1),
 * (0,-1), (+1,0), (-1,0), (-1,+1), (+1,-1), (0,+2), (0,-2), (+2,0), (-2,0), other meshes their
 * struts in strut order): strain_scale * strain * d/|d| and (v_nbr - v) * H / L.  Velocities of
 * the surrogate state are position differences over H.  A big step is redone with `ratio` plain
 * steps for every mesh where a strut strain changed by more than strain_limit, or the energy
 * (kinetic + spring + gravity, net of the face forces' work) rose by more than energy_limit times
 * the kinetic + spring energy before and after.  A mesh whose step was redone runs plain for the
 * next 1, 2, 4, ... 16 big steps (doubling on each further rejection, reset by an accepted step)
 * before the network is tried on it again; when every mesh is backing off no network step is
 * attempted.
 * ------------------------------------------------------------------------------------------- */
typedef struct fp_surrogate fp_surrogate;
typedef struct fp_surrogate_config {
  float dt;            /* small step the network replaces `ratio` of (s) */
  int32_t ratio;       /* big step H = ratio * dt */
  float strain_scale;  /* position-feature scale */
  float out_scale;     /* network output unit (m/s^2) */
  int32_t base_struts; /* 1: the big step also applies the plain strut forces */
  float strain_limit;  /* fallback threshold on a strut's strain change per big step */
  float energy_limit;  /* fallback threshold on the relative energy rise per big step */
} fp_surrogate_config;
typedef struct fp_surrogate_stats {
  uint64_t big_steps;      /* big steps taken */
  uint64_t fallbacks;      /* big steps in which at least one mesh fell back to plain steps */
  uint64_t mesh_fallbacks; /* (mesh, big step) pairs that fell back */
  uint64_t plain_only;     /* big steps run plain without a network attempt (all meshes backing off) */
  float max_strain_residual;
  float max_energy_residual;
} fp_surrogate_stats;
/* Parameter vector: W1[128][72], b1[128], W2[128][128], b2[128], W3[128][128], b3[128], W4[3][128],
 * b4[3] (f32, row-major [out][in]). */
#define FP_SURROGATE_PARAMS 42755
#define FP_SURROGATE_FEATURES 72
int fp_surrogate_config_default(fp_surrogate_config* out);
/* New network, weights drawn uniform in +-1/sqrt(fan_in) from `seed` (params NULL) or copied. */
int fp_surrogate_new(const fp_surrogate_config* cfg, uint64_t seed, const float* params, fp_surrogate** out);
int fp_surrogate_get_config(const fp_surrogate* s, fp_surrogate_config* out);
int fp_surrogate_get_params(fp_surrogate* s, float* params);
/* Replaces the weights and resets the optimiser state. */
int fp_surrogate_set_params(fp_surrogate* s, const float* params);
/* One forward and backward pass on the GPU over `rows` samples (features X[rows][72], targets
 * T[rows][3] in output units): the mean squared error, its gradient with respect to every
 * parameter and the outputs Y[rows][3] (any of loss, grad, Y may be NULL). */
int fp_surrogate_grad(fp_surrogate* s, const float* X, const float* T, uint64_t rows, float* loss, float* grad,
                      float* Y);
/* Training set held on the device. */
int fp_surrogate_reserve(fp_surrogate* s, uint64_t capacity);
int fp_surrogate_sample_count(const fp_surrogate* s, uint64_t* out);
int fp_surrogate_clear_samples(fp_surrogate* s);
int fp_surrogate_add_samples(fp_surrogate* s, const float* X, const float* T, uint64_t rows);
/* `steps` Adam steps (beta 0.9 / 0.999) on batches of `batch` samples drawn from the training set
 * by a hash of (seed, step); losses (may be NULL) receives each step's batch loss. */
int fp_surrogate_train(fp_surrogate* s, uint64_t steps, uint32_t batch, float lr, uint64_t seed, float* losses);
void fp_surrogate_free(fp_surrogate* s);
/* Runs n_big * ratio plain steps of the simulation and adds one sample per free vertex and big
 * step b = 0 ... n_big - 1 to the network's training set: the features of the state at big step b
 * (positions and position-difference velocities, perturbed by normal noise of `noise` m; the state
 * before the first step is taken to be at rest) and the acceleration that carries it to the plain
 * solver's positions at big step b + 1.  A mesh whose plain state stops being finite or reaches
 * 30 m/s or 100 % strain contributes nothing from then on.  All meshes must use the network's dt. */
int fp_sim_collect_samples(fp_sim* sim, fp_surrogate* s, uint64_t n_big, float noise, uint64_t seed);
/* Copies the network into the simulation (NULL detaches).  All meshes must use its dt. */
int fp_sim_set_surrogate(fp_sim* sim, const fp_surrogate* s);
/* Shape of the network's inference kernel: 32 vertices per wave (1024-thread workgroups, four
 * waves per SIMD; the default) or 64 (512-thread workgroups, two waves per SIMD). */
int fp_sim_set_surrogate_shape(fp_sim* sim, int32_t vertices_per_wave);
/* n_big surrogate steps (each ratio * dt), with per-mesh fallback to plain steps.  Blocks. */
int fp_sim_step_surrogate(fp_sim* sim, uint64_t n_big, fp_surrogate_stats* stats);
/* The same, timed with HIP events on the simulation's stream: total_ms for the whole run,
 * mlp_ms for the network kernel's launches alone. */
int fp_sim_time_surrogate(fp_sim* sim, uint64_t n_big, float* total_ms, float* mlp_ms, fp_surrogate_stats* stats);
/* The network's input features (n x 72) and output accelerations (n x 3, m/s^2) for one mesh at
 * the current state, and its slot table (n x 12 local neighbour ids, -1 empty; rest lengths). */
int fp_sim_surrogate_features(fp_sim* sim, uint64_t mesh, float* features);
int fp_sim_surrogate_accel(fp_sim* sim, uint64_t mesh, float* accel);
int fp_sim_surrogate_slots(fp_sim* sim, uint64_t mesh, int64_t* nbr, float* rest);

/* ---- rigid bodies (rigidbody/rigidbody.rs, rigidbody/simulation.rs) ----------------------- */
/* A body is the reference's unit cube (graphics/forms.rs:91-146) with a 29-float Stateful record
 * (rigidbody.rs:62-93): position, rotation (v.x, v.y, v.z, s), linear momentum, angular momentum,
 * mass, the initial inverse inertia by columns, accumulated force, accumulated torque.  A
 * simulation steps every body on the GPU, one thread per body, bit for bit the reference's f32
 * arithmetic. */
typedef struct fp_rigid fp_rigid;         /* RigidBody (rigidbody.rs:201-206)                      */
typedef struct fp_rigid_sim fp_rigid_sim; /* rigidbody::Simulation (rigidbody/simulation.rs:10-14) */
#define FP_RIGID_RECORD 29
typedef struct fp_rigid_config { /* rigidbody/config.rs:7-13 */
  int32_t integration;           /* FP_EULER / FP_RK4 */
  float dt;
  float coefficient_of_restitution;
  float gravity[3]; /* added to the accumulated force every step */
  float torque[3];  /* added to the accumulated torque every step */
} fp_rigid_config;
/* Config::default() (rigidbody/config.rs:15-25): RK4, dt 1 ms, restitution 0.7, no gravity or torque. */
void fp_rigid_config_default(fp_rigid_config* out);
/* RigidBody::new (rigidbody.rs:213-251): a unit cube at `position`, at rest, inertia mass/6 on the
 * diagonal (inverted as cgmath's Matrix3::invert does); FP_EINVAL when it is not invertible. */
int fp_rigid_new(const float position[3], float mass, fp_rigid** out);
/* get_state().as_state() (rigidbody.rs:253-255, 62-93): the 29-float record. */
int fp_rigid_get_state(const fp_rigid* body, float* record);
/* RigidBody::apply_impulse (rigidbody.rs:345-352): impulse at `position` relative to the centre. */
int fp_rigid_apply_impulse(fp_rigid* body, const float impulse[3], const float position[3]);
void fp_rigid_free(fp_rigid* body);
/* Simulation::new (rigidbody/simulation.rs:17-24): takes ownership of the bodies and obstacles
 * (their handles are freed by this call, on success and on failure); the config starts as
 * fp_rigid_config_default. */
int fp_rigid_sim_new(fp_rigid* const* bodies, uint64_t n_bodies, fp_obstacle* const* obstacles, uint64_t n_obstacles,
                     fp_rigid_sim** out);
void fp_rigid_sim_free(fp_rigid_sim* sim);
int fp_rigid_sim_count(const fp_rigid_sim* sim, uint64_t* n_bodies);
/* The reference sets its Config only from its GUI (rigidbody/simulation.rs:64-76); these set and
 * read it directly. */
int fp_rigid_sim_set_config(fp_rigid_sim* sim, const fp_rigid_config* config);
int fp_rigid_sim_get_config(const fp_rigid_sim* sim, fp_rigid_config* config);
/* Simulation::step (rigidbody/simulation.rs:26-52) n_steps times; *timestep (may be NULL) receives
 * Duration::from_secs_f32(dt) as f32. */
int fp_rigid_sim_step(fp_rigid_sim* sim, uint64_t n_steps, float* timestep);
/* No counterpart in the reference: n_steps steps bracketed by HIP events on the simulation's
 * stream, in milliseconds. */
int fp_rigid_sim_time_steps(fp_rigid_sim* sim, uint64_t n_steps, float* ms);
/* get_rigidbodies()[i].get_state() for every body: n_bodies x 29 floats. */
int fp_rigid_sim_get_states(fp_rigid_sim* sim, float* records);
/* No counterpart in the reference: restores every body's record (a checkpoint written by
 * fp_rigid_sim_get_states); n must be the body count. */
int fp_rigid_sim_set_states(fp_rigid_sim* sim, const float* records, uint64_t n);
/* rigidbodies[body].apply_impulse (rigidbody/simulation.rs:73-75 applies the GUI's impulse to body 0). */
int fp_rigid_sim_apply_impulse(fp_rigid_sim* sim, uint64_t body, const float impulse[3], const float position[3]);

#ifdef __cplusplus
}
#endif
#endif /* FERIPHYS_GPU_H */
