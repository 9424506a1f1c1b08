/* novaphy_hip.h -- C ABI of the MI355X PBF fluid path (libnpx_hip.so).
 *
 * Plain C types only: pointers, sizes, POD structs.  Every entry point returns
 * 0 on success or a negative NPX_E* code; npx_last_error() gives the message
 * for the context (or the last failed create when ctx is NULL).
 *
 * Each entry point replaces one function of the reference's fluid path
 * (FantasyVR/NovaPhy); the reference interface it stands in for is cited on
 * the declaration.  Vectors are AoS float triples in the caller's (original)
 * particle order; the device keeps its own cell-sorted SoA copy.
 */
#ifndef NOVAPHY_HIP_H
#define NOVAPHY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NPX_OK 0
#define NPX_EINVAL (-1)     /* bad argument (std::invalid_argument in the reference) */
#define NPX_EDEVICE (-2)    /* HIP runtime error / no device                          */
#define NPX_ENOMEM (-3)     /* capacity or device memory exhausted                    */
#define NPX_EDOMAIN (-4)    /* a particle left the cell window (see npx_last_error)  */
#define NPX_ESTATE (-5)     /* call out of order (e.g. no particles set)              */
#define NPX_EINTERNAL (-6)  /* unexpected internal failure (an exception inside the
                               library, which never crosses this interface; the
                               reference would throw std::runtime_error)            */

/* Creation flags. */
#define NPX_FLAG_FAST_MATH 1u   /* FMA / rsqrt arithmetic instead of the reference's
                                   operation order (results within tolerance, not
                                   bitwise).  Default: reference operation order.  */
#define NPX_FLAG_NO_GRAPH 2u    /* launch kernels directly, no hipGraph replay      */

typedef struct npx_ctx npx_ctx;

/* Mirror of PBFSettings, include/novaphy/fluid/pbf_solver.h:13-38 (same field
 * order and meaning; corr_delta_q is a fraction of kernel_radius). */
typedef struct npx_pbf_params {
  float rest_density;
  float kernel_radius;
  float particle_radius;
  int32_t solver_iterations;
  float epsilon;
  float xsph_viscosity;
  float vorticity_epsilon;
  float corr_delta_q;
  float corr_k;
  int32_t corr_n;
  int32_t use_domain_bounds;
  float domain_lower[3];
  float domain_upper[3];
} npx_pbf_params;

/* ---------------------------------------------------------------- context */

/* Creates a device context for up to `capacity` particles of one scene
 * (PBFSolver(const PBFSettings&), src/fluid/pbf_solver.cpp:14-16, plus the
 * device buffers).  Uses the current HIP device. */
int npx_create(const npx_pbf_params* params, int32_t capacity, uint32_t flags, npx_ctx** out);

/* Destroys a context and frees its device memory. */
int npx_destroy(npx_ctx* ctx);

/* Message for the last error on ctx (or of the last failed npx_create when
 * ctx is NULL).  Never NULL. */
const char* npx_last_error(const npx_ctx* ctx);

/* Number of visible HIP devices (0 when there is none). */
int32_t npx_device_count(void);

/* Replaces the solver settings (PBFSolver::settings(), pbf_solver.h:79). */
int npx_set_params(npx_ctx* ctx, const npx_pbf_params* params);

/* ParticleState::init(positions, v0) with per-particle velocities
 * (src/fluid/particle_state.cpp:11-20): uploads n particles, zeroes density
 * and lambda.  vel3 may be NULL (zero velocity). */
int npx_set_particles(npx_ctx* ctx, const float* pos3, const float* vel3, int32_t n);

/* Reads the particle state back in original order; any pointer may be NULL.
 * (ParticleState fields, include/novaphy/fluid/particle_state.h:17-22.) */
int npx_get_particles(npx_ctx* ctx, float* pos3, float* vel3, float* rho, float* lambda);

/* Number of particles currently held. */
int32_t npx_num_particles(const npx_ctx* ctx);

/* One PBF step: PBFSolver::step(state, dt, gravity, particle_mass),
 * src/fluid/pbf_solver.cpp:18-127 (predict, grid, I x (density, lambda,
 * delta-p with s_corr, apply), velocity update with speed clamp, XSPH,
 * vorticity confinement when vorticity_epsilon != 0, commit with clamp). */
int npx_step(npx_ctx* ctx, float dt, const float gravity[3], float particle_mass);

/* nsteps steps with the same dt / gravity / mass, without host round trips. */
int npx_step_n(npx_ctx* ctx, float dt, const float gravity[3], float particle_mass, int32_t nsteps);

/* Blocks until all work queued on the context's stream has finished. */
int npx_synchronize(npx_ctx* ctx);

/* Grid diagnostics of the last step's neighbour grid (built on the predicted
 * positions, src/fluid/pbf_solver.cpp:37-51), in original order:
 *   keys      Teschner key of the particle's cell (neighbor_search.cpp:89-97)
 *   cand_cnt  length of its candidate list (query_neighbors, duplicates kept)
 *   nbr_cnt   candidates j != i with squaredNorm(x*_i - x*_j) < h^2
 * Any pointer may be NULL. */
int npx_debug_grid(npx_ctx* ctx, uint64_t* keys, int32_t* cand_cnt, int32_t* nbr_cnt);

/* Flattened candidate lists (original indices) of the last step's grid, in
 * the reference's visiting order; returns the total length in *total (lists
 * are written only when cap >= total). */
int npx_debug_candidates(npx_ctx* ctx, int32_t* lists, int64_t cap, int64_t* total);

/* ------------------------------------------------------ standalone grid */

/* SpatialHashGrid::build(positions) + query_neighbors(point, radius) for a
 * batch of query points (neighbor_search.cpp:14-55): writes per-query
 * candidate counts and, when cap suffices, the concatenated lists. */
int npx_grid_query(const float* pos3, int32_t n, float cell_size, const float* query3,
                   int32_t nq, int32_t* counts, int32_t* lists, int64_t cap, int64_t* total);

/* npx_grid_query without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as npx_grid_query, on the
 * current device's null stream; every device buffer lives in one static
 * per-GPU scratch array of about 23.0 MiB.  Limits: n <= 65536 points,
 * nq <= 65536 queries, a window of <= 2^18 cells (NPX_EDOMAIN beyond),
 * <= 2^21 list entries when lists are read (NPX_ENOMEM beyond).  Calls that
 * overlap on one GPU share the scratch array and race. */
int npx_grid_query_noalloc(const float* pos3, int32_t n, float cell_size, const float* query3,
                           int32_t nq, int32_t* counts, int32_t* lists, int64_t cap, int64_t* total);

/* ------------------------------------------------------ Akinci boundary */

/* Boundary shapes for npx_sample_boundaries: the fields of CollisionShape
   (include/novaphy/core/shape.h) that the samplers read. */
#define NPX_SHAPE_BOX 0
#define NPX_SHAPE_SPHERE 1
#define NPX_SHAPE_PLANE 2
typedef struct npx_shape {
  int32_t type;            /* NPX_SHAPE_*                                     */
  int32_t body;            /* owning body index, -1 for world-owned shapes    */
  float half_extents[3];   /* box                                             */
  float radius;            /* sphere                                          */
  float normal[3];         /* plane (unit)                                    */
  float offset;            /* plane                                           */
  float position[3];       /* shape local transform: translation              */
  float rotation[4];       /* shape local transform: unit quaternion w, x, y, z */
} npx_shape;

/* Replaces sample_model_boundaries (src/fluid/boundary.cpp:128-150) and the
   box / sphere / plane samplers it calls (:15-127): the boundary particles of
   every shape, shape by shape in the reference's loop order, in body-local
   coordinates (world coordinates for planes), computed on the device.
   *count receives the particle count; nothing is written when it exceeds cap
   (call with cap = 0 to size local3 (3 floats each) and body).  Every
   sample equals the reference's float operations bit for bit: the sphere
   samples' acos / sin / cos are the host libm's float functions, evaluated
   in this call before the launch (4 floats per sphere sample). */
int npx_sample_boundaries(const npx_shape* shapes, int32_t nshapes, float spacing, float plane_extent,
                          float* local3, int32_t* body, int64_t cap, int64_t* count);

/* Uploads boundary particles (BoundaryParticle, boundary.h:21-25): body-local
 * positions, owning body (-1 = world) and psi volumes (may be NULL = 0). */
int npx_set_boundary(npx_ctx* ctx, const float* local3, const int32_t* body, const float* psi,
                     int32_t nb, int32_t nbodies);

/* boundary_world_positions + compute_boundary_volumes (boundary.cpp:152-192):
 * psi_b = 1 / sum_b' W(x_b - x_b') at the given body poses (7 floats each:
 * position xyz, quaternion wxyz).  Stores psi on the device; psi_out may be
 * NULL. */
int npx_compute_boundary_volumes(npx_ctx* ctx, const float* pose7, int32_t nbodies,
                                 float* psi_out);

/* boundary_world_positions (boundary.cpp:178-192) for this step's poses;
 * out3 may be NULL. */
int npx_boundary_world_positions(npx_ctx* ctx, const float* pose7, int32_t nbodies, float* out3);

/* FluidWorld::apply_boundary_density (fluid_world.cpp:101-123):
 * rho_i += sum_b rho0 psi_b W(x_i - x_b) on committed positions. */
int npx_apply_boundary_density(npx_ctx* ctx);

/* FluidWorld::apply_coupling_forces (fluid_world.cpp:125-174): per-body
 * force and torque (about the pose position) summed in boundary-particle
 * order, skipping world-owned and static bodies and |F_b|^2 <= 1e-20.
 * Writes nbodies x 3 floats each. */
int npx_coupling_forces(npx_ctx* ctx, const float* pose7, const uint8_t* is_static,
                        int32_t nbodies, float* force3, float* torque3);

/* ------------------------------------------------------ batched scenes */

/* Creates a context that holds many independent scenes in one set of device
 * arrays; every step advances all of them with one launch per pass (no
 * reference counterpart: the reference steps one FluidWorld at a time). */
int npx_batch_create(int32_t slot_capacity, int32_t cell_capacity, int32_t max_scenes,
                     uint32_t flags, npx_ctx** out);

/* Adds a scene (its own PBFSettings, dt, gravity, particle mass and initial
 * state); scene ids are stable until removed.  Bounded domains only. */
int npx_scene_add(npx_ctx* ctx, const npx_pbf_params* params, float dt, const float gravity[3],
                  float particle_mass, const float* pos3, const float* vel3, int32_t n,
                  int32_t* scene_id);

/* Removes a scene and frees its slots and cells. */
int npx_scene_remove(npx_ctx* ctx, int32_t scene_id);

/* Reads one scene's state back in its original particle order. */
int npx_scene_get(npx_ctx* ctx, int32_t scene_id, float* pos3, float* vel3, float* rho,
                  float* lambda);

/* Akinci boundary of a batched scene (FluidWorld's boundary particles for
 * that scene: local positions, owning body or -1 for world-owned, volumes
 * psi as npx_compute_boundary_volumes gives them, and the static flag of each
 * of its nbodies bodies).  Replaces the scene's previous boundary; nb = 0
 * removes it.  Batched counterpart of npx_set_boundary (boundary.cpp,
 * fluid_world.cpp:36-44). */
int npx_scene_set_boundary(npx_ctx* ctx, int32_t scene_id, const float* local3, const int32_t* body,
                           const float* psi, int32_t nb, int32_t nbodies, const uint8_t* is_static);

/* Number of bodies over all scenes with a boundary (the length of the pose
 * and force tables of npx_batch_coupling), or -1. */
int32_t npx_batch_num_bodies(npx_ctx* ctx);

/* After npx_batch_step: boundary world positions from the bodies' poses,
 * boundary density and coupling forces for every scene with a boundary, in
 * one launch per pass over all scenes (FluidWorld::step after the solve,
 * fluid_world.cpp:58-82).  pose7 / out6 hold the bodies of the scenes with a
 * boundary by ascending scene id, each scene's bodies in its own order:
 * position + quaternion (w, x, y, z) in, force + torque out.  Each scene's
 * result is the single-scene FluidWorld's, bit for bit. */
int npx_batch_coupling(npx_ctx* ctx, const float* pose7, int32_t nbodies, float* out6);

/* Advances every scene by nsteps (each with its own dt / gravity / mass). */
int npx_batch_step(npx_ctx* ctx, int32_t nsteps);

/* Launch tunables of a context (any context kind): block size of the
 * neighbour passes (64, 128 or 256; default 128), the XCD-aware block order
 * (1, the default: each of the 8 XCDs sweeps one contiguous run of
 * particles, keeping shared neighbour rows in its own L2, or, in a step
 * whose computing particles fall unevenly over those runs, chunks of 8192
 * slots dealt to the XCDs in turn; C in [2, 4096]: always chunks of C
 * blocks; 0: hardware round-robin order) and the skin of the per-step
 * candidate lists as a fraction of the kernel radius ([0.02, 0.5], default
 * 0.1: a larger skin keeps the lists valid through larger per-step motion, a
 * smaller one visits fewer candidates).  Results are bitwise independent of
 * all three; only speed changes.  No reference counterpart (used by the
 * autotuner). */
int npx_set_tuning(npx_ctx* ctx, int32_t block_size, int32_t xcd_remap, float skin_fraction);

/* Persistent neighbour passes: each pass launches at most `blocks` per-lane
 * workgroups (a multiple of 8, at least 8), each looping over particle blocks
 * in the same XCD order; 0 (the default) launches one workgroup per particle
 * block.  Results are bitwise independent of it; only speed changes.  No
 * reference counterpart (a launch tunable for the autotuner). */
int npx_set_persistent(npx_ctx* ctx, int32_t blocks);

/* Cell classes spanning more than 2^huge_levels 256-slot blocks (default 6;
 * 0..24) are ranked by original index in index space (a histogram, scan and
 * placement over the scene's indices) instead of by merge launches, up to 64
 * such classes per step.  Results are bitwise independent of it; only speed
 * changes.  No reference counterpart (a launch tunable, set low by tests to
 * exercise the index-space rank on small scenes). */
int npx_set_rank_split(npx_ctx* ctx, int32_t huge_levels);

/* Skin lists: 0 (the default) builds none, and every testing pass sweeps the
 * 27 cells inline; 1 builds per-step skin lists in the steps where at least
 * a quarter of the testing lanes sampled in the previous step found their
 * list bound holding, and the testing passes walk them where the bound
 * holds.  Results are bitwise independent of it; only speed changes (with
 * the inline sweep of this library, lists measured slower on every workload
 * tried, moving or at rest).  No reference counterpart (a launch tunable). */
int npx_set_list_policy(npx_ctx* ctx, int32_t policy);

/* npx_batch_step with a device timestamp after every step: step_ms[k] is the
 * duration of step k between consecutive HIP events on the context's stream
 * (may be NULL).  Blocks until the steps are done. */
int npx_batch_step_timed(npx_ctx* ctx, int32_t nsteps, float* step_ms);

/* Per-scene status since the last call, as NPX_STATUS_* bits, and clears it:
 * lets a caller retire one failing scene without failing the batch. */
#define NPX_STATUS_WINDOW 1     /* a particle left the scene's cell window  */
#define NPX_STATUS_NONFINITE 2  /* NaN / inf in positions or velocities      */
int npx_scene_status(npx_ctx* ctx, int32_t* status, int32_t max_scenes);

/* Device bytes needed by a scene of n particles and params (admission). */
int64_t npx_scene_bytes(const npx_pbf_params* params, int32_t n);

/* ------------------------------------------------------ measurement */

/* Enables per-kernel HIP-event timing on the context's stream; while on,
 * kernels launch directly (no graph).  npx_kernel_times returns, per kernel
 * name, the accumulated milliseconds and launch count since enabling. */
int npx_set_kernel_timing(npx_ctx* ctx, int32_t enable);
int npx_kernel_times(npx_ctx* ctx, char* names, int32_t names_cap, float* ms, int32_t* launches,
                     int32_t max_kernels, int32_t* count);

/* The same timed launches one by one, in launch order: names ('\n'-joined),
 * start offset (ms, from the first timed launch's start event) and duration
 * of each; *count is the number of launches recorded.  Used for the GPU
 * track of PerformanceMonitor's trace export.  No reference counterpart. */
int npx_kernel_trace(npx_ctx* ctx, char* names, int32_t names_cap, float* start_ms, float* dur_ms,
                     int32_t max_launches, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* NOVAPHY_HIP_H */
