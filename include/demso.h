/*
 * demso.h -- C ABI of the MI355X (gfx950) DemSo time step.
 *
 * The five stages of DemSo's step (uniform-grid hash, key/index sort,
 * reorder + cellStart/cellEnd, collide, integrate) as hand-written HIP
 * kernels behind plain `extern "C"` functions: device pointers and sizes in,
 * an int status out.  Reference paths below are relative to the DemSo source
 * tree (src/...).
 *
 * Conventions
 *   - Every function returns DEM_OK (0) or a negative dem_status; the text of
 *     the last failure is available from dem_last_error().  Nothing exits.
 *   - Pointers named pos/vel/acc are float2 arrays passed as float* (x,y
 *     interleaved, 8 bytes per particle), as in src/functions.cu:326.  All
 *     other per-particle arrays are 4-byte float or uint32.  All device
 *     pointers must be 4-byte aligned (8 for float2 arrays).
 *   - All stage calls are asynchronous on the context's HIP stream and do no
 *     host synchronisation and no allocation once the context has reserved
 *     scratch for the sizes in use (dem_reserve), so a step can be captured
 *     into a hipGraph.  The sort's scratch is grown on demand if a call needs
 *     more than was reserved.
 *   - A context is not thread-safe; use one context per host thread/stream.
 */
#ifndef DEMSO_H
#define DEMSO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEMSO_ABI_VERSION 1
#define DEM_EMPTY_CELL 0xFFFFFFFFu
#define DEM_MAX_TYPES 4096u /* material rows per context (a table of per-pair contact parameters is kept) */

typedef enum dem_status {
  DEM_OK = 0,
  DEM_ERR_INVALID = -1,     /* bad argument or inconsistent sizes */
  DEM_ERR_NO_DEVICE = -2,   /* no HIP device / wrong device index */
  DEM_ERR_ALLOC = -3,       /* device or host allocation failed */
  DEM_ERR_LAUNCH = -4,      /* kernel launch or HIP runtime call failed */
  DEM_ERR_STATE = -5,       /* call order: parameters not set, etc. */
  DEM_ERR_NONFINITE = -6    /* finite-state guard tripped */
} dem_status;

/* dem_system_props.flags */
#define DEM_GRID_CLAMP 0x1u /* clamp-to-edge hashing; out-of-range neighbour
                               cells are skipped.  Without it the grid wraps
                               with `& (G-1)` as src/particles_kernel.cuh:127-128
                               does, which needs power-of-two grid sizes. */
#define DEM_OPEN_TOP 0x2u   /* no top wall (src/particles_kernel.cuh:611-613 skipped) */

/* Global parameters: SystemProperties (src/main.cuh:108-122), field for
 * field, followed by two additions. */
typedef struct dem_system_props {
  uint32_t num_particles;
  float cube_dimension[2];   /* wall box [0,Lx] x [0,Ly] */
  uint32_t grid_size[2];     /* cells per axis */
  uint32_t num_cells;        /* grid_size[0]*grid_size[1] */
  float time_step;
  float gravity[2];
  float boundary_normal_stiffness;
  float boundary_shear_stiffness;
  /* additions */
  float grid_extent[2];      /* hashing domain: cell = floor(p*G/extent); {0,0} = cube_dimension */
  uint32_t flags;            /* DEM_GRID_CLAMP | DEM_OPEN_TOP */
} dem_system_props;

/* Per-type material: ParticleProperties (src/main.cuh:39-54), field for
 * field, followed by two additions. */
typedef struct dem_particle_props {
  float radius;
  float mass;
  float inertia;
  float normal_stiffness;
  float shear_stiffness;
  float normal_damping;
  float boundary_damping;
  float friction_coefficient;
  float attract_coefficient;
  float color_r, color_g, color_b; /* carried, unused by the step */
  /* additions */
  float areal_density;  /* with a per-particle radius array: m = (areal*r)*r, I = 2*m*r*r/5 */
  uint32_t instance;    /* ensemble instance this type belongs to (0 for a single system) */
} dem_particle_props;

/* One independent system of an ensemble: its own wall box and grid.  Its
 * cells occupy keys [cell_base, cell_base + grid_size[0]*grid_size[1]). */
typedef struct dem_instance {
  float cube_dimension[2];
  float grid_extent[2];      /* {0,0} = cube_dimension */
  uint32_t grid_size[2];
  uint32_t cell_base;
  uint32_t reserved;
} dem_instance;

/* The controlled ("big") particle of DemSo's default build (USE_BIG_PARTICLE,
 * src/main.cuh:29): a disc of material row `type` that every particle
 * collides with after its neighbours (src/particles_kernel.cuh:525-543) and
 * that is moved kinematically after each step (src/main.cu:510-514): pos +=
 * move (computed in double), then pos.y < wrap_below_y -> reset_y and
 * pos.x > wrap_above_x -> reset_x.  DemSo's values: pos (5, 9.5), vel 0,
 * move (0, -0.005), wrap_below_y -1 -> 10.5, wrap_above_x Lx + 0.5 -> -0.5. */
typedef struct dem_controlled {
  float pos[2];
  float vel[2];
  float theta;
  float omega;
  uint32_t type;
  uint32_t enabled;          /* 0: no controlled particle */
  double move[2];
  double wrap_below_y;
  double wrap_above_x;
  float reset_y;
  float reset_x;
} dem_controlled;

typedef struct dem_context dem_context;

/* ---- context, parameters ------------------------------------------------ */

/* Create a context on HIP device `device`, launching on `hip_stream`
 * (a hipStream_t; NULL = the null stream). */
int dem_create(dem_context** out, int device, void* hip_stream);
int dem_destroy(dem_context* ctx);
/* Last error text of ctx (ctx NULL: of the last failed dem_create). */
const char* dem_last_error(const dem_context* ctx);
int dem_abi_version(void);
int dem_set_stream(dem_context* ctx, void* hip_stream);
int dem_synchronize(dem_context* ctx);

/* Replaces the constant-memory upload of allocateVectors
 * (src/functions.cu:145-147): global parameters plus `num_types` material
 * rows (the reference allows 10; here up to DEM_MAX_TYPES).  Defines a single
 * instance from sys (box, grid, extent, cell_base 0). */
int dem_set_params(dem_context* ctx, const dem_system_props* sys,
                   const dem_particle_props* types, uint32_t num_types);
/* Ensemble geometry; replaces the single instance set by dem_set_params.
 * Material rows name their instance in dem_particle_props.instance. */
int dem_set_instances(dem_context* ctx, const dem_instance* inst, uint32_t num_instances);
/* Replace the material rows (same layout and limits as dem_set_params);
 * each row's `instance` must name an existing instance.  Used to give every
 * ensemble instance its own material after dem_set_instances. */
int dem_set_materials(dem_context* ctx, const dem_particle_props* types, uint32_t num_types);
/* set_gravity (src/functions.cu:663-668). */
int dem_set_gravity(dem_context* ctx, float gx, float gy);
/* Pre-allocate sort scratch for up to max_particles (call before capture):
 * two pairs of key/value arrays (16 bytes a particle) and the per-tile digit
 * counts (256 words per 8,192 keys). */
int dem_reserve(dem_context* ctx, uint32_t max_particles);

/* ---- the five stages, reference signatures ------------------------------ */

/* calcHash (src/functions.cu:315-327; kernel src/particles_kernel.cuh:133-151):
 * gridParticleHash[i] = cell key of pos[i], gridParticleIndex[i] = i. */
int dem_calc_hash(dem_context* ctx, const float* pos, uint32_t* gridParticleIndex,
                  uint32_t* gridParticleHash, uint32_t numParticles);

/* sortParticles (src/functions.cu:336-341): ascending sort of the hash keys
 * with the index array as payload, in place.  Stable: equal keys keep their
 * input order (the reference's thrust::sort_by_key leaves ties unspecified). */
int dem_sort_particles(dem_context* ctx, uint32_t* gridParticleHash, uint32_t* gridParticleIndex,
                       uint32_t numParticles);

/* No counterpart in the reference.  dem_sort_particles for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the same sort and result, with
 * its scratch in a fixed per-GPU array of the library (32.5 MiB, held until
 * the process ends) instead of buffers the context allocates.  Sorts at
 * most DEM_SORT_NOALLOC_MAX keys (larger counts fail with DEM_ERR_INVALID
 * before anything is launched).  All calls on one GPU share that array:
 * calls that overlap on one GPU (other streams, other contexts, other host
 * threads) race and must be serialised by the caller. */
#define DEM_SORT_NOALLOC_MAX 4194304u
int dem_sort_particles_noalloc(dem_context* ctx, uint32_t* gridParticleHash, uint32_t* gridParticleIndex,
                               uint32_t numParticles);

/* reorderDataAndFindCellStart (src/functions.cu:347-393; kernel
 * src/particles_kernel.cuh:156-233): cellStart := EMPTY for all numCells,
 * then cellStart/cellEnd of every occupied cell, the sorted copies of the
 * state and sortedLoc[ID] = i.  cellEnd of empty cells is left untouched, as
 * in the reference. */
int dem_reorder_data_and_find_cell_start(
    dem_context* ctx, uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos, float* sortedVel,
    float* sortedTheta, float* sortedOmega, uint32_t* sortedID, uint32_t* sortedLoc,
    uint32_t* sortedType, const uint32_t* gridParticleHash, const uint32_t* gridParticleIndex,
    const float* oldPos, const float* oldVel, const float* oldTheta, const float* oldOmega,
    const uint32_t* oldID, const uint32_t* oldType, uint32_t numParticles, uint32_t numCells);

/* collide (src/functions.cu:398-468; kernel src/particles_kernel.cuh:459-549
 * without the controlled-particle term): pressure := 0, then for every
 * particle the spring-dashpot + Coulomb-capped shear force of all contacts in
 * the 3x3 cell neighbourhood; acc = F/m, alpha = M/I. */
int dem_collide(dem_context* ctx, const float* pos, const float* vel, float* acc,
                const float* omega, float* alpha, const uint32_t* type, const uint32_t* cellStart,
                const uint32_t* cellEnd, uint32_t numParticles, uint32_t numCells, float* pressure);

/* integrateSystem (src/functions.cu:529-551; kernel
 * src/particles_kernel.cuh:580-627): symplectic Euler, theta wrap, the
 * reflect-clamp side/top walls and the bottom-wall law, in place. */
int dem_integrate_system(dem_context* ctx, float* pos, float* vel, const float* acc, float* theta,
                         float* omega, const float* alpha, const uint32_t* type, float* pressure,
                         uint32_t numParticles);

/* ---- extended stages: per-particle radius and ensembles ----------------- */
/* Same operations; `radius` arrays (may be NULL = radius from the type row)
 * travel with the state; `type` in dem_calc_hash_ex selects the instance
 * (may be NULL when there is one instance). */
int dem_calc_hash_ex(dem_context* ctx, const float* pos, const uint32_t* type,
                     uint32_t* gridParticleIndex, uint32_t* gridParticleHash, uint32_t numParticles);
int dem_reorder_ex(dem_context* ctx, uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos,
                   float* sortedVel, float* sortedTheta, float* sortedOmega, uint32_t* sortedID,
                   uint32_t* sortedLoc, uint32_t* sortedType, float* sortedRadius,
                   const uint32_t* gridParticleHash, const uint32_t* gridParticleIndex,
                   const float* oldPos, const float* oldVel, const float* oldTheta,
                   const float* oldOmega, const uint32_t* oldID, const uint32_t* oldType,
                   const float* oldRadius, uint32_t numParticles, uint32_t numCells);
int dem_collide_ex(dem_context* ctx, const float* pos, const float* vel, float* acc,
                   const float* omega, float* alpha, const uint32_t* type, const float* radius,
                   const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t numParticles,
                   uint32_t numCells, float* pressure);
int dem_integrate_ex(dem_context* ctx, float* pos, float* vel, const float* acc, float* theta,
                     float* omega, const float* alpha, const uint32_t* type, const float* radius,
                     float* pressure, uint32_t numParticles);

/* ---- owned simulation: the step loop of src/main.cu:396-507 -------------- */

typedef struct dem_sim dem_sim;

/* Device state for n particles on a grid of num_cells total keys, double
 * buffered as in the reference (src/functions.cu:68-99).  with_radius != 0
 * adds a per-particle radius array.  Parameters come from ctx, which must
 * outlive the sim. */
int dem_sim_create(dem_context* ctx, uint32_t n, uint32_t num_cells, int with_radius, dem_sim** out);
int dem_sim_destroy(dem_sim* sim);
/* Host or device arrays -> current device state (radius may be NULL without
 * radius). */
int dem_sim_upload(dem_sim* sim, const float* pos, const float* vel, const float* theta,
                   const float* omega, const uint32_t* id, const uint32_t* type, const uint32_t* loc,
                   const float* radius);
/* Current device state -> host arrays; any pointer may be NULL. */
int dem_sim_download(dem_sim* sim, float* pos, float* vel, float* theta, float* omega, uint32_t* id,
                     uint32_t* type, uint32_t* loc, float* radius, float* acc, float* alpha,
                     float* pressure, uint32_t* hash, uint32_t* index, uint32_t* cellStart,
                     uint32_t* cellEnd);
/* Run `steps` steps with eager launches on the context stream.  Under the
 * automatic sort choice (dem_sim_set_sort_mode 0) and from 262,144
 * particles up, the call queues the steps in groups of 4 and, before a
 * group, waits for the group two back to finish, so that each step's sort
 * method is picked from recent counts; it returns with up to 8 steps still
 * running (not synchronized). */
int dem_sim_step(dem_sim* sim, uint32_t steps);
/* Run `steps` steps by replaying a captured hipGraph of the step sequence
 * (two steps a replay; under the automatic sort choice the replays are
 * queued in groups of 4, as dem_sim_step queues steps). */
int dem_sim_step_graph(dem_sim* sim, uint32_t steps);
/* Average device time per stage over `reps` eager steps, measured with HIP
 * events on the context stream: ms[0]=hash ms[1]=sort ms[2]=cell fill
 * ms[3]=reorder ms[4]=collide ms[5]=integrate ms[6]=whole step.  The owned
 * simulation computes the next step's hash in integrate (ms[0] is 0 for
 * re-sorted steps) and, if dem_sim_autotune chose so, runs collide and
 * integrate as one kernel (timed in ms[4]; ms[5] is then 0). */
int dem_sim_stage_times(dem_sim* sim, uint32_t reps, float* ms);
/* Choose block sizes per kernel, whether the sim runs collide and
 * integrate as one kernel or two, and how collide loads its neighbour
 * candidates (DEM_STEP_PREFETCH, DEM_STEP_BATCH, DEM_STEP_LIST or neither), and
 * how many particles (1, 2 or 4) a lane of hash, reorder and integrate
 * takes, by timing candidates on the current state (the state is not
 * modified).  Writes the choice to block[5] (hash, reorder, collide,
 * integrate, flags: the DEM_STEP_* bits chosen, and in bits 8-9, 10-11 and
 * 12-13 log2 of the particles a lane of hash, reorder and integrate) if
 * non-NULL. */
int dem_sim_autotune(dem_sim* sim, uint32_t* block);
/* Sort method of the sim's steps.  After a step the state is in the sorted
 * order of its keys, so the next step can re-sort: only particles whose cell
 * changed are sorted and merged with the rest (same result as the radix
 * sort, bit for bit).  mode 0 (default): re-sort while the last re-sort's
 * counts (dem_sim_resort_counts) stay small (movers at most max(4096, n/64);
 * in place, at most n/4 particles listed), else radix sort and try the
 * re-sort again after 1, 2, 4 ... 64 radix-sorted steps (8, 16 ... 64 when a
 * count was twice its limit or more); 1: radix sort every
 * step; 2: re-sort whenever the state allows it.  The first step after
 * dem_sim_upload, dem_sim_set_by_id (positions), dem_sim_autotune or
 * dem_adj_backward always radix sorts. */
int dem_sim_set_sort_mode(dem_sim* sim, int mode);
/* Particles whose cell changed in the last re-sorted step (exact once the
 * stream is synchronized). */
int dem_sim_mover_count(dem_sim* sim, uint32_t* count);
/* The last re-sort's counts (exact once the stream is synchronized): movers
 * as dem_sim_mover_count, and listed, the particles the last in-place
 * re-sort moved to another slot or gave a new key (the movers plus the
 * particles their moves shift), whose state it staged and wrote back.  No
 * counterpart in the reference (its steps always radix sort); for callers
 * that choose the sort mode themselves. */
int dem_sim_resort_counts(dem_sim* sim, uint32_t* movers, uint32_t* listed);
/* Kernel variants of the owned simulations of ctx (dem_sim_autotune sets
 * the fastest): DEM_STEP_FUSED runs collide and integrate as one kernel,
 * DEM_STEP_PREFETCH makes collide fetch each neighbour candidate one
 * iteration ahead, DEM_STEP_BATCH makes it load the first candidate of all
 * three neighbour rows at once (not together with PREFETCH; the fused
 * kernel loads one at a time instead).  Results are the same bit for bit. */
#define DEM_STEP_FUSED 0x1
#define DEM_STEP_PREFETCH 0x2
#define DEM_STEP_BATCH 0x4
/* Re-sorted steps of the split layout normally keep the state in its buffer
 * and move only the particles whose sorted slot changes (plus the cell
 * ranges around them); DEM_STEP_COPY makes them reorder the whole state
 * into the other buffer instead, as radix-sorted steps do.  Same results. */
#define DEM_STEP_COPY 0x8
/* Collide normally addresses its arrays by 32-bit offsets while they stay
 * under 4 GiB (up to 2^29 particles and 2^30 cells) and with 64-bit
 * addresses and one candidate at a time beyond; DEM_STEP_WIDE makes it use
 * the latter at every size.  Same results. */
#define DEM_STEP_WIDE 0x10
/* DEM_STEP_LIST makes collide first mark, row by row, the candidates that
 * may touch (position and radius only) and then run the contact law over
 * the marked ones: fewer idle lanes on contact-dense states (not together
 * with PREFETCH or BATCH).  Same results. */
#define DEM_STEP_LIST 0x20
/* Particles a lane of hash, reorder and integrate: log2 of 1, 2 or 4 in bits
 * 8-9 (hash), 10-11 (reorder) and 12-13 (integrate); 0 is one a lane.
 * Same results. */
#define DEM_STEP_PER_HASH(lg) (((lg) & 3) << 8)
#define DEM_STEP_PER_REORDER(lg) (((lg) & 3) << 10)
#define DEM_STEP_PER_INTEGRATE(lg) (((lg) & 3) << 12)
#define DEM_STEP_PER_MASK 0x3F00
int dem_set_step_variant(dem_context* ctx, int flags);
/* Number of radix passes the sort uses for this sim's key range. */
int dem_sim_sort_passes(dem_sim* sim);
/* No counterpart in the reference (its particle count is a global,
 * src/main.cuh:109).  The sim's particle count and the HIP device of its
 * context, so callers can check array sizes before handing them over. */
int dem_sim_info(dem_sim* sim, uint32_t* n, int* device);
/* Current positions / velocities gathered by particle ID into device arrays
 * (n float2 each; either may be NULL). */
int dem_sim_get_by_id(dem_sim* sim, float* pos_by_id, float* vel_by_id);
/* Overwrite current positions / velocities from device arrays indexed by
 * particle ID (either may be NULL). */
int dem_sim_set_by_id(dem_sim* sim, const float* pos_by_id, const float* vel_by_id);

/* Output rows of followed particles, replacing the per-value copies of
 * writeOutputFile (src/functions.cu:608-661): for each of `count` particle
 * IDs (device array follow_ids), row f of `rows` (device, count x
 * DEM_ROW_WORDS uint32) receives id, type, pos.x, pos.y, vel.x, vel.y,
 * acc.x, acc.y, theta, omega, alpha, pressure of the particle at sorted
 * index loc[ID] (floats as their bit patterns; an ID >= numParticles gives
 * a row of 0xFFFFFFFF).  Asynchronous on the context stream. */
#define DEM_ROW_WORDS 12
int dem_gather_rows(dem_context* ctx, const uint32_t* follow_ids, uint32_t count, const uint32_t* loc,
                    const float* pos, const float* vel, const float* acc, const float* theta, const float* omega,
                    const float* alpha, const uint32_t* id, const uint32_t* type, const float* pressure,
                    uint32_t numParticles, uint32_t* rows);
/* The same rows from the sim's current state. */
int dem_sim_gather_rows(dem_sim* sim, const uint32_t* follow_ids, uint32_t count, uint32_t* rows);
/* ---- scene generators (src/functions.cu:198-311) ---------------------------- */
/* Stream of the x draws of dem_create_rectangles (y: DEM_RECT_STREAM + 1). */
#define DEM_RECT_STREAM 1000u
/* createRectangles: side[0] x side[1] particles p = x + y*side[0] at
 * start + sideLength/(side-1) * ((x, y) + (normal - 0.5)/100), ID = loc =
 * startID + p, type = particleTypeVec[(x+y) % numParticleTypes] (host array).
 * The reference's curand_normal draws (seeded with time(NULL)) are replaced
 * by reproducible normals: draw p of streams DEM_RECT_STREAM (x) and
 * DEM_RECT_STREAM + 1 (y) under `seed`.  pos/ID/loc/type are device arrays
 * that start at the rectangle's first particle.  Synchronizes. */
int dem_create_rectangles(dem_context* ctx, float* pos, uint32_t* ID, uint32_t* loc, uint32_t* type,
                          const float start[2], const float sideLength[2], const uint32_t side[2], uint32_t startID,
                          uint32_t numParticleTypes, const uint32_t* particleTypeVec, uint64_t seed);
/* The same with the x draws from `stream` and the y draws from stream + 1
 * (the driver gives rectangle r the streams 1000 + 2r, 1001 + 2r). */
int dem_create_rectangles_ex(dem_context* ctx, float* pos, uint32_t* ID, uint32_t* loc, uint32_t* type,
                             const float start[2], const float sideLength[2], const uint32_t side[2],
                             uint32_t startID, uint32_t numParticleTypes, const uint32_t* particleTypeVec,
                             uint64_t seed, uint64_t stream);
/* createTriangles: numParticles (at most N(N+1)/2) particles of an N-row
 * triangle, row spacing `height`, particle spacing `space`; type =
 * particleTypeVec[(C+L) % numParticleTypes].  Synchronizes. */
int dem_create_triangles(dem_context* ctx, float* pos, uint32_t* ID, uint32_t* loc, uint32_t* type,
                         const float start[2], uint32_t N, uint32_t numParticleTypes, const uint32_t* particleTypeVec,
                         float space, float height, uint32_t startID, uint32_t numParticles);
/* createSingleParticles: copies the given particles (host or device arrays)
 * and numbers them startID, startID + 1, ...  Synchronizes. */
int dem_create_single_particles(dem_context* ctx, float* pos, float* vel, float* theta, float* omega, uint32_t* ID,
                                uint32_t* loc, uint32_t* type, const float* usrPos, const float* usrVel,
                                const float* usrTheta, const float* usrOmega, const uint32_t* usrType,
                                uint32_t numParticles, uint32_t startID);
/* ---- images (src/functions.cu:557-600, 684-693) ---------------------------- */
/* RenderParameters (src/main.cuh:124-135).  DemSo's values: image_dim =
 * (Lx/Ly * imageheight, imageheight), control_dim = 2 ceil(pixels per metre *
 * controlled radius) (at least 2), bg_color 0, view_rotations 1,
 * color_by_pressure 0. */
typedef struct dem_render_params {
  int32_t image_dim[2];      /* imageDIMx, imageDIMy */
  int32_t control_dim[2];    /* dimx, dimy */
  int32_t bg_color;          /* byte written to every channel of the background */
  uint32_t view_rotations;   /* a black line from each centre along theta */
  uint32_t color_by_pressure;
  float max_pressure;        /* pressure scale (dem_update_pressure_scale) */
} dem_render_params;
/* plotParticles: an RGBA image (device, image_dim[0] * image_dim[1] * 4 bytes,
 * row 0 at the bottom) of the particles in array order, coloured by type or
 * by pressure / max_pressure, shaded towards the rim, with rotation lines and
 * the controlled particle (host state, NULL or not enabled: none) on top.
 * Where discs overlap, the higher array index wins (the reference leaves it
 * to thread timing).  radius may be NULL (type radii).  Synchronizes. */
int dem_plot_particles(dem_context* ctx, uint8_t* image, const float* pos, const float* theta, const uint32_t* type,
                       const float* radius, uint32_t numParticles, const dem_render_params* rp,
                       const dem_controlled* control, const float* pressure);
/* The same image of an owned simulation's current state (sorted order, its
 * controlled particle if enabled). */
int dem_sim_plot(dem_sim* sim, uint8_t* image, const dem_render_params* rp);
/* updatePressureScale: the largest pressure of the n values (NaN ignored),
 * or 1 if none is positive, to the host.  Synchronizes. */
int dem_update_pressure_scale(dem_context* ctx, const float* pressure, uint32_t numParticles, float* maxPressure);
/* The same over the owned simulation's current pressure. */
int dem_sim_pressure_scale(dem_sim* sim, float* maxPressure);
/* ---- controlled particle -------------------------------------------------- */
/* Owned simulations of ctx: with state->enabled, every step adds the pair
 * with the controlled particle to each particle's force, reduces the
 * reaction on it (the reference's reduceVecD order: a pairwise tree over
 * blocks of 512 sorted particles, the block sums added in order) and then
 * moves it.  All on the device, graph-capturable. */
int dem_set_controlled(dem_context* ctx, const dem_controlled* state);
/* Current state of the controlled particle (after the steps run so far) and
 * the last step's reaction: force x, y and moment on it.  Synchronizes. */
int dem_get_controlled(dem_context* ctx, dem_controlled* state, float reaction[3]);
/* collide with the controlled particle, the reference's USE_BIG_PARTICLE
 * collide (src/functions.cu:398-521, USE_ATOMIC 0): control_state = (pos.x,
 * pos.y, vel.x, vel.y, theta, omega) on the host; forceVecX/Y, momentVec
 * (device, n floats) receive the per-particle reaction and blockX/Y/M
 * (device, ceil(n/512) floats) its block sums; ctrlF (2) and ctrlM (1) on the
 * host receive the totals.  Synchronizes, as the reference does. */
int dem_collide_controlled(dem_context* ctx, const float* pos, const float* vel, float* acc, const float* omega,
                           float* alpha, const uint32_t* type, const uint32_t* cellStart, const uint32_t* cellEnd,
                           uint32_t numParticles, uint32_t numCells, const float control_state[6],
                           uint32_t control_type, float* forceVecX, float* blockX, float* forceVecY, float* blockY,
                           float* momentVec, float* blockM, float* ctrlF, float* ctrlM, float* pressure);
/* Update the contact parameters of every material from a host array of
 * num_types rows (kn, kt, cn, mu), leaving the other fields unchanged. */
int dem_set_contact_params(dem_context* ctx, const double* kn_kt_cn_mu, uint32_t num_types);

/* ---- adjoint of collide + integrate ---------------------------------------
 * No counterpart in the reference.  Reverse mode through `steps` steps with
 * the contact set, approach-only damping and stick/slip frozen at the
 * forward's decisions (hashing and sorting are not differentiated: the
 * permutation is a constant), checkpoints every `checkpoint_every` steps and
 * recomputation of each segment (its neighbour lists are rebuilt, not
 * stored).  FP32 forward, FP64 adjoint state and parameter gradients.
 * Attraction is not covered (dem_adj_create fails if a material attracts). */
typedef struct dem_adjoint dem_adjoint;
int dem_adj_create(dem_sim* sim, uint32_t checkpoint_every, dem_adjoint** out);
int dem_adj_destroy(dem_adjoint* adj);
/* Run `steps` steps from the sim's current state, keeping checkpoints. */
int dem_adj_forward(dem_adjoint* adj, uint32_t steps);
/* How many dem_adj_forward calls have started on this adjoint (successful
 * or not).  dem_adj_backward differentiates the latest of them, so a caller
 * that holds an earlier count knows its trajectory was replaced. */
int dem_adj_generation(dem_adjoint* adj, uint64_t* generation);
/* Given dL/dpos and dL/dvel of the final state (device, n float2 each,
 * indexed by particle ID; either may be NULL), write dL/dpos0, dL/dvel0 of
 * the state dem_adj_forward started from (device, by ID; either may be
 * NULL) and dL/d(kn, kt, cn, mu) of every material (device, num_types x 4
 * doubles; may be NULL).  The sim is left in the final state. */
int dem_adj_backward(dem_adjoint* adj, const float* grad_pos, const float* grad_vel, float* grad_pos0,
                     float* grad_vel0, double* grad_params);

#ifdef __cplusplus
}
#endif

#endif /* DEMSO_H */
