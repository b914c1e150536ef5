/*
 * sph_abi.h - C ABI of libsph_hip.so, the MI355X (gfx950) implementation of
 * TinySPH's per-step compute: the Clavet double-density relaxation 2D SPH
 * step of AdamSimpson/SPH.
 *
 * Three layers, all plain C linkage, plain pointers and sizes:
 *
 *   1. The reference's records, layout-identical, so AoS buffers that the
 *      reference's host code fills can be handed over unchanged
 *      (src/fluid.h:56-106, src/hash.h:35-48, src/geometry.h:33-40).
 *   2. A device-resident context API (sph_create / sph_upload / sph_step /
 *      sph_download ...).  The whole domain lives in one SoA particle array
 *      in HBM, sorted by grid cell; it replaces the reference's per-rank
 *      AoS + pointer array + MPI halo exchange (src/communication.c).
 *   3. Reference-named entry points with the exact prototypes of
 *      src/fluid.h:112-126 and src/hash.h:50-52, so the reference's main
 *      loop (src/fluid.c:270-372) links against this library unchanged.
 *
 * Error convention of layer 2: 0 on success, a negative SPH_E* code on
 * failure; sph_last_error() returns a message for the calling thread's last
 * failure.  Layer 3 keeps the reference's void returns and records failures
 * the same way.
 *
 * Threading: one host thread per context; every operation is ordered on the
 * context's own HIP stream.  Calls on one context from several threads race.
 */
#ifndef SPH_ABI_H
#define SPH_ABI_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define SPH_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define SPH_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

/* ------------------------------------------------------------------------
 * 1. Reference records (layout-identical)
 * ---------------------------------------------------------------------- */

/* Included after the reference's own src/fluid.h (include guard
 * fluid_fluid_h), the records below come from there and are only checked;
 * otherwise they are defined here with the same layout. */
#ifdef fluid_fluid_h
#define SPH_ABI_HAVE_REFERENCE_TYPES 1
#else

#define SPHERE_MOVER 0    /* src/fluid.h:48 */
#define RECTANGLE_MOVER 1 /* src/fluid.h:49 */

typedef struct FLUID_PARTICLE fluid_particle;
typedef struct NEIGHBOR neighbor;
typedef struct PARAM param;
typedef struct TUNABLE_PARAMETERS tunable_parameters;
typedef struct AABB_T AABB_t;
typedef struct BUCKET_T bucket_t;
typedef struct NEIGHBOR_GRID_T neighbor_grid_t;
/* MPI-only records (src/communication.h:45-63); only passed by pointer. */
typedef struct EDGE_T edge_t;
typedef struct OOB_T oob_t;

/* src/fluid.h:56-70 - 52 bytes. id is the particle's index in the
 * reference's pointer array; on one domain it never changes. */
struct FLUID_PARTICLE {
    float x_prev;
    float y_prev;
    float x;
    float y;
    float v_x;
    float v_y;
    float a_x;
    float a_y;
    float density;
    float density_near;
    float pressure;
    float pressure_near;
    int id;
};

/* src/fluid.h:72-75 */
struct NEIGHBOR {
    fluid_particle **fluid_neighbors;
    int number_fluid_neighbors;
};

/* src/fluid.h:78-97 - 64 bytes */
struct TUNABLE_PARAMETERS {
    float rest_density;
    float smoothing_radius;
    float g;
    float k;
    float k_near;
    float k_spring;
    float sigma;
    float beta;
    float time_step;
    float node_start_x;
    float node_end_x;
    float mover_center_x;
    float mover_center_y;
    float mover_width;
    float mover_height;
    char mover_type;
    char kill_sim;
    char active;
};

/* src/fluid.h:100-106 - 80 bytes */
struct PARAM {
    tunable_parameters tunable_params;
    int number_fluid_particles_global;
    int number_fluid_particles_local;
    int max_fluid_particle_index;
    int number_halo_particles;
};

/* src/geometry.h:33-40 - 24 bytes */
struct AABB_T {
    float min_x;
    float max_x;
    float min_y;
    float max_y;
    float min_z;
    float max_z;
};

/* src/hash.h:35-38 */
struct BUCKET_T {
    fluid_particle **fluid_particles;
    unsigned int number_fluid;
};

/* src/hash.h:40-48 */
struct NEIGHBOR_GRID_T {
    float spacing;
    unsigned int size_x;
    unsigned int size_y;
    neighbor *neighbors;
    bucket_t *grid_buckets;
    unsigned int max_neighbors;
    unsigned int max_bucket_size;
};

#endif /* fluid_fluid_h */

SPH_STATIC_ASSERT(sizeof(fluid_particle) == 52, "fluid_particle must be 52 bytes");
SPH_STATIC_ASSERT(offsetof(fluid_particle, x) == 8, "fluid_particle.x offset");
SPH_STATIC_ASSERT(offsetof(fluid_particle, v_x) == 16, "fluid_particle.v_x offset");
SPH_STATIC_ASSERT(offsetof(fluid_particle, density) == 32, "fluid_particle.density offset");
SPH_STATIC_ASSERT(offsetof(fluid_particle, id) == 48, "fluid_particle.id offset");
SPH_STATIC_ASSERT(sizeof(tunable_parameters) == 64, "tunable_parameters must be 64 bytes");
SPH_STATIC_ASSERT(offsetof(tunable_parameters, mover_type) == 60, "mover_type offset");
SPH_STATIC_ASSERT(offsetof(tunable_parameters, active) == 62, "active offset");
SPH_STATIC_ASSERT(sizeof(param) == 80, "param must be 80 bytes");
SPH_STATIC_ASSERT(offsetof(param, number_fluid_particles_global) == 64, "param.global offset");
SPH_STATIC_ASSERT(offsetof(param, number_halo_particles) == 76, "param.halo offset");
SPH_STATIC_ASSERT(sizeof(AABB_t) == 24, "AABB_t must be 24 bytes");
SPH_STATIC_ASSERT(sizeof(bucket_t) == 16, "bucket_t must be 16 bytes");
SPH_STATIC_ASSERT(sizeof(neighbor) == 16, "neighbor must be 16 bytes");
SPH_STATIC_ASSERT(sizeof(neighbor_grid_t) == 40, "neighbor_grid_t must be 40 bytes");

/* ------------------------------------------------------------------------
 * 2. Device-resident context API
 * ---------------------------------------------------------------------- */

#define SPH_OK 0
#define SPH_EINVAL (-1)    /* bad argument (null pointer, size, id set ...) */
#define SPH_ENOMEM (-2)    /* device or host allocation failed             */
#define SPH_EHIP (-3)      /* a HIP runtime call failed                     */
#define SPH_ESTATE (-4)    /* call out of order (e.g. step before upload)   */
#define SPH_ENODEV (-5)    /* no usable gfx950 device                       */

/* Capacity limits the reference silently applies (src/fluid.c:173-175).
 * By default the device path has no cap; it counts how often the
 * reference's caps would have dropped work, because parity with the
 * reference's lists holds only while those counts stay zero.
 * sph_set_bucket_cap reproduces the bucket cap on request. */
#define SPH_REF_MAX_BUCKET 100
#define SPH_REF_MAX_NEIGHBORS 400

typedef struct sph_ctx sph_ctx;

/* Device-side health counters, accumulated since the last reset.  The
 * neighbour counts (neighbor_overflow, max_neighbors) of a step's second hash,
 * the lists the reference's next viscosity pass uses (fluid.c:341), are taken
 * by this library's next viscosity pass, which walks exactly that set: after
 * T steps they cover both hashes of steps 1..T-1 and the first of step T. */
typedef struct {
    uint64_t steps;             /* full steps taken                               */
    uint64_t hashes;            /* cell sorts performed                           */
    uint64_t bucket_overflow;   /* cells that held more than 100 particles        */
    uint64_t neighbor_overflow; /* particles with more than 400 half-neighbours   */
    uint64_t key_out_of_range;  /* keys >= cell count (x==max_x on the top row,
                                   y==max_y, NaN); clamped to the last cell       */
    uint64_t nonfinite;         /* particles with a NaN/inf coordinate after a step */
    uint64_t max_cell_count;    /* largest cell occupancy seen                    */
    uint64_t max_neighbors;     /* largest full (both-way) neighbour count seen   */
} sph_counters_t;

/* Create a context for up to `capacity` particles on the current device.
 * `params` supplies the tunables (smoothing radius = grid spacing, as in
 * src/fluid.c:173-176); `boundary` is the simulation box.  The box must
 * start at the origin (the reference assumes so: src/fluid.c:117, hash.c:41). */
int sph_create(sph_ctx **ctx, const param *params, const AABB_t *boundary, int capacity);
void sph_destroy(sph_ctx *ctx);

/* Several independent scenes stepped together by the same launches (the
 * batched configuration).  Each scene has its own box, tunables and
 * particle capacity; scenes never interact.  sph_create is the one-scene
 * case.  Scenes are numbered 0..nscenes-1 in the order given. */
typedef struct {
    tunable_parameters tunables;
    AABB_t boundary; /* must start at the origin */
    int capacity;
} sph_scene_desc;

int sph_create_batch(sph_ctx **ctx, const sph_scene_desc *scenes, int nscenes);
int sph_num_scenes(const sph_ctx *ctx);
/* Particles of scene s, grid size of its box. */
int sph_scene_info(const sph_ctx *ctx, int scene, int *n, unsigned *size_x, unsigned *size_y);
/* Upload all scenes: particles holds counts[0] records of scene 0, then
 * counts[1] of scene 1, ...; ids are local to each scene (a permutation of
 * 0..counts[s]-1).  sph_download writes back in the same concatenated
 * layout (scene s's particle id at offset sum(counts[0..s)) + id). */
int sph_upload_batch(sph_ctx *ctx, const fluid_particle *particles, const int *counts);
/* Tunables of one scene (sph_set_tunables sets every scene). */
int sph_set_scene_tunables(sph_ctx *ctx, int scene, const tunable_parameters *tunables);

/* Extensions with no counterpart in the reference (off by default):
 * sloshing - a horizontal box acceleration A sin(2 pi t / period), applied
 * like gravity (v_x += -A sin(2 pi t / period) dt) at the start of every
 * fused step, t = steps since the last upload times dt;
 * frozen scenes - flags[s] != 0 makes sph_step leave scene s unchanged
 * (used to re-take single scenes of a batch). */
int sph_set_scene_forcing(sph_ctx *ctx, int scene, float amplitude, float period);
int sph_set_scene_frozen(sph_ctx *ctx, const int *flags);

/* The reference's bucket cap (src/hash.c:153-164, max_bucket_size = 100 at
 * src/fluid.c:174): each cell's bucket holds its first `cap` particles in
 * insertion order, and a particle left out is in no bucket of that hash -
 * it has no neighbours, is nobody's neighbour, gets density 0 and no
 * relaxation displacement, and its next viscosity pass sees no pairs;
 * gravity, prediction, boundary and velocity update still move it.  The
 * rank is taken by id within the cell, which is the reference's insertion
 * order when its pointer array is in id order (a single process:
 * src/geometry.c:56 sets id = array index and nothing reorders the array;
 * the multi-rank exchange renumbers at src/communication.c:240, 428).  cap = 0 (the default) turns the cap off; SPH_REF_MAX_BUCKET is
 * the reference's value.  Applies to every scene of the context and to the
 * solver's walks (viscosity, density, relaxation, sph_read_neighbors); the
 * surrogate's learned step does not apply it.  A change applies at once:
 * the next viscosity pass walks the last hash's sets under the new cap (the
 * reference's cap never changes).  The 400-neighbour list cap is not
 * reproduced. */
int sph_set_bucket_cap(sph_ctx *ctx, int cap);

/* Launch shape, with no counterpart in the reference: lanes per particle of
 * the fused step's gathers.  0 (the default) chooses by size - 4 when the
 * context has at most two 256-slot tiles per CU of the device (a scene of
 * up to 131,072 particles on 256 CUs), where a launch's time is set by its
 * slowest workgroups, else 1; 1 and 4 force a shape.  Results are bit
 * for bit the same either way: four lanes split a particle's candidate tests
 * and pair terms and apply the terms in the same order.  Any other value is
 * SPH_EINVAL. */
int sph_set_gather_lanes(sph_ctx *ctx, int lanes);
/* Sloshing clock offset of one scene: its forcing is evaluated at
 * t = (steps since the last upload or load + steps) * dt, so a scene loaded
 * mid-trajectory keeps its phase. */
int sph_set_scene_time_offset(sph_ctx *ctx, int scene, unsigned steps);

/* State snapshots (no counterpart in the reference; used to re-take
 * macro-steps).  save copies the whole device state and the step clock;
 * restore copies back the scenes whose flag is set, and the clock.  Both
 * need a hashed state (after sph_step or sph_phase_hash). */
int sph_snapshot_save(sph_ctx *ctx);
int sph_snapshot_restore(sph_ctx *ctx, const int *scene_flags);

/* Device-resident state exchange (no counterpart in the reference): xv is a
 * DEVICE pointer to n_total float4 records (x, y, v_x, v_y) in caller order
 * (scene offset + id), 16-byte aligned.  load replaces positions and
 * velocities (x_prev = x, ids renumbered in storage order, clock reset, the
 * state unhashed: call sph_phase_hash before stepping to keep viscosity on). */
int sph_gather_state_dev(sph_ctx *ctx, float *xv_by_id);
int sph_load_state_dev(sph_ctx *ctx, const float *xv_by_id);

/* Upload n particles (AoS, reference layout).  Particle ids must be a
 * permutation of 0..n-1; id k is the particle's pointer-array index in the
 * reference.  Marks the neighbour lists invalid, as a fresh reference run
 * has them (src/fluid.c:202 calloc): the next step's viscosity is a no-op. */
int sph_upload(sph_ctx *ctx, const fluid_particle *particles, int n);
/* Same, from the reference's pointer array (src/fluid.c:197). */
int sph_upload_pointers(sph_ctx *ctx, fluid_particle *const *fluid_particle_pointers, int n);

/* Replace the tunables used by subsequent steps (the per-frame
 * MPI_Scatterv of src/fluid.c:293-294).  Never triggers graph re-capture. */
int sph_set_tunables(sph_ctx *ctx, const tunable_parameters *tunables);

/* Run nsteps full steps (src/fluid.c:273-348 with MPI removed) through the
 * captured HIP graph.  Returns once the work is enqueued; see sph_sync. */
int sph_step(sph_ctx *ctx, int nsteps);
/* One full step whose first half (gravity, viscosity, predict:
 * src/fluid.c:273-279) uses `first` and whose second half (hash, density,
 * relaxation, velocity update, hash: src/fluid.c:310-341) uses `second`,
 * applied to every scene.  This is the step on which the reference loop
 * receives new tunables from its render rank (src/fluid.c:293-294, between
 * predict_positions and hash_fluid).  Afterwards the context's tunables are
 * `second`.  smoothing_radius cannot change. */
int sph_step_tunables2(sph_ctx *ctx, const tunable_parameters *first,
                       const tunable_parameters *second);
/* Same, launching kernels one by one instead of replaying the graph. */
int sph_step_eager(sph_ctx *ctx, int nsteps);
int sph_sync(sph_ctx *ctx);

/* Copy the state back into reference-layout AoS, indexed by id:
 * out[id] for id in 0..n-1.  Fills every field the reference's loop leaves
 * behind after a step (x_prev/density/pressure from that step). */
int sph_download(sph_ctx *ctx, fluid_particle *out);
/* Positions only: xy_by_id[2*id] = x, xy_by_id[2*id + 1] = y (caller order
 * for batches: scene offset + id).  One 8-byte record per particle crosses
 * to the host, the fields the reference's loop reads between steps
 * (src/fluid.c:356-363). */
int sph_download_positions(sph_ctx *ctx, float *xy_by_id);
/* Pack coordinates as the reference does for its renderer
 * (src/fluid.c:356-363): coords[2*id] = (2x/max_x - 1)*SHRT_MAX, same for y. */
int sph_pack_coords_i16(sph_ctx *ctx, short *coords);

int sph_get_counters(sph_ctx *ctx, sph_counters_t *out);
int sph_reset_counters(sph_ctx *ctx);
int sph_num_particles(const sph_ctx *ctx);
const char *sph_last_error(void);

/* Phase-level calls.  A full step is exactly
 *   gravity, viscosity, predict, hash(1), relax, update_velocities, hash(0)
 * matching src/fluid.c:273-341.  The fused step computes the same values
 * with fewer kernels. */
int sph_phase_gravity(sph_ctx *ctx);            /* src/fluid.c:398-413 */
int sph_phase_viscosity(sph_ctx *ctx);          /* src/fluid.c:416-478 */
int sph_phase_predict(sph_ctx *ctx);            /* src/fluid.c:507-523 */
int sph_phase_hash(sph_ctx *ctx, int compute_density); /* src/hash.c:127-242 */
int sph_phase_relax(sph_ctx *ctx);              /* src/fluid.c:541-611 */
int sph_phase_update_velocities(sph_ctx *ctx);  /* src/fluid.c:642-653 */

/* Read-back of the device state in its current (cell, id) sorted order,
 * for tests and tools. field: one of SPH_FIELD_*; out has n elements
 * (float, or int for SPH_FIELD_ID / unsigned for SPH_FIELD_KEY). */
#define SPH_FIELD_X 0
#define SPH_FIELD_Y 1
#define SPH_FIELD_VX 2
#define SPH_FIELD_VY 3
#define SPH_FIELD_XPREV 4
#define SPH_FIELD_YPREV 5
#define SPH_FIELD_DENSITY 6
#define SPH_FIELD_DENSITY_NEAR 7
#define SPH_FIELD_ID 8
#define SPH_FIELD_KEY 9
int sph_read_sorted(sph_ctx *ctx, int field, void *out);
/* Cell table of the last hash: start[c], count[c] for c < cells. */
int sph_read_cells(sph_ctx *ctx, unsigned int *start, unsigned int *count);
int sph_num_cells(const sph_ctx *ctx, unsigned int *size_x, unsigned int *size_y);
/* Neighbour sets of the last hash, by caller index (scene offset + id), for
 * tests and tools: full_by_id[i] = number of particles within h of i (self
 * excluded; the reference's pair test r*r <= h*h, src/hash.c:99, 185, 221);
 * half_by_id[i] = length of i's half list in the reference
 * (src/hash.c:168-240: same-cell partners of higher id and partners in the
 * forward cells); ids_by_id[i*stride + k] = id of i's k-th neighbour in the
 * canonical walk order, for k < min(full, stride).  The context must have
 * been hashed since the last upload. */
int sph_read_neighbors(sph_ctx *ctx, int stride, int *full_by_id, int *half_by_id,
                       int *ids_by_id);

/* Per-kernel timing: enqueue nsteps full steps kernel by kernel, with
 * pre-created HIP events recorded on the context stream before every kernel
 * and after each step's last one, then synchronise once and return the mean
 * duration (ms) of each kernel kind.  The steps advance the simulation.
 * names/ms have room for max_kinds entries; *nkinds receives the count. */
int sph_time_kernels(sph_ctx *ctx, int nsteps, int max_kinds, const char **names,
                     float *ms, int *nkinds);

/* Host-side scene builder: the reference's lattice (src/geometry.c:29-67,
 * row-major, y outer, id = index) of nx*ny particles at `spacing`, lower
 * left corner (x0, y0), plus a seeded uniform jitter of +-jitter per
 * coordinate (SplitMix64, see scene.c).  v = a = 0.  out has nx*ny slots. */
int sph_scene_lattice(fluid_particle *out, int nx, int ny, float x0, float y0, float spacing,
                      float jitter, uint64_t seed);

/* Self-check of the kernels' correctly rounded square root (no counterpart
 * in the reference): counts the float bit patterns in [lo_bits, hi_bits)
 * whose sqrt differs from IEEE sqrtf (all NaNs count as equal). */
int sph_selftest_sqrt(unsigned lo_bits, unsigned hi_bits, unsigned long long *mismatches);

/* The same for the kernels' correctly rounded reciprocal against IEEE
 * division 1.0f / x over its domain (+-0, +-inf, NaN and 2^-126 <= |x| <=
 * 2^125; patterns outside it are skipped); no counterpart in the reference. */
int sph_selftest_rcp(unsigned lo_bits, unsigned hi_bits, unsigned long long *mismatches);

/* ------------------------------------------------------------------------
 * 3. Reference-named entry points (drop-in for src/fluid.h:112-126 and
 *    src/hash.h:50-52).  They drive one process-global context that is
 *    created from the first call's arguments and uploads lazily from the
 *    caller's pointer array.  After hash_fluid(..., false) (the last call
 *    of a step, src/fluid.c:341) the state is copied back into the caller's
 *    particles every sph_shim_set_sync_interval() steps (default 1), so the
 *    caller's coordinate packing (src/fluid.c:356-363) reads current data.
 *    neighbor / neighbor_grid_t list storage is accepted and left untouched:
 *    neighbour sets live on the device.
 * ---------------------------------------------------------------------- */
void boundaryConditions(fluid_particle *p, AABB_t *boundary, param *params);
void initParticles(fluid_particle **fluid_particle_pointers, fluid_particle *fluid_particles,
                   AABB_t *water, int start_x, int number_particles_x, edge_t *edges,
                   int max_fluid_particles_local, float spacing, param *params);
void calculate_density(fluid_particle *p, fluid_particle *q, float ratio);
void apply_gravity(fluid_particle **fluid_particle_pointers, param *params);
void viscosity_impluses(fluid_particle **fluid_particle_pointers, neighbor *neighbors,
                        param *params);
void predict_positions(fluid_particle **fluid_particle_pointers, AABB_t *boundary_global,
                       param *params);
void double_density_relaxation(fluid_particle **fluid_particle_pointers, neighbor *neighbors,
                               param *params);
void updateVelocity(fluid_particle *p, param *params);
void updateVelocities(fluid_particle **fluid_particle_pointers, edge_t *edges,
                      AABB_t *boundary_global, param *params);
void checkVelocity(float *v_x, float *v_y);
void identify_oob_particles(fluid_particle **fluid_particle_pointers,
                            fluid_particle *fluid_particles, oob_t *out_of_bounds,
                            AABB_t *boundary_global, param *params);
unsigned int hash_val(float x, float y, neighbor_grid_t *grid, param *params);
void hash_fluid(fluid_particle **fluid_particle_pointers, neighbor_grid_t *grid, param *params,
                bool compute_density);
void hash_halo(fluid_particle **fluid_particle_pointers, neighbor_grid_t *grid, param *params,
               bool compute_density);

/* Shim controls (no reference counterpart). */
int sph_shim_set_sync_interval(int steps); /* 0 = never copy back automatically */
/* Which fields the copy back writes into the caller's particles:
 * SPH_SYNC_ALL (default) or SPH_SYNC_POSITIONS (x and y only, the fields the
 * reference's loop reads between steps, src/fluid.c:356-363). */
#define SPH_SYNC_ALL 0
#define SPH_SYNC_POSITIONS 1
int sph_shim_set_sync_fields(int fields);
int sph_shim_sync(void);                   /* copy device state to the caller now */
/* The caller changed its particles (or their number, up to any size): the
 * next call re-uploads them and hashes them, keeping viscosity on. */
int sph_shim_mark_host_dirty(void);
void sph_shim_release(void); /* destroy the process-global context, reset the controls */
sph_ctx *sph_shim_context(void);

/* ------------------------------------------------------------------------
 * 4. The learned macro-step (no counterpart in the reference).
 *
 * A particle-interaction network predicts each particle's displacement over
 * four plain steps from its neighbourhood (all particles within h):
 *   edge features f = (dx/h, dy/h, dv_x/5, dv_y/5, r/h)
 *   edge MLP  5 -> 64 -> 64 (SiLU after the first layer), sum over edges,
 *   node MLP  [sum, v/5, a/10] (68) -> 64 -> 2 (SiLU),
 *   d = 4 dt v + 10 dt^2 a + 0.05 out          (a: sloshing and gravity).
 * BF16 weights and activations on MFMA with FP32 accumulation; FP32 master
 * weights and Adam.  The parameter vector (sph_surrogate_num_params floats)
 * is laid out as W1[64][16] (columns 0-4 features, 5 bias, rest zero),
 * W2[64][64], b2[64], W3[64][80] (columns 0-63 sum, 64-67 node features,
 * 68 bias, rest zero), W4[2][64], b4[2].
 * Arrays "by id" are in caller order (scene offset + id), 2 floats each.
 * ---------------------------------------------------------------------- */
typedef struct sph_surrogate sph_surrogate;

const char *sph_surrogate_last_error(void);
int sph_surrogate_num_params(void);
int sph_surrogate_create(sph_surrogate **m, uint64_t seed);
void sph_surrogate_destroy(sph_surrogate *m);
int sph_surrogate_get_params(sph_surrogate *m, float *out);
int sph_surrogate_set_params(sph_surrogate *m, const float *in);
int sph_surrogate_get_grads(sph_surrogate *m, float *out);
int sph_surrogate_zero_grads(sph_surrogate *m);
/* Forward on the context's hashed state; disp_by_id (host, optional)
 * receives the predicted displacements. */
int sph_surrogate_predict(sph_surrogate *m, sph_ctx *ctx, float *disp_by_id);
/* Forward + loss + backward on the hashed state against target
 * displacements target_by_id (host, caller order).  Gradients accumulate;
 * loss (host, optional) = mean |d - target|^2 / 0.05^2. */
int sph_surrogate_backward(sph_surrogate *m, sph_ctx *ctx, const float *target_by_id,
                           float *loss);
/* Adam step on the accumulated gradients (which it then clears). */
int sph_surrogate_adam(sph_surrogate *m, float lr, float beta1, float beta2, float eps);
/* Move every unfrozen scene by the last prediction (x += d with the box
 * clamp, v = d / (4 dt) + 1.5 a dt with the velocity clamp); the state is
 * unhashed afterwards and the clock advances by four steps. */
int sph_surrogate_apply(sph_surrogate *m, sph_ctx *ctx);
/* One judged macro-step of every scene (see sph_mlp.hip); flags (optional,
 * nscenes ints) marks the scenes re-taken by four plain steps: 1 judged
 * off the plain solver's density trend by more than tol, 2 the scheduled
 * plain macro-step (the first one, then one in 16), 3 off the plain shadow
 * (below) by more than tol. */
int sph_surrogate_macro_step(sph_surrogate *m, sph_ctx *ctx, float tol, int *flags,
                             int *nfallback);
/* Shadow judging: on each macro-step one scene in every `every` (rotating;
 * 0 = off, the default) first takes the macro-step plainly from the same
 * start.  Its surrogate result is then also judged against that plain
 * result, and for the steps the trend judge accepted the error against it
 * is recorded.  Resets the statistics. */
int sph_surrogate_set_shadow(sph_surrogate *m, int every);
/* Statistics since sph_surrogate_set_shadow: out[0] shadowed scene
 * macro-steps, out[1] of them accepted by the trend judge, out[2] of those
 * with |e| > tol, out[3] sum |e|, out[4] max |e|, out[5] sum e, out[6] sum
 * and out[7] max of the per-particle relative RMS density error, where e is
 * the relative error of the scene's mean density against the plain
 * result. */
int sph_surrogate_shadow_stats(sph_surrogate *m, double *out);
/* D = A B of one 16x16x16 BF16 MFMA with fixed small-integer operands,
 * row-major (checks the lane layout the kernels assume). */
int sph_surrogate_mfma_selftest(float *out256);
/* Per-scene mean and maximum density of the last density pass. */
int sph_scene_density(sph_ctx *ctx, float *mean, float *maxv);
/* The same without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory - code under HIP graph capture,
 * serving loops or harnesses that forbid hipMalloc.  The per-scene partial
 * results use a fixed 32 KiB device array per GPU (at most 4096 scenes; more
 * is rejected with SPH_EINVAL).  Calls that overlap on one GPU race; the
 * context's device must be current. */
int sph_scene_density_noalloc(sph_ctx *ctx, float *mean, float *maxv);

/* Trajectory store for training, resident in device memory: frames of the
 * state (x, y, v_x, v_y per particle) of nscenes scenes of n particles.
 * record stores a context's state (nscenes x n) as frame f; load puts
 * count samples (scene[k] at frame[k]) into the count scenes of a training
 * context (n particles each; x_prev = x, the state unhashed) and keeps their
 * 4-step targets (frame + 1 minus frame positions) for
 * sph_surrogate_backward_trajectory; read copies a frame to the host. */
typedef struct sph_trajectory sph_trajectory;
int sph_trajectory_create(sph_trajectory **tr, int nscenes, int n, int frames);
void sph_trajectory_destroy(sph_trajectory *tr);
int sph_trajectory_record(sph_trajectory *tr, sph_ctx *ctx, int frame);
int sph_trajectory_read(sph_trajectory *tr, int frame, float *out);
int sph_trajectory_load(sph_trajectory *tr, sph_ctx *ctx, int count, const int *scene,
                        const int *frame);
int sph_surrogate_backward_trajectory(sph_surrogate *m, sph_ctx *ctx, sph_trajectory *tr,
                                      float *loss);

#ifdef __cplusplus
}
#endif

#endif /* SPH_ABI_H */
