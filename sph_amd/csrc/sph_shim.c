/*
 * sph_shim.c - the reference's step-function API (src/fluid.h:112-126,
 * src/hash.h:50-52) implemented on the device context, so the reference's
 * main loop (src/fluid.c:270-372) links against libsph_hip.so in place of
 * src/hash.c and the step functions of src/fluid.c.
 *
 * One process-global context serves the calls, as the reference keeps one
 * simulation per process.  It is created by the first call that carries the
 * box (predict_positions), with room for the max_fluid_particles_local that
 * initParticles was given (the reference sizes it at twice the global count,
 * src/fluid.c:156), uploading the caller's particles through the pointer
 * array; calls before that (the first apply_gravity of a run) run on the
 * host with the reference's arithmetic, and the first viscosity pass is a
 * no-op exactly as in the reference, whose lists start empty.
 *
 * The step's calls are recorded, not run one by one: when a step's calls
 * arrive in the reference's order (apply_gravity, viscosity_impluses,
 * predict_positions, hash_fluid(true), double_density_relaxation,
 * updateVelocities, hash_fluid(false); identify_oob_particles and hash_halo
 * are no-ops on one domain), the last one runs the whole step as one fused,
 * graph-replayed step - with the tunables of the first three calls for the
 * first half and those of the last four for the second, because the
 * reference receives new tunables between predict_positions and hash_fluid
 * (src/fluid.c:293-294).  Any other order, or a call that needs the device
 * state mid-step, replays the recorded calls phase by phase (the same
 * values) and runs the rest of that step the same way.
 *
 * hash_fluid(..., false) ends a step (src/fluid.c:341); every sync_interval
 * steps it copies the device state back into the caller's particles, so
 * code that reads p->x afterwards (src/fluid.c:356-363) sees the current
 * state: all fields, or only x and y (sph_shim_set_sync_fields), which is
 * all the reference's loop reads.  neighbor / bucket storage is left
 * untouched: the neighbour sets live on the device.  The grid's
 * max_bucket_size is applied as the reference's bucket cap
 * (sph_set_bucket_cap); 0 runs uncapped.
 *
 * After sph_shim_mark_host_dirty() the next call re-uploads the caller's
 * particles (growing the context when their number exceeds its capacity)
 * and hashes them, so the next viscosity pass has neighbour lists, as the
 * reference keeps its lists across such a change.  initParticles starts a
 * new run: its upload leaves the lists empty.
 *
 * Errors keep the reference's void returns; the message is available from
 * sph_last_error() and is also printed once to stderr.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sph_abi.h"

enum { DIRTY_NONE = 0, DIRTY_KEEP_LISTS = 1, DIRTY_NEW_RUN = 2 };
/* The recorded calls of a step, in the reference's order. */
enum { CALL_GRAVITY = 1, CALL_VISCOSITY, CALL_PREDICT, CALL_HASH_DENSITY, CALL_RELAX,
       CALL_UPDATE_VELOCITIES, CALL_HASH };

static sph_ctx *g_ctx;
static fluid_particle **g_ptrs;
static int g_n;
static int g_capacity; /* max_fluid_particles_local given to initParticles */
static AABB_t g_box;   /* the box the context was created with */
static int g_cap_set = -1;  /* bucket cap last given to the context (-1: none yet) */
static int g_grid_cap = 0;  /* the cap of the last grid record hash_fluid received */
/* Do the caller's pointers address one array in order (ptrs[i] == ptrs[0] + i,
 * as the reference's initParticles lays them out)?  Then the state is copied
 * back into it in one piece. */
static int g_contiguous;
static int g_dirty = DIRTY_NEW_RUN;
static int g_sync_interval = 1;
static int g_sync_fields = SPH_SYNC_ALL;
static int g_steps_since_sync;
static int g_reported;
/* calls of the current step recorded so far (0: none; -1: this step runs
 * phase by phase), and the tunables of its two halves */
static int g_seq;
static tunable_parameters g_first, g_second;

static void shim_fail(const char *where, int rc)
{
    if (rc && !g_reported) {
        fprintf(stderr, "libsph_hip: %s failed (%d): %s\n", where, rc, sph_last_error());
        g_reported = 1;
    }
}

/* Make the context hold the caller's particles: create it (needs the box),
 * grow it, or re-upload, as needed. */
static int shim_upload(fluid_particle **ptrs, param *params, AABB_t *box)
{
    int n = params->number_fluid_particles_local;
    int rc = SPH_OK;
    if (g_ctx && n > g_capacity) { /* more particles than the context holds: recreate it */
        sph_destroy(g_ctx);
        g_ctx = NULL;
        if (!box)
            box = &g_box;
    }
    if (!g_ctx) {
        if (!box)
            return SPH_ESTATE;
        int cap = g_capacity > n ? g_capacity : n;
        AABB_t b = *box;
        rc = sph_create(&g_ctx, params, &b, cap);
        if (rc)
            return rc;
        /* a context made again (growth) keeps the run's cap from its first pass */
        g_cap_set = -1;
        if (g_grid_cap) {
            rc = sph_set_bucket_cap(g_ctx, g_grid_cap);
            if (rc)
                return rc;
            g_cap_set = g_grid_cap;
        }
        g_capacity = cap;
        g_box = b;
        if (g_dirty == DIRTY_NONE)
            g_dirty = DIRTY_KEEP_LISTS;
    }
    if (n != g_n || ptrs != g_ptrs)
        if (g_dirty == DIRTY_NONE)
            g_dirty = DIRTY_KEEP_LISTS;
    if (g_dirty != DIRTY_NONE) {
        int keep = g_dirty == DIRTY_KEEP_LISTS;
        rc = sph_set_tunables(g_ctx, &params->tunable_params);
        if (!rc)
            rc = sph_upload_pointers(g_ctx, ptrs, n);
        if (!rc && keep) /* lists of the uploaded positions for the next viscosity pass */
            rc = sph_phase_hash(g_ctx, 0);
        if (rc)
            return rc;
        g_ptrs = ptrs;
        g_n = n;
        g_dirty = DIRTY_NONE;
        g_steps_since_sync = 0;
        g_contiguous = 1;
        for (int i = 1; i < n && g_contiguous; i++)
            g_contiguous = ptrs[i] == ptrs[0] + i;
    }
    return SPH_OK;
}

/* Run recorded call `call` as its phase, with tunables t. */
static int run_phase(int call, const tunable_parameters *t)
{
    int rc = sph_set_tunables(g_ctx, t);
    if (rc)
        return rc;
    switch (call) {
    case CALL_GRAVITY: return sph_phase_gravity(g_ctx);
    case CALL_VISCOSITY: return sph_phase_viscosity(g_ctx);
    case CALL_PREDICT: return sph_phase_predict(g_ctx);
    case CALL_HASH_DENSITY: return sph_phase_hash(g_ctx, 1);
    case CALL_RELAX: return sph_phase_relax(g_ctx);
    case CALL_UPDATE_VELOCITIES: return sph_phase_update_velocities(g_ctx);
    case CALL_HASH: return sph_phase_hash(g_ctx, 0);
    default: return SPH_EINVAL;
    }
}

/* Run the calls recorded so far phase by phase; the rest of the step
 * follows the same way. */
static int shim_flush(void)
{
    int rc = SPH_OK;
    for (int call = CALL_GRAVITY; call <= g_seq && !rc; call++)
        rc = run_phase(call, call <= CALL_PREDICT ? &g_first : &g_second);
    g_seq = -1;
    return rc;
}

/* A step function's entry: make the context current, then either record
 * the call (returns 1 in *deferred) or leave it to run as a phase. */
static int shim_enter(int call, fluid_particle **ptrs, param *params, AABB_t *box, int *deferred)
{
    *deferred = 0;
    if (g_seq > 0 && g_dirty != DIRTY_NONE) { /* the host changed data mid-step */
        int rc = shim_flush();
        if (rc)
            return rc;
    }
    int rc = shim_upload(ptrs, params, box);
    if (rc)
        return rc;
    const tunable_parameters *t = &params->tunable_params;
    if (call == CALL_GRAVITY) {
        if (g_seq > 0 && (rc = shim_flush()))
            return rc;
        g_seq = 0;
    }
    if (g_seq < 0)
        return sph_set_tunables(g_ctx, t);
    int same = call <= CALL_PREDICT ? call == CALL_GRAVITY || !memcmp(t, &g_first, sizeof *t)
                                    : call == CALL_HASH_DENSITY || !memcmp(t, &g_second, sizeof *t);
    if (call != g_seq + 1 || !same) {
        if ((rc = shim_flush()))
            return rc;
        return sph_set_tunables(g_ctx, t);
    }
    if (call == CALL_GRAVITY)
        g_first = *t;
    if (call == CALL_HASH_DENSITY)
        g_second = *t;
    g_seq = call;
    *deferred = 1;
    return SPH_OK;
}

/* Host buffer reused by every copy back (grown as needed). */
static void *g_stage;
static size_t g_stage_bytes;

static void *stage(size_t bytes)
{
    if (bytes > g_stage_bytes) {
        void *p = realloc(g_stage, bytes);
        if (!p)
            return NULL;
        g_stage = p;
        g_stage_bytes = bytes;
    }
    return g_stage;
}

static int shim_sync_now(void)
{
    if (!g_ctx || !g_ptrs || g_n == 0)
        return SPH_OK;
    int rc = SPH_OK;
    size_t n = (size_t)g_n;
    if (g_sync_fields == SPH_SYNC_POSITIONS) {
        float *xy = (float *)stage(sizeof(float) * 2 * n);
        if (!xy)
            return SPH_ENOMEM;
        rc = sph_download_positions(g_ctx, xy);
        if (!rc)
            for (size_t i = 0; i < n; i++) { /* pointer index == id on one domain */
                g_ptrs[i]->x = xy[2 * i];
                g_ptrs[i]->y = xy[2 * i + 1];
            }
    } else if (g_contiguous) {
        rc = sph_download(g_ctx, g_ptrs[0]);
    } else {
        fluid_particle *tmp = (fluid_particle *)stage(sizeof(fluid_particle) * n);
        if (!tmp)
            return SPH_ENOMEM;
        rc = sph_download(g_ctx, tmp);
        if (!rc)
            for (size_t i = 0; i < n; i++)
                *g_ptrs[i] = tmp[i];
    }
    g_steps_since_sync = 0;
    return rc;
}

int sph_shim_sync(void)
{
    if (g_ctx && g_seq > 0) { /* recorded calls first: the host sees the state they leave */
        int rc = shim_flush();
        if (rc)
            return rc;
    }
    return shim_sync_now();
}

int sph_shim_set_sync_interval(int steps)
{
    if (steps < 0)
        return SPH_EINVAL;
    g_sync_interval = steps;
    return SPH_OK;
}

int sph_shim_set_sync_fields(int fields)
{
    if (fields != SPH_SYNC_ALL && fields != SPH_SYNC_POSITIONS)
        return SPH_EINVAL;
    g_sync_fields = fields;
    return SPH_OK;
}

int sph_shim_mark_host_dirty(void)
{
    if (g_dirty == DIRTY_NONE)
        g_dirty = DIRTY_KEEP_LISTS;
    return SPH_OK;
}

void sph_shim_release(void)
{
    sph_destroy(g_ctx);
    g_ctx = NULL;
    g_ptrs = NULL;
    g_n = 0;
    g_capacity = 0;
    g_dirty = DIRTY_NEW_RUN;
    g_seq = 0;
    g_steps_since_sync = 0;
    g_sync_interval = 1;
    g_sync_fields = SPH_SYNC_ALL;
    g_reported = 0;
    g_grid_cap = 0;
    g_cap_set = -1;
    free(g_stage);
    g_stage = NULL;
    g_stage_bytes = 0;
}

sph_ctx *sph_shim_context(void) { return g_ctx; }

/* ---- host helpers with the reference's arithmetic ---------------------- */

/* src/fluid.c:613-625 */
void checkVelocity(float *v_x, float *v_y)
{
    const float v_max = 5.0f;
    if (*v_x > v_max)
        *v_x = v_max;
    else if (*v_x < -v_max)
        *v_x = -v_max;
    if (*v_y > v_max)
        *v_y = v_max;
    else if (*v_y < -v_max)
        *v_y = -v_max;
}

/* src/fluid.c:656-744 */
void boundaryConditions(fluid_particle *p, AABB_t *boundary, param *params)
{
    const tunable_parameters *t = &params->tunable_params;
    float cx = t->mover_center_x, cy = t->mover_center_y;
    if (t->mover_type == SPHERE_MOVER) {
        float radius = t->mover_width * 0.5f;
        float d2 = (p->x - cx) * (p->x - cx) + (p->y - cy) * (p->y - cy);
        if (d2 <= radius * radius && d2 > 0.0f) {
            float d = sqrtf(d2);
            float nx = (cx - p->x) / d;
            float ny = (cy - p->y) / d;
            float pen = radius - d;
            p->x -= pen * nx;
            p->y -= pen * ny;
        }
    } else if (t->mover_type == RECTANGLE_MOVER) {
        float hw = (float)(t->mover_width * 0.5);
        float hh = (float)(t->mover_height * 0.5);
        float rx = p->x - cx, ry = p->y - cy;
        float ax = fabsf(rx), ay = fabsf(ry);
        if (ax < hw && ay < hh) {
            float penx = hw - ax, peny = hh - ay;
            if (penx < peny)
                p->x += rx < 0.0f ? -penx : penx;
            else
                p->y += ry < 0.0f ? -peny : peny;
        }
    }
    if (p->x < boundary->min_x)
        p->x = boundary->min_x;
    else if (p->x > boundary->max_x)
        p->x = boundary->max_x - 0.001f;
    if (p->y < boundary->min_y)
        p->y = boundary->min_y;
    else if (p->y > boundary->max_y)
        p->y = boundary->max_y - 0.001f;
}

/* src/fluid.c:527-539 */
void calculate_density(fluid_particle *p, fluid_particle *q, float ratio)
{
    float OmR2 = (1.0f - ratio) * (1.0f - ratio);
    if (ratio < 1.0f) {
        p->density += OmR2;
        p->density_near += OmR2 * (1.0f - ratio);
        q->density += OmR2;
        q->density_near += OmR2 * (1.0f - ratio);
    }
}

/* src/fluid.c:627-639 */
void updateVelocity(fluid_particle *p, param *params)
{
    float dt = params->tunable_params.time_step;
    float v_x = (p->x - p->x_prev) / dt;
    float v_y = (p->y - p->y_prev) / dt;
    checkVelocity(&v_x, &v_y);
    p->v_x = v_x;
    p->v_y = v_y;
}

/* src/hash.c:35-47 */
unsigned int hash_val(float x, float y, neighbor_grid_t *grid, param *params)
{
    (void)params;
    unsigned gx = (unsigned)floorf(x / grid->spacing);
    unsigned gy = (unsigned)floorf(y / grid->spacing);
    return gy * grid->size_x + gx;
}

/* src/fluid.c:747-768 with src/geometry.c:29-67 on one domain. */
void initParticles(fluid_particle **fluid_particle_pointers, fluid_particle *fluid_particles,
                   AABB_t *water, int start_x, int number_particles_x, edge_t *edges,
                   int max_fluid_particles_local, float spacing, param *params)
{
    (void)edges; /* halo bookkeeping: MPI only */
    int num_y = (int)floor((water->max_y - water->min_y) / spacing);
    int i = 0;
    for (int ny = 0; ny < num_y; ny++) {
        float y = water->min_y + ny * spacing;
        for (int nx = 0; nx < number_particles_x; nx++) {
            fluid_particle *p = fluid_particles + i;
            p->x = water->min_x + (start_x + nx) * spacing;
            p->y = y;
            fluid_particle_pointers[i] = p;
            p->id = i;
            i++;
        }
    }
    params->number_fluid_particles_local = i;
    params->max_fluid_particle_index = i - 1;
    for (int k = i; k < max_fluid_particles_local; k++)
        fluid_particle_pointers[k] = NULL;
    for (int k = 0; k < i; k++) {
        fluid_particle_pointers[k]->a_x = 0.0f;
        fluid_particle_pointers[k]->a_y = 0.0f;
        fluid_particle_pointers[k]->v_x = 0.0f;
        fluid_particle_pointers[k]->v_y = 0.0f;
    }
    /* a new run: room for max_fluid_particles_local, lists empty */
    if (max_fluid_particles_local > g_capacity) {
        sph_destroy(g_ctx);
        g_ctx = NULL;
        g_capacity = max_fluid_particles_local;
    }
    g_dirty = DIRTY_NEW_RUN;
    g_seq = 0;
}

/* ---- step functions on the device -------------------------------------- */

/* src/fluid.c:398-413 */
void apply_gravity(fluid_particle **fluid_particle_pointers, param *params)
{
    if (!g_ctx) { /* before the context exists: host pass, reference arithmetic */
        float dt = params->tunable_params.time_step, g = -params->tunable_params.g;
        for (int i = 0; i < params->number_fluid_particles_local; i++) {
            fluid_particle *p = fluid_particle_pointers[i];
            p->v_y += g * dt;
            p->density = 0.0f;
            p->density_near = 0.0f;
        }
        g_seq = -1;
        return;
    }
    int deferred;
    int rc = shim_enter(CALL_GRAVITY, fluid_particle_pointers, params, NULL, &deferred);
    if (!rc && !deferred)
        rc = sph_phase_gravity(g_ctx);
    shim_fail("apply_gravity", rc);
}

/* src/fluid.c:416-478 */
void viscosity_impluses(fluid_particle **fluid_particle_pointers, neighbor *neighbors,
                        param *params)
{
    (void)neighbors;
    if (!g_ctx)
        return; /* no hash yet: the reference's lists are empty */
    int deferred;
    int rc = shim_enter(CALL_VISCOSITY, fluid_particle_pointers, params, NULL, &deferred);
    if (!rc && !deferred)
        rc = sph_phase_viscosity(g_ctx);
    shim_fail("viscosity_impluses", rc);
}

/* src/fluid.c:507-523 */
void predict_positions(fluid_particle **fluid_particle_pointers, AABB_t *boundary_global,
                       param *params)
{
    if (!g_ctx)
        g_seq = -1; /* the run's first step: phase by phase */
    int deferred;
    int rc = shim_enter(CALL_PREDICT, fluid_particle_pointers, params, boundary_global, &deferred);
    if (!rc && !deferred)
        rc = sph_phase_predict(g_ctx);
    shim_fail("predict_positions", rc);
}

/* src/fluid.c:481-502: on one domain nothing leaves the box (every
 * particle was clamped into it by predict_positions) and the pointer array
 * stays compact with id = index, so there is nothing to do. */
void identify_oob_particles(fluid_particle **fluid_particle_pointers,
                            fluid_particle *fluid_particles, oob_t *out_of_bounds,
                            AABB_t *boundary_global, param *params)
{
    (void)fluid_particle_pointers;
    (void)fluid_particles;
    (void)out_of_bounds;
    (void)boundary_global;
    (void)params;
}

/* The grid's max_bucket_size is the reference's bucket cap (src/hash.c:137,
 * 153-164; src/fluid.c:174 sets 100): the context applies it from this hash
 * on.  0 leaves the buckets uncapped. */
static int shim_apply_cap(const neighbor_grid_t *grid)
{
    const unsigned v = grid ? grid->max_bucket_size : 0u;
    const int cap = v > 0x7fffffffu ? 0x7fffffff : (int)v;
    g_grid_cap = cap;
    if (!g_ctx || cap == g_cap_set)
        return SPH_OK;
    int rc = sph_set_bucket_cap(g_ctx, cap);
    if (!rc)
        g_cap_set = cap;
    return rc;
}

/* src/hash.c:127-242 */
void hash_fluid(fluid_particle **fluid_particle_pointers, neighbor_grid_t *grid, param *params,
                bool compute_density)
{
    int deferred;
    int call = compute_density ? CALL_HASH_DENSITY : CALL_HASH;
    int rc = shim_enter(call, fluid_particle_pointers, params, NULL, &deferred);
    if (!rc)
        rc = shim_apply_cap(grid);
    if (!rc && deferred && call == CALL_HASH) { /* a whole step recorded: run it fused */
        rc = sph_step_tunables2(g_ctx, &g_first, &g_second);
        g_seq = 0;
    } else if (!rc && !deferred) {
        rc = sph_phase_hash(g_ctx, compute_density ? 1 : 0);
    }
    if (!rc && !compute_density) { /* end of a step (src/fluid.c:341) */
        if (g_seq < 0)
            g_seq = 0;
        ++g_steps_since_sync;
        if (g_sync_interval > 0 && g_steps_since_sync >= g_sync_interval)
            rc = shim_sync_now();
    }
    shim_fail("hash_fluid", rc);
}

/* src/hash.c:51-122: there are no halo particles on one domain. */
void hash_halo(fluid_particle **fluid_particle_pointers, neighbor_grid_t *grid, param *params,
               bool compute_density)
{
    (void)fluid_particle_pointers;
    (void)grid;
    (void)params;
    (void)compute_density;
}

/* src/fluid.c:541-611 */
void double_density_relaxation(fluid_particle **fluid_particle_pointers, neighbor *neighbors,
                               param *params)
{
    (void)neighbors;
    int deferred;
    int rc = shim_enter(CALL_RELAX, fluid_particle_pointers, params, NULL, &deferred);
    if (!rc && !deferred)
        rc = sph_phase_relax(g_ctx);
    shim_fail("double_density_relaxation", rc);
}

/* src/fluid.c:642-653 */
void updateVelocities(fluid_particle **fluid_particle_pointers, edge_t *edges,
                      AABB_t *boundary_global, param *params)
{
    (void)edges;
    int deferred;
    int rc = shim_enter(CALL_UPDATE_VELOCITIES, fluid_particle_pointers, params, boundary_global,
                        &deferred);
    if (!rc && !deferred)
        rc = sph_phase_update_velocities(g_ctx);
    shim_fail("updateVelocities", rc);
}
