/*
 * ref_harness.c - drives the reference's own compiled step functions
 * (src/hash.c, src/fluid.c, src/geometry.c, src/communication.c of
 * AdamSimpson/SPH, built unmodified from /root/reference by oracle/Makefile)
 * on one domain, for golden fixtures and the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, never into the product.
 *
 * It owns the buffers that start_simulation allocates (src/fluid.c:181-233):
 * the AoS particle array, the pointer array (id = index), the per-particle
 * half-neighbour lists (max_bucket_size*4 slots, calloc'd so the first
 * viscosity pass sees empty lists) and the bucket grid
 * (ceil(extent/h) cells per axis, max_bucket_size slots each), and calls the
 * loop body of src/fluid.c:273-348 without its MPI lines.
 * identify_oob_particles is not called: with one compute rank the node
 * bounds are the box (src/geometry.c:144-147), every particle is clamped
 * into the box before it (src/fluid.c:279), so nothing leaves and the
 * pointer array is already compact with id = index.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fluid.h"
#include "geometry.h"
#include "hash.h"

typedef struct {
    int n;
    param params;
    AABB_t box;
    edge_t edges;
    neighbor_grid_t grid;
    fluid_particle *parts;
    fluid_particle **ptrs;
    neighbor *neighbors;
    fluid_particle **list_store;
    bucket_t *buckets;
    fluid_particle **bucket_store;
} refh;

refh *refh_create(const AABB_t *box, const tunable_parameters *tp, int n, unsigned max_bucket)
{
    refh *s = calloc(1, sizeof(refh));
    if (!s)
        return NULL;
    s->n = n;
    s->box = *box;
    s->params.tunable_params = *tp;
    s->params.tunable_params.node_start_x = box->min_x;
    s->params.tunable_params.node_end_x = box->max_x;
    s->params.number_fluid_particles_global = n;
    s->params.number_fluid_particles_local = n;
    s->params.max_fluid_particle_index = n - 1;
    s->params.number_halo_particles = 0;

    s->grid.max_bucket_size = max_bucket;
    s->grid.max_neighbors = max_bucket * 4;
    s->grid.spacing = tp->smoothing_radius;
    s->grid.size_x = ceil((box->max_x - box->min_x) / s->grid.spacing);
    s->grid.size_y = ceil((box->max_y - box->min_y) / s->grid.spacing);
    size_t cells = (size_t)s->grid.size_x * s->grid.size_y;

    s->parts = calloc((size_t)n, sizeof(fluid_particle));
    s->ptrs = calloc((size_t)n, sizeof(fluid_particle *));
    s->neighbors = calloc((size_t)n, sizeof(neighbor));
    s->list_store = calloc((size_t)n * s->grid.max_neighbors, sizeof(fluid_particle *));
    s->buckets = calloc(cells, sizeof(bucket_t));
    s->bucket_store = calloc(cells * max_bucket, sizeof(fluid_particle *));
    if (!s->parts || !s->ptrs || !s->neighbors || !s->list_store || !s->buckets ||
        !s->bucket_store) {
        free(s->parts);
        free(s->ptrs);
        free(s->neighbors);
        free(s->list_store);
        free(s->buckets);
        free(s->bucket_store);
        free(s);
        return NULL;
    }
    for (int i = 0; i < n; i++) {
        s->ptrs[i] = &s->parts[i];
        s->neighbors[i].fluid_neighbors = &s->list_store[(size_t)i * s->grid.max_neighbors];
    }
    for (size_t c = 0; c < cells; c++)
        s->buckets[c].fluid_particles = &s->bucket_store[c * max_bucket];
    s->grid.neighbors = s->neighbors;
    s->grid.grid_buckets = s->buckets;
    return s;
}

void refh_destroy(refh *s)
{
    if (!s)
        return;
    free(s->parts);
    free(s->ptrs);
    free(s->neighbors);
    free(s->list_store);
    free(s->buckets);
    free(s->bucket_store);
    free(s);
}

void refh_set_tunables(refh *s, const tunable_parameters *tp)
{
    float a = s->params.tunable_params.node_start_x, b = s->params.tunable_params.node_end_x;
    s->params.tunable_params = *tp;
    s->params.tunable_params.node_start_x = a;
    s->params.tunable_params.node_end_x = b;
}

/* Particle i of `in` must carry id i (the reference's pointer index). */
int refh_load(refh *s, const fluid_particle *in)
{
    for (int i = 0; i < s->n; i++)
        if (in[i].id != i)
            return -1;
    memcpy(s->parts, in, sizeof(fluid_particle) * (size_t)s->n);
    return 0;
}

void refh_store(const refh *s, fluid_particle *out)
{
    memcpy(out, s->parts, sizeof(fluid_particle) * (size_t)s->n);
}

/* Phase ids match the oracle's: 0 gravity, 1 viscosity, 2 predict,
 * 3 hash with density, 4 relaxation, 5 velocity update, 6 hash. */
int refh_phase(refh *s, int phase)
{
    switch (phase) {
    case 0: apply_gravity(s->ptrs, &s->params); break;
    case 1: viscosity_impluses(s->ptrs, s->neighbors, &s->params); break;
    case 2: predict_positions(s->ptrs, &s->box, &s->params); break;
    case 3:
        hash_fluid(s->ptrs, &s->grid, &s->params, true);
        hash_halo(s->ptrs, &s->grid, &s->params, true);
        break;
    case 4: double_density_relaxation(s->ptrs, s->neighbors, &s->params); break;
    case 5: updateVelocities(s->ptrs, &s->edges, &s->box, &s->params); break;
    case 6:
        hash_fluid(s->ptrs, &s->grid, &s->params, false);
        hash_halo(s->ptrs, &s->grid, &s->params, false);
        break;
    default: return -1;
    }
    return 0;
}

void refh_step(refh *s, int nsteps)
{
    for (int it = 0; it < nsteps; it++)
        for (int ph = 0; ph <= 6; ph++)
            refh_phase(s, ph);
}

/* Half lists of the last hash as ids: counts[i], lists[i*stride + j]. */
void refh_lists(const refh *s, int *counts, int *lists, int stride)
{
    for (int i = 0; i < s->n; i++) {
        const neighbor *ne = &s->neighbors[i];
        counts[i] = ne->number_fluid_neighbors;
        if (lists)
            for (int j = 0; j < ne->number_fluid_neighbors && j < stride; j++)
                lists[(size_t)i * stride + j] = ne->fluid_neighbors[j]->id;
    }
}

/* Bucket occupancy of the last hash (the reference stops at max_bucket). */
void refh_bucket_counts(const refh *s, unsigned *counts)
{
    size_t cells = (size_t)s->grid.size_x * s->grid.size_y;
    for (size_t c = 0; c < cells; c++)
        counts[c] = s->buckets[c].number_fluid;
}
