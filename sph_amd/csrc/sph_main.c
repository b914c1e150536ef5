/*
 * sph_main.c - a start_simulation-shaped driver (src/fluid.c:71-395 with
 * the MPI render/halo plumbing removed) that calls the reference-named step
 * functions exported by libsph_hip.so, exactly in the reference's order.
 * It demonstrates the drop-in: the loop body below is the reference's loop
 * body (src/fluid.c:273-365) minus its MPI lines.
 *
 *   sph_main [-s N] [-p] [-t] [-u] [-b W,H] nx ny steps [out.bin]
 *   sph_main [-s N] [-p] [-t] [-u] -f frames.bin steps [out.bin]
 *
 * Lattice mode: an nx x ny lattice at spacing 0.5 (h = 1) filling the
 * lower-left of a box sized as the benchmark scenes (60 x 40 for 64 x 64)
 * unless -b gives it.
 * Frame mode (-f): the scene and the render rank's per-frame tunables come
 * from a file (format below).  initParticles builds the fluid as the
 * reference does, and at the last substep of every frame the next record
 * replaces the tunables between predict_positions and hash_fluid, where the
 * reference receives them by MPI_Scatterv (src/fluid.c:293-294).
 *
 *   -s N  copy the state back to the host every N steps (default 1; 0 never)
 *   -p    copy back positions only (the fields the loop reads)
 *   -t    print the loop's wall time per step after its first two steps
 *   -u    uncapped buckets (grid max_bucket_size 0); by default the grid
 *         record carries the reference's 100 (src/fluid.c:174), which the
 *         entry points apply as its bucket cap
 *
 * The final particles are written as raw fluid_particle records to out.bin.
 *
 * Frame file (little-endian): char magic[8] = "SPHFRAME"; int32 version (1);
 * AABB_t box; AABB_t water; param params; float spacing;
 * int32 number_particles_x; int32 max_fluid_particles_local;
 * int32 steps_per_frame; int32 nframes; tunable_parameters frames[nframes].
 */
#define _POSIX_C_SOURCE 200809L
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "../../include/sph_abi.h"

typedef struct {
    AABB_t box, water;
    param params;
    float spacing;
    int number_particles_x, max_local, steps_per_frame, nframes;
    tunable_parameters *frames;
} scene_file;

static int read_frames(const char *path, scene_file *sf)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return -1;
    char magic[8];
    int version = 0, ok = 1;
    ok &= fread(magic, 1, 8, f) == 8 && memcmp(magic, "SPHFRAME", 8) == 0;
    ok &= fread(&version, sizeof version, 1, f) == 1 && version == 1;
    ok &= fread(&sf->box, sizeof sf->box, 1, f) == 1;
    ok &= fread(&sf->water, sizeof sf->water, 1, f) == 1;
    ok &= fread(&sf->params, sizeof sf->params, 1, f) == 1;
    ok &= fread(&sf->spacing, sizeof sf->spacing, 1, f) == 1;
    ok &= fread(&sf->number_particles_x, sizeof(int), 1, f) == 1;
    ok &= fread(&sf->max_local, sizeof(int), 1, f) == 1;
    ok &= fread(&sf->steps_per_frame, sizeof(int), 1, f) == 1;
    ok &= fread(&sf->nframes, sizeof(int), 1, f) == 1;
    ok &= sf->nframes >= 0 && sf->steps_per_frame > 0 && sf->max_local > 0;
    sf->frames = NULL;
    if (ok && sf->nframes > 0) {
        sf->frames = malloc(sizeof(tunable_parameters) * (size_t)sf->nframes);
        ok &= sf->frames != NULL &&
              fread(sf->frames, sizeof(tunable_parameters), (size_t)sf->nframes, f) ==
                  (size_t)sf->nframes;
    }
    fclose(f);
    return ok ? 0 : -1;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void usage(const char *argv0)
{
    fprintf(stderr,
            "usage: %s [-s N] [-p] [-t] [-u] [-b W,H] nx ny steps [out.bin]\n"
            "       %s [-s N] [-p] [-t] [-u] -f frames.bin steps [out.bin]\n",
            argv0, argv0);
}

int main(int argc, char **argv)
{
    int sync_interval = 1, positions_only = 0, timing = 0, uncapped = 0, opt;
    const char *frames_path = NULL;
    float bw = 0.0f, bh = 0.0f;
    while ((opt = getopt(argc, argv, "s:ptub:f:")) != -1) {
        switch (opt) {
        case 's': sync_interval = atoi(optarg); break;
        case 'p': positions_only = 1; break;
        case 't': timing = 1; break;
        case 'u': uncapped = 1; break;
        case 'b':
            if (sscanf(optarg, "%f,%f", &bw, &bh) != 2 || bw <= 0.0f || bh <= 0.0f) {
                usage(argv[0]);
                return 2;
            }
            break;
        case 'f': frames_path = optarg; break;
        default: usage(argv[0]); return 2;
        }
    }
    int npos = argc - optind;
    char **pos = argv + optind;

    param params;
    AABB_t boundary, water;
    scene_file sf;
    memset(&sf, 0, sizeof sf);
    int nx, max_local, steps_per_frame = 4, steps;
    float spacing;
    const char *out = NULL;
    if (frames_path) {
        if (npos < 1 || read_frames(frames_path, &sf) != 0) {
            fprintf(stderr, "cannot read frame file %s\n", frames_path);
            usage(argv[0]);
            return 2;
        }
        steps = atoi(pos[0]);
        out = npos > 1 ? pos[1] : NULL;
        params = sf.params;
        boundary = sf.box;
        water = sf.water;
        spacing = sf.spacing;
        nx = sf.number_particles_x;
        max_local = sf.max_local;
        steps_per_frame = sf.steps_per_frame;
    } else {
        if (npos < 3) {
            usage(argv[0]);
            return 2;
        }
        nx = atoi(pos[0]);
        int ny = atoi(pos[1]);
        steps = atoi(pos[2]);
        out = npos > 3 ? pos[3] : NULL;
        if (nx <= 0 || ny <= 0) {
            usage(argv[0]);
            return 2;
        }
        memset(&params, 0, sizeof params);
        params.tunable_params.rest_density = 10.0f;
        params.tunable_params.smoothing_radius = 1.0f;
        params.tunable_params.g = 9.8f;
        params.tunable_params.k = 0.004f;
        params.tunable_params.k_near = 0.01f;
        params.tunable_params.k_spring = 10.0f;
        params.tunable_params.sigma = 0.0f;
        params.tunable_params.beta = 0.3f;
        params.tunable_params.time_step = 1.0f / 60.0f;
        params.tunable_params.mover_type = SPHERE_MOVER;
        params.tunable_params.mover_width = 0.0f;
        params.tunable_params.mover_height = 0.0f;
        params.tunable_params.active = 1;
        AABB_t b = {0.0f, bw > 0.0f ? bw : (float)(nx * 60 / 64), 0.0f,
                    bh > 0.0f ? bh : (float)(ny * 40 / 64), 0.0f, 0.0f};
        AABB_t w = {0.0f, nx * 0.5f - 0.25f, 0.0f, ny * 0.5f, 0.0f, 0.0f};
        boundary = b;
        water = w;
        spacing = 0.5f;
        params.tunable_params.node_start_x = boundary.min_x;
        params.tunable_params.node_end_x = boundary.max_x;
        params.number_fluid_particles_global = nx * ny;
        max_local = nx * ny;
    }
    if (steps < 0) {
        usage(argv[0]);
        return 2;
    }

    fluid_particle *particles = calloc((size_t)max_local, sizeof *particles);
    fluid_particle **ptrs = calloc((size_t)max_local, sizeof *ptrs);
    short *coords = malloc(sizeof(short) * 2 * (size_t)max_local);
    neighbor_grid_t grid; /* the reference's grid record; lists stay on the device */
    memset(&grid, 0, sizeof grid);
    grid.spacing = params.tunable_params.smoothing_radius;
    grid.max_bucket_size = uncapped ? 0u : SPH_REF_MAX_BUCKET; /* src/fluid.c:174-175 */
    grid.max_neighbors = SPH_REF_MAX_NEIGHBORS;
    grid.size_x = (unsigned)(boundary.max_x / grid.spacing + 0.999999f);
    grid.size_y = (unsigned)(boundary.max_y / grid.spacing + 0.999999f);
    if (!particles || !ptrs || !coords) {
        fprintf(stderr, "allocation failed\n");
        return 1;
    }
    initParticles(ptrs, particles, &water, 0, nx, NULL, max_local, spacing, &params);
    sph_shim_set_sync_interval(sync_interval);
    sph_shim_set_sync_fields(positions_only ? SPH_SYNC_POSITIONS : SPH_SYNC_ALL);

    double t_start = 0.0;
    int frame = 0;
    for (int step = 0, sub_step = 0; step < steps; step++) {
        if (step == 2 && timing) {
            sph_sync(sph_shim_context());
            t_start = now_s();
        }
        apply_gravity(ptrs, &params);
        viscosity_impluses(ptrs, NULL, &params);
        predict_positions(ptrs, &boundary, &params);
        if (sub_step == steps_per_frame - 1 && frame < sf.nframes) /* src/fluid.c:293-294 */
            params.tunable_params = sf.frames[frame++];
        if (params.tunable_params.kill_sim)
            break;
        identify_oob_particles(ptrs, particles, NULL, &boundary, &params);
        hash_fluid(ptrs, &grid, &params, true);
        hash_halo(ptrs, &grid, &params, true);
        double_density_relaxation(ptrs, NULL, &params);
        updateVelocities(ptrs, NULL, &boundary, &params);
        hash_fluid(ptrs, &grid, &params, false);
        hash_halo(ptrs, &grid, &params, false);
        if (sub_step == steps_per_frame - 1) { /* src/fluid.c:356-363 */
            for (int i = 0; i < params.number_fluid_particles_local; i++) {
                fluid_particle *p = ptrs[i];
                coords[i * 2] = (2.0f * p->x / boundary.max_x - 1.0f) * SHRT_MAX;
                coords[i * 2 + 1] = (2.0f * p->y / boundary.max_y - 1.0f) * SHRT_MAX;
            }
        }
        sub_step = sub_step == steps_per_frame - 1 ? 0 : sub_step + 1;
    }
    if (timing && steps > 2) {
        sph_sync(sph_shim_context());
        double el = now_s() - t_start;
        int n = params.number_fluid_particles_local;
        printf("sph_main timing: %d particles, %d timed steps, %.4f ms/step, %.4g particle-steps/s, "
               "sync every %d step(s), %s\n",
               n, steps - 2, el * 1e3 / (steps - 2), (double)n * (steps - 2) / el, sync_interval,
               positions_only ? "positions only" : "all fields");
    }
    if (sph_last_error()[0]) {
        fprintf(stderr, "libsph_hip: %s\n", sph_last_error());
        return 1;
    }
    sph_shim_set_sync_fields(SPH_SYNC_ALL);
    sph_shim_sync();
    int n = params.number_fluid_particles_local;
    if (out) {
        FILE *f = fopen(out, "wb");
        if (!f || fwrite(particles, sizeof *particles, (size_t)n, f) != (size_t)n) {
            fprintf(stderr, "cannot write %s\n", out);
            return 1;
        }
        fclose(f);
    }
    double ysum = 0.0;
    for (int i = 0; i < n; i++)
        ysum += particles[i].y;
    printf("sph_main: %d particles, %d steps, mean y %.6f\n", n, steps, n ? ysum / n : 0.0);
    sph_shim_release();
    free(particles);
    free(ptrs);
    free(coords);
    free(sf.frames);
    return 0;
}
