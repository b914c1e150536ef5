/*
 * sph_main.c - a start_simulation-shaped driver (src/fluid.c:71-395 with
 * the MPI render/halo plumbing removed) that calls the reference-named step
 * functions exported by libsph_hip.so, exactly in the reference's order.
 * It demonstrates the drop-in: the loop body below is the reference's loop
 * body (src/fluid.c:273-348) minus its MPI lines.
 *
 *   sph_main [nx ny steps [out.bin]]
 *
 * Scene: an nx x ny lattice at spacing 0.5 (h = 1) filling the lower-left
 * of a box sized as the benchmark scenes (60 x 40 for 64 x 64).  The final
 * particles are written as raw fluid_particle records to out.bin.
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sph_abi.h"

int main(int argc, char **argv)
{
    int nx = argc > 1 ? atoi(argv[1]) : 64;
    int ny = argc > 2 ? atoi(argv[2]) : 64;
    int steps = argc > 3 ? atoi(argv[3]) : 100;
    const char *out = argc > 4 ? argv[4] : NULL;
    if (nx <= 0 || ny <= 0 || steps < 0) {
        fprintf(stderr, "usage: %s [nx ny steps [out.bin]]\n", argv[0]);
        return 2;
    }

    param params;
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

    AABB_t boundary = {0.0f, (float)(nx * 60 / 64), 0.0f, (float)(ny * 40 / 64), 0.0f, 0.0f};
    AABB_t water = {0.0f, nx * 0.5f - 0.25f, 0.0f, ny * 0.5f, 0.0f, 0.0f};
    params.tunable_params.node_start_x = boundary.min_x;
    params.tunable_params.node_end_x = boundary.max_x;
    int n = nx * ny;
    params.number_fluid_particles_global = n;

    fluid_particle *particles = calloc((size_t)n, sizeof *particles);
    fluid_particle **ptrs = calloc((size_t)n, sizeof *ptrs);
    short *coords = malloc(sizeof(short) * 2 * (size_t)n);
    neighbor_grid_t grid; /* the reference's grid record; lists stay on the device */
    memset(&grid, 0, sizeof grid);
    grid.spacing = params.tunable_params.smoothing_radius;
    grid.size_x = (unsigned)(boundary.max_x / grid.spacing + 0.999999f);
    grid.size_y = (unsigned)(boundary.max_y / grid.spacing + 0.999999f);
    if (!particles || !ptrs || !coords) {
        fprintf(stderr, "allocation failed\n");
        return 1;
    }
    initParticles(ptrs, particles, &water, 0, nx, NULL, n, 0.5f, &params);

    int steps_per_frame = 4;
    for (int step = 0, sub_step = 0; step < steps; step++) {
        apply_gravity(ptrs, &params);
        viscosity_impluses(ptrs, NULL, &params);
        predict_positions(ptrs, &boundary, &params);
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
    if (sph_last_error()[0]) {
        fprintf(stderr, "libsph_hip: %s\n", sph_last_error());
        return 1;
    }
    sph_shim_sync();
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
    printf("sph_main: %d particles, %d steps, mean y %.6f\n", n, steps, ysum / n);
    sph_shim_release();
    free(particles);
    free(ptrs);
    free(coords);
    return 0;
}
