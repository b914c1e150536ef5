/*
 * scene.c - host-side scene builder of libsph_hip.so.
 *
 * The reference fills its fluid block as a lattice, row-major with y in the
 * outer loop and id = index (src/geometry.c:29-67):
 *     x = min_x + (start_x + nx) * spacing,  y = min_y + ny * spacing.
 * The benchmark scenes add a seeded uniform jitter of +-jitter to each
 * coordinate.  The generator is SplitMix64 (Steele, Lea, Flood 2014):
 *     s += 0x9E3779B97F4A7C15; z = s;
 *     z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9;
 *     z = (z ^ (z >> 27)) * 0x94D049BB133111EB;
 *     return z ^ (z >> 31);
 * A draw u = (z >> 40) * 2^-24 is exact in float; the jitter is
 * jitter * (2u - 1), drawn for x then y, particle by particle.
 */
#include <stdint.h>
#include <string.h>

#include "../../include/sph_abi.h"

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static float draw_sym(uint64_t *s, float amp)
{
    float u = (float)(splitmix64(s) >> 40) * 0x1p-24f;
    return amp * (2.0f * u - 1.0f);
}

int sph_scene_lattice(fluid_particle *out, int nx, int ny, float x0, float y0, float spacing,
                      float jitter, uint64_t seed)
{
    if (!out || nx < 0 || ny < 0 || !(spacing > 0.0f) || !(jitter >= 0.0f))
        return SPH_EINVAL;
    uint64_t s = seed;
    int i = 0;
    for (int jy = 0; jy < ny; jy++) {
        float y = y0 + (float)jy * spacing;
        for (int jx = 0; jx < nx; jx++) {
            fluid_particle *p = &out[i];
            memset(p, 0, sizeof *p);
            float x = x0 + (float)jx * spacing;
            float dx = jitter > 0.0f ? draw_sym(&s, jitter) : 0.0f;
            float dy = jitter > 0.0f ? draw_sym(&s, jitter) : 0.0f;
            p->x = x + dx;
            p->y = y + dy;
            p->x_prev = p->x;
            p->y_prev = p->y;
            p->id = i;
            i++;
        }
    }
    return SPH_OK;
}
