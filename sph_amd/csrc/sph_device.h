// sph_device.h - device-side math of the TinySPH step, shared by every
// kernel of libsph_hip.so.  Each helper restates one reference expression
// with the same operation order, so that with -ffp-contract=off and
// IEEE-correct division / sqrt (HIP's defaults) the device reproduces the
// reference's float arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_abi.h"

// Accepted neighbour pairs evaluated per iteration of a particle's walk
// (independent dependency chains for latency hiding).
#ifndef SPH_PAIR_ILP
#define SPH_PAIR_ILP 1
#endif

// Step parameters of one scene as the kernels read them, derived on the
// host from tunable_parameters exactly as the reference derives them
// (h_recip = 1.0f/h: src/fluid.c:427, 555; g = -g: src/fluid.c:403;
// h2 = h*h: src/hash.c:133), plus the scene's place in the shared arrays:
// its particles occupy slots [poff, poff + n) of the sorted arrays (the
// range is padded to a multiple of 256 slots, so every workgroup belongs to
// one scene) and its cells are global cells [cell_base, cell_base + ncells),
// followed by one padding cell that absorbs the pad slots in the sort.
struct sph_dev_params {
    float h, h_recip, h2, dt, g_neg;
    float rest_density, k, k_near, k_spring, sigma, beta;
    float min_x, max_x, min_y, max_y;
    float mover_cx, mover_cy, mover_w, mover_h;
    int mover_type;
    unsigned sx, sy, ncells; // local grid: ncells = sx * sy
    unsigned cell_base;      // first global cell of the scene
    unsigned poff;           // first slot of the scene in the sorted arrays
    unsigned uoff;           // first index of the scene in caller (unpadded) order
    int n;                   // particles of the scene
    unsigned pad;            // padding slots after the scene's particles
    // Extensions beyond the reference (both off by default):
    float slosh_amp;   // horizontal box acceleration A sin(omega t) ("sloshing")
    float slosh_omega; //   applied, like gravity, as v_x += -A sin(omega t) dt
    unsigned slosh_t0; //   with t = (steps since upload + slosh_t0) * dt
    int frozen;        // 1: the fused step leaves this scene's state unchanged
};

// Horizontal box acceleration of a sloshing scene at the step counted by
// `clock` (0 when the scene does not slosh).
__device__ __forceinline__ float sph_slosh_accel(const sph_dev_params &P, unsigned long long clock)
{
    if (P.slosh_amp == 0.0f)
        return 0.0f;
    return -P.slosh_amp * sinf(P.slosh_omega * (float)(clock + P.slosh_t0) * P.dt);
}

// Device health counters (mirrors sph_counters_t from the key counters on).
struct sph_dev_counters {
    unsigned long long bucket_overflow;
    unsigned long long neighbor_overflow;
    unsigned long long key_out_of_range;
    unsigned long long nonfinite;
    unsigned long long max_cell_count;
    unsigned long long max_neighbors;
};

// src/hash.c:35-47.  key = floor(y/h)*size_x + floor(x/h); keys the
// reference could not index (>= cells, negative, NaN: undefined behaviour
// there) are counted and clamped to the scene's last cell.  Returns the
// global cell (cell_base + local key).
__device__ __forceinline__ unsigned sph_cell_key(float x, float y, const sph_dev_params &P,
                                                 unsigned long long *oor)
{
    float qx = x / P.h, qy = y / P.h;
    if (!(qx >= 0.0f && qy >= 0.0f && qx < 4294967296.0f && qy < 4294967296.0f)) {
        atomicAdd(oor, 1ull);
        return P.cell_base + P.ncells - 1;
    }
    unsigned long long k = (unsigned long long)(unsigned)floorf(qy) * P.sx +
                           (unsigned long long)(unsigned)floorf(qx);
    if (k >= P.ncells) {
        atomicAdd(oor, 1ull);
        return P.cell_base + P.ncells - 1;
    }
    return P.cell_base + (unsigned)k;
}

// Correctly rounded square root.  For 2^-96 <= x < inf, v_sqrt_f32 is
// within one ulp of sqrt(x); the rounded result is the candidate among
// s - ulp, s, s + ulp whose square brackets x: s - ulp when
// x - (s - ulp)*s <= 0, s + ulp when x - (s + ulp)*s > 0 (each residual
// formed with one rounding by an FMA), else s.  Smaller, zero and
// non-finite inputs take the general sqrtf path.  Results equal sqrtf's
// (IEEE round-to-nearest) for every input: checked exhaustively over all
// floats (sph_selftest_sqrt).
__device__ __forceinline__ float sph_sqrt(float x)
{
    if (!(x >= 0x1p-96f && x < __builtin_inff()))
        return sqrtf(x);
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1);
    const float su = __int_as_float(__float_as_int(s) + 1);
    float r = s;
    if (__builtin_fmaf(-sd, s, x) <= 0.0f)
        r = sd;
    if (__builtin_fmaf(-su, s, x) > 0.0f)
        r = su;
    return r;
}

// src/fluid.c:613-625
__device__ __forceinline__ void sph_check_velocity(float &vx, float &vy)
{
    const float v_max = 5.0f;
    if (vx > v_max)
        vx = v_max;
    else if (vx < -v_max)
        vx = -v_max;
    if (vy > v_max)
        vy = v_max;
    else if (vy < -v_max)
        vy = -v_max;
}

// src/fluid.c:656-744: mover push-out then clamp into the box.
__device__ __forceinline__ void sph_boundary(float &px, float &py, const sph_dev_params &P)
{
    const float cx = P.mover_cx, cy = P.mover_cy;
    if (P.mover_type == SPHERE_MOVER) {
        float radius = P.mover_w * 0.5f;
        float d2 = (px - cx) * (px - cx) + (py - cy) * (py - cy);
        if (d2 <= radius * radius && d2 > 0.0f) {
            float d = sqrtf(d2);
            float nx = (cx - px) / d;
            float ny = (cy - py) / d;
            float pen = radius - d;
            px -= pen * nx;
            py -= pen * ny;
        }
    } else if (P.mover_type == RECTANGLE_MOVER) {
        float hw = (float)((double)P.mover_w * 0.5);
        float hh = (float)((double)P.mover_h * 0.5);
        float rx = px - cx, ry = py - cy;
        float ax = fabsf(rx), ay = fabsf(ry);
        if (ax < hw && ay < hh) {
            float penx = hw - ax, peny = hh - ay;
            if (penx < peny) {
                if (rx < 0.0f)
                    px -= penx;
                else
                    px += penx;
            } else {
                if (ry < 0.0f)
                    py -= peny;
                else
                    py += peny;
            }
        }
    }
    if (px < P.min_x)
        px = P.min_x;
    else if (px > P.max_x)
        px = P.max_x - 0.001f;
    if (py < P.min_y)
        py = P.min_y;
    else if (py > P.max_y)
        py = P.max_y - 0.001f;
}

// src/fluid.c:591.  The literal 0.5 is a double, so the spring term, its
// sum with the pressure terms and the product with dt*dt are double
// operations, rounded to float once.
__device__ __forceinline__ float sph_relax_D(float dt, float Psum, float Pnsum, float OmR,
                                             float k_spring, float h, float r)
{
    float a = Psum * OmR + Pnsum * OmR * OmR;
    float s = k_spring * (h - r);
    float dt2 = dt * dt;
    return (float)((double)dt2 * ((double)a + (double)s * 0.5));
}

// src/fluid.c:442-459: impulse of pair (p, q) with p in the owner role.
// Exactly antisymmetric in (p, q), so a particle in either role subtracts
// half of the impulse computed with itself as p.
__device__ __forceinline__ bool sph_visc_impulse(float px, float py, float pvx, float pvy, float qx,
                                                 float qy, float qvx, float qvy,
                                                 const sph_dev_params &P, float &ix, float &iy)
{
    float QmP_x = qx - px;
    float QmP_y = qy - py;
    float r = sph_sqrt(QmP_x * QmP_x + QmP_y * QmP_y);
    float r_recip = 1.0f / r;
    float ratio = r * P.h_recip;
    float u = ((pvx - qvx) * QmP_x + (pvy - qvy) * QmP_y) * r_recip;
    if (u > 0.0f) {
        float imp = P.dt * (1.0f - ratio) * (P.sigma * u + P.beta * u * u);
        ix = imp * QmP_x * r_recip;
        iy = imp * QmP_y * r_recip;
        sph_check_velocity(ix, iy);
        return true;
    }
    return false;
}

// Half-list owner rule of src/hash.c:176-237 for the pair (self at sorted
// position s, candidate at sorted position k in the row dy = r - j, where
// cells i-1 / i / i+1 of that row occupy [b0,b1) / [b1,b2) / [b2,b3)): the
// pair is in self's half list when the candidate's cell is a forward cell
// (0,+1), (+1,-1), (+1,0), (+1,+1), or shares self's cell with a higher id
// (cells are sorted by id, so: a later sorted position).
__device__ __forceinline__ bool sph_owns(unsigned k, unsigned s, int dy, unsigned b1, unsigned b2)
{
    if (k >= b2)
        return true; // dx = +1
    if (k < b1)
        return false; // dx = -1
    if (dy == 1)
        return true; // (0, +1)
    return dy == 0 && k > s;
}

// Walk the 3x3 neighbourhood of the particle at sorted position s (global
// cell `key`) in canonical order: rows j-1, j, j+1; in each row the contiguous
// sorted range of cells i-1..i+1 (clipped to the grid), ascending.  Self
// (k == s) is visited too; every body rejects it together with its distance
// test, so the wave takes one branch per candidate.  `src.row(dy)` prepares
// a row (an LDS offset for staged tiles); `src.at(k, row)` returns candidate
// k's staged record; body(k, dy, b1, b2, rec).
template <class Src, class Body>
__device__ __forceinline__ void sph_walk(unsigned s, unsigned key, const sph_dev_params &P,
                                         const unsigned *__restrict__ start, const Src &src,
                                         Body &&body)
{
    const unsigned sx = P.sx, sy = P.sy;
    const unsigned local = key - P.cell_base;
    const unsigned i = local % sx, j = local / sx;
    const unsigned i0 = i > 0 ? i - 1 : 0;
    const unsigned i1 = i + 1 < sx ? i + 1 : sx - 1;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
        const long r = (long)j + dy;
        if (r < 0 || r >= (long)sy)
            continue;
        const unsigned base = P.cell_base + (unsigned)r * sx;
        const unsigned b0 = start[base + i0], b1 = start[base + i];
        const unsigned b2 = start[base + i + 1], b3 = start[base + i1 + 1];
        const int row = src.row(dy);
        for (unsigned k = b0; k < b3; k++)
            body(k, dy, b1, b2, src.at(k, row));
    }
}

// The candidate rows of a particle in global cell `key`: row dy = -1, 0, 1
// holds its cells i-1 / i / i+1 (clipped to the grid) at sorted positions
// [b0,b1) / [b1,b2) / [b2,b3); a row outside the grid is empty.  Kernels
// load them as soon as the key is known, so that the walk itself reads no
// global memory but its candidates.
struct sph_rows {
    unsigned b0[3], b1[3], b2[3], b3[3];
};

__device__ __forceinline__ sph_rows sph_row_bounds(unsigned key, const sph_dev_params &P,
                                                   const unsigned *__restrict__ start)
{
    sph_rows R;
    const unsigned sx = P.sx, sy = P.sy;
    const unsigned local = key - P.cell_base;
    const unsigned i = local % sx, j = local / sx;
    const unsigned i0 = i > 0 ? i - 1 : 0;
    const unsigned i1 = i + 1 < sx ? i + 1 : sx - 1;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
        const long r = (long)j + dy;
        const bool in = r >= 0 && r < (long)sy;
        const unsigned base = P.cell_base + (in ? (unsigned)r : j) * sx;
        const unsigned b0 = start[base + i0], b1 = start[base + i];
        const unsigned b2 = start[base + i + 1], b3 = start[base + i1 + 1];
        R.b0[dy + 1] = in ? b0 : 0;
        R.b1[dy + 1] = in ? b1 : 0;
        R.b2[dy + 1] = in ? b2 : 0;
        R.b3[dy + 1] = in ? b3 : 0;
    }
    return R;
}

// The same walk in two passes per row, for SIMD efficiency: a branch-free
// pass tests up to W (32 or 64) candidates at a time (test(k, rec) -> within
// h) and records the accepted ones in a W-bit mask, from which the
// particle's own slot `self` is then cleared; a second pass runs body() on
// the set bits only, lowest first.  The visiting order of accepted
// candidates is exactly sph_walk's, so sums are bit-identical, but a wave no
// longer executes the pair work for every candidate that any one of its
// lanes accepts.  W = 32 makes the mask one register (cheapest per
// candidate: the density pass, whose pair work is light); W = 64 halves the
// number of second-pass loops (the heavy relaxation and viscosity pairs).
// body(k, dy, b1, b2, rec, use) applies its updates only when `use` is
// true.  Returns the number of accepted candidates.
template <int W> struct sph_mask_t;
template <> struct sph_mask_t<32> {
    typedef unsigned type;
    static __device__ __forceinline__ unsigned ctz(unsigned m) { return (unsigned)__builtin_ctz(m); }
    static __device__ __forceinline__ unsigned popc(unsigned m) { return (unsigned)__popc(m); }
};
template <> struct sph_mask_t<64> {
    typedef unsigned long long type;
    static __device__ __forceinline__ unsigned ctz(unsigned long long m)
    {
        return (unsigned)__builtin_ctzll(m);
    }
    static __device__ __forceinline__ unsigned popc(unsigned long long m)
    {
        return (unsigned)__popcll(m);
    }
};

template <int W = 64, class Src, class Test, class Body>
__device__ __forceinline__ unsigned sph_walk_masked(unsigned self, const sph_rows &R,
                                                    const Src &src, Test &&test, Body &&body)
{
    typedef sph_mask_t<W> M;
    typedef typename M::type mask_t;
    unsigned accepted = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
        const unsigned b0 = R.b0[dy + 1], b1 = R.b1[dy + 1];
        const unsigned b2 = R.b2[dy + 1], b3 = R.b3[dy + 1];
        const int row = src.row(dy);
        for (unsigned c = b0; c < b3; c += W) {
            const unsigned m = b3 - c < (unsigned)W ? b3 - c : (unsigned)W;
            mask_t mask = 0;
            for (unsigned t = 0; t < m; t++)
                mask |= (mask_t)test(c + t, src.at(c + t, row)) << t;
            if (self - c < m) // the particle itself is no neighbour
                mask &= ~((mask_t)1 << (self - c));
            accepted += M::popc(mask);
#if SPH_PAIR_ILP >= 2
            // two accepted candidates per iteration, as independent chains;
            // the second is applied only if it exists (use = false turns
            // its updates into no-ops), so the order of updates is unchanged
            while (mask) {
                const unsigned t0 = M::ctz(mask);
                mask &= mask - 1;
                const bool two = mask != 0;
                const unsigned t1 = two ? M::ctz(mask) : t0;
                mask &= two ? mask - 1 : mask;
                const auto r0 = src.at(c + t0, row);
                const auto r1 = src.at(c + t1, row);
                body(c + t0, dy, b1, b2, r0, true);
                body(c + t1, dy, b1, b2, r1, two);
            }
#else
            while (mask) {
                const unsigned t = M::ctz(mask);
                mask &= mask - 1;
                body(c + t, dy, b1, b2, src.at(c + t, row), true);
            }
#endif
        }
    }
    return accepted;
}
