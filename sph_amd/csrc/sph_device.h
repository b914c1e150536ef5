// sph_device.h - device-side math of the TinySPH step, shared by every
// kernel of libsph_hip.so.  Each helper restates one reference expression
// with the same operation order, so that with -ffp-contract=off and
// IEEE-correct division / sqrt (HIP's defaults) the device reproduces the
// reference's float arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_abi.h"

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
    int h_pow2;        // 1: h is a power of two with a normal reciprocal, so that
                       //    x / h == x * h_recip for every x (both round x 2^-k once)
    unsigned bucket_cap; // 0, or the reference's per-cell bucket cap: a particle
                         // ranked bucket_cap or later by id in its cell is in
                         // no neighbour set of the hash (src/hash.c:153-164)
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
    const float qx = P.h_pow2 ? x * P.h_recip : x / P.h;
    const float qy = P.h_pow2 ? y * P.h_recip : y / P.h;
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

// Correctly rounded square root, branch-free.  For 2^-96 <= x < inf,
// v_sqrt_f32 is within one ulp of sqrt(x); the rounded result is the
// candidate among s - ulp, s, s + ulp whose square brackets x: s - ulp when
// x - (s - ulp)*s <= 0, s + ulp when x - (s + ulp)*s > 0 (each residual
// formed with one rounding by an FMA), else s.  Smaller inputs (denormals
// included) are scaled by 2^64 into that range and the root by 2^-32, both
// exactly; zero, negative numbers, infinity and NaN pass through the same
// arithmetic with the right result (the residuals are NaN or zero there).
// Results equal sqrtf's (IEEE round-to-nearest) for every input: checked
// exhaustively over all floats (sph_selftest_sqrt).
__device__ __forceinline__ float sph_sqrt_rounded(float xs)
{
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float sd = __int_as_float(__float_as_int(s) - 1);
    const float su = __int_as_float(__float_as_int(s) + 1);
    float r = s;
    r = __builtin_fmaf(-sd, s, xs) <= 0.0f ? sd : r;
    r = __builtin_fmaf(-su, s, xs) > 0.0f ? su : r;
    return r;
}

// The scaling is taken only when some active lane of the wave needs it: a
// pair closer than 2^-48 (coincident particles) is rare, and the common
// path then costs the root and its correction alone.
__device__ __forceinline__ float sph_sqrt(float x)
{
    if (__builtin_expect(__ballot(x < 0x1p-96f) != 0ull, 0)) {
        const bool tiny = x < 0x1p-96f;
        const float r = sph_sqrt_rounded(tiny ? x * 0x1p64f : x);
        return tiny ? r * 0x1p-32f : r;
    }
    return sph_sqrt_rounded(x);
}

// Records a walk may read past the end of a candidate row (the test loop
// works in whole groups of 8): every array a walk reads from global memory,
// and every LDS tile, has this many spare records at its end.
constexpr unsigned kSphWalkPad = 8;

// One candidate's bit of an accepted-candidate mask: mask = 2 mask + !(r2 >
// h2), the reference's pair test (src/hash.c:99, 185, 221; a NaN distance is
// accepted, as there) as one compare into VCC and one add with carry.
__device__ __forceinline__ unsigned sph_mask_push(unsigned mask, float r2, float h2)
{
    asm("v_cmp_nlt_f32_e32 vcc, %1, %2\n\t"
        "v_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
        : "+v"(mask)
        : "s"(h2), "v"(r2)
        : "vcc");
    return mask;
}

// Correctly rounded 1/x, branch-free, for x = +-0, +-inf, NaN and
// 2^-126 <= |x| <= 2^125.  There v_rcp_f32 gives y within one ulp of 1/x;
// the residual e = 1 - x*y is formed with one rounding by an FMA and the
// corrected y + e*y rounded once (a second FMA) is the correctly rounded
// reciprocal.  e is NaN exactly for zero, infinite and NaN x, where y itself
// is 1/x.  Every r the pair terms divide by lies in that domain: r =
// sqrt(r2) is 0 or at least 2^-74.5 and at most h, which the host bounds
// by 2^60.  Results equal 1.0f / x over the whole domain: checked
// exhaustively (sph_selftest_rcp).
__device__ __forceinline__ float sph_rcp(float x)
{
    const float y = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, y, 1.0f);
    const float r = __builtin_fmaf(e, y, y);
    return e == e ? r : y;
}

// The domain of sph_rcp.
__device__ __forceinline__ bool sph_rcp_domain(float x)
{
    const float ax = __builtin_fabsf(x);
    return !(ax > 0.0f && ax < 0x1p-126f) && !(ax > 0x1p125f && ax < __builtin_inff());
}

// src/fluid.c:613-625 without branches: the median of (v, -5, 5) is v
// clamped to [-5, 5] (v_med3_f32); a NaN passes through unchanged, as in
// the reference's if / else if.
__device__ __forceinline__ float sph_clamp_v(float v)
{
    const float v_max = 5.0f;
    return v != v ? v : __builtin_amdgcn_fmed3f(v, -v_max, v_max);
}

__device__ __forceinline__ void sph_check_velocity(float &vx, float &vy)
{
    vx = sph_clamp_v(vx);
    vy = sph_clamp_v(vy);
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
// Evaluated unconditionally and returned with the test u > 0, so that the
// caller applies it with selects (no branch in the pair loop).
__device__ __forceinline__ bool sph_visc_impulse(float px, float py, float pvx, float pvy, float qx,
                                                 float qy, float qvx, float qvy,
                                                 const sph_dev_params &P, float &ix, float &iy)
{
    float QmP_x = qx - px;
    float QmP_y = qy - py;
    float r = sph_sqrt(QmP_x * QmP_x + QmP_y * QmP_y);
    float r_recip = sph_rcp(r); // 1.0f / r
    float ratio = r * P.h_recip;
    float u = ((pvx - qvx) * QmP_x + (pvy - qvy) * QmP_y) * r_recip;
    float imp = P.dt * (1.0f - ratio) * (P.sigma * u + P.beta * u * u);
    ix = imp * QmP_x * r_recip;
    iy = imp * QmP_y * r_recip;
    sph_check_velocity(ix, iy);
    return u > 0.0f;
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
// test, so the wave takes one branch per candidate.  `src.row(dy)` is the
// offset from sorted position to source index in row dy (the LDS offset of
// a staged tile, 0 for global memory); `src.at(idx)` returns the record at a
// source index; body(k, dy, b1, b2, rec).
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
        const unsigned off = (unsigned)src.row(dy);
        for (unsigned k = b0; k < b3; k++)
            body(k, dy, b1, b2, src.at(k + off));
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

// Accepted-candidate masks carried from one walk to a later walk over the
// same neighbour set: the density pass and the relaxation after it walk the
// same sorted positions, so the relaxation replays the density pass's masks
// of a particle's first kMaskMemo chunks (in walk order) instead of testing
// those candidates again.  A policy supplies (have/get) or keeps (put) them.
constexpr unsigned kMaskMemo = 4;

struct sph_masks_none {
    __device__ bool have(unsigned) const { return false; }
    __device__ unsigned get(unsigned) const { return 0u; }
    __device__ void put(unsigned, unsigned) {}
};

struct sph_masks_record {
    uint4 m = make_uint4(0u, 0u, 0u, 0u);
    __device__ bool have(unsigned) const { return false; }
    __device__ unsigned get(unsigned) const { return 0u; }
    __device__ void put(unsigned ci, unsigned mask)
    {
        m.x = ci == 0u ? mask : m.x;
        m.y = ci == 1u ? mask : m.y;
        m.z = ci == 2u ? mask : m.z;
        m.w = ci == 3u ? mask : m.w;
    }
};

struct sph_masks_replay {
    uint4 m;
    __device__ bool have(unsigned ci) const { return ci < kMaskMemo; }
    __device__ unsigned get(unsigned ci) const
    {
        return ci == 0u ? m.x : ci == 1u ? m.y : ci == 2u ? m.z : m.w;
    }
    __device__ void put(unsigned, unsigned) {}
};

// Four lanes per particle (the walk's LP = 4): lanes 4p .. 4p + 3 of a wave
// hold the same particle, and quad-wide DPP moves share values among them.
// Each of the four takes one pair of a drain step and the pair terms are
// then applied in canonical order by all four, so every sum is formed in
// exactly the order of one lane per particle.
struct sph_pair_term {
    float a, b; // the term's two components (zero: no update)
    int f;      // an extra flag of the term (the relaxation's 1e-6 nudge)
};

// A pair body split in two for LP = 4: term(idx, own, rec) evaluates one
// pair's contribution, apply(t) adds it to the particle's sums; apply of a
// zero term with f = 0 must leave the sums unchanged bit for bit.
template <class Term, class Apply>
struct sph_split_body {
    Term term;
    Apply apply;
};

__device__ __forceinline__ unsigned sph_quad_lane() { return __lane_id() & 3u; }

template <int CTRL>
__device__ __forceinline__ unsigned sph_dpp(unsigned x)
{
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}

// Lane j's term of this lane's quad (quad_perm j, j, j, j).
template <int J>
__device__ __forceinline__ sph_pair_term sph_quad_term(const sph_pair_term &t)
{
    constexpr int kCtrl = J * 0x55;
    return sph_pair_term{__uint_as_float(sph_dpp<kCtrl>(__float_as_uint(t.a))),
                         __uint_as_float(sph_dpp<kCtrl>(__float_as_uint(t.b))),
                         (int)sph_dpp<kCtrl>((unsigned)t.f)};
}

// The same walk in two passes, for SIMD efficiency.  Pass 1 tests the
// candidates 32 at a time (dist2(rec): squared distance, within h when not
// above h2), branch-free, into a mask
// of accepted ones, from which the particle's own slot `self` is cleared;
// non-empty masks wait in a per-lane queue of up to three chunks (mask,
// source index of bit 0, owner threshold).  Pass 2 drains the queue in the
// order the chunks were found, lowest bit first, running body() on each
// accepted candidate: exactly sph_walk's visiting order, so sums are
// bit-identical.  The whole wave drains together, at the end and whenever
// some lane's queue is full, so a lane runs the accepted pairs of up to
// three chunks (usually all three rows) in one loop: a wave executes as many
// pair iterations as its busiest lane has pairs, where a drain per row would
// execute the sum over rows of each row's busiest lane.
//
// The half-list count of src/hash.c:176-237 is taken per chunk: the pair is
// in self's half list when the candidate lies in a forward cell (0,+1),
// (+1,-1), (+1,0), (+1,+1) or shares self's cell at a later sorted position
// (cells are sorted by id), i.e. when its sorted position is at least b1 in
// row +1, self + 1 in row 0 and b2 in row -1.  body(idx, own, rec) receives
// that test for the candidate when OWN is set (false otherwise); idx is the
// candidate's index in the source.  mk supplies or keeps the masks of the
// lane's first chunks (sph_masks_*).  Returns (accepted, in the half list).
//
// CAP applies the reference's bucket cap `cap`: a candidate ranked cap or
// later in its cell (its sorted position minus the cell's start) is masked
// out, and a lane whose own rank is cap or more walks nothing.  A separate
// instantiation, so that the uncapped walk keeps its registers.
//
// DY0..DY1 limit the walk to those candidate rows, and ci counts the lane's
// chunks across calls: walking the rows one call at a time (a tile staged row
// by row) visits the same chunks in the same order as one call over all three.
//
// LP = 4: four lanes per particle (sph_split_body).  Lane q of the quad tests
// the chunk's candidates 8q .. 8q + 7 and the quad ORs the four bytes into
// the chunk mask, so all four hold the same masks and queue; a drain step
// gives lane q the q-th pending pair of the head chunk, and the four terms
// are applied by every lane in pair order.  Same chunks, same masks, same
// sums as LP = 1, with a quarter of the serial work per lane.
template <bool OWN = false, bool CAP = false, int DY0 = -1, int DY1 = 1, int LP = 1, class Src,
          class Dist2, class Body, class Masks>
__device__ __forceinline__ uint2 sph_walk_rows(unsigned self, const sph_rows &R, const Src &src,
                                               float h2, unsigned cap, Dist2 &&dist2, Body &&body,
                                               Masks &mk, unsigned &ci)
{
    // bits t of a chunk starting at c with c + t < x
    auto below = [](unsigned c, unsigned x) {
        return x <= c ? 0u : x - c >= 32u ? ~0u : (1u << (x - c)) - 1u;
    };
    const bool dropped = CAP && self - R.b1[1] >= cap;
    unsigned m0 = 0, m1 = 0, m2 = 0; // pending masks, m0 drained first
    unsigned x0 = 0, x1 = 0, x2 = 0; // source index of each mask's bit 0
    unsigned t0 = 0, t1 = 0, t2 = 0; // owner threshold, source index space
    unsigned nq = 0;
    unsigned full = 0, half = 0;
    auto drain = [&]() {
        // the queue in locals: selects between captured variables can be
        // folded into a load from a selected address, which puts the queue
        // in scratch memory
        unsigned q0 = m0, q1 = m1, q2 = m2, y0 = x0, y1 = x1, y2 = x2, u0 = t0, u1 = t1,
                 u2 = t2;
        while (q0) {
            if constexpr (LP == 1) {
                const unsigned t = (unsigned)__builtin_ctz(q0);
                q0 &= q0 - 1u;
                const unsigned idx = y0 + t;
                body(idx, OWN && idx >= u0, src.at(idx));
            } else {
                // lane q: the q-th lowest pending pair of the head chunk
                const unsigned q = sph_quad_lane();
                unsigned r = q0;
                r = q >= 1u ? r & (r - 1u) : r;
                r = q >= 2u ? r & (r - 1u) : r;
                r = q >= 3u ? r & (r - 1u) : r;
                const bool have = r != 0u;
                const unsigned idx = y0 + (have ? (unsigned)__builtin_ctz(r) : 0u);
                sph_pair_term tm = body.term(idx, OWN && idx >= u0, src.at(idx));
                tm.a = have ? tm.a : 0.0f;
                tm.b = have ? tm.b : 0.0f;
                tm.f = have ? tm.f : 0;
                body.apply(sph_quad_term<0>(tm));
                body.apply(sph_quad_term<1>(tm));
                body.apply(sph_quad_term<2>(tm));
                body.apply(sph_quad_term<3>(tm));
                q0 &= q0 - 1u;
                q0 &= q0 - 1u;
                q0 &= q0 - 1u;
                q0 &= q0 - 1u;
            }
            const bool next = q0 == 0; // chunk done: the queue moves up
            q0 = next ? q1 : q0;
            y0 = next ? y1 : y0;
            u0 = next ? u1 : u0;
            q1 = next ? q2 : q1;
            y1 = next ? y2 : y1;
            u1 = next ? u2 : u1;
            q2 = next ? 0u : q2;
        }
        m0 = m1 = m2 = 0u;
        nq = 0;
    };
#pragma unroll
    for (int dy = DY0; dy <= DY1; dy++) {
        // the row, its owner threshold and self in source index space (each
        // a true index in the row's range where it matters: the arithmetic
        // is modular, so a negative offset cancels)
        const unsigned off = (unsigned)src.row(dy);
        const unsigned thr = dy == 1 ? R.b1[dy + 1] : dy == 0 ? self + 1u : R.b2[dy + 1];
        const unsigned tt = thr + off, me = self + off;
        const unsigned b3 = R.b3[dy + 1] + off;
        unsigned c = dropped ? b3 : R.b0[dy + 1] + off;
        // does the cap cut one of the row's three cells?  (rarely: then the
        // chunk masks below are cut to each cell's first cap positions)
        const bool cut = CAP && (R.b1[dy + 1] - R.b0[dy + 1] > cap ||
                                 R.b2[dy + 1] - R.b1[dy + 1] > cap ||
                                 R.b3[dy + 1] - R.b2[dy + 1] > cap);
        // The chunk loop runs while any lane of the wave has candidates left
        // in this row (lanes that are done contribute empty masks), so a
        // drain is always taken by the whole wave at once: lanes' pair work
        // overlaps instead of running lane after lane.
        while (__any(c < b3)) {
            unsigned mask = 0, x = 0;
            if (c < b3) {
                const unsigned m = b3 - c < 32u ? b3 - c : 32u;
                if (mk.have(ci)) {
                    mask = mk.get(ci);
                } else {
                    // groups of 8 candidates, pushed without predicates: the
                    // last group reads up to 7 records past the row (sources
                    // are padded by kSphWalkPad); the first candidate ends in
                    // the top pushed bit, so a bit reversal and a shift give
                    // candidate t bit t, and the padding's bits are cut off
                    if constexpr (LP == 1) {
                        const unsigned groups = (m + 7u) >> 3;
                        unsigned acc = 0;
                        for (unsigned g = 0; g < groups; g++) {
                            const unsigned base = c + 8u * g;
#pragma unroll
                            for (unsigned k = 0; k < 8u; k++)
                                acc = sph_mask_push(acc, dist2(src.at(base + k)), h2);
                        }
                        mask = __builtin_bitreverse32(acc) >> (32u - 8u * groups);
                    } else {
                        // lane q: group q; the quad ORs the four bytes
                        const unsigned q = sph_quad_lane();
                        unsigned acc = 0;
                        if (8u * q < m) {
                            const unsigned base = c + 8u * q;
#pragma unroll
                            for (unsigned k = 0; k < 8u; k++)
                                acc = sph_mask_push(acc, dist2(src.at(base + k)), h2);
                        }
                        unsigned part = (__builtin_bitreverse32(acc) >> 24) << (8u * q);
                        part |= sph_dpp<0xB1>(part); // quad_perm 1, 0, 3, 2
                        part |= sph_dpp<0x4E>(part); // quad_perm 2, 3, 0, 1
                        mask = part;
                    }
                    mask &= m < 32u ? (1u << m) - 1u : ~0u;
                    if (me - c < m) // the particle itself is no neighbour
                        mask &= ~(1u << (me - c));
                    if (cut) {
                        // the capped part [a, e) of each cell, source index space
                        const unsigned a0 = R.b0[dy + 1] + off, a1 = R.b1[dy + 1] + off,
                                       a2 = R.b2[dy + 1] + off;
                        const unsigned e0 = a1 - a0 > cap ? a0 + cap : a1;
                        const unsigned e1 = a2 - a1 > cap ? a1 + cap : a2;
                        const unsigned e2 = b3 - a2 > cap ? a2 + cap : b3;
                        mask &= (below(c, e0) & ~below(c, a0)) | (below(c, e1) & ~below(c, a1)) |
                                (below(c, e2) & ~below(c, a2));
                    }
                }
                mk.put(ci, mask);
                ci++;
                full += (unsigned)__popc(mask);
                const unsigned own_from = tt > c ? tt - c : 0u;
                half += (unsigned)__popc(own_from < 32u ? mask & (~0u << own_from) : 0u);
                x = c;
                c += 32u;
            }
            if (__any(nq == 3u && mask != 0u))
                drain();
            if (mask) {
                if (nq == 0u) {
                    m0 = mask;
                    x0 = x;
                    t0 = tt;
                } else if (nq == 1u) {
                    m1 = mask;
                    x1 = x;
                    t1 = tt;
                } else {
                    m2 = mask;
                    x2 = x;
                    t2 = tt;
                }
                nq++;
            }
        }
    }
    drain();
    return make_uint2(full, half);
}

// All three candidate rows in one call.
template <bool OWN = false, bool CAP = false, int LP = 1, class Src, class Dist2, class Body,
          class Masks>
__device__ __forceinline__ uint2 sph_walk_masked(unsigned self, const sph_rows &R, const Src &src,
                                                 float h2, unsigned cap, Dist2 &&dist2,
                                                 Body &&body, Masks &mk)
{
    unsigned ci = 0; // chunks of this lane so far (all rows, walk order)
    return sph_walk_rows<OWN, CAP, -1, 1, LP>(self, R, src, h2, cap, dist2, body, mk, ci);
}
