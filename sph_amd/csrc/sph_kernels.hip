// sph_kernels.hip - gfx950 kernels of the TinySPH step (AdamSimpson/SPH).
//
// Data layout: one SoA particle array in HBM, kept sorted by (cell, id)
// after every hash, with paired fields interleaved (pos = (x, y),
// vel = (v_x, v_y), prev = (x_prev, y_prev), rho = (density,
// density_near)) so a gather reads a pair with one 8-byte load.
//
// Physics kernels are gathers over the 3x3 cells around each particle's
// cell.  A workgroup owns 256 consecutive sorted particles; their
// candidates are at most six contiguous ranges of the sorted arrays (three
// rows around each of the one or two cell rows the group touches), which
// the group stages into LDS with coalesced loads before each particle walks
// its own candidates in canonical order (rows ascending, sorted order within
// a row).  A group whose tile would not fit (it spans more than two cell
// rows, or the staged ranges exceed the tile) reads the same candidates from
// global memory in the same order, so both paths give identical results.
// No kernel writes another particle's state: no atomics on particle data.
//
// The cell sort is a counting sort: the producing kernel computes each
// particle's new key and claims a slot in its cell with an atomic counter;
// a two-kernel scan turns counts into cell starts (and clears the counts
// for the next sort); a scatter drops (id, source) pairs into the cell
// ranges; the reorder kernel ranks each particle by id inside its cell and
// gathers its fields there - the canonical (cell, id) order of
// src/hash.c:153-163.
//
// Several independent scenes can share one launch: each scene owns a
// 256-aligned range of slots of the sorted arrays and a range of global
// cells (plus a padding cell that takes the slack of its slot range in the
// sort), so a workgroup always belongs to exactly one scene and reads that
// scene's parameters as uniform values (BS maps workgroup -> scene).
//
// Workgroups are numbered XCD-aware: the hardware deals workgroups to the
// 8 XCDs round-robin, and logical tiles are remapped so that each XCD takes
// runs of 16 consecutive tiles (which share candidate rows in its L2), the
// runs dealt to the XCDs in turn.  Whole eighths per XCD would keep more rows
// in one L2, but the work per tile follows the fluid's density, which varies
// with height: in a settling column the densest eighth costs twice the mean,
// and that XCD would finish last.
#include "sph_kernels.h"

#include <hip/hip_ext.h>

#include <type_traits>

namespace {

constexpr int kWave = 64;
constexpr int kTPB = 256;
// The gathers ask the compiler for at least 6 resident workgroups per CU
// (6 waves per SIMD, at most 80 VGPRs): it then schedules them in 62-80
// VGPRs without scratch, where left alone it took 72-84 (5-6 waves).  LDS
// tiles allow 7 workgroups.  Measured on the 1M dam break at steps 221-230:
// density -2%, relaxation -4%, the capped variants -3 to -7%, viscosity
// unchanged.  The uncapped relaxation asks for 7, which keeps it at 72 VGPRs
// (7 waves) with its row-by-row staging path; 6 let it grow to 74 (6 waves).
// The four-lanes-per-particle variants (contexts of at most two tiles per
// CU, so at most eight of their 64-particle workgroups per CU) set no bound:
// their 66-88 VGPRs still allow 5-7 waves per SIMD, without scratch.
constexpr int kGatherMinBlocks = 6;
#ifndef SPH_TILE_CAP
#define SPH_TILE_CAP 1280
#endif
constexpr int kTileCap = SPH_TILE_CAP; // staged candidates per workgroup
constexpr unsigned kXcds = 8;
#ifndef SPH_XCD_CHUNK
#define SPH_XCD_CHUNK 16
#endif
constexpr unsigned kXcdChunk = SPH_XCD_CHUNK; // consecutive tiles per XCD run

__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v = max(v, (unsigned)__shfl_xor((int)v, o, kWave));
    return v;
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v += (unsigned)__shfl_xor((int)v, o, kWave);
    return v;
}

__device__ __forceinline__ bool lane0() { return (threadIdx.x & (kWave - 1)) == 0; }

// Logical workgroup index: runs of kXcdChunk consecutive tiles go to one
// XCD, and consecutive runs to consecutive XCDs.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned b = blockIdx.x, nb = gridDim.x;
    const unsigned full = nb / (kXcds * kXcdChunk) * (kXcds * kXcdChunk);
    if (b >= full)
        return b; // the last, partial round keeps dispatch order
    const unsigned xcd = b % kXcds, local = b / kXcds;
    return ((local / kXcdChunk) * kXcds + xcd) * kXcdChunk + local % kXcdChunk;
}

// Raise a running maximum.  Every wave of a launch reports to the same
// address, so the atomic is issued only when the value would change it:
// same-address atomics serialise in L2, and the maximum settles after the
// first few waves.
__device__ __forceinline__ void raise_max(unsigned long long *p, unsigned long long v)
{
    if (v > *(volatile unsigned long long *)p)
        atomicMax(p, v);
}

// Count-only record of a particle's neighbourhood (overflow guard for the
// reference's 400-slot half lists, and the largest full set seen).
__device__ __forceinline__ void report_neighbors(unsigned full, unsigned half,
                                                 sph_dev_counters *C)
{
    if (half > SPH_REF_MAX_NEIGHBORS)
        atomicAdd(&C->neighbor_overflow, 1ull);
    unsigned m = wave_max_u32(full);
    if (lane0())
        raise_max(&C->max_neighbors, m);
}

// A scene's slot range is padded to whole workgroups; the padding slots are
// counted into the scene's padding cell by its first workgroup, so the
// scan places every scene's particles at its fixed slot offset.
__device__ __forceinline__ void add_pad(const sph_dev_params &P, unsigned s0, unsigned *cnt)
{
    if (threadIdx.x == 0 && s0 == P.poff && P.pad)
        atomicAdd(&cnt[P.cell_base + P.ncells], P.pad);
}

// ---------------------------------------------------------------------------
// Tiles: the candidate ranges of one workgroup of sorted particles.
struct TilePlan {
    int ok;           // 1: staged in LDS; 0: read candidates from global memory
    int rows_ok;      // 1 (when not ok): each candidate row (dy) fits on its own
    unsigned jA;      // first cell row of the group
    unsigned g0[6];   // segment sorted begin, segment = (row - jA) * 3 + dy + 1
    unsigned len[6];  // segment length
    int off[6];       // LDS index = k + off[segment]
    unsigned tb[6];   // LDS index of the segment's first record (~0u: no segment)
    unsigned total;   // staged records
    unsigned sv[12];  // the segments' bounding cell starts, read by 12 threads
};

// Slot of this lane's particle within its cell `key` for the counting sort.
// The wavefront's live lanes are grouped by key first (ballots, no memory),
// then each group's first lane reserves the group's slots with one atomic,
// all groups in the same instruction: one round trip per wavefront instead
// of one per particle (or per key).  Lanes of a group take consecutive slots
// in lane order; the reorder ranks a cell by particle id, so any assignment
// of distinct slots gives the same sorted order.  Call with every live lane.
// The grouping loop reads each group's key with a scalar lane read (the lane
// index is wave-uniform) and keeps only the lane's own group mask, from which
// rank, count and leader follow after the loop.
__device__ __forceinline__ unsigned wave_slot(unsigned *cnt, unsigned key)
{
    const unsigned lane = threadIdx.x & 63u;
    unsigned long long pending = __ballot(1);
    unsigned long long mine = 0; // live lanes with this lane's key
    while (pending) {
        const int first = __ffsll((long long)pending) - 1;
        const unsigned k = (unsigned)__builtin_amdgcn_readlane((int)key, first);
        const unsigned long long same = __ballot(key == k);
        mine = key == k ? same : mine;
        pending &= ~same;
    }
    const int leader = __ffsll((long long)mine) - 1; // the group's lowest lane
    const unsigned rank = (unsigned)__popcll(mine & ((1ull << lane) - 1ull));
    unsigned base = 0;
    if ((int)lane == leader)
        base = atomicAdd(&cnt[key], (unsigned)__popcll(mine));
    return __shfl(base, leader, 64) + rank;
}

// The tile plan of the sorted particles [s0, s1) of a workgroup, given the
// keys of s0 and s1 - 1 (the caller reads them early): the segments of the
// candidate rows of its (at most two) cell rows, i.e. for each row the
// sorted range of cells c0-1 .. c1+1.  Made by the whole workgroup: threads
// 0..11 read one segment-bounding cell start each, and thread 0 assembles
// the plan from those with no memory access of its own.  Ends with a
// barrier.  The group is staged when its rows hold at most `cap` records.
__device__ void plan_tile_par(unsigned key_first, unsigned key_last,
                              const unsigned *__restrict__ start, const sph_dev_params &P,
                              TilePlan &T, unsigned cap)
{
    const unsigned sx = P.sx, sy = P.sy;
    const unsigned kf = key_first - P.cell_base, kl = key_last - P.cell_base;
    const unsigned jA = kf / sx, jB = kl / sx;
    const unsigned t = threadIdx.x;
    if (t < 12u && jB <= jA + 1) {
        const unsigned q = t / 2u, rho = q / 3u;
        const long r = (long)(jA + rho) + (long)(q % 3u) - 1;
        if (rho <= jB - jA && r >= 0 && r < (long)sy) {
            const unsigned j = jA + rho;
            const unsigned lo = rho == 0 ? kf % sx : 0;
            const unsigned hi = j == jB ? kl % sx : sx - 1;
            const unsigned c0 = lo > 0 ? lo - 1 : 0;
            const unsigned c1 = hi + 1 < sx ? hi + 1 : sx - 1;
            T.sv[t] = start[P.cell_base + (unsigned)r * sx + ((t & 1u) ? c1 + 1 : c0)];
        }
    }
    __syncthreads();
    if (t == 0) {
        T.jA = jA;
        T.ok = 0;
        T.rows_ok = 0;
        for (int q = 0; q < 6; q++) {
            T.g0[q] = 0;
            T.len[q] = 0;
            T.off[q] = 0;
            T.tb[q] = ~0u;
        }
        T.total = 0;
        if (jB <= jA + 1) {
            unsigned total = 0;
            for (unsigned rho = 0; rho <= jB - jA; rho++)
                for (int dy = -1; dy <= 1; dy++) {
                    const long r = (long)(jA + rho) + dy;
                    if (r < 0 || r >= (long)sy)
                        continue;
                    const int q = (int)rho * 3 + dy + 1;
                    const unsigned a = T.sv[2 * q], b = T.sv[2 * q + 1];
                    T.g0[q] = a;
                    T.len[q] = b - a;
                    T.off[q] = (int)total - (int)a;
                    T.tb[q] = total;
                    total += b - a;
                }
            T.ok = total <= cap;
            T.total = total;
            T.rows_ok = T.len[0] + T.len[3] <= cap && T.len[1] + T.len[4] <= cap &&
                        T.len[2] + T.len[5] <= cap;
        }
    }
    __syncthreads();
}

// LDS storage of a tile of CAPACITY records (plus the walk's padding): as
// the records are, or - float4 records, SPH_SPLIT_TILES - as two planes of
// float2, (x, y) and (z, w), in the same bytes, so that the candidate tests
// read 8-byte positions at an 8-byte stride.
#ifndef SPH_SPLIT_TILES
#define SPH_SPLIT_TILES 1
#endif
template <class R>
struct TileMem {
    using Rec = R;
    R *p;
    __device__ void put(unsigned i, R v) const { p[i] = v; }
    __device__ R get(unsigned i) const { return p[i]; }
};
struct TileMemSplit {
    using Rec = float4;
    float2 *xy, *zw;
    __device__ void put(unsigned i, float4 v) const
    {
        xy[i] = make_float2(v.x, v.y);
        zw[i] = make_float2(v.z, v.w);
    }
    __device__ float4 get(unsigned i) const
    {
        const float2 a = xy[i], b = zw[i];
        return make_float4(a.x, a.y, b.x, b.y);
    }
};
template <int CAPACITY, class R>
__device__ __forceinline__ TileMem<R> tile_mem(R *tile)
{
    return TileMem<R>{tile};
}
template <int CAPACITY>
__device__ __forceinline__ auto tile_mem(float4 *tile)
{
    if constexpr (SPH_SPLIT_TILES) {
        float2 *xy = reinterpret_cast<float2 *>(tile);
        return TileMemSplit{xy, xy + (CAPACITY + kSphWalkPad)};
    } else {
        return TileMem<float4>{tile};
    }
}

// Copy the planned segments into LDS (all threads), then barrier.  The
// segments lie back to back in the tile, so thread t stages tile records
// t, t + 256, ...: it finds each record's segment (the last one beginning at
// or before it), issues all of its loads, and only then stores them - one
// memory round trip per workgroup, not one per segment.
template <int CAPACITY, class Mem, class Load>
__device__ __forceinline__ void stage_tile(const TilePlan &T, const Mem &tile, Load load)
{
    using Rec = typename Mem::Rec;
    constexpr int kPer = (CAPACITY + kTPB - 1) / kTPB;
    const unsigned total = T.total;
    Rec v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const unsigned i = threadIdx.x + (unsigned)u * kTPB;
        if (i < total) {
            int off = T.off[0];
#pragma unroll
            for (int q = 1; q < 6; q++)
                off = T.tb[q] <= i ? T.off[q] : off;
            v[u] = load((unsigned)((int)i - off));
        }
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const unsigned i = threadIdx.x + (unsigned)u * kTPB;
        if (i < total)
            tile.put(i, v[u]);
    }
    __syncthreads();
}

// Stage only candidate row dy (segments dy + 1 of the first cell row and
// dy + 4 of the second, back to back) into the tile, then barrier: for groups
// whose rows fit one at a time but not all together.
template <int CAPACITY, class Mem, class Load>
__device__ __forceinline__ void stage_row(const TilePlan &T, int dy, const Mem &tile, Load load)
{
    using Rec = typename Mem::Rec;
    constexpr int kPer = (CAPACITY + kTPB - 1) / kTPB;
    const int qa = dy + 1, qb = dy + 4;
    const unsigned la = T.len[qa], total = la + T.len[qb];
    const unsigned ga = T.g0[qa], gb = T.g0[qb];
    Rec v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const unsigned i = threadIdx.x + (unsigned)u * kTPB;
        if (i < total)
            v[u] = load(i < la ? ga + i : gb + (i - la));
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const unsigned i = threadIdx.x + (unsigned)u * kTPB;
        if (i < total)
            tile.put(i, v[u]);
    }
    __syncthreads();
}

// Candidate source reading the LDS tile: the tile index of sorted position
// k of row dy is k + row(dy).
template <class Mem>
struct LdsSrc {
    Mem tile;
    int offA[3], offB[3];
    bool rowB;
    __device__ int row(int dy) const { return rowB ? offB[dy + 1] : offA[dy + 1]; }
    __device__ typename Mem::Rec at(unsigned idx) const { return tile.get(idx); }
};

template <class Mem>
__device__ __forceinline__ LdsSrc<Mem> lds_src(const Mem &tile, const TilePlan &T, unsigned key,
                                              const sph_dev_params &P)
{
    LdsSrc<Mem> r;
    r.tile = tile;
    for (int q = 0; q < 3; q++) {
        r.offA[q] = T.off[q];
        r.offB[q] = T.off[3 + q];
    }
    r.rowB = (key - P.cell_base) / P.sx != T.jA;
    return r;
}

// The source of a tile staged by stage_row(dy): only row dy's offsets matter.
template <class Mem>
__device__ __forceinline__ LdsSrc<Mem> lds_row_src(const Mem &tile, const TilePlan &T, int dy,
                                                  unsigned key, const sph_dev_params &P)
{
    LdsSrc<Mem> r;
    r.tile = tile;
    const int qa = dy + 1, qb = dy + 4;
    for (int q = 0; q < 3; q++) {
        r.offA[q] = -(int)T.g0[qa];
        r.offB[q] = (int)T.len[qa] - (int)T.g0[qb];
    }
    r.rowB = (key - P.cell_base) / P.sx != T.jA;
    return r;
}

// Candidate source reading global memory.
template <class Rec, class Load>
struct GlobalSrc {
    Load load;
    __device__ int row(int) const { return 0; }
    __device__ Rec at(unsigned idx) const { return load(idx); }
};

template <class Rec, class Load>
__device__ __forceinline__ GlobalSrc<Rec, Load> global_src(Load l)
{
    return GlobalSrc<Rec, Load>{l};
}

// Walk a workgroup's candidates from its LDS tile (CAPACITY records) if the
// plan staged one, else from global memory (same order, same results).  Every
// thread of the workgroup calls this (the staging and its barrier need all of
// them); `live` lanes walk.
template <bool OWN, bool CAP, int CAPACITY, int LP = 1, class Rec, class Load, class Test,
          class Body, class Masks>
__device__ __forceinline__ uint2 walk_group(const TilePlan &T, Rec *tile, Load load, bool live,
                                            unsigned self, const sph_rows &R, unsigned key,
                                            const sph_dev_params &P, Test &test, Body &body,
                                            Masks &mk)
{
    uint2 fh = make_uint2(0u, 0u);
    const auto mem = tile_mem<CAPACITY>(tile);
    if (T.ok) {
        stage_tile<CAPACITY>(T, mem, load);
        if (live)
            fh = sph_walk_masked<OWN, CAP, LP>(self, R, lds_src(mem, T, key, P), P.h2,
                                               P.bucket_cap, test, body, mk);
    } else if (T.rows_ok) {
        // one candidate row at a time: same chunks, same order
        unsigned ci = 0;
        auto pass = [&](auto row) {
            constexpr int dy = decltype(row)::value;
            stage_row<CAPACITY>(T, dy, mem, load);
            if (live) {
                const uint2 r = sph_walk_rows<OWN, CAP, dy, dy, LP>(
                    self, R, lds_row_src(mem, T, dy, key, P), P.h2, P.bucket_cap, test, body,
                    mk, ci);
                fh.x += r.x;
                fh.y += r.y;
            }
            __syncthreads(); // the next row overwrites the tile
        };
        pass(std::integral_constant<int, -1>{});
        pass(std::integral_constant<int, 0>{});
        pass(std::integral_constant<int, 1>{});
    } else if (live) {
        fh = sph_walk_masked<OWN, CAP, LP>(self, R, global_src<Rec>(load), P.h2, P.bucket_cap,
                                           test, body, mk);
    }
    return fh;
}

} // namespace

// ---------------------------------------------------------------------------
// Gravity + viscosity + prediction (src/fluid.c:398-413, 416-478, 507-523)
//
// GRAV: apply v_y += (-g)*dt first (to self and, on the fly, to every
//       neighbour read, so the result equals a separate gravity pass).
// VISC: gather the viscosity impulses over the neighbour set of the last
//       hash (positions have not moved since, so the set is the
//       reference's half lists seen from both ends).
// PRED: x_prev = x; x += v*dt; boundary; then the new cell key and a slot
//       in that cell (histogram of the next sort).  Without PRED the
//       post-viscosity velocity is written to out_pos.
template <bool GRAV, bool VISC, bool PRED, bool CAP, int LP = 1>
__device__ __forceinline__ void
visc_predict_tile(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                  const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
                  const unsigned *__restrict__ K, const unsigned *__restrict__ start,
                  float2 *__restrict__ out_pos, unsigned *__restrict__ out_key,
                  unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
                  sph_dev_counters *__restrict__ C, const unsigned long long *__restrict__ clock,
                  float2 *__restrict__ held_vel)
{

    __shared__ TilePlan T;
    __shared__ float4 tile[VISC ? kTileCap + kSphWalkPad : 1];
    constexpr unsigned kPer = kTPB / LP; // particles per workgroup
    const unsigned blk = xcd_block();
    const unsigned s0 = blk * kPer;
    const unsigned s = s0 + threadIdx.x / LP;
    // loads that need no scene parameter go out before the parameters (every
    // slot of the grid is inside the arrays; a value is used where its slot
    // is live)
    const float2 a0 = POS[s], av0 = VEL[s];
    const unsigned ks = VISC ? K[s] : 0u;
    const unsigned kf0 = VISC ? K[s0] : 0u, kl0 = VISC ? K[s0 + kPer - 1] : 0u;
    const sph_dev_params P = Pp[BS[blk / LP]];
    const unsigned s1 = min(s0 + kPer, P.poff + (unsigned)P.n);
    const bool lead = threadIdx.x % LP == 0; // the lane that stores the particle's results
    const float gdt = P.g_neg * P.dt;
    if (PRED)
        add_pad(P, s0, cnt);
    if (s0 >= s1)
        return; // a padding workgroup
    if (PRED && P.frozen) {
        // frozen scene: same position, same cell; its velocity waits in
        // held_vel for the relaxation kernel (the sort of a frozen scene is
        // the identity)
        if (s < s1 && lead) {
            const float2 a = POS[s];
            out_pos[s] = a;
            held_vel[s] = VEL[s];
            const unsigned key = sph_cell_key(a.x, a.y, P, &C->key_out_of_range);
            out_key[s] = key;
            out_slot[s] = wave_slot(cnt, key);
        }
        return;
    }
    // sloshing: horizontal box acceleration, applied like gravity
    const bool slosh = GRAV && P.slosh_amp != 0.0f;
    const float sdt = slosh ? sph_slosh_accel(P, *clock) * P.dt : 0.0f;
    // the lane's own particle and candidate rows, loaded before the planning
    // barriers
    const float2 a = a0, av = av0;
    const bool live = s < s1;
    const unsigned key = VISC && live ? ks : P.cell_base; // pad lanes: any valid cell
    const sph_rows R = VISC ? sph_row_bounds(key, P, start) : sph_rows{};
    const float avx = slosh ? av.x + sdt : av.x;
    const float avy = GRAV ? av.y + gdt : av.y;
    float nvx = avx, nvy = avy;
    unsigned full = 0, half = 0;
    if (VISC) {
        // a neighbour's record carries its velocity after gravity and
        // sloshing (the same expressions as the particle's own), applied
        // once per record rather than once per pair
        auto load = [&](unsigned g) {
            const float2 p = POS[g], v = VEL[g];
            return make_float4(p.x, p.y, slosh ? v.x + sdt : v.x, GRAV ? v.y + gdt : v.y);
        };
        auto test = [&](float4 b) {
            const float dx = b.x - a.x, ddy = b.y - a.y;
            return dx * dx + ddy * ddy;
        };
        auto body = [&](unsigned, bool, float4 b) {
            float ix, iy;
            const bool push = sph_visc_impulse(a.x, a.y, avx, avy, b.x, b.y, b.z, b.w, P, ix, iy);
            nvx = push ? nvx - ix * 0.5f : nvx;
            nvy = push ? nvy - iy * 0.5f : nvy;
        };
        plan_tile_par(kf0, s1 - s0 == kPer ? kl0 : K[s1 - 1], start, P, T, kTileCap);
        sph_masks_none mk;
        uint2 fh;
        if constexpr (LP == 1) {
            fh = walk_group<false, CAP, kTileCap>(T, tile, load, live, s, R, key, P, test, body,
                                                  mk);
        } else {
            // the same impulse as a term (zero when not pushed: nv - 0 == nv)
            auto term = [&](unsigned, bool, float4 b) {
                float ix, iy;
                const bool push =
                    sph_visc_impulse(a.x, a.y, avx, avy, b.x, b.y, b.z, b.w, P, ix, iy);
                return sph_pair_term{push ? ix : 0.0f, push ? iy : 0.0f, 0};
            };
            auto apply = [&](const sph_pair_term &t) {
                nvx = nvx - t.a * 0.5f;
                nvy = nvy - t.b * 0.5f;
            };
            sph_split_body<decltype(term), decltype(apply)> sb{term, apply};
            fh = walk_group<false, CAP, kTileCap, LP>(T, tile, load, live, s, R, key, P, test, sb,
                                                      mk);
        }
        full = fh.x;
        half = lead ? fh.y : 0u; // a particle's half list counted once
    }
    if (live && lead) {
        if (PRED) {
            float x = a.x + nvx * P.dt;
            float y = a.y + nvy * P.dt;
            sph_boundary(x, y, P);
            out_pos[s] = make_float2(x, y);
            const unsigned nk = sph_cell_key(x, y, P, &C->key_out_of_range);
            out_key[s] = nk;
            out_slot[s] = wave_slot(cnt, nk);
        } else {
            out_pos[s] = make_float2(nvx, nvy);
        }
    }
    if (VISC)
        report_neighbors(full, half, C);
}

template <bool GRAV, bool VISC, bool PRED, bool CAP>
__global__ void __launch_bounds__(kTPB, kGatherMinBlocks)
k_visc_predict(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
               const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
               const unsigned *__restrict__ K, const unsigned *__restrict__ start,
               float2 *__restrict__ out_pos, unsigned *__restrict__ out_key,
               unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
               sph_dev_counters *__restrict__ C, const unsigned long long *__restrict__ clock,
               float2 *__restrict__ held_vel)
{
    visc_predict_tile<GRAV, VISC, PRED, CAP, 1>(Pp, BS, POS, VEL, K, start, out_pos, out_key,
                                                out_slot, cnt, C, clock, held_vel);
}

// The fused step's variant with four lanes per particle (sph_gather_lanes).
template <bool CAP>
__global__ void __launch_bounds__(kTPB, 1)
k_visc_predict_quad(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                    const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
                    const unsigned *__restrict__ K, const unsigned *__restrict__ start,
                    float2 *__restrict__ out_pos, unsigned *__restrict__ out_key,
                    unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
                    sph_dev_counters *__restrict__ C, const unsigned long long *__restrict__ clock,
                    float2 *__restrict__ held_vel)
{
    visc_predict_tile<true, true, true, CAP, 4>(Pp, BS, POS, VEL, K, start, out_pos, out_key,
                                                out_slot, cnt, C, clock, held_vel);
}

// src/fluid.c:398-413 as a separate pass (phase API).
__global__ void __launch_bounds__(kTPB)
k_gravity(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
          float2 *__restrict__ VEL, float2 *__restrict__ RHO)
{
    const unsigned blk = xcd_block();
    const sph_dev_params P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    VEL[s].y += P.g_neg * P.dt;
    RHO[s] = make_float2(0.0f, 0.0f);
}

// src/fluid.c:507-523 as a separate, in-place pass (phase API).
__global__ void __launch_bounds__(kTPB)
k_predict(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
          float2 *__restrict__ POS, const float2 *__restrict__ VEL, float2 *__restrict__ PREV)
{
    const unsigned blk = xcd_block();
    const sph_dev_params P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    const float2 p = POS[s], v = VEL[s];
    PREV[s] = p;
    float x = p.x + v.x * P.dt;
    float y = p.y + v.y * P.dt;
    sph_boundary(x, y, P);
    POS[s] = make_float2(x, y);
}

// src/fluid.c:642-653 as a separate, in-place pass (phase API).
__global__ void __launch_bounds__(kTPB)
k_update_velocities(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                    float2 *__restrict__ POS, float2 *__restrict__ VEL,
                    const float2 *__restrict__ PREV)
{
    const unsigned blk = xcd_block();
    const sph_dev_params P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    float2 p = POS[s];
    const float2 q = PREV[s];
    sph_boundary(p.x, p.y, P);
    float vx = (p.x - q.x) / P.dt;
    float vy = (p.y - q.y) / P.dt;
    sph_check_velocity(vx, vy);
    POS[s] = p;
    VEL[s] = make_float2(vx, vy);
}

// Keys and slots of the current positions (histogram pass of a standalone
// hash, phase API; the fused step folds this into its producers).
__global__ void __launch_bounds__(kTPB)
k_key_hist(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
           const float2 *__restrict__ POS, unsigned *__restrict__ out_key,
           unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
           sph_dev_counters *__restrict__ C)
{
    const unsigned blk = xcd_block();
    const sph_dev_params P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    add_pad(P, blk * kTPB, cnt);
    if (s >= P.poff + (unsigned)P.n)
        return;
    const float2 p = POS[s];
    const unsigned key = sph_cell_key(p.x, p.y, P, &C->key_out_of_range);
    out_key[s] = key;
    out_slot[s] = wave_slot(cnt, key);
}

// ---------------------------------------------------------------------------
// Cell-count scan: start[c] = sum of cnt[0..c), over m = cells + 1 entries
// (cnt[cells] is always 0, so start[cells] = n).  Reduce-then-scan over tiles
// of kScanTile counts; the apply kernel adds the sum of all earlier tiles,
// writes the starts and clears the counts for the next sort.  Both arrays hold
// whole tiles (sph_scan_padded(m) entries, the padding's counts zero), so
// nothing is bounds-checked: a tile is two halves of 1024 counts, thread t
// holding counts [4t, 4t+4) of each half, and every wave's load or store
// covers 1 KiB of consecutive memory.
constexpr int kScanThreads = 256;
constexpr int kScanTile = 2 * 4 * kScanThreads;

__device__ __forceinline__ unsigned sum4(uint4 v) { return v.x + v.y + v.z + v.w; }

__global__ void __launch_bounds__(kScanThreads)
k_scan_reduce(const uint4 *__restrict__ cnt, unsigned *__restrict__ bsum)
{
    __shared__ unsigned red[kScanThreads / kWave];
    const uint4 *t = cnt + blockIdx.x * (kScanTile / 4);
    unsigned sum = sum4(t[threadIdx.x]) + sum4(t[threadIdx.x + kScanThreads]);
    sum = wave_sum_u32(sum);
    if (lane0())
        red[threadIdx.x / kWave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned a = 0;
        for (int i = 0; i < kScanThreads / kWave; i++)
            a += red[i];
        bsum[blockIdx.x] = a;
    }
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_apply(uint4 *__restrict__ cnt, const unsigned *__restrict__ bsum,
             uint4 *__restrict__ start)
{
    constexpr int kW = kScanThreads / kWave;
    __shared__ unsigned wpre[kW], wa[kW], wb[kW];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    uint4 *t = cnt + blockIdx.x * (kScanTile / 4);
    const uint4 a = t[threadIdx.x], b = t[threadIdx.x + kScanThreads];
    // prefix of all earlier tiles
    unsigned pre = 0;
    for (unsigned i = threadIdx.x; i < blockIdx.x; i += kScanThreads)
        pre += bsum[i];
    pre = wave_sum_u32(pre);
    // inclusive wave scans of both halves' per-thread totals
    const unsigned ta = sum4(a), tb = sum4(b);
    unsigned ia = ta, ib = tb;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned ya = (unsigned)__shfl_up((int)ia, o, kWave);
        const unsigned yb = (unsigned)__shfl_up((int)ib, o, kWave);
        ia += lane >= o ? ya : 0u;
        ib += lane >= o ? yb : 0u;
    }
    if (lane == kWave - 1) {
        wpre[w] = pre;
        wa[w] = ia;
        wb[w] = ib;
    }
    __syncthreads();
    unsigned tile_pre = 0, half_a = 0, pa = 0, pb = 0;
#pragma unroll
    for (int i = 0; i < kW; i++) {
        tile_pre += wpre[i];
        half_a += wa[i];
        pa += i < w ? wa[i] : 0u;
        pb += i < w ? wb[i] : 0u;
    }
    const unsigned ra = tile_pre + pa + (ia - ta);
    const unsigned rb = tile_pre + half_a + pb + (ib - tb);
    uint4 *o = start + blockIdx.x * (kScanTile / 4);
    o[threadIdx.x] = make_uint4(ra, ra + a.x, ra + a.x + a.y, ra + a.x + a.y + a.z);
    o[threadIdx.x + kScanThreads] = make_uint4(rb, rb + b.x, rb + b.x + b.y, rb + b.x + b.y + b.z);
    t[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
    t[threadIdx.x + kScanThreads] = make_uint4(0u, 0u, 0u, 0u);
}

// The whole scan of a small count array (sph_scan_single_fits: at most
// kScanSingle counts, whole tiles) in one workgroup: the same starts and the
// same cleared counts as k_scan_reduce and k_scan_apply, in one launch where
// they take two.  Thread t holds counts [4t, 4t + 4).
constexpr unsigned kScanSingleThreads = 1024;
constexpr unsigned kScanSingle = 4 * kScanSingleThreads;

__global__ void __launch_bounds__(kScanSingleThreads)
k_scan_single(uint4 *__restrict__ cnt, unsigned mp, uint4 *__restrict__ start)
{
    constexpr unsigned kWaves = kScanSingleThreads / kWave;
    __shared__ unsigned wsum[kWaves];
    const unsigned t = threadIdx.x, lane = t & (kWave - 1u), w = t / kWave;
    const bool live = 4u * t < mp;
    const uint4 c = live ? cnt[t] : make_uint4(0u, 0u, 0u, 0u);
    const unsigned tot = sum4(c);
    unsigned inc = tot;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned y = (unsigned)__shfl_up((int)inc, o, kWave);
        inc += lane >= (unsigned)o ? y : 0u;
    }
    if (lane == kWave - 1u)
        wsum[w] = inc;
    __syncthreads();
    unsigned r = inc - tot;
    for (unsigned i = 0; i < w; i++)
        r += wsum[i];
    if (live) {
        start[t] = make_uint4(r, r + c.x, r + c.x + c.y, r + c.x + c.y + c.z);
        cnt[t] = make_uint4(0u, 0u, 0u, 0u); // cleared for the next sort
    }
}

// Drop (id, source index) into the cell ranges in slot order.
__global__ void __launch_bounds__(kTPB)
k_scatter(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
          const unsigned *__restrict__ key, const unsigned *__restrict__ slot,
          const unsigned *__restrict__ start, const int *__restrict__ id_src,
          int *__restrict__ perm_src, int *__restrict__ perm_id, unsigned *__restrict__ key_dst)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned i = blk * kTPB + threadIdx.x;
    if (i >= P.poff + (unsigned)P.n)
        return;
    const unsigned k = key[i];
    const unsigned pos = start[k] + slot[i];
    perm_src[pos] = (int)i;
    perm_id[pos] = id_src[i];
    key_dst[pos] = k; // every position of a cell holds its key, whatever the order inside
}

// Rank each particle by id inside its cell and gather its fields to
// (cell, id) order.  The slot order within a cell is arbitrary (atomics);
// ranking restores the reference's bucket order (ascending id).  The rank
// also feeds the occupancy counters: a particle of rank 100 means its cell
// holds more than the reference's 100 bucket slots (src/fluid.c:174).
constexpr unsigned kRankBatch = 8;
constexpr unsigned kRankHalo = 64; // staged ids either side of a workgroup's positions

template <int NF>
__global__ void __launch_bounds__(kTPB)
k_reorder(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
          const unsigned *__restrict__ start, const int *__restrict__ perm_src,
          const int *__restrict__ perm_id, const unsigned *__restrict__ skey,
          sph_reorder_fields<NF> f, int *__restrict__ id_dst, sph_dev_counters *__restrict__ C)
{
    __shared__ int wid[kTPB + 2 * kRankHalo];
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s0 = blk * kTPB;
    const unsigned end = P.poff + (unsigned)P.n;
    if (s0 >= end)
        return; // a padding workgroup
    const unsigned k = s0 + threadIdx.x;
    const bool live = k < end;
    // the ids of sorted positions [w0, w1): this workgroup's and up to
    // kRankHalo either side, inside the scene, staged for the ranking
    static_assert(2 * kRankHalo <= kTPB, "two staged ids per thread");
    const unsigned w0 = s0 >= P.poff + kRankHalo ? s0 - kRankHalo : P.poff;
    const unsigned w1 = min(s0 + kTPB + kRankHalo, end);
    const unsigned g0 = w0 + threadIdx.x, g1 = g0 + kTPB;
    const int id0 = g0 < w1 ? perm_id[g0] : 0;
    const int id1 = g1 < w1 ? perm_id[g1] : 0;
    // the lane's own record, its fields and its cell's range, all in flight
    // with the staging loads, before the barrier
    int i = 0, myid = 0;
    unsigned b = 0, e = 0;
    float2 v[NF];
#pragma unroll
    for (int q = 0; q < NF; q++)
        v[q] = make_float2(0.0f, 0.0f);
    if (live) {
        i = perm_src[k];
        myid = perm_id[k];
        const unsigned c = skey[k]; // the sorted keys, written by the scatter
#pragma unroll
        for (int q = 0; q < NF; q++)
            v[q] = f.src[q][i];
        b = start[c];
        e = start[c + 1];
    }
    if (g0 < w1)
        wid[g0 - w0] = id0;
    if (g1 < w1)
        wid[g1 - w0] = id1;
    __syncthreads();
    unsigned occupancy = 0;
    if (live) {
        // rank = members of the cell with a smaller id, kRankBatch at a time
        // (indices clamped into the cell, so every read is in bounds and none
        // is predicated): from the staged ids when the cell lies inside them
        unsigned rank = 0;
        if (b >= w0 && e <= w1) {
            for (unsigned m = b; m < e; m += kRankBatch) {
                int idv[kRankBatch];
#pragma unroll
                for (int t = 0; t < kRankBatch; t++)
                    idv[t] = wid[min(m + t, e - 1u) - w0];
#pragma unroll
                for (int t = 0; t < kRankBatch; t++)
                    rank += (m + t < e && idv[t] < myid) ? 1u : 0u;
            }
        } else {
            for (unsigned m = b; m < e; m += kRankBatch) {
                int idv[kRankBatch];
#pragma unroll
                for (int t = 0; t < kRankBatch; t++)
                    idv[t] = perm_id[min(m + t, e - 1u)];
#pragma unroll
                for (int t = 0; t < kRankBatch; t++)
                    rank += (m + t < e && idv[t] < myid) ? 1u : 0u;
            }
        }
        const unsigned d = b + rank;
#pragma unroll
        for (int q = 0; q < NF; q++)
            f.dst[q][d] = v[q];
        id_dst[d] = myid;
        if (rank == SPH_REF_MAX_BUCKET)
            atomicAdd(&C->bucket_overflow, 1ull);
        occupancy = rank + 1;
    }
    const unsigned m = wave_max_u32(occupancy);
    if (lane0())
        raise_max(&C->max_cell_count, m);
}

// ---------------------------------------------------------------------------
// Density (src/hash.c:190-194, 225-229 with src/fluid.c:527-539) in gather
// form: each particle sums (1-r/h)^2 and (1-r/h)^3 over its neighbours.
template <bool CAP, int LP = 1>
__device__ __forceinline__ void
density_tile(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
             const float2 *__restrict__ POS, const unsigned *__restrict__ K,
             const unsigned *__restrict__ start, float2 *__restrict__ RHO,
             uint4 *__restrict__ masks, sph_dev_counters *__restrict__ C)
{

    __shared__ TilePlan T;
    // positions only: twice the candidates of the other gathers' float4
    // tiles in the same LDS bytes, so denser rows stay in LDS at no cost in
    // occupancy (registers allow 7 workgroups per CU either way)
    __shared__ float2 tile[2 * kTileCap + kSphWalkPad];
    constexpr unsigned kPer = kTPB / LP; // particles per workgroup
    const unsigned blk = xcd_block();
    const unsigned s0 = blk * kPer;
    const unsigned s = s0 + threadIdx.x / LP;
    // the lane's own particle and the tile's end keys, before the scene
    // parameters (every slot of the grid is inside the arrays)
    const float2 a = POS[s];
    const unsigned ks = K[s], kf0 = K[s0], kl0 = K[s0 + kPer - 1];
    const sph_dev_params P = Pp[BS[blk / LP]];
    const unsigned s1 = min(s0 + kPer, P.poff + (unsigned)P.n);
    const bool lead = threadIdx.x % LP == 0; // the lane that stores the particle's results
    if (s0 >= s1 || P.frozen)
        return; // a padding workgroup, or a frozen scene
    // the candidate rows, before the barriers
    const bool live = s < s1;
    const unsigned key = live ? ks : P.cell_base; // pad lanes: any valid cell
    const sph_rows R = sph_row_bounds(key, P, start);
    plan_tile_par(kf0, s1 - s0 == kPer ? kl0 : K[s1 - 1], start, P, T, 2 * kTileCap);
    float rho = 0.0f, rhon = 0.0f;
    auto load = [&](unsigned g) { return POS[g]; };
    auto test = [&](float2 b) {
        const float dx = a.x - b.x, ddy = a.y - b.y;
        return dx * dx + ddy * ddy;
    };
    auto body = [&](unsigned, bool, float2 b) {
        const float dx = a.x - b.x, ddy = a.y - b.y;
        const float r2 = dx * dx + ddy * ddy;
        const float ratio = sph_sqrt(r2) * P.h_recip;
        const float OmR2 = (1.0f - ratio) * (1.0f - ratio);
        const bool in = ratio < 1.0f;
        rho = in ? rho + OmR2 : rho;
        rhon = in ? rhon + OmR2 * (1.0f - ratio) : rhon;
    };
    sph_masks_record mk;
    uint2 fh;
    if constexpr (LP == 1) {
        fh = walk_group<false, CAP, 2 * kTileCap>(T, tile, load, live, s, R, key, P, test, body,
                                                  mk);
    } else {
        // the same sums as terms (zero outside h: rho + 0 == rho, as rho >= +0)
        auto term = [&](unsigned, bool, float2 b) {
            const float dx = a.x - b.x, ddy = a.y - b.y;
            const float r2 = dx * dx + ddy * ddy;
            const float ratio = sph_sqrt(r2) * P.h_recip;
            const float OmR2 = (1.0f - ratio) * (1.0f - ratio);
            const bool in = ratio < 1.0f;
            return sph_pair_term{in ? OmR2 : 0.0f, in ? OmR2 * (1.0f - ratio) : 0.0f, 0};
        };
        auto apply = [&](const sph_pair_term &t) {
            rho = rho + t.a;
            rhon = rhon + t.b;
        };
        sph_split_body<decltype(term), decltype(apply)> sb{term, apply};
        fh = walk_group<false, CAP, 2 * kTileCap, LP>(T, tile, load, live, s, R, key, P, test, sb,
                                                      mk);
    }
    if (live && lead) {
        RHO[s] = make_float2(rho, rhon);
        masks[s] = mk.m; // for the relaxation of the same step
    }
    const unsigned full = fh.x, half = lead ? fh.y : 0u;
    report_neighbors(full, half, C);
}

template <bool CAP>
__global__ void __launch_bounds__(kTPB, kGatherMinBlocks)
k_density(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
          const float2 *__restrict__ POS, const unsigned *__restrict__ K,
          const unsigned *__restrict__ start, float2 *__restrict__ RHO,
          uint4 *__restrict__ masks, sph_dev_counters *__restrict__ C)
{
    density_tile<CAP, 1>(Pp, BS, POS, K, start, RHO, masks, C);
}

// Four lanes per particle.
template <bool CAP>
__global__ void __launch_bounds__(kTPB, 1)
k_density_quad(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
               const float2 *__restrict__ POS, const unsigned *__restrict__ K,
               const unsigned *__restrict__ start, float2 *__restrict__ RHO,
               uint4 *__restrict__ masks, sph_dev_counters *__restrict__ C)
{
    density_tile<CAP, 4>(Pp, BS, POS, K, start, RHO, masks, C);
}

// Double-density relaxation (src/fluid.c:541-611) in gather form.
// FUSED adds the velocity update (src/fluid.c:642-653): boundary, then
// v = (x - x_prev)/dt clamped, and the histogram of the next sort; without
// FUSED only the relaxed positions are written.
template <bool FUSED, bool CAP, int LP = 1>
__device__ __forceinline__ void
relax_tile(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
           const float2 *__restrict__ POS, const float2 *__restrict__ PREV,
           const float2 *__restrict__ RHO, const unsigned *__restrict__ K,
           const unsigned *__restrict__ start, float2 *__restrict__ out_pos,
           float2 *__restrict__ out_vel, unsigned *__restrict__ out_key,
           unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
           const uint4 *__restrict__ masks, sph_dev_counters *__restrict__ C,
           unsigned long long *__restrict__ clock)
{

    __shared__ TilePlan T;
    __shared__ float4 tile[kTileCap + kSphWalkPad];
    constexpr unsigned kPer = kTPB / LP; // particles per workgroup
    const unsigned blk = xcd_block();
    const unsigned s0 = blk * kPer;
    const unsigned s = s0 + threadIdx.x / LP;
    // the lane's own key and the tile's end keys, before the scene
    // parameters (every slot of the grid is inside the arrays)
    const unsigned ks = K[s], kf0 = K[s0], kl0 = K[s0 + kPer - 1];
    const sph_dev_params P = Pp[BS[blk / LP]];
    const unsigned s1 = min(s0 + kPer, P.poff + (unsigned)P.n);
    const bool lead = threadIdx.x % LP == 0; // the lane that stores the particle's results
    if (FUSED) {
        add_pad(P, s0, cnt);
        if (blockIdx.x == 0 && threadIdx.x == 0)
            atomicAdd(clock, 1ull); // one step of simulated time
    }
    if (s0 >= s1)
        return; // a padding workgroup
    if (FUSED && P.frozen) {
        // frozen scene: keep the position; out_vel already holds the
        // velocity parked by the prediction kernel at this same slot
        if (s < s1 && lead) {
            const float2 a = POS[s];
            out_pos[s] = a;
            const unsigned nk = sph_cell_key(a.x, a.y, P, &C->key_out_of_range);
            out_key[s] = nk;
            out_slot[s] = wave_slot(cnt, nk);
        }
        return;
    }
    // the lane's own particle and candidate rows, loaded before the planning
    // barriers
    const float2 a = POS[s];
    const float2 ar = RHO[s];
    const bool live = s < s1;
    const unsigned key = live ? ks : P.cell_base; // pad lanes: any valid cell
    const sph_rows R = sph_row_bounds(key, P, start);
    plan_tile_par(kf0, s1 - s0 == kPer ? kl0 : K[s1 - 1], start, P, T, kTileCap);
    // a neighbour's record carries its pressure and near-pressure
    // (src/fluid.c:563-564), evaluated once per record rather than per pair
    auto load = [&](unsigned g) {
        const float2 p = POS[g], r = RHO[g];
        return make_float4(p.x, p.y, P.k * (r.x - P.rest_density), P.k_near * r.y);
    };
    const float aP = P.k * (ar.x - P.rest_density);
    const float aPn = P.k_near * ar.y;
    float nx = a.x, ny = a.y;
    auto test = [&](float4 b) {
        const float dx = a.x - b.x, ddy = a.y - b.y;
        return dx * dx + ddy * ddy;
    };
    auto body = [&](unsigned, bool own, float4 b) {
        const float dx = a.x - b.x, ddy = a.y - b.y;
        const float r2 = dx * dx + ddy * ddy;
        const float rr = sph_sqrt(r2);
        const float r_recip = sph_rcp(rr); // 1.0f / rr
        const float ratio = rr * P.h_recip;
        const float OmR = 1.0f - ratio;
        // both updates as selects (no branch in the pair loop)
        const bool nudge = rr <= 0.000001f && own;
        nx = nudge ? nx + 0.000001f : nx;
        ny = nudge ? ny + 0.000001f : ny;
        const float D = sph_relax_D(P.dt, aP + b.z, aPn + b.w, OmR, P.k_spring, P.h, rr);
        // b - a = -(a - b) exactly in IEEE arithmetic
        const float Dx = D * -dx * r_recip;
        const float Dy = D * -ddy * r_recip;
        const bool push = ratio < 1.0f && rr > 0.0f;
        nx = push ? nx - Dx : nx;
        ny = push ? ny - Dy : ny;
    };
    if constexpr (LP == 1) {
        if (FUSED) { // the density pass of this step walked the same neighbour set
            sph_masks_replay mk{live ? masks[s] : make_uint4(0u, 0u, 0u, 0u)};
            walk_group<true, CAP, kTileCap>(T, tile, load, live, s, R, key, P, test, body, mk);
        } else {
            sph_masks_none mk;
            walk_group<true, CAP, kTileCap>(T, tile, load, live, s, R, key, P, test, body, mk);
        }
    } else {
        // the same updates as terms: the displacement (zero when not pushed:
        // n - 0 == n) and the nudge flag, applied nudge first as in body
        auto term = [&](unsigned, bool own, float4 b) {
            const float dx = a.x - b.x, ddy = a.y - b.y;
            const float r2 = dx * dx + ddy * ddy;
            const float rr = sph_sqrt(r2);
            const float r_recip = sph_rcp(rr); // 1.0f / rr
            const float ratio = rr * P.h_recip;
            const float OmR = 1.0f - ratio;
            const bool nudge = rr <= 0.000001f && own;
            const float D = sph_relax_D(P.dt, aP + b.z, aPn + b.w, OmR, P.k_spring, P.h, rr);
            const float Dx = D * -dx * r_recip;
            const float Dy = D * -ddy * r_recip;
            const bool push = ratio < 1.0f && rr > 0.0f;
            return sph_pair_term{push ? Dx : 0.0f, push ? Dy : 0.0f, nudge ? 1 : 0};
        };
        auto apply = [&](const sph_pair_term &t) {
            nx = t.f ? nx + 0.000001f : nx;
            ny = t.f ? ny + 0.000001f : ny;
            nx = nx - t.a;
            ny = ny - t.b;
        };
        sph_split_body<decltype(term), decltype(apply)> sb{term, apply};
        if (FUSED) {
            sph_masks_replay mk{live ? masks[s] : make_uint4(0u, 0u, 0u, 0u)};
            walk_group<true, CAP, kTileCap, LP>(T, tile, load, live, s, R, key, P, test, sb, mk);
        } else {
            sph_masks_none mk;
            walk_group<true, CAP, kTileCap, LP>(T, tile, load, live, s, R, key, P, test, sb, mk);
        }
    }
    if (!live || !lead)
        return;
    if (FUSED) {
        sph_boundary(nx, ny, P);
        const float2 q = PREV[s];
        float vx = (nx - q.x) / P.dt;
        float vy = (ny - q.y) / P.dt;
        sph_check_velocity(vx, vy);
        out_vel[s] = make_float2(vx, vy);
        if (!(isfinite(nx) && isfinite(ny)))
            atomicAdd(&C->nonfinite, 1ull);
        const unsigned nk = sph_cell_key(nx, ny, P, &C->key_out_of_range);
        out_key[s] = nk;
        out_slot[s] = wave_slot(cnt, nk);
    }
    out_pos[s] = make_float2(nx, ny);
}

template <bool FUSED, bool CAP>
__global__ void __launch_bounds__(kTPB, CAP ? kGatherMinBlocks : kGatherMinBlocks + 1)
k_relax(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
        const float2 *__restrict__ POS, const float2 *__restrict__ PREV,
        const float2 *__restrict__ RHO, const unsigned *__restrict__ K,
        const unsigned *__restrict__ start, float2 *__restrict__ out_pos,
        float2 *__restrict__ out_vel, unsigned *__restrict__ out_key,
        unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
        const uint4 *__restrict__ masks, sph_dev_counters *__restrict__ C,
        unsigned long long *__restrict__ clock)
{
    relax_tile<FUSED, CAP, 1>(Pp, BS, POS, PREV, RHO, K, start, out_pos, out_vel, out_key, out_slot,
                              cnt, masks, C, clock);
}

// The fused variant with four lanes per particle.
template <bool CAP>
__global__ void __launch_bounds__(kTPB, 1)
k_relax_quad(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
             const float2 *__restrict__ POS, const float2 *__restrict__ PREV,
             const float2 *__restrict__ RHO, const unsigned *__restrict__ K,
             const unsigned *__restrict__ start, float2 *__restrict__ out_pos,
             float2 *__restrict__ out_vel, unsigned *__restrict__ out_key,
             unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
             const uint4 *__restrict__ masks, sph_dev_counters *__restrict__ C,
             unsigned long long *__restrict__ clock)
{
    relax_tile<true, CAP, 4>(Pp, BS, POS, PREV, RHO, K, start, out_pos, out_vel, out_key, out_slot,
                             cnt, masks, C, clock);
}

// Reference coordinate packing (src/fluid.c:356-363), written by id.
__global__ void __launch_bounds__(kTPB)
k_pack_coords(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
              const float2 *__restrict__ POS, const int *__restrict__ ID,
              short *__restrict__ out)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    const size_t id = P.uoff + (unsigned)ID[s];
    const float2 p = POS[s];
    out[2 * id] = (short)((2.0f * p.x / P.max_x - 1.0f) * 32767.0f);
    out[2 * id + 1] = (short)((2.0f * p.y / P.max_y - 1.0f) * 32767.0f);
}

// Neighbour sets of the current hash, written by id (small-N read-back for
// tests): full set in canonical order, and the half-list length the
// reference would hold for each particle.
__global__ void __launch_bounds__(kTPB)
k_neighbor_lists(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                 const float2 *__restrict__ POS, const unsigned *__restrict__ K,
                 const int *__restrict__ ID, const unsigned *__restrict__ start, int stride,
                 int *__restrict__ full_by_id, int *__restrict__ half_by_id,
                 int *__restrict__ ids_by_id)
{
    const unsigned blk = xcd_block();
    const sph_dev_params P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    const int me = (int)P.uoff + ID[s];
    const float2 a = POS[s];
    int full = 0, half = 0;
    // the bucket cap: rank in the cell = sorted position - the cell's start
    auto dropped = [&](unsigned k) { return P.bucket_cap && k - start[K[k]] >= P.bucket_cap; };
    if (!dropped(s))
        sph_walk(s, K[s], P, start, global_src<float2>([&](unsigned g) { return POS[g]; }),
                 [&](unsigned k, int dy, unsigned b1, unsigned b2, float2 b) {
                     const float dx = a.x - b.x, ddy = a.y - b.y;
                     if (k == s || dx * dx + ddy * ddy > P.h2 || dropped(k))
                         return;
                     if (full < stride)
                         ids_by_id[(size_t)me * stride + full] = ID[k];
                     full++;
                     half += sph_owns(k, s, dy, b1, b2);
                 });
    full_by_id[me] = full;
    half_by_id[me] = half;
}

// dst[pos_of_id[id_src[k]]] = src[k]: re-canonicalise auxiliary fields.
__global__ void __launch_bounds__(kTPB)
k_pos_of_id(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
            const int *__restrict__ ID, int *__restrict__ pos_of_id)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s < P.poff + (unsigned)P.n)
        pos_of_id[P.uoff + (unsigned)ID[s]] = (int)s;
}

__global__ void __launch_bounds__(kTPB)
k_permute_by_id(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                const int *__restrict__ id_src, const int *__restrict__ pos_of_id,
                sph_reorder_fields<2> f)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned k = blk * kTPB + threadIdx.x;
    if (k >= P.poff + (unsigned)P.n)
        return;
    const int d = pos_of_id[P.uoff + (unsigned)id_src[k]];
#pragma unroll
    for (int q = 0; q < 2; q++)
        f.dst[q][d] = f.src[q][k];
}

// ---------------------------------------------------------------------------
// Host-side launchers.  `slots` is the padded slot count of all scenes
// (a multiple of kTPB); BS maps each logical workgroup to its scene.

#define GRID(slots) dim3((unsigned)((slots) / kTPB))

namespace {

// Start/stop events for the next launch (sph_set_next_launch_events).
thread_local hipEvent_t g_next_start = nullptr, g_next_stop = nullptr;

// Launch on `st`; when timing events are pending, through
// hipExtLaunchKernelGGL, whose events carry the kernel's own start and end
// times (no dispatch latency).
template <class F, class... Args>
void launch(F kernel, dim3 grid, dim3 block, hipStream_t st, Args... args)
{
    if (g_next_start) {
        hipEvent_t a = g_next_start, b = g_next_stop;
        g_next_start = g_next_stop = nullptr;
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, a, b, 0, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    }
}

} // namespace

// Exhaustive checks of sph_sqrt (FN 0) against sqrtf and sph_rcp (FN 1)
// against IEEE division: every float with bit pattern in [lo, hi)
// (grid-stride; for sph_rcp, every one in its domain), bit for bit (any NaN
// equals any NaN).
__device__ __attribute__((noinline)) float ieee_sqrt(float x) { return sqrtf(x); }
__device__ __attribute__((noinline)) float ieee_rcp(float x) { return 1.0f / x; }

template <int FN>
__global__ void __launch_bounds__(256)
k_selftest_math(unsigned lo, unsigned hi, unsigned long long *__restrict__ bad)
{
    unsigned long long nbad = 0;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long b = lo + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         b < hi; b += stride) {
        const float x = __uint_as_float((unsigned)b);
        if (FN == 1 && !sph_rcp_domain(x))
            continue;
        const float got = FN == 0 ? sph_sqrt(x) : sph_rcp(x);
        const float want = FN == 0 ? ieee_sqrt(x) : ieee_rcp(x);
        const bool same = __float_as_uint(got) == __float_as_uint(want) ||
                          (got != got && want != want);
        nbad += same ? 0u : 1u;
    }
    for (int o = 32; o > 0; o >>= 1)
        nbad += __shfl_xor(nbad, o, 64);
    if ((threadIdx.x & 63u) == 0 && nbad)
        atomicAdd(bad, nbad);
}

void sph_launch_selftest_math(hipStream_t st, int fn, unsigned lo, unsigned hi,
                              unsigned long long *bad)
{
    if (fn == 0)
        launch(k_selftest_math<0>, dim3(4096), dim3(256), st, lo, hi, bad);
    else
        launch(k_selftest_math<1>, dim3(4096), dim3(256), st, lo, hi, bad);
}

void sph_set_next_launch_events(hipEvent_t start, hipEvent_t stop)
{
    g_next_start = start;
    g_next_stop = stop;
}

__global__ void __launch_bounds__(kTPB)
k_restore(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
          const int *__restrict__ flags, sph_restore_args a)
{
    const unsigned blk = blockIdx.x;
    const int scene = BS[blk];
    if (!flags[scene])
        return;
    const sph_dev_params &P = Pp[scene];
    const unsigned s = blk * kTPB + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++)
        a.dst[q][s] = a.src[q][s];
    a.id_dst[s] = a.id_src[s];
    a.key_dst[s] = a.key_src[s];
    if (blk * kTPB == P.poff)
        for (unsigned c = threadIdx.x; c <= P.ncells; c += kTPB)
            a.start_dst[P.cell_base + c] = a.start_src[P.cell_base + c];
}

__global__ void __launch_bounds__(kTPB)
k_gather_state(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
               const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
               const int *__restrict__ ID, float4 *__restrict__ out)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    const float2 p = POS[s], v = VEL[s];
    out[P.uoff + (unsigned)ID[s]] = make_float4(p.x, p.y, v.x, v.y);
}

// The reference's particle records (fluid_particle, 13 words) by caller
// index: position, velocity, a and id from the canonical order (ID0); x_prev,
// density and the pressures (src/fluid.c:563-564) from the auxiliary fields'
// order (IDA: ID0, or ID1 after fused steps).
__global__ void __launch_bounds__(kTPB)
k_gather_aos(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
             const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
             const int *__restrict__ ID0, const float2 *__restrict__ PREV,
             const float2 *__restrict__ RHO, const int *__restrict__ IDA,
             const float2 *__restrict__ ACC, float *__restrict__ out)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    const int id = ID0[s];
    const size_t g = P.uoff + (unsigned)id;
    const float2 p = POS[s], v = VEL[s], a = ACC[g];
    float *r = out + 13 * g;
    r[2] = p.x;
    r[3] = p.y;
    r[4] = v.x;
    r[5] = v.y;
    r[6] = a.x;
    r[7] = a.y;
    r[12] = __int_as_float(id);
    const float2 q = PREV[s], rho = RHO[s];
    float *r2 = out + 13 * ((size_t)P.uoff + (unsigned)IDA[s]);
    r2[0] = q.x;
    r2[1] = q.y;
    r2[8] = rho.x;
    r2[9] = rho.y;
    r2[10] = P.k * (rho.x - P.rest_density);
    r2[11] = P.k_near * rho.y;
}

// Positions by caller index (scene offset + id), for position-only read-back.
__global__ void __launch_bounds__(kTPB)
k_gather_xy(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
            const float2 *__restrict__ POS, const int *__restrict__ ID, float2 *__restrict__ out)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    out[P.uoff + (unsigned)ID[s]] = POS[s];
}

__global__ void __launch_bounds__(kTPB)
k_load_state(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
             const float4 *__restrict__ in, float2 *__restrict__ POS, float2 *__restrict__ VEL,
             float2 *__restrict__ PREV, float2 *__restrict__ RHO, int *__restrict__ ID)
{
    const unsigned blk = xcd_block();
    const sph_dev_params &P = Pp[BS[blk]];
    const unsigned s = blk * kTPB + threadIdx.x;
    if (s >= P.poff + (unsigned)P.n)
        return;
    const unsigned id = s - P.poff;
    const float4 q = in[P.uoff + id];
    POS[s] = make_float2(q.x, q.y);
    PREV[s] = make_float2(q.x, q.y);
    VEL[s] = make_float2(q.z, q.w);
    RHO[s] = make_float2(0.0f, 0.0f);
    ID[s] = (int)id;
}

// Four lanes per particle when the context has at most two 256-slot tiles
// per CU (a single scene of up to 131,072 particles on 256 CUs).  There the
// gather launches are short of work, and a launch's time is set by its
// slowest workgroups: the dense bottom rows of a settling column, whose walks
// four lanes cut to about a quarter.  Measured on dam breaks of 4k-1M
// particles (graph steps 21-220 / 501-600, one lane against four): 4k 164 /
// 87 us; 65k 164 / 91; 131k 96 against 77 and 239 against 135; 262k 102
// against 112 and 247 against 175; 524k 121 against 183; 1M 170 against 323.
// Past the crossover the quad's repeated per-particle work and four times
// the tiles (staging, plans) cost more than the shorter walks gain.
// SPH_SMALL_SCENE_LANES=1 turns the four-lane path off.
#ifndef SPH_SMALL_SCENE_LANES
#define SPH_SMALL_SCENE_LANES 4
#endif
int sph_gather_lanes(unsigned slots)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    const unsigned long long tiles = slots / kTPB;
    return SPH_SMALL_SCENE_LANES == 4 && cus > 0 && tiles <= 2ull * (unsigned long long)cus ? 4
                                                                                             : 1;
}

void sph_launch_visc_predict(hipStream_t st, unsigned slots, bool grav, bool visc, bool pred,
                             bool cap, const sph_dev_params *P, const int *BS, const float2 *POS,
                             const float2 *VEL, const unsigned *K, const unsigned *start,
                             float2 *opos, unsigned *okey, unsigned *oslot, unsigned *cnt,
                             sph_dev_counters *C, const unsigned long long *clock,
                             float2 *held_vel, int lanes)
{
#define VP(G, V, R, CAP)                                                                         \
    launch(k_visc_predict<G, V, R, CAP>, GRID(slots), kTPB, st, P, BS, POS, VEL, K, start, opos, \
           okey, oslot, cnt, C, clock, held_vel)
#define VP4(CAP)                                                                                 \
    launch(k_visc_predict_quad<CAP>, GRID(4 * (slots)), kTPB, st, P, BS, POS,    \
           VEL, K, start, opos, okey, oslot, cnt, C, clock, held_vel)
#define VPC(G, V, R)                                                                             \
    do {                                                                                         \
        if (cap)                                                                                 \
            VP(G, V, R, true);                                                                   \
        else                                                                                     \
            VP(G, V, R, false);                                                                  \
    } while (0)
    // without viscosity the walk does not run: no capped variant; four lanes
    // per particle for the fused step's variant only
    if (grav && visc && pred && lanes == 4) {
        if (cap)
            VP4(true);
        else
            VP4(false);
    } else if (grav && visc && pred)
        VPC(true, true, true);
    else if (grav && !visc && pred)
        VP(true, false, true, false);
    else if (!grav && visc && !pred)
        VPC(false, true, false);
    else if (!grav && visc && pred)
        VPC(false, true, true);
    else if (!grav && !visc && pred)
        VP(false, false, true, false);
    else
        VPC(true, true, false);
#undef VP4
#undef VPC
#undef VP
}

void sph_launch_gravity(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        float2 *VEL, float2 *RHO)
{
    launch(k_gravity, GRID(slots), kTPB, st, P, BS, VEL, RHO);
}

void sph_launch_predict(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        float2 *POS, const float2 *VEL, float2 *PREV)
{
    launch(k_predict, GRID(slots), kTPB, st, P, BS, POS, VEL, PREV);
}

void sph_launch_update_velocities(hipStream_t st, unsigned slots, const sph_dev_params *P,
                                  const int *BS, float2 *POS, float2 *VEL, const float2 *PREV)
{
    launch(k_update_velocities, GRID(slots), kTPB, st, P, BS, POS, VEL, PREV);
}

void sph_launch_key_hist(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                         const float2 *POS, unsigned *okey, unsigned *oslot, unsigned *cnt,
                         sph_dev_counters *C)
{
    launch(k_key_hist, GRID(slots), kTPB, st, P, BS, POS, okey, oslot, cnt, C);
}

unsigned sph_scan_blocks(unsigned m)
{
    return (unsigned)(((unsigned long long)m + kScanTile - 1) / kScanTile); // m near 2^32: no wrap
}

size_t sph_scan_padded(unsigned m) { return (size_t)sph_scan_blocks(m) * kScanTile; }

void sph_launch_scan_reduce(hipStream_t st, const unsigned *cnt, unsigned m, unsigned *bsum)
{
    launch(k_scan_reduce, dim3(sph_scan_blocks(m)), kScanThreads, st,
           reinterpret_cast<const uint4 *>(cnt), bsum);
}

void sph_launch_scan_apply(hipStream_t st, unsigned *cnt, unsigned m, const unsigned *bsum,
                           unsigned *start)
{
    launch(k_scan_apply, dim3(sph_scan_blocks(m)), kScanThreads, st,
           reinterpret_cast<uint4 *>(cnt), bsum, reinterpret_cast<uint4 *>(start));
}

#ifndef SPH_SCAN_SINGLE
#define SPH_SCAN_SINGLE 1
#endif
bool sph_scan_single_fits(unsigned m)
{
    return SPH_SCAN_SINGLE && sph_scan_padded(m) <= kScanSingle;
}

void sph_launch_scan_single(hipStream_t st, unsigned *cnt, unsigned m, unsigned *start)
{
    launch(k_scan_single, dim3(1), kScanSingleThreads, st, reinterpret_cast<uint4 *>(cnt),
           (unsigned)sph_scan_padded(m), reinterpret_cast<uint4 *>(start));
}

void sph_launch_scatter(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        const unsigned *key, const unsigned *slot, const unsigned *start,
                        const int *id_src, int *perm_src, int *perm_id, unsigned *key_dst)
{
    launch(k_scatter, GRID(slots), kTPB, st, P, BS, key, slot, start, id_src, perm_src, perm_id,
           key_dst);
}

void sph_launch_reorder2(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                         const unsigned *start, const int *perm_src, const int *perm_id,
                         const unsigned *skey, sph_reorder_fields<2> f, int *id_dst,
                         sph_dev_counters *C)
{
    launch(k_reorder<2>, GRID(slots), kTPB, st, P, BS, start, perm_src, perm_id, skey, f, id_dst,
                                               C);
}

void sph_launch_reorder4(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                         const unsigned *start, const int *perm_src, const int *perm_id,
                         const unsigned *skey, sph_reorder_fields<4> f, int *id_dst,
                         sph_dev_counters *C)
{
    launch(k_reorder<4>, GRID(slots), kTPB, st, P, BS, start, perm_src, perm_id, skey, f, id_dst,
                                               C);
}

void sph_launch_density(hipStream_t st, unsigned slots, bool cap, const sph_dev_params *P,
                        const int *BS, const float2 *POS, const unsigned *K,
                        const unsigned *start, float2 *RHO, uint4 *masks, sph_dev_counters *C,
                        int lanes)
{
    if (lanes == 4 && cap)
        launch(k_density_quad<true>, GRID(4 * slots), kTPB, st, P, BS, POS, K, start, RHO, masks,
               C);
    else if (lanes == 4)
        launch(k_density_quad<false>, GRID(4 * slots), kTPB, st, P, BS, POS, K, start, RHO, masks,
               C);
    else if (cap)
        launch(k_density<true>, GRID(slots), kTPB, st, P, BS, POS, K, start, RHO, masks, C);
    else
        launch(k_density<false>, GRID(slots), kTPB, st, P, BS, POS, K, start, RHO, masks, C);
}

void sph_launch_relax(hipStream_t st, unsigned slots, bool fused, bool cap,
                      const sph_dev_params *P, const int *BS, const float2 *POS,
                      const float2 *PREV, const float2 *RHO, const unsigned *K,
                      const unsigned *start, float2 *opos, float2 *ovel, unsigned *okey,
                      unsigned *oslot, unsigned *cnt, const uint4 *masks, sph_dev_counters *C,
                      unsigned long long *clock, int lanes)
{
#define RL(F, CAP)                                                                               \
    launch(k_relax<F, CAP>, GRID(slots), kTPB, st, P, BS, POS, PREV, RHO, K, start, opos, ovel,   \
           okey, oslot, cnt, masks, C, clock)
#define RL4(CAP)                                                                                 \
    launch(k_relax_quad<CAP>, GRID(4 * (slots)), kTPB, st, P, BS, POS, PREV, RHO, K, start,  \
           opos, ovel, okey, oslot, cnt, masks, C, clock)
    // four lanes per particle for the fused step's variant only
    if (fused && lanes == 4 && cap)
        RL4(true);
    else if (fused && lanes == 4)
        RL4(false);
    else if (fused && cap)
        RL(true, true);
    else if (fused)
        RL(true, false);
    else if (cap)
        RL(false, true);
    else
        RL(false, false);
#undef RL4
#undef RL
}

void sph_launch_pack_coords(hipStream_t st, unsigned slots, const sph_dev_params *P,
                            const int *BS, const float2 *POS, const int *ID, short *out)
{
    launch(k_pack_coords, GRID(slots), kTPB, st, P, BS, POS, ID, out);
}

void sph_launch_neighbor_lists(hipStream_t st, unsigned slots, const sph_dev_params *P,
                               const int *BS, const float2 *POS, const unsigned *K, const int *ID,
                               const unsigned *start, int stride, int *full_by_id,
                               int *half_by_id, int *ids_by_id)
{
    launch(k_neighbor_lists, GRID(slots), kTPB, st, P, BS, POS, K, ID, start, stride, full_by_id,
                                                   half_by_id, ids_by_id);
}

void sph_launch_pos_of_id(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                          const int *ID, int *pos_of_id)
{
    launch(k_pos_of_id, GRID(slots), kTPB, st, P, BS, ID, pos_of_id);
}

void sph_launch_permute_by_id(hipStream_t st, unsigned slots, const sph_dev_params *P,
                              const int *BS, const int *id_src, const int *pos_of_id,
                              sph_reorder_fields<2> f)
{
    launch(k_permute_by_id, GRID(slots), kTPB, st, P, BS, id_src, pos_of_id, f);
}

void sph_launch_restore(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        const int *flags, sph_restore_args a)
{
    launch(k_restore, dim3(slots / kTPB), kTPB, st, P, BS, flags, a);
}

void sph_launch_gather_state(hipStream_t st, unsigned slots, const sph_dev_params *P,
                             const int *BS, const float2 *POS, const float2 *VEL, const int *ID,
                             float4 *out)
{
    launch(k_gather_state, GRID(slots), kTPB, st, P, BS, POS, VEL, ID, out);
}

void sph_launch_gather_aos(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                           const float2 *POS, const float2 *VEL, const int *ID0,
                           const float2 *PREV, const float2 *RHO, const int *IDA,
                           const float2 *ACC, float *out)
{
    launch(k_gather_aos, GRID(slots), kTPB, st, P, BS, POS, VEL, ID0, PREV, RHO, IDA, ACC, out);
}

void sph_launch_gather_xy(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                          const float2 *POS, const int *ID, float2 *out)
{
    launch(k_gather_xy, GRID(slots), kTPB, st, P, BS, POS, ID, out);
}

void sph_launch_load_state(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                           const float4 *in, float2 *POS, float2 *VEL, float2 *PREV, float2 *RHO,
                           int *ID)
{
    launch(k_load_state, GRID(slots), kTPB, st, P, BS, in, POS, VEL, PREV, RHO, ID);
}
