// sph_kernels.hip - gfx950 kernels of the TinySPH step (AdamSimpson/SPH).
//
// Data layout: one SoA particle array in HBM, kept sorted by (cell, id)
// after every hash.  Every physics kernel is a gather over the 3x3 cells
// around the particle's cell: the candidates are three contiguous row
// ranges of the sorted arrays, visited in ascending sorted order, so each
// particle sums its pair terms in one canonical order and no kernel scatters
// into another particle's state (no atomics on particle data).
//
// The cell sort is a counting sort: the producing kernel computes each
// particle's new key and claims a slot in its cell with an atomic counter,
// a two-kernel scan turns the counts into cell starts (and clears them for
// the next sort), a scatter drops (id, source) pairs into the cell ranges,
// and the reorder kernel ranks each particle by id inside its cell and
// gathers its fields there - the canonical (cell, id) order of
// src/hash.c:153-163.
#include "sph_kernels.h"

namespace {

constexpr int kWave = 64;

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

// Count-only record of a particle's neighbourhood (overflow guard for the
// reference's 400-slot half lists, and the largest full set seen).
__device__ __forceinline__ void report_neighbors(unsigned full, unsigned half,
                                                 sph_dev_counters *C)
{
    if (half > SPH_REF_MAX_NEIGHBORS)
        atomicAdd(&C->neighbor_overflow, 1ull);
    unsigned m = wave_max_u32(full);
    if (lane0())
        atomicMax(&C->max_neighbors, (unsigned long long)m);
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
//       post-viscosity velocity is written to out_vx/out_vy.
template <bool GRAV, bool VISC, bool PRED>
__global__ void __launch_bounds__(256)
k_visc_predict(const sph_dev_params *__restrict__ Pp, const float *__restrict__ X,
               const float *__restrict__ Y, const float *__restrict__ VX,
               const float *__restrict__ VY, const unsigned *__restrict__ K,
               const unsigned *__restrict__ start, float *__restrict__ out_x,
               float *__restrict__ out_y, unsigned *__restrict__ out_key,
               unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
               sph_dev_counters *__restrict__ C)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = s < (unsigned)P.n;
    const float gdt = P.g_neg * P.dt;
    unsigned full = 0, half = 0;
    if (live) {
        const float ax = X[s], ay = Y[s];
        const float avx = VX[s];
        float avy = VY[s];
        if (GRAV)
            avy += gdt;
        float nvx = avx, nvy = avy;
        if (VISC) {
            sph_rows R;
            sph_row_ranges(K[s], P, start, R);
            for (int r = 0; r < R.nr; r++) {
                for (unsigned k = R.b0[r]; k < R.b3[r]; k++) {
                    if (k == s)
                        continue;
                    const float bx = X[k], by = Y[k];
                    const float dx = bx - ax, dy = by - ay;
                    if (dx * dx + dy * dy > P.h2)
                        continue;
                    full++;
                    half += sph_owns(k, s, R.dy[r], R.b1[r], R.b2[r]);
                    float bvy = VY[k];
                    if (GRAV)
                        bvy += gdt;
                    float ix, iy;
                    if (sph_visc_impulse(ax, ay, avx, avy, bx, by, VX[k], bvy, P, ix, iy)) {
                        nvx -= ix * 0.5f;
                        nvy -= iy * 0.5f;
                    }
                }
            }
        }
        if (PRED) {
            float x = ax + nvx * P.dt;
            float y = ay + nvy * P.dt;
            sph_boundary(x, y, P);
            out_x[s] = x;
            out_y[s] = y;
            const unsigned key = sph_cell_key(x, y, P, &C->key_out_of_range);
            out_key[s] = key;
            out_slot[s] = atomicAdd(&cnt[key], 1u);
        } else {
            out_x[s] = nvx;
            out_y[s] = nvy;
        }
    }
    if (VISC)
        report_neighbors(full, half, C);
}

// src/fluid.c:398-413 as a separate pass (phase API).
__global__ void __launch_bounds__(256)
k_gravity(const sph_dev_params *__restrict__ Pp, float *__restrict__ VY,
          float *__restrict__ RHO, float *__restrict__ RHON)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)Pp->n)
        return;
    VY[s] += Pp->g_neg * Pp->dt;
    RHO[s] = 0.0f;
    RHON[s] = 0.0f;
}

// src/fluid.c:507-523 as a separate, in-place pass (phase API).
__global__ void __launch_bounds__(256)
k_predict(const sph_dev_params *__restrict__ Pp, float *__restrict__ X, float *__restrict__ Y,
          const float *__restrict__ VX, const float *__restrict__ VY, float *__restrict__ XP,
          float *__restrict__ YP)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)P.n)
        return;
    float x = X[s], y = Y[s];
    XP[s] = x;
    YP[s] = y;
    x += VX[s] * P.dt;
    y += VY[s] * P.dt;
    sph_boundary(x, y, P);
    X[s] = x;
    Y[s] = y;
}

// src/fluid.c:642-653 as a separate, in-place pass (phase API).
__global__ void __launch_bounds__(256)
k_update_velocities(const sph_dev_params *__restrict__ Pp, float *__restrict__ X,
                    float *__restrict__ Y, float *__restrict__ VX, float *__restrict__ VY,
                    const float *__restrict__ XP, const float *__restrict__ YP)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)P.n)
        return;
    float x = X[s], y = Y[s];
    sph_boundary(x, y, P);
    float vx = (x - XP[s]) / P.dt;
    float vy = (y - YP[s]) / P.dt;
    sph_check_velocity(vx, vy);
    X[s] = x;
    Y[s] = y;
    VX[s] = vx;
    VY[s] = vy;
}

// Keys and slots of the current positions (histogram pass of a standalone
// hash, phase API; the fused step folds this into its producers).
__global__ void __launch_bounds__(256)
k_key_hist(const sph_dev_params *__restrict__ Pp, const float *__restrict__ X,
           const float *__restrict__ Y, unsigned *__restrict__ out_key,
           unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
           sph_dev_counters *__restrict__ C)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)P.n)
        return;
    const unsigned key = sph_cell_key(X[s], Y[s], P, &C->key_out_of_range);
    out_key[s] = key;
    out_slot[s] = atomicAdd(&cnt[key], 1u);
}

// ---------------------------------------------------------------------------
// Cell-count scan: start[c] = sum of cnt[0..c), over m = cells + 1 entries
// (cnt[cells] is always 0, so start[cells] = n).  Reduce-then-scan: tiles of
// kScanTile counts; the apply kernel adds the sum of all earlier tiles,
// writes the starts and clears the counts for the next sort.
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanTile = kScanThreads * kScanPerThread;

__global__ void __launch_bounds__(kScanThreads)
k_scan_reduce(const unsigned *__restrict__ cnt, unsigned m, unsigned *__restrict__ bsum,
              sph_dev_counters *__restrict__ C)
{
    __shared__ unsigned red[kScanThreads / kWave][3];
    const unsigned base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    unsigned sum = 0, over = 0, mx = 0;
#pragma unroll
    for (int e = 0; e < kScanPerThread; e++) {
        unsigned c = base + e < m ? cnt[base + e] : 0u;
        sum += c;
        over += c > SPH_REF_MAX_BUCKET;
        mx = max(mx, c);
    }
    sum = wave_sum_u32(sum);
    over = wave_sum_u32(over);
    mx = wave_max_u32(mx);
    const int w = threadIdx.x / kWave;
    if (lane0()) {
        red[w][0] = sum;
        red[w][1] = over;
        red[w][2] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0, o = 0, x = 0;
        for (int i = 0; i < kScanThreads / kWave; i++) {
            t += red[i][0];
            o += red[i][1];
            x = max(x, red[i][2]);
        }
        bsum[blockIdx.x] = t;
        if (o)
            atomicAdd(&C->bucket_overflow, (unsigned long long)o);
        atomicMax(&C->max_cell_count, (unsigned long long)x);
    }
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_apply(unsigned *__restrict__ cnt, unsigned m, const unsigned *__restrict__ bsum,
             unsigned *__restrict__ start)
{
    __shared__ unsigned wsum[kScanThreads / kWave];
    __shared__ unsigned wpre[kScanThreads / kWave];
    // prefix of all earlier tiles
    unsigned pre = 0;
    for (unsigned b = threadIdx.x; b < blockIdx.x; b += kScanThreads)
        pre += bsum[b];
    pre = wave_sum_u32(pre);
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (lane == 0)
        wsum[w] = pre;
    __syncthreads();
    unsigned tile_pre = 0;
    for (int i = 0; i < kScanThreads / kWave; i++)
        tile_pre += wsum[i];
    __syncthreads();

    const unsigned base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    unsigned v[kScanPerThread];
    unsigned t = 0;
#pragma unroll
    for (int e = 0; e < kScanPerThread; e++) {
        v[e] = base + e < m ? cnt[base + e] : 0u;
        t += v[e];
    }
    // inclusive wave scan of per-thread totals
    unsigned incl = t;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        unsigned y = (unsigned)__shfl_up((int)incl, o, kWave);
        if (lane >= o)
            incl += y;
    }
    if (lane == kWave - 1)
        wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned a = 0;
        for (int i = 0; i < kScanThreads / kWave; i++) {
            wpre[i] = a;
            a += wsum[i];
        }
    }
    __syncthreads();
    unsigned run = tile_pre + wpre[w] + (incl - t);
#pragma unroll
    for (int e = 0; e < kScanPerThread; e++) {
        if (base + e < m) {
            start[base + e] = run;
            cnt[base + e] = 0u;
        }
        run += v[e];
    }
}

// Drop (id, source index) into the cell ranges in slot order.
__global__ void __launch_bounds__(256)
k_scatter(int n, const unsigned *__restrict__ key, const unsigned *__restrict__ slot,
          const unsigned *__restrict__ start, const int *__restrict__ id_src,
          int *__restrict__ perm_src, int *__restrict__ perm_id)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned)n)
        return;
    const unsigned pos = start[key[i]] + slot[i];
    perm_src[pos] = (int)i;
    perm_id[pos] = id_src[i];
}

// Rank each particle by id inside its cell and gather its fields to
// (cell, id) order.  The slot order within a cell is arbitrary (atomics);
// ranking restores the reference's bucket order (ascending id).
template <int NF>
__global__ void __launch_bounds__(256)
k_reorder(int n, const unsigned *__restrict__ start, const int *__restrict__ perm_src,
          const int *__restrict__ perm_id, const unsigned *__restrict__ key,
          sph_reorder_fields<NF> f, int *__restrict__ id_dst, unsigned *__restrict__ key_dst)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (unsigned)n)
        return;
    const int i = perm_src[k];
    const int myid = perm_id[k];
    const unsigned c = key[i];
    const unsigned b = start[c], e = start[c + 1];
    unsigned rank = 0;
    for (unsigned m = b; m < e; m++)
        rank += perm_id[m] < myid;
    const unsigned d = b + rank;
#pragma unroll
    for (int q = 0; q < NF; q++)
        f.dst[q][d] = f.src[q][i];
    id_dst[d] = myid;
    key_dst[d] = c;
}

// ---------------------------------------------------------------------------
// Density (src/hash.c:190-194, 225-229 with src/fluid.c:527-539) in gather
// form: each particle sums (1-r/h)^2 and (1-r/h)^3 over its neighbours.
__global__ void __launch_bounds__(256)
k_density(const sph_dev_params *__restrict__ Pp, const float *__restrict__ X,
          const float *__restrict__ Y, const unsigned *__restrict__ K,
          const unsigned *__restrict__ start, float *__restrict__ RHO, float *__restrict__ RHON,
          sph_dev_counters *__restrict__ C)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned full = 0, half = 0;
    if (s < (unsigned)P.n) {
        const float ax = X[s], ay = Y[s];
        float rho = 0.0f, rhon = 0.0f;
        sph_rows R;
        sph_row_ranges(K[s], P, start, R);
        for (int r = 0; r < R.nr; r++) {
            for (unsigned k = R.b0[r]; k < R.b3[r]; k++) {
                if (k == s)
                    continue;
                const float dx = ax - X[k], dy = ay - Y[k];
                const float r2 = dx * dx + dy * dy;
                if (r2 > P.h2)
                    continue;
                full++;
                half += sph_owns(k, s, R.dy[r], R.b1[r], R.b2[r]);
                const float ratio = sqrtf(r2) * P.h_recip;
                const float OmR2 = (1.0f - ratio) * (1.0f - ratio);
                if (ratio < 1.0f) {
                    rho += OmR2;
                    rhon += OmR2 * (1.0f - ratio);
                }
            }
        }
        RHO[s] = rho;
        RHON[s] = rhon;
    }
    report_neighbors(full, half, C);
}

// Double-density relaxation (src/fluid.c:541-611) in gather form.
// FUSED adds the velocity update (src/fluid.c:642-653): boundary, then
// v = (x - x_prev)/dt clamped, and the histogram of the next sort; without
// FUSED only the relaxed positions are written.
template <bool FUSED>
__global__ void __launch_bounds__(256)
k_relax(const sph_dev_params *__restrict__ Pp, const float *__restrict__ X,
        const float *__restrict__ Y, const float *__restrict__ XP, const float *__restrict__ YP,
        const float *__restrict__ RHO, const float *__restrict__ RHON,
        const unsigned *__restrict__ K, const unsigned *__restrict__ start,
        float *__restrict__ out_x, float *__restrict__ out_y, float *__restrict__ out_vx,
        float *__restrict__ out_vy, unsigned *__restrict__ out_key,
        unsigned *__restrict__ out_slot, unsigned *__restrict__ cnt,
        sph_dev_counters *__restrict__ C)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)P.n)
        return;
    const float ax = X[s], ay = Y[s];
    const float aP = P.k * (RHO[s] - P.rest_density);
    const float aPn = P.k_near * RHON[s];
    float nx = ax, ny = ay;
    sph_rows R;
    sph_row_ranges(K[s], P, start, R);
    for (int r = 0; r < R.nr; r++) {
        for (unsigned k = R.b0[r]; k < R.b3[r]; k++) {
            if (k == s)
                continue;
            const float bx = X[k], by = Y[k];
            const float dx = ax - bx, dy = ay - by;
            const float r2 = dx * dx + dy * dy;
            if (r2 > P.h2)
                continue;
            const float rr = sqrtf(r2);
            const float r_recip = 1.0f / rr;
            const float ratio = rr * P.h_recip;
            const float OmR = 1.0f - ratio;
            if (rr <= 0.000001f && sph_owns(k, s, R.dy[r], R.b1[r], R.b2[r])) {
                nx += 0.000001f;
                ny += 0.000001f;
            }
            if (ratio < 1.0f && rr > 0.0f) {
                const float bP = P.k * (RHO[k] - P.rest_density);
                const float bPn = P.k_near * RHON[k];
                const float D = sph_relax_D(P.dt, aP + bP, aPn + bPn, OmR, P.k_spring, P.h, rr);
                const float Dx = D * (bx - ax) * r_recip;
                const float Dy = D * (by - ay) * r_recip;
                nx -= Dx;
                ny -= Dy;
            }
        }
    }
    if (FUSED) {
        sph_boundary(nx, ny, P);
        float vx = (nx - XP[s]) / P.dt;
        float vy = (ny - YP[s]) / P.dt;
        sph_check_velocity(vx, vy);
        out_vx[s] = vx;
        out_vy[s] = vy;
        if (!(isfinite(nx) && isfinite(ny)))
            atomicAdd(&C->nonfinite, 1ull);
        const unsigned key = sph_cell_key(nx, ny, P, &C->key_out_of_range);
        out_key[s] = key;
        out_slot[s] = atomicAdd(&cnt[key], 1u);
    }
    out_x[s] = nx;
    out_y[s] = ny;
}

// Neighbour sets of the current hash, written by id (small-N read-back for
// tests): full set in canonical order, and the half-list length the
// reference would hold for each particle.
__global__ void __launch_bounds__(256)
k_neighbor_lists(const sph_dev_params *__restrict__ Pp, const float *__restrict__ X,
                 const float *__restrict__ Y, const unsigned *__restrict__ K,
                 const int *__restrict__ ID, const unsigned *__restrict__ start, int stride,
                 int *__restrict__ full_by_id, int *__restrict__ half_by_id,
                 int *__restrict__ ids_by_id)
{
    const sph_dev_params P = *Pp;
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)P.n)
        return;
    const int me = ID[s];
    const float ax = X[s], ay = Y[s];
    int full = 0, half = 0;
    sph_rows R;
    sph_row_ranges(K[s], P, start, R);
    for (int r = 0; r < R.nr; r++)
        for (unsigned k = R.b0[r]; k < R.b3[r]; k++) {
            if (k == s)
                continue;
            const float dx = ax - X[k], dy = ay - Y[k];
            if (dx * dx + dy * dy > P.h2)
                continue;
            if (full < stride)
                ids_by_id[(size_t)me * stride + full] = ID[k];
            full++;
            half += sph_owns(k, s, R.dy[r], R.b1[r], R.b2[r]);
        }
    full_by_id[me] = full;
    half_by_id[me] = half;
}

// Reference coordinate packing (src/fluid.c:356-363), written by id.
__global__ void __launch_bounds__(256)
k_pack_coords(int n, const float *__restrict__ X, const float *__restrict__ Y,
              const int *__restrict__ ID, float max_x, float max_y, short *__restrict__ out)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (unsigned)n)
        return;
    const int id = ID[s];
    out[2 * id] = (short)((2.0f * X[s] / max_x - 1.0f) * 32767.0f);
    out[2 * id + 1] = (short)((2.0f * Y[s] / max_y - 1.0f) * 32767.0f);
}

// Scatter a field from one sorted order to another through ids:
// dst[pos_of_id[id_src[k]]] = src[k] (re-canonicalises auxiliary fields).
__global__ void __launch_bounds__(256)
k_pos_of_id(int n, const int *__restrict__ ID, int *__restrict__ pos_of_id)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < (unsigned)n)
        pos_of_id[ID[s]] = (int)s;
}

__global__ void __launch_bounds__(256)
k_permute_by_id(int n, const int *__restrict__ id_src, const int *__restrict__ pos_of_id,
                sph_reorder_fields<4> f)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (unsigned)n)
        return;
    const int d = pos_of_id[id_src[k]];
#pragma unroll
    for (int q = 0; q < 4; q++)
        f.dst[q][d] = f.src[q][k];
}

// ---------------------------------------------------------------------------
// Host-side launchers (C linkage inside the library).

#define GRID(n, b) dim3((unsigned)(((n) + (b)-1) / (b)))

void sph_launch_visc_predict(hipStream_t st, int n, bool grav, bool visc, bool pred,
                             const sph_dev_params *P, const float *X, const float *Y,
                             const float *VX, const float *VY, const unsigned *K,
                             const unsigned *start, float *ox, float *oy, unsigned *okey,
                             unsigned *oslot, unsigned *cnt, sph_dev_counters *C)
{
#define VP(G, V, R)                                                                              \
    k_visc_predict<G, V, R><<<GRID(n, 256), 256, 0, st>>>(P, X, Y, VX, VY, K, start, ox, oy,    \
                                                          okey, oslot, cnt, C)
    if (grav && visc && pred)
        VP(true, true, true);
    else if (grav && !visc && pred)
        VP(true, false, true);
    else if (!grav && visc && !pred)
        VP(false, true, false);
    else if (!grav && visc && pred)
        VP(false, true, true);
    else if (!grav && !visc && pred)
        VP(false, false, true);
    else
        VP(true, true, false);
#undef VP
}

void sph_launch_gravity(hipStream_t st, int n, const sph_dev_params *P, float *VY, float *RHO,
                        float *RHON)
{
    k_gravity<<<GRID(n, 256), 256, 0, st>>>(P, VY, RHO, RHON);
}

void sph_launch_predict(hipStream_t st, int n, const sph_dev_params *P, float *X, float *Y,
                        const float *VX, const float *VY, float *XP, float *YP)
{
    k_predict<<<GRID(n, 256), 256, 0, st>>>(P, X, Y, VX, VY, XP, YP);
}

void sph_launch_update_velocities(hipStream_t st, int n, const sph_dev_params *P, float *X,
                                  float *Y, float *VX, float *VY, const float *XP,
                                  const float *YP)
{
    k_update_velocities<<<GRID(n, 256), 256, 0, st>>>(P, X, Y, VX, VY, XP, YP);
}

void sph_launch_key_hist(hipStream_t st, int n, const sph_dev_params *P, const float *X,
                         const float *Y, unsigned *okey, unsigned *oslot, unsigned *cnt,
                         sph_dev_counters *C)
{
    k_key_hist<<<GRID(n, 256), 256, 0, st>>>(P, X, Y, okey, oslot, cnt, C);
}

unsigned sph_scan_blocks(unsigned m) { return (m + kScanTile - 1) / kScanTile; }

void sph_launch_scan_reduce(hipStream_t st, const unsigned *cnt, unsigned m, unsigned *bsum,
                            sph_dev_counters *C)
{
    k_scan_reduce<<<dim3(sph_scan_blocks(m)), kScanThreads, 0, st>>>(cnt, m, bsum, C);
}

void sph_launch_scan_apply(hipStream_t st, unsigned *cnt, unsigned m, const unsigned *bsum,
                           unsigned *start)
{
    k_scan_apply<<<dim3(sph_scan_blocks(m)), kScanThreads, 0, st>>>(cnt, m, bsum, start);
}

void sph_launch_scatter(hipStream_t st, int n, const unsigned *key, const unsigned *slot,
                        const unsigned *start, const int *id_src, int *perm_src, int *perm_id)
{
    k_scatter<<<GRID(n, 256), 256, 0, st>>>(n, key, slot, start, id_src, perm_src, perm_id);
}

void sph_launch_reorder4(hipStream_t st, int n, const unsigned *start, const int *perm_src,
                         const int *perm_id, const unsigned *key, sph_reorder_fields<4> f,
                         int *id_dst, unsigned *key_dst)
{
    k_reorder<4><<<GRID(n, 256), 256, 0, st>>>(n, start, perm_src, perm_id, key, f, id_dst,
                                               key_dst);
}

void sph_launch_reorder8(hipStream_t st, int n, const unsigned *start, const int *perm_src,
                         const int *perm_id, const unsigned *key, sph_reorder_fields<8> f,
                         int *id_dst, unsigned *key_dst)
{
    k_reorder<8><<<GRID(n, 256), 256, 0, st>>>(n, start, perm_src, perm_id, key, f, id_dst,
                                               key_dst);
}

void sph_launch_density(hipStream_t st, int n, const sph_dev_params *P, const float *X,
                        const float *Y, const unsigned *K, const unsigned *start, float *RHO,
                        float *RHON, sph_dev_counters *C)
{
    k_density<<<GRID(n, 256), 256, 0, st>>>(P, X, Y, K, start, RHO, RHON, C);
}

void sph_launch_relax(hipStream_t st, int n, bool fused, const sph_dev_params *P, const float *X,
                      const float *Y, const float *XP, const float *YP, const float *RHO,
                      const float *RHON, const unsigned *K, const unsigned *start, float *ox,
                      float *oy, float *ovx, float *ovy, unsigned *okey, unsigned *oslot,
                      unsigned *cnt, sph_dev_counters *C)
{
    if (fused)
        k_relax<true><<<GRID(n, 256), 256, 0, st>>>(P, X, Y, XP, YP, RHO, RHON, K, start, ox, oy,
                                                    ovx, ovy, okey, oslot, cnt, C);
    else
        k_relax<false><<<GRID(n, 256), 256, 0, st>>>(P, X, Y, XP, YP, RHO, RHON, K, start, ox,
                                                     oy, ovx, ovy, okey, oslot, cnt, C);
}

void sph_launch_pack_coords(hipStream_t st, int n, const float *X, const float *Y, const int *ID,
                            float max_x, float max_y, short *out)
{
    k_pack_coords<<<GRID(n, 256), 256, 0, st>>>(n, X, Y, ID, max_x, max_y, out);
}

void sph_launch_neighbor_lists(hipStream_t st, int n, const sph_dev_params *P, const float *X,
                               const float *Y, const unsigned *K, const int *ID,
                               const unsigned *start, int stride, int *full_by_id,
                               int *half_by_id, int *ids_by_id)
{
    k_neighbor_lists<<<GRID(n, 256), 256, 0, st>>>(P, X, Y, K, ID, start, stride, full_by_id,
                                                    half_by_id, ids_by_id);
}

void sph_launch_pos_of_id(hipStream_t st, int n, const int *ID, int *pos_of_id)
{
    k_pos_of_id<<<GRID(n, 256), 256, 0, st>>>(n, ID, pos_of_id);
}

void sph_launch_permute_by_id(hipStream_t st, int n, const int *id_src, const int *pos_of_id,
                              sph_reorder_fields<4> f)
{
    k_permute_by_id<<<GRID(n, 256), 256, 0, st>>>(n, id_src, pos_of_id, f);
}
