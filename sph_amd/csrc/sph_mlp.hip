// sph_mlp.hip - the learned macro-step ("surrogate") of libsph_hip.so.
//
// A particle-interaction network predicts each particle's displacement over
// four plain steps from its neighbourhood at the start of the macro-step:
//
//   edge features  f_ij = (dx/h, dy/h, dvx/5, dvy/5, r/h, 1)   (j within h)
//   edge MLP       e_ij = W2 SiLU(W1 f_ij)      (5 -> 64 -> 64, b1 folded
//                                                into W1 via the constant 1)
//   aggregation    m_i  = sum_j e_ij = W2 S_i + deg_i b2,  S_i = sum_j SiLU(W1 f_ij)
//   node MLP       o_i  = W4 SiLU(W3 [m_i, v_i/5, a_i/10, 1]) + b4   (68 -> 64 -> 2)
//   displacement   d_i  = 4 dt v_i + 10 dt^2 a_i + 0.05 o_i
//
// a_i is the external acceleration (sloshing, -g).  The edge MLP's second
// layer is linear, so it is applied once per particle after the sum: the
// per-edge work is W1 f_ij on MFMA plus one SiLU per hidden unit.
//
// Precision: weights and activations bf16, MFMA 16x16x16 bf16 with fp32
// accumulation, SiLU / sums / losses / Adam in fp32 (fp32 master weights).
//
// Work layout: one wavefront per particle for the edge passes (its
// neighbours are compacted into LDS, 16 edges per MFMA tile), one wavefront
// per 16 consecutive particles for the node passes, split-K MFMA reductions
// for the weight gradients.  Particles are the solver context's sorted slots.
#include "sph_ctx_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) short bf16x4;
typedef __attribute__((ext_vector_type(4))) __bf16 bfx4;

namespace {

constexpr int kH = 64;      // hidden width
constexpr int kKE = 16;     // edge input width (6 used)
constexpr int kKN = 80;     // node input width (64 + 4 + bias, padded)
constexpr unsigned kGroup = 16;         // consecutive slots per wavefront in the edge passes
constexpr unsigned kEdgeBuf = 256;      // LDS edge rows per wavefront (flushed as it fills)
constexpr float kVScale = 0.2f, kAScale = 0.1f, kDScale = 0.05f;
constexpr int kTrendRefresh = 16; // forced plain macro-step per scene every this many

// parameter block (fp32 master copy)
constexpr int oW1 = 0;                  // [64][16]
constexpr int oW2 = oW1 + kH * kKE;     // [64][64]
constexpr int ob2 = oW2 + kH * kH;      // [64]
constexpr int oW3 = ob2 + kH;           // [64][80]
constexpr int oW4 = oW3 + kH * kKN;     // [2][64]
constexpr int ob4 = oW4 + 2 * kH;       // [2]
constexpr int kNParams = ob4 + 2;

// float -> bf16 bits, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short f2bf(float f)
{
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}

__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// Logistic function from the hardware exp2 and reciprocal (1 ulp each): the
// network is trained and run with this definition on both passes.
__device__ __forceinline__ float sigm(float x)
{
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * -1.44269504088896341f));
}

__device__ __forceinline__ float silu(float x) { return x * sigm(x); }

__device__ __forceinline__ float silu_grad(float x)
{
    const float s = sigm(x);
    return s * (1.0f + x * (1.0f - s));
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    const f32x4 v = {a, b, c, d};
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bfx4));
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ unsigned lane() { return threadIdx.x & 63u; }

// Particle-level parameters of a scene needed by the network.
struct NodeCtx {
    float h_recip, dt, g_neg, fx;
};

__device__ __forceinline__ NodeCtx node_ctx(const sph_dev_params &P, unsigned long long clock)
{
    NodeCtx c;
    c.h_recip = P.h_recip;
    c.dt = P.dt;
    c.g_neg = P.g_neg;
    c.fx = sph_slosh_accel(P, clock);
    return c;
}

// LDS edge list of one wavefront's group of kGroup consecutive slots:
// features (dx/h, dy/h, 0.2 dvx, 0.2 dvy, r/h, 1, 0, 0) as two bf16x4, and
// the owning particle (0..15 within the group; 0xff for padding rows).
struct EdgeBuf {
    bf16x4 f[kEdgeBuf][2];
    unsigned char owner[kEdgeBuf];
};

// Stream every edge of the particles [n0, n0 + 16) through L (one
// wavefront).  Lane l tests particle l % 16 against every fourth candidate
// (offset l / 16) of each of the particle's three candidate rows (cells
// i-1..i+1 of rows j-1..j+1, contiguous in the sorted arrays), so all 64
// lanes test useful pairs.  Accepted edges are appended to L in arbitrary
// order; whenever L is nearly full, flush(nedges) consumes its first nedges
// (a multiple of 16) and the remainder moves to the front.  The final call
// gets the rest, padded to a multiple of 16 with zero rows owned by 0xff.
// Lanes 0..15 get the degree and, with DENS, the density sum
// (src/fluid.c:527-539: (1 - r/h)^2 for r/h < 1) of particle n0 + lane.
template <bool DENS, class Flush>
__device__ void walk_group(unsigned n0, const sph_dev_params &P, const float2 *POS,
                           const float2 *VEL, const unsigned *K, const unsigned *start,
                           EdgeBuf &L, Flush &&flush, float &deg_out, float &rho_out)
{
    const unsigned l = lane(), p = l % 16u, sub = l / 16u;
    const unsigned me = n0 + p;
    const bool valid = me < P.poff + (unsigned)P.n;
    float2 a = make_float2(0, 0), av = make_float2(0, 0);
    unsigned i = 0, j = 0;
    if (valid) {
        a = POS[me];
        av = VEL[me];
        const unsigned key = K[me] - P.cell_base;
        i = key % P.sx;
        j = key / P.sx;
    }
    const unsigned lo = i > 0 ? i - 1 : 0, hi = i + 1 < P.sx ? i + 1 : P.sx - 1;
    unsigned count = 0, mine = 0;
    float dens = 0.0f;
    for (int dy = -1; dy <= 1; dy++) {
        const long r = (long)j + dy;
        unsigned b0 = 0, b3 = 0;
        if (valid && r >= 0 && r < (long)P.sy) {
            const unsigned base = P.cell_base + (unsigned)r * P.sx;
            b0 = start[base + lo];
            b3 = start[base + hi + 1];
        }
        for (unsigned c = b0 + sub;; c += 4) {
            const bool act = c < b3;
            if (__ballot(act) == 0ull)
                break;
            bool acc = false;
            float dx = 0, dyy = 0, r2 = 0;
            if (act && c != me) {
                const float2 q = POS[c];
                dx = q.x - a.x;
                dyy = q.y - a.y;
                r2 = dx * dx + dyy * dyy;
                acc = !(r2 > P.h2);
            }
            const unsigned long long m = __ballot(acc);
            if (acc) {
                const unsigned pos = count + (unsigned)__popcll(m & ((1ull << l) - 1ull));
                const float rr = sqrtf(r2);
                if (DENS) {
                    const float ratio = rr * P.h_recip;
                    if (ratio < 1.0f)
                        dens += (1.0f - ratio) * (1.0f - ratio);
                }
                mine++;
                const float2 qv = VEL[c];
                L.f[pos][0] = pack4(dx * P.h_recip, dyy * P.h_recip, (qv.x - av.x) * kVScale,
                                    (qv.y - av.y) * kVScale);
                L.f[pos][1] = pack4(rr * P.h_recip, 1.0f, 0.0f, 0.0f);
                L.owner[pos] = (unsigned char)p;
            }
            count += (unsigned)__popcll(m);
            if (count > kEdgeBuf - 64u) { // the next batch might not fit: flush full tiles
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                const unsigned full = count & ~15u, rem = count - full;
                flush(full);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                bf16x4 f0 = pack4(0, 0, 0, 0), f1 = f0;
                unsigned char ow = 0xffu;
                if (l < rem) {
                    f0 = L.f[full + l][0];
                    f1 = L.f[full + l][1];
                    ow = L.owner[full + l];
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                if (l < rem) {
                    L.f[l][0] = f0;
                    L.f[l][1] = f1;
                    L.owner[l] = ow;
                }
                count = rem;
            }
        }
    }
    // the rest, padded to a whole 16-edge tile
    const unsigned padded = (count + 15u) & ~15u;
    if (count + l < padded) {
        L.f[count + l][0] = pack4(0, 0, 0, 0);
        L.f[count + l][1] = pack4(0, 0, 0, 0);
        L.owner[count + l] = 0xffu;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (padded)
        flush(padded);
    float fm = (float)mine;
    fm += __shfl_xor(fm, 16, 64);
    fm += __shfl_xor(fm, 32, 64);
    deg_out = fm;
    if (DENS) {
        dens += __shfl_xor(dens, 16, 64);
        dens += __shfl_xor(dens, 32, 64);
        rho_out = dens;
    }
    __builtin_amdgcn_wave_barrier();
}

// A-operand of an edge tile: lane l holds edge e0 + l%16, features
// 4*(l/16) .. +3 (features 8..15 are zero).
__device__ __forceinline__ bf16x4 edge_tile_a(const EdgeBuf &L, unsigned e0)
{
    const unsigned sub = lane() / 16u;
    return sub < 2u ? L.f[e0 + lane() % 16u][sub] : pack4(0, 0, 0, 0);
}

// A-operand of the group membership matrix of an edge tile: lane l holds
// row (particle) l%16, columns (edges) e0 + 4*(l/16) .. +3: 1 where the edge
// belongs to the particle.
__device__ __forceinline__ bf16x4 member_a(const EdgeBuf &L, unsigned e0)
{
    const unsigned p = lane() % 16u, e = e0 + 4u * (lane() / 16u);
    bf16x4 r;
    for (int q = 0; q < 4; q++)
        r[q] = L.owner[e + q] == p ? (short)0x3f80 : (short)0;
    return r;
}

// B-operand of W1^T for hidden-unit tile t: B[k][c] = W1[16t + c][k].
__device__ __forceinline__ bf16x4 w1_b(const unsigned short *W1b, int t)
{
    const unsigned c = 16u * t + lane() % 16u, k = 4u * (lane() / 16u);
    bf16x4 r;
    for (int q = 0; q < 4; q++)
        r[q] = (short)W1b[c * kKE + k + q];
    return r;
}

} // namespace

struct sph_surrogate {
    size_t slots = 0;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    unsigned short *W1b = nullptr, *W2b = nullptr, *W3b = nullptr;
    float *S = nullptr, *deg = nullptr, *dS = nullptr, *disp = nullptr, *tgt = nullptr;
    unsigned short *U = nullptr, *dP3 = nullptr, *dM = nullptr; // tile-major bf16
    float *loss = nullptr;
    float2 *tmp2 = nullptr;
    int adam_t = 0;
    bool have_pred = false;
    // judge state (per scene)
    size_t ns = 0;
    float *ref = nullptr, *jsum = nullptr;
    float *jmax = nullptr;
    int *jflags = nullptr;
    const sph_ctx *ref_ctx = nullptr;
    bool ref_ok = false;
    unsigned long long macro = 0; // judged macro-steps so far (the refresh / shadow schedule)
    // shadow judging: 1 of every shadow_every scenes (rotating) also takes
    // the macro-step plainly from the same start; 0 = off
    int shadow_every = 0;
    float *sh_sum = nullptr;              // per scene: the plain shadow's density sum
    float *rho_plain = nullptr, *rho_sur = nullptr; // density by caller index
    float *sh_err = nullptr;              // per scene: sum (sur - plain)^2, sum plain^2
    double sh_stats[8] = {0};             // see sph_surrogate_shadow_stats
};

struct sph_trajectory {
    int nscenes = 0, n = 0, frames = 0;
    float4 *data = nullptr;    // [frames][nscenes][n]
    float4 *staging = nullptr; // [cap_samples * n]
    float2 *target = nullptr;  // [cap_samples * n]
    int *idx = nullptr;        // scene[cap], frame[cap]
    int cap = 0, loaded = 0;
};

// ---------------------------------------------------------------------------
// kernels

// MFMA lane-layout self-test: D = A B for fixed small-integer A, B
// (exact in bf16), written row-major as the layout assumed above reads it.
__global__ void k_mfma_selftest(float *D)
{
    const unsigned l = lane();
    bf16x4 a, b;
    for (int q = 0; q < 4; q++) {
        const unsigned row = l % 16u, k = 4u * (l / 16u) + q;
        a[q] = (short)f2bf((float)((int)((row * 16u + k) % 7u) - 3));
        const unsigned kk = 4u * (l / 16u) + q, col = l % 16u;
        b[q] = (short)f2bf((float)((int)((kk * 3u + col) % 5u) - 2));
    }
    f32x4 d = {0, 0, 0, 0};
    d = mfma(a, b, d);
    for (int q = 0; q < 4; q++)
        D[(4u * (l / 16u) + q) * 16u + l % 16u] = d[q];
}

// Edge forward: S_i = sum_j SiLU(W1 f_ij) and deg_i (and, with DENS, the
// density sum of the judge).  One wavefront per group of 16 slots: the
// group's edges are compacted into LDS, each 16-edge tile runs
// W1 on MFMA (4 hidden tiles), SiLU, and the per-particle sum as a second
// MFMA against the tile's membership matrix.
template <bool DENS>
__global__ void __launch_bounds__(256)
k_edge_fwd(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS, unsigned slots,
           const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
           const unsigned *__restrict__ K, const unsigned *__restrict__ start,
           const unsigned short *__restrict__ W1b, float *__restrict__ S,
           float *__restrict__ deg, float *__restrict__ rho)
{
    __shared__ EdgeBuf Lb[4];
    const unsigned w = threadIdx.x / 64u, l = lane(), p = l % 16u, sub = l / 16u;
    EdgeBuf &L = Lb[w];
    bf16x4 B[4];
    for (int t = 0; t < 4; t++)
        B[t] = w1_b(W1b, t);
    const unsigned ngroups = slots / kGroup, nw = gridDim.x * 4u;
    for (unsigned g = blockIdx.x * 4u + w; g < ngroups; g += nw) {
        const unsigned n0 = g * kGroup;
        const sph_dev_params &P = Pp[BS[n0 / kSphTPB]];
        f32x4 acc[4];
        for (int t = 0; t < 4; t++)
            acc[t] = {0, 0, 0, 0};
        float dg = 0.0f, rh = 0.0f;
        if (n0 < P.poff + (unsigned)P.n) {
            auto tiles = [&](unsigned ne) {
                for (unsigned e0 = 0; e0 < ne; e0 += 16) {
                    const bf16x4 a = edge_tile_a(L, e0);
                    const bf16x4 mem = member_a(L, e0);
                    for (int t = 0; t < 4; t++) {
                        f32x4 d = {0, 0, 0, 0};
                        d = mfma(a, B[t], d);
                        // D layout (edge 4*sub+q, unit 16t+p) is the B layout of the sum
                        acc[t] = mfma(mem, pack4(silu(d[0]), silu(d[1]), silu(d[2]), silu(d[3])),
                                      acc[t]);
                    }
                }
            };
            walk_group<DENS>(n0, P, POS, VEL, K, start, L, tiles, dg, rh);
        }
        for (int t = 0; t < 4; t++)
            for (int q = 0; q < 4; q++)
                S[(size_t)(n0 + 4u * sub + q) * kH + 16u * t + p] = acc[t][q];
        if (l < 16u) {
            deg[n0 + l] = dg;
            if (DENS)
                rho[n0 + l] = rh;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

namespace {

// Node-pass state of one 16-particle tile held by one wavefront.
struct NodeTile {
    f32x4 pre3[4]; // D layout: row 4*(l/16)+q, unit 16t + l%16
    float out[4][2];
};

// Node-layer weights staged in LDS once per workgroup: the B operands of
// every tile's MFMAs are read from here as 8-byte rows.
struct NodeW {
    unsigned short W2[kH * kH];
    unsigned short W3[kH * kKN];
    float b2[kH], W4[2 * kH], b4[2];
};

__device__ void stage_node_weights(NodeW &sw, const unsigned short *W2b,
                                   const unsigned short *W3b, const float *params)
{
    for (unsigned i = threadIdx.x; i < kH * kH / 4u; i += blockDim.x)
        ((uint2 *)sw.W2)[i] = ((const uint2 *)W2b)[i];
    for (unsigned i = threadIdx.x; i < kH * kKN / 4u; i += blockDim.x)
        ((uint2 *)sw.W3)[i] = ((const uint2 *)W3b)[i];
    for (unsigned i = threadIdx.x; i < kH; i += blockDim.x)
        sw.b2[i] = params[ob2 + i];
    for (unsigned i = threadIdx.x; i < 2u * kH; i += blockDim.x)
        sw.W4[i] = params[oW4 + i];
    if (threadIdx.x < 2u)
        sw.b4[threadIdx.x] = params[ob4 + threadIdx.x];
    __syncthreads();
}

__device__ __forceinline__ bf16x4 lds_row4(const unsigned short *base)
{
    return *(const bf16x4 *)base;
}

// Build u = [m, feats, 1, 0..] for a tile in LDS (bf16 [16][80]) and run the
// node forward.  Rows of invalid slots are zero.
// A operands of a tile's S rows (row l%16, columns 16kc + 4(l/16) ..) from
// the fp32 S array, and the degrees of rows 4(l/16) + q.
__device__ __forceinline__ void load_s_tile(unsigned n0, unsigned vend, const float *S,
                                            const float *deg, bf16x4 A[4], float degv[4])
{
    const unsigned l = lane();
    const unsigned row = n0 + l % 16u;
    for (int kc = 0; kc < 4; kc++) {
        const unsigned k = 16u * kc + 4u * (l / 16u);
        if (row < vend) {
            const float4 sv = *(const float4 *)(S + (size_t)row * kH + k);
            A[kc] = pack4(sv.x, sv.y, sv.z, sv.w);
        } else {
            A[kc] = pack4(0, 0, 0, 0);
        }
    }
    for (int q = 0; q < 4; q++) {
        const unsigned r = n0 + 4u * (l / 16u) + q;
        degv[q] = r < vend ? deg[r] : 0.0f;
    }
}

__device__ __forceinline__ void node_forward(unsigned n0, const sph_dev_params &P,
                                             unsigned long long clock, const float2 *VEL,
                                             const bf16x4 A[4], const float degv[4],
                                             const NodeW &sw, unsigned short (*ul)[kKN],
                                             NodeTile &T)
{
    const unsigned l = lane();
    const unsigned nvalid_end = P.poff + (unsigned)P.n;
    const NodeCtx nc = node_ctx(P, clock);
    // m = S W2^T + deg b2
    for (int t = 0; t < 4; t++) {
        f32x4 d = {0, 0, 0, 0};
        for (int kc = 0; kc < 4; kc++) {
            // B[k][c] = W2[16t + c][k]
            const unsigned c = 16u * t + l % 16u, k = 16u * kc + 4u * (l / 16u);
            d = mfma(A[kc], lds_row4(&sw.W2[c * kH + k]), d);
        }
        const float b2 = sw.b2[16u * t + l % 16u];
        for (int q = 0; q < 4; q++) {
            const unsigned r = 4u * (l / 16u) + q;
            const unsigned row = n0 + r;
            const float m = row < nvalid_end ? d[q] + degv[q] * b2 : 0.0f;
            ul[r][16u * t + l % 16u] = f2bf(m);
        }
    }
    // node features, bias column, padding
    if (l < 16u) {
        const unsigned row = n0 + l;
        const bool ok = row < nvalid_end;
        const float2 v = ok ? VEL[row] : make_float2(0, 0);
        ul[l][64] = f2bf(v.x * kVScale);
        ul[l][65] = f2bf(v.y * kVScale);
        ul[l][66] = f2bf(ok ? nc.fx * kAScale : 0.0f);
        ul[l][67] = f2bf(ok ? nc.g_neg * kAScale : 0.0f);
        ul[l][68] = f2bf(ok ? 1.0f : 0.0f);
        for (int q = 69; q < kKN; q++)
            ul[l][q] = 0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    // pre3 = u W3^T
    bf16x4 Au[5];
    for (int kc = 0; kc < 5; kc++) {
        const unsigned k = 16u * kc + 4u * (l / 16u);
        for (int q = 0; q < 4; q++)
            Au[kc][q] = (short)ul[l % 16u][k + q];
    }
    for (int t = 0; t < 4; t++) {
        f32x4 d = {0, 0, 0, 0};
        for (int kc = 0; kc < 5; kc++) {
            // B[k][c] = W3[16t + c][k]
            const unsigned c = 16u * t + l % 16u, k = 16u * kc + 4u * (l / 16u);
            d = mfma(Au[kc], lds_row4(&sw.W3[c * kKN + k]), d);
        }
        T.pre3[t] = d;
    }
    // out = W4 SiLU(pre3) + b4, reduced over the 16 lanes of each row group
    for (int q = 0; q < 4; q++)
        for (int o = 0; o < 2; o++) {
            float p = 0.0f;
            for (int t = 0; t < 4; t++)
                p += silu(T.pre3[t][q]) * sw.W4[o * kH + 16u * t + l % 16u];
            for (int x = 1; x < 16; x <<= 1)
                p += __shfl_xor(p, x, 64);
            T.out[q][o] = p + sw.b4[o];
        }
}

__device__ __forceinline__ float2 ballistic(const NodeCtx &nc, float2 v)
{
    // four steps of v += a dt; x += v dt from rest of interactions
    return make_float2(4.0f * nc.dt * v.x + 10.0f * nc.dt * nc.dt * nc.fx,
                       4.0f * nc.dt * v.y + 10.0f * nc.dt * nc.dt * nc.g_neg);
}

} // namespace

// Inference: the edge pass and the node MLP of each 16-slot group in one
// kernel (the group is one node tile).  S and deg stay on chip: the edge
// pass's per-particle sums are transposed through LDS into the node MLP's A
// operands, in the LDS the edge buffer used.
union InferLds {
    EdgeBuf e;
    struct {
        unsigned short ul[16][kKN];
        unsigned short sS[16][kH];
        float deg[16];
    } n;
};

__global__ void __launch_bounds__(256)
k_infer(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS, unsigned slots,
        const unsigned long long *__restrict__ clock, const float2 *__restrict__ POS,
        const float2 *__restrict__ VEL, const unsigned *__restrict__ K,
        const unsigned *__restrict__ start, const unsigned short *__restrict__ W1b,
        const unsigned short *__restrict__ W2b, const float *__restrict__ params,
        const unsigned short *__restrict__ W3b, float2 *__restrict__ disp)
{
    __shared__ __align__(16) InferLds Lb[4];
    __shared__ __align__(16) NodeW sw;
    stage_node_weights(sw, W2b, W3b, params);
    const unsigned w = threadIdx.x / 64u, l = lane(), p = l % 16u, sub = l / 16u;
    InferLds &L = Lb[w];
    bf16x4 B[4];
    for (int t = 0; t < 4; t++)
        B[t] = w1_b(W1b, t);
    const unsigned ngroups = slots / kGroup, nw = gridDim.x * 4u;
    const unsigned long long ck = *clock;
    for (unsigned g = blockIdx.x * 4u + w; g < ngroups; g += nw) {
        const unsigned n0 = g * kGroup;
        const sph_dev_params &P = Pp[BS[n0 / kSphTPB]];
        const unsigned vend = P.poff + (unsigned)P.n;
        if (n0 >= vend)
            continue;
        f32x4 acc[4];
        for (int t = 0; t < 4; t++)
            acc[t] = {0, 0, 0, 0};
        float dg = 0.0f, rh = 0.0f;
        auto tiles = [&](unsigned ne) {
            for (unsigned e0 = 0; e0 < ne; e0 += 16) {
                const bf16x4 a = edge_tile_a(L.e, e0);
                const bf16x4 mem = member_a(L.e, e0);
                for (int t = 0; t < 4; t++) {
                    f32x4 d = {0, 0, 0, 0};
                    d = mfma(a, B[t], d);
                    acc[t] = mfma(mem, pack4(silu(d[0]), silu(d[1]), silu(d[2]), silu(d[3])),
                                  acc[t]);
                }
            }
        };
        walk_group<false>(n0, P, POS, VEL, K, start, L.e, tiles, dg, rh);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // S tile (rows 4*sub+q, units 16t+p) -> LDS, transposed on read below
        for (int t = 0; t < 4; t++)
            for (int q = 0; q < 4; q++)
                L.n.sS[4u * sub + q][16u * t + p] = f2bf(acc[t][q]);
        if (l < 16u)
            L.n.deg[l] = dg;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        bf16x4 A[4];
        for (int kc = 0; kc < 4; kc++)
            A[kc] = lds_row4(&L.n.sS[p][16u * kc + 4u * sub]);
        float degv[4];
        for (int q = 0; q < 4; q++)
            degv[q] = L.n.deg[4u * sub + q];
        NodeTile T;
        node_forward(n0, P, ck, VEL, A, degv, sw, L.n.ul, T);
        const NodeCtx nc = node_ctx(P, ck);
        if (p == 0)
            for (int q = 0; q < 4; q++) {
                const unsigned row = n0 + 4u * sub + q;
                if (row < vend) {
                    const float2 b = ballistic(nc, VEL[row]);
                    disp[row] = make_float2(b.x + kDScale * T.out[q][0],
                                            b.y + kDScale * T.out[q][1]);
                }
            }
        __builtin_amdgcn_wave_barrier();
    }
}

// Node backward: loss, dout, dW4/db4 (+ atomics), and per particle dS (fp32),
// dPre3, dM, U (tile-major bf16 [tile][col][16]) for the weight-gradient
// reduction.  tgt holds the target displacement per slot.
__global__ void __launch_bounds__(256)
k_node_bwd(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS, unsigned slots,
           const unsigned long long *__restrict__ clock, const float2 *__restrict__ VEL,
           const float *__restrict__ S, const float *__restrict__ deg,
           const unsigned short *__restrict__ W2b, const float *__restrict__ params,
           const unsigned short *__restrict__ W3b, const float2 *__restrict__ tgt, float inv_n,
           float *__restrict__ dS, unsigned short *__restrict__ U,
           unsigned short *__restrict__ dP3, unsigned short *__restrict__ dM,
           float *__restrict__ grads, float *__restrict__ loss)
{
    __shared__ unsigned short ul[4][16][kKN];
    __shared__ unsigned short dl[4][16][kH];
    __shared__ float du_l[4][16][kH];
    __shared__ __align__(16) NodeW sw;
    stage_node_weights(sw, W2b, W3b, params);
    const unsigned w = threadIdx.x / 64u, l = lane();
    const unsigned ntiles = slots / 16u, nw = gridDim.x * 4u;
    const unsigned long long ck = *clock;
    float gW4[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    float gb4[2] = {0, 0}, gb2[4] = {0, 0, 0, 0}, lsum = 0.0f;
    for (unsigned tile = blockIdx.x * 4u + w; tile < ntiles; tile += nw) {
        const unsigned n0 = tile * 16u;
        const sph_dev_params &P = Pp[BS[n0 / kSphTPB]];
        const unsigned vend = P.poff + (unsigned)P.n;
        unsigned short *Ut = U + (size_t)tile * kKN * 16u;
        unsigned short *Pt = dP3 + (size_t)tile * kH * 16u;
        unsigned short *Mt = dM + (size_t)tile * kH * 16u;
        if (n0 >= vend) { // padding tile: zero operands for the reduction
            for (unsigned x = l; x < kKN * 16u; x += 64)
                Ut[x] = 0;
            for (unsigned x = l; x < kH * 16u; x += 64) {
                Pt[x] = 0;
                Mt[x] = 0;
            }
            continue;
        }
        NodeTile T;
        bf16x4 A[4];
        float degv[4];
        load_s_tile(n0, vend, S, deg, A, degv);
        node_forward(n0, P, ck, VEL, A, degv, sw, ul[w], T);
        const NodeCtx nc = node_ctx(P, ck);
        // dout for the 4 rows of this lane group
        float dout[4][2];
        for (int q = 0; q < 4; q++) {
            const unsigned row = n0 + 4u * (l / 16u) + q;
            const bool ok = row < vend;
            float2 pred = make_float2(0, 0), y = make_float2(0, 0);
            if (ok) {
                const float2 b = ballistic(nc, VEL[row]);
                pred = make_float2(b.x + kDScale * T.out[q][0], b.y + kDScale * T.out[q][1]);
                y = tgt[row];
            }
            // loss = mean over particles of |d - y|^2 / kDScale^2
            const float ex = (pred.x - y.x) / kDScale, ey = (pred.y - y.y) / kDScale;
            dout[q][0] = ok ? 2.0f * ex * inv_n : 0.0f;
            dout[q][1] = ok ? 2.0f * ey * inv_n : 0.0f;
            if (ok && l % 16u == 0)
                lsum += (ex * ex + ey * ey) * inv_n;
        }
        // dW4, db4, dz, dpre3
        for (int t = 0; t < 4; t++) {
            const unsigned c = 16u * t + l % 16u;
            for (int q = 0; q < 4; q++) {
                const float x = T.pre3[t][q];
                const float z = silu(x);
                gW4[0][t] += dout[q][0] * z;
                gW4[1][t] += dout[q][1] * z;
                const float dz = dout[q][0] * sw.W4[c] + dout[q][1] * sw.W4[kH + c];
                const float dp = dz * silu_grad(x);
                dl[w][4u * (l / 16u) + q][c] = f2bf(dp);
            }
            // tile-major dPre3: [unit][16 rows]
            bf16x4 v;
            for (int q = 0; q < 4; q++)
                v[q] = (short)dl[w][4u * (l / 16u) + q][c];
            *(bf16x4 *)(Pt + (size_t)c * 16u + 4u * (l / 16u)) = v;
        }
        if (l % 16u == 0)
            for (int q = 0; q < 4; q++) {
                gb4[0] += dout[q][0];
                gb4[1] += dout[q][1];
            }
        // U tile-major [col][16 rows]
        for (unsigned x = l; x < kKN * 4u; x += 64) {
            const unsigned col = x / 4u, r0 = 4u * (x % 4u);
            bf16x4 v;
            for (int q = 0; q < 4; q++)
                v[q] = (short)ul[w][r0 + q][col];
            *(bf16x4 *)(Ut + (size_t)col * 16u + r0) = v;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // du = dpre3 W3  (16 x 80); only the first 64 columns (dm) are needed
        bf16x4 Ad[4];
        for (int kc = 0; kc < 4; kc++)
            for (int q = 0; q < 4; q++)
                Ad[kc][q] = (short)dl[w][l % 16u][16u * kc + 4u * (l / 16u) + q];
        for (int t = 0; t < 4; t++) {
            f32x4 d = {0, 0, 0, 0};
            for (int kc = 0; kc < 4; kc++) {
                bf16x4 b; // B[k][c] = W3[k][16t + c]  (k: hidden unit)
                const unsigned c = 16u * t + l % 16u, k = 16u * kc + 4u * (l / 16u);
                for (int q = 0; q < 4; q++)
                    b[q] = (short)sw.W3[(k + q) * kKN + c];
                d = mfma(Ad[kc], b, d);
            }
            for (int q = 0; q < 4; q++) {
                const unsigned r = 4u * (l / 16u) + q;
                du_l[w][r][16u * t + l % 16u] = n0 + r < vend ? d[q] : 0.0f;
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // dM tile-major, db2 partials, dS = dm W2
        for (int t = 0; t < 4; t++) {
            const unsigned c = 16u * t + l % 16u;
            bf16x4 v;
            for (int q = 0; q < 4; q++) {
                const unsigned r = 4u * (l / 16u) + q;
                const float dm = du_l[w][r][c];
                v[q] = (short)f2bf(dm);
                if (n0 + r < vend)
                    gb2[t] += deg[n0 + r] * dm;
            }
            *(bf16x4 *)(Mt + (size_t)c * 16u + 4u * (l / 16u)) = v;
        }
        bf16x4 Am[4];
        for (int kc = 0; kc < 4; kc++)
            for (int q = 0; q < 4; q++)
                Am[kc][q] = (short)f2bf(du_l[w][l % 16u][16u * kc + 4u * (l / 16u) + q]);
        for (int t = 0; t < 4; t++) {
            f32x4 d = {0, 0, 0, 0};
            for (int kc = 0; kc < 4; kc++) {
                bf16x4 b; // B[k][c] = W2[k][16t + c]  (k: output unit of layer 2)
                const unsigned c = 16u * t + l % 16u, k = 16u * kc + 4u * (l / 16u);
                for (int q = 0; q < 4; q++)
                    b[q] = (short)sw.W2[(k + q) * kH + c];
                d = mfma(Am[kc], b, d);
            }
            for (int q = 0; q < 4; q++) { // rows past the scene's end get 0
                const unsigned row = n0 + 4u * (l / 16u) + q;
                dS[(size_t)row * kH + 16u * t + l % 16u] = d[q];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // dW4 partials: sum over the 4 row groups
    for (int o = 0; o < 2; o++)
        for (int t = 0; t < 4; t++) {
            float v = gW4[o][t];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (l < 16u)
                atomicAdd(&grads[oW4 + o * kH + 16u * t + l], v);
        }
    for (int t = 0; t < 4; t++) {
        float v = gb2[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (l < 16u)
            atomicAdd(&grads[ob2 + 16u * t + l], v);
    }
    if (l % 16u == 0) {
        atomicAdd(&grads[ob4], gb4[0]);
        atomicAdd(&grads[ob4 + 1], gb4[1]);
        atomicAdd(loss, lsum);
    }
}

// Weight gradients of the node layers by split-K MFMA over particle tiles:
// dW3 (64x80) = sum dPre3^T U, dW2 (64x64) = sum dM^T S.  36 output tiles
// of 16x16; wave w of a workgroup owns output tiles w, w+4, ...
__global__ void __launch_bounds__(256)
k_node_wgrad(unsigned ntiles, const unsigned short *__restrict__ U,
             const unsigned short *__restrict__ dP3, const unsigned short *__restrict__ dM,
             const float *__restrict__ S, float *__restrict__ grads)
{
    const unsigned w = threadIdx.x / 64u, l = lane();
    f32x4 acc[9];
    for (int i = 0; i < 9; i++)
        acc[i] = {0, 0, 0, 0};
    for (unsigned tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const unsigned short *Ut = U + (size_t)tile * kKN * 16u;
        const unsigned short *Pt = dP3 + (size_t)tile * kH * 16u;
        const unsigned short *Mt = dM + (size_t)tile * kH * 16u;
        for (int i = 0; i < 9; i++) {
            const unsigned ot = w + 4u * i; // 0..35
            const unsigned mt = ot < 20u ? ot / 5u : (ot - 20u) / 4u;
            const unsigned nt = ot < 20u ? ot % 5u : (ot - 20u) % 4u;
            // A[r][k] = dX[node k][unit 16mt + r]; B[k][c] = Y[node k][col 16nt + c]
            const unsigned short *Ab = ot < 20u ? Pt : Mt;
            const bf16x4 a = *(const bf16x4 *)(Ab + (size_t)(16u * mt + l % 16u) * 16u + 4u * (l / 16u));
            bf16x4 b;
            if (ot < 20u) {
                b = *(const bf16x4 *)(Ut + (size_t)(16u * nt + l % 16u) * 16u + 4u * (l / 16u));
            } else {
                const unsigned col = 16u * nt + l % 16u;
                const size_t n = (size_t)tile * 16u + 4u * (l / 16u);
                b = pack4(S[n * kH + col], S[(n + 1) * kH + col], S[(n + 2) * kH + col],
                          S[(n + 3) * kH + col]);
            }
            acc[i] = mfma(a, b, acc[i]);
        }
    }
    for (int i = 0; i < 9; i++) {
        const unsigned ot = w + 4u * i;
        const unsigned mt = ot < 20u ? ot / 5u : (ot - 20u) / 4u;
        const unsigned nt = ot < 20u ? ot % 5u : (ot - 20u) % 4u;
        for (int q = 0; q < 4; q++) {
            const unsigned r = 16u * mt + 4u * (l / 16u) + q, c = 16u * nt + l % 16u;
            if (ot < 20u)
                atomicAdd(&grads[oW3 + r * kKN + c], acc[i][q]);
            else
                atomicAdd(&grads[oW2 + r * kH + c], acc[i][q]);
        }
    }
}

// Edge backward: dW1 += sum over edges of (dS_i * SiLU'(W1 f_ij)) f_ij^T.
// Same groups and tiles as the forward; dS_i is expanded to the tile's edges
// by an MFMA of the one-hot edge->particle matrix with the group's dS, whose
// D layout equals that of the recomputed pre-activations, and the product
// with SiLU' is, in that layout, the A operand of dW1 += dPre1^T F.
__global__ void __launch_bounds__(256)
k_edge_bwd(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS, unsigned slots,
           const float2 *__restrict__ POS, const float2 *__restrict__ VEL,
           const unsigned *__restrict__ K, const unsigned *__restrict__ start,
           const unsigned short *__restrict__ W1b, const float *__restrict__ dS,
           float *__restrict__ grads)
{
    __shared__ EdgeBuf Lb[4];
    const unsigned w = threadIdx.x / 64u, l = lane(), p = l % 16u, sub = l / 16u;
    EdgeBuf &L = Lb[w];
    bf16x4 B[4];
    for (int t = 0; t < 4; t++)
        B[t] = w1_b(W1b, t);
    f32x4 g[4];
    for (int t = 0; t < 4; t++)
        g[t] = {0, 0, 0, 0};
    const unsigned ngroups = slots / kGroup, nw = gridDim.x * 4u;
    for (unsigned gi = blockIdx.x * 4u + w; gi < ngroups; gi += nw) {
        const unsigned n0 = gi * kGroup;
        const sph_dev_params &P = Pp[BS[n0 / kSphTPB]];
        if (n0 >= P.poff + (unsigned)P.n)
            continue;
        // B operand of the expansion: dS[particle 4*sub+i][unit 16t+p]
        bf16x4 Bds[4];
        for (int t = 0; t < 4; t++) {
            const float *d0 = dS + (size_t)(n0 + 4u * sub) * kH + 16u * t + p;
            Bds[t] = pack4(d0[0], d0[kH], d0[2 * kH], d0[3 * kH]);
        }
        auto tiles = [&](unsigned ne) {
            for (unsigned e0 = 0; e0 < ne; e0 += 16) {
                const bf16x4 a = edge_tile_a(L, e0);
                // one-hot A: row edge e0 + p, columns particles 4*sub .. +3
                bf16x4 oh;
                const unsigned own = L.owner[e0 + p];
                for (int q = 0; q < 4; q++)
                    oh[q] = own == 4u * sub + q ? (short)0x3f80 : (short)0;
                // B operand of the gradient product: F[edge 4*sub+i][feature p]
                bf16x4 f;
                for (int q = 0; q < 4; q++) {
                    const bf16x4 fr = L.f[e0 + 4u * sub + q][(p / 4u) & 1u];
                    f[q] = p < 8u ? fr[p % 4u] : (short)0;
                }
                for (int t = 0; t < 4; t++) {
                    f32x4 d = {0, 0, 0, 0};
                    d = mfma(a, B[t], d);
                    f32x4 ds = {0, 0, 0, 0};
                    ds = mfma(oh, Bds[t], ds);
                    float gp[4];
                    for (int q = 0; q < 4; q++)
                        gp[q] = ds[q] * silu_grad(d[q]);
                    g[t] = mfma(pack4(gp[0], gp[1], gp[2], gp[3]), f, g[t]);
                }
            }
        };
        float dg, rh;
        walk_group<false>(n0, P, POS, VEL, K, start, L, tiles, dg, rh);
        __builtin_amdgcn_wave_barrier();
    }
    if (p < 6u)
        for (int t = 0; t < 4; t++)
            for (int q = 0; q < 4; q++) {
                const unsigned unit = 16u * t + 4u * sub + q;
                atomicAdd(&grads[oW1 + unit * kKE + p], g[t][q]);
            }
}

// Adam on the fp32 master weights, then refresh the bf16 copies.  Weights
// of the padding columns (W1 beyond 6 inputs, W3 beyond 69) stay zero.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, unsigned short *__restrict__ W1b,
                       unsigned short *__restrict__ W2b, unsigned short *__restrict__ W3b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    bool live = true;
    if (i >= oW1 && i < oW2)
        live = (i - oW1) % kKE < 6;
    else if (i >= oW3 && i < oW4)
        live = (i - oW3) % kKN < 69;
    float x = p[i];
    if (live) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        x -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = x;
    }
    if (i >= oW1 && i < oW2)
        W1b[i - oW1] = f2bf(x);
    else if (i >= oW2 && i < ob2)
        W2b[i - oW2] = f2bf(x);
    else if (i >= oW3 && i < oW4)
        W3b[i - oW3] = f2bf(x);
}

__global__ void k_to_bf16(const float *__restrict__ p, unsigned short *__restrict__ W1b,
                          unsigned short *__restrict__ W2b, unsigned short *__restrict__ W3b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= oW1 && i < oW2)
        W1b[i - oW1] = f2bf(p[i]);
    else if (i >= oW2 && i < ob2)
        W2b[i - oW2] = f2bf(p[i]);
    else if (i >= oW3 && i < oW4)
        W3b[i - oW3] = f2bf(p[i]);
}

// by-id (concatenated caller order) -> slot order, and back.
__global__ void k_by_id_to_slot(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                                unsigned slots, const int *__restrict__ ID,
                                const float2 *__restrict__ by_id, float2 *__restrict__ by_slot)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots)
        return;
    const sph_dev_params &P = Pp[BS[s / kSphTPB]];
    if (s < P.poff + (unsigned)P.n)
        by_slot[s] = by_id[P.uoff + (unsigned)ID[s]];
}

__global__ void k_slot_to_by_id(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                                unsigned slots, const int *__restrict__ ID,
                                const float2 *__restrict__ by_slot, float2 *__restrict__ by_id)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots)
        return;
    const sph_dev_params &P = Pp[BS[s / kSphTPB]];
    if (s < P.poff + (unsigned)P.n)
        by_id[P.uoff + (unsigned)ID[s]] = by_slot[s];
}

// Apply a predicted macro-step: x += d (then the box clamp of
// src/fluid.c:729-743), v = d / (4 dt) + 1.5 a dt (the last step's velocity
// under constant acceleration), clamped like src/fluid.c:613-625.
__global__ void k_apply(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS,
                        unsigned slots, const unsigned long long *__restrict__ clock,
                        const float2 *__restrict__ disp, float2 *__restrict__ POS,
                        float2 *__restrict__ VEL)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots)
        return;
    const sph_dev_params &P = Pp[BS[s / kSphTPB]];
    if (s >= P.poff + (unsigned)P.n || P.frozen)
        return;
    const NodeCtx nc = node_ctx(P, *clock);
    const float2 d = disp[s];
    float2 p = POS[s];
    p.x += d.x;
    p.y += d.y;
    if (p.x < P.min_x)
        p.x = P.min_x;
    else if (p.x > P.max_x)
        p.x = P.max_x - 0.001f;
    if (p.y < P.min_y)
        p.y = P.min_y;
    else if (p.y > P.max_y)
        p.y = P.max_y - 0.001f;
    float vx = d.x / (4.0f * nc.dt) + 1.5f * nc.fx * nc.dt;
    float vy = d.y / (4.0f * nc.dt) + 1.5f * nc.g_neg * nc.dt;
    sph_check_velocity(vx, vy);
    POS[s] = p;
    VEL[s] = make_float2(vx, vy);
}

__global__ void k_clock_add(unsigned long long *clock, unsigned long long n)
{
    *clock += n; // one thread: a plain vector load/store pair
}

// Per-scene density sums and maxima of the last density pass: one
// workgroup per scene over its slots, no atomics (deterministic).
__global__ void __launch_bounds__(256)
k_scene_density(const sph_dev_params *__restrict__ Pp, const float2 *__restrict__ RHO,
                float *__restrict__ sum, float *__restrict__ maxv)
{
    __shared__ float ps[4], pm[4];
    const sph_dev_params &P = Pp[blockIdx.x];
    float t = 0.0f, m = 0.0f;
    for (unsigned s = P.poff + threadIdx.x; s < P.poff + (unsigned)P.n; s += blockDim.x) {
        const float r = RHO[s].x;
        t += r;
        m = fmaxf(m, r);
    }
    for (int x = 32; x > 0; x >>= 1) {
        t += __shfl_xor(t, x, 64);
        m = fmaxf(m, __shfl_xor(m, x, 64));
    }
    if (lane() == 0) {
        ps[threadIdx.x / 64u] = t;
        pm[threadIdx.x / 64u] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sum[blockIdx.x] = (ps[0] + ps[1]) + (ps[2] + ps[3]);
        maxv[blockIdx.x] = fmaxf(fmaxf(pm[0], pm[1]), fmaxf(pm[2], pm[3]));
    }
}

// Density of every particle by caller index (the state is canonical: RHO
// in ID0 order).
__global__ void __launch_bounds__(256)
k_rho_by_id(const sph_dev_params *__restrict__ Pp, const int *__restrict__ BS, unsigned slots,
            const float2 *__restrict__ RHO, const int *__restrict__ ID, float *__restrict__ out)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots)
        return;
    const sph_dev_params &P = Pp[BS[s / 256u]];
    if (s >= P.poff + (unsigned)P.n)
        return;
    out[P.uoff + (unsigned)ID[s]] = RHO[s].x;
}

// Per shadowed scene (flags[s] != 0): sum over its particles of
// (sur - plain)^2 and of plain^2, for the relative RMS density error.
__global__ void __launch_bounds__(256)
k_shadow_err(const sph_dev_params *__restrict__ Pp, const int *__restrict__ flags,
             const float *__restrict__ plain, const float *__restrict__ sur,
             float *__restrict__ out)
{
    __shared__ float pa[4], pb[4];
    const sph_dev_params &P = Pp[blockIdx.x];
    if (!flags[blockIdx.x])
        return;
    float a = 0.0f, b = 0.0f;
    for (unsigned i = threadIdx.x; i < (unsigned)P.n; i += blockDim.x) {
        const float x = plain[P.uoff + i], y = sur[P.uoff + i];
        a += (y - x) * (y - x);
        b += x * x;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if (lane() == 0) {
        pa[threadIdx.x / 64u] = a;
        pb[threadIdx.x / 64u] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = (pa[0] + pa[1]) + (pa[2] + pa[3]);
        out[2 * blockIdx.x + 1] = (pb[0] + pb[1]) + (pb[2] + pb[3]);
    }
}

// Judge of a macro-step.  Per scene, ref[s] is the mean density of the
// current state and dref[s] = ref[nscenes + s] the change of the scene's last
// macro-step computed by the plain solver (its density trend).  A predicted
// state passes when its mean density is within tol (relative) of
// ref + dref, the plain solver's trend extrapolated by one macro-step; a
// passing scene takes the new mean as ref.  On input flags[s] = 1 forces a
// re-take (a periodic plain macro-step that refreshes the trend); on output
// flags[s] = 1 for a failed scene, 2 for a forced one.  mode 1 (set) starts
// every scene at the current mean with no trend; mode 2 (after re-takes)
// gives each re-taken scene (flags[s] != 0) the plain change as its trend
// and the plain mean as ref.
__global__ void k_judge(const sph_dev_params *__restrict__ Pp, int nscenes,
                        const float *__restrict__ sum, float *__restrict__ ref, float tol,
                        int mode, int *__restrict__ flags)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscenes)
        return;
    float *dref = ref + nscenes;
    const int n = Pp[s].n;
    const float mean = n > 0 ? sum[s] / (float)n : 0.0f;
    if (mode == 1) {
        ref[s] = mean;
        dref[s] = 0.0f;
        return;
    }
    if (mode == 2) {
        if (flags[s]) {
            dref[s] = mean - ref[s];
            ref[s] = mean;
        }
        return;
    }
    const float expect = ref[s] + dref[s];
    const bool bad = n > 0 && !(fabsf(mean - expect) <= tol * expect);
    const bool forced = flags[s] != 0;
    flags[s] = forced ? 2 : (bad ? 1 : 0);
    if (!bad && !forced)
        ref[s] = mean;
}

// Trajectory sampling: sample k takes frame f = frame[k] of scene s =
// scene[k]; staging gets its (x, y, v_x, v_y) records, target the 4-step
// displacement to frame f + 1.  data is [frames][nscenes][n] float4.
__global__ void k_traj_sample(const float4 *__restrict__ data, unsigned nscenes, unsigned n,
                              unsigned count, const int *__restrict__ scene,
                              const int *__restrict__ frame, float4 *__restrict__ staging,
                              float2 *__restrict__ target)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)count * n)
        return;
    const unsigned k = (unsigned)(i / n), j = (unsigned)(i % n);
    const size_t src = ((size_t)frame[k] * nscenes + (unsigned)scene[k]) * n + j;
    const float4 a = data[src], b = data[src + (size_t)nscenes * n];
    staging[i] = a;
    target[i] = make_float2(b.x - a.x, b.y - a.y);
}

// ---------------------------------------------------------------------------
// host side

namespace {

thread_local char m_err[256] = "";

int merr(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(m_err, sizeof m_err, fmt, ap);
    va_end(ap);
    return code;
}

#define MCHK(call)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return merr(SPH_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));                \
    } while (0)

template <class T>
int malloc_dev(T **p, size_t n)
{
    if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess)
        return merr(SPH_ENOMEM, "hipMalloc of %zu x %zu B failed", n, sizeof(T));
    return SPH_OK;
}

int ensure_buffers(sph_surrogate *m, size_t slots)
{
    if (m->slots >= slots)
        return SPH_OK;
    void *old[] = {m->S, m->deg, m->dS, m->disp, m->tgt, m->U, m->dP3, m->dM, m->tmp2};
    for (void *p : old)
        if (p)
            (void)hipFree(p);
    m->S = m->deg = m->dS = m->disp = m->tgt = nullptr;
    m->U = m->dP3 = m->dM = nullptr;
    m->tmp2 = nullptr;
    int rc = SPH_OK;
    if (!rc) rc = malloc_dev(&m->S, slots * kH);
    if (!rc) rc = malloc_dev(&m->deg, slots);
    if (!rc) rc = malloc_dev(&m->dS, slots * kH);
    if (!rc) rc = malloc_dev(&m->disp, slots * 2);
    if (!rc) rc = malloc_dev(&m->tgt, slots * 2);
    if (!rc) rc = malloc_dev(&m->U, slots * kKN);
    if (!rc) rc = malloc_dev(&m->dP3, slots * kH);
    if (!rc) rc = malloc_dev(&m->dM, slots * kH);
    if (!rc) rc = malloc_dev(&m->tmp2, slots);
    m->slots = rc ? 0 : slots;
    return rc;
}

int ensure_judge(sph_surrogate *m, const sph_ctx *c)
{
    const size_t ns = c->sc.size();
    if (m->ns >= ns)
        return SPH_OK;
    void *old[] = {m->ref, m->jsum, m->jmax, m->jflags, m->sh_sum, m->sh_err};
    for (void *p : old)
        if (p)
            (void)hipFree(p);
    m->ref = m->jsum = m->jmax = m->sh_sum = m->sh_err = nullptr;
    m->jflags = nullptr;
    m->ref_ok = false;
    int rc = SPH_OK;
    if (!rc) rc = malloc_dev(&m->ref, 2 * ns); // ref, then the per-scene trend
    if (!rc) rc = malloc_dev(&m->jsum, ns);
    if (!rc) rc = malloc_dev(&m->jmax, ns);
    if (!rc) rc = malloc_dev(&m->jflags, ns);
    if (!rc) rc = malloc_dev(&m->sh_sum, ns);
    if (!rc) rc = malloc_dev(&m->sh_err, 2 * ns);
    m->ns = rc ? 0 : ns;
    return rc;
}

// Per-scene density sums (and maxima) of RHO into m->jsum / m->jmax.
int scene_sums(sph_surrogate *m, sph_ctx *c)
{
    k_scene_density<<<dim3((unsigned)c->sc.size()), 256, 0, c->st>>>(c->dP, c->RHO, m->jsum,
                                                                       m->jmax);
    MCHK(hipGetLastError());
    return SPH_OK;
}

// Density of the current (hashed) state into RHO.
int density_now(sph_ctx *c)
{
    if (int rc = sph_internal_canonicalize(c))
        return rc;
    if (!c->lists_valid)
        return merr(SPH_ESTATE, "surrogate: state not hashed");
    sph_launch_density(c->st, c->slots, c->bucket_cap != 0, c->dP, c->BS, c->POS0, c->K0,
                       c->start, c->RHO, c->MASK, c->dC, c->lanes);
    MCHK(hipGetLastError());
    return SPH_OK;
}

// Prediction of the 4-step displacement of every particle (fused kernel).
int forward(sph_surrogate *m, sph_ctx *c)
{
    if (!c->lists_valid)
        return merr(SPH_ESTATE, "surrogate: the context has no hashed state (step or hash first)");
    if (int rc = ensure_buffers(m, c->slots))
        return rc;
    k_infer<<<dim3(2048), 256, 0, c->st>>>(c->dP, c->BS, c->slots, c->dclock, c->POS0, c->VEL,
                                           c->K0, c->start, m->W1b, m->W2b, m->params, m->W3b,
                                           (float2 *)m->disp);
    MCHK(hipGetLastError());
    m->have_pred = true;
    return SPH_OK;
}

// The edge pass of training: per-particle sums S and degrees for the node
// backward (which recomputes the node forward itself).
int forward_edges(sph_surrogate *m, sph_ctx *c)
{
    if (!c->lists_valid)
        return merr(SPH_ESTATE, "surrogate: the context has no hashed state (step or hash first)");
    if (int rc = ensure_buffers(m, c->slots))
        return rc;
    k_edge_fwd<false><<<dim3(2048), 256, 0, c->st>>>(c->dP, c->BS, c->slots, c->POS0, c->VEL,
                                                     c->K0, c->start, m->W1b, m->S, m->deg,
                                                     nullptr);
    MCHK(hipGetLastError());
    return SPH_OK;
}

} // namespace

extern "C" {

const char *sph_surrogate_last_error(void) { return m_err; }

int sph_surrogate_num_params(void) { return kNParams; }

int sph_surrogate_create(sph_surrogate **out, uint64_t seed)
{
    if (!out)
        return merr(SPH_EINVAL, "null output");
    *out = nullptr;
    sph_surrogate *m = new (std::nothrow) sph_surrogate();
    if (!m)
        return merr(SPH_ENOMEM, "host allocation failed");
    int rc = SPH_OK;
    if (!rc) rc = malloc_dev(&m->params, kNParams);
    if (!rc) rc = malloc_dev(&m->grads, kNParams);
    if (!rc) rc = malloc_dev(&m->adam_m, kNParams);
    if (!rc) rc = malloc_dev(&m->adam_v, kNParams);
    if (!rc) rc = malloc_dev(&m->W1b, kH * kKE);
    if (!rc) rc = malloc_dev(&m->W2b, kH * kH);
    if (!rc) rc = malloc_dev(&m->W3b, kH * kKN);
    if (!rc) rc = malloc_dev(&m->loss, 1);
    if (rc) {
        sph_surrogate_destroy(m);
        return rc;
    }
    // Initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from SplitMix64,
    // padding columns zero, the output layer scaled down so the untrained
    // network starts close to the ballistic estimate.
    std::vector<float> p(kNParams, 0.0f);
    uint64_t st = seed;
    auto uni = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)(z >> 40) * 0x1p-24f * 2.0f - 1.0f;
    };
    for (int r = 0; r < kH; r++)
        for (int k = 0; k < 6; k++)
            p[oW1 + r * kKE + k] = uni() / sqrtf(6.0f);
    for (int i = 0; i < kH * kH; i++)
        p[oW2 + i] = uni() / sqrtf((float)kH);
    for (int r = 0; r < kH; r++)
        for (int k = 0; k < 69; k++)
            p[oW3 + r * kKN + k] = uni() / sqrtf(69.0f);
    for (int i = 0; i < 2 * kH; i++)
        p[oW4 + i] = 0.1f * uni() / sqrtf((float)kH);
    if (hipMemcpy(m->params, p.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->grads, 0, sizeof(float) * kNParams) != hipSuccess ||
        hipMemset(m->adam_m, 0, sizeof(float) * kNParams) != hipSuccess ||
        hipMemset(m->adam_v, 0, sizeof(float) * kNParams) != hipSuccess) {
        sph_surrogate_destroy(m);
        return merr(SPH_EHIP, "surrogate initialisation failed");
    }
    k_to_bf16<<<dim3((kNParams + 255) / 256), 256>>>(m->params, m->W1b, m->W2b, m->W3b);
    if (hipDeviceSynchronize() != hipSuccess) {
        sph_surrogate_destroy(m);
        return merr(SPH_EHIP, "surrogate initialisation failed");
    }
    *out = m;
    return SPH_OK;
}

void sph_surrogate_destroy(sph_surrogate *m)
{
    if (!m)
        return;
    void *ptrs[] = {m->params, m->grads, m->adam_m, m->adam_v, m->W1b, m->W2b, m->W3b, m->S,
                    m->deg,    m->dS,    m->disp,   m->tgt,    m->U,   m->dP3, m->dM,  m->loss,
                    m->tmp2, m->ref, m->jsum, m->jmax, m->jflags, m->sh_sum,
                    m->rho_plain, m->rho_sur, m->sh_err};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    delete m;
}

int sph_surrogate_get_params(sph_surrogate *m, float *out)
{
    if (!m || !out)
        return merr(SPH_EINVAL, "null argument");
    MCHK(hipDeviceSynchronize());
    MCHK(hipMemcpy(out, m->params, sizeof(float) * kNParams, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_surrogate_set_params(sph_surrogate *m, const float *in)
{
    if (!m || !in)
        return merr(SPH_EINVAL, "null argument");
    MCHK(hipDeviceSynchronize());
    MCHK(hipMemcpy(m->params, in, sizeof(float) * kNParams, hipMemcpyHostToDevice));
    k_to_bf16<<<dim3((kNParams + 255) / 256), 256>>>(m->params, m->W1b, m->W2b, m->W3b);
    MCHK(hipDeviceSynchronize());
    return SPH_OK;
}

int sph_surrogate_get_grads(sph_surrogate *m, float *out)
{
    if (!m || !out)
        return merr(SPH_EINVAL, "null argument");
    MCHK(hipDeviceSynchronize());
    MCHK(hipMemcpy(out, m->grads, sizeof(float) * kNParams, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_surrogate_zero_grads(sph_surrogate *m)
{
    if (!m)
        return merr(SPH_EINVAL, "null argument");
    MCHK(hipMemset(m->grads, 0, sizeof(float) * kNParams));
    return SPH_OK;
}

// Forward on the context's current hashed state; the predicted displacement
// per particle (caller order, 2 floats each) is copied out if disp_by_id.
int sph_surrogate_predict(sph_surrogate *m, sph_ctx *c, float *disp_by_id)
{
    if (!m || !c)
        return merr(SPH_EINVAL, "null argument");
    if (int rc = forward(m, c))
        return rc;
    if (disp_by_id && c->n_total) {
        k_slot_to_by_id<<<dim3(c->slots / 256), 256, 0, c->st>>>(
            c->dP, c->BS, c->slots, c->ID0, (const float2 *)m->disp, m->tmp2);
        MCHK(hipMemcpyAsync(disp_by_id, m->tmp2, sizeof(float2) * (size_t)c->n_total,
                            hipMemcpyDeviceToHost, c->st));
        MCHK(hipStreamSynchronize(c->st));
    }
    return SPH_OK;
}

// One gradient evaluation: forward, loss against target displacements
// (a device buffer in caller order, 2 floats per particle), backward; gradients are accumulated into the gradient
// buffer (zero it with sph_surrogate_zero_grads).  *loss gets the mean of
// |d - target|^2 / 0.05^2 over the particles.
static int backward_dev(sph_surrogate *m, sph_ctx *c, const float2 *target_dev, float *loss)
{
    if (c->n_total == 0)
        return SPH_OK;
    if (int rc = forward_edges(m, c))
        return rc;
    const unsigned S = c->slots;
    k_by_id_to_slot<<<dim3(S / 256), 256, 0, c->st>>>(c->dP, c->BS, S, c->ID0, target_dev,
                                                     (float2 *)m->tgt);
    MCHK(hipMemsetAsync(m->loss, 0, sizeof(float), c->st));
    k_node_bwd<<<dim3(2048), 256, 0, c->st>>>(c->dP, c->BS, S, c->dclock, c->VEL, m->S, m->deg,
                                              m->W2b, m->params, m->W3b, (const float2 *)m->tgt,
                                              1.0f / (float)c->n_total, m->dS, m->U, m->dP3,
                                              m->dM, m->grads, m->loss);
    k_node_wgrad<<<dim3(1024), 256, 0, c->st>>>(S / 16u, m->U, m->dP3, m->dM, m->S, m->grads);
    k_edge_bwd<<<dim3(2048), 256, 0, c->st>>>(c->dP, c->BS, S, c->POS0, c->VEL, c->K0, c->start,
                                              m->W1b, m->dS, m->grads);
    MCHK(hipGetLastError());
    if (loss) {
        MCHK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, c->st));
        MCHK(hipStreamSynchronize(c->st));
    }
    return SPH_OK;
}

int sph_surrogate_backward(sph_surrogate *m, sph_ctx *c, const float *target_by_id, float *loss)
{
    if (!m || !c || (!target_by_id && c->n_total))
        return merr(SPH_EINVAL, "null argument");
    if (c->n_total == 0)
        return SPH_OK;
    if (int rc = ensure_buffers(m, c->slots))
        return rc;
    MCHK(hipStreamSynchronize(c->st)); // tmp2 may still be read by earlier work
    MCHK(hipMemcpy(m->tmp2, target_by_id, sizeof(float2) * (size_t)c->n_total,
                   hipMemcpyHostToDevice));
    return backward_dev(m, c, m->tmp2, loss);
}

int sph_surrogate_backward_trajectory(sph_surrogate *m, sph_ctx *c, sph_trajectory *tr,
                                      float *loss)
{
    if (!m || !c || !tr)
        return merr(SPH_EINVAL, "null argument");
    if (tr->loaded != (int)c->sc.size() || (long)tr->loaded * tr->n != (long)c->n_total)
        return merr(SPH_ESTATE, "the context does not hold the trajectory's last loaded samples");
    return backward_dev(m, c, tr->target, loss);
}

// ---- trajectory store ------------------------------------------------------

int sph_trajectory_create(sph_trajectory **out, int nscenes, int n, int frames)
{
    if (!out || nscenes <= 0 || n <= 0 || frames < 2)
        return merr(SPH_EINVAL, "sph_trajectory_create: bad sizes");
    *out = nullptr;
    sph_trajectory *t = new (std::nothrow) sph_trajectory();
    if (!t)
        return merr(SPH_ENOMEM, "host allocation failed");
    t->nscenes = nscenes;
    t->n = n;
    t->frames = frames;
    if (int rc = malloc_dev(&t->data, (size_t)frames * nscenes * n)) {
        sph_trajectory_destroy(t);
        return rc;
    }
    *out = t;
    return SPH_OK;
}

void sph_trajectory_destroy(sph_trajectory *t)
{
    if (!t)
        return;
    void *ptrs[] = {t->data, t->staging, t->target, t->idx};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    delete t;
}

int sph_trajectory_record(sph_trajectory *t, sph_ctx *c, int frame)
{
    if (!t || !c || frame < 0 || frame >= t->frames)
        return merr(SPH_EINVAL, "sph_trajectory_record: bad argument");
    if ((int)c->sc.size() != t->nscenes || (long)c->n_total != (long)t->nscenes * t->n)
        return merr(SPH_EINVAL, "sph_trajectory_record: context is not %d scenes x %d particles",
                    t->nscenes, t->n);
    for (const auto &S : c->sc)
        if (S.n != t->n)
            return merr(SPH_EINVAL, "sph_trajectory_record: scenes of unequal size");
    if (sph_gather_state_dev(c, (float *)(t->data + (size_t)frame * t->nscenes * t->n)))
        return merr(SPH_EHIP, "gather failed: %s", sph_last_error());
    return SPH_OK;
}

int sph_trajectory_read(sph_trajectory *t, int frame, float *out)
{
    if (!t || !out || frame < 0 || frame >= t->frames)
        return merr(SPH_EINVAL, "sph_trajectory_read: bad argument");
    MCHK(hipDeviceSynchronize());
    MCHK(hipMemcpy(out, t->data + (size_t)frame * t->nscenes * t->n,
                   sizeof(float4) * (size_t)t->nscenes * t->n, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_trajectory_load(sph_trajectory *t, sph_ctx *c, int count, const int *scene,
                        const int *frame)
{
    if (!t || !c || !scene || !frame || count <= 0)
        return merr(SPH_EINVAL, "sph_trajectory_load: bad argument");
    if ((int)c->sc.size() != count || (long)c->n_total != (long)count * t->n)
        return merr(SPH_EINVAL, "sph_trajectory_load: context must hold %d scenes of %d particles",
                    count, t->n);
    for (const auto &S : c->sc)
        if (S.n != t->n)
            return merr(SPH_EINVAL, "sph_trajectory_load: scenes of unequal size");
    for (int k = 0; k < count; k++)
        if (scene[k] < 0 || scene[k] >= t->nscenes || frame[k] < 0 || frame[k] + 1 >= t->frames)
            return merr(SPH_EINVAL, "sph_trajectory_load: sample %d (scene %d, frame %d) out of range",
                        k, scene[k], frame[k]);
    if (t->cap < count) {
        void *old[] = {t->staging, t->target, t->idx};
        for (void *p : old)
            if (p)
                (void)hipFree(p);
        t->staging = nullptr;
        t->target = nullptr;
        t->idx = nullptr;
        t->cap = 0;
        int rc = malloc_dev(&t->staging, (size_t)count * t->n);
        if (!rc) rc = malloc_dev(&t->target, (size_t)count * t->n);
        if (!rc) rc = malloc_dev(&t->idx, 2 * (size_t)count);
        if (rc)
            return rc;
        t->cap = count;
    }
    std::vector<int> hidx((size_t)2 * count);
    memcpy(hidx.data(), scene, sizeof(int) * count);
    memcpy(hidx.data() + count, frame, sizeof(int) * count);
    MCHK(hipStreamSynchronize(c->st));
    MCHK(hipMemcpyAsync(t->idx, hidx.data(), sizeof(int) * 2 * count, hipMemcpyHostToDevice, c->st));
    const size_t tot = (size_t)count * t->n;
    k_traj_sample<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, c->st>>>(
        t->data, (unsigned)t->nscenes, (unsigned)t->n, (unsigned)count, t->idx, t->idx + count,
        t->staging, t->target);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(c->st));
    if (sph_load_state_dev(c, (const float *)t->staging))
        return merr(SPH_EHIP, "load failed: %s", sph_last_error());
    t->loaded = count;
    return SPH_OK;
}

int sph_surrogate_adam(sph_surrogate *m, float lr, float beta1, float beta2, float eps)
{
    if (!m)
        return merr(SPH_EINVAL, "null argument");
    m->adam_t++;
    const float c1 = 1.0f - powf(beta1, (float)m->adam_t);
    const float c2 = 1.0f - powf(beta2, (float)m->adam_t);
    k_adam<<<dim3((kNParams + 255) / 256), 256>>>(m->params, m->grads, m->adam_m, m->adam_v,
                                                  kNParams, lr, beta1, beta2, eps, c1, c2, m->W1b,
                                                  m->W2b, m->W3b);
    MCHK(hipGetLastError());
    MCHK(hipMemset(m->grads, 0, sizeof(float) * kNParams));
    MCHK(hipDeviceSynchronize());
    return SPH_OK;
}

// Apply the last prediction to the context (positions and velocities move
// four steps); frozen scenes are left alone.  The state must be hashed
// again before the next step (the context knows this).
int sph_surrogate_apply(sph_surrogate *m, sph_ctx *c)
{
    if (!m || !c)
        return merr(SPH_EINVAL, "null argument");
    if (!m->have_pred)
        return merr(SPH_ESTATE, "sph_surrogate_apply: no prediction");
    if (int rc = sph_internal_canonicalize(c))
        return rc;
    const unsigned S = c->slots;
    k_apply<<<dim3(S / 256), 256, 0, c->st>>>(c->dP, c->BS, S, c->dclock,
                                             (const float2 *)m->disp, c->POS0, c->VEL);
    MCHK(hipGetLastError());
    k_clock_add<<<1, 1, 0, c->st>>>(c->dclock, 4ull);
    MCHK(hipGetLastError());
    c->lists_valid = false;
    m->have_pred = false;
    return SPH_OK;
}

// One surrogate macro-step (four plain steps of simulated time) of every
// scene, judged by the plain solver.  The predicted state is hashed and its
// density computed; a scene whose mean density is off by more than tol
// (relative) from the plain solver's trend (its mean density plus the change
// of its last plain macro-step) is restored and re-taken with four plain
// steps.  Every scene is also re-taken plainly on its first macro-step and
// then on a fixed schedule, one macro-step in kTrendRefresh (staggered over
// the scenes), which keeps its trend current whatever the failures.
// With shadow judging on (sph_surrogate_set_shadow), one scene in every
// shadow_every (rotating) first takes the macro-step plainly from the same
// start: its surrogate result is then judged against that plain result as
// well (error above tol: re-taken, flag 3), and the error of the steps the
// trend judge accepted is recorded (sph_surrogate_shadow_stats).
// flags (optional, nscenes ints) receives 1 for failed scenes, 2 for
// scheduled re-takes, 3 for shadow failures, 0 for accepted ones;
// *nfallback the number re-taken.
int sph_surrogate_macro_step(sph_surrogate *m, sph_ctx *c, float tol, int *flags, int *nfallback)
{
    if (!m || !c)
        return merr(SPH_EINVAL, "null argument");
    if (!(tol >= 0.0f))
        return merr(SPH_EINVAL, "tolerance must be >= 0");
    const int ns = (int)c->sc.size();
    if (nfallback)
        *nfallback = 0;
    if (c->n_total == 0)
        return SPH_OK;
    if (int rc = ensure_judge(m, c))
        return rc;
    if (!c->lists_valid)
        if (sph_phase_hash(c, 0))
            return merr(SPH_EHIP, "hash failed: %s", sph_last_error());
    const dim3 jg((ns + 255) / 256);
    bool first = false;
    if (!m->ref_ok || m->ref_ctx != c) {
        if (int rc = density_now(c))
            return rc;
        if (int rc = scene_sums(m, c))
            return rc;
        k_judge<<<jg, 256, 0, c->st>>>(c->dP, ns, m->jsum, m->ref, tol, 1, nullptr);
        MCHK(hipGetLastError());
        m->ref_ok = true;
        m->ref_ctx = c;
        first = true;
    }
    const unsigned long long k = m->macro++;
    std::vector<int> hf((size_t)ns), shadow((size_t)ns, 0);
    int nshadow = 0;
    for (int s = 0; s < ns; s++) {
        hf[s] = first || (k + (unsigned)s) % kTrendRefresh == 0 ? 1 : 0;
        if (m->shadow_every > 0 && !hf[s] && (k + (unsigned)s) % (unsigned)m->shadow_every == 0 &&
            c->sc[s].n > 0) {
            shadow[s] = 1;
            nshadow++;
        }
    }
    if (sph_snapshot_save(c))
        return merr(SPH_EHIP, "snapshot failed: %s", sph_last_error());
    const unsigned S = c->slots;
    if (nshadow) { // the shadowed scenes' plain macro-step, then back to the start
        if (!m->rho_plain) {
            if (int rc = malloc_dev(&m->rho_plain, S))
                return rc;
            if (int rc = malloc_dev(&m->rho_sur, S))
                return rc;
        }
        std::vector<int> frozen((size_t)ns), none((size_t)ns, 0);
        for (int s = 0; s < ns; s++)
            frozen[s] = shadow[s] ? 0 : 1;
        if (sph_set_scene_frozen(c, frozen.data()) || sph_step(c, 4) ||
            sph_set_scene_frozen(c, none.data()))
            return merr(SPH_EHIP, "shadow step failed: %s", sph_last_error());
        if (int rc = density_now(c))
            return rc;
        k_scene_density<<<dim3((unsigned)ns), 256, 0, c->st>>>(c->dP, c->RHO, m->sh_sum, m->jmax);
        k_rho_by_id<<<dim3(S / 256), 256, 0, c->st>>>(c->dP, c->BS, S, c->RHO, c->ID0,
                                                      m->rho_plain);
        MCHK(hipGetLastError());
        if (sph_snapshot_restore(c, shadow.data()))
            return merr(SPH_EHIP, "shadow restore failed: %s", sph_last_error());
    }
    if (int rc = forward(m, c))
        return rc;
    if (int rc = sph_surrogate_apply(m, c))
        return rc;
    if (sph_phase_hash(c, 1))
        return merr(SPH_EHIP, "hash failed: %s", sph_last_error());
    if (int rc = scene_sums(m, c))
        return rc;
    MCHK(hipMemcpyAsync(m->jflags, hf.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->st));
    k_judge<<<jg, 256, 0, c->st>>>(c->dP, ns, m->jsum, m->ref, tol, 0, m->jflags);
    MCHK(hipGetLastError());
    MCHK(hipMemcpyAsync(hf.data(), m->jflags, sizeof(int) * ns, hipMemcpyDeviceToHost, c->st));
    if (nshadow) {
        std::vector<float> sur((size_t)ns), pl((size_t)ns), err(2 * (size_t)ns);
        k_rho_by_id<<<dim3(S / 256), 256, 0, c->st>>>(c->dP, c->BS, S, c->RHO, c->ID0, m->rho_sur);
        MCHK(hipMemcpyAsync(m->jflags, shadow.data(), sizeof(int) * ns, hipMemcpyHostToDevice,
                            c->st));
        k_shadow_err<<<dim3((unsigned)ns), 256, 0, c->st>>>(c->dP, m->jflags, m->rho_plain,
                                                            m->rho_sur, m->sh_err);
        MCHK(hipGetLastError());
        MCHK(hipMemcpyAsync(sur.data(), m->jsum, sizeof(float) * ns, hipMemcpyDeviceToHost, c->st));
        MCHK(hipMemcpyAsync(pl.data(), m->sh_sum, sizeof(float) * ns, hipMemcpyDeviceToHost, c->st));
        MCHK(hipMemcpyAsync(err.data(), m->sh_err, sizeof(float) * 2 * ns, hipMemcpyDeviceToHost,
                            c->st));
        MCHK(hipStreamSynchronize(c->st));
        for (int s = 0; s < ns; s++) {
            if (!shadow[s])
                continue;
            const double e = ((double)sur[s] - (double)pl[s]) / (double)pl[s];
            const double rms = sqrt((double)err[2 * s] / (double)err[2 * s + 1]);
            m->sh_stats[0] += 1.0;
            if (hf[s] == 0) { // accepted by the trend judge: the error it let through
                m->sh_stats[1] += 1.0;
                m->sh_stats[2] += fabs(e) > tol ? 1.0 : 0.0;
                m->sh_stats[3] += fabs(e);
                m->sh_stats[4] = fmax(m->sh_stats[4], fabs(e));
                m->sh_stats[5] += e;
                m->sh_stats[6] += rms;
                m->sh_stats[7] = fmax(m->sh_stats[7], rms);
                if (fabs(e) > tol)
                    hf[s] = 3;
            }
        }
        // the judge's mean for a shadow failure is the surrogate's; the
        // re-take below replaces it with the plain one (mode 2)
    } else {
        MCHK(hipStreamSynchronize(c->st));
    }
    int nf = 0;
    for (int s = 0; s < ns; s++)
        nf += hf[s] != 0;
    if (nf) {
        std::vector<int> retake((size_t)ns), frozen((size_t)ns), none((size_t)ns, 0);
        for (int s = 0; s < ns; s++) {
            retake[s] = hf[s] ? 1 : 0;
            frozen[s] = hf[s] ? 0 : 1;
        }
        if (sph_snapshot_restore(c, retake.data()) || sph_set_scene_frozen(c, frozen.data()) ||
            sph_step(c, 4) || sph_set_scene_frozen(c, none.data()))
            return merr(SPH_EHIP, "fallback failed: %s", sph_last_error());
        if (int rc = density_now(c))
            return rc;
        if (int rc = scene_sums(m, c))
            return rc;
        MCHK(hipMemcpyAsync(m->jflags, retake.data(), sizeof(int) * ns, hipMemcpyHostToDevice,
                            c->st));
        k_judge<<<jg, 256, 0, c->st>>>(c->dP, ns, m->jsum, m->ref, tol, 2, m->jflags);
        MCHK(hipGetLastError());
        MCHK(hipStreamSynchronize(c->st));
    }
    if (flags)
        memcpy(flags, hf.data(), sizeof(int) * ns);
    if (nfallback)
        *nfallback = nf;
    return SPH_OK;
}

int sph_surrogate_set_shadow(sph_surrogate *m, int every)
{
    if (!m || every < 0)
        return merr(SPH_EINVAL, "sph_surrogate_set_shadow: null model or negative period");
    m->shadow_every = every;
    for (double &x : m->sh_stats)
        x = 0.0;
    return SPH_OK;
}

int sph_surrogate_shadow_stats(sph_surrogate *m, double *out)
{
    if (!m || !out)
        return merr(SPH_EINVAL, "null argument");
    for (int i = 0; i < 8; i++)
        out[i] = m->sh_stats[i];
    return SPH_OK;
}

int sph_surrogate_mfma_selftest(float *out256)
{
    if (!out256)
        return merr(SPH_EINVAL, "null argument");
    float *d = nullptr;
    MCHK(hipMalloc(&d, 256 * sizeof(float)));
    k_mfma_selftest<<<1, 64>>>(d);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(out256, d, 256 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess)
        return merr(SPH_EHIP, "mfma selftest: %s", hipGetErrorString(e));
    return SPH_OK;
}

// Per-scene mean and maximum density of the last density pass (sph_step's
// or sph_phase_hash(ctx, 1)'s).
int sph_scene_density(sph_ctx *c, float *mean, float *maxv)
{
    if (!c || !mean || !maxv)
        return merr(SPH_EINVAL, "null argument");
    if (int rc = sph_internal_canonicalize(c))
        return rc;
    const size_t ns = c->sc.size();
    float *buf = nullptr;
    if (int rc = malloc_dev(&buf, 2 * ns))
        return rc;
    k_scene_density<<<dim3((unsigned)ns), 256, 0, c->st>>>(c->dP, c->RHO, buf, buf + ns);
    hipError_t e = hipGetLastError();
    std::vector<float> h(2 * ns);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->st);
    if (e == hipSuccess)
        e = hipMemcpy(h.data(), buf, sizeof(float) * 2 * ns, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return merr(SPH_EHIP, "sph_scene_density: %s", hipGetErrorString(e));
    for (size_t s = 0; s < ns; s++) {
        mean[s] = c->sc[s].n ? h[s] / (float)c->sc[s].n : 0.0f;
        maxv[s] = h[ns + s];
    }
    return SPH_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// sph_scene_density without allocation.  The per-scene partial results (the
// density sum and maximum of every scene) go to a fixed device array instead
// of a buffer allocated per call.  The array's name is a host-side stand-in
// for its device copy: host code only ever uses the address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on the first call there.  Every call on a device shares that
// device's array and overwrites the entries it reads back (one sum and one
// maximum per scene, each written by that scene's workgroup), so nothing
// needs clearing; calls that overlap on one device race.

constexpr int kNoAllocMaxScenes = 4096; // 32 KiB of partial results per device
static __device__ __align__(256) float scratch[2 * kNoAllocMaxScenes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices]; // per device: the device address of `scratch`

int scratch_for_current_device(float **out)
{
    int dev = 0;
    MCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return merr(SPH_ENODEV, "sph_scene_density_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return merr(SPH_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return SPH_OK;
}

} // namespace

extern "C" int sph_scene_density_noalloc(sph_ctx *c, float *mean, float *maxv)
{
    if (!c || !mean || !maxv)
        return merr(SPH_EINVAL, "null argument");
    const size_t ns = c->sc.size();
    if (ns > (size_t)kNoAllocMaxScenes)
        return merr(SPH_EINVAL,
                    "sph_scene_density_noalloc: %zu scenes exceed the fixed scratch of %d scenes "
                    "(use sph_scene_density)",
                    ns, kNoAllocMaxScenes);
    int dev = 0;
    MCHK(hipGetDevice(&dev));
    if (dev != c->device)
        return merr(SPH_EINVAL, "sph_scene_density_noalloc: the context lives on device %d, "
                                "device %d is current", c->device, dev);
    float *buf = nullptr;
    if (int rc = scratch_for_current_device(&buf))
        return rc;
    if (int rc = sph_internal_canonicalize(c))
        return rc;
    k_scene_density<<<dim3((unsigned)ns), 256, 0, c->st>>>(c->dP, c->RHO, buf, buf + ns);
    MCHK(hipGetLastError());
    float h[2 * kNoAllocMaxScenes];
    MCHK(hipMemcpyAsync(h, buf, sizeof(float) * 2 * ns, hipMemcpyDeviceToHost, c->st));
    MCHK(hipStreamSynchronize(c->st));
    for (size_t s = 0; s < ns; s++) {
        mean[s] = c->sc[s].n ? h[s] / (float)c->sc[s].n : 0.0f;
        maxv[s] = h[ns + s];
    }
    return SPH_OK;
}
