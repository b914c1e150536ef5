// sph_kernels.h - launchers of the gfx950 kernels (sph_kernels.hip).
#pragma once

#include "sph_device.h"

// Source/destination pairs of interleaved (float2) fields for one reorder
// or permutation pass.
template <int NF>
struct sph_reorder_fields {
    const float2 *src[NF];
    float2 *dst[NF];
};

// Workgroup size of every particle kernel; scenes are padded to multiples.
constexpr unsigned kSphTPB = 256;

// Lanes per particle (1 or 4) of the gathers of a context with `slots`
// padded slots, on the current device.
int sph_gather_lanes(unsigned slots);

// cap: the reference's bucket cap is on (sph_dev_params::bucket_cap != 0).
// lanes: lanes per particle of the gathers, 1 or 4 (sph_gather_lanes); 4
// applies to the fused step's variants (grav, visc and pred; fused relax).
void sph_launch_visc_predict(hipStream_t st, unsigned slots, bool grav, bool visc, bool pred,
                             bool cap, const sph_dev_params *P, const int *BS, const float2 *POS,
                             const float2 *VEL, const unsigned *K, const unsigned *start,
                             float2 *opos, unsigned *okey, unsigned *oslot, unsigned *cnt,
                             sph_dev_counters *C, const unsigned long long *clock,
                             float2 *held_vel, int lanes);
void sph_launch_gravity(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        float2 *VEL, float2 *RHO);
void sph_launch_predict(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        float2 *POS, const float2 *VEL, float2 *PREV);
void sph_launch_update_velocities(hipStream_t st, unsigned slots, const sph_dev_params *P,
                                  const int *BS, float2 *POS, float2 *VEL, const float2 *PREV);
void sph_launch_key_hist(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                         const float2 *POS, unsigned *okey, unsigned *oslot, unsigned *cnt,
                         sph_dev_counters *C);
// The scan over m counts reads and writes whole tiles: cnt and start must hold
// sph_scan_padded(m) entries (16-byte aligned), cnt's beyond m zero.
unsigned sph_scan_blocks(unsigned m);
size_t sph_scan_padded(unsigned m);
void sph_launch_scan_reduce(hipStream_t st, const unsigned *cnt, unsigned m, unsigned *bsum);
// Small count arrays (sph_scan_single_fits(m)): the same scan in one launch.
bool sph_scan_single_fits(unsigned m);
void sph_launch_scan_single(hipStream_t st, unsigned *cnt, unsigned m, unsigned *start);
void sph_launch_scan_apply(hipStream_t st, unsigned *cnt, unsigned m, const unsigned *bsum,
                           unsigned *start);
void sph_launch_scatter(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        const unsigned *key, const unsigned *slot, const unsigned *start,
                        const int *id_src, int *perm_src, int *perm_id, unsigned *key_dst);
// skey: the sorted keys the scatter wrote (its key_dst)
void sph_launch_reorder2(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                         const unsigned *start, const int *perm_src, const int *perm_id,
                         const unsigned *skey, sph_reorder_fields<2> f, int *id_dst,
                         sph_dev_counters *C);
void sph_launch_reorder4(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                         const unsigned *start, const int *perm_src, const int *perm_id,
                         const unsigned *skey, sph_reorder_fields<4> f, int *id_dst,
                         sph_dev_counters *C);
void sph_launch_density(hipStream_t st, unsigned slots, bool cap, const sph_dev_params *P,
                        const int *BS, const float2 *POS, const unsigned *K,
                        const unsigned *start, float2 *RHO, uint4 *masks, sph_dev_counters *C,
                        int lanes);
void sph_launch_relax(hipStream_t st, unsigned slots, bool fused, bool cap,
                      const sph_dev_params *P, const int *BS, const float2 *POS,
                      const float2 *PREV, const float2 *RHO, const unsigned *K,
                      const unsigned *start, float2 *opos, float2 *ovel, unsigned *okey,
                      unsigned *oslot, unsigned *cnt, const uint4 *masks, sph_dev_counters *C,
                      unsigned long long *clock, int lanes);
void sph_launch_pack_coords(hipStream_t st, unsigned slots, const sph_dev_params *P,
                            const int *BS, const float2 *POS, const int *ID, short *out);
void sph_launch_neighbor_lists(hipStream_t st, unsigned slots, const sph_dev_params *P,
                               const int *BS, const float2 *POS, const unsigned *K, const int *ID,
                               const unsigned *start, int stride, int *full_by_id,
                               int *half_by_id, int *ids_by_id);
void sph_launch_pos_of_id(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                          const int *ID, int *pos_of_id);
void sph_launch_permute_by_id(hipStream_t st, unsigned slots, const sph_dev_params *P,
                              const int *BS, const int *id_src, const int *pos_of_id,
                              sph_reorder_fields<2> f);

// Per-scene copy of a state snapshot back into the live arrays: the slot
// range of every scene whose flag is set, and its cell starts (incl. the
// scene's padding cell).  F[0..3] = POS, VEL, PREV, RHO.
struct sph_restore_args {
    const float2 *src[4];
    float2 *dst[4];
    const int *id_src;
    int *id_dst;
    const unsigned *key_src, *start_src;
    unsigned *key_dst, *start_dst;
};
void sph_launch_restore(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                        const int *flags, sph_restore_args a);

// Caller-order state exchange with device buffers of float4 (x, y, v_x, v_y)
// per particle at uoff + id.  Load sets ids to the slot order (id = slot -
// poff), PREV = POS and RHO = 0.
void sph_launch_gather_state(hipStream_t st, unsigned slots, const sph_dev_params *P,
                             const int *BS, const float2 *POS, const float2 *VEL, const int *ID,
                             float4 *out);
void sph_launch_load_state(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                           const float4 *in, float2 *POS, float2 *VEL, float2 *PREV, float2 *RHO,
                           int *ID);
// fluid_particle records (13 words) by caller index uoff + id; IDA orders
// PREV and RHO, ACC holds (a_x, a_y) by caller index.
void sph_launch_gather_aos(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                           const float2 *POS, const float2 *VEL, const int *ID0,
                           const float2 *PREV, const float2 *RHO, const int *IDA,
                           const float2 *ACC, float *out);
// Positions (x, y) by caller index uoff + id.
void sph_launch_gather_xy(hipStream_t st, unsigned slots, const sph_dev_params *P, const int *BS,
                          const float2 *POS, const int *ID, float2 *out);

// Time the next kernel launch of this thread with a start/stop event pair
// (recorded by the launch itself: the kernel's own duration).
void sph_set_next_launch_events(hipEvent_t start, hipEvent_t stop);

// Exhaustive bit-for-bit check of sph_sqrt against sqrtf over float bit
// patterns [lo, hi).
void sph_launch_selftest_math(hipStream_t st, int fn, unsigned lo, unsigned hi,
                              unsigned long long *bad);
