// sph_ctx_internal.h - the device-resident context record of libsph_hip.so,
// shared by the solver (sph_ctx.hip) and the surrogate (sph_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "sph_kernels.h"

namespace sph_internal {

struct Scene {
    tunable_parameters tun{};
    AABB_t box{};
    int cap = 0, n = 0;
    unsigned sx = 0, sy = 0, ncells = 0;
    unsigned cell_base = 0, poff = 0, slots = 0, uoff = 0;
    float slosh_amp = 0.0f, slosh_omega = 0.0f;
    unsigned slosh_t0 = 0;
    int frozen = 0;
};

} // namespace sph_internal

struct sph_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::vector<sph_internal::Scene> sc;
    unsigned slots = 0;       // padded slots of all scenes
    unsigned cells = 0;       // global cells incl. one padding cell per scene
    int n_total = 0;
    std::vector<sph_dev_params> hp;
    sph_dev_params *dP = nullptr;
    // parameters of the fused step's first kernel (gravity, viscosity,
    // predict); equal to dP except after sph_step_tunables2, which gives the
    // two halves of a step different tunables (dPA_stale: dPA still holds
    // the first half's and must be refreshed before the next fused step)
    sph_dev_params *dPA = nullptr;
    bool dPA_stale = false;
    unsigned bucket_cap = 0; // the reference's bucket cap (0: none)
    int lanes = 1;           // lanes per particle of the gathers (sph_gather_lanes)
    float2 *XY = nullptr; // positions by id for sph_download_positions (allocated on first use)
    int *BS = nullptr; // workgroup -> scene
    sph_dev_counters *dC = nullptr;
    unsigned long long *dclock = nullptr; // fused steps since upload (sloshing time)
    // state (canonical: sorted by (cell, id); ID0/K0 give id and key);
    // pairs interleaved: POS=(x,y) VEL=(v_x,v_y) PREV=(x_prev,y_prev)
    // RHO=(density,density_near)
    float2 *POS0 = nullptr, *VEL = nullptr, *PREV = nullptr, *RHO = nullptr;
    // work buffers
    float2 *POS1 = nullptr, *UPOS = nullptr, *UVEL = nullptr, *T2a = nullptr, *T2b = nullptr;
    int *ID0 = nullptr, *ID1 = nullptr, *perm_src = nullptr, *perm_id = nullptr,
        *pos_of_id = nullptr;
    unsigned *K0 = nullptr, *K1 = nullptr, *ukey = nullptr, *uslot = nullptr, *cnt = nullptr,
             *start = nullptr, *bsum = nullptr;
    short *coords = nullptr;
    uint4 *MASK = nullptr; // per slot: the density pass's first accepted-candidate masks
    float2 *ACC = nullptr;  // (a_x, a_y) by caller index: carried, never changed (src/fluid.c)
    float *AOS = nullptr;   // fluid_particle records by caller index (allocated on first download)
    bool lists_valid = false; // a hash ran since upload: K0/start describe POS0
    int aux_order = 0;        // PREV, RHO ordered like ID0 (0) or ID1 (1)
    // snapshot of the canonical state (sph_snapshot_save / sph_snapshot_restore)
    float2 *SPOS = nullptr, *SVEL = nullptr, *SPREV = nullptr, *SRHO = nullptr;
    int *SID = nullptr;
    unsigned *SK = nullptr, *Sstart = nullptr;
    unsigned long long *Sclock = nullptr;
    int *dflags = nullptr; // per-scene flags for the restore kernel
    bool snap_valid = false, snap_lists = false;
    hipGraphExec_t g1 = nullptr, gK = nullptr;
    uint64_t steps = 0, hashes = 0;
    // per-kernel timing (only while sph_time_kernels enqueues)
    std::vector<hipEvent_t> *tev = nullptr;
    std::vector<int> *tkind = nullptr;
    size_t tidx = 0;
};

// Bring PREV/RHO into the canonical (ID0) order; defined in sph_ctx.hip.
int sph_internal_canonicalize(sph_ctx *c);
