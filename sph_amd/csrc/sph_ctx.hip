// sph_ctx.hip - host side of libsph_hip.so: the device-resident context
// behind the C ABI of include/sph_abi.h.
//
// One context = one or more independent scenes on one GPU, stepped
// together.  All scenes share one SoA particle array sorted by
// (global cell, id): scene s owns slots [poff_s, poff_s + n_s) (its range
// padded to whole 256-thread workgroups) and global cells
// [cell_base_s, cell_base_s + ncells_s) plus one padding cell.  sph_step
// replays a captured HIP graph of the fused step; the phase calls run the
// reference's functions one by one.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "sph_ctx_internal.h"

namespace {

thread_local char g_err[512] = "";

int set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_err(SPH_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                           __FILE__, __LINE__);                                                  \
    } while (0)

// Kernel kinds of the fused step, for per-kernel timing.
enum Kind { KV = 0, KSR, KSA, KSC, KRO, KDE, KRE, NKIND };
const char *kKindNames[NKIND] = {"visc_predict", "scan_reduce", "scan_apply", "scatter",
                                 "reorder",      "density",     "relax"};

constexpr int kGraphSteps = 16; // steps per replay of the long graph

} // namespace

using sph_internal::Scene;

namespace {

void derive_params(sph_ctx *c)
{
    c->hp.resize(c->sc.size());
    for (size_t s = 0; s < c->sc.size(); s++) {
        const Scene &S = c->sc[s];
        const tunable_parameters &t = S.tun;
        sph_dev_params &p = c->hp[s];
        p.h = t.smoothing_radius;
        p.h_recip = 1.0f / t.smoothing_radius;
        p.h2 = t.smoothing_radius * t.smoothing_radius;
        p.dt = t.time_step;
        p.g_neg = -t.g;
        p.rest_density = t.rest_density;
        p.k = t.k;
        p.k_near = t.k_near;
        p.k_spring = t.k_spring;
        p.sigma = t.sigma;
        p.beta = t.beta;
        p.min_x = S.box.min_x;
        p.max_x = S.box.max_x;
        p.min_y = S.box.min_y;
        p.max_y = S.box.max_y;
        p.mover_cx = t.mover_center_x;
        p.mover_cy = t.mover_center_y;
        p.mover_w = t.mover_width;
        p.mover_h = t.mover_height;
        p.mover_type = t.mover_type;
        p.sx = S.sx;
        p.sy = S.sy;
        p.ncells = S.ncells;
        p.cell_base = S.cell_base;
        p.poff = S.poff;
        p.uoff = S.uoff;
        p.n = S.n;
        p.pad = S.slots - (unsigned)S.n;
        p.slosh_amp = S.slosh_amp;
        p.slosh_omega = S.slosh_omega;
        p.slosh_t0 = S.slosh_t0;
        p.frozen = S.frozen;
        p.bucket_cap = c->bucket_cap;
        int e = 0;
        p.h_pow2 = std::frexp(p.h, &e) == 0.5f && std::fpclassify(p.h_recip) == FP_NORMAL;
    }
}

// Derive and copy the tunables of every scene to both device parameter
// arrays.  The caller has synchronised the context stream (the copy runs on
// the null stream, which does not wait for the non-blocking context stream).
int push_params(sph_ctx *c)
{
    derive_params(c);
    const size_t b = sizeof(sph_dev_params) * c->hp.size();
    HIPCHK(hipMemcpy(c->dP, c->hp.data(), b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dPA, c->hp.data(), b, hipMemcpyHostToDevice));
    c->dPA_stale = false;
    return SPH_OK;
}

// After a step with split tunables the first-half array still holds the
// first half's values: bring it back to the current ones before the next
// fused step.
int refresh_first_half_params(sph_ctx *c)
{
    if (!c->dPA_stale)
        return SPH_OK;
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(c->dPA, c->hp.data(), sizeof(sph_dev_params) * c->hp.size(),
                     hipMemcpyHostToDevice));
    c->dPA_stale = false;
    return SPH_OK;
}

void free_graphs(sph_ctx *c)
{
    if (c->g1)
        (void)hipGraphExecDestroy(c->g1);
    if (c->gK)
        (void)hipGraphExecDestroy(c->gK);
    c->g1 = c->gK = nullptr;
}

// Timing hook before each launch (a no-op unless sph_time_kernels is
// enqueueing): hands the next start/stop event pair to the launch, which
// records the kernel's own start and end.
struct Tick {
    Tick(sph_ctx *c, int kind)
    {
        if (c->tev) {
            sph_set_next_launch_events((*c->tev)[c->tidx], (*c->tev)[c->tidx + 1]);
            c->tidx += 2;
            c->tkind->push_back(kind);
        }
    }
};

void enqueue_sort_tail(sph_ctx *c, const int *id_src, unsigned *key_dst)
{
    const unsigned m = c->cells + 1;
    if (sph_scan_single_fits(m)) { // a small grid: the whole scan in one launch
        Tick t(c, KSA);
        sph_launch_scan_single(c->st, c->cnt, m, c->start);
    } else {
        {
            Tick t(c, KSR);
            sph_launch_scan_reduce(c->st, c->cnt, m, c->bsum);
        }
        {
            Tick t(c, KSA);
            sph_launch_scan_apply(c->st, c->cnt, m, c->bsum, c->start);
        }
    }
    {
        Tick t(c, KSC);
        sph_launch_scatter(c->st, c->slots, c->dP, c->BS, c->ukey, c->uslot, c->start, id_src,
                           c->perm_src, c->perm_id, key_dst);
    }
}

// The fused step: 11 launches, no host synchronisation.
//   visc_predict: gravity + viscosity + predict + boundary + key/slot
//   scan, scan, scatter, reorder   -> sorted predicted state (POS1, PREV)
//   density                        -> RHO
//   relax: relaxation + boundary + velocity update + key/slot
//   scan, scan, scatter, reorder   -> sorted relaxed state (POS0, VEL)
void enqueue_fused_step(sph_ctx *c, bool visc)
{
    const unsigned S = c->slots;
    const bool cap = c->bucket_cap != 0;
    {
        Tick t(c, KV);
        sph_launch_visc_predict(c->st, S, true, visc, true, cap, c->dPA, c->BS, c->POS0, c->VEL,
                                c->K0, c->start, c->UPOS, c->ukey, c->uslot, c->cnt, c->dC,
                                c->dclock, c->UVEL, c->lanes);
    }
    enqueue_sort_tail(c, c->ID0, c->K1);
    {
        Tick t(c, KRO);
        sph_reorder_fields<2> f{{c->UPOS, c->POS0}, {c->POS1, c->PREV}};
        sph_launch_reorder2(c->st, S, c->dP, c->BS, c->start, c->perm_src, c->perm_id, c->K1,
                            f, c->ID1, c->dC);
    }
    {
        Tick t(c, KDE);
        sph_launch_density(c->st, S, cap, c->dP, c->BS, c->POS1, c->K1, c->start, c->RHO,
                           c->MASK, c->dC, c->lanes);
    }
    {
        Tick t(c, KRE);
        sph_launch_relax(c->st, S, true, cap, c->dP, c->BS, c->POS1, c->PREV, c->RHO, c->K1,
                         c->start, c->UPOS, c->UVEL, c->ukey, c->uslot, c->cnt, c->MASK, c->dC,
                         c->dclock, c->lanes);
    }
    enqueue_sort_tail(c, c->ID1, c->K0);
    {
        Tick t(c, KRO);
        sph_reorder_fields<2> f{{c->UPOS, c->UVEL}, {c->POS0, c->VEL}};
        sph_launch_reorder2(c->st, S, c->dP, c->BS, c->start, c->perm_src, c->perm_id, c->K0,
                            f, c->ID0, c->dC);
    }
}

int build_graph(sph_ctx *c, int steps, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; i++)
        enqueue_fused_step(c, true);
    HIPCHK(hipStreamEndCapture(c->st, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return set_err(SPH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return SPH_OK;
}

int ensure_graphs(sph_ctx *c)
{
    if (c->g1 && c->gK)
        return SPH_OK;
    free_graphs(c);
    int rc = build_graph(c, 1, &c->g1);
    if (rc)
        return rc;
    return build_graph(c, kGraphSteps, &c->gK);
}

// Bring PREV/RHO into ID0 order (after fused steps they follow ID1).
int canonicalize(sph_ctx *c)
{
    if (c->aux_order == 0 || c->n_total == 0) {
        c->aux_order = 0;
        return SPH_OK;
    }
    sph_launch_pos_of_id(c->st, c->slots, c->dP, c->BS, c->ID0, c->pos_of_id);
    sph_reorder_fields<2> f{{c->PREV, c->RHO}, {c->T2a, c->T2b}};
    sph_launch_permute_by_id(c->st, c->slots, c->dP, c->BS, c->ID1, c->pos_of_id, f);
    size_t b = sizeof(float2) * (size_t)c->slots;
    HIPCHK(hipMemcpyAsync(c->PREV, c->T2a, b, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->RHO, c->T2b, b, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipGetLastError());
    c->aux_order = 0;
    return SPH_OK;
}

int check_ctx(const sph_ctx *c)
{
    if (!c)
        return set_err(SPH_EINVAL, "null context");
    return SPH_OK;
}

template <class T>
int dalloc(T **p, size_t count)
{
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess)
        return set_err(SPH_ENOMEM, "hipMalloc(%zu x %zu B) failed: %s", count, sizeof(T),
                       hipGetErrorString(e));
    return SPH_OK;
}

int check_scene_desc(const sph_scene_desc &d, int idx)
{
    const float h = d.tunables.smoothing_radius;
    if (!(h > 0.0f && h <= 0x1p60f)) // the kernels' reciprocal needs r <= h <= 2^60
        return set_err(SPH_EINVAL, "scene %d: smoothing_radius must be in (0, 2^60]", idx);
    if (d.boundary.min_x != 0.0f || d.boundary.min_y != 0.0f)
        return set_err(SPH_EINVAL, "scene %d: the box must start at the origin", idx);
    if (!(d.boundary.max_x > 0.0f && d.boundary.max_y > 0.0f))
        return set_err(SPH_EINVAL, "scene %d: empty box", idx);
    if (d.capacity < 0)
        return set_err(SPH_EINVAL, "scene %d: negative capacity", idx);
    return SPH_OK;
}

} // namespace

extern "C" {

const char *sph_last_error(void) { return g_err; }

int sph_create_batch(sph_ctx **out, const sph_scene_desc *scenes, int nscenes)
{
    if (!out || !scenes || nscenes <= 0)
        return set_err(SPH_EINVAL, "sph_create_batch: null argument or no scenes");
    *out = nullptr;
    for (int s = 0; s < nscenes; s++)
        if (int rc = check_scene_desc(scenes[s], s))
            return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(SPH_ENODEV, "no HIP device available");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(SPH_ENODEV, "device %d is %s; this library is built for gfx950", dev,
                       prop.gcnArchName);

    sph_ctx *c = new (std::nothrow) sph_ctx();
    if (!c)
        return set_err(SPH_ENOMEM, "host allocation failed");
    c->device = dev;
    c->sc.resize((size_t)nscenes);
    unsigned long long slots = 0, cells = 0;
    for (int s = 0; s < nscenes; s++) {
        Scene &S = c->sc[s];
        S.tun = scenes[s].tunables;
        S.box = scenes[s].boundary;
        S.cap = scenes[s].capacity;
        const float h = S.tun.smoothing_radius;
        // grid extent, src/fluid.c:214-215
        const double gx = ceil((S.box.max_x - S.box.min_x) / h);
        const double gy = ceil((S.box.max_y - S.box.min_y) / h);
        if (gx * gy >= 4294967295.0) {
            delete c;
            return set_err(SPH_EINVAL, "scene %d: %g x %g cells exceed 32-bit keys", s, gx, gy);
        }
        S.sx = (unsigned)gx;
        S.sy = (unsigned)gy;
        S.ncells = S.sx * S.sy;
        S.cell_base = (unsigned)cells;
        S.poff = (unsigned)slots;
        S.slots = ((unsigned)(S.cap > 0 ? S.cap : 1) + kSphTPB - 1) / kSphTPB * kSphTPB;
        cells += (unsigned long long)S.ncells + 1; // + the scene's padding cell
        slots += S.slots;
    }
    if (cells >= 4294967295ull || slots >= 2147483647ull) {
        delete c;
        return set_err(SPH_EINVAL, "sph_create_batch: %llu cells / %llu slots exceed 32-bit indexing",
                       cells, slots);
    }
    c->cells = (unsigned)cells;
    c->slots = (unsigned)slots;
    c->lanes = sph_gather_lanes(c->slots);
    int rc = SPH_OK;
    const size_t N = c->slots, M = (size_t)c->cells + 1;
    const size_t MP = sph_scan_padded((unsigned)M); // the scan works on whole tiles
#define A(p, count)                                                                              \
    if (!rc)                                                                                     \
    rc = dalloc(&c->p, count)
    // the arrays the gathers walk carry kSphWalkPad spare records
    A(POS0, N + kSphWalkPad); A(VEL, N + kSphWalkPad); A(PREV, N); A(RHO, N + kSphWalkPad);
    A(POS1, N + kSphWalkPad); A(UPOS, N); A(UVEL, N); A(T2a, N); A(T2b, N);
    A(ID0, N); A(ID1, N); A(perm_src, N); A(perm_id, N); A(pos_of_id, N);
    A(K0, N); A(K1, N); A(ukey, N); A(uslot, N);
    A(cnt, MP); A(start, MP); A(bsum, sph_scan_blocks((unsigned)M));
    A(coords, 2 * N); A(MASK, N); A(dP, (size_t)nscenes); A(dPA, (size_t)nscenes); A(ACC, N);
    A(BS, N / kSphTPB); A(dC, 1); A(dclock, 1);
#undef A
    if (!rc) {
        hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
        if (e != hipSuccess)
            rc = set_err(SPH_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    if (!rc) {
        std::vector<int> bs(N / kSphTPB);
        for (int s = 0; s < nscenes; s++)
            for (unsigned b = 0; b < c->sc[s].slots / kSphTPB; b++)
                bs[c->sc[s].poff / kSphTPB + b] = s;
        hipError_t e = hipMemcpy(c->BS, bs.data(), sizeof(int) * bs.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemset(c->cnt, 0, sizeof(unsigned) * MP);
        if (e == hipSuccess)
            e = hipMemset(c->start, 0, sizeof(unsigned) * MP);
        if (e == hipSuccess)
            e = hipMemset(c->dC, 0, sizeof(sph_dev_counters));
        if (e == hipSuccess)
            e = hipMemset(c->dclock, 0, sizeof(unsigned long long));
        if (e == hipSuccess)
            e = hipMemset(c->POS0, 0, sizeof(float2) * N);
        if (e != hipSuccess)
            rc = set_err(SPH_EHIP, "context set-up failed: %s", hipGetErrorString(e));
    }
    if (!rc)
        rc = push_params(c);
    if (rc) {
        sph_destroy(c);
        return rc;
    }
    *out = c;
    return SPH_OK;
}

int sph_create(sph_ctx **out, const param *params, const AABB_t *boundary, int capacity)
{
    if (!out || !params || !boundary || capacity < 0)
        return set_err(SPH_EINVAL, "sph_create: null argument or negative capacity");
    sph_scene_desc d;
    d.tunables = params->tunable_params;
    d.boundary = *boundary;
    d.capacity = capacity;
    return sph_create_batch(out, &d, 1);
}

void sph_destroy(sph_ctx *c)
{
    if (!c)
        return;
    if (c->st)
        (void)hipStreamSynchronize(c->st);
    free_graphs(c);
    void *ptrs[] = {c->POS0, c->VEL,  c->PREV,     c->RHO,     c->POS1,      c->UPOS,
                    c->UVEL, c->T2a,  c->T2b,      c->ID0,     c->ID1,       c->perm_src,
                    c->perm_id, c->pos_of_id, c->K0, c->K1,    c->ukey,      c->uslot,
                    c->cnt,  c->start, c->bsum,    c->coords,  c->dP,        c->BS,
                    c->dC,   c->dclock, c->SPOS,    c->SVEL,    c->SPREV,     c->SRHO,
                    c->SID,  c->SK,    c->Sstart,  c->Sclock,  c->dflags, c->MASK,
                    c->dPA,  c->XY,    c->ACC,     c->AOS};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (c->st)
        (void)hipStreamDestroy(c->st);
    delete c;
}

int sph_num_particles(const sph_ctx *c) { return c ? c->n_total : 0; }

int sph_num_scenes(const sph_ctx *c) { return c ? (int)c->sc.size() : 0; }

int sph_num_cells(const sph_ctx *c, unsigned *size_x, unsigned *size_y)
{
    if (!c)
        return set_err(SPH_EINVAL, "null context");
    if (size_x)
        *size_x = c->sc[0].sx;
    if (size_y)
        *size_y = c->sc[0].sy;
    return SPH_OK;
}

int sph_scene_info(const sph_ctx *c, int scene, int *n, unsigned *size_x, unsigned *size_y)
{
    if (!c || scene < 0 || scene >= (int)c->sc.size())
        return set_err(SPH_EINVAL, "sph_scene_info: bad context or scene");
    if (n)
        *n = c->sc[scene].n;
    if (size_x)
        *size_x = c->sc[scene].sx;
    if (size_y)
        *size_y = c->sc[scene].sy;
    return SPH_OK;
}

// particles: scenes back to back, counts[s] records each, ids local to the
// scene; pp (pointer array) is the one-scene form.
static int upload_impl(sph_ctx *c, const fluid_particle *const *pp, const fluid_particle *arr,
                       const int *counts)
{
    if (int rc = check_ctx(c))
        return rc;
    const int ns = (int)c->sc.size();
    long total = 0;
    for (int s = 0; s < ns; s++) {
        if (counts[s] < 0 || counts[s] > c->sc[s].cap)
            return set_err(SPH_EINVAL, "sph_upload: scene %d count %d outside [0, capacity=%d]", s,
                           counts[s], c->sc[s].cap);
        total += counts[s];
    }
    if (total > 0 && !pp && !arr)
        return set_err(SPH_EINVAL, "sph_upload: null particles");
    const size_t N = c->slots;
    std::vector<float2> f[4];
    for (auto &v : f)
        v.assign(N, make_float2(0.0f, 0.0f));
    std::vector<int> ids(N, 0);
    std::vector<float2> acc((size_t)total, make_float2(0.0f, 0.0f));
    std::vector<unsigned> uoff((size_t)ns);
    long u = 0;
    for (int s = 0; s < ns; s++) {
        const int n = counts[s];
        std::vector<char> seen((size_t)n, 0);
        for (int i = 0; i < n; i++) {
            const fluid_particle *p = pp ? pp[u + i] : &arr[u + i];
            if (!p)
                return set_err(SPH_EINVAL, "sph_upload: null particle pointer at %ld", u + i);
            if (p->id < 0 || p->id >= n || seen[p->id])
                return set_err(SPH_EINVAL,
                               "sph_upload: scene %d ids must be a permutation of 0..%d (bad id %d at %d)",
                               s, n - 1, p->id, i);
            seen[p->id] = 1;
            const size_t slot = c->sc[s].poff + (size_t)i;
            ids[slot] = p->id;
            f[0][slot] = make_float2(p->x, p->y);
            f[1][slot] = make_float2(p->v_x, p->v_y);
            f[2][slot] = make_float2(p->x_prev, p->y_prev);
            f[3][slot] = make_float2(p->density, p->density_near);
            acc[u + p->id] = make_float2(p->a_x, p->a_y);
        }
        uoff[s] = (unsigned)u;
        u += n;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    float2 *dst[4] = {c->POS0, c->VEL, c->PREV, c->RHO};
    for (int k = 0; k < 4; k++)
        HIPCHK(hipMemcpy(dst[k], f[k].data(), sizeof(float2) * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->ID0, ids.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    if (total > 0)
        HIPCHK(hipMemcpy(c->ACC, acc.data(), sizeof(float2) * (size_t)total, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->dclock, 0, sizeof(unsigned long long)));
    for (int s = 0; s < ns; s++) {
        c->sc[s].n = counts[s];
        c->sc[s].uoff = uoff[s];
    }
    c->n_total = (int)total;
    c->lists_valid = false;
    c->aux_order = 0;
    c->snap_valid = false;
    return push_params(c);
}

int sph_upload(sph_ctx *c, const fluid_particle *particles, int n)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->sc.size() != 1)
        return set_err(SPH_EINVAL, "sph_upload: context holds %zu scenes; use sph_upload_batch",
                       c->sc.size());
    return upload_impl(c, nullptr, particles, &n);
}

int sph_upload_pointers(sph_ctx *c, fluid_particle *const *fluid_particle_pointers, int n)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->sc.size() != 1)
        return set_err(SPH_EINVAL, "sph_upload_pointers: one-scene contexts only");
    return upload_impl(c, (const fluid_particle *const *)fluid_particle_pointers, nullptr, &n);
}

int sph_upload_batch(sph_ctx *c, const fluid_particle *particles, const int *counts)
{
    if (int rc = check_ctx(c))
        return rc;
    if (!counts)
        return set_err(SPH_EINVAL, "sph_upload_batch: null counts");
    return upload_impl(c, nullptr, particles, counts);
}

static int set_tunables_impl(sph_ctx *c, int scene, const tunable_parameters *t)
{
    if (!t)
        return set_err(SPH_EINVAL, "sph_set_tunables: null");
    bool changed = false;
    for (int s = 0; s < (int)c->sc.size(); s++) {
        if (scene >= 0 && s != scene)
            continue;
        if (t->smoothing_radius != c->sc[s].tun.smoothing_radius)
            return set_err(SPH_EINVAL,
                           "sph_set_tunables: smoothing_radius fixes the grid and cannot change");
    }
    for (int s = 0; s < (int)c->sc.size(); s++) {
        if (scene >= 0 && s != scene)
            continue;
        if (memcmp(t, &c->sc[s].tun, sizeof *t) != 0) {
            c->sc[s].tun = *t;
            changed = true;
        }
    }
    if (!changed)
        return SPH_OK;
    HIPCHK(hipStreamSynchronize(c->st));
    return push_params(c);
}

int sph_set_tunables(sph_ctx *c, const tunable_parameters *t)
{
    if (int rc = check_ctx(c))
        return rc;
    return set_tunables_impl(c, -1, t);
}

int sph_set_scene_tunables(sph_ctx *c, int scene, const tunable_parameters *t)
{
    if (int rc = check_ctx(c))
        return rc;
    if (scene < 0 || scene >= (int)c->sc.size())
        return set_err(SPH_EINVAL, "sph_set_scene_tunables: scene %d out of range", scene);
    return set_tunables_impl(c, scene, t);
}

int sph_set_scene_forcing(sph_ctx *c, int scene, float amplitude, float period)
{
    if (int rc = check_ctx(c))
        return rc;
    if (scene < 0 || scene >= (int)c->sc.size() || !(period > 0.0f) || !(amplitude >= 0.0f))
        return set_err(SPH_EINVAL, "sph_set_scene_forcing: bad scene, amplitude or period");
    c->sc[scene].slosh_amp = amplitude;
    c->sc[scene].slosh_omega = (float)(2.0 * 3.14159265358979323846 / period);
    HIPCHK(hipStreamSynchronize(c->st));
    return push_params(c);
}

int sph_set_scene_time_offset(sph_ctx *c, int scene, unsigned steps)
{
    if (int rc = check_ctx(c))
        return rc;
    if (scene < 0 || scene >= (int)c->sc.size())
        return set_err(SPH_EINVAL, "sph_set_scene_time_offset: scene %d out of range", scene);
    c->sc[scene].slosh_t0 = steps;
    HIPCHK(hipStreamSynchronize(c->st));
    return push_params(c);
}

int sph_set_scene_frozen(sph_ctx *c, const int *frozen)
{
    if (int rc = check_ctx(c))
        return rc;
    if (!frozen)
        return set_err(SPH_EINVAL, "sph_set_scene_frozen: null flags");
    for (size_t s = 0; s < c->sc.size(); s++)
        c->sc[s].frozen = frozen[s] ? 1 : 0;
    HIPCHK(hipStreamSynchronize(c->st));
    return push_params(c);
}

int sph_set_bucket_cap(sph_ctx *c, int cap)
{
    if (int rc = check_ctx(c))
        return rc;
    if (cap < 0)
        return set_err(SPH_EINVAL, "sph_set_bucket_cap: cap %d < 0", cap);
    HIPCHK(hipStreamSynchronize(c->st));
    if ((c->bucket_cap != 0) != (cap != 0))
        free_graphs(c); // the captured steps launch the other kernel variants
    c->bucket_cap = (unsigned)cap;
    return push_params(c);
}

int sph_set_gather_lanes(sph_ctx *c, int lanes)
{
    if (int rc = check_ctx(c))
        return rc;
    if (lanes != 0 && lanes != 1 && lanes != 4)
        return set_err(SPH_EINVAL, "sph_set_gather_lanes: lanes %d is not 0, 1 or 4", lanes);
    const int want = lanes ? lanes : sph_gather_lanes(c->slots);
    HIPCHK(hipStreamSynchronize(c->st));
    if (want != c->lanes)
        free_graphs(c); // the captured steps launch the other kernel variants
    c->lanes = want;
    return SPH_OK;
}

int sph_step_eager(sph_ctx *c, int nsteps)
{
    if (int rc = check_ctx(c))
        return rc;
    if (nsteps < 0)
        return set_err(SPH_EINVAL, "nsteps < 0");
    if (c->n_total == 0)
        return SPH_OK;
    if (int rc = refresh_first_half_params(c))
        return rc;
    for (int i = 0; i < nsteps; i++) {
        enqueue_fused_step(c, c->lists_valid);
        c->lists_valid = true;
        c->aux_order = 1;
        c->steps++;
        c->hashes += 2;
    }
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

int sph_step(sph_ctx *c, int nsteps)
{
    if (int rc = check_ctx(c))
        return rc;
    if (nsteps < 0)
        return set_err(SPH_EINVAL, "nsteps < 0");
    if (c->n_total == 0 || nsteps == 0)
        return SPH_OK;
    if (!c->lists_valid) { // first step after upload: empty lists, no viscosity
        int rc = sph_step_eager(c, 1);
        if (rc)
            return rc;
        nsteps--;
    }
    if (nsteps == 0)
        return SPH_OK;
    if (int rc = refresh_first_half_params(c))
        return rc;
    if (int rc = ensure_graphs(c))
        return rc;
    while (nsteps >= kGraphSteps) {
        HIPCHK(hipGraphLaunch(c->gK, c->st));
        nsteps -= kGraphSteps;
        c->steps += kGraphSteps;
        c->hashes += 2 * kGraphSteps;
    }
    while (nsteps-- > 0) {
        HIPCHK(hipGraphLaunch(c->g1, c->st));
        c->steps++;
        c->hashes += 2;
    }
    c->aux_order = 1;
    return SPH_OK;
}

int sph_sync(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return SPH_OK;
}

// One fused step whose first kernel (gravity, viscosity, predict) reads
// `first` and the rest (hash, density, relaxation, velocity update) reads
// `second`: the first-half array is loaded with `first`, the second with
// `second`, one step is enqueued, and the context keeps `second` (the
// first-half array is refreshed before the next fused step).
int sph_step_tunables2(sph_ctx *c, const tunable_parameters *first,
                       const tunable_parameters *second)
{
    if (int rc = check_ctx(c))
        return rc;
    if (!first || !second)
        return set_err(SPH_EINVAL, "sph_step_tunables2: null tunables");
    for (const Scene &S : c->sc)
        if (first->smoothing_radius != S.tun.smoothing_radius ||
            second->smoothing_radius != S.tun.smoothing_radius)
            return set_err(SPH_EINVAL,
                           "sph_step_tunables2: smoothing_radius fixes the grid and cannot change");
    if (memcmp(first, second, sizeof *first) == 0) {
        if (int rc = set_tunables_impl(c, -1, second))
            return rc;
        return sph_step(c, 1);
    }
    if (c->n_total == 0) {
        return set_tunables_impl(c, -1, second);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    for (Scene &S : c->sc)
        S.tun = *first;
    derive_params(c);
    const size_t b = sizeof(sph_dev_params) * c->hp.size();
    HIPCHK(hipMemcpy(c->dPA, c->hp.data(), b, hipMemcpyHostToDevice));
    for (Scene &S : c->sc)
        S.tun = *second;
    derive_params(c);
    HIPCHK(hipMemcpy(c->dP, c->hp.data(), b, hipMemcpyHostToDevice));
    c->dPA_stale = false; // dPA holds `first` on purpose for this one step
    int rc;
    if (!c->lists_valid) {
        rc = sph_step_eager(c, 1);
    } else {
        rc = ensure_graphs(c);
        if (!rc) {
            hipError_t e = hipGraphLaunch(c->g1, c->st);
            if (e != hipSuccess)
                rc = set_err(SPH_EHIP, "hipGraphLaunch failed: %s", hipGetErrorString(e));
            c->steps++;
            c->hashes += 2;
            c->aux_order = 1;
        }
    }
    c->dPA_stale = true; // hp now holds `second`; dPA must follow before the next step
    return rc;
}

// ---- phase-level calls (reference functions one by one) -------------------

int sph_phase_gravity(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    if (c->n_total)
        sph_launch_gravity(c->st, c->slots, c->dP, c->BS, c->VEL, c->RHO);
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

int sph_phase_viscosity(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    if (!c->lists_valid || c->n_total == 0)
        return SPH_OK; // empty lists (src/fluid.c:202)
    sph_launch_visc_predict(c->st, c->slots, false, true, false, c->bucket_cap != 0, c->dP, c->BS,
                            c->POS0, c->VEL, c->K0, c->start, c->UVEL, nullptr, nullptr, c->cnt,
                            c->dC, c->dclock, nullptr, c->lanes);
    HIPCHK(hipMemcpyAsync(c->VEL, c->UVEL, sizeof(float2) * (size_t)c->slots,
                          hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

int sph_phase_predict(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    if (c->n_total)
        sph_launch_predict(c->st, c->slots, c->dP, c->BS, c->POS0, c->VEL, c->PREV);
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

int sph_phase_hash(sph_ctx *c, int compute_density)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    if (c->n_total == 0)
        return SPH_OK;
    const unsigned S = c->slots;
    sph_launch_key_hist(c->st, S, c->dP, c->BS, c->POS0, c->ukey, c->uslot, c->cnt, c->dC);
    enqueue_sort_tail(c, c->ID0, c->K1);
    sph_reorder_fields<4> f{{c->POS0, c->VEL, c->PREV, c->RHO},
                            {c->POS1, c->UVEL, c->T2a, c->T2b}};
    sph_launch_reorder4(c->st, S, c->dP, c->BS, c->start, c->perm_src, c->perm_id, c->K1, f,
                        c->ID1, c->dC);
    const size_t b = sizeof(float2) * (size_t)S;
    for (int k = 0; k < 4; k++)
        HIPCHK(hipMemcpyAsync((void *)f.src[k], f.dst[k], b, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->ID0, c->ID1, sizeof(int) * (size_t)S, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->K0, c->K1, sizeof(unsigned) * (size_t)S, hipMemcpyDeviceToDevice,
                          c->st));
    if (compute_density)
        sph_launch_density(c->st, S, c->bucket_cap != 0, c->dP, c->BS, c->POS0, c->K0, c->start,
                           c->RHO, c->MASK, c->dC, c->lanes);
    HIPCHK(hipGetLastError());
    c->lists_valid = true;
    c->hashes++;
    return SPH_OK;
}

int sph_phase_relax(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    if (c->n_total == 0)
        return SPH_OK;
    if (!c->lists_valid)
        return set_err(SPH_ESTATE, "sph_phase_relax: no hash since upload");
    sph_launch_relax(c->st, c->slots, false, c->bucket_cap != 0, c->dP, c->BS, c->POS0, c->PREV,
                     c->RHO, c->K0, c->start, c->UPOS, nullptr, nullptr, nullptr, nullptr,
                     nullptr, c->dC, c->dclock, c->lanes);
    HIPCHK(hipMemcpyAsync(c->POS0, c->UPOS, sizeof(float2) * (size_t)c->slots,
                          hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

int sph_phase_update_velocities(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    if (c->n_total)
        sph_launch_update_velocities(c->st, c->slots, c->dP, c->BS, c->POS0, c->VEL, c->PREV);
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

// ---- state snapshots (the surrogate's fallback) ---------------------------

int sph_snapshot_save(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    if (int rc = canonicalize(c))
        return rc;
    const size_t N = c->slots, M = (size_t)c->cells + 1;
    if (!c->SPOS) {
        int rc = SPH_OK;
#define A(p, count)                                                                              \
    if (!rc)                                                                                     \
    rc = dalloc(&c->p, count)
        A(SPOS, N); A(SVEL, N); A(SPREV, N); A(SRHO, N); A(SID, N); A(SK, N); A(Sstart, M);
        A(Sclock, 1); A(dflags, c->sc.size());
#undef A
        if (rc)
            return rc;
    }
    const size_t b2 = sizeof(float2) * N, b1 = sizeof(unsigned) * N;
    HIPCHK(hipMemcpyAsync(c->SPOS, c->POS0, b2, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->SVEL, c->VEL, b2, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->SPREV, c->PREV, b2, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->SRHO, c->RHO, b2, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->SID, c->ID0, b1, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->SK, c->K0, b1, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->Sstart, c->start, sizeof(unsigned) * M, hipMemcpyDeviceToDevice,
                          c->st));
    HIPCHK(hipMemcpyAsync(c->Sclock, c->dclock, sizeof(unsigned long long),
                          hipMemcpyDeviceToDevice, c->st));
    c->snap_valid = true;
    c->snap_lists = c->lists_valid;
    return SPH_OK;
}

int sph_snapshot_restore(sph_ctx *c, const int *scene_flags)
{
    if (int rc = check_ctx(c))
        return rc;
    if (!scene_flags)
        return set_err(SPH_EINVAL, "sph_snapshot_restore: null flags");
    if (!c->snap_valid)
        return set_err(SPH_ESTATE, "sph_snapshot_restore: no snapshot");
    if (int rc = canonicalize(c))
        return rc;
    const size_t ns = c->sc.size();
    bool any = false;
    for (size_t s = 0; s < ns; s++)
        any |= scene_flags[s] != 0;
    if (!any)
        return SPH_OK;
    if (!c->snap_lists || !c->lists_valid)
        return set_err(SPH_ESTATE, "sph_snapshot_restore: snapshot or state not hashed");
    HIPCHK(hipMemcpyAsync(c->dflags, scene_flags, sizeof(int) * ns, hipMemcpyHostToDevice, c->st));
    sph_restore_args a{{c->SPOS, c->SVEL, c->SPREV, c->SRHO},
                       {c->POS0, c->VEL, c->PREV, c->RHO},
                       c->SID, c->ID0, c->SK, c->Sstart, c->K0, c->start};
    sph_launch_restore(c->st, c->slots, c->dP, c->BS, c->dflags, a);
    HIPCHK(hipMemcpyAsync(c->dclock, c->Sclock, sizeof(unsigned long long),
                          hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipGetLastError());
    // the host copy of the flags must outlive the asynchronous upload
    HIPCHK(hipStreamSynchronize(c->st));
    return SPH_OK;
}

// ---- device-resident state exchange ----------------------------------------

int sph_gather_state_dev(sph_ctx *c, float *xv_by_id)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->n_total > 0 && !xv_by_id)
        return set_err(SPH_EINVAL, "sph_gather_state_dev: null output");
    if (c->n_total)
        sph_launch_gather_state(c->st, c->slots, c->dP, c->BS, c->POS0, c->VEL, c->ID0,
                                (float4 *)xv_by_id);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return SPH_OK;
}

int sph_load_state_dev(sph_ctx *c, const float *xv_by_id)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->n_total > 0 && !xv_by_id)
        return set_err(SPH_EINVAL, "sph_load_state_dev: null input");
    if (((uintptr_t)xv_by_id & 15u) != 0)
        return set_err(SPH_EINVAL, "sph_load_state_dev: input not 16-byte aligned");
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->n_total)
        sph_launch_load_state(c->st, c->slots, c->dP, c->BS, (const float4 *)xv_by_id, c->POS0,
                              c->VEL, c->PREV, c->RHO, c->ID0);
    HIPCHK(hipMemsetAsync(c->dclock, 0, sizeof(unsigned long long), c->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    c->lists_valid = false;
    c->aux_order = 0;
    c->snap_valid = false;
    return SPH_OK;
}

// ---- read-back -------------------------------------------------------------

int sph_download(sph_ctx *c, fluid_particle *out)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->n_total > 0 && !out)
        return set_err(SPH_EINVAL, "sph_download: null output");
    if (c->n_total == 0) {
        HIPCHK(hipStreamSynchronize(c->st));
        return SPH_OK;
    }
    if (!c->AOS)
        if (int rc = dalloc(&c->AOS, (size_t)c->slots * 13))
            return rc;
    sph_launch_gather_aos(c->st, c->slots, c->dP, c->BS, c->POS0, c->VEL, c->ID0, c->PREV, c->RHO,
                          c->aux_order ? c->ID1 : c->ID0, c->ACC, c->AOS);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->AOS, sizeof(fluid_particle) * (size_t)c->n_total,
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return SPH_OK;
}

int sph_download_positions(sph_ctx *c, float *xy_by_id)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->n_total > 0 && !xy_by_id)
        return set_err(SPH_EINVAL, "sph_download_positions: null output");
    if (c->n_total == 0)
        return SPH_OK;
    if (!c->XY)
        if (int rc = dalloc(&c->XY, (size_t)c->slots))
            return rc;
    sph_launch_gather_xy(c->st, c->slots, c->dP, c->BS, c->POS0, c->ID0, c->XY);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(xy_by_id, c->XY, sizeof(float2) * (size_t)c->n_total,
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return SPH_OK;
}

int sph_pack_coords_i16(sph_ctx *c, short *coords)
{
    if (int rc = check_ctx(c))
        return rc;
    if (c->n_total > 0 && !coords)
        return set_err(SPH_EINVAL, "sph_pack_coords_i16: null output");
    if (c->n_total == 0)
        return SPH_OK;
    sph_launch_pack_coords(c->st, c->slots, c->dP, c->BS, c->POS0, c->ID0, c->coords);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(coords, c->coords, sizeof(short) * 2 * (size_t)c->n_total,
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return SPH_OK;
}

int sph_read_sorted(sph_ctx *c, int field, void *out)
{
    if (int rc = check_ctx(c))
        return rc;
    if (!out && c->n_total)
        return set_err(SPH_EINVAL, "sph_read_sorted: null output");
    if (int rc = canonicalize(c))
        return rc;
    const void *src = nullptr;
    int comp = -1; // component of an interleaved pair, -1 for 4-byte fields
    switch (field) {
    case SPH_FIELD_X: src = c->POS0; comp = 0; break;
    case SPH_FIELD_Y: src = c->POS0; comp = 1; break;
    case SPH_FIELD_VX: src = c->VEL; comp = 0; break;
    case SPH_FIELD_VY: src = c->VEL; comp = 1; break;
    case SPH_FIELD_XPREV: src = c->PREV; comp = 0; break;
    case SPH_FIELD_YPREV: src = c->PREV; comp = 1; break;
    case SPH_FIELD_DENSITY: src = c->RHO; comp = 0; break;
    case SPH_FIELD_DENSITY_NEAR: src = c->RHO; comp = 1; break;
    case SPH_FIELD_ID: src = c->ID0; break;
    case SPH_FIELD_KEY: src = c->K0; break;
    default: return set_err(SPH_EINVAL, "sph_read_sorted: unknown field %d", field);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->n_total == 0)
        return SPH_OK;
    const size_t N = c->slots;
    std::vector<float2> tmp(N);
    HIPCHK(hipMemcpy(tmp.data(), src, (comp < 0 ? 4 : 8) * N, hipMemcpyDeviceToHost));
    const uint32_t *w = (const uint32_t *)tmp.data();
    uint32_t *o = (uint32_t *)out;
    size_t k = 0;
    for (const Scene &S : c->sc)
        for (int i = 0; i < S.n; i++) {
            const size_t slot = S.poff + (size_t)i;
            o[k++] = comp < 0 ? w[slot] : w[2 * slot + (size_t)comp];
        }
    return SPH_OK;
}

int sph_read_cells(sph_ctx *c, unsigned *start, unsigned *count)
{
    if (int rc = check_ctx(c))
        return rc;
    const Scene &S = c->sc[0];
    std::vector<unsigned> s((size_t)S.ncells + 1);
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(s.data(), c->start + S.cell_base, sizeof(unsigned) * s.size(),
                     hipMemcpyDeviceToHost));
    for (unsigned i = 0; i < S.ncells; i++) {
        if (start)
            start[i] = s[i] - S.poff;
        if (count)
            count[i] = s[i + 1] - s[i];
    }
    return SPH_OK;
}

int sph_read_neighbors(sph_ctx *c, int stride, int *full_by_id, int *half_by_id, int *ids_by_id)
{
    if (int rc = check_ctx(c))
        return rc;
    if (stride < 0 || !full_by_id || !half_by_id || (stride > 0 && !ids_by_id))
        return set_err(SPH_EINVAL, "sph_read_neighbors: bad arguments");
    if (!c->lists_valid)
        return set_err(SPH_ESTATE, "sph_read_neighbors: no hash since upload");
    const int n = c->n_total;
    if (n == 0)
        return SPH_OK;
    int *dfull = nullptr, *dhalf = nullptr, *dids = nullptr;
    int rc = dalloc(&dfull, (size_t)n);
    if (!rc)
        rc = dalloc(&dhalf, (size_t)n);
    if (!rc)
        rc = dalloc(&dids, (size_t)n * (size_t)(stride ? stride : 1));
    if (!rc) {
        sph_launch_neighbor_lists(c->st, c->slots, c->dP, c->BS, c->POS0, c->K0, c->ID0, c->start,
                                  stride, dfull, dhalf, dids);
        hipError_t e = hipStreamSynchronize(c->st);
        if (e == hipSuccess)
            e = hipMemcpy(full_by_id, dfull, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(half_by_id, dhalf, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && stride > 0)
            e = hipMemcpy(ids_by_id, dids, sizeof(int) * (size_t)n * stride, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            rc = set_err(SPH_EHIP, "sph_read_neighbors: %s", hipGetErrorString(e));
    }
    if (dfull)
        (void)hipFree(dfull);
    if (dhalf)
        (void)hipFree(dhalf);
    if (dids)
        (void)hipFree(dids);
    return rc;
}

int sph_get_counters(sph_ctx *c, sph_counters_t *out)
{
    if (int rc = check_ctx(c))
        return rc;
    if (!out)
        return set_err(SPH_EINVAL, "sph_get_counters: null output");
    sph_dev_counters d;
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(&d, c->dC, sizeof d, hipMemcpyDeviceToHost));
    out->steps = c->steps;
    out->hashes = c->hashes;
    out->bucket_overflow = d.bucket_overflow;
    out->neighbor_overflow = d.neighbor_overflow;
    out->key_out_of_range = d.key_out_of_range;
    out->nonfinite = d.nonfinite;
    out->max_cell_count = d.max_cell_count;
    out->max_neighbors = d.max_neighbors;
    return SPH_OK;
}

int sph_reset_counters(sph_ctx *c)
{
    if (int rc = check_ctx(c))
        return rc;
    HIPCHK(hipMemsetAsync(c->dC, 0, sizeof(sph_dev_counters), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->steps = c->hashes = 0;
    return SPH_OK;
}

int sph_time_kernels(sph_ctx *c, int nsteps, int max_kinds, const char **names, float *ms,
                     int *nkinds)
{
    if (int rc = check_ctx(c))
        return rc;
    if (nsteps <= 0 || max_kinds < NKIND || !names || !ms || !nkinds)
        return set_err(SPH_EINVAL, "sph_time_kernels: need nsteps > 0 and room for %d kinds",
                       (int)NKIND);
    if (c->n_total == 0)
        return set_err(SPH_ESTATE, "sph_time_kernels: no particles");
    if (int rc = refresh_first_half_params(c))
        return rc;
    if (!c->lists_valid) {
        int rc = sph_step_eager(c, 1);
        if (rc)
            return rc;
    }
    // One start/stop event pair per launch, created up front; the launches
    // record them around the kernels themselves (hipExtLaunchKernelGGL), and
    // all nsteps steps are enqueued before the one synchronisation.
    constexpr int kLaunches = 11;
    const size_t nev = (size_t)nsteps * kLaunches * 2;
    std::vector<hipEvent_t> ev(nev, nullptr);
    std::vector<int> kind;
    hipError_t e = hipSuccess;
    for (auto &x : ev)
        if (e == hipSuccess)
            e = hipEventCreate(&x);
    if (e == hipSuccess) {
        c->tev = &ev;
        c->tkind = &kind;
        c->tidx = 0;
        for (int it = 0; it < nsteps; it++) {
            enqueue_fused_step(c, true);
            c->steps++;
            c->hashes += 2;
        }
        c->tev = nullptr;
        c->tkind = nullptr;
        c->aux_order = 1;
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->st);
    }
    double tot[NKIND] = {0};
    int cntk[NKIND] = {0};
    for (size_t i = 0; e == hipSuccess && i < kind.size(); i++) {
        float t = 0.0f;
        e = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]);
        tot[kind[i]] += t;
        cntk[kind[i]]++;
    }
    for (auto &x : ev)
        if (x)
            (void)hipEventDestroy(x);
    if (e != hipSuccess)
        return set_err(SPH_EHIP, "sph_time_kernels: %s", hipGetErrorString(e));
    for (int k = 0; k < NKIND; k++) {
        names[k] = kKindNames[k];
        ms[k] = cntk[k] ? (float)(tot[k] / cntk[k]) : 0.0f;
    }
    *nkinds = NKIND;
    return SPH_OK;
}

static int selftest_math(int fn, const char *name, unsigned lo_bits, unsigned hi_bits,
                         unsigned long long *mismatches)
{
    if (!mismatches || hi_bits < lo_bits)
        return set_err(SPH_EINVAL, "%s: bad argument", name);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(SPH_ENODEV, "no HIP device available");
    unsigned long long *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof *d));
    hipError_t e = hipMemset(d, 0, sizeof *d);
    if (e == hipSuccess) {
        sph_launch_selftest_math(nullptr, fn, lo_bits, hi_bits, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(mismatches, d, sizeof *d, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess)
        return set_err(SPH_EHIP, "%s: %s", name, hipGetErrorString(e));
    return SPH_OK;
}

int sph_selftest_sqrt(unsigned lo_bits, unsigned hi_bits, unsigned long long *mismatches)
{
    return selftest_math(0, "sph_selftest_sqrt", lo_bits, hi_bits, mismatches);
}

int sph_selftest_rcp(unsigned lo_bits, unsigned hi_bits, unsigned long long *mismatches)
{
    return selftest_math(1, "sph_selftest_rcp", lo_bits, hi_bits, mismatches);
}

} // extern "C"

int sph_internal_canonicalize(sph_ctx *c) { return canonicalize(c); }
