// gfx950 kernels of the volumetric light-transport path.
//
// k_render is a persistent path tracer: every lane owns one path at a time
// and runs a flat state machine (find next leaf -> one Woodcock step -> ray
// end / scattering event), so lanes do not wait for each other at leaf, ray
// or sample boundaries; the one deliberate wait gathers lanes whose ray has
// ended until enough of them run the ray-end code together (kEndBatch),
// and new paths and ray starts each run as one converged pass per round.
// Two long ray-end code blocks that only a few lanes of a wave need at a
// time, the phase-function half of a scattering event and the sun and sky
// radiance of an escape, are regrouped across the four waves of a
// workgroup: lanes queue them in LDS (ballot + mbcnt ranks, one LDS atomic
// per wave) and a wave that finds enough queued runs them with full lanes,
// while the queuing lanes go on (the scattering lane traces its shadow ray
// meanwhile; the escaping lane takes its next work unit).
// A lane that finishes its work unit (one pixel x a
// chunk of samples, one sample by default) refills from its wave's reserve
// of units, which one wave-aggregated atomic on a unit counter (ballot +
// mbcnt prefix) tops up, so the wave's lanes stay busy until the whole tile
// is drained. The tile's pixels are split into one row band per
// XCD, each with its own counter: the waves of an XCD work on their own band
// while it lasts (its rays see a smaller part of the volume, so the density
// and gradient bricks they touch stay in that XCD's L2), then help drain the
// others.
//
// The estimator is the reference's SampleVolume recursion
// (src/kernel.cpp:128-445) unrolled into a loop: each scattering event
// spawns one shadow ray toward lights[0] and one continuation ray whose
// throughput carries the balance-heuristic weight and the albedo.
//
// The same state machine, instantiated in adjoint mode, computes the
// gradient of <dL/dimage, image> with respect to the FOG brick densities by
// path replay: pass 0 re-traces the path to get its adjoint-weighted total
// and which shadow rays escape; pass 1 re-traces it again and, for each
// tracked segment with downstream adjoint radiance W, adds W * w_v(x)/rho(x)
// at the real collision and -W * sigma_e * integral of w_v over the leaf
// intervals the tracker covered (one uniform point per leaf, from a separate
// replayable stream). The majorants are held fixed. No path vertex is
// stored; a forward launch in record mode may keep each path's radiance and
// shadow-escape bits (40 B per sample), and an adjoint launch given that
// record starts at pass 1.
#include <hip/hip_runtime.h>

#include <mutex>

#include "drk_common.h"
#include "estimator.h"
#include "render_launch.h"

namespace drk {

// Workgroups of eight waves, two per CU (four waves per SIMD): the
// workgroup services draw their batches from twice the lanes of a
// four-wave workgroup (256^3 frame 25.5 -> 25.3 ms, 25.1 with a scattering batch of
// 48; sixteen waves per workgroup: 25.4 ms).
constexpr int kBlock = 512;

// Lane states: needs a work unit, needs its next camera path, needs its
// pending ray started, traversing to the next leaf, tracking inside a leaf,
// ray finished (escape, collision or scattering event pending), drained,
// waiting for a workgroup service (scattering or escape radiance) to return.
enum : uint32_t { ST_WORK = 0, ST_PATH = 1, ST_RAY = 2, ST_TRAV = 3, ST_LEAF = 4, ST_END = 5, ST_DONE = 6, ST_WAIT = 7 };

// Scheduling of the state machine. A traversal call runs at most
// kTravBudget node steps before the wave moves on (a lane with a longer walk
// resumes it in the next round instead of holding up the whole wave), and
// the ray-end phase, which carries the expensive scattering, sky and camera
// code, runs only once kEndBatch lanes (at most half the wave's live lanes)
// are waiting for it or no lane can make progress otherwise. Neither changes
// what any lane computes, only when; both raise the share of active lanes
// per instruction. (The macros exist so that a build can sweep them.)
#ifndef DRK_TRAV_BUDGET
#define DRK_TRAV_BUDGET 4
#endif
// The leaf-grid walk takes one cell step per round (grid_step): a step
// costs far less than a Revelles node visit, and a longer loop only holds
// the wave's other lanes (256^3 frame with 1/2/3/4 steps per round:
// 29.5/31.2/32.8/33.6 ms).
#ifndef DRK_UNIT_RESERVE
#define DRK_UNIT_RESERVE 32
#endif
#ifndef DRK_END_BATCH
#define DRK_END_BATCH 32
#endif
// Workgroup scattering service (forward modes): at a scattering event a
// lane samples only the shadow ray's direction (Philox and the sun cone) and
// starts tracing it at once; the rest of the event (the phase-function
// sample, both phase values and sun pdfs, the MIS weights) is queued in LDS
// and computed by whichever wave of the workgroup next finds kScatterBatch
// jobs queued (or none of its own lanes able to progress), so that this
// long code block runs with full waves instead of the few lanes of one wave
// whose rays end in the same round. The results (the continuation and the
// shadow ray's weight) are needed only at the shadow ray's end, where a lane
// whose job has not come back yet waits for it.
// Workgroup escape service (forward modes without the path record): when a
// path that escaped after scattering ends a work unit whose sums are still
// zero (every unit of one sample), the lane writes the unit's sums without
// the escape radiance and moves on; the sun and sky radiance along the
// escape direction (the longest code of the ray end, needed by a few lanes
// at a time) is computed by the wave that next finds kEscapeBatch such jobs
// queued, which adds it to the unit's sums in memory, in the order the lane
// would have (path radiance first, then the unit's sums).
#ifndef DRK_SCATTER_SERVICE
#define DRK_SCATTER_SERVICE 1
#endif
#ifndef DRK_ESCAPE_SERVICE
#define DRK_ESCAPE_SERVICE 1
#endif
#ifndef DRK_SCATTER_BATCH
#define DRK_SCATTER_BATCH 48
#endif
#ifndef DRK_ESCAPE_BATCH
#define DRK_ESCAPE_BATCH 48
#endif
constexpr uint32_t kScatterBatch = DRK_SCATTER_BATCH;
constexpr uint32_t kEscapeBatch = DRK_ESCAPE_BATCH;
constexpr uint32_t kQueueScatter = 0, kQueueEscape = 1;
// Per-lane service flags: the lane's scattering job is back; the lane's
// escape job (in its mailbox in memory) has not been taken yet.
constexpr uint32_t kFlagScatterDone = 1u, kFlagEscapeBusy = 2u;
// Bounds of the waits of the queue protocol: a slot or job wait spins at
// most kSpinLimit times, and a wave whose live lanes all wait gives up after
// kStallLimit rounds without a returned job (status word 2). Neither is
// reached unless the protocol is broken; they keep every wave finite.
constexpr uint32_t kSpinLimit = 1u << 20;
constexpr uint32_t kStallLimit = 1u << 22;
constexpr uint16_t kSlotFree = 0xFFFFu;
constexpr uint32_t kTravBudget = DRK_TRAV_BUDGET;
constexpr uint32_t kEndBatch = DRK_END_BATCH;
enum : uint32_t { RK_PRIMARY = 0, RK_CONT = 1, RK_SHADOW = 2, RK_SHADOW_PASS = 3 };

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// The XCD (accelerator complex die) the wave runs on: hardware register
// XCC_ID (id 20), bits 3:0, on gfx942/gfx950. Each XCD has its own L2, so
// float atomics into a per-XCD copy of the gradient never contend for a
// cache line across dies; a wave stays on its XCD for its lifetime.
__device__ __forceinline__ uint32_t xcc_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
}

// The wave's compute unit within its XCD: hardware register HW_ID (id 4),
// CU_ID bits 11:8 and SE_ID bits 14:13 on gfx942/gfx950.
__device__ __forceinline__ uint32_t cu_slot() {
    uint32_t cu = __builtin_amdgcn_s_getreg((3 << 11) | (8 << 6) | 4);
    uint32_t se = __builtin_amdgcn_s_getreg((1 << 11) | (13 << 6) | 4);
    return cu | (se << 4);
}

// Work units for the lanes that need one, from the wave's reserve first; a
// fetch (one atomic for the wave) takes what the reserve lacks plus r_extra
// more units for later refills, so that most refills do not wait for a
// device-scope atomic. The reserve is kept while the band hands out full
// chunks (chunk index < nfull) and not in the tail of short chunks, where
// units held by one wave would lengthen the drain at the end of a launch.
// Units are indices into the band's [0, total); a unit >= total means the
// band is drained. r_base, r_cnt and r_extra are wave-uniform.
constexpr uint32_t kUnitReserve = DRK_UNIT_RESERVE;
__device__ __forceinline__ uint32_t take_units(uint32_t *counter, bool need, uint32_t bp, uint32_t nfull,
                                               uint32_t total, uint32_t &r_base, uint32_t &r_cnt,
                                               uint32_t &r_extra) {
    const uint64_t mask = __ballot(need);
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    uint32_t u;
    if (cnt <= r_cnt) {
        u = r_base + rank;
        r_base += cnt;
        r_cnt -= cnt;
    } else {
        const uint32_t req = cnt - r_cnt + r_extra;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t fresh = 0;
        if (lane_id() == leader) fresh = atomicAdd(counter, req);
        fresh = __builtin_amdgcn_readfirstlane(__shfl(fresh, (int)leader));
        u = rank < r_cnt ? r_base + rank : fresh + (rank - r_cnt);
        r_base = fresh + (cnt - r_cnt);
        r_cnt = r_extra;
        r_extra = fresh < nfull * bp ? kUnitReserve : 0u;   // still in the full chunks
    }
    // Units past the band's end are not held.
    r_cnt = r_base < total ? min(r_cnt, total - r_base) : 0u;
    return u;
}

// Workgroup-shared state of the path tracer.
struct RenderLds {
    MieLds mie;
    uint32_t stk[kMaxDepth * kBlock];   // per-lane traversal stacks, lane-interleaved
    float root[6 * kBlock];             // per-lane root interval of the walk
    // Per lane, from a scattering event to the end of its shadow ray: the
    // continuation's direction and throughput and the shadow ray's weight.
    float cont[9 * kBlock];
    // Adjoint: per wave, the next free slot and the end of the chunk of the
    // event buffer it has reserved (in LDS, so that every lane sees the
    // updates made under divergent control flow).
    unsigned long long ev_cur[kBlock / 64][2];
    // Workgroup services (scattering, escape): per queue a ring of queued
    // lanes' workgroup indices (kSlotFree: empty), the next job to serve and
    // the next free ring position (both counting up); per lane the service
    // flags. A scattering job's inputs and results are in the lane's cont[]
    // slots (a lane queues at most one, and reads cont[] only at the end of
    // the shadow ray it traces meanwhile); an escape job is in the lane's
    // mailbox in memory (WorkParams::esc_mail).
    uint32_t q_flag[kBlock];
    uint32_t q_head[2], q_tail[2];
    uint16_t q_ring[2][kBlock];
};

__device__ __forceinline__ void render_lds_init(RenderLds &S) {
    if ((threadIdx.x & 63) == 0) S.ev_cur[threadIdx.x >> 6][0] = S.ev_cur[threadIdx.x >> 6][1] = 0;
    S.q_flag[threadIdx.x] = 0u;
    S.q_ring[0][threadIdx.x] = kSlotFree;
    S.q_ring[1][threadIdx.x] = kSlotFree;
    if (threadIdx.x < 2) S.q_head[threadIdx.x] = S.q_tail[threadIdx.x] = 0u;
    load_mie(S.mie);
    __syncthreads();
}

// Scattering event of a path: MIS between the phase function and lights[0]
// (src/kernel.cpp:387-430). scatter_phase: from the incoming direction, the
// shadow ray's direction lrd (sampled from the sun cone) and the phase
// function's two uniforms, the continuation's direction srd and the
// balance-heuristic weights w1 (continuation) and w2 (shadow ray); the
// throughputs are beta * w * albedo.
__device__ __forceinline__ void scatter_phase(const KernelConsts &K, const RenderParams &P, const MieLds &mie, F3 rd,
                                              F3 lrd, float u1, float u2, F3 &srd, F3 &w1, F3 &w2) {
    F3 p1, p3;
    if (P.phase == DRK_PHASE_MIE) {
        srd = mie_sample(mie, rd, u1, u2);
        p1 = mie_eval_rgb(mie, dot3(srd, rd));
        p3 = mie_eval_rgb(mie, dot3(lrd, rd));
    } else {
        srd = hg_sample(P.hg_g, rd, u1, u2);
        float h1 = hg_eval(P.hg_g, dot3(srd, rd)), h3 = hg_eval(P.hg_g, dot3(lrd, rd));
        p1 = mk3(h1, h1, h1);
        p3 = mk3(h3, h3, h3);
    }
    float pd1 = sun_pdf(K, srd), pd2 = sun_pdf(K, lrd);
    w1 = mk3(div_approx(p1.x, p1.x + pd1), div_approx(p1.y, p1.y + pd1), div_approx(p1.z, p1.z + pd1));
    w2 = mk3(div_approx(p3.x, p3.x + pd2), div_approx(p3.y, p3.y + pd2), div_approx(p3.z, p3.z + pd2));
}

// The event's four uniforms (counter: sample, event; key: frame pixel) and
// the shadow ray's direction from the last two.
__device__ __forceinline__ void scatter_uniforms(const KernelConsts &K, const RenderParams &P, uint32_t key0,
                                                 uint32_t sample, uint32_t event, float &u1, float &u2, F3 &lrd) {
    U4 r4 = philox4x32_10(U4{sample, event, 0u, kRngScatter}, key0, P.key1);
    u1 = u01(r4.x);
    u2 = u01(r4.y);
    lrd = sun_sample(K, u01(r4.z), u01(r4.w));
}

// Runs one launch's work (P, W) to completion in the calling wave: its lanes
// take work units until the units are drained, then the wave returns.
template <int kMode, bool kQueue>
__device__ __forceinline__ void render_body(const RenderParams &P, const WorkParams &W, RenderLds &S) {
    constexpr bool kSumSq = (kMode == kModeForwardSq || kMode == kModeCount);
    constexpr bool kCount = (kMode == kModeCount);
    constexpr bool kNull = (kMode == kModeAdjointNull);   // adjoint with null-collision terms
    constexpr bool kAdj = (kMode == kModeAdjoint) || kNull;
    constexpr bool kRec = (kMode == kModeRecord);
    constexpr bool kWt = (kMode == kModeWeighted);       // weighted delta tracking
    // The adjoint needs the shadow ray's weight when the ray starts, so it
    // samples its scattering events in place.
    static_assert(!(kQueue && kAdj), "the adjoint does not use the scattering service");
    constexpr bool kEscSvc = kQueue && DRK_ESCAPE_SERVICE;
    constexpr bool kScatSvc = kQueue && DRK_SCATTER_SERVICE;
    MieLds &mie = S.mie;
    unsigned long long (*ev_cur)[2] = S.ev_cur;

    uint32_t *stk = S.stk + threadIdx.x;
    float *root = S.root + threadIdx.x;
    float *cont = S.cont + threadIdx.x;
    const uint32_t stride = kBlock;
    const KernelConsts &K = *P.kc;
    const float sig_s = P.sigma_s;
    const float sig_e = P.sigma_e;
    const float albedo = sig_s / sig_e;

    // Work unit.
    uint32_t st = ST_WORK;
    uint32_t chunk_cur = 0, pix_local = 0, key0 = 0;   // key0: frame pixel index
    const uint32_t band0 = xcc_id() % W.nbands;   // first band: this wave's XCD's
    uint32_t band_step = 0;            // bands this wave has found drained (wave-uniform)
    // The wave's reserve of fetched units of the current band, [r_base,
    // r_base + r_cnt), and how many extra units its next fetch asks for
    // (wave-uniform).
    uint32_t r_base = 0, r_cnt = 0, r_extra = 0;
    uint32_t s_cur = 0, s_end = 0;
    float pix_depth = 3.402823466e38f;
    float accL[3] = {0, 0, 0}, accE[3] = {0, 0, 0}, accA = 0.0f;
    float sqL[3] = {0, 0, 0}, sqE[3] = {0, 0, 0};
    // Path.
    F3 beta = mk3(1, 1, 1), pathL = mk3(0, 0, 0), pathE = mk3(0, 0, 0);
    float alpha = 0.0f;
    uint32_t depth_r = 0, kind = RK_PRIMARY, event = 0;
    // Ray.
    F3 ro = mk3(0, 0, 0), rd = mk3(0, 0, 1);
    bool hit = false;
    // Weighted tracking: the ray's weight, the product of its collisions'
    // density ratios (exactly 1 where the densities equal the reference).
    float rw = 1.0f;
    Traversal T;
    T.lvl = -1;
    RayRng rng;
    rng.s0 = rng.s1 = rng.s2 = rng.s3 = 1u;
    // Woodcock state of the current leaf.
    float4 ce = make_float4(0, 0, 0, 1);
    uint32_t leaf_brick = 0;              // FOG brick of the current leaf (DRK_NO_BRICK: none)
    uint32_t leaf_sdf = DRK_NO_BRICK;     // SDF brick of the current leaf
    float smax = 1.0f, kmaj = 1.0f, lo = 0.0f, s0 = 0.0f, tr0 = 0.0f, tr1 = 0.0f, trb = 0.0f, sr = 0.0f, td = 0.0f;
    bool first_step = true;
    uint32_t cnt_nodes = 0, cnt_lookups = 0;
    // Scattering service: the lane waited for its job at a shadow ray's end
    // and the job is back (the ray end runs at once); rounds in which every
    // live lane of the wave waited (wave-uniform).
    bool svc = false;
    uint32_t stall_rounds = 0;
    // Adjoint state.
    uint32_t adj_d = 0;   // the pixel's offset in the image-gradient tiles
    float ltot = 0.0f, collected = 0.0f, wseg = 0.0f, seg_a = 0.0f;
    uint32_t pass = 0;
    uint64_t vis = 0;
    // The path's last contribution: the scattering event whose shadow ray
    // escaped, or the depth of the escaping continuation (-1: none). Kept by
    // the forward pass (record or adjoint pass 0); after it the replay stops.
    int32_t last_c = -1;
    RayRng arng;
    arng.s0 = arng.s1 = arng.s2 = arng.s3 = 1u;
    // Gradient copy of this wave: spread over XCDs first, then over CUs.
    float *const grad_copy =
        kAdj ? W.grad + (size_t)((xcc_id() + W.grad_nxcc * cu_slot()) % W.grad_copies) * W.grad_stride : nullptr;

    auto start_ray = [&](uint32_t ray_id) {
#ifdef DRK_POISON_LDS
        // Debug builds: the lane's stack and root interval read as all-ones
        // or NaN until the new ray writes them.
        for (uint32_t i = 0; i < kMaxDepth; ++i) stk[i * stride] = 0xffffffffu;
        for (uint32_t i = 0; i < 6; ++i) root[i * stride] = __uint_as_float(0xffffffffu);
#endif
        td = 0.0f;
        hit = false;
        T.lvl = -1;
        if (kWt) rw = 1.0f;
        rng.seed(key0, P.key1, P.sample_base + s_cur, ray_id, kRngTrack);
        if (kAdj && !kNull && pass == 1) arng.seed(key0, P.key1, P.sample_base + s_cur, ray_id, kRngAdjoint);
        F3 t0, t1;
        uint32_t a;
        if (!octree_init(P.ob, ro, rd, t0, t1, a) || t1.x < 0.0f || t1.y < 0.0f || t1.z < 0.0f) {
            st = ST_END;   // misses the root: escapes
            return;
        }
        trav_set_root(root, stride, t0, t1);
        T.t0 = t0; T.t1 = t1; T.a = a;
        if (kCount) cnt_nodes += 1;
        if (is_leaf(P.ob[0])) {
            T.lvl = -2;    // root leaf pending
        } else if (P.grid_levels && grid_start(t0, t1, P.grid_levels, T, stk)) {
            // walking the leaf grid from the start cell
        } else {
            F3 tm = mk3(0.5f * (t0.x + t1.x), 0.5f * (t0.y + t1.y), 0.5f * (t0.z + t1.z));
            stk[0] = 0u | (first_node(t0, tm) << 24);
            T.lvl = 0;
        }
        st = ST_TRAV;
    };

    auto start_path = [&]() {
#ifdef DRK_POISON_LDS
        // Debug builds: the lane's continuation slots read as NaN until an
        // event of the new path writes them.
        for (uint32_t i = 0; i < 9; ++i) cont[i * stride] = __uint_as_float(0xffffffffu);
#endif
        bool gm;
        if (W.cam) {
            const float4 c = W.cam[pix_local];
            ro = mk3(K.viewi[12], K.viewi[13], K.viewi[14]);
            rd = mk3(c.x, c.y, c.z);
            gm = c.w != 0.0f;
        } else {
            camera_ray(K, P.w, P.h, key0 % P.w, key0 / P.w, ro, rd, gm);
        }
        beta = mk3(1, 1, 1);
        pathL = mk3(0, 0, 0);
        pathE = mk3(0, 0, 0);
        alpha = 0.0f;
        depth_r = 0;
        kind = RK_PRIMARY;
        if (kAdj && pass == 1) {
            collected = 0.0f;
            wseg = ltot;   // the camera segment carries the whole path
        }
        if (kRec || (kAdj && pass == 0)) last_c = -1;
        if (P.shadow_mode) {
            // RenderKernel::Shadow (src/kernel.cpp:569-607): a sun-sampled
            // shadow ray from the surface point at the compositing depth.
            kind = RK_SHADOW_PASS;
            const bool gm1 = gm && pix_depth < K.clip_end;
            const F3 surf = mk3(ro.x + rd.x * pix_depth, ro.y + rd.y * pix_depth, ro.z + rd.z * pix_depth);
            U4 r4 = philox4x32_10(U4{P.sample_base + s_cur, 0u, 0u, kRngShadowPass}, key0, P.key1);
            rd = sun_sample(K, u01(r4.x), u01(r4.y));
            if (!gm1) { st = ST_END; hit = false; return; }   // untraced lane: unshadowed
            ro = surf;
            st = ST_RAY;
            return;
        }
        if (!gm) { st = ST_END; hit = false; return; }
        st = ST_RAY;
    };

    // Adjoint: set up the next sample of the current pixel. With a path
    // record, its total and shadow-escape bits come from the forward launch
    // and the replay starts directly at pass 1.
    // The pixel's image gradient (lights rgb, env rgb; alpha appears in both
    // buffers), re-read from the tiles where it is needed rather than held
    // in registers for the whole unit.
    auto adj_grad = [&](F3 &gL, F3 &gE, float &gA) {
        const float4 a = reinterpret_cast<const float4 *>(W.adj_l)[adj_d];
        const float4 c = reinterpret_cast<const float4 *>(W.adj_e)[adj_d];
        gL = mk3(a.x, a.y, a.z);
        gE = mk3(c.x, c.y, c.z);
        gA = a.w + c.w;
    };

    auto adj_begin = [&]() {
        if (W.rec) {
            const size_t n = (size_t)W.npix * P.samples;
            const size_t ri = (size_t)s_cur * W.npix + pix_local;
            const float4 *R = reinterpret_cast<const float4 *>(W.rec);
            float4 a = R[ri], b = R[n + ri];
            uint2 v = reinterpret_cast<const uint2 *>(R + 2 * n)[ri];
            F3 gL, gE;
            float gA;
            adj_grad(gL, gE, gA);
            ltot = gL.x * a.x + gL.y * a.y + gL.z * a.z + gE.x * b.x + gE.y * b.y + gE.z * b.z + gA * a.w;
            vis = ((uint64_t)v.y << 32) | (uint64_t)v.x;
            last_c = __float_as_int(b.w);
            pass = 1;
        } else {
            pass = 0;
            vis = 0;
        }
    };

    // Adds weight * (trilinear weight of vertex v at x) to every corner v of
    // the current leaf's gradient brick (adjoint pass 1 only): as one event
    // {weight, x; brick} in the event buffer (one 64-bit atomic per wave for
    // the slots; the reduction finds the lookup cell), or, once the buffer is
    // full, as 8 float atomics into the brick.
    auto grad_at = [&](F3 x, float weight) {
        if (weight == 0.0f) return;
        bool queued = false;
        if (W.ev_cap) {
            // Slots come from the wave's reserved chunk; when it runs out, the
            // wave's first active lane reserves the next kEvChunk slots with
            // one atomic (a single counter updated per event would serialise).
            const uint64_t m = __ballot(true);
            const uint32_t lane = lane_id();
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            const uint32_t cnt = (uint32_t)__popcll(m);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            unsigned long long *cur = ev_cur[threadIdx.x >> 6];
            const unsigned long long c0 = cur[0], avail = cur[1] - c0;
            unsigned long long fresh = 0;
            if (cnt > avail) {
                if (lane == leader) fresh = atomicAdd(W.ev_count, (unsigned long long)kEvChunk);
                const uint32_t lo = __shfl((uint32_t)fresh, (int)leader), hi = __shfl((uint32_t)(fresh >> 32), (int)leader);
                fresh = ((unsigned long long)hi << 32) | lo;
            }
            const unsigned long long slot = rank < avail ? c0 + rank : fresh + (rank - avail);
            if (lane == leader) {
                if (cnt > avail) {
                    cur[0] = fresh + (cnt - avail);
                    cur[1] = fresh + kEvChunk;
                } else {
                    cur[0] = c0 + cnt;
                }
            }
            if (slot < W.ev_cap) {
                const EventPlanes ev = event_planes(W.ev_base, W.ev_cap);
                ev.wx[slot] = make_float4(weight, x.x, x.y, x.z);
                ev.b[slot] = leaf_brick;
                queued = true;
            }
        }
        if (!queued) {
            float nl[3];
            const uint32_t cell = brick_cell(x, ce, P.lvoxc, nl);
            float *g = grad_copy + (size_t)leaf_brick * P.lvoxc3;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                uint32_t idx;
                float w;
                cell_corner(cell, nl, P.lvoxc, v, idx, w);
                unsafeAtomicAdd(g + idx, weight * w);
            }
        }
    };
    // Workgroup services. A lane queues itself (its workgroup index) after
    // writing its job: one LDS atomic per wave reserves ring positions, and
    // each lane waits until its position's previous entry has been taken
    // before it writes its index.
    auto queue_push = [&](uint32_t q) {
        const uint64_t m = __ballot(true);
        const uint32_t cnt = (uint32_t)__popcll(m);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
        uint32_t base = 0;
        if (lane_id() == leader)
            base = __hip_atomic_fetch_add(&S.q_tail[q], cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        base = __shfl(base, (int)leader);
        const uint32_t idx = (base + rank) & (kBlock - 1);
        for (uint32_t it = 0;
             __hip_atomic_load(&S.q_ring[q][idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != kSlotFree; ++it) {
            if (it >= kSpinLimit) {
                W.counter[kStatusWord] = 2u;
                break;
            }
        }
        // Release: the job (in LDS or in memory) is written before the entry
        // is seen.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __hip_atomic_store(&S.q_ring[q][idx], (uint16_t)threadIdx.x, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    // The service for lane e: in its cont[0..2] the incoming direction,
    // cont[3..5] the shadow ray's direction, cont[6..7] the phase function's
    // uniforms; out in cont[0..8]: the continuation's direction and the MIS
    // weights w1 and w2 (the lane applies its throughput and the albedo).
    auto scatter_service = [&](uint32_t e) {
        float *c = S.cont + e;
        const F3 rd_in = mk3(c[0], c[stride], c[2 * stride]);
        const F3 lrd_in = mk3(c[3 * stride], c[4 * stride], c[5 * stride]);
        F3 srd, w1, w2;
        scatter_phase(K, P, mie, rd_in, lrd_in, c[6 * stride], c[7 * stride], srd, w1, w2);
        c[0] = srd.x; c[stride] = srd.y; c[2 * stride] = srd.z;
        c[3 * stride] = w1.x; c[4 * stride] = w1.y; c[5 * stride] = w1.z;
        c[6 * stride] = w2.x; c[7 * stride] = w2.y; c[8 * stride] = w2.z;
    };
    // The escape service for lane e: its mailbox holds the escape direction,
    // the unit's sums' index, the path throughput and (record mode) the
    // record index; the unit's sums (and the record) in memory hold the
    // path's radiance without the escape. Adds the sun's
    // and the sky's radiance along the direction as the lane would have
    // (path radiance += throughput * radiance, then the unit's sums, zero,
    // + path radiance) and writes the sums of squares where they are kept.
    auto escape_service = [&](uint32_t e) {
        constexpr uint32_t kNb = kSumSq ? 4u : 2u;
        const float4 *mb = W.esc_mail + 2 * ((size_t)blockIdx.x * kBlock + e);
        const float4 a = mb[0];
        const F3 d = mk3(a.x, a.y, a.z);
        const F3 lc = lights_eval(K, d);
        const F3 le = env_radiance(P.env, K, d);
        // (The rest of the job is read after the radiance, which keeps it
        // out of the registers the sky model needs.)
        asm volatile("" ::: "memory");
        const float4 b = mb[1];
        const uint32_t dst = __float_as_uint(a.w);
        float4 *oL, *oE, *qL, *qE;
        if (W.partial) {
            float4 *o = reinterpret_cast<float4 *>(W.partial) + (size_t)dst * kNb;
            oL = o; oE = o + 1; qL = o + 2; qE = o + 3;
        } else {
            oL = reinterpret_cast<float4 *>(W.out_l) + dst;
            oE = reinterpret_cast<float4 *>(W.out_e) + dst;
            qL = reinterpret_cast<float4 *>(W.sq_l) + dst;
            qE = reinterpret_cast<float4 *>(W.sq_e) + dst;
        }
        const float4 l = *oL, v = *oE;
        const F3 pL = mk3(l.x + b.x * lc.x, l.y + b.y * lc.y, l.z + b.z * lc.z);
        const F3 pE = mk3(v.x + b.x * le.x, v.y + b.y * le.y, v.z + b.z * le.z);
        if (kRec) {
            // The path record's radiance and alpha (index in the mailbox's
            // fourth word): the unit's sums were zero before this path, so
            // they hold the path's radiance and alpha exactly, and with the
            // escape added they are the entry the lane would have written.
            // The lane writes the entry's fourth word of the second half.
            float4 *R = reinterpret_cast<float4 *>(W.rec);
            const size_t n = (size_t)W.npix * P.samples, ri = __float_as_uint(b.w);
            R[ri] = make_float4(pL.x, pL.y, pL.z, l.w);
            float *e = reinterpret_cast<float *>(R + n + ri);
            e[0] = pE.x;
            e[1] = pE.y;
            e[2] = pE.z;
        }
        *oL = make_float4(0.0f + pL.x, 0.0f + pL.y, 0.0f + pL.z, l.w);
        *oE = make_float4(0.0f + pE.x, 0.0f + pE.y, 0.0f + pE.z, v.w);
        if (kSumSq) {
            *qL = make_float4(0.0f + pL.x * pL.x, 0.0f + pL.y * pL.y, 0.0f + pL.z * pL.z, l.w);
            *qE = make_float4(0.0f + pE.x * pE.x, 0.0f + pE.y * pE.y, 0.0f + pE.z * pE.z, v.w);
        }
    };
    // Unclaimed jobs of queue q (wave-uniform).
    auto queued = [&](uint32_t q, uint32_t &head) -> uint32_t {
        head = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&S.q_head[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        const uint32_t tail = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&S.q_tail[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        return tail - head;
    };
    // A wave serves up to 64 queued jobs of one queue at once, claimed with
    // one LDS compare-and-swap on the queue head: scattering jobs once
    // kScatterBatch are queued, else escape jobs once kEscapeBatch are, or
    // either when none of its own lanes can progress. Called with every lane
    // active.
    auto serve = [&](bool stalled) {
        uint32_t h0, h1;
        const uint32_t a0 = queued(kQueueScatter, h0), a1 = queued(kQueueEscape, h1);
        uint32_t q, head, avail;
        if (kScatSvc && (a0 >= kScatterBatch || (stalled && a0 > 0))) {
            q = kQueueScatter; head = h0; avail = a0;
        } else if (kEscSvc && (a1 >= kEscapeBatch || (stalled && a1 > 0))) {
            q = kQueueEscape; head = h1; avail = a1;
        } else {
            return;
        }
        const uint32_t n = min(64u, avail);
        uint32_t ok = 0;
        if (lane_id() == 0) {
            uint32_t expect = head;
            ok = __hip_atomic_compare_exchange_strong(&S.q_head[q], &expect, head + n, __ATOMIC_RELAXED,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                     ? 1u : 0u;
        }
        if (!__builtin_amdgcn_readfirstlane(ok)) return;
        if (lane_id() < n) {
            const uint32_t idx = (head + lane_id()) & (kBlock - 1);
            uint32_t e = kSlotFree;
            for (uint32_t it = 0; it < kSpinLimit; ++it) {
                e = __hip_atomic_load(&S.q_ring[q][idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (e != kSlotFree) break;
            }
            if (e != kSlotFree) {
                __hip_atomic_store(&S.q_ring[q][idx], kSlotFree, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                if (kScatSvc && q == kQueueScatter) {
                    scatter_service(e);
                    __hip_atomic_fetch_or(&S.q_flag[e], kFlagScatterDone, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else if (kEscSvc) {
                    escape_service(e);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __hip_atomic_fetch_and(&S.q_flag[e], ~kFlagEscapeBusy, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            } else {
                W.counter[kStatusWord] = 2u;
            }
        }
    };

    // -W * sigma_e * integral over [a, b] of the trilinear weights, by one
    // uniform point of the adjoint stream (drawn for every covered leaf with a
    // non-zero majorant, so the stream stays aligned with the oracle).
    // Returns the event's point and weight (0: no event) for grad_at.
    auto grad_segment = [&](float a, float b, F3 &x) -> float {
        float u = arng.uniform();
        if (!(b > a)) return 0.0f;
        float t = a + u * (b - a);
        x = mk3(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
        return -wseg * sig_e * (b - a);
    };

    // Watchdog: a lane makes at most one step per round of this loop, so a
    // limit on the wave's rounds bounds every lane's steps. A wave past the
    // limit records it in the launch's status word and leaves; the call
    // then reports DRK_ELIMIT (the sums of unfinished units are missing).
    uint32_t rounds = 0;
    for (;;) {
        if (W.round_limit != 0 && ++rounds > W.round_limit) {
            if (lane_id() == 0) W.counter[kStatusWord] = 1u;
            break;
        }
        // ---- lanes at a shadow ray's end whose scattering job is back ------
        if (kScatSvc && st == ST_WAIT &&
            (__hip_atomic_load(&S.q_flag[threadIdx.x], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) &
             kFlagScatterDone)) {
            st = ST_END;
            svc = true;
        }
        // ---- refill lanes that finished their unit ------------------------
        // The pixels are split into W.nbands contiguous row bands; a wave
        // drains the band of its own XCD first, then the others in turn.
        bool need = (st == ST_WORK);
        while (__ballot(need)) {
            if (band_step >= W.nbands) {
                if (need) st = ST_DONE;   // every band drained
                break;
            }
            const uint32_t b = band0 + band_step < W.nbands ? band0 + band_step : band0 + band_step - W.nbands;
            const uint32_t p0 = W.band_p0[b];
            const uint32_t bp = W.band_p0[b + 1] - p0;
            const uint32_t total = bp * W.nchunks;
            const uint32_t u = take_units(W.counter + kCounterStride * b, need, bp, W.nfull, total, r_base, r_cnt,
                                          r_extra);
            if (need && u < total) {
                chunk_cur = udiv(u, W.band_div[b]);
                pix_local = p0 + (u - chunk_cur * bp);
                const uint32_t ly = udiv(pix_local, W.rx_div), lx = pix_local - ly * P.rx;
                key0 = (P.y0 + ly) * P.w + P.x0 + lx;
                s_cur = chunk_start(W.chunk, W.nfull, P.samples, chunk_cur);
                s_end = chunk_start(W.chunk, W.nfull, P.samples, chunk_cur + 1);
                pix_depth = P.depth ? P.depth[key0] : 3.402823466e38f;
                accL[0] = accL[1] = accL[2] = 0.0f;
                accE[0] = accE[1] = accE[2] = 0.0f;
                accA = 0.0f;
                sqL[0] = sqL[1] = sqL[2] = 0.0f;
                sqE[0] = sqE[1] = sqE[2] = 0.0f;
                if (kAdj) {
                    adj_d = ly * P.tilex + lx;
                    F3 gL, gE;
                    float gA;
                    adj_grad(gL, gE, gA);
                    if (gL.x == 0.0f && gL.y == 0.0f && gL.z == 0.0f && gE.x == 0.0f && gE.y == 0.0f &&
                        gE.z == 0.0f && gA == 0.0f)
                        s_cur = s_end;   // no adjoint on this pixel: nothing to differentiate
                    else
                        adj_begin();
                }
                if (s_cur < s_end) {
                    st = ST_PATH;
                    need = false;
                }
            }
            // Lanes that drew past the band's end leave it for the next one.
            if (__ballot(need && !(u < total))) {
                ++band_step;
                r_cnt = 0;
                r_extra = 0;
            }
        }
        // A wave leaves when its lanes are drained and no escape job waits
        // (those have no waiting lane; the last wave to leave serves them).
        if (!__ballot(st != ST_DONE)) {
            uint32_t h;
            if (!kEscSvc || queued(kQueueEscape, h) == 0) break;
        }

        // ---- new camera paths and ray starts, one converged pass each -----
        if (st == ST_PATH) start_path();
        if (st == ST_RAY) {
            // The ray a lane asked for follows from what it is: the camera
            // ray, the surface shadow ray, or the shadow or continuation ray
            // of scattering event `event`.
            start_ray(kind == RK_PRIMARY ? 0u
                      : kind == RK_SHADOW_PASS ? shadow_ray_id(0)
                      : kind == RK_SHADOW ? shadow_ray_id(event) : cont_ray_id(event));
        }

        // ---- A: find the next leaf along the ray --------------------------
        if (st == ST_TRAV) {
            uint32_t leaf = 0;
            float tra = 0.0f, trbn = 0.0f;
            bool found, continue_trav = false;
            if (T.lvl == -2) {
                leaf = 0;
                F3 r0, r1;
                trav_root(root, stride, r0, r1);
                tra = fmaxf(fmaxf(r0.x, r0.y), r0.z);
                trbn = fminf(fminf(r1.x, r1.y), r1.z);
                T.lvl = -1;
                found = true;
            } else {
                // One cell of the grid walk, or up to kTravBudget node
                // visits of the Revelles walk (which also takes over from a
                // grid step that met a tie); one inlined copy of each.
                int r = kTravRevelles;
                if (T.lvl == kTravGrid)
                    r = grid_step(P.cells, P.grid_levels, P.chx, T, stk, root, stride, leaf, tra, trbn, cnt_nodes);
                if (r == kTravRevelles)
                    r = trav_next(P.chx, T, stk, root, stride, leaf, tra, trbn, cnt_nodes, kTravBudget);
                if (r == kTravPending) continue_trav = true;
                found = r == kTravLeaf;
            }
            // Compositing depth stops primary rays (src/kernel.cpp:217-222).
            if (found && kind == RK_PRIMARY && !(td < pix_depth)) found = false;
            if (continue_trav) {
                // walk budget spent: stay in ST_TRAV
            } else if (!found) {
                st = ST_END;
                hit = false;
            } else {
                const drk_octree_node &n = P.ob[leaf];
                ce = *reinterpret_cast<const float4 *>(n.ce);
                leaf_brick = n.volx[DRK_VOLUME_FOG];
                leaf_sdf = n.volx[DRK_VOLUME_SDF];
                if (kCount) cnt_nodes += 1;
                // Majorant and empty-space skip, src/kernel.cpp:233-269: a leaf
                // with an SDF brick has majorant 1 and is entered at its entry
                // point only when that point lies ahead and outside the surface;
                // a fog-only leaf is always entered at its entry point (even one
                // behind the ray's position).
                bool skip = true;
                if (leaf_sdf != DRK_NO_BRICK) {
                    smax = 1.0f;
                    skip = false;
                    if (tra > td) {
                        F3 r0 = mk3(ro.x + rd.x * tra, ro.y + rd.y * tra, ro.z + rd.z * tra);
                        skip = brick_lookup(P.pvol_sdf + (size_t)leaf_sdf * P.lvoxc3, r0, ce, P.lvoxc, P.lvoxm1, P.leaf_rcp) > 0.0f;
                        if (kCount) cnt_lookups += 1;
                    }
                } else {
                    smax = n.qval[DRK_VOLUME_FOG];
                }
                lo = td;
                tr0 = tra - td;
                tr1 = trbn - td;
                trb = trbn;
                seg_a = tra;
                if (skip) {
                    td = tra;
                    s0 = tr0;
                } else {
                    s0 = 0.0f;
                }
                kmaj = sig_e * smax;
                first_step = true;
                st = ST_LEAF;
            }
        }

        // ---- B: one Woodcock step inside the current leaf -----------------
        // Adjoint pass 1: the step's gradient events (a tracked segment, and a
        // real collision), emitted after the branches so that the lanes of
        // every branch share one pass through the event code.
        F3 ev_x0 = mk3(0, 0, 0), ev_x1 = mk3(0, 0, 0);
        float ev_w0 = 0.0f, ev_w1 = 0.0f;
        if (st == ST_LEAF) {
            float step = log_unit(rng.uniform()) / kmaj;
            sr = first_step ? -step : sr - step;
            first_step = false;
            float sc = s0 + sr;
            td = lo + sc;
            if (!(sc < tr1)) {
                td = trb;     // left the leaf: memoryless restart in the next one
                st = ST_TRAV;
                if (kAdj && !kNull && pass == 1 && kmaj > 0.0f) ev_w0 = grad_segment(seg_a, trb, ev_x0);
            } else if (!(sc > tr0)) {
                hit = true;   // collision before the leaf entry counts as a real one (src/kernel.cpp:287)
                st = ST_END;
            } else {
                F3 rc = mk3(ro.x + rd.x * td, ro.y + rd.y * td, ro.z + rd.z * td);
                // Inside the surface (distance <= 0) is a collision; outside,
                // the fog density if the leaf has fog, else none (-1)
                // (src/kernel.cpp:292-315).
                float d = 1.0f, p;
                if (leaf_sdf != DRK_NO_BRICK) {
                    d = brick_lookup(P.pvol_sdf + (size_t)leaf_sdf * P.lvoxc3, rc, ce, P.lvoxc, P.lvoxm1, P.leaf_rcp);
                    p = (d > 0.0f && leaf_brick != DRK_NO_BRICK)
                            ? brick_lookup(P.pvol_fog + (size_t)leaf_brick * P.lvoxc3, rc, ce, P.lvoxc, P.lvoxm1, P.leaf_rcp)
                            : -1.0f;
                    if (kCount) cnt_lookups += (d > 0.0f && leaf_brick != DRK_NO_BRICK) ? 2 : 1;
                } else {
                    p = brick_lookup(P.pvol_fog + (size_t)leaf_brick * P.lvoxc3, rc, ce, P.lvoxc, P.lvoxm1, P.leaf_rcp);
                    if (kCount) cnt_lookups += 1;
                }
                float q = rng.uniform();
                // Weighted tracking: the reference density decides, with the
                // same comparison; the ray's weight takes the ratio of the
                // probabilities of that decision under the two densities
                // (fog-only leaves: the weighted entry point refuses SDF).
                const float pd = kWt ? brick_lookup(P.pvol_ref + (size_t)leaf_brick * P.lvoxc3, rc, ce, P.lvoxc, P.lvoxm1, P.leaf_rcp) : p;
                if (!(pd / smax < q) || !(d > 0.0f)) {
                    hit = true;
                    st = ST_END;
                    if (kWt) rw *= p / pd;
                    if (kAdj && pass == 1) {
                        if (!kNull) ev_w0 = grad_segment(seg_a, td, ev_x0);
                        ev_x1 = rc;
                        ev_w1 = wseg / p;   // d sigma / sigma at the real collision
                    }
                } else {
                    if (kWt) rw *= (smax - p) / (smax - pd);
                    if (kNull && pass == 1) {
                        // d/d rho of the null collision's probability 1 - rho/smax, relative
                        ev_x0 = rc;
                        ev_w0 = -wseg / (smax - p);
                    }
                }
            }
        }
        if (kAdj) {
            grad_at(ev_x0, ev_w0);
            grad_at(ev_x1, ev_w1);
        }

        // A shadow ray's end: its contribution, then the path's continuation
        // (run in the ray-end batch below). Returns whether the adjoint's
        // replay ends here: past the path's last contribution every segment
        // carries a zero weight (what remains of the path's total), so the
        // rest of the path has no gradient to give.
        auto shadow_end = [&]() -> bool {
            if (kScatSvc) {
                // The service left the MIS weights: the shadow weight and the
                // continuation's throughput from the path throughput at the
                // event (beta is unchanged while the shadow ray is traced).
                cont[6 * stride] = beta.x * cont[6 * stride] * albedo;
                cont[7 * stride] = beta.y * cont[7 * stride] * albedo;
                cont[8 * stride] = beta.z * cont[8 * stride] * albedo;
                cont[3 * stride] = beta.x * cont[3 * stride] * albedo;
                cont[4 * stride] = beta.y * cont[4 * stride] * albedo;
                cont[5 * stride] = beta.z * cont[5 * stride] * albedo;
            }
            if (!hit) {
                F3 le = lights_eval(K, rd);
                if (kWt) le = mk3(le.x * rw, le.y * rw, le.z * rw);
                pathL.x += cont[6 * stride] * le.x;
                pathL.y += cont[7 * stride] * le.y;
                pathL.z += cont[8 * stride] * le.z;
                if ((kAdj && pass == 0) || kRec) {
                    vis |= 1ull << event;
                    last_c = (int32_t)event;
                }
            }
            // Continue the path along the phase-sampled direction.
            rd = mk3(cont[0], cont[stride], cont[2 * stride]);
            beta = mk3(cont[3 * stride], cont[4 * stride], cont[5 * stride]);
            depth_r += 1;
            kind = RK_CONT;
            if (kAdj && pass == 1) wseg = ltot - collected;
            st = ST_RAY;
            return kAdj && pass == 1 && (int32_t)depth_r > last_c;
        };

        // ---- C: ray finished (escaped or collided) -------------------------
        // (The batch shrinks with the wave's remaining lanes, so that in the
        // drain at the end of a launch the last paths do not wait for each
        // other's ray ends.)
        const uint32_t n_live = (uint32_t)__popcll(__ballot(st != ST_DONE));
        const uint32_t end_batch = min(kEndBatch, max(1u, n_live / 2));
        // Lanes that waited for their scattering job end their ray at once.
        const bool run_end = (uint32_t)__popcll(__ballot(st == ST_END)) >= end_batch ||
                             !__ballot(st == ST_TRAV || st == ST_LEAF);
        if ((run_end || svc) && st == ST_END) {
            bool path_done = false;
            bool esc_defer = false;   // the escape radiance goes to the escape service
            svc = false;
            // Weighted tracking: a camera or continuation ray's weight joins
            // the path throughput before the ray's end uses it.
            if (kWt && kind != RK_SHADOW && kind != RK_SHADOW_PASS) beta = mk3(beta.x * rw, beta.y * rw, beta.z * rw);
            if (kind == RK_SHADOW_PASS) {
                // SampleVolume at r = scattevs returns the lights' radiance when
                // the ray escapes (r > 0); Shadow divides by lights[0]'s colour.
                if (!hit && P.scattevs > 0) {
                    F3 le = lights_eval(K, rd);
                    pathL = mk3(div_approx(le.x, K.sun_color[0][0]), div_approx(le.y, K.sun_color[0][1]),
                                div_approx(le.z, K.sun_color[0][2]));
                }
                alpha = 1.0f;
                path_done = true;
            } else if (kind == RK_SHADOW) {
                // The continuation and the shadow weight come from the
                // scattering service; wait for them if they are not back.
                if (kScatSvc && !(__hip_atomic_load(&S.q_flag[threadIdx.x], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) &
                                kFlagScatterDone))
                    st = ST_WAIT;
                else
                    path_done = shadow_end();
            } else if (!hit) {
                if (depth_r > 0 && ((kAdj && pass == 0) || kRec)) last_c = (int32_t)depth_r;
                // The adjoint's gradient pass takes the path's radiance from
                // pass 0 or the record, so it does not evaluate the escape's
                // sun and sky again.
                if (depth_r > 0 && !(kAdj && pass == 1)) {
                    // The escape service takes the radiance when this path
                    // ends a unit whose sums are zero so far and the lane's
                    // previous escape job has been taken.
                    if (kEscSvc && W.esc_mail && s_cur + 1 == s_end && accA == 0.0f && accL[0] == 0.0f &&
                        (!kRec || (uint64_t)s_cur * W.npix + pix_local <= 0xffffffffull) &&
                        accL[1] == 0.0f && accL[2] == 0.0f && accE[0] == 0.0f && accE[1] == 0.0f &&
                        accE[2] == 0.0f &&
                        !(__hip_atomic_load(&S.q_flag[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) &
                          kFlagEscapeBusy)) {
                        esc_defer = true;
                        uint32_t dst;
                        if (W.partial) {
                            dst = chunk_cur * W.npix + pix_local;
                        } else {
                            const uint32_t ly = udiv(pix_local, W.rx_div), lx = pix_local - ly * P.rx;
                            dst = ly * P.tilex + lx;
                        }
                        float4 *mb = W.esc_mail + 2 * ((size_t)blockIdx.x * kBlock + threadIdx.x);
                        mb[0] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(dst));
                        mb[1] = make_float4(beta.x, beta.y, beta.z,
                                            __uint_as_float(kRec ? s_cur * W.npix + pix_local : 0u));
                    } else {
                        F3 lc = lights_eval(K, rd);
                        F3 le = env_radiance(P.env, K, rd);
                        pathL.x += beta.x * lc.x; pathL.y += beta.y * lc.y; pathL.z += beta.z * lc.z;
                        pathE.x += beta.x * le.x; pathE.y += beta.y * le.y; pathE.z += beta.z * le.z;
                    }
                }
                path_done = true;
            } else {
                if (depth_r == 0) {
                    alpha = kWt ? rw : 1.0f;
                    if (kAdj && pass == 1) {
                        F3 gL, gE;
                        float gA;
                        adj_grad(gL, gE, gA);
                        collected += gA;
                    }
                }
                if (depth_r >= P.scattevs) {
                    path_done = true;
                } else {
                    // Scattering event (src/kernel.cpp:387-430): the
                    // continuation waits in cont[] for the shadow ray's end.
                    event = depth_r;
                    F3 lrd, shadow_w;
                    if (kScatSvc) {
                        shadow_w = mk3(0, 0, 0);   // (used by the adjoint only, which does not queue)
                        // The shadow ray's direction now; the rest of the
                        // event from the workgroup service, meanwhile.
                        float u1, u2;
                        scatter_uniforms(K, P, key0, P.sample_base + s_cur, event, u1, u2, lrd);
                        cont[0] = rd.x; cont[stride] = rd.y; cont[2 * stride] = rd.z;
                        cont[3 * stride] = lrd.x; cont[4 * stride] = lrd.y; cont[5 * stride] = lrd.z;
                        cont[6 * stride] = u1; cont[7 * stride] = u2;
                        __hip_atomic_fetch_and(&S.q_flag[threadIdx.x], ~kFlagScatterDone, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        queue_push(kQueueScatter);
                    } else {
                        F3 srd, w1, w2;
                        float u1, u2;
                        scatter_uniforms(K, P, key0, P.sample_base + s_cur, event, u1, u2, lrd);
                        scatter_phase(K, P, mie, rd, lrd, u1, u2, srd, w1, w2);
                        shadow_w = mk3(beta.x * w2.x * albedo, beta.y * w2.y * albedo, beta.z * w2.z * albedo);
                        cont[0] = srd.x; cont[stride] = srd.y; cont[2 * stride] = srd.z;
                        cont[3 * stride] = beta.x * w1.x * albedo;
                        cont[4 * stride] = beta.y * w1.y * albedo;
                        cont[5 * stride] = beta.z * w1.z * albedo;
                        cont[6 * stride] = shadow_w.x; cont[7 * stride] = shadow_w.y; cont[8 * stride] = shadow_w.z;
                    }
                    F3 x = mk3(ro.x + rd.x * td, ro.y + rd.y * td, ro.z + rd.z * td);
                    ro = x;
                    rd = lrd;
                    kind = RK_SHADOW;
                    if (kAdj && pass == 1) {
                        // The shadow ray's segment carries only its own contribution.
                        wseg = 0.0f;
                        if ((vis >> event) & 1ull) {
                            F3 le = lights_eval(K, lrd);
                            F3 gL, gE;
                            float gA;
                            adj_grad(gL, gE, gA);
                            wseg = gL.x * (shadow_w.x * le.x) + gL.y * (shadow_w.y * le.y) + gL.z * (shadow_w.z * le.z);
                            collected += wseg;
                        }
                    }
                    st = ST_RAY;
                    if (kAdj && pass == 1 && !((vis >> event) & 1ull)) {
                        // Blocked in the forward pass: no contribution, no
                        // gradient and no random numbers, so the continuation
                        // starts at once instead of after an empty shadow ray
                        // (256^3 frame fwd+adjoint 100.5 -> 90.9 ms).
                        hit = true;
                        path_done = shadow_end();
                    }
                }
            }
            if (path_done && kAdj) {
                if (pass == 0) {
                    F3 gL, gE;
                    float gA;
                    adj_grad(gL, gE, gA);
                    ltot = gL.x * pathL.x + gL.y * pathL.y + gL.z * pathL.z + gE.x * pathE.x + gE.y * pathE.y +
                           gE.z * pathE.z + gA * alpha;
                    // Replay the same path for the gradients. (A zero total does not
                    // mean zero sub-path weights when adjoints have mixed signs.)
                    pass = 1;
                    st = ST_PATH;
                    path_done = false;
                }
                if (path_done) {
                    ++s_cur;
                    if (s_cur < s_end) {
                        adj_begin();
                        st = ST_PATH;
                    } else {
                        st = ST_WORK;
                    }
                }
            } else if (path_done) {
                if (kRec) {
                    const size_t n = (size_t)W.npix * P.samples;
                    const size_t ri = (size_t)s_cur * W.npix + pix_local;
                    float4 *R = reinterpret_cast<float4 *>(W.rec);
                    if (kEscSvc && esc_defer) {
                        // The escape service writes the entry's radiance and
                        // alpha (from the unit's sums, which are this path's);
                        // the lane writes where its last contribution arose.
                        reinterpret_cast<float *>(R + n + ri)[3] = __int_as_float(last_c);
                    } else {
                        R[ri] = make_float4(pathL.x, pathL.y, pathL.z, alpha);
                        R[n + ri] = make_float4(pathE.x, pathE.y, pathE.z, __int_as_float(last_c));
                    }
                    reinterpret_cast<uint2 *>(R + 2 * n)[ri] = make_uint2((uint32_t)vis, (uint32_t)(vis >> 32));
                    vis = 0;
                }
                accL[0] += pathL.x; accL[1] += pathL.y; accL[2] += pathL.z;
                accE[0] += pathE.x; accE[1] += pathE.y; accE[2] += pathE.z;
                accA += alpha;
                if (kSumSq) {
                    sqL[0] += pathL.x * pathL.x; sqL[1] += pathL.y * pathL.y; sqL[2] += pathL.z * pathL.z;
                    sqE[0] += pathE.x * pathE.x; sqE[1] += pathE.y * pathE.y; sqE[2] += pathE.z * pathE.z;
                }
                ++s_cur;
                if (s_cur < s_end) {
                    st = ST_PATH;
                } else {
                    // Unit complete: write its sums.
                    if (W.partial) {
                        // The unit's buffers side by side: one lane writes
                        // whole 32-byte (64 with squares) spans.
                        constexpr uint32_t kNb = kSumSq ? 4u : 2u;
                        float4 *oP = reinterpret_cast<float4 *>(W.partial) +
                                     ((size_t)chunk_cur * W.npix + pix_local) * kNb;
                        oP[0] = make_float4(accL[0], accL[1], accL[2], accA);
                        oP[1] = make_float4(accE[0], accE[1], accE[2], accA);
                        if (kSumSq) {
                            oP[2] = make_float4(sqL[0], sqL[1], sqL[2], accA);
                            oP[3] = make_float4(sqE[0], sqE[1], sqE[2], accA);
                        }
                    } else {
                        const uint32_t ly = udiv(pix_local, W.rx_div), lx = pix_local - ly * P.rx;
                        size_t d = (size_t)ly * P.tilex + lx;
                        reinterpret_cast<float4 *>(W.out_l)[d] = make_float4(accL[0], accL[1], accL[2], accA);
                        reinterpret_cast<float4 *>(W.out_e)[d] = make_float4(accE[0], accE[1], accE[2], accA);
                        if (kSumSq) {
                            reinterpret_cast<float4 *>(W.sq_l)[d] = make_float4(sqL[0], sqL[1], sqL[2], accA);
                            reinterpret_cast<float4 *>(W.sq_e)[d] = make_float4(sqE[0], sqE[1], sqE[2], accA);
                        }
                    }
                    if (kEscSvc && esc_defer) {
                        // The sums above lack the escape radiance: queue the
                        // job that adds it (after the sums and the mailbox
                        // are written).
                        __hip_atomic_fetch_or(&S.q_flag[threadIdx.x], kFlagEscapeBusy, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
                        queue_push(kQueueEscape);
                    }
                    st = ST_WORK;
                }
            }
        }

        // ---- workgroup services ------------------------------------------
        if (kScatSvc || kEscSvc) {
            const bool stalled = !__ballot(st != ST_WAIT && st != ST_DONE);
            stall_rounds = stalled ? stall_rounds + 1u : 0u;
            if (stall_rounds > kStallLimit) {
                // A job did not come back: leave with the status word set.
                if (lane_id() == 0) W.counter[kStatusWord] = 2u;
                if (st == ST_WAIT) st = ST_DONE;
            }
            // Jobs are queued by the ray-end batch, so a wave looks for a
            // batch to serve only in rounds where it ran one (or when it
            // cannot progress otherwise): 256^3 frame 25.8 -> 25.5 ms.
            if (stalled || run_end) serve(stalled);
        }
    }
    if (kAdj && W.ev_cap) {
        // The unused tail of the wave's last chunk gets zero-weight events.
        const unsigned long long *cur = ev_cur[threadIdx.x >> 6];
        const unsigned long long end = min(cur[1], (unsigned long long)W.ev_cap);
        const EventPlanes ev = event_planes(W.ev_base, W.ev_cap);
        for (unsigned long long i = cur[0] + lane_id(); i < end; i += 64) {
            ev.wx[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            ev.b[i] = 0u;
        }
    }
    if (kCount) {
        atomicAdd(reinterpret_cast<unsigned long long *>(W.counts) + 0, (unsigned long long)cnt_nodes);
        atomicAdd(reinterpret_cast<unsigned long long *>(W.counts) + 1, (unsigned long long)cnt_lookups);
    }
}

template <int kMode>
// Four waves per SIMD (at most 128 VGPRs; the root interval and the
// continuation state wait in LDS so that nothing spills): the fourth wave
// hides more of the traversal's dependent loads (256^3 frame: 52.9 -> 46.9 ms).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_render(
    const LaunchParams *__restrict__ L) {
    __shared__ RenderLds S;
    render_lds_init(S);
    render_body<kMode, true>(L->P, L->W, S);
}

// The adjoint, like the forward, at four waves per SIMD: its image gradient
// is re-read from the tiles where needed, not kept in registers, so it fits
// 128 VGPRs without spills (256^3 frame forward+adjoint step 132.8 -> 124.3 ms).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_render_adjoint(
    const LaunchParams *__restrict__ L) {
    __shared__ RenderLds S;
    render_lds_init(S);
    render_body<kModeAdjoint, false>(L->P, L->W, S);
}

// The adjoint with null-collision terms (the pathwise derivative of the
// weighted forward), same occupancy as the default adjoint.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_render_adjoint_null(
    const LaunchParams *__restrict__ L) {
    __shared__ RenderLds S;
    render_lds_init(S);
    render_body<kModeAdjointNull, false>(L->P, L->W, S);
}

// Many launches (scenes, views, tiles) in one persistent grid. Workgroup b
// starts at job b % njobs and visits every job once; each of its waves drains
// the job's work units (shared with every other workgroup on that job) before
// moving on, so the job index is wave-uniform and the job's parameters are
// read with scalar loads.
template <int kMode>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_render_batch(const RenderParams *__restrict__ Ps,
                                                         const WorkParams *__restrict__ Ws, uint32_t njobs) {
    __shared__ RenderLds S;
    render_lds_init(S);
    const uint32_t j0 = blockIdx.x % njobs;
    for (uint32_t i = 0; i < njobs; ++i) {
        const uint32_t j = (j0 + i < njobs) ? j0 + i : j0 + i - njobs;
        render_body<kMode, false>(Ps[j], Ws[j], S);
    }
}

// Sums the per-chunk partials of each pixel in chunk order (deterministic).
template <uint32_t kNb>
__global__ void k_reduce(WorkParams W, uint32_t rx, uint32_t tilex) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W.npix) return;
    const float4 *part = reinterpret_cast<const float4 *>(W.partial) + (size_t)i * kNb;
    const size_t cstride = (size_t)W.npix * kNb;
    float4 s[kNb];
#pragma unroll
    for (uint32_t b = 0; b < kNb; ++b) s[b] = make_float4(0, 0, 0, 0);
    for (uint32_t c = 0; c < W.nchunks; ++c) {
#pragma unroll
        for (uint32_t b = 0; b < kNb; ++b) {
            const float4 v = part[(size_t)c * cstride + b];
            s[b].x += v.x; s[b].y += v.y; s[b].z += v.z; s[b].w += v.w;
        }
    }
    float *outs[4] = {W.out_l, W.out_e, W.sq_l, W.sq_e};
    uint32_t lx = i % rx, ly = i / rx;
    size_t d = (size_t)ly * tilex + lx;
#pragma unroll
    for (uint32_t b = 0; b < kNb; ++b) reinterpret_cast<float4 *>(outs[b])[d] = s[b];
}

static void launch_reduce_kernel(const WorkParams &W, uint32_t rx, uint32_t tilex, uint32_t buffers, hipStream_t s) {
    const uint32_t nb = (W.npix + 255) / 256;
    if (buffers == 4) k_reduce<4><<<nb, 256, 0, s>>>(W, rx, tilex);
    else k_reduce<2><<<nb, 256, 0, s>>>(W, rx, tilex);
}

// Launch setup: clears the counter block and stores the launch's parameters
// after it (one workgroup).
__global__ void k_launch_setup(RenderParams P, WorkParams W) {
    for (uint32_t i = threadIdx.x; i < kCounterWords; i += blockDim.x) W.counter[i] = 0u;
    if (threadIdx.x == 0) {
        LaunchParams *L = reinterpret_cast<LaunchParams *>(reinterpret_cast<char *>(W.counter) + kParamOffset);
        L->P = P;
        L->W = W;
    }
}

// The camera ray of every pixel of the rendered extent (W.cam): one launch,
// or one row of blocks per job of a batch.
__device__ __forceinline__ void camera_ray_of(const RenderParams &P, const WorkParams &W, uint32_t i) {
    if (i >= W.npix || !W.cam) return;
    const uint32_t ly = udiv(i, W.rx_div), lx = i - ly * P.rx;
    F3 ro, rd;
    bool gm;
    camera_ray(*P.kc, P.w, P.h, P.x0 + lx, P.y0 + ly, ro, rd, gm);
    W.cam[i] = make_float4(rd.x, rd.y, rd.z, gm ? 1.0f : 0.0f);
}
__global__ void k_camera_rays(const LaunchParams *__restrict__ L) {
    camera_ray_of(L->P, L->W, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ void k_camera_rays_batch(const RenderParams *__restrict__ Ps, const WorkParams *__restrict__ Ws) {
    camera_ray_of(Ps[blockIdx.y], Ws[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float *p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host-side launchers.
// ---------------------------------------------------------------------------
int render_block_size() { return kBlock; }

int render_blocks_per_cu(int mode) {
    int nb = 0;
    hipError_t e;
    switch (mode) {
    case kModeForward: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kModeForward>, kBlock, 0); break;
    case kModeForwardSq: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kModeForwardSq>, kBlock, 0); break;
    case kModeCount: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kModeCount>, kBlock, 0); break;
    case kModeRecord: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kModeRecord>, kBlock, 0); break;
    case kModeWeighted: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kModeWeighted>, kBlock, 0); break;
    case kModeAdjointNull: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_adjoint_null, kBlock, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_adjoint, kBlock, 0); break;
    }
    if (e != hipSuccess || nb <= 0) return 1;
    return nb;
}

int render_batch_blocks_per_cu(int mode) {
    int nb = 0;
    hipError_t e;
    switch (mode) {
    case kModeForwardSq: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<kModeForwardSq>, kBlock, 0); break;
    case kModeRecord: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<kModeRecord>, kBlock, 0); break;
    case kModeAdjoint: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<kModeAdjoint>, kBlock, 0); break;
    case kModeAdjointNull:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<kModeAdjointNull>, kBlock, 0);
        break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<kModeForward>, kBlock, 0); break;
    }
    if (e != hipSuccess || nb <= 0) {
        (void)hipGetLastError();
        return 1;
    }
    return nb;
}

hipError_t launch_render_batch(const RenderParams *dP, const WorkParams *dW, uint32_t njobs, uint32_t *counters,
                               uint32_t ncounters, uint32_t grid, int mode, uint32_t cam_pix, hipStream_t s) {
    if (njobs == 0) return hipSuccess;
    if (counters) k_fill_u32<<<(ncounters + 255) / 256, 256, 0, s>>>(counters, ncounters, 0u);
    for (uint32_t j0 = 0; cam_pix && j0 < njobs; j0 += 65535u)
        k_camera_rays_batch<<<dim3((cam_pix + 255) / 256, min(65535u, njobs - j0)), 256, 0, s>>>(dP + j0, dW + j0);
    switch (mode) {
    case kModeForwardSq: k_render_batch<kModeForwardSq><<<grid, kBlock, 0, s>>>(dP, dW, njobs); break;
    case kModeRecord: k_render_batch<kModeRecord><<<grid, kBlock, 0, s>>>(dP, dW, njobs); break;
    case kModeAdjoint: k_render_batch<kModeAdjoint><<<grid, kBlock, 0, s>>>(dP, dW, njobs); break;
    case kModeAdjointNull: k_render_batch<kModeAdjointNull><<<grid, kBlock, 0, s>>>(dP, dW, njobs); break;
    default: k_render_batch<kModeForward><<<grid, kBlock, 0, s>>>(dP, dW, njobs); break;
    }
    return hipGetLastError();
}

__global__ void k_job_setup(RenderParams P, WorkParams W, RenderParams *dP, WorkParams *dW) {
    for (uint32_t i = threadIdx.x; i < kCounterWords; i += blockDim.x) W.counter[i] = 0u;
    if (threadIdx.x == 0) {
        *dP = P;
        *dW = W;
    }
}

hipError_t launch_job_setup(const RenderParams &P, const WorkParams &W, RenderParams *dP, WorkParams *dW,
                            hipStream_t s) {
    k_job_setup<<<1, 256, 0, s>>>(P, W, dP, dW);
    return hipGetLastError();
}

hipError_t launch_reduce(const WorkParams &W, uint32_t rx, uint32_t tilex, uint32_t buffers, hipStream_t s) {
    if (!W.partial || W.npix == 0) return hipSuccess;
    launch_reduce_kernel(W, rx, tilex, buffers, s);
    return hipGetLastError();
}

hipError_t launch_render(const RenderParams &P, const WorkParams &W, uint32_t grid, int mode, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1) {
    k_launch_setup<<<1, 256, 0, s>>>(P, W);
    const LaunchParams *L = reinterpret_cast<const LaunchParams *>(reinterpret_cast<char *>(W.counter) + kParamOffset);
    if (W.cam && W.npix) k_camera_rays<<<(W.npix + 255) / 256, 256, 0, s>>>(L);
    if (ev0) (void)hipEventRecord(ev0, s);
    switch (mode) {
    case kModeForward: k_render<kModeForward><<<grid, kBlock, 0, s>>>(L); break;
    case kModeForwardSq: k_render<kModeForwardSq><<<grid, kBlock, 0, s>>>(L); break;
    case kModeCount: k_render<kModeCount><<<grid, kBlock, 0, s>>>(L); break;
    case kModeRecord: k_render<kModeRecord><<<grid, kBlock, 0, s>>>(L); break;
    case kModeWeighted: k_render<kModeWeighted><<<grid, kBlock, 0, s>>>(L); break;
    case kModeAdjointNull: k_render_adjoint_null<<<grid, kBlock, 0, s>>>(L); break;
    default: k_render_adjoint<<<grid, kBlock, 0, s>>>(L); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev1) (void)hipEventRecord(ev1, s);
    if (W.partial && mode != kModeAdjoint && mode != kModeAdjointNull) {
        launch_reduce_kernel(W, P.rx, P.tilex, partial_buffers(mode), s);
        e = hipGetLastError();
    }
    return e;
}

__global__ void k_grad_sum(const float *part, uint32_t copies, size_t n, float *out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float acc = out[i];
        for (uint32_t c = 0; c < copies; ++c) acc += part[(size_t)c * n + i];
        out[i] = acc;
    }
}

hipError_t launch_grad_sum(const float *part, uint32_t copies, size_t n, float *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t nb = std::min<size_t>((n + 255) / 256, 65536);
    k_grad_sum<<<(uint32_t)nb, 256, 0, s>>>(part, copies, n, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradient events: counting sort by bin, then LDS accumulation per bin.
// Every kernel reads the event count from device memory and clamps it to the
// capacity, so the launches are sized by capacity and need no host sync.
// ---------------------------------------------------------------------------
constexpr uint32_t kEvSortChunk = 65536;   // events per sort block
constexpr uint32_t kEvSortBlock = 1024;
constexpr uint32_t kEvAccBlock = 256;

__device__ __forceinline__ uint32_t ev_n(const EventReduce &E) {
    const unsigned long long c = *reinterpret_cast<const unsigned long long *>(E.bins);
    return c < E.ev_cap ? (uint32_t)c : E.ev_cap;
}

// Bin tables after the 2-word count.
// bin_max: the largest |weight| of each bin's events, as float bits (which
// order like the values for non-negative floats; above 0x7f7fffff: an
// infinite or NaN weight), gathered by the scatter for the accumulation.
struct BinTables {
    uint32_t *hist, *bin_start, *bin_fill, *blk_start, *bin_max;
};
__device__ __forceinline__ BinTables bin_tables(const EventReduce &E) {
    BinTables t;
    t.hist = E.bins + 2;
    t.bin_start = t.hist + E.nbins;
    t.bin_fill = t.bin_start + E.nbins + 1;
    t.blk_start = t.bin_fill + E.nbins;
    t.bin_max = t.blk_start + E.nbins + 1;
    return t;
}

__global__ __launch_bounds__(kEvSortBlock) void k_ev_hist(EventReduce E) {
    extern __shared__ uint32_t lh[];
    const uint32_t n = ev_n(E);
    const uint64_t start = (uint64_t)blockIdx.x * kEvSortChunk;
    if (start >= n) return;
    const uint32_t end = (uint32_t)min<uint64_t>(start + kEvSortChunk, n);
    const EventPlanes ev = event_planes(E.ev_base, E.ev_cap);
    const BinTables t = bin_tables(E);
    for (uint32_t b = threadIdx.x; b < E.nbins; b += kEvSortBlock) lh[b] = 0;
    __syncthreads();
    for (uint32_t i = (uint32_t)start + threadIdx.x; i < end; i += kEvSortBlock) atomicAdd(&lh[ev.b[i] >> E.shift], 1u);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < E.nbins; b += kEvSortBlock)
        if (lh[b]) atomicAdd(&t.hist[b], lh[b]);
}

// Exclusive scans of the bin sizes (event offsets) and of their chunk counts
// (accumulation block offsets); resets the scatter cursors.
__global__ __launch_bounds__(1024) void k_ev_scan(EventReduce E) {
    __shared__ uint32_t part_e[1024], part_b[1024];
    const BinTables T = bin_tables(E);
    const uint32_t per = (E.nbins + 1023) / 1024;
    const uint32_t b0 = min(threadIdx.x * per, E.nbins), b1 = min(b0 + per, E.nbins);
    uint32_t se = 0, sb = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        se += T.hist[b];
        sb += (T.hist[b] + (1u << E.acc_log2) - 1) >> E.acc_log2;
    }
    part_e[threadIdx.x] = se;
    part_b[threadIdx.x] = sb;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ae = 0, ab = 0;
        for (uint32_t t = 0; t < 1024; ++t) {
            const uint32_t e = part_e[t], b = part_b[t];
            part_e[t] = ae;
            part_b[t] = ab;
            ae += e;
            ab += b;
        }
        T.bin_start[E.nbins] = ae;
        T.blk_start[E.nbins] = ab;
    }
    __syncthreads();
    se = part_e[threadIdx.x];
    sb = part_b[threadIdx.x];
    for (uint32_t b = b0; b < b1; ++b) {
        T.bin_start[b] = se;
        T.blk_start[b] = sb;
        T.bin_fill[b] = 0;
        se += T.hist[b];
        sb += (T.hist[b] + (1u << E.acc_log2) - 1) >> E.acc_log2;
    }
}

// Exclusive prefix sum of a[0, n) in place by the whole block (kEvSortBlock
// threads; wsum: 16 words of LDS). Ends with a barrier.
__device__ __forceinline__ void block_exscan(uint32_t *a, uint32_t n, uint32_t *wsum) {
    const uint32_t per = (n + kEvSortBlock - 1) / kEvSortBlock;
    const uint32_t b0 = min(threadIdx.x * per, n), b1 = min(b0 + per, n);
    uint32_t s = 0;
    for (uint32_t b = b0; b < b1; ++b) s += a[b];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = s;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = x - s;
    for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t v = a[b];
        a[b] = base;
        base += v;
    }
    __syncthreads();
}

// Writes every event of the block's kEvSortChunk to its slot in the copy
// sorted by bin. Each bin's run of the block is reserved with one atomic;
// the events then go through LDS kEvTile at a time, ordered by bin (a
// counting sort of the tile), so that consecutive lanes store consecutive
// slots of a bin's run. Stored straight from their lanes (one cache line per
// lane), the 16-byte events kept the texture addresser busy 80% of the
// kernel's time (256^3 frame: 9.4 ms; through the tiles 7.2 ms; tiles of 3072 or 6144
// events measured no faster). The scatter also gathers each bin's largest
// |weight| (an LDS maximum per event, one global maximum per bin and block)
// so that the accumulation needs no pass of its own for the fixed-point
// scale. LDS: three words per bin (<= 8192 bins) and 20 B per tile slot, at
// most 136 KB.
constexpr uint32_t kEvTile = 2048;
__host__ __device__ constexpr size_t ev_scatter_lds(uint32_t nbins) {
    return ((size_t)3 * nbins + 16 + 3) / 4 * 16 + (size_t)kEvTile * (16 + 4);
}
__global__ __launch_bounds__(kEvSortBlock) void k_ev_scatter(EventReduce E) {
    extern __shared__ uint32_t lds[];
    const uint32_t nb = E.nbins;
    uint32_t *cur = lds, *th = lds + nb, *tmx = lds + 2 * nb, *wsum = lds + 3 * nb;
    float4 *tw = reinterpret_cast<float4 *>(lds + (3 * nb + 16 + 3) / 4 * 4);
    uint32_t *tb = reinterpret_cast<uint32_t *>(tw + kEvTile);
    const uint32_t n = ev_n(E);
    const uint64_t start = (uint64_t)blockIdx.x * kEvSortChunk;
    if (start >= n) return;
    const uint32_t end = (uint32_t)min<uint64_t>(start + kEvSortChunk, n);
    const EventPlanes ev = event_planes(E.ev_base, E.ev_cap);
    const BinTables t = bin_tables(E);
    for (uint32_t b = threadIdx.x; b < nb; b += kEvSortBlock) th[b] = tmx[b] = 0;
    __syncthreads();
    for (uint32_t i = (uint32_t)start + threadIdx.x; i < end; i += kEvSortBlock) atomicAdd(&th[ev.b[i] >> E.shift], 1u);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += kEvSortBlock)
        cur[b] = th[b] ? t.bin_start[b] + atomicAdd(&t.bin_fill[b], th[b]) : 0u;
    constexpr uint32_t kPer = kEvTile / kEvSortBlock;
    for (uint32_t t0 = (uint32_t)start; t0 < end; t0 += kEvTile) {
        const uint32_t tn = min(kEvTile, end - t0);
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < nb; b += kEvSortBlock) th[b] = 0;
        __syncthreads();
        uint32_t br[kPer], rk[kPer];
        float4 w[kPer];
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kEvSortBlock;
            if (j < tn) {
                br[u] = ev.b[t0 + j];
                w[u] = ev.wx[t0 + j];
                rk[u] = atomicAdd(&th[br[u] >> E.shift], 1u);
                atomicMax(&tmx[br[u] >> E.shift], __float_as_uint(fabsf(w[u].x)));
            }
        }
        __syncthreads();
        block_exscan(th, nb, wsum);   // th: the tile's first slot of every bin
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kEvSortBlock;
            if (j < tn) {
                const uint32_t p = th[br[u] >> E.shift] + rk[u];
                tb[p] = br[u];
                tw[p] = w[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            const uint32_t p = threadIdx.x + u * kEvSortBlock;
            if (p < tn) {
                const uint32_t brick = tb[p];
                const uint32_t b = brick >> E.shift;
                const uint32_t g = cur[b] + (p - th[b]);
                ev.swx[g] = tw[p];
                ev.sb[g] = brick;
            }
        }
        __syncthreads();
        // Advance the block's cursors by the tile's counts.
        for (uint32_t b = threadIdx.x; b < nb; b += kEvSortBlock)
            cur[b] += (b + 1 < nb ? th[b + 1] : tn) - th[b];
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += kEvSortBlock)
        if (tmx[b]) atomicMax(&t.bin_max[b], tmx[b]);
}

// One block per (bin, chunk of at most 2^acc_log2 events): sums the chunk's
// trilinear splats into an LDS copy of the bin's bricks, then adds the
// non-zero sums to the gradient. The LDS sums are 64-bit fixed point: LDS
// integer atomics run about ten times faster than ds_add_f32, and integer
// sums do not depend on the order the lanes add in. The scale is a power of
// two from the largest |weight| M of the chunk's bin (gathered by the
// scatter): every splat (|weight * w| <= M, w <= 1) is rounded to a multiple
// of 2^-S with M * 2^S < 2^(59 - acc_log2), so the at most 2^(acc_log2 + 3)
// splats of a chunk stay below 2^62, and a splat loses at most
// 2^(-60 + acc_log2) M to rounding (2^-46 M for chunks of 2^14 events; float
// sums lose 2^-24 of the running sum per add). A bin holding a non-finite weight takes
// a float path instead, so that the NaN or infinity reaches the gradient.
__global__ __launch_bounds__(kEvAccBlock) void k_ev_accum(EventReduce E) {
    extern __shared__ unsigned long long lq[];
    const EventPlanes ev = event_planes(E.ev_base, E.ev_cap);
    const BinTables t = bin_tables(E);
    const uint32_t total = t.blk_start[E.nbins];
    if (blockIdx.x >= total) return;
    // Bin of this block: the last b with blk_start[b] <= blockIdx.x.
    uint32_t lo = 0, hi = E.nbins;   // blk_start[lo] <= blockIdx.x < blk_start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (t.blk_start[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const uint32_t b = lo;
    const uint32_t e0 = t.bin_start[b] + ((blockIdx.x - t.blk_start[b]) << E.acc_log2);
    const uint32_t e1 = min(e0 + (1u << E.acc_log2), t.bin_start[b + 1]);
    const uint32_t brick0 = b << E.shift;
    const uint32_t span = min(1u << E.shift, E.nbricks - brick0) * E.lvoxc3;
    for (uint32_t j = threadIdx.x; j < span; j += kEvAccBlock) lq[j] = 0ull;
    // Largest |weight| of the bin (from the scatter) and whether any weight
    // of the bin is not finite.
    const uint32_t mb = t.bin_max[b];
    if (mb == 0u) return;   // all weights zero: nothing to add
    const bool bad = mb > 0x7f7fffffu;
    __syncthreads();
    float *out = E.grad + (size_t)brick0 * E.lvoxc3;
    if (bad) {
        float *lg = reinterpret_cast<float *>(lq);
        for (uint32_t j = threadIdx.x; j < span; j += kEvAccBlock) lg[j] = 0.0f;
        __syncthreads();
        for (uint32_t i = e0 + threadIdx.x; i < e1; i += kEvAccBlock) {
            const float4 wx = ev.swx[i];
            const uint32_t brick = ev.sb[i];
            float nl[3];
            const uint32_t cell = brick_cell(mk3(wx.y, wx.z, wx.w), E.bce[brick], E.lvoxc, nl);
            float *g = lg + (brick - brick0) * E.lvoxc3;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                uint32_t idx;
                float w;
                cell_corner(cell, nl, E.lvoxc, v, idx, w);
                atomicAdd(g + idx, wx.x * w);
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < span; j += kEvAccBlock) {
            const float v = lg[j];
            if (v != 0.0f) unsafeAtomicAdd(out + j, v);
        }
        return;
    }
    const int sh = 58 - (int)E.acc_log2 - ilogbf(__uint_as_float(mb));   // M * 2^sh < 2^(59 - acc_log2)
    for (uint32_t i = e0 + threadIdx.x; i < e1; i += kEvAccBlock) {
        const float4 wx = ev.swx[i];
        const uint32_t brick = ev.sb[i];
        float nl[3];
        const uint32_t cell = brick_cell(mk3(wx.y, wx.z, wx.w), E.bce[brick], E.lvoxc, nl);
        unsigned long long *g = lq + (brick - brick0) * E.lvoxc3;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            uint32_t idx;
            float w;
            cell_corner(cell, nl, E.lvoxc, v, idx, w);
            // |splat| * 2^sh < 2^45: scaled exactly in double and rounded to
            // the nearest integer (ties to even) by adding 1.5 * 2^52, whose
            // ulp is 1; the integer is then the low bits of the sum. Equal to
            // (long long)rintf(ldexpf(splat, sh)) in fewer instructions.
            const double r = ldexp((double)(wx.x * w), sh) + 6755399441055744.0;
            const long long q = __double_as_longlong(r) - 0x4338000000000000ll;
            atomicAdd(g + idx, (unsigned long long)q);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < span; j += kEvAccBlock) {
        const long long q = (long long)lq[j];
        if (q != 0) unsafeAtomicAdd(out + j, (float)ldexp((double)q, -sh));
    }
}

static uint32_t *bin_max_of(const EventReduce &E) { return E.bins + 2 + 4 * (size_t)E.nbins + 2; }

bool event_bins(uint32_t nbricks, uint32_t lvoxc3, uint32_t &shift, uint32_t &nbins) {
    shift = 0;
    while (((uint64_t)nbricks + (1ull << shift) - 1) >> shift > kEvMaxBins) ++shift;
    // Up to 8 bricks per bin while a bin still fits the accumulation LDS:
    // fewer, fuller bins let the scatter write the sorted copy in longer runs
    // (256^3 frame: 13.9 -> 11.1 ms), at a little more clearing and flushing per chunk.
    while (shift < 3 && ((uint64_t)lvoxc3 * sizeof(unsigned long long) << (shift + 1)) <= kEvAccumLds) ++shift;
    nbins = (uint32_t)(((uint64_t)nbricks + (1ull << shift) - 1) >> shift);
    return nbricks > 0 && ((uint64_t)lvoxc3 * sizeof(unsigned long long) << shift) <= kEvAccumLds;
}

hipError_t launch_event_reduce(const EventReduce &E, hipStream_t s) {
    if (E.ev_cap == 0 || E.nbins == 0) return hipSuccess;
    const uint32_t sort_blocks = (uint32_t)(((uint64_t)E.ev_cap + kEvSortChunk - 1) / kEvSortChunk);
    const uint32_t acc_blocks = (uint32_t)(((uint64_t)E.ev_cap + (1u << E.acc_log2) - 1) >> E.acc_log2) + E.nbins;
    k_fill_u32<<<(E.nbins + 255) / 256, 256, 0, s>>>(E.bins + 2, E.nbins, 0u);
    k_fill_u32<<<(E.nbins + 255) / 256, 256, 0, s>>>(bin_max_of(E), E.nbins, 0u);
    k_ev_hist<<<sort_blocks, kEvSortBlock, E.nbins * sizeof(uint32_t), s>>>(E);
    k_ev_scan<<<1, 1024, 0, s>>>(E);
    k_ev_scatter<<<sort_blocks, kEvSortBlock, ev_scatter_lds(E.nbins), s>>>(E);
    const size_t lds = (size_t)E.lvoxc3 * sizeof(unsigned long long) << E.shift;
    k_ev_accum<<<acc_blocks, kEvAccBlock, lds, s>>>(E);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fixed scratch for the allocation-free render entry point: the launch's
// counter block and parameters and its per-chunk partial sums live in this
// file-scope device array instead of in buffers the library allocates. Every GPU the
// library runs on holds its own copy (kScratchBytes) until the process ends.
// In host code the array's name is not a device address: scratch_address
// looks the address up once per device with hipGetSymbolAddress and the
// launches use that pointer.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kScratchDevices];

hipError_t scratch_address(void **ptr) {
    *ptr = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= (int)kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t nb = (n + 255) / 256;
    k_fill_f32<<<(uint32_t)nb, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace drk
