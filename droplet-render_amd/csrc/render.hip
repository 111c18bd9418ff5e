// gfx950 kernels of the volumetric light-transport path.
//
// k_render is a persistent path tracer: every lane owns one path at a time
// and runs a flat state machine (find next leaf -> one Woodcock step -> ray
// end / scattering event), so lanes never wait for each other at leaf, ray
// or sample boundaries. A lane that finishes its work unit (one pixel x a
// chunk of samples) refills from a unit counter with one wave-aggregated
// atomic (ballot + mbcnt prefix), which keeps the wave's lanes busy until the
// whole tile is drained. The tile's pixels are split into one row band per
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

#include "drk_common.h"
#include "estimator.h"
#include "render_launch.h"

namespace drk {

constexpr int kBlock = 256;

enum : uint32_t { ST_WORK = 0, ST_TRAV = 1, ST_LEAF = 2, ST_END = 3, ST_DONE = 4 };
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

// Wave-aggregated fetch of work units for the lanes that need one.
__device__ __forceinline__ uint32_t fetch_units(uint32_t *counter, bool need) {
    uint64_t mask = __ballot(need);
    uint32_t lane = lane_id();
    uint32_t base = 0;
    uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    return base + below;
}

template <int kMode>
__global__ __launch_bounds__(kBlock) void k_render(RenderParams P, WorkParams W) {
    constexpr bool kSumSq = (kMode == kModeForwardSq || kMode == kModeCount);
    constexpr bool kCount = (kMode == kModeCount);
    constexpr bool kAdj = (kMode == kModeAdjoint);
    constexpr bool kRec = (kMode == kModeRecord);
    __shared__ MieLds mie;
    __shared__ uint32_t stk_all[kMaxDepth * kBlock];
    load_mie(mie);
    __syncthreads();

    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t stride = kBlock;
    const KernelConsts &K = *P.kc;
    const float sig_s = P.sigma_s;
    const float sig_e = P.sigma_e;

    // Work unit.
    uint32_t st = ST_WORK;
    uint32_t chunk_cur = 0, pix_local = 0, px = 0, py = 0, key0 = 0;
    const uint32_t band0 = xcc_id();   // first band: this wave's XCD
    uint32_t band_step = 0;            // bands this wave has found drained (wave-uniform)
    uint32_t s_cur = 0, s_end = 0;
    float pix_depth = 3.402823466e38f;
    float accL[3] = {0, 0, 0}, accE[3] = {0, 0, 0}, accA = 0.0f;
    float sqL[3] = {0, 0, 0}, sqE[3] = {0, 0, 0};
    // Path.
    F3 beta = mk3(1, 1, 1), pathL = mk3(0, 0, 0), pathE = mk3(0, 0, 0);
    float alpha = 0.0f;
    uint32_t depth_r = 0, kind = RK_PRIMARY, event = 0;
    F3 cont_dir = mk3(0, 0, 0), cont_beta = mk3(0, 0, 0), shadow_w = mk3(0, 0, 0);
    // Ray.
    F3 ro = mk3(0, 0, 0), rd = mk3(0, 0, 1);
    bool hit = false;
    Traversal T;
    T.lvl = -1;
    RayRng rng;
    rng.s0 = rng.s1 = rng.s2 = rng.s3 = 1u;
    // Woodcock state of the current leaf.
    float4 ce = make_float4(0, 0, 0, 1);
    const float *brick = nullptr;
    size_t brick_off = 0;
    float smax = 1.0f, kmaj = 1.0f, lo = 0.0f, s0 = 0.0f, tr0 = 0.0f, tr1 = 0.0f, trb = 0.0f, sr = 0.0f, td = 0.0f;
    bool first_step = true;
    uint32_t cnt_nodes = 0, cnt_lookups = 0;
    // Adjoint state.
    F3 gL = mk3(0, 0, 0), gE = mk3(0, 0, 0);
    float gA = 0.0f, ltot = 0.0f, collected = 0.0f, wseg = 0.0f, seg_a = 0.0f;
    uint32_t pass = 0;
    uint64_t vis = 0;
    RayRng arng;
    arng.s0 = arng.s1 = arng.s2 = arng.s3 = 1u;
    // Gradient copy of this wave: spread over XCDs first, then over CUs.
    float *const grad_copy =
        kAdj ? W.grad + (size_t)((xcc_id() + W.grad_nxcc * cu_slot()) % W.grad_copies) * W.grad_stride : nullptr;

    auto start_ray = [&](uint32_t ray_id) {
        td = 0.0f;
        hit = false;
        T.lvl = -1;
        if (kAdj && pass == 1 && kind == RK_SHADOW && !((vis >> event) & 1ull)) {
            hit = true;   // blocked in pass 0: no contribution, no gradient, not re-traced
            st = ST_END;
            return;
        }
        rng.seed(key0, P.key1, P.sample_base + s_cur, ray_id, kRngTrack);
        if (kAdj && pass == 1) arng.seed(key0, P.key1, P.sample_base + s_cur, ray_id, kRngAdjoint);
        F3 t0, t1;
        uint32_t a;
        if (!octree_init(P.ob, ro, rd, t0, t1, a) || t1.x < 0.0f || t1.y < 0.0f || t1.z < 0.0f) {
            st = ST_END;   // misses the root: escapes
            return;
        }
        T.rt0 = t0; T.rt1 = t1; T.t0 = t0; T.t1 = t1; T.a = a;
        if (kCount) cnt_nodes += 1;
        if (is_leaf(P.ob[0])) {
            T.lvl = -2;    // root leaf pending
        } else {
            F3 tm = mk3(0.5f * (t0.x + t1.x), 0.5f * (t0.y + t1.y), 0.5f * (t0.z + t1.z));
            stk[0] = 0u | (first_node(t0, tm) << 24);
            T.lvl = 0;
        }
        st = ST_TRAV;
    };

    auto start_path = [&]() {
        bool gm;
        camera_ray(K, P.w, P.h, px, py, ro, rd, gm);
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
            start_ray(shadow_ray_id(0));
            return;
        }
        if (!gm) { st = ST_END; hit = false; return; }
        start_ray(0);
    };

    // Adjoint: set up the next sample of the current pixel. With a path
    // record, its total and shadow-escape bits come from the forward launch
    // and the replay starts directly at pass 1.
    auto adj_begin = [&]() {
        if (W.rec) {
            const size_t n = (size_t)W.npix * P.samples;
            const size_t ri = (size_t)s_cur * W.npix + pix_local;
            const float4 *R = reinterpret_cast<const float4 *>(W.rec);
            float4 a = R[ri], b = R[n + ri];
            uint2 v = reinterpret_cast<const uint2 *>(R + 2 * n)[ri];
            ltot = gL.x * a.x + gL.y * a.y + gL.z * a.z + gE.x * b.x + gE.y * b.y + gE.z * b.z + gA * a.w;
            vis = ((uint64_t)v.y << 32) | (uint64_t)v.x;
            pass = 1;
        } else {
            pass = 0;
            vis = 0;
        }
    };

    // Adds weight * (trilinear weight of vertex v at x) to every corner v of
    // the current leaf's gradient brick (adjoint pass 1 only).
    auto grad_at = [&](F3 x, float weight) {
        if (weight == 0.0f) return;
        uint32_t idx[8];
        float w8[8];
        brick_weights(x, ce, P.lvoxc, idx, w8);
        float *g = grad_copy + brick_off;
#pragma unroll
        for (int v = 0; v < 8; ++v) unsafeAtomicAdd(g + idx[v], weight * w8[v]);
    };
    // -W * sigma_e * integral over [a, b] of the trilinear weights, by one
    // uniform point of the adjoint stream (drawn for every covered leaf with a
    // non-zero majorant, so the stream stays aligned with the oracle).
    auto grad_segment = [&](float a, float b) {
        float u = arng.uniform();
        if (!(b > a)) return;
        float t = a + u * (b - a);
        F3 x = mk3(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
        grad_at(x, -wseg * sig_e * (b - a));
    };

    for (;;) {
        // ---- refill lanes that finished their unit ------------------------
        // The pixels are split into W.nbands contiguous row bands; a wave
        // drains the band of its own XCD first, then the others in turn.
        bool need = (st == ST_WORK);
        while (__ballot(need)) {
            if (band_step >= W.nbands) {
                if (need) st = ST_DONE;   // every band drained
                break;
            }
            const uint32_t b = (band0 + band_step) % W.nbands;
            const uint32_t p0 = (uint32_t)((uint64_t)b * W.npix / W.nbands);
            const uint32_t bp = (uint32_t)((uint64_t)(b + 1) * W.npix / W.nbands) - p0;
            const uint32_t u = fetch_units(W.counter + kCounterStride * b, need);
            if (need && u < bp * W.nchunks) {
                chunk_cur = u / bp;
                pix_local = p0 + (u - chunk_cur * bp);
                uint32_t lx = pix_local % P.rx, ly = pix_local / P.rx;
                px = P.x0 + lx;
                py = P.y0 + ly;
                key0 = py * P.w + px;
                s_cur = chunk_cur * W.chunk;
                s_end = min(s_cur + W.chunk, P.samples);
                pix_depth = P.depth ? P.depth[key0] : 3.402823466e38f;
                accL[0] = accL[1] = accL[2] = 0.0f;
                accE[0] = accE[1] = accE[2] = 0.0f;
                accA = 0.0f;
                sqL[0] = sqL[1] = sqL[2] = 0.0f;
                sqE[0] = sqE[1] = sqE[2] = 0.0f;
                if (kAdj) {
                    size_t d = (size_t)ly * P.tilex + lx;
                    float4 a = reinterpret_cast<const float4 *>(W.adj_l)[d];
                    float4 c = reinterpret_cast<const float4 *>(W.adj_e)[d];
                    gL = mk3(a.x, a.y, a.z);
                    gE = mk3(c.x, c.y, c.z);
                    gA = a.w + c.w;   // alpha appears in both buffers
                    if (gL.x == 0.0f && gL.y == 0.0f && gL.z == 0.0f && gE.x == 0.0f && gE.y == 0.0f &&
                        gE.z == 0.0f && gA == 0.0f)
                        s_cur = s_end;   // no adjoint on this pixel: nothing to differentiate
                    else
                        adj_begin();
                }
                if (s_cur < s_end) {
                    start_path();
                    need = false;
                }
            }
            // Lanes that drew past the band's end leave it for the next one.
            if (__ballot(need && !(u < bp * W.nchunks))) ++band_step;
        }
        if (!__ballot(st != ST_DONE)) break;

        // ---- A: find the next leaf along the ray --------------------------
        if (st == ST_TRAV) {
            uint32_t leaf = 0;
            float tra = 0.0f, trbn = 0.0f;
            bool found;
            if (T.lvl == -2) {
                leaf = 0;
                tra = fmaxf(fmaxf(T.rt0.x, T.rt0.y), T.rt0.z);
                trbn = fminf(fminf(T.rt1.x, T.rt1.y), T.rt1.z);
                T.lvl = -1;
                found = true;
            } else {
                found = trav_next(P.chx, T, stk, stride, leaf, tra, trbn, cnt_nodes);
            }
            // Compositing depth stops primary rays (src/kernel.cpp:217-222).
            if (found && kind == RK_PRIMARY && !(td < pix_depth)) found = false;
            if (!found) {
                st = ST_END;
                hit = false;
            } else {
                const drk_octree_node &n = P.ob[leaf];
                ce = *reinterpret_cast<const float4 *>(n.ce);
                smax = n.qval[DRK_VOLUME_FOG];
                brick_off = (size_t)n.volx[DRK_VOLUME_FOG] * P.lvoxc3;
                brick = P.pvol_fog + brick_off;
                if (kCount) cnt_nodes += 1;
                // Fog-only leaf: always skip to the leaf entry (src/kernel.cpp:233-269).
                lo = td;
                tr0 = tra - td;
                tr1 = trbn - td;
                trb = trbn;
                td = tra;
                seg_a = tra;
                s0 = tr0;
                kmaj = sig_e * smax;
                first_step = true;
                st = ST_LEAF;
            }
        }

        // ---- B: one Woodcock step inside the current leaf -----------------
        if (st == ST_LEAF) {
            float step = logf(rng.uniform()) / kmaj;
            sr = first_step ? -step : sr - step;
            first_step = false;
            float sc = s0 + sr;
            td = lo + sc;
            if (!(sc < tr1)) {
                td = trb;     // left the leaf: memoryless restart in the next one
                st = ST_TRAV;
                if (kAdj && pass == 1 && kmaj > 0.0f) grad_segment(seg_a, trb);
            } else if (!(sc > tr0)) {
                hit = true;   // collision before the leaf entry counts as a real one (src/kernel.cpp:287)
                st = ST_END;
            } else {
                F3 rc = mk3(ro.x + rd.x * td, ro.y + rd.y * td, ro.z + rd.z * td);
                float p = brick_lookup(brick, rc, ce, P.lvoxc);
                if (kCount) cnt_lookups += 1;
                float q = rng.uniform();
                if (!(p / smax < q)) {
                    hit = true;
                    st = ST_END;
                    if (kAdj && pass == 1) {
                        grad_segment(seg_a, td);
                        grad_at(rc, wseg / p);   // d sigma / sigma at the real collision
                    }
                }
            }
        }

        // ---- C: ray finished (escaped or collided) -------------------------
        if (st == ST_END) {
            bool path_done = false;
            if (kind == RK_SHADOW_PASS) {
                // SampleVolume at r = scattevs returns the lights' radiance when
                // the ray escapes (r > 0); Shadow divides by lights[0]'s colour.
                if (!hit && P.scattevs > 0) {
                    F3 le = lights_eval(K, rd);
                    pathL = mk3(le.x / K.sun_color[0][0], le.y / K.sun_color[0][1], le.z / K.sun_color[0][2]);
                }
                alpha = 1.0f;
                path_done = true;
            } else if (kind == RK_SHADOW) {
                if (!hit) {
                    F3 le = lights_eval(K, rd);
                    pathL.x += shadow_w.x * le.x;
                    pathL.y += shadow_w.y * le.y;
                    pathL.z += shadow_w.z * le.z;
                    if ((kAdj && pass == 0) || kRec) vis |= 1ull << event;
                }
                // Continue the path along the phase-sampled direction.
                rd = cont_dir;
                beta = cont_beta;
                depth_r += 1;
                kind = RK_CONT;
                if (kAdj && pass == 1) wseg = ltot - collected;
                start_ray(cont_ray_id(event));
            } else if (!hit) {
                if (depth_r > 0) {
                    F3 lc = lights_eval(K, rd);
                    F3 le = env_radiance(P.env, K, rd);
                    pathL.x += beta.x * lc.x; pathL.y += beta.y * lc.y; pathL.z += beta.z * lc.z;
                    pathE.x += beta.x * le.x; pathE.y += beta.y * le.y; pathE.z += beta.z * le.z;
                }
                path_done = true;
            } else {
                if (depth_r == 0) {
                    alpha = 1.0f;
                    if (kAdj && pass == 1) collected += gA;
                }
                if (depth_r >= P.scattevs) {
                    path_done = true;
                } else {
                    // Scattering event: MIS between the phase function and lights[0]
                    // (src/kernel.cpp:387-430).
                    event = depth_r;
                    U4 r4 = philox4x32_10(U4{P.sample_base + s_cur, event, 0u, kRngScatter}, key0, P.key1);
                    float u1 = u01(r4.x), u2 = u01(r4.y), u3 = u01(r4.z), u4 = u01(r4.w);
                    F3 srd, p1, p3;
                    F3 lrd = sun_sample(K, u3, u4);
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
                    F3 w1 = mk3(p1.x / (p1.x + pd1), p1.y / (p1.y + pd1), p1.z / (p1.z + pd1));
                    F3 w2 = mk3(p3.x / (p3.x + pd2), p3.y / (p3.y + pd2), p3.z / (p3.z + pd2));
                    F3 x = mk3(ro.x + rd.x * td, ro.y + rd.y * td, ro.z + rd.z * td);
                    shadow_w = mk3(beta.x * w2.x * sig_s / sig_e,
                                   beta.y * w2.y * sig_s / sig_e,
                                   beta.z * w2.z * sig_s / sig_e);
                    cont_beta = mk3(beta.x * w1.x * sig_s / sig_e,
                                    beta.y * w1.y * sig_s / sig_e,
                                    beta.z * w1.z * sig_s / sig_e);
                    cont_dir = srd;
                    ro = x;
                    rd = lrd;
                    kind = RK_SHADOW;
                    if (kAdj && pass == 1) {
                        // The shadow ray's segment carries only its own contribution.
                        wseg = 0.0f;
                        if ((vis >> event) & 1ull) {
                            F3 le = lights_eval(K, lrd);
                            wseg = gL.x * (shadow_w.x * le.x) + gL.y * (shadow_w.y * le.y) + gL.z * (shadow_w.z * le.z);
                            collected += wseg;
                        }
                    }
                    start_ray(shadow_ray_id(event));
                }
            }
            if (path_done && kAdj) {
                if (pass == 0) {
                    ltot = gL.x * pathL.x + gL.y * pathL.y + gL.z * pathL.z + gE.x * pathE.x + gE.y * pathE.y +
                           gE.z * pathE.z + gA * alpha;
                    // Replay the same path for the gradients. (A zero total does not
                    // mean zero sub-path weights when adjoints have mixed signs.)
                    pass = 1;
                    start_path();
                    path_done = false;
                }
                if (path_done) {
                    ++s_cur;
                    if (s_cur < s_end) {
                        adj_begin();
                        start_path();
                    } else {
                        st = ST_WORK;
                    }
                }
            } else if (path_done) {
                if (kRec) {
                    const size_t n = (size_t)W.npix * P.samples;
                    const size_t ri = (size_t)s_cur * W.npix + pix_local;
                    float4 *R = reinterpret_cast<float4 *>(W.rec);
                    R[ri] = make_float4(pathL.x, pathL.y, pathL.z, alpha);
                    R[n + ri] = make_float4(pathE.x, pathE.y, pathE.z, 0.0f);
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
                    start_path();
                } else {
                    // Unit complete: write its sums.
                    size_t o = (size_t)chunk_cur * W.npix + pix_local;
                    float4 *oL = W.partial ? reinterpret_cast<float4 *>(W.partial) : nullptr;
                    if (W.partial) {
                        size_t plane = (size_t)W.nchunks * W.npix;
                        oL[o] = make_float4(accL[0], accL[1], accL[2], accA);
                        oL[plane + o] = make_float4(accE[0], accE[1], accE[2], accA);
                        if (kSumSq) {
                            oL[2 * plane + o] = make_float4(sqL[0], sqL[1], sqL[2], accA);
                            oL[3 * plane + o] = make_float4(sqE[0], sqE[1], sqE[2], accA);
                        }
                    } else {
                        uint32_t lx = pix_local % P.rx, ly = pix_local / P.rx;
                        size_t d = (size_t)ly * P.tilex + lx;
                        reinterpret_cast<float4 *>(W.out_l)[d] = make_float4(accL[0], accL[1], accL[2], accA);
                        reinterpret_cast<float4 *>(W.out_e)[d] = make_float4(accE[0], accE[1], accE[2], accA);
                        if (kSumSq) {
                            reinterpret_cast<float4 *>(W.sq_l)[d] = make_float4(sqL[0], sqL[1], sqL[2], accA);
                            reinterpret_cast<float4 *>(W.sq_e)[d] = make_float4(sqE[0], sqE[1], sqE[2], accA);
                        }
                    }
                    st = ST_WORK;
                }
            }
        }
    }
    if (kCount) {
        atomicAdd(reinterpret_cast<unsigned long long *>(W.counts) + 0, (unsigned long long)cnt_nodes);
        atomicAdd(reinterpret_cast<unsigned long long *>(W.counts) + 1, (unsigned long long)cnt_lookups);
    }
}

// Sums the per-chunk partials of each pixel in chunk order (deterministic).
__global__ void k_reduce(WorkParams W, uint32_t rx, uint32_t tilex, uint32_t buffers) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W.npix) return;
    const float4 *part = reinterpret_cast<const float4 *>(W.partial);
    size_t plane = (size_t)W.nchunks * W.npix;
    float *outs[4] = {W.out_l, W.out_e, W.sq_l, W.sq_e};
    uint32_t lx = i % rx, ly = i / rx;
    size_t d = (size_t)ly * tilex + lx;
    for (uint32_t b = 0; b < buffers; ++b) {
        float4 s = make_float4(0, 0, 0, 0);
        for (uint32_t c = 0; c < W.nchunks; ++c) {
            float4 v = part[b * plane + (size_t)c * W.npix + i];
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
        reinterpret_cast<float4 *>(outs[b])[d] = s;
    }
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
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kModeAdjoint>, kBlock, 0); break;
    }
    if (e != hipSuccess || nb <= 0) return 1;
    return nb;
}

hipError_t launch_render(const RenderParams &P, const WorkParams &W, uint32_t grid, int mode, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1) {
    k_fill_u32<<<1, 64 * kCounterStride / 4, 0, s>>>(W.counter, W.nbands * kCounterStride, 0u);
    if (ev0) (void)hipEventRecord(ev0, s);
    switch (mode) {
    case kModeForward: k_render<kModeForward><<<grid, kBlock, 0, s>>>(P, W); break;
    case kModeForwardSq: k_render<kModeForwardSq><<<grid, kBlock, 0, s>>>(P, W); break;
    case kModeCount: k_render<kModeCount><<<grid, kBlock, 0, s>>>(P, W); break;
    case kModeRecord: k_render<kModeRecord><<<grid, kBlock, 0, s>>>(P, W); break;
    default: k_render<kModeAdjoint><<<grid, kBlock, 0, s>>>(P, W); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev1) (void)hipEventRecord(ev1, s);
    if (W.partial && mode != kModeAdjoint) {
        uint32_t nb = (W.npix + 255) / 256;
        bool sq = (mode == kModeForwardSq || mode == kModeCount);
        k_reduce<<<nb, 256, 0, s>>>(W, P.rx, P.tilex, sq ? 4u : 2u);
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

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t nb = (n + 255) / 256;
    k_fill_f32<<<(uint32_t)nb, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace drk
