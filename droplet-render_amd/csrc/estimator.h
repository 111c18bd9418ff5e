// Device-side pieces of droplet-render's volumetric estimator, restated for
// gfx950: camera rays, Revelles octree traversal, trilinear brick lookup,
// Mie / Henyey-Greenstein phase, sun cone, Hosek sky and equirect env.
// Each function cites the reference code whose behaviour it reproduces.
#pragma once

#include "drk_common.h"

namespace drk {

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 mk3(float x, float y, float z) { return {x, y, z}; }

// Mie tables (data, see data/mie_*.inc), one copy per translation unit.
static __constant__ float c_mie_pdf[kMieBins * 3] = {
#include "../data/mie_pdf_rgb.inc"
};
static __constant__ float c_mie_cdf_rgb[kMieBins * 3] = {
#include "../data/mie_cdf_rgb.inc"
};

// Tables staged in LDS once per workgroup.
constexpr uint32_t kMieLutBuckets = 1024;   // u1 buckets of mie_sample's search bounds

struct MieLds {
    float pdf[kMieBins * 3];   // miepdf RGB, KernelSampler.cpp:8-73
    float cdf[kMieBins];       // miecdf R channel, the only one Sample reads (KernelSampler.cpp:262-263)
    uint16_t lut[kMieLutBuckets];   // per u1 bucket: bounds lo | hi << 8 of mie_sample's search
};

__device__ __forceinline__ uint32_t mie_search(const MieLds &T, float u1, uint32_t lo, uint32_t hi);

// Stages the tables and the search bounds; every thread of the block calls
// it (it synchronises the block once inside), and the caller synchronises
// again before use.
__device__ __forceinline__ void load_mie(MieLds &m) {
    for (uint32_t i = threadIdx.x; i < kMieBins * 3; i += blockDim.x) m.pdf[i] = c_mie_pdf[i];
    for (uint32_t i = threadIdx.x; i < kMieBins; i += blockDim.x) m.cdf[i] = c_mie_cdf_rgb[i * 3];
    __syncthreads();
    // mie_sample's answer grows with u1 except for two steps back by one
    // (the keys themselves move with u1), so the answers at the smallest and
    // the largest uniform of a bucket, widened by one each way, bound the
    // answer for all of it: checked over every uniform the generator can
    // return (odd multiples of 2^-24). Bucket b holds the uniforms in
    // [b, b + 1) / kMieLutBuckets.
    constexpr uint32_t kPer = (1u << 24) / kMieLutBuckets;
    for (uint32_t b = threadIdx.x; b < kMieLutBuckets; b += blockDim.x) {
        const float u_lo = (float)(b * kPer + 1u) * 5.9604644775390625e-8f;
        const float u_hi = (float)((b + 1u) * kPer - 1u) * 5.9604644775390625e-8f;
        const uint32_t lo = mie_search(m, u_lo, 0u, 255u), hi = mie_search(m, u_hi, 0u, 255u);
        m.lut[b] = (uint16_t)((lo > 0u ? lo - 1u : 0u) | (min(hi + 1u, 255u) << 8));
    }
}

// ---------------------------------------------------------------------------
// xoshiro128** (Blackman & Vigna) per-ray stream, seeded from Philox so that
// every (pixel, sample, ray) has a replayable sequence of uniforms.
// ---------------------------------------------------------------------------
struct RayRng {
    uint32_t s0, s1, s2, s3;
    __device__ __forceinline__ void seed(uint32_t key0, uint32_t key1, uint32_t sample, uint32_t ray,
                                         uint32_t purpose = kRngTrack) {
        U4 v = philox4x32_10(U4{sample, ray, 0u, purpose}, key0, key1);
        s0 = v.x; s1 = v.y; s2 = v.z; s3 = v.w | ((v.x | v.y | v.z) == 0u ? 1u : 0u);
    }
    __device__ __forceinline__ uint32_t next() {
        uint32_t r = s1 * 5u;
        r = ((r << 7) | (r >> 25)) * 9u;
        uint32_t t = s1 << 9;
        s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t;
        s3 = (s3 << 11) | (s3 >> 21);
        return r;
    }
    __device__ __forceinline__ float uniform() { return u01(next()); }
};

// logf(u) for the tracker's uniforms u in [2^-24, 1): the device library's
// logf for normal, finite arguments -- the hardware log2 (v_log_f32) scaled
// by ln 2 split into a head and a tail with one fma each -- without its
// denormal rescaling and non-finite selects, which these arguments never
// take (a GPU test checks every possible uniform against logf bit for bit).
__device__ __forceinline__ float log_unit(float u) {
    const float r = __builtin_amdgcn_logf(u);
    const float y = r * 0x1.62e42ep-1f;
    const float t = __builtin_fmaf(r, 0x1.62e42ep-1f, -y);
    return y + __builtin_fmaf(r, 0x1.efa39ep-25f, t);
}

// ---------------------------------------------------------------------------
// Primary ray, K_ParallelRender (src/kernel.cpp:461-474): NDC with the
// reference's sign flips, ro = viewi.r[3], rd = normalize((posh*proji).xyz0 * viewi).
// Row-vector convention: v*M = sum_k v[k]*M.r[k].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void camera_ray(const KernelConsts &K, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                                           F3 &ro, F3 &rd, bool &gm) {
#pragma clang fp contract(off)
    float hx = -(2.0f * ((float)px + 0.5f) / (float)w - 1.0f);
    float hy = -(2.0f * ((float)py + 0.5f) / (float)h - 1.0f);
    const float *pi = K.proji;
    float rv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) rv[j] = hx * pi[0 * 4 + j] + hy * pi[1 * 4 + j] + 0.0f * pi[2 * 4 + j] + 1.0f * pi[3 * 4 + j];
    const float *vi = K.viewi;
    float d[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) d[j] = rv[0] * vi[0 * 4 + j] + rv[1] * vi[1 * 4 + j] + rv[2] * vi[2 * 4 + j] + 0.0f * vi[3 * 4 + j];
    float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    rd = mk3(d[0] / len, d[1] / len, d[2] / len);
    ro = mk3(vi[12], vi[13], vi[14]);
    gm = hx > -1.0f && hx < 1.0f && hy > -1.0f && hy < 1.0f;
}

// ---------------------------------------------------------------------------
// Revelles et al. parametric octree traversal (KernelOctree.cpp:9-110 and the
// step traverser :222-349). Per-lane stack lives in LDS, one 32-bit word per
// level: node index (bits 0-23), next child tn (bits 24-27), child being
// visited c (bits 28-30). The t-intervals of a level are recomputed from the
// root interval along the stored child path with the reference's exact
// operations (tm = 0.5f*(t0+t1)), so they are bit-identical to the values the
// reference keeps on its stack.
// ---------------------------------------------------------------------------
// First child the ray enters (OctreeFirstNode, KernelOctree.cpp:9-34): the
// entry plane is the largest of t0 (ties resolved as the reference's nested
// comparisons do); a bit is set for each other axis whose midpoint is
// crossed before the entry. Written with selects, not branches: the lanes of
// a wave then run one path instead of each branch in turn.
__device__ __forceinline__ uint32_t first_node(F3 t0, F3 tm) {
    const bool gxy = t0.x > t0.y;
    const bool ex = gxy && t0.x > t0.z;          // entry plane x
    const bool ey = !gxy && t0.y > t0.z;         // entry plane y (else z)
    const float m = ex ? t0.x : (ey ? t0.y : t0.z);
    return (!ex && tm.x < m ? 4u : 0u) | (!ey && tm.y < m ? 2u : 0u) | ((ex || ey) && tm.z < m ? 1u : 0u);
}

// Exit plane of child c whose far corner is t1: x -> c|4, y -> c|2, z -> c|1,
// or 8 when that bit is already set (the successor table at KernelOctree.cpp:165-196,
// with OctreeNextNode's tie order :62-71).
__device__ __forceinline__ uint32_t next_node(F3 t1, uint32_t c) {
    const uint32_t nx = (c & 4) ? 8u : (c | 4), ny = (c & 2) ? 8u : (c | 2), nz = (c & 1) ? 8u : (c | 1);
    const bool lxy = t1.x < t1.y, lxz = t1.x < t1.z, lyz = t1.y < t1.z;
    return lxy ? (lxz ? nx : nz) : (lyz ? ny : nz);
}

__device__ __forceinline__ void child_interval(F3 t0, F3 t1, uint32_t c, F3 &ct0, F3 &ct1) {
#pragma clang fp contract(off)
    F3 tm = mk3(0.5f * (t0.x + t1.x), 0.5f * (t0.y + t1.y), 0.5f * (t0.z + t1.z));
    ct0.x = (c & 4) ? tm.x : t0.x; ct1.x = (c & 4) ? t1.x : tm.x;
    ct0.y = (c & 2) ? tm.y : t0.y; ct1.y = (c & 2) ? t1.y : tm.y;
    ct0.z = (c & 1) ? tm.z : t0.z; ct1.z = (c & 1) ? t1.z : tm.z;
}

__device__ __forceinline__ bool is_leaf(const drk_octree_node &n) {
    return n.volx[DRK_VOLUME_SDF] != DRK_NO_BRICK || n.volx[DRK_VOLUME_FOG] != DRK_NO_BRICK;
}

// OctreeInitialize (KernelOctree.cpp:78-110): mirror to a positive direction
// in the root-local frame; returns false when the ray misses the root box
// (the max(t0) < min(t1) test of :137/:238).
__device__ __forceinline__ bool octree_init(const drk_octree_node *ob, F3 ro, F3 rd, F3 &t0, F3 &t1, uint32_t &a) {
#pragma clang fp contract(off)
    float4 ce = *reinterpret_cast<const float4 *>(ob[0].ce);
    float e = ce.w;
    F3 o = mk3((ro.x - ce.x) + e, (ro.y - ce.y) + e, (ro.z - ce.z) + e);
    F3 d = rd;
    a = 0;
    if (d.x < 0.0f) { o.x = 2.0f * e - o.x; d.x = -d.x; a |= 4; }
    if (d.y < 0.0f) { o.y = 2.0f * e - o.y; d.y = -d.y; a |= 2; }
    if (d.z < 0.0f) { o.z = 2.0f * e - o.z; d.z = -d.z; a |= 1; }
    float e2 = 2.0f * e;
    t0 = mk3((0.0f - o.x) / d.x, (0.0f - o.y) / d.y, (0.0f - o.z) / d.z);
    t1 = mk3((e2 - o.x) / d.x, (e2 - o.y) / d.y, (e2 - o.z) / d.z);
    return fmaxf(fmaxf(t0.x, t0.y), t0.z) < fminf(fminf(t1.x, t1.y), t1.z);
}

// Child order slot map nla = {0,4,2,6,1,5,3,7} (KernelOctree.cpp:164): Revelles
// index (x=4,y=2,z=1) -> build slot (x=1,y=2,z=4), a 3-bit reversal.
__device__ __forceinline__ uint32_t build_slot(uint32_t r) { return ((r & 4) >> 2) | (r & 2) | ((r & 1) << 2); }

// The root interval lives next to the stack in LDS (six floats, `stride`
// apart), not in registers: it is read only when a pop recomputes an
// interval from the root.
struct Traversal {
    F3 t0, t1;        // interval of the node on top of the stack
    uint32_t a;       // mirror flags
    int lvl;          // top of stack, -1 = exhausted
};

constexpr uint32_t kStkNodeMask = 0x00ffffffu;

enum : int { kTravExhausted = 0, kTravLeaf = 1, kTravPending = 2 };

// Advances the walk to the next leaf in ray order (kTravLeaf: its index and
// [tra, trb] interval), to its end (kTravExhausted), or by at most `budget`
// node steps (kTravPending: call again to resume where it stopped).
// `touches` counts node-record reads (for the algorithmic-bytes model).
__device__ __forceinline__ void trav_set_root(float *root, uint32_t stride, F3 t0, F3 t1) {
    root[0] = t0.x; root[stride] = t0.y; root[2 * stride] = t0.z;
    root[3 * stride] = t1.x; root[4 * stride] = t1.y; root[5 * stride] = t1.z;
}
__device__ __forceinline__ void trav_root(const float *root, uint32_t stride, F3 &t0, F3 &t1) {
    t0 = mk3(root[0], root[stride], root[2 * stride]);
    t1 = mk3(root[3 * stride], root[4 * stride], root[5 * stride]);
}

__device__ __forceinline__ int trav_next(const uint32_t *__restrict__ chx, Traversal &T, uint32_t *stk,
                                         const float *root, uint32_t stride, uint32_t &leaf, float &tra, float &trb,
                                         uint32_t &touches, uint32_t budget) {
    for (uint32_t steps = 0;; ++steps) {
        if (steps >= budget) return kTravPending;
        if (T.lvl < 0) return kTravExhausted;
        uint32_t f = stk[T.lvl * stride];
        uint32_t tn = (f >> 24) & 15u;
        if (tn >= 8) {
            // Pop every exhausted level, then recompute the new top's interval
            // from the root along the path and visit its next child in this
            // same step.
            do {
                if (--T.lvl < 0) return kTravExhausted;
                f = stk[T.lvl * stride];
                tn = (f >> 24) & 15u;
            } while (tn >= 8);
            F3 a0, a1;
            trav_root(root, stride, a0, a1);
#pragma nounroll   // usually one or two levels; unrolled copies only cost issue slots
            for (int k = 0; k < T.lvl; ++k) {
                uint32_t ck = (stk[k * stride] >> 28) & 7u;
                F3 b0, b1;
                child_interval(a0, a1, ck, b0, b1);
                a0 = b0; a1 = b1;
            }
            T.t0 = a0; T.t1 = a1;
        }
        uint32_t node = f & kStkNodeMask;
        F3 ct0, ct1;
        child_interval(T.t0, T.t1, tn, ct0, ct1);
        uint32_t nxt = (tn == 7) ? 8u : next_node(ct1, tn);
        stk[T.lvl * stride] = node | (nxt << 24) | (tn << 28);
        const uint32_t link = chx[node * 8u + build_slot(tn ^ T.a)];
        const uint32_t ch = link & kStkNodeMask;
        touches += 1;
        if (ct1.x < 0.0f || ct1.y < 0.0f || ct1.z < 0.0f || ch == 0) continue;
        touches += 1;
        if (link & kChildLeaf) {
            leaf = ch;
            tra = fmaxf(fmaxf(ct0.x, ct0.y), ct0.z);
            trb = fminf(fminf(ct1.x, ct1.y), ct1.z);
            return kTravLeaf;
        }
        // Interior child: push.
        F3 tm;
        {
#pragma clang fp contract(off)
            tm = mk3(0.5f * (ct0.x + ct1.x), 0.5f * (ct0.y + ct1.y), 0.5f * (ct0.z + ct1.z));
        }
        ++T.lvl;
        stk[T.lvl * stride] = ch | (first_node(ct0, tm) << 24);
        T.t0 = ct0; T.t1 = ct1;
    }
}

// ---------------------------------------------------------------------------
// Leaf-grid walk. The reference's build puts every leaf at the same depth L
// (Octree::BuildPath, src/scene.cpp:273), so the leaves are the occupied
// cells of a 2^L-per-axis grid over the root cube. The t-interval of a cell
// along one axis is a pair of boundaries of the binary subdivision of the
// root's [t0, t1] on that axis, B[k], each the midpoint 0.5f*(lo+hi) the
// Revelles walk computes at the level where k is first a midpoint; so a
// cell's [max entries, min exits] is bit-identical to the interval the
// Revelles walk gives its leaf. Stepping from cell to cell across the exit
// plane chosen with next_node's comparisons visits the leaves in the
// Revelles order, with one step per cell instead of a climb and descent
// through the tree per leaf (empty cells are stepped over). Two situations
// are handed to the Revelles walk, so that the sequence stays exact:
// non-finite root intervals (a zero direction component) or a start point
// on a cell boundary (ties in the descent), and a step where the new entry
// time equals another axis's entry time (the Revelles descent into the next
// subtree may then pick a lower cell). The hand-over rebuilds the Revelles
// stack for the current cell from the root (grid_to_stack).
//
// State while walking the grid: T.lvl = kTravGrid, T.t0/T.t1 = the cell's
// entry and exit times per axis (mirrored frame), stk[0] = the cell index
// per axis (8 bits each) and bit 31 set while the start cell is unvisited.
// ---------------------------------------------------------------------------
constexpr int kTravGrid = -3;
constexpr uint32_t kGridPending = 0x80000000u;

// B[k] of the subdivision of [t0, t1] into 2^L cells (0 < k < 2^L).
__device__ __forceinline__ float grid_boundary(float t0, float t1, uint32_t L, uint32_t k) {
#pragma clang fp contract(off)
    float lo = t0, hi = t1, res = t1;
    uint32_t base = 0;
    for (uint32_t span = (1u << L) >> 1; span; span >>= 1) {
        const float mid = 0.5f * (lo + hi);
        const uint32_t mi = base + span;
        res = k == mi ? mid : res;
        const bool up = k >= mi;
        lo = up ? mid : lo;
        hi = up ? hi : mid;
        base = up ? mi : base;
    }
    return k == 0 ? t0 : res;
}

// Start of a ray's grid walk: the cell the Revelles walk reaches first (the
// one holding the root entry point, or the ray origin when that lies inside
// the root). False when the walk must be the Revelles one (see above).
__device__ __forceinline__ bool grid_start(F3 t0, F3 t1, uint32_t L, Traversal &T, uint32_t *stk) {
#pragma clang fp contract(off)
    if (!(isfinite(t0.x) && isfinite(t0.y) && isfinite(t0.z) && isfinite(t1.x) && isfinite(t1.y) && isfinite(t1.z)))
        return false;
    const float m = fmaxf(fmaxf(t0.x, t0.y), t0.z);
    const float s = fmaxf(m, 0.0f);
    const float a0[3] = {t0.x, t0.y, t0.z}, a1[3] = {t1.x, t1.y, t1.z};
    float E[3], X[3];
    uint32_t packed = kGridPending;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = a0[k], hi = a1[k];
        uint32_t id = 0;
        for (uint32_t span = (1u << L) >> 1; span; span >>= 1) {
            const float mid = 0.5f * (lo + hi);
            const bool up = mid < s;
            lo = up ? mid : lo;
            hi = up ? hi : mid;
            id += up ? span : 0u;
        }
        E[k] = lo;
        X[k] = hi;
        packed |= id << (8 * k);
        // the start point on a boundary of this axis (the root entry plane excepted)
        ok = ok && !(X[k] == s) && !(E[k] == s && !(m >= 0.0f && a0[k] == m));
    }
    if (!ok) return false;
    T.t0 = mk3(E[0], E[1], E[2]);
    T.t1 = mk3(X[0], X[1], X[2]);
    T.lvl = kTravGrid;
    stk[0] = packed;
    return true;
}

// Rebuilds the Revelles stack for the walk standing at grid cell idx (the
// state after visiting it, or after visiting the missing subtree that
// contains it), so that trav_next continues from there.
__device__ __forceinline__ void grid_to_stack(const uint32_t *__restrict__ chx, Traversal &T, uint32_t *stk,
                                              const float *root, uint32_t stride, uint32_t L, uint32_t packed) {
    F3 t0, t1;
    trav_root(root, stride, t0, t1);
    uint32_t node = 0;
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t sh = L - 1 - l;
        const uint32_t c = (((packed >> sh) & 1u) << 2) | (((packed >> (8 + sh)) & 1u) << 1) | ((packed >> (16 + sh)) & 1u);
        F3 ct0, ct1;
        child_interval(t0, t1, c, ct0, ct1);
        const uint32_t nxt = (c == 7) ? 8u : next_node(ct1, c);
        stk[l * stride] = node | (nxt << 24) | (c << 28);
        T.lvl = (int)l;
        T.t0 = t0;
        T.t1 = t1;
        const uint32_t link = chx[node * 8u + build_slot(c ^ T.a)];
        const uint32_t ch = link & kStkNodeMask;
        if (ch == 0 || l + 1 == L) return;
        node = ch;
        t0 = ct0;
        t1 = ct1;
    }
}

// One cell of the grid walk: kTravLeaf (its leaf and [tra, trb]),
// kTravExhausted, kTravPending (an empty cell was stepped over), or
// kTravRevelles (a tie: the Revelles stack has been rebuilt for the current
// position and the walk continues in trav_next). One exit per outcome and no
// loop, so that the walk's state is not shuffled between registers at the
// merges of an inlined loop.
constexpr int kTravRevelles = 3;
__device__ __forceinline__ int grid_step(const uint32_t *__restrict__ cells, uint32_t L,
                                         const uint32_t *__restrict__ chx, Traversal &T, uint32_t *stk,
                                         const float *root, uint32_t stride, uint32_t &leaf, float &tra, float &trb,
                                         uint32_t &touches) {
    const uint32_t n = 1u << L;
    uint32_t packed = stk[0];
    if (packed & kGridPending) {
        packed &= ~kGridPending;
    } else {
        // next_node's choice of the exit plane (ties as the reference breaks them)
        const bool lxy = T.t1.x < T.t1.y, lxz = T.t1.x < T.t1.z, lyz = T.t1.y < T.t1.z;
        const uint32_t ax = lxy ? (lxz ? 0u : 2u) : (lyz ? 1u : 2u);
        const float m = ax == 0 ? T.t1.x : (ax == 1 ? T.t1.y : T.t1.z);
        const bool tie = (ax != 0 && T.t0.x >= m) || (ax != 1 && T.t0.y >= m) || (ax != 2 && T.t0.z >= m);
        if (tie) {
            grid_to_stack(chx, T, stk, root, stride, L, packed);
            return kTravRevelles;
        }
        const uint32_t id = (packed >> (8 * ax)) & 0xffu;
        if (id + 1 == n) {
            T.lvl = -1;
            return kTravExhausted;
        }
        packed += 1u << (8 * ax);
        // the root interval of that axis (see trav_set_root's layout)
        const float lo = root[ax * stride], hi = root[(3 + ax) * stride];
        const float nx = grid_boundary(lo, hi, L, id + 2);
        if (ax == 0) { T.t0.x = m; T.t1.x = nx; }
        else if (ax == 1) { T.t0.y = m; T.t1.y = nx; }
        else { T.t0.z = m; T.t1.z = nx; }
    }
    // the cell's leaf, if any (grid indices are in the unmirrored frame)
    const uint32_t ix = packed & 0xffu, iy = (packed >> 8) & 0xffu, iz = (packed >> 16) & 0xffu;
    const uint32_t ux = (T.a & 4) ? n - 1 - ix : ix, uy = (T.a & 2) ? n - 1 - iy : iy, uz = (T.a & 1) ? n - 1 - iz : iz;
    const uint32_t cell = cells[(uz * n + uy) * n + ux];
    touches += 1;
    stk[0] = packed;
    if (!cell) return kTravPending;
    leaf = cell;
    tra = fmaxf(fmaxf(T.t0.x, T.t0.y), T.t0.z);
    trb = fminf(fminf(T.t1.x, T.t1.y), T.t1.z);
    return kTravLeaf;
}

// The grid walk with the contract of trav_next: at most `budget` cell
// steps, then kTravPending; a hand-over continues in trav_next within the
// same budget.
__device__ __forceinline__ int grid_next(const uint32_t *__restrict__ cells, uint32_t L,
                                         const uint32_t *__restrict__ chx, Traversal &T, uint32_t *stk,
                                         const float *root, uint32_t stride, uint32_t &leaf, float &tra, float &trb,
                                         uint32_t &touches, uint32_t budget) {
    for (uint32_t steps = 0;; ++steps) {
        if (steps >= budget) return kTravPending;
        const int r = grid_step(cells, L, chx, T, stk, root, stride, leaf, tra, trb, touches);
        if (r == kTravRevelles) return trav_next(chx, T, stk, root, stride, leaf, tra, trb, touches, budget - steps);
        if (r != kTravPending) return r;
    }
}

// ---------------------------------------------------------------------------
// SampleVoxelSpace (src/kernel.cpp:96-126): trilinear lookup in one
// lvoxc^3 vertex-centred brick, lerping x, then y, then z; lerp(a,b,t) is
// t*(b-a)+a with separate rounding, as the reference's SSE3 build computes it.
// ---------------------------------------------------------------------------
// The lookup's cell coordinate a/e along one axis, a = (p - (c - e))*(n-1)/2
// as the reference writes it and e the leaf's half-extent, correctly rounded
// either way: by the division, or, given the correctly rounded reciprocal
// y = RN(1/e) (rcp != 0: every leaf of the scene has the same extent), by
// Markstein's correction of q = RN(a*y): the residual r = a - e*q is exact in
// one fma, and RN(q + r*y) = RN(a/e) when y is within half an ulp of 1/e and q
// within an ulp of a/e, unless the residual underflows (P. Markstein, "Computation
// of elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
// Dev. 34(1), 1990). It underflows only for |a| below about 2^-103, so
// numerators under 2^-100 (zero included, which keeps its sign) take the
// division. Three fma-class operations per axis instead of a division's ten.
__device__ __forceinline__ float markstein_div(float a, float e, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-e, q, a);
    return __builtin_fmaf(r, y, q);
}

__device__ __forceinline__ float brick_lookup(const float *__restrict__ brick, F3 p, float4 ce, uint32_t lvoxc,
                                              float nvm, float rcp) {
#pragma clang fp contract(off)
    // nvm = lvoxc - 1 (exact as a float), passed in rather than converted per lookup
    const float ax = (-0.5f * ((ce.x - ce.w) - p.x)) * nvm, ay = (-0.5f * ((ce.y - ce.w) - p.y)) * nvm,
                az = (-0.5f * ((ce.z - ce.w) - p.z)) * nvm;
    float ni[3];
    if (rcp != 0.0f && fminf(fminf(fabsf(ax), fabsf(ay)), fabsf(az)) >= 0x1p-100f) {
        ni[0] = markstein_div(ax, ce.w, rcp);
        ni[1] = markstein_div(ay, ce.w, rcp);
        ni[2] = markstein_div(az, ce.w, rcp);
    } else {
        ni[0] = ax / ce.w;
        ni[1] = ay / ce.w;
        ni[2] = az / ce.w;
    }
    uint32_t ia[3], ib[3];
    float nl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float nf = fmaxf(floorf(ni[k]), 0.0f);
        float nc = fminf(ceilf(ni[k]), nvm);
        nl[k] = ni[k] - nf;
        nc = fmaxf(fminf(nc, nvm), 0.0f);
        nf = fmaxf(fminf(nf, nvm), 0.0f);
        ia[k] = (uint32_t)nf;
        ib[k] = (uint32_t)nc;
    }
    const uint32_t l2 = lvoxc * lvoxc;
    const uint32_t za = ia[2] * l2, zb = ib[2] * l2, ya = ia[1] * lvoxc, yb = ib[1] * lvoxc;
    float a00 = brick[za + ya + ia[0]], a10 = brick[za + yb + ia[0]], a01 = brick[zb + ya + ia[0]], a11 = brick[zb + yb + ia[0]];
    float b00 = brick[za + ya + ib[0]], b10 = brick[za + yb + ib[0]], b01 = brick[zb + ya + ib[0]], b11 = brick[zb + yb + ib[0]];
    float u00 = nl[0] * (b00 - a00) + a00;
    float u10 = nl[0] * (b10 - a10) + a10;
    float u01 = nl[0] * (b01 - a01) + a01;
    float u11 = nl[0] * (b11 - a11) + a11;
    float v0 = nl[1] * (u10 - u00) + u00;
    float v1 = nl[1] * (u11 - u01) + u01;
    return nl[2] * (v1 - v0) + v0;
}

// Derivative of brick_lookup with respect to the 8 corner values it reads,
// with exactly brick_lookup's index and fraction arithmetic; clamped corners
// may repeat, their weights then add up.
// Cell of the trilinear lookup at p: lower corner ia (per axis), whether the
// upper corner differs (d = ib - ia, 0 or 1 after the clamps) and the lerp
// fractions nl (not clamped, as in the lookup). Packed as
// ia.x | ia.y << 8 | ia.z << 16 | d.x << 24 | d.y << 25 | d.z << 26
// (lvoxc <= 256).
__device__ __forceinline__ uint32_t brick_cell(F3 p, float4 ce, uint32_t lvoxc, float nl[3]) {
    const float nv = (float)lvoxc;
    const float nvm = nv - 1.0f;
    float ni[3] = {((-0.5f * ((ce.x - ce.w) - p.x)) * nvm) / ce.w,
                   ((-0.5f * ((ce.y - ce.w) - p.y)) * nvm) / ce.w,
                   ((-0.5f * ((ce.z - ce.w) - p.z)) * nvm) / ce.w};
    uint32_t cell = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float nf = fmaxf(floorf(ni[k]), 0.0f);
        float nc = fminf(ceilf(ni[k]), nvm);
        nl[k] = ni[k] - nf;
        nc = fmaxf(fminf(nc, nvm), 0.0f);
        nf = fmaxf(fminf(nf, nvm), 0.0f);
        const uint32_t ia = (uint32_t)nf, ib = (uint32_t)nc;
        cell |= (ia << (8 * k)) | ((ib - ia) << (24 + k));
    }
    return cell;
}

// Corner v (bit 0: x, 1: y, 2: z) of a packed cell: its index in the brick
// and its trilinear weight, d(lookup)/d(corner value) (lerp(a,b,t) has
// derivatives 1-t and t).
__device__ __forceinline__ void cell_corner(uint32_t cell, const float nl[3], uint32_t lvoxc, int v, uint32_t &idx,
                                            float &w) {
    const uint32_t bx = v & 1, by = (v >> 1) & 1, bz = (v >> 2) & 1;
    const uint32_t x = (cell & 255u) + (bx & (cell >> 24));
    const uint32_t y = ((cell >> 8) & 255u) + (by & (cell >> 25));
    const uint32_t z = ((cell >> 16) & 255u) + (bz & (cell >> 26));
    idx = z * (lvoxc * lvoxc) + y * lvoxc + x;
    w = (bx ? nl[0] : 1.0f - nl[0]) * (by ? nl[1] : 1.0f - nl[1]) * (bz ? nl[2] : 1.0f - nl[2]);
}

// ---------------------------------------------------------------------------
// Phase functions.
// ---------------------------------------------------------------------------
// SamplingBasis (KernelSampler.cpp:145-153); degenerate for iv = +-y like the reference.
__device__ __forceinline__ void sampling_basis(F3 iv, F3 &b1, F3 &b2) {
#pragma clang fp contract(off)
    float len = sqrtf(iv.x * iv.x + 0.0f * 0.0f + iv.z * iv.z);
    b1 = mk3(-iv.z / len, 0.0f / len, iv.x / len);
    b2 = mk3(iv.y * b1.z - iv.z * b1.y, iv.z * b1.x - iv.x * b1.z, iv.x * b1.y - iv.y * b1.x);
}

__device__ __forceinline__ F3 basis_dir(F3 b1, F3 b2, F3 iv, float st, float ct, float ph) {
    float sph, cph;
    sincosf(ph, &sph, &cph);
    return mk3(b1.x * st * cph + b2.x * st * sph + iv.x * ct,
               b1.y * st * cph + b2.y * st * sph + iv.y * ct,
               b1.z * st * cph + b2.z * st * sph + iv.z * ct);
}

// MiePhase::EvaluateRGB (KernelSampler.cpp:233-256): table lerp on
// a = acos(-ct)/pi with ct clamped to [-1,1]. Uses an exact acos where the
// reference uses a 4-term polynomial approximation.
__device__ __forceinline__ F3 mie_eval_rgb(const MieLds &T, float ct) {
#pragma clang fp contract(off)
    float ct1 = fmaxf(fminf(ct, 1.0f), -1.0f);
    float a = acosf(-ct1) * kInvPi;
    float b = 256.0f * a;
    float c = fmaxf(fminf(floorf(b), 254.0f), 0.0f);
    float t = b - c;
    uint32_t i = (uint32_t)c * 3u;
    return mk3(t * (T.pdf[i + 3] - T.pdf[i + 0]) + T.pdf[i + 0],
               t * (T.pdf[i + 4] - T.pdf[i + 1]) + T.pdf[i + 1],
               t * (T.pdf[i + 5] - T.pdf[i + 2]) + T.pdf[i + 2]);
}

// Key k_i of MiePhase::Sample's scan (KernelSampler.cpp:261-271) and the
// angle it would assign: b = max(ub,i), c = min(floor b,255), t = b-c.
__device__ __forceinline__ float mie_key(const MieLds &T, float ub, uint32_t i, float &t) {
#pragma clang fp contract(off)
    float b = fmaxf(ub, (float)i);
    float c = fminf(floorf(b), 255.0f);
    t = b - c;
    float c0 = T.cdf[i], c1 = T.cdf[i + 1];
    return t * (c1 - c0) + c0;
}

// MiePhase::Sample (KernelSampler.cpp:258-285). The reference scans
// i = 0..254 and keeps the angle of the last i whose key is below u1; the keys
// are non-decreasing in i, so a binary search for the first key >= u1 gives
// the same i.
// First i in [lo, hi) whose key is >= u1, else hi (keys non-decreasing in i).
__device__ __forceinline__ uint32_t mie_search(const MieLds &T, float u1, uint32_t lo, uint32_t hi) {
#pragma clang fp contract(off)
    const float ub = 256.0f * u1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        float t;
        if (mie_key(T, ub, mid, t) < u1) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ F3 mie_sample(const MieLds &T, F3 iv, float u1, float u2) {
#pragma clang fp contract(off)
    float ub = 256.0f * u1;
    // The full search over [0, 255] narrowed to the bounds of u1's bucket
    // (u1 * kMieLutBuckets is exact: a power-of-two scaling).
    const uint32_t lb = T.lut[(uint32_t)(u1 * (float)kMieLutBuckets)];
    const uint32_t lo = mie_search(T, u1, lb & 255u, lb >> 8);
    float th = 0.0f;
    if (lo > 0) {
        float t;
        uint32_t i = lo - 1;
        (void)mie_key(T, ub, i, t);
        th = (float)i / 256.0f + t / 256.0f;
    }
    th *= kPi;
    float ph = (2.0f * kPi) * u2;
    F3 b1, b2;
    sampling_basis(iv, b1, b2);
    float ct = -cosf(th);
    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
    return basis_dir(b1, b2, iv, st, ct, ph);
}

// x / y through the hardware reciprocal (within about 2 ulp), for values
// that only scale radiance (phase and sky values, path weights) and never
// steer a ray, so they need not round exactly as the oracle does.
__device__ __forceinline__ float div_approx(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }

// x^1.5 as x * sqrt(x): two correctly rounded operations, so within about
// 2^-23 relative of the exact power (the general powf costs ~20x more).
__device__ __forceinline__ float pow_1p5(float x) { return x * sqrtf(x); }

// x^2.2 as exp2(2.2 * log2(x)) on the hardware base-2 logarithm and
// exponential: the error of log2 is scaled by 2.2 * ln 2 into the result,
// a few 1e-7 relative for radiance-sized x; 0, infinities and NaNs map as
// powf maps them.
__device__ __forceinline__ float pow_2p2(float x) { return exp2f(2.2f * log2f(x)); }

// HGPhase (KernelSampler.cpp:165-204).
__device__ __forceinline__ float hg_eval(float g, float ct) {
    return div_approx(1.0f - g * g, 4.0f * kPi * pow_1p5(1.0f + g * g - 2.0f * g * ct));
}
__device__ __forceinline__ F3 hg_sample(float g, F3 iv, float u1, float u2) {
    float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
    float ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
    float ph = (2.0f * kPi) * u2;
    F3 b1, b2;
    sampling_basis(iv, b1, b2);
    return basis_dir(b1, b2, iv, st, ct, ph);
}

__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ---------------------------------------------------------------------------
// Sky: Hosek-Wilkie radiance with the reference's gamma construction and
// exposure (src/kernel.cpp:329-360).
// ---------------------------------------------------------------------------
__device__ __forceinline__ F3 hosek_sky(const SkyState &S, F3 rd) {
    float rdz = rd.z;
    float sth = sqrtf(1.0f - rd.z * rd.z);
    float cth = rdz;
    float slth = sqrtf(1.0f - S.dir[2] * S.dir[2]);
    float clth = S.dir[2];
    float slph = S.dir[1];
    float clph = S.dir[0];
    float rdd = div_approx(rd.x, rd.y);
    float acr = __builtin_amdgcn_rsqf(rdd * rdd + 1.0f);
    // rd.y == +-0 makes rdd infinite and acr*(...) 0*inf (NaN): take the limit
    // from the side of the zero's sign, acr*rdd -> sign(rd.x)*sign(rd.y).
    float czp;
    if (rd.y == 0.0f) {
        float sx = rd.x > 0.0f ? 1.0f : (rd.x < 0.0f ? -1.0f : 0.0f);
        czp = (signbit(rd.y) ? -sx : sx) * slph;
    } else {
        czp = acr * (rdd * slph + clph);
    }
    float cph = sth * slth * czp + cth * clth;
    // cos(gamma) can exceed 1 by rounding for rd at the sun (acos would be NaN).
    float gmma = acosf(fmaxf(fminf(cph, 1.0f), -1.0f));
    float gacs = fmaxf(fminf(cph, 1.0f), 0.0f);
    float thcs = fmaxf(rdz, 0.0f);
    float raym = gacs * gacs;
    float zenh = sqrtf(thcs);
    float out[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float *c = S.cfg[i];
        float expm = expf(c[4] * gmma);
        float miem = div_approx(1.0f + raym, pow_1p5(1.0f + c[8] * c[8] - 2.0f * c[8] * gacs));
        float ca = (1.0f + c[0] * expf(div_approx(c[1], thcs + 0.01f))) * (c[2] + c[3] * expm + c[5] * raym + c[6] * miem + c[7] * zenh);
        ca *= S.rad[i];
        out[i] = 1e-3f * pow_2p2(ca);
    }
    return mk3(out[0], out[1], out[2]);
}

// MapEnv::Evaluate (KernelSampler.cpp:369-421), followed by the pow(.,2.2)
// of src/kernel.cpp:362-364.
__device__ __forceinline__ F3 map_env(const float *__restrict__ tex, uint32_t w, uint32_t h, F3 rd) {
    float tha = acosf(-rd.z) * kInvPi;
    float thb = (float)h * fmaxf(fminf(tha, 1.0f), 0.0f);
    float thc = fmaxf(fminf(floorf(thb), (float)(h - 2)), 0.0f);
    float tht = thb - thc;
    float pha = 0.5f + atan2f(rd.y, -rd.x) * (0.5f * kInvPi);
    float phb = (float)w * fmaxf(fminf(pha, 1.0f), 0.0f);
    float phc = fmaxf(fminf(floorf(phb), (float)(w - 2)), 0.0f);
    float pht = phb - phc;
    uint32_t ti = (uint32_t)thc, pj = (uint32_t)phc;
    float out[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ca = tex[(ti * w + pj) * 4 + i];
        float cb = tex[((ti + 1) * w + pj) * 4 + i];
        float da = tex[(ti * w + pj + 1) * 4 + i];
        float db = tex[((ti + 1) * w + pj + 1) * 4 + i];
        float l0 = tht * (cb - ca) + ca;
        float l1 = tht * (db - da) + da;
        out[i] = pow_2p2(pht * (l1 - l0) + l0);
    }
    return mk3(out[0], out[1], out[2]);
}

// Environment radiance for an escaping continuation ray.
__device__ __forceinline__ F3 env_radiance(uint32_t env, const KernelConsts &K, F3 rd) {
    if (env == DRK_ENV_HOSEK) return hosek_sky(K.sky, rd);
    if (env == DRK_ENV_MAP) return map_env(K.env_map, K.env_w, K.env_h, rd);
    return mk3(0.0f, 0.0f, 0.0f);   // NullEnv: pow(0,2.2) = 0
}

// Sum of SunLight::Evaluate over all lights (src/kernel.cpp:326-327).
__device__ __forceinline__ F3 lights_eval(const KernelConsts &P, F3 rd) {
    F3 s = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t l = 0; l < DRK_MAX_LIGHTS; ++l) {
        if (l >= P.light_count) break;
        float d = rd.x * P.sun_dir[l][0] + rd.y * P.sun_dir[l][1] + rd.z * P.sun_dir[l][2];
        if (d > P.sun_cos[l]) { s.x += P.sun_color[l][0]; s.y += P.sun_color[l][1]; s.z += P.sun_color[l][2]; }
    }
    return s;
}

// SunLight::Pdf of lights[0] (KernelSampler.cpp:326-330).
__device__ __forceinline__ float sun_pdf(const KernelConsts &P, F3 v) {
    float d = v.x * P.sun_dir[0][0] + v.y * P.sun_dir[0][1] + v.z * P.sun_dir[0][2];
    return d > P.sun_cos[0] ? P.sun_pdf0 : 0.0f;
}

// SunLight::Sample of lights[0] (KernelSampler.cpp:332-345): cos uniform in [cos angle, 1].
__device__ __forceinline__ F3 sun_sample(const KernelConsts &P, float u1, float u2) {
    float ctm = P.sun_cos[0];
    float ct = (1.0f - u1) + u1 * ctm;
    float st = sqrtf(1.0f - ct * ct);
    float ph = (2.0f * kPi) * u2;
    F3 D = mk3(P.sun_dir[0][0], P.sun_dir[0][1], P.sun_dir[0][2]);
    F3 b1 = mk3(P.sun_b1[0], P.sun_b1[1], P.sun_b1[2]);
    F3 b2 = mk3(P.sun_b2[0], P.sun_b2[1], P.sun_b2[2]);
    return basis_dir(b1, b2, D, st, ct, ph);
}

}  // namespace drk
