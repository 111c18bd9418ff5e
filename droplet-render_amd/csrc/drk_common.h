// Definitions shared by the host library and the gfx950 kernels: node and
// parameter layouts, the counter-based RNG and the small deterministic math
// helpers whose exact operation order the bit-exact parts rely on.
#pragma once

#include <stdint.h>
#include "droplet_render.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DRK_HD __host__ __device__ __forceinline__
#else
#define DRK_HD inline
#endif

namespace drk {

constexpr uint32_t kMaxDepth = 16;       // traversal stack levels (reference MAX_DEPTH, KernelOctree.h:35)
constexpr uint32_t kMaxNodes = 1u << 24; // node index field of a traversal stack word
constexpr uint32_t kMieBins = 256;       // miedl, KernelSampler.cpp:143
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;   // 1/pi rounded to float

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11). Keyed by (pixel, seed); the counter names the
// sample, the ray within the path, the step within the ray and the purpose.
// Replaces the reference's per-packet Tausworthe/LCG streams
// (src/kernel.cpp:23-84), which cannot be replayed.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

DRK_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

DRK_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c.x, hi0, lo0);
        mulhilo32(0xCD9E8D57u, c.z, hi1, lo1);
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Uniform in the open interval (0,1): the odd multiples of 2^-24, all exactly
// representable, so -log(u) is finite and u < 1 (the reference generator can
// return 0 and 1, src/kernel.cpp:61-63).
DRK_HD float u01(uint32_t x) {
    return (float)((x >> 8) | 1u) * 5.9604644775390625e-8f;
}

// RNG key from the global pixel index and the 64-bit seed.
DRK_HD uint32_t rng_key1(uint64_t seed) {
    return (uint32_t)seed ^ ((uint32_t)(seed >> 32) * 0x9E3779B9u);
}

// Counter purposes (counter word w).
constexpr uint32_t kRngTrack = 0;    // free-flight / acceptance pairs of a tracked ray
constexpr uint32_t kRngScatter = 1;  // phase + light direction samples at a scattering event
constexpr uint32_t kRngShadowPass = 2;  // sun direction of a Shadow-pass sample
constexpr uint32_t kRngAdjoint = 3;     // adjoint integral sample points of a replayed ray

// Ray identifiers inside one path: the camera ray is 0; scattering event e
// (at depth e) spawns the shadow ray 2e+1 and the continuation ray 2e+2.
DRK_HD uint32_t shadow_ray_id(uint32_t e) { return 2u * e + 1u; }
DRK_HD uint32_t cont_ray_id(uint32_t e) { return 2u * e + 2u; }

// ---------------------------------------------------------------------------
// Kernel parameters (one struct per launch, passed by value).
// ---------------------------------------------------------------------------
struct SkyState {
    float cfg[3][9];   // ArHosekSkyModelState::configs[0..2], cast to float (src/kernel.cpp:413)
    float rad[3];      // ::radiances[0..2]
    float dir[3];      // skydir = lights[0]->direction (src/kernel.cpp:522)
};

// Per-kernel constants (camera, lights, sky, environment): device-resident,
// read through a pointer so the kernels load them on use instead of holding
// them in scalar registers for the whole launch.
struct KernelConsts {
    float viewi[16], proji[16];
    uint32_t light_count;
    float sun_dir[DRK_MAX_LIGHTS][3];
    float sun_color[DRK_MAX_LIGHTS][3];
    float sun_cos[DRK_MAX_LIGHTS];
    float sun_pdf0;            // lights[0]->Pdf value inside the cone
    float sun_b1[3], sun_b2[3];// SamplingBasis(lights[0]->direction)
    SkyState sky;
    const float *env_map;      // RGBA texels for DRK_ENV_MAP
    uint32_t env_w, env_h;
    float clip_end;            // Shadow pass far plane
};

// Child link word: node index (bits 23:0) and the child's leaf flag.
constexpr uint32_t kChildLeaf = 0x80000000u;

struct RenderParams {
    const drk_octree_node *ob;
    const uint32_t *chx;      // node_count x 8 child links (index | kChildLeaf)
    const uint32_t *cells;    // leaf grid: (2^grid_levels)^3 leaf node indices (0: empty), x fastest; or null
    uint32_t grid_levels;     // depth of every leaf (0: walk the octree)
    const float *pvol_fog;
    const float *pvol_sdf;
    const float *pvol_ref;    // weighted tracking: the reference FOG bricks that decide collisions
    const float *depth;       // w*h compositing depth or nullptr
    const KernelConsts *kc;
    uint32_t lvoxc, lvoxc3;
    float lvoxm1;             // lvoxc - 1 as a float (the lookup's cell scale)
    float leaf_rcp;           // RN(1 / leaf half-extent) when every leaf has the same extent, else 0
    uint32_t w, h;            // full image size (NDC)
    uint32_t x0, y0, rx, ry;  // tile origin and the rendered extent (even)
    uint32_t tilex;           // output row stride
    uint32_t samples;
    uint32_t sample_base;
    uint32_t key1;
    uint32_t scattevs;
    float sigma_s, sigma_e;   // msigmas, msigmas+msigmaa
    uint32_t phase;
    float hg_g;
    uint32_t env;
    uint32_t shadow_mode;      // 0: Render, 1: Shadow pass
};

}  // namespace drk
