// Synthetic procedural clouds used as benchmark and test inputs (the reference
// builds its grids from Blender node trees through OpenVDB, which is outside
// this path). Deterministic for a given seed on any machine.
//
//   kind 0, fBm cloud:      f(x) = sum_{o<oct} gain^o * noise(freq*lac^o*x + off_o)
//                           rho  = max(f - threshold, 0) / max over the grid
//   kind 1, sphere union:   phi(x) = min_k(|x-c_k| - r_k) - amp * fbm(x)
//                           rho  = clamp(-phi / band, 0, 1)
// noise() is Perlin's improved gradient noise (Perlin 2002) with a
// permutation and offsets drawn from Philox-4x32-10 keyed by the seed.
// Voxel (i,j,k) sits at world (i,j,k)/n.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "drk_common.h"
#include "drk_internal.h"

namespace drk {
namespace {

struct Noise {
    uint8_t perm[512];
    explicit Noise(uint64_t seed) {
        uint8_t p[256];
        for (int i = 0; i < 256; ++i) p[i] = (uint8_t)i;
        uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32) ^ 0x6e6f6973u;
        for (int i = 255; i > 0; --i) {   // Fisher-Yates with Philox draws
            U4 r = philox4x32_10(U4{(uint32_t)i, 0x7065726du, 0u, 0u}, k0, k1);
            int j = (int)(r.x % (uint32_t)(i + 1));
            std::swap(p[i], p[j]);
        }
        for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
    }
    static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
    static float grad(int h, float x, float y, float z) {
        h &= 15;
        float u = h < 8 ? x : y;
        float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
        return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
    }
    static float lerp(float t, float a, float b) { return a + t * (b - a); }
    float operator()(float x, float y, float z) const {
        float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
        x -= fx; y -= fy; z -= fz;
        float u = fade(x), v = fade(y), w = fade(z);
        int A = perm[X] + Y, AA = perm[A] + Z, AB = perm[A + 1] + Z;
        int B = perm[X + 1] + Y, BA = perm[B] + Z, BB = perm[B + 1] + Z;
        return lerp(w, lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                               lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                    lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                         lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))));
    }
};

struct Fbm {
    Noise noise;
    uint32_t octaves;
    float freq, lac, gain;
    float off[16][3];
    Fbm(uint64_t seed, uint32_t oct, float f, float l, float g) : noise(seed), octaves(std::min(oct, 16u)), freq(f), lac(l), gain(g) {
        for (uint32_t o = 0; o < 16; ++o) {
            U4 r = philox4x32_10(U4{o, 0x6f666673u, 0u, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
            off[o][0] = 256.0f * u01(r.x);
            off[o][1] = 256.0f * u01(r.y);
            off[o][2] = 256.0f * u01(r.z);
        }
    }
    float operator()(float x, float y, float z) const {
        float s = 0.0f, a = 1.0f, f = freq;
        for (uint32_t o = 0; o < octaves; ++o) {
            s += a * noise(f * x + off[o][0], f * y + off[o][1], f * z + off[o][2]);
            f *= lac;
            a *= gain;
        }
        return s;
    }
};

template <class F>
void parallel_slices(uint32_t n, F &&f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (uint32_t z = t; z < n; z += nt) f(z);
        });
    for (auto &x : th) x.join();
}

}  // namespace
}  // namespace drk

using namespace drk;

extern "C" int drk_synth_cloud(uint32_t kind, uint32_t n, uint64_t seed, const float *prm, uint32_t nprm, float *out) {
    if (!out || !prm || n == 0 || n > 4096) return set_error(DRK_EINVAL, "drk_synth_cloud: bad arguments");
    const float inv = 1.0f / (float)n;
    const size_t plane = (size_t)n * n;
    if (kind == 0) {
        // prm: octaves, base frequency, lacunarity, gain, threshold
        if (nprm < 5) return set_error(DRK_EINVAL, "drk_synth_cloud: fBm needs 5 parameters");
        Fbm fb(seed, (uint32_t)prm[0], prm[1], prm[2], prm[3]);
        const float thr = prm[4];
        std::vector<float> zmax(n, 0.0f);
        parallel_slices(n, [&](uint32_t z) {
            float m = 0.0f;
            for (uint32_t y = 0; y < n; ++y)
                for (uint32_t x = 0; x < n; ++x) {
                    float v = std::max(fb(x * inv, y * inv, z * inv) - thr, 0.0f);
                    out[z * plane + (size_t)y * n + x] = v;
                    m = std::max(m, v);
                }
            zmax[z] = m;
        });
        float m = *std::max_element(zmax.begin(), zmax.end());
        if (m > 0.0f) {
            const float s = 1.0f / m;
            parallel_slices(n, [&](uint32_t z) {
                for (size_t i = 0; i < plane; ++i) out[z * plane + i] *= s;
            });
        }
        return DRK_OK;
    }
    if (kind == 1) {
        // prm: spheres, centre min, centre max, radius min, radius max,
        //      displacement amplitude, octaves, base frequency, band width
        if (nprm < 9) return set_error(DRK_EINVAL, "drk_synth_cloud: sphere union needs 9 parameters");
        const uint32_t ns = std::max(1u, (uint32_t)prm[0]);
        std::vector<float> sp(4 * (size_t)ns);
        for (uint32_t k = 0; k < ns; ++k) {
            U4 r = philox4x32_10(U4{k, 0x73706872u, 0u, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
            sp[4 * k + 0] = prm[1] + (prm[2] - prm[1]) * u01(r.x);
            sp[4 * k + 1] = prm[1] + (prm[2] - prm[1]) * u01(r.y);
            sp[4 * k + 2] = prm[1] + (prm[2] - prm[1]) * u01(r.z);
            sp[4 * k + 3] = prm[3] + (prm[4] - prm[3]) * u01(r.w);
        }
        Fbm fb(seed ^ 0x9e3779b97f4a7c15ull, (uint32_t)prm[6], prm[7], 2.0f, 0.5f);
        const float amp = prm[5], band = prm[8];
        parallel_slices(n, [&](uint32_t z) {
            for (uint32_t y = 0; y < n; ++y)
                for (uint32_t x = 0; x < n; ++x) {
                    float px = x * inv, py = y * inv, pz = z * inv;
                    float d = 1e30f;
                    for (uint32_t k = 0; k < ns; ++k) {
                        float dx = px - sp[4 * k], dy = py - sp[4 * k + 1], dz = pz - sp[4 * k + 2];
                        d = std::min(d, std::sqrt(dx * dx + dy * dy + dz * dz) - sp[4 * k + 3]);
                    }
                    float v = 0.0f;
                    // Far outside every sphere the displacement cannot bring the
                    // surface back (|fbm| < 2): skip the noise there.
                    if (d < 2.0f * amp + band) {
                        float phi = d - amp * fb(px, py, pz);
                        v = std::min(std::max(-phi / band, 0.0f), 1.0f);
                    }
                    out[z * plane + (size_t)y * n + x] = v;
                }
        });
        return DRK_OK;
    }
    return set_error(DRK_EINVAL, "drk_synth_cloud: unknown kind %u", kind);
}
