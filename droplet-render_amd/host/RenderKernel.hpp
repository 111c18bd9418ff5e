// RenderKernel.hpp - C++ render entry points of droplet-render, over the C ABI.
//
// Header-only mirror of the reference's kernel-side interface:
//   RenderKernel                 src/kernel.h:11-59
//   KernelSampler::PhaseFunction src/KernelSampler.h:6-35  (MiePhase, HGPhase)
//   KernelSampler::BaseLight     src/KernelSampler.h:37-59 (SunLight, lights registry)
//   KernelSampler::BaseEnv       src/KernelSampler.h:61-97 (NullEnv, MapEnv)
//   Scene                        src/scene.h:146-159 (ob, pvol, lvoxc, leafx)
// with the same method names, argument order and meaning. Value types are
// layout-compatible with the reference's (dfloat3/dfloat4 are plain floats,
// dmatrix44 is four float4 rows). The work runs in libdroplet_amd.so
// (include/droplet_render.h); this header only owns handles and buffers.
// Errors: Initialize returns false (as the reference does on allocation
// failure, src/kernel.cpp:500-502); Render/Shadow leave the error text in
// drk_last_error() and zero phb.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "droplet_render.h"

namespace droplet {

typedef unsigned int uint;
struct dfloat3 { float x, y, z; };
struct alignas(16) dfloat4 { float x, y, z, w; };
struct alignas(16) dmatrix44 { dfloat4 r[4]; };

enum VOLUME_BUFFER { VOLUME_BUFFER_SDF, VOLUME_BUFFER_FOG, VOLUME_BUFFER_COUNT };

// Scene: the octree records and bricks the kernel reads (src/scene.h:146-159).
class Scene {
public:
    std::vector<drk_octree_node> ob;
    float *pvol[VOLUME_BUFFER_COUNT] = {nullptr, nullptr};
    uint lvoxc = 8;
    uint index = 0;
    uint leafx[VOLUME_BUFFER_COUNT] = {0, 0};
    uint lvoxc3 = 512;

    // Scene::Initialize's role for a dense grid (x fastest, voxel (i,j,k) at
    // world (i,j,k)*s): octree + bricks via drk_scene_build_dense.
    bool Initialize(const float *fog, uint nx, uint ny, uint nz, float s, uint maxd) {
        Destroy();
        drk_host_scene *hs = nullptr;
        if (drk_scene_build_dense(fog, nullptr, nx, ny, nz, s, maxd, &hs) != DRK_OK) return false;
        ob.assign(hs->ob, hs->ob + hs->node_count);
        lvoxc = hs->lvoxc;
        lvoxc3 = lvoxc * lvoxc * lvoxc;
        index = hs->node_count - 1;
        for (int b = 0; b < VOLUME_BUFFER_COUNT; ++b) {
            leafx[b] = hs->leafx[b];
            size_t n = (size_t)leafx[b] * lvoxc3;
            pvol[b] = n ? new float[n] : nullptr;
            if (n) std::memcpy(pvol[b], hs->pvol[b], n * sizeof(float));
        }
        drk_host_scene_free(hs);
        return true;
    }
    void Destroy() {
        for (auto &p : pvol) {
            delete[] p;
            p = nullptr;
        }
        ob.clear();
        if (dev) drk_scene_destroy(dev);
        dev = nullptr;
    }
    ~Scene() { Destroy(); }
    // Device copy, created on first use by a kernel.
    drk_scene *device() const {
        if (!dev) drk_scene_create(ob.data(), (uint32_t)ob.size(), nullptr, 0, pvol[VOLUME_BUFFER_FOG],
                                   leafx[VOLUME_BUFFER_FOG], lvoxc, &dev);
        return dev;
    }

private:
    mutable drk_scene *dev = nullptr;
};

class SceneOcclusion;   // Embree holdouts: compiled out in the reference by default

namespace KernelSampler {

class PhaseFunction {
public:
    virtual ~PhaseFunction() {}
    virtual uint32_t kind() const = 0;
    virtual float g() const { return 0.0f; }
};

class MiePhase : public PhaseFunction {
public:
    uint32_t kind() const override { return DRK_PHASE_MIE; }
    static MiePhase gmie;
};

class HGPhase : public PhaseFunction {
public:
    explicit HGPhase(float _g) : g1(_g) {}
    uint32_t kind() const override { return DRK_PHASE_HG; }
    float g() const override { return g1; }
    float g1;
    static HGPhase ghg;
};

class BaseLight {
public:
    BaseLight() { lights.push_back(this); }
    virtual ~BaseLight() {}
    static void DeleteAll() {
        for (BaseLight *l : lights) delete l;
        lights.clear();
    }
    virtual drk_sun desc() const = 0;
    static std::vector<BaseLight *> lights;
};

class SunLight : public BaseLight {
public:
    SunLight(const dfloat3 *pd, const dfloat3 *pc, float _angle) : direction(*pd), color(*pc), angle(_angle) {
        cosAngle = cosf(_angle);
    }
    drk_sun desc() const override {
        drk_sun s;
        s.direction[0] = direction.x; s.direction[1] = direction.y; s.direction[2] = direction.z;
        s.color[0] = color.x; s.color[1] = color.y; s.color[2] = color.z;
        s.angle = angle;
        return s;
    }
    dfloat3 direction;
    dfloat3 color;
    float angle;
    float cosAngle;
};

class BaseEnv {
public:
    virtual ~BaseEnv() {}
    virtual uint32_t kind() const = 0;
};

class NullEnv : public BaseEnv {
public:
    uint32_t kind() const override { return DRK_ENV_NULL; }
    static NullEnv nenv;
};

class MapEnv : public BaseEnv {
public:
    uint32_t kind() const override { return DRK_ENV_MAP; }
    dfloat4 *Initialize(uint _w, uint _h) {
        Destroy();
        ptex = new dfloat4[(size_t)_w * _h];
        w = _w;
        h = _h;
        return ptex;
    }
    void Destroy() {
        delete[] ptex;
        ptex = nullptr;
    }
    ~MapEnv() { Destroy(); }
    dfloat4 *ptex = nullptr;
    uint w = 0, h = 0;
    static MapEnv genv;
};

// The Hosek-Wilkie sky of USE_ARHOSEK_SKYMODEL builds (src/kernel.cpp:329-360),
// with turbidity and ground albedo as parameters.
class HosekSky : public BaseEnv {
public:
    HosekSky(float t = 7.0f, float a = 0.15f) : turbidity(t), albedo(a) {}
    uint32_t kind() const override { return DRK_ENV_HOSEK; }
    float turbidity, albedo;
};

inline MiePhase MiePhase::gmie;
inline HGPhase HGPhase::ghg(0.75f);
inline std::vector<BaseLight *> BaseLight::lights;
inline NullEnv NullEnv::nenv;
inline MapEnv MapEnv::genv;

}  // namespace KernelSampler

#define KERNEL_DEPTHCOMP DRK_KERNEL_DEPTHCOMP

class RenderKernel {
public:
    RenderKernel() {}
    ~RenderKernel() { Destroy(); }

    bool Initialize(const Scene *pscene, const SceneOcclusion *psceneocc, const dmatrix44 *pviewi,
                    const dmatrix44 *pproji, KernelSampler::PhaseFunction *ppf, KernelSampler::BaseEnv *penv,
                    float *_pdepth, uint _scattevs, float _msigmas, float _msigmaa, uint tilex, uint tiley, uint _w,
                    uint _h, uint _flags) {
        (void)psceneocc;
        Destroy();
        if (!pscene || !pviewi || !pproji || !ppf || !penv) return false;
        drk_scene *ds = pscene->device();
        if (!ds) return false;
        std::vector<drk_sun> suns;
        for (const KernelSampler::BaseLight *l : KernelSampler::BaseLight::lights) suns.push_back(l->desc());
        drk_sampler_desc d;
        std::memset(&d, 0, sizeof(d));
        d.phase = ppf->kind();
        d.hg_g = ppf->g();
        d.env = penv->kind();
        if (auto *sky = dynamic_cast<KernelSampler::HosekSky *>(penv)) {
            d.sky_turbidity = sky->turbidity;
            d.sky_albedo = sky->albedo;
        }
        if (auto *map = dynamic_cast<KernelSampler::MapEnv *>(penv)) {
            d.env_map = reinterpret_cast<const float *>(map->ptex);
            d.env_w = map->w;
            d.env_h = map->h;
        }
        d.lights = suns.data();
        d.light_count = (uint32_t)suns.size();
        if (drk_kernel_create(ds, &pviewi->r[0].x, &pproji->r[0].x, &d, _pdepth, _scattevs, _msigmas, _msigmaa,
                              tilex, tiley, _w, _h, _flags, &k) != DRK_OK)
            return false;
        for (uint i = 0; i < BUFFER_COUNT; ++i) {
            phb[i] = static_cast<dfloat4 *>(std::aligned_alloc(16, (size_t)tilex * tiley * 16 + 16));
            if (!phb[i]) return false;
        }
        pscene_ = pscene;
        pdepth = _pdepth;
        viewi = *pviewi;
        proji = *pproji;
        scattevs = _scattevs;
        msigmas = _msigmas;
        msigmaa = _msigmaa;
        w = _w;
        h = _h;
        tilex_ = tilex;
        tiley_ = tiley;
        flags = _flags;
        return true;
    }

    void Render(uint x0, uint y0, uint tilex, uint tiley, uint samples) { run(false, x0, y0, tilex, tiley, samples); }
    void Shadow(uint x0, uint y0, uint tilex, uint tiley, uint samples) { run(true, x0, y0, tilex, tiley, samples); }

    void Destroy() {
        if (k) drk_kernel_destroy(k);
        k = nullptr;
        for (auto &p : phb) {
            std::free(p);
            p = nullptr;
        }
    }

    enum BUFFER { BUFFER_LIGHTS, BUFFER_ENVIRONMENT, BUFFER_COUNT };
    dfloat4 *phb[BUFFER_COUNT] = {nullptr, nullptr};
    float *pdepth = nullptr;
    dmatrix44 viewi, proji;
    uint scattevs = 0;
    float msigmas = 0, msigmaa = 0;
    uint w = 0, h = 0;
    uint tilew = 0, tileh = 0;
    uint flags = 0;
    drk_kernel *handle() const { return k; }

private:
    void run(bool shadow, uint x0, uint y0, uint tilex, uint tiley, uint samples) {
        tilew = tilex;
        tileh = tiley;
        int rc = shadow ? drk_kernel_shadow(k, x0, y0, tilex, tiley, samples)
                        : drk_kernel_render(k, x0, y0, tilex, tiley, samples);
        // The library's buffers are packed with this call's tilex, like phb.
        const size_t n = (size_t)std::min(tilex, tilex_) * std::min(tiley, tiley_);
        for (uint b = 0; b < BUFFER_COUNT; ++b) {
            if (!phb[b]) continue;
            const float *src = rc == DRK_OK ? drk_kernel_phb(k, b) : nullptr;
            if (src) std::memcpy(phb[b], src, n * 16);
            else std::memset(phb[b], 0, n * 16);
        }
    }
    drk_kernel *k = nullptr;
    const Scene *pscene_ = nullptr;
    uint tilex_ = 0, tiley_ = 0;
};

}  // namespace droplet
