// Test hooks: run single device functions of the renderer over batches of
// inputs (declared in droplet_render_probe.h). Not used by the render path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "drk_common.h"
#include "drk_internal.h"
#include "droplet_render_probe.h"
#include "estimator.h"
#include "render_launch.h"

namespace drk {
namespace {

constexpr int kProbeBlock = 256;

// Device scratch for one probe call; every buffer is freed on return.
struct DevBufs {
    std::vector<void *> ptrs;
    ~DevBufs() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T *&p, size_t n) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(q);
        p = static_cast<T *>(q);
        return e;
    }
    template <class T>
    hipError_t upload(T *&p, const T *h, size_t n) {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess && n) e = hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice);
        return e;
    }
};

int fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? DRK_ENOMEM : DRK_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define PTRY(expr, what)                          \
    do {                                          \
        hipError_t _e = (expr);                   \
        if (_e != hipSuccess) return fail(_e, what); \
    } while (0)

uint32_t blocks(uint32_t n) { return (n + kProbeBlock - 1) / kProbeBlock; }

__global__ void k_philox(const uint32_t *ctr, const uint32_t *key, uint32_t n, uint32_t *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    U4 r = philox4x32_10(U4{ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]}, key[2 * i], key[2 * i + 1]);
    out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
}

__global__ void k_stream(const uint32_t *key0, const uint32_t *sample, const uint32_t *ray, uint32_t n, uint32_t key1,
                         uint32_t m, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RayRng r;
    r.seed(key0[i], key1, sample[i], ray[i]);
    for (uint32_t j = 0; j < m; ++j) out[(size_t)i * m + j] = r.uniform();
}

__global__ void k_mie_eval(const float *ct, uint32_t n, float *out) {
    __shared__ MieLds mie;
    load_mie(mie);
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 v = mie_eval_rgb(mie, ct[i]);
    out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
}

__global__ void k_mie_sample(const float *iv, const float *u1, const float *u2, uint32_t n, float *out) {
    __shared__ MieLds mie;
    load_mie(mie);
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 d = mie_sample(mie, mk3(iv[3 * i], iv[3 * i + 1], iv[3 * i + 2]), u1[i], u2[i]);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
}

__global__ void k_log_unit(uint32_t n, float *out_logf, float *out_fast) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float u = (float)((i << 1) | 1u) * 5.9604644775390625e-8f;
    out_logf[i] = logf(u);
    out_fast[i] = log_unit(u);
}

__global__ void k_hg_eval(float g, const float *ct, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = hg_eval(g, ct[i]);
}

__global__ void k_hg_sample(float g, const float *iv, const float *u1, const float *u2, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 d = hg_sample(g, mk3(iv[3 * i], iv[3 * i + 1], iv[3 * i + 2]), u1[i], u2[i]);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
}

__global__ void k_sun(KernelConsts P, const float *u1, const float *u2, const float *dirs, uint32_t n, float *osd,
                      float *opdf, float *oev) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 d = sun_sample(P, u1[i], u2[i]);
    osd[3 * i] = d.x; osd[3 * i + 1] = d.y; osd[3 * i + 2] = d.z;
    F3 v = mk3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    opdf[i] = sun_pdf(P, v);
    F3 e = lights_eval(P, v);
    oev[3 * i] = e.x; oev[3 * i + 1] = e.y; oev[3 * i + 2] = e.z;
}

__global__ void k_map_env(const float *tex, uint32_t w, uint32_t h, const float *rd, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 c = map_env(tex, w, h, mk3(rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]));
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
}

__global__ void k_sky(SkyState S, const float *rd, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 c = hosek_sky(S, mk3(rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]));
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
}

__global__ void k_trilinear(const float *brick, uint32_t lvoxc, float4 ce, float rcp, const float *p, uint32_t n,
                            float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = brick_lookup(brick, mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), ce, lvoxc, (float)lvoxc - 1.0f, rcp);
}

__global__ __launch_bounds__(kProbeBlock) void k_traverse(const drk_octree_node *ob, const uint32_t *chx,
                                                          const uint32_t *cells, uint32_t grid_levels,
                                                          const float *ro, const float *rd,
                                                          uint32_t n, uint32_t max_leaves, uint32_t *nodes,
                                                          float *t0o, float *t1o, uint32_t *count) {
    __shared__ uint32_t stk_all[kMaxDepth * kProbeBlock];
    __shared__ float root_all[6 * kProbeBlock];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t *stk = stk_all + threadIdx.x;
    float *root = root_all + threadIdx.x;
    F3 o = mk3(ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]);
    F3 d = mk3(rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]);
    Traversal T;
    uint32_t cnt = 0, touches = 0;
    F3 t0, t1;
    uint32_t a;
    if (octree_init(ob, o, d, t0, t1, a) && !(t1.x < 0.0f || t1.y < 0.0f || t1.z < 0.0f)) {
        trav_set_root(root, kProbeBlock, t0, t1);
        T.t0 = t0; T.t1 = t1; T.a = a;
        if (is_leaf(ob[0])) {
            if (max_leaves > 0) {
                nodes[(size_t)i * max_leaves] = 0;
                t0o[(size_t)i * max_leaves] = fmaxf(fmaxf(t0.x, t0.y), t0.z);
                t1o[(size_t)i * max_leaves] = fminf(fminf(t1.x, t1.y), t1.z);
            }
            cnt = 1;
        } else {
            if (!(grid_levels && grid_start(t0, t1, grid_levels, T, stk))) {
                F3 tm;
                {
#pragma clang fp contract(off)
                    tm = mk3(0.5f * (t0.x + t1.x), 0.5f * (t0.y + t1.y), 0.5f * (t0.z + t1.z));
                }
                stk[0] = first_node(t0, tm) << 24;
                T.lvl = 0;
            }
            uint32_t leaf;
            float a0, a1;
            while ((T.lvl == kTravGrid
                        ? grid_next(cells, grid_levels, chx, T, stk, root, kProbeBlock, leaf, a0, a1, touches, ~0u)
                        : trav_next(chx, T, stk, root, kProbeBlock, leaf, a0, a1, touches, ~0u)) == kTravLeaf) {
                if (cnt < max_leaves) {
                    nodes[(size_t)i * max_leaves + cnt] = leaf;
                    t0o[(size_t)i * max_leaves + cnt] = a0;
                    t1o[(size_t)i * max_leaves + cnt] = a1;
                }
                ++cnt;
            }
        }
    }
    count[i] = cnt;
}

__global__ void k_camera(KernelConsts K, uint32_t w, uint32_t h, const uint32_t *px, const uint32_t *py, uint32_t n,
                         float *ro, float *rd, uint32_t *gm) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 o, d;
    bool g;
    camera_ray(K, w, h, px[i], py[i], o, d, g);
    ro[3 * i] = o.x; ro[3 * i + 1] = o.y; ro[3 * i + 2] = o.z;
    rd[3 * i] = d.x; rd[3 * i + 1] = d.y; rd[3 * i + 2] = d.z;
    gm[i] = g ? 1u : 0u;
}

// Host mirror of SamplingBasis for the sun (same as api.cpp).
void basis_host(const float iv[3], float b1[3], float b2[3]) {
#pragma clang fp contract(off)
    float len = std::sqrt(iv[0] * iv[0] + 0.0f * 0.0f + iv[2] * iv[2]);
    b1[0] = -iv[2] / len;
    b1[1] = 0.0f / len;
    b1[2] = iv[0] / len;
    b2[0] = iv[1] * b1[2] - iv[2] * b1[1];
    b2[1] = iv[2] * b1[0] - iv[0] * b1[2];
    b2[2] = iv[0] * b1[1] - iv[1] * b1[0];
}

}  // namespace
}  // namespace drk

using namespace drk;

// The scene's device octree lives in the opaque handle; api.cpp exposes it here.
extern "C" const drk_octree_node *drk_internal_scene_nodes(const drk_scene *s);
extern "C" const uint32_t *drk_internal_scene_links(const drk_scene *s);
extern "C" const uint32_t *drk_internal_scene_cells(const drk_scene *s);
extern "C" uint32_t drk_internal_scene_grid_levels(const drk_scene *s);

extern "C" {

int drk_probe_philox(const uint32_t *ctr, const uint32_t *key, uint32_t n, uint32_t *out) {
    if (!ctr || !key || !out) return set_error(DRK_EINVAL, "drk_probe_philox: NULL argument");
    if (!n) return DRK_OK;
    DevBufs B;
    uint32_t *dc, *dk, *dout;
    PTRY(B.upload(dc, ctr, 4 * (size_t)n), "probe upload");
    PTRY(B.upload(dk, key, 2 * (size_t)n), "probe upload");
    PTRY(B.alloc(dout, 4 * (size_t)n), "probe alloc");
    k_philox<<<blocks(n), kProbeBlock>>>(dc, dk, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out, dout, 16 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_ray_stream(const uint32_t *key0, const uint32_t *sample, const uint32_t *ray, uint32_t n, uint32_t key1,
                         uint32_t m, float *out) {
    if (!key0 || !sample || !ray || !out) return set_error(DRK_EINVAL, "drk_probe_ray_stream: NULL argument");
    if (!n || !m) return DRK_OK;
    DevBufs B;
    uint32_t *dk, *ds, *dr;
    float *dout;
    PTRY(B.upload(dk, key0, n), "probe upload");
    PTRY(B.upload(ds, sample, n), "probe upload");
    PTRY(B.upload(dr, ray, n), "probe upload");
    PTRY(B.alloc(dout, (size_t)n * m), "probe alloc");
    k_stream<<<blocks(n), kProbeBlock>>>(dk, ds, dr, n, key1, m, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out, dout, sizeof(float) * (size_t)n * m, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_mie_eval(const float *ct, uint32_t n, float *out_rgb) {
    if (!ct || !out_rgb) return set_error(DRK_EINVAL, "drk_probe_mie_eval: NULL argument");
    if (!n) return DRK_OK;
    DevBufs B;
    float *dct, *dout;
    PTRY(B.upload(dct, ct, n), "probe upload");
    PTRY(B.alloc(dout, 3 * (size_t)n), "probe alloc");
    k_mie_eval<<<blocks(n), kProbeBlock>>>(dct, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_rgb, dout, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_mie_sample(const float *iv, const float *u1, const float *u2, uint32_t n, float *out_dir) {
    if (!iv || !u1 || !u2 || !out_dir) return set_error(DRK_EINVAL, "drk_probe_mie_sample: NULL argument");
    if (!n) return DRK_OK;
    DevBufs B;
    float *div, *du1, *du2, *dout;
    PTRY(B.upload(div, iv, 3 * (size_t)n), "probe upload");
    PTRY(B.upload(du1, u1, n), "probe upload");
    PTRY(B.upload(du2, u2, n), "probe upload");
    PTRY(B.alloc(dout, 3 * (size_t)n), "probe alloc");
    k_mie_sample<<<blocks(n), kProbeBlock>>>(div, du1, du2, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_dir, dout, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_hg_eval(float g, const float *ct, uint32_t n, float *out) {
    if (!ct || !out) return set_error(DRK_EINVAL, "drk_probe_hg_eval: NULL argument");
    if (!n) return DRK_OK;
    DevBufs B;
    float *dct, *dout;
    PTRY(B.upload(dct, ct, n), "probe upload");
    PTRY(B.alloc(dout, n), "probe alloc");
    k_hg_eval<<<blocks(n), kProbeBlock>>>(g, dct, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out, dout, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_hg_sample(float g, const float *iv, const float *u1, const float *u2, uint32_t n, float *out_dir) {
    if (!iv || !u1 || !u2 || !out_dir) return set_error(DRK_EINVAL, "drk_probe_hg_sample: NULL argument");
    if (g == 0.0f) return set_error(DRK_EINVAL, "drk_probe_hg_sample: g must be non-zero");
    if (!n) return DRK_OK;
    DevBufs B;
    float *div, *du1, *du2, *dout;
    PTRY(B.upload(div, iv, 3 * (size_t)n), "probe upload");
    PTRY(B.upload(du1, u1, n), "probe upload");
    PTRY(B.upload(du2, u2, n), "probe upload");
    PTRY(B.alloc(dout, 3 * (size_t)n), "probe alloc");
    k_hg_sample<<<blocks(n), kProbeBlock>>>(g, div, du1, du2, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_dir, dout, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_sun(const drk_sun *sun, const float *u1, const float *u2, const float *dirs, uint32_t n,
                  float *out_sample_dir, float *out_pdf, float *out_eval_rgb) {
    if (!sun || !u1 || !u2 || !dirs || !out_sample_dir || !out_pdf || !out_eval_rgb)
        return set_error(DRK_EINVAL, "drk_probe_sun: NULL argument");
    if (!n) return DRK_OK;
    KernelConsts P;
    std::memset(&P, 0, sizeof(P));
    P.light_count = 1;
    for (int c = 0; c < 3; ++c) {
        P.sun_dir[0][c] = sun->direction[c];
        P.sun_color[0][c] = sun->color[c];
    }
    P.sun_cos[0] = cosf(sun->angle);
    {
#pragma clang fp contract(off)
        P.sun_pdf0 = 1.0f / (2.0f * kPi * (1.0f - P.sun_cos[0] * P.sun_cos[0]));
    }
    basis_host(P.sun_dir[0], P.sun_b1, P.sun_b2);
    DevBufs B;
    float *du1, *du2, *ddir, *osd, *opdf, *oev;
    PTRY(B.upload(du1, u1, n), "probe upload");
    PTRY(B.upload(du2, u2, n), "probe upload");
    PTRY(B.upload(ddir, dirs, 3 * (size_t)n), "probe upload");
    PTRY(B.alloc(osd, 3 * (size_t)n), "probe alloc");
    PTRY(B.alloc(opdf, n), "probe alloc");
    PTRY(B.alloc(oev, 3 * (size_t)n), "probe alloc");
    k_sun<<<blocks(n), kProbeBlock>>>(P, du1, du2, ddir, n, osd, opdf, oev);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_sample_dir, osd, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    PTRY(hipMemcpy(out_pdf, opdf, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    PTRY(hipMemcpy(out_eval_rgb, oev, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_hosek_state(double turbidity, double albedo, double elevation, double *cfg27, double *rad3) {
    if (!cfg27 || !rad3) return set_error(DRK_EINVAL, "drk_probe_hosek_state: NULL argument");
    if (!(turbidity >= 1.0 && turbidity <= 10.0) || !(albedo >= 0.0 && albedo <= 1.0))
        return set_error(DRK_EINVAL, "drk_probe_hosek_state: turbidity in [1,10], albedo in [0,1]");
    double cfg[3][9], rad[3];
    hosek_state(turbidity, albedo, elevation, cfg, rad);
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 9; ++j) cfg27[i * 9 + j] = cfg[i][j];
        rad3[i] = rad[i];
    }
    return DRK_OK;
}

int drk_probe_sky(const drk_sun *sun, float turbidity, float albedo, const float *rd, uint32_t n, float *out_rgb) {
    if (!sun || !rd || !out_rgb) return set_error(DRK_EINVAL, "drk_probe_sky: NULL argument");
    if (!(turbidity >= 1.0f && turbidity <= 10.0f) || !(albedo >= 0.0f && albedo <= 1.0f))
        return set_error(DRK_EINVAL, "drk_probe_sky: turbidity in [1,10], albedo in [0,1]");
    if (!n) return DRK_OK;
    SkyState S;
    float th = acosf(sun->direction[2]);
    float se = 1.57079632679489661923f - th;
    double cfg[3][9], rad[3];
    hosek_state((double)turbidity, (double)albedo, (double)se, cfg, rad);
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 9; ++j) S.cfg[i][j] = (float)cfg[i][j];
        S.rad[i] = (float)rad[i];
        S.dir[i] = sun->direction[i];
    }
    DevBufs B;
    float *drd, *dout;
    PTRY(B.upload(drd, rd, 3 * (size_t)n), "probe upload");
    PTRY(B.alloc(dout, 3 * (size_t)n), "probe alloc");
    k_sky<<<blocks(n), kProbeBlock>>>(S, drd, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_rgb, dout, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_log_unit(uint32_t n, float *out_logf, float *out_fast) {
    if (!out_logf || !out_fast) return set_error(DRK_EINVAL, "drk_probe_log_unit: NULL argument");
    if (n > (1u << 23)) return set_error(DRK_EINVAL, "drk_probe_log_unit: n <= 2^23");
    if (!n) return DRK_OK;
    DevBufs B;
    float *da, *db;
    PTRY(B.alloc(da, n), "probe alloc");
    PTRY(B.alloc(db, n), "probe alloc");
    k_log_unit<<<blocks(n), kProbeBlock>>>(n, da, db);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_logf, da, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    PTRY(hipMemcpy(out_fast, db, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_trilinear(const float *brick, uint32_t lvoxc, const float ce[4], const float *p, uint32_t n,
                        int leaf_rcp, float *out) {
    if (!brick || !ce || !p || !out) return set_error(DRK_EINVAL, "drk_probe_trilinear: NULL argument");
    if (lvoxc < 2 || lvoxc > 256) return set_error(DRK_EINVAL, "drk_probe_trilinear: lvoxc in [2,256]");
    if (!(ce[3] > 0.0f)) return set_error(DRK_EINVAL, "drk_probe_trilinear: extent must be positive");
    if (!n) return DRK_OK;
    DevBufs B;
    float *db, *dp, *dout;
    PTRY(B.upload(db, brick, (size_t)lvoxc * lvoxc * lvoxc), "probe upload");
    PTRY(B.upload(dp, p, 3 * (size_t)n), "probe upload");
    PTRY(B.alloc(dout, n), "probe alloc");
    // leaf_rcp: the division by the extent through its reciprocal, as the
    // render kernels do for scenes whose leaves share one extent.
    const float rcp = leaf_rcp && ce[3] >= 0x1p-60f && ce[3] <= 0x1p20f ? 1.0f / ce[3] : 0.0f;
    k_trilinear<<<blocks(n), kProbeBlock>>>(db, lvoxc, make_float4(ce[0], ce[1], ce[2], ce[3]), rcp, dp, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out, dout, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_traverse(const drk_scene *scene, const float *ro, const float *rd, uint32_t n, uint32_t max_leaves,
                       uint32_t *nodes, float *tr0, float *tr1, uint32_t *count) {
    if (!scene || !ro || !rd || !count || (max_leaves && (!nodes || !tr0 || !tr1)))
        return set_error(DRK_EINVAL, "drk_probe_traverse: NULL argument");
    if (!n) return DRK_OK;
    DevBufs B;
    float *dro, *drd, *dt0, *dt1;
    uint32_t *dn, *dc;
    PTRY(B.upload(dro, ro, 3 * (size_t)n), "probe upload");
    PTRY(B.upload(drd, rd, 3 * (size_t)n), "probe upload");
    PTRY(B.alloc(dn, (size_t)n * max_leaves), "probe alloc");
    PTRY(B.alloc(dt0, (size_t)n * max_leaves), "probe alloc");
    PTRY(B.alloc(dt1, (size_t)n * max_leaves), "probe alloc");
    PTRY(B.alloc(dc, n), "probe alloc");
    // entries past a ray's leaf count stay zero, as the oracle leaves them
    if (max_leaves) {
        PTRY(hipMemset(dn, 0, 4 * (size_t)n * max_leaves), "probe clear");
        PTRY(hipMemset(dt0, 0, 4 * (size_t)n * max_leaves), "probe clear");
        PTRY(hipMemset(dt1, 0, 4 * (size_t)n * max_leaves), "probe clear");
    }
    k_traverse<<<blocks(n), kProbeBlock>>>(drk_internal_scene_nodes(scene), drk_internal_scene_links(scene),
                                           drk_internal_scene_cells(scene), drk_internal_scene_grid_levels(scene), dro,
                                           drd, n, max_leaves, dn, dt0, dt1, dc);
    PTRY(hipGetLastError(), "probe launch");
    if (max_leaves) {
        PTRY(hipMemcpy(nodes, dn, 4 * (size_t)n * max_leaves, hipMemcpyDeviceToHost), "probe download");
        PTRY(hipMemcpy(tr0, dt0, 4 * (size_t)n * max_leaves, hipMemcpyDeviceToHost), "probe download");
        PTRY(hipMemcpy(tr1, dt1, 4 * (size_t)n * max_leaves, hipMemcpyDeviceToHost), "probe download");
    }
    PTRY(hipMemcpy(count, dc, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_camera(const float viewi[16], const float proji[16], uint32_t w, uint32_t h, const uint32_t *px,
                     const uint32_t *py, uint32_t n, float *ro, float *rd, uint32_t *gm) {
    if (!viewi || !proji || !px || !py || !ro || !rd || !gm) return set_error(DRK_EINVAL, "drk_probe_camera: NULL argument");
    if (!w || !h) return set_error(DRK_EINVAL, "drk_probe_camera: empty image");
    if (!n) return DRK_OK;
    KernelConsts K;
    std::memset(&K, 0, sizeof(K));
    std::memcpy(K.viewi, viewi, sizeof(K.viewi));
    std::memcpy(K.proji, proji, sizeof(K.proji));
    DevBufs B;
    uint32_t *dpx, *dpy, *dgm;
    float *dro, *drd;
    PTRY(B.upload(dpx, px, n), "probe upload");
    PTRY(B.upload(dpy, py, n), "probe upload");
    PTRY(B.alloc(dro, 3 * (size_t)n), "probe alloc");
    PTRY(B.alloc(drd, 3 * (size_t)n), "probe alloc");
    PTRY(B.alloc(dgm, n), "probe alloc");
    k_camera<<<blocks(n), kProbeBlock>>>(K, w, h, dpx, dpy, n, dro, drd, dgm);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(ro, dro, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    PTRY(hipMemcpy(rd, drd, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    PTRY(hipMemcpy(gm, dgm, 4 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_map_env(const float *tex, uint32_t w, uint32_t h, const float *rd, uint32_t n, float *out_rgb) {
    if (!tex || !rd || !out_rgb) return set_error(DRK_EINVAL, "drk_probe_map_env: NULL argument");
    if (w < 2 || h < 2) return set_error(DRK_EINVAL, "drk_probe_map_env: needs >= 2x2 texels");
    if (!n) return DRK_OK;
    DevBufs B;
    float *dtex, *drd, *dout;
    PTRY(B.upload(dtex, tex, 4 * (size_t)w * h), "probe upload");
    PTRY(B.upload(drd, rd, 3 * (size_t)n), "probe upload");
    PTRY(B.alloc(dout, 3 * (size_t)n), "probe alloc");
    k_map_env<<<blocks(n), kProbeBlock>>>(dtex, w, h, drd, n, dout);
    PTRY(hipGetLastError(), "probe launch");
    PTRY(hipMemcpy(out_rgb, dout, 12 * (size_t)n, hipMemcpyDeviceToHost), "probe download");
    return DRK_OK;
}

int drk_probe_udiv(uint32_t d, uint32_t n0, uint32_t step, uint64_t count, uint64_t *mismatches) {
    if (!mismatches || d == 0) return set_error(DRK_EINVAL, "drk_probe_udiv: NULL argument or zero divisor");
    const UDiv D = udiv_make(d);
    uint64_t bad = 0;
    uint32_t n = n0;
    for (uint64_t i = 0; i < count; ++i, n += step) bad += udiv(n, D) != n / d;
    for (uint32_t j = 0; j < 64; ++j) bad += udiv(0xffffffffu - j, D) != (0xffffffffu - j) / d;
    *mismatches = bad;
    return DRK_OK;
}

}  // extern "C"
