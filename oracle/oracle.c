/*
 * CPU oracle of droplet-render's volumetric light-transport path.
 *
 * TEST INFRASTRUCTURE ONLY: this restatement is the checker the parity tests
 * compare the gfx950 kernels against, and the CPU baseline that bench.py
 * times beside them. The product path never links or calls it.
 *
 * Plain C, one lane per path (or_render), written from the reference's
 * behaviour (jaelpark/droplet-render); or_render_packets traces the same
 * estimator in the reference's 2x2-packet, 4-lane structure (same sums):
 *   ray generation            src/kernel.cpp:447-487 (K_ParallelRender)
 *   estimator (recursive)     src/kernel.cpp:128-445 (SampleVolume), followed
 *                             structurally: the recursion, the masks per lane,
 *                             the leaf loop and the Woodcock loop
 *   trilinear lookup          src/kernel.cpp:96-126 (SampleVoxelSpace)
 *   octree traversal          src/KernelOctree.cpp:9-220 (recursive full traverser)
 *   Mie / HG / sun            src/KernelSampler.cpp:145-345
 *   sky                       src/kernel.cpp:329-360, src/ArHosekSkyModel.cpp:147-339
 *   octree build              src/scene.cpp:234-305, :676-762, :840-917
 * Deliberate, shared deviations from the reference (also in the kernels):
 *   - RNG: Philox-4x32-10 keyed by (pixel, seed) seeds one xoshiro128**
 *     stream per (sample, ray); uniforms lie in (0,1) strictly;
 *   - acos/cos/sin/log/exp/pow are the C library's, not the reference's
 *     polynomial/SSE approximations;
 *   - Hosek turbidity/albedo are parameters (the reference hard-codes 7/0.15).
 * Compiled with -ffp-contract=off so every multiply-add rounds twice, as the
 * reference's default SSE3 build does.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ONO 0xffffffffu

typedef struct {
    float ce[4];
    uint32_t chn[8];
    uint32_t volx[2];
    float qval[2];
} onode;

/* ------------------------------------------------------------------ RNG -- */
static void philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void or_philox(const uint32_t *ctr, const uint32_t *key, uint32_t n, uint32_t *out) {
    for (uint32_t i = 0; i < n; ++i) philox(ctr + 4 * i, key[2 * i], key[2 * i + 1], out + 4 * i);
}

static float u01(uint32_t x) { return (float)((x >> 8) | 1u) * 5.9604644775390625e-8f; }

typedef struct { uint32_t s[4]; } ostream;

static void stream_seed(ostream *r, uint32_t key0, uint32_t key1, uint32_t sample, uint32_t ray) {
    uint32_t c[4] = {sample, ray, 0u, 0u}, v[4];
    philox(c, key0, key1, v);
    r->s[0] = v[0]; r->s[1] = v[1]; r->s[2] = v[2];
    r->s[3] = v[3] | ((v[0] | v[1] | v[2]) == 0u ? 1u : 0u);
}

static uint32_t stream_next(ostream *r) {
    uint32_t *s = r->s;
    uint32_t x = s[1] * 5u;
    uint32_t res = ((x << 7) | (x >> 25)) * 9u;
    uint32_t t = s[1] << 9;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t;
    s[3] = (s[3] << 11) | (s[3] >> 21);
    return res;
}

static float stream_u(ostream *r) { return u01(stream_next(r)); }

void or_ray_stream(const uint32_t *key0, const uint32_t *sample, const uint32_t *ray, uint32_t n, uint32_t key1,
                   uint32_t m, float *out) {
    for (uint32_t i = 0; i < n; ++i) {
        ostream r;
        stream_seed(&r, key0[i], key1, sample[i], ray[i]);
        for (uint32_t j = 0; j < m; ++j) out[(size_t)i * m + j] = stream_u(&r);
    }
}

/* ------------------------------------------------------------- octree -- */
typedef struct { uint32_t node; float t0, t1; } oleaf;

typedef struct {
    const onode *ob;
    oleaf *ls;
    uint32_t n, cap;
} olist;

static void push_leaf(olist *L, uint32_t node, float a, float b) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 64;
        L->ls = (oleaf *)realloc(L->ls, sizeof(oleaf) * L->cap);
    }
    L->ls[L->n].node = node; L->ls[L->n].t0 = a; L->ls[L->n].t1 = b;
    L->n++;
}

static uint32_t first_node(const float t0[3], const float tm[3]) {
    uint32_t a = 0;
    if (t0[0] > t0[1]) {
        if (t0[0] > t0[2]) {
            if (tm[1] < t0[0]) a |= 2;
            if (tm[2] < t0[0]) a |= 1;
            return a;
        }
    } else {
        if (t0[1] > t0[2]) {
            if (tm[0] < t0[1]) a |= 4;
            if (tm[2] < t0[1]) a |= 1;
            return a;
        }
    }
    if (tm[0] < t0[2]) a |= 4;
    if (tm[1] < t0[2]) a |= 2;
    return a;
}

static uint32_t next_node(const float tm[3], uint32_t x, uint32_t y, uint32_t z) {
    if (tm[0] < tm[1]) {
        if (tm[0] < tm[2]) return x;
    } else {
        if (tm[1] < tm[2]) return y;
    }
    return z;
}

static float fmax3(const float v[3]) { return fmaxf(fmaxf(v[0], v[1]), v[2]); }
static float fmin3(const float v[3]) { return fminf(fminf(v[0], v[1]), v[2]); }
static int is_leaf(const onode *n) { return n->volx[0] != ONO || n->volx[1] != ONO; }

/* OctreeFullTraverser::OctreeProcessSubtree (KernelOctree.cpp:143-200). */
static void process_subtree(olist *L, const float t0[3], const float t1[3], uint32_t a, uint32_t n, uint32_t l) {
    static const uint32_t nla[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    if (t1[0] < 0.0f || t1[1] < 0.0f || t1[2] < 0.0f || (n == 0 && l > 0)) return;
    const onode *N = &L->ob[n];
    if (is_leaf(N)) {
        push_leaf(L, n, fmax3(t0), fmin3(t1));
        return;
    }
    float tm[3] = {0.5f * (t0[0] + t1[0]), 0.5f * (t0[1] + t1[1]), 0.5f * (t0[2] + t1[2])};
    uint32_t tn = first_node(t0, tm);
    do {
        float a0[3], a1[3];
        switch (tn) {
        case 0:
            process_subtree(L, t0, tm, a, N->chn[nla[a]], l + 1);
            tn = next_node(tm, 4, 2, 1);
            break;
        case 1:
            a0[0] = t0[0]; a0[1] = t0[1]; a0[2] = tm[2]; a1[0] = tm[0]; a1[1] = tm[1]; a1[2] = t1[2];
            process_subtree(L, a0, a1, a, N->chn[nla[1 ^ a]], l + 1);
            tn = next_node(a1, 5, 3, 8);
            break;
        case 2:
            a0[0] = t0[0]; a0[1] = tm[1]; a0[2] = t0[2]; a1[0] = tm[0]; a1[1] = t1[1]; a1[2] = tm[2];
            process_subtree(L, a0, a1, a, N->chn[nla[2 ^ a]], l + 1);
            tn = next_node(a1, 6, 8, 3);
            break;
        case 3:
            a0[0] = t0[0]; a0[1] = tm[1]; a0[2] = tm[2]; a1[0] = tm[0]; a1[1] = t1[1]; a1[2] = t1[2];
            process_subtree(L, a0, a1, a, N->chn[nla[3 ^ a]], l + 1);
            tn = next_node(a1, 7, 8, 8);
            break;
        case 4:
            a0[0] = tm[0]; a0[1] = t0[1]; a0[2] = t0[2]; a1[0] = t1[0]; a1[1] = tm[1]; a1[2] = tm[2];
            process_subtree(L, a0, a1, a, N->chn[nla[4 ^ a]], l + 1);
            tn = next_node(a1, 8, 6, 5);
            break;
        case 5:
            a0[0] = tm[0]; a0[1] = t0[1]; a0[2] = tm[2]; a1[0] = t1[0]; a1[1] = tm[1]; a1[2] = t1[2];
            process_subtree(L, a0, a1, a, N->chn[nla[5 ^ a]], l + 1);
            tn = next_node(a1, 8, 7, 8);
            break;
        case 6:
            a0[0] = tm[0]; a0[1] = tm[1]; a0[2] = t0[2]; a1[0] = t1[0]; a1[1] = t1[1]; a1[2] = tm[2];
            process_subtree(L, a0, a1, a, N->chn[nla[6 ^ a]], l + 1);
            tn = next_node(a1, 8, 8, 7);
            break;
        case 7:
            process_subtree(L, tm, t1, a, N->chn[nla[7 ^ a]], l + 1);
            tn = 8;
            break;
        }
    } while (tn < 8);
}

/* OctreeInitialize (KernelOctree.cpp:78-110) + the root test of :137. */
static void traverse(olist *L, const float ro[3], const float rd[3]) {
    L->n = 0;
    const float *ce = L->ob[0].ce;
    float o[3], d[3];
    uint32_t a = 0;
    for (int k = 0; k < 3; ++k) {
        o[k] = (ro[k] - ce[k]) + ce[3];
        d[k] = rd[k];
    }
    if (d[0] < 0.0f) { o[0] = 2.0f * ce[3] - o[0]; d[0] = -d[0]; a |= 4; }
    if (d[1] < 0.0f) { o[1] = 2.0f * ce[3] - o[1]; d[1] = -d[1]; a |= 2; }
    if (d[2] < 0.0f) { o[2] = 2.0f * ce[3] - o[2]; d[2] = -d[2]; a |= 1; }
    float t0[3], t1[3], e2 = 2.0f * ce[3];
    for (int k = 0; k < 3; ++k) {
        t0[k] = (0.0f - o[k]) / d[k];
        t1[k] = (e2 - o[k]) / d[k];
    }
    if (fmax3(t0) < fmin3(t1)) process_subtree(L, t0, t1, a, 0, 0);
}

uint32_t or_traverse(const onode *ob, const float *ro, const float *rd, uint32_t n, uint32_t max_leaves,
                     uint32_t *nodes, float *tr0, float *tr1, uint32_t *count) {
    olist L = {ob, NULL, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        traverse(&L, ro + 3 * i, rd + 3 * i);
        count[i] = L.n;
        for (uint32_t j = 0; j < L.n && j < max_leaves; ++j) {
            nodes[(size_t)i * max_leaves + j] = L.ls[j].node;
            tr0[(size_t)i * max_leaves + j] = L.ls[j].t0;
            tr1[(size_t)i * max_leaves + j] = L.ls[j].t1;
        }
    }
    free(L.ls);
    return 0;
}

/* ------------------------------------------------------- brick lookup -- */
/* SampleVoxelSpace (kernel.cpp:96-126). */
float or_voxel(const float *pvol, const float p[3], const float ce[4], uint32_t lvoxc) {
    float nv = (float)lvoxc, nvm = nv - 1.0f, ni[3], nl[3];
    uint32_t a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        ni[k] = ((-0.5f * ((ce[k] - ce[3]) - p[k])) * nvm) / ce[3];
        float nf = fmaxf(floorf(ni[k]), 0.0f);
        float nc = fminf(ceilf(ni[k]), nvm);
        nl[k] = ni[k] - nf;
        nc = fmaxf(fminf(nc, (float)(lvoxc - 1)), 0.0f);
        nf = fmaxf(fminf(nf, (float)(lvoxc - 1)), 0.0f);
        a[k] = (uint32_t)nf;
        b[k] = (uint32_t)nc;
    }
    uint32_t l2 = lvoxc * lvoxc;
    float ua[4] = {pvol[a[2] * l2 + a[1] * lvoxc + a[0]], pvol[a[2] * l2 + b[1] * lvoxc + a[0]],
                   pvol[b[2] * l2 + a[1] * lvoxc + a[0]], pvol[b[2] * l2 + b[1] * lvoxc + a[0]]};
    float ub[4] = {pvol[a[2] * l2 + a[1] * lvoxc + b[0]], pvol[a[2] * l2 + b[1] * lvoxc + b[0]],
                   pvol[b[2] * l2 + a[1] * lvoxc + b[0]], pvol[b[2] * l2 + b[1] * lvoxc + b[0]]};
    float u[4];
    for (int k = 0; k < 4; ++k) u[k] = nl[0] * (ub[k] - ua[k]) + ua[k];
    float v0 = nl[1] * (u[1] - u[0]) + u[0];
    float v1 = nl[1] * (u[3] - u[2]) + u[2];
    return nl[2] * (v1 - v0) + v0;
}

/* ------------------------------------------------------------ samplers -- */
typedef struct {
    float pdf[256][3];
    float cdf[256][3];
} omie;

static void basis(const float iv[3], float b1[3], float b2[3]) {
    float len = sqrtf(iv[0] * iv[0] + 0.0f * 0.0f + iv[2] * iv[2]);
    b1[0] = -iv[2] / len; b1[1] = 0.0f / len; b1[2] = iv[0] / len;
    b2[0] = iv[1] * b1[2] - iv[2] * b1[1];
    b2[1] = iv[2] * b1[0] - iv[0] * b1[2];
    b2[2] = iv[0] * b1[1] - iv[1] * b1[0];
}

static void basis_dir(const float iv[3], float st, float ct, float ph, float out[3]) {
    float b1[3], b2[3];
    basis(iv, b1, b2);
    float sph = sinf(ph), cph = cosf(ph);
    for (int k = 0; k < 3; ++k) out[k] = b1[k] * st * cph + b2[k] * st * sph + iv[k] * ct;
}

/* MiePhase::EvaluateRGB (KernelSampler.cpp:233-256). */
void or_mie_eval(const omie *M, float ct, float out[3]) {
    float ct1 = fmaxf(fminf(ct, 1.0f), -1.0f);
    float a = acosf(-ct1) / 3.14159265358979f;
    float b = 256.0f * a;
    float c = floorf(b);
    c = fminf(c, 254.0f);
    c = fmaxf(c, 0.0f);
    float t = b - c;
    uint32_t i = (uint32_t)c;
    for (int k = 0; k < 3; ++k) out[k] = t * (M->pdf[i + 1][k] - M->pdf[i][k]) + M->pdf[i][k];
}

/* MiePhase::Sample's quantised inverse CDF (KernelSampler.cpp:258-273): the
 * linear scan over i = 0..254, keeping s for every i whose key is below u1
 * and stopping at the first that is not. Returns th/pi. */
float or_mie_theta(const omie *M, float u1) {
    float th = 0.0f;
    float ub = 256.0f * u1;
    for (uint32_t i = 0; i < 255; ++i) {
        float c0 = M->cdf[i][0], c1 = M->cdf[i + 1][0];
        float b = fmaxf(ub, (float)i);
        float c = fminf(floorf(b), 255.0f);
        float t = b - c;
        float cdf = t * (c1 - c0) + c0;
        if (!(cdf < u1)) break;
        th = (float)i / 256.0f + t / 256.0f;
    }
    return th;
}

void or_mie_sample(const omie *M, const float iv[3], float u1, float u2, float out[3]) {
    float th = or_mie_theta(M, u1) * 3.14159265358979f;
    float ph = (2.0f * 3.14159265358979f) * u2;
    float ct = -cosf(th);
    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
    basis_dir(iv, st, ct, ph, out);
}

float or_hg_eval(float g, float ct) {
    return (1.0f - g * g) / (4.0f * 3.14159265358979f * powf(1.0f + g * g - 2.0f * g * ct, 1.5f));
}

void or_hg_sample(float g, const float iv[3], float u1, float u2, float out[3]) {
    float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
    float ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
    float ph = (2.0f * 3.14159265358979f) * u2;
    basis_dir(iv, st, ct, ph, out);
}

typedef struct {
    float dir[3], color[3], cosang;
} olight;

static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* SunLight::Pdf (KernelSampler.cpp:326-330). */
static float sun_pdf(const olight *l, const float v[3]) {
    float ctm = l->cosang;
    return dot3(v, l->dir) > ctm ? 1.0f / (2.0f * 3.14159265358979f * (1.0f - ctm * ctm)) : 0.0f;
}

/* SunLight::Sample (KernelSampler.cpp:332-345). */
static void sun_sample(const olight *l, float u1, float u2, float out[3]) {
    float ctm = l->cosang;
    float ct = (1.0f - u1) + u1 * ctm;
    float st = sqrtf(1.0f - ct * ct);
    float ph = (2.0f * 3.14159265358979f) * u2;
    basis_dir(l->dir, st, ct, ph, out);
}

void or_sun(const float dir[3], const float color[3], float angle, const float *u1, const float *u2,
            const float *dirs, uint32_t n, float *osd, float *opdf, float *oev) {
    olight l;
    memcpy(l.dir, dir, 12);
    memcpy(l.color, color, 12);
    l.cosang = cosf(angle);
    for (uint32_t i = 0; i < n; ++i) {
        sun_sample(&l, u1[i], u2[i], osd + 3 * i);
        opdf[i] = sun_pdf(&l, dirs + 3 * i);
        int in = dot3(dirs + 3 * i, l.dir) > l.cosang;
        for (int k = 0; k < 3; ++k) oev[3 * i + k] = in ? l.color[k] : 0.0f;
    }
}

/* --------------------------------------------------------------- sky --- */
/* Hosek-Wilkie state (ArHosekSkyModel.cpp:147-289), double. */
static void hosek_cook(const double *dataset, double *config, double turbidity, double albedo, double elev) {
    int it = (int)turbidity;
    double tr = turbidity - (double)it;
    double x = pow(elev / (3.141592653589793 / 2.0), 1.0 / 3.0);
    double wa[4] = {(1.0 - albedo) * (1.0 - tr), albedo * (1.0 - tr), (1.0 - albedo) * tr, albedo * tr};
    int ai[4] = {0, 1, 0, 1}, ti[4] = {it - 1, it - 1, it, it};
    for (int i = 0; i < 9; ++i) config[i] = 0.0;
    for (int k = 0; k < (it == 10 ? 2 : 4); ++k) {
        const double *m = dataset + 9 * 6 * 10 * ai[k] + 9 * 6 * ti[k];
        for (int i = 0; i < 9; ++i)
            config[i] += wa[k] * (pow(1.0 - x, 5.0) * m[i] + 5.0 * pow(1.0 - x, 4.0) * x * m[i + 9] +
                                  10.0 * pow(1.0 - x, 3.0) * pow(x, 2.0) * m[i + 18] +
                                  10.0 * pow(1.0 - x, 2.0) * pow(x, 3.0) * m[i + 27] +
                                  5.0 * (1.0 - x) * pow(x, 4.0) * m[i + 36] + pow(x, 5.0) * m[i + 45]);
    }
}

static double hosek_cook_rad(const double *dataset, double turbidity, double albedo, double elev) {
    int it = (int)turbidity;
    double tr = turbidity - (double)it;
    double x = pow(elev / (3.141592653589793 / 2.0), 1.0 / 3.0);
    double wa[4] = {(1.0 - albedo) * (1.0 - tr), albedo * (1.0 - tr), (1.0 - albedo) * tr, albedo * tr};
    int ai[4] = {0, 1, 0, 1}, ti[4] = {it - 1, it - 1, it, it};
    double res = 0.0;
    for (int k = 0; k < (it == 10 ? 2 : 4); ++k) {
        const double *m = dataset + 6 * 10 * ai[k] + 6 * ti[k];
        res += wa[k] * (pow(1.0 - x, 5.0) * m[0] + 5.0 * pow(1.0 - x, 4.0) * x * m[1] +
                        10.0 * pow(1.0 - x, 3.0) * pow(x, 2.0) * m[2] + 10.0 * pow(1.0 - x, 2.0) * pow(x, 3.0) * m[3] +
                        5.0 * (1.0 - x) * pow(x, 4.0) * m[4] + pow(x, 5.0) * m[5]);
    }
    return res;
}

/* coeffs: 3 x 1080 doubles, rad: 3 x 120 doubles (the RGB dataset). */
void or_hosek_state(const double *coeffs, const double *rad, double turbidity, double albedo, double elev,
                    double *cfg27, double *rad3) {
    for (int ch = 0; ch < 3; ++ch) {
        hosek_cook(coeffs + 1080 * ch, cfg27 + 9 * ch, turbidity, albedo, elev);
        rad3[ch] = hosek_cook_rad(rad + 120 * ch, turbidity, albedo, elev);
    }
}

typedef struct {
    float cfg[3][9], rad[3], dir[3];
} osky;

/* src/kernel.cpp:329-360 for one direction, float. */
static void sky_eval(const osky *S, const float rd[3], float out[3]) {
    float rdz = rd[2];
    float sth = sqrtf(1.0f - rd[2] * rd[2]);
    float cth = rdz;
    float slth = sqrtf(1.0f - S->dir[2] * S->dir[2]);
    float clth = S->dir[2], slph = S->dir[1], clph = S->dir[0];
    float rdd = rd[0] / rd[1];
    float acr = 1.0f / sqrtf(rdd * rdd + 1.0f);
    /* Deviation from src/kernel.cpp:336-341, which is NaN here: at rd.y = +-0
       take the limit acr*rdd -> sign(rd.x)*sign(rd.y), and clamp the acos
       argument, which rounding can push past 1 for rd at the sun. */
    float czp;
    if (rd[1] == 0.0f) {
        float sx = rd[0] > 0.0f ? 1.0f : (rd[0] < 0.0f ? -1.0f : 0.0f);
        czp = (signbit(rd[1]) ? -sx : sx) * slph;
    } else {
        czp = acr * (rdd * slph + clph);
    }
    float cph = sth * slth * czp + cth * clth;
    float gmma = acosf(fmaxf(fminf(cph, 1.0f), -1.0f));
    float gacs = fmaxf(fminf(cph, 1.0f), 0.0f);
    float thcs = fmaxf(rdz, 0.0f);
    float raym = gacs * gacs;
    for (int i = 0; i < 3; ++i) {
        const float *c = S->cfg[i];
        float expm = expf(c[4] * gmma);
        float miem = (1.0f + raym) / powf(1.0f + c[8] * c[8] - 2.0f * c[8] * gacs, 1.5f);
        float zenh = sqrtf(thcs);
        float ca = (1.0f + c[0] * expf(c[1] / (thcs + 0.01f))) * (c[2] + c[3] * expm + c[5] * raym + c[6] * miem + c[7] * zenh);
        ca *= S->rad[i];
        out[i] = 1e-3f * powf(ca, 2.2f);
    }
}

void or_sky(const float cfg27[27], const float rad3[3], const float dir[3], const float *rd, uint32_t n, float *out) {
    osky S;
    memcpy(S.cfg, cfg27, sizeof(S.cfg));
    memcpy(S.rad, rad3, 12);
    memcpy(S.dir, dir, 12);
    for (uint32_t i = 0; i < n; ++i) sky_eval(&S, rd + 3 * i, out + 3 * i);
}

/* MapEnv::Evaluate (KernelSampler.cpp:369-421) then pow 2.2. */
static void map_env(const float *tex, uint32_t w, uint32_t h, const float rd[3], float out[3]) {
    float tha = acosf(-rd[2]) / 3.14159265358979f;
    float thb = (float)h * fmaxf(fminf(tha, 1.0f), 0.0f);
    float thc = fmaxf(fminf(floorf(thb), (float)(h - 2)), 0.0f);
    float tht = thb - thc;
    float pha = 0.5f + atan2f(rd[1], -rd[0]) / (2.0f * 3.14159265358979f);
    float phb = (float)w * fmaxf(fminf(pha, 1.0f), 0.0f);
    float phc = fmaxf(fminf(floorf(phb), (float)(w - 2)), 0.0f);
    float pht = phb - phc;
    uint32_t ti = (uint32_t)thc, pj = (uint32_t)phc;
    for (int i = 0; i < 3; ++i) {
        float ca = tex[(ti * w + pj) * 4 + i], cb = tex[((ti + 1) * w + pj) * 4 + i];
        float da = tex[(ti * w + pj + 1) * 4 + i], db = tex[((ti + 1) * w + pj + 1) * 4 + i];
        float l0 = tht * (cb - ca) + ca, l1 = tht * (db - da) + da;
        out[i] = powf(pht * (l1 - l0) + l0, 2.2f);
    }
}

void or_map_env(const float *tex, uint32_t w, uint32_t h, const float *rd, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) map_env(tex, w, h, rd + 3 * i, out + 3 * i);
}

/* ----------------------------------------------------------- renderer -- */
typedef struct {
    const onode *ob;
    const float *pvol_fog;
    const float *pvol_sdf;   /* NULL without SDF leaves */
    uint32_t lvoxc, lvoxc3;
    float viewi[16], proji[16];
    uint32_t w, h;
    uint32_t scattevs;
    float sigma_s, sigma_a;
    uint32_t phase;          /* 0 Mie, 1 HG */
    float hg_g;
    uint32_t env;            /* 0 Hosek, 1 null, 2 map */
    const float *env_map;
    uint32_t env_w, env_h;
    uint32_t light_count;
    olight lights[8];
    osky sky;
    const omie *mie;
    const float *depth;      /* w*h or NULL */
    uint32_t key1;
} oparams;

typedef struct {
    const oparams *P;
    olist L;
    uint32_t key0, sample;
} octx;

/* Result of one SampleVolume call for one lane: cl/cs RGBA sums. */
typedef struct { float cl[4], cs[4]; } ores;

static void env_eval(const oparams *P, const float rd[3], float le[3]) {
    if (P->env == 0) sky_eval(&P->sky, rd, le);
    else if (P->env == 2) map_env(P->env_map, P->env_w, P->env_h, rd, le);
    else le[0] = le[1] = le[2] = 0.0f;
}

/* Leaf loop + Woodcock loop of SampleVolume (kernel.cpp:197-317) for one lane
 * with FOG-only leaves. Returns 1 if the ray collided (rm == false), with the
 * distance in *ptd. */
static int track(octx *X, const float ro[3], const float rd[3], uint32_t ray_id, float depth, float *ptd) {
    const oparams *P = X->P;
    ostream rng;
    stream_seed(&rng, X->key0, P->key1, X->sample, ray_id);
    traverse(&X->L, ro, rd);
    const float sige = P->sigma_a + P->sigma_s;
    float td = 0.0f;
    int rm = 1;
    for (uint32_t i = 0;; ++i) {
        if (i >= X->L.n) break;          /* lm == 0 */
        if (!(td < depth)) break;        /* mm == 0 */
        const oleaf *lf = &X->L.ls[i];
        const onode *N = &P->ob[lf->node];
        float tra = lf->t0, trb = lf->t1;
        float lo = td;
        float tr0 = tra - td, tr1 = trb - td;
        const int sdf = N->volx[0] != 0xFFFFFFFFu;
        const float *fogb = N->volx[1] != 0xFFFFFFFFu ? P->pvol_fog + (size_t)N->volx[1] * P->lvoxc3 : NULL;
        const float *sdfb = sdf ? P->pvol_sdf + (size_t)N->volx[0] * P->lvoxc3 : NULL;
        /* kernel.cpp:245-269: an SDF leaf has majorant 1 and is entered at
           its entry point only when that lies ahead (tra > td) and outside
           the surface (distance there > 0); a sole fog leaf always is. */
        float smax = sdf ? 1.0f : N->qval[1];
        int skip = 1;
        if (sdf) {
            skip = 0;
            if (tra > td) {
                float r0[3] = {ro[0] + rd[0] * tra, ro[1] + rd[1] * tra, ro[2] + rd[2] * tra};
                skip = or_voxel(sdfb, r0, N->ce, P->lvoxc) > 0.0f;
            }
        }
        float s0 = 0.0f;
        if (skip) {
            td = tra;
            s0 = tr0;
        }
        float sr = -(logf(stream_u(&rng)) / (sige * smax));
        int sm = 1;
        for (;;) {
            if (!(sm && rm)) break;
            float sc = s0 + sr;
            td = lo + sc;
            sm = sc < tr1;
            if (!sm) td = trb;
            if (sm) rm = sc > tr0;
            if (sm && rm) {
                float rc[3] = {ro[0] + rd[0] * td, ro[1] + rd[1] * td, ro[2] + rd[2] * td};
                /* kernel.cpp:292-315: inside the surface is a collision;
                   outside, the fog if any (else -1: never accepted). */
                float d = 1.0f, p;
                if (sdf) {
                    d = or_voxel(sdfb, rc, N->ce, P->lvoxc);
                    p = (d > 0.0f && fogb) ? or_voxel(fogb, rc, N->ce, P->lvoxc) : -1.0f;
                } else {
                    p = or_voxel(fogb, rc, N->ce, P->lvoxc);
                }
                float q = stream_u(&rng);
                rm = (p / smax < q) && (d > 0.0f);
                if (rm) sr = sr - logf(stream_u(&rng)) / (sige * smax);
            }
        }
        if (!rm) break;
    }
    *ptd = td;
    return !rm;
}

/* SampleVolume (kernel.cpp:128-445) for one lane, samples = 1, depth r,
 * tracking ray `ray_id`. */
static void sample_volume(octx *X, const float ro[3], const float rd[3], uint32_t r, uint32_t ray_id, float depth,
                          ores *out) {
    const oparams *P = X->P;
    memset(out, 0, sizeof(*out));
    float td;
    int scattered = track(X, ro, rd, ray_id, depth, &td);
    int rm = !scattered;
    float lc[4] = {0, 0, 0, 0}, le[4] = {0, 0, 0, 0};
    if (rm && r > 0) {
        for (uint32_t l = 0; l < P->light_count; ++l)
            if (dot3(rd, P->lights[l].dir) > P->lights[l].cosang)
                for (int k = 0; k < 3; ++k) lc[k] += P->lights[l].color[k];
        env_eval(P, rd, le);
        lc[3] = le[3] = 1.0f;
    } else {
        lc[3] = le[3] = rm ? 0.0f : 1.0f;
    }
    if (r < P->scattevs && !rm) {
        uint32_t c[4] = {X->sample, r, 0u, 1u}, v[4];
        philox(c, X->key0, P->key1, v);
        float u1 = u01(v[0]), u2 = u01(v[1]), u3 = u01(v[2]), u4 = u01(v[3]);
        float srd[3], lrd[3], p1[3], p3[3];
        if (P->phase == 0) or_mie_sample(P->mie, rd, u1, u2, srd);
        else or_hg_sample(P->hg_g, rd, u1, u2, srd);
        sun_sample(&P->lights[0], u3, u4, lrd);
        if (P->phase == 0) {
            or_mie_eval(P->mie, dot3(srd, rd), p1);
            or_mie_eval(P->mie, dot3(lrd, rd), p3);
        } else {
            float h1 = or_hg_eval(P->hg_g, dot3(srd, rd)), h3 = or_hg_eval(P->hg_g, dot3(lrd, rd));
            p1[0] = p1[1] = p1[2] = h1;
            p3[0] = p3[1] = p3[2] = h3;
        }
        float p2 = sun_pdf(&P->lights[0], lrd);
        float rc[3] = {ro[0] + rd[0] * td, ro[1] + rd[1] * td, ro[2] + rd[2] * td};
        ores S1, S2;
        sample_volume(X, rc, srd, r + 1, 2u * r + 2u, FLT_MAX, &S1);
        sample_volume(X, rc, lrd, P->scattevs, 2u * r + 1u, FLT_MAX, &S2);
        float pd1 = sun_pdf(&P->lights[0], srd);
        float sige = P->sigma_a + P->sigma_s;
        for (int k = 0; k < 3; ++k) {
            float w1 = p1[k] / (p1[k] + pd1);
            float w2 = p3[k] / (p3[k] + p2);
            float cl1 = ((S1.cl[k] * w1 + S2.cl[k] * w2) * P->sigma_s) / sige;
            float cs1 = ((S1.cs[k] * w1 + 0.0f * w2) * P->sigma_s) / sige;
            out->cl[k] += cl1;
            out->cs[k] += cs1;
        }
        out->cl[3] += lc[3];
        out->cs[3] += le[3];
    } else {
        for (int k = 0; k < 3; ++k) {
            out->cl[k] += rm ? lc[k] : 0.0f;
            out->cs[k] += rm ? le[k] : 0.0f;
        }
        out->cl[3] += lc[3];
        out->cs[3] += le[3];
    }
}

/* Primary ray of pixel (px,py), K_ParallelRender (kernel.cpp:461-474). */
static int camera_ray(const oparams *P, uint32_t px, uint32_t py, float ro[3], float rd[3]) {
    float hx = -(2.0f * ((float)px + 0.5f) / (float)P->w - 1.0f);
    float hy = -(2.0f * ((float)py + 0.5f) / (float)P->h - 1.0f);
    const float *pi = P->proji, *vi = P->viewi;
    float rv[3], d[3];
    for (int j = 0; j < 3; ++j) rv[j] = hx * pi[j] + hy * pi[4 + j] + 0.0f * pi[8 + j] + 1.0f * pi[12 + j];
    for (int j = 0; j < 3; ++j) d[j] = rv[0] * vi[j] + rv[1] * vi[4 + j] + rv[2] * vi[8 + j] + 0.0f * vi[12 + j];
    float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int j = 0; j < 3; ++j) {
        rd[j] = d[j] / len;
        ro[j] = vi[12 + j];
    }
    return hx > -1.0f && hx < 1.0f && hy > -1.0f && hy < 1.0f;
}

void or_camera(const float viewi[16], const float proji[16], uint32_t w, uint32_t h, const uint32_t *px,
               const uint32_t *py, uint32_t n, float *ro, float *rd, uint32_t *gm) {
    oparams P;
    memset(&P, 0, sizeof(P));
    memcpy(P.viewi, viewi, 64);
    memcpy(P.proji, proji, 64);
    P.w = w;
    P.h = h;
    for (uint32_t i = 0; i < n; ++i) gm[i] = (uint32_t)camera_ray(&P, px[i], py[i], ro + 3 * i, rd + 3 * i);
}

typedef struct {
    const oparams *P;
    uint32_t x0, y0, rx, ry, tilex, samples, sample_base;
    float *out;              /* 4 buffers of tilex*tiley*4: lights, env, sumsq lights, sumsq env */
    size_t plane;
    uint32_t next_row;
    pthread_mutex_t mu;
} ojob;

static void render_pixel(ojob *J, octx *X, uint32_t lx, uint32_t ly) {
    const oparams *P = J->P;
    uint32_t px = J->x0 + lx, py = J->y0 + ly;
    float ro[3], rd[3];
    int gm = camera_ray(P, px, py, ro, rd);
    size_t o = ((size_t)ly * J->tilex + lx) * 4;
    if (!gm) return;
    X->key0 = py * P->w + px;
    float depth = P->depth ? P->depth[X->key0] : FLT_MAX;
    double acc[4][4] = {{0}};
    for (uint32_t s = 0; s < J->samples; ++s) {
        X->sample = J->sample_base + s;
        ores R;
        sample_volume(X, ro, rd, 0, 0u, depth, &R);
        for (int k = 0; k < 4; ++k) {
            acc[0][k] += R.cl[k];
            acc[1][k] += R.cs[k];
            acc[2][k] += k < 3 ? (double)R.cl[k] * R.cl[k] : R.cl[k];
            acc[3][k] += k < 3 ? (double)R.cs[k] * R.cs[k] : R.cs[k];
        }
    }
    for (int b = 0; b < 4; ++b)
        for (int k = 0; k < 4; ++k) J->out[b * J->plane + o + k] = (float)acc[b][k];
}

static void *render_worker(void *arg) {
    ojob *J = (ojob *)arg;
    octx X;
    memset(&X, 0, sizeof(X));
    X.P = J->P;
    X.L.ob = J->P->ob;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        uint32_t row = J->next_row++;
        pthread_mutex_unlock(&J->mu);
        if (row >= J->ry) break;
        for (uint32_t lx = 0; lx < J->rx; ++lx) render_pixel(J, &X, lx, row);
    }
    free(X.L.ls);
    return NULL;
}

/*
 * Renders tile (x0,y0,tilex,tiley) with `samples` paths per pixel into out
 * (4 * tilex*tiley*4 floats: lights RGBA sums, environment RGBA sums, and the
 * per-pixel sums of squares of the lights / environment RGB, alpha repeated).
 * Pixels of the dropped odd row/column stay 0. Uses `threads` POSIX threads.
 * Parameter block layout matches oparams above (filled by pyoracle.py).
 */
int or_render(const oparams *P, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
              uint32_t sample_base, uint32_t threads, float *out) {
    ojob J;
    memset(&J, 0, sizeof(J));
    J.P = P;
    J.x0 = x0; J.y0 = y0;
    J.rx = tilex & ~1u; J.ry = tiley & ~1u;
    if (x0 + J.rx > P->w) J.rx = x0 < P->w ? P->w - x0 : 0;
    if (y0 + J.ry > P->h) J.ry = y0 < P->h ? P->h - y0 : 0;
    J.tilex = tilex;
    J.samples = samples;
    J.sample_base = sample_base;
    J.out = out;
    J.plane = (size_t)tilex * tiley * 4;
    memset(out, 0, sizeof(float) * 4 * J.plane);
    pthread_mutex_init(&J.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (uint32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, render_worker, &J);
    for (uint32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J.mu);
    return 0;
}

/* -------------------------------------------------------- packet path -- */
/*
 * The reference's CPU structure for the same estimator. K_ParallelRender
 * (kernel.cpp:447-487) traces 2x2 pixel packets in 4-wide SSE lanes
 * (BLCLOUD_VSIZE 4, BLCLOUD_VX = BLCLOUD_VY = 2, smmath.inl:4-6), and
 * SampleVolume (kernel.cpp:128-445) runs the leaf loop, the Woodcock loop
 * and both recursive calls for the whole packet in lockstep under lane
 * masks: a lane that has collided, left its last leaf or is masked off
 * idles until the packet's last lane is done (qm/sm/rm, kernel.cpp:206-316),
 * and the continuation and shadow calls run for the packet with the mask of
 * the lanes that scattered (gm1, kernel.cpp:393-406).
 *
 * Arithmetic that is the same for every lane (free-flight lengths, the
 * positions, the xoshiro128** streams, the MIS combination) is 4-wide; the
 * per-lane work (traversal lists, brick gathers, the Mie scan, libm
 * transcendentals) is a loop over the active lanes, as the reference's
 * gathers are (kernel.cpp:223-258, 287-310). A lane draws from its own
 * stream only while it is active and rounds every operation as the
 * one-lane restatement above does, so or_render_packets returns or_render's
 * sums bit for bit; it is the CPU baseline in the reference's structure.
 */
typedef float v4f __attribute__((vector_size(16)));
typedef int32_t v4i __attribute__((vector_size(16)));
typedef uint32_t v4u __attribute__((vector_size(16)));

static inline v4f vsel(v4i m, v4f a, v4f b) { return (v4f)(((v4i)a & m) | ((v4i)b & ~m)); }
static inline v4i isel(v4i m, v4i a, v4i b) { return (a & m) | (b & ~m); }
static inline v4u usel(v4i m, v4u a, v4u b) { return (a & (v4u)m) | (b & ~(v4u)m); }
static inline int vany(v4i m) { return (m[0] | m[1] | m[2] | m[3]) != 0; }
static inline v4f vsplat(float x) { return (v4f){x, x, x, x}; }

typedef struct { v4u s[4]; } pstream;

static void pstream_seed(pstream *r, const uint32_t key0[4], uint32_t key1, uint32_t sample, uint32_t ray) {
    for (int j = 0; j < 4; ++j) {
        ostream o;
        stream_seed(&o, key0[j], key1, sample, ray);
        for (int k = 0; k < 4; ++k) r->s[k][j] = o.s[k];
    }
}

/* One xoshiro128** step (stream_next) of the lanes in m; the other lanes keep
 * their state. Returns the uniforms of every lane (u01). */
static v4f pstream_u(pstream *r, v4i m) {
    v4u *s = r->s;
    v4u x = s[1] * 5u;
    v4u res = ((x << 7) | (x >> 25)) * 9u;
    v4u t = s[1] << 9;
    v4u n2 = s[2] ^ s[0], n3 = s[3] ^ s[1];
    v4u n1 = s[1] ^ n2, n0 = s[0] ^ n3;
    n2 ^= t;
    n3 = (n3 << 11) | (n3 >> 21);
    s[0] = usel(m, n0, s[0]);
    s[1] = usel(m, n1, s[1]);
    s[2] = usel(m, n2, s[2]);
    s[3] = usel(m, n3, s[3]);
    v4i q = (v4i)((res >> 8) | 1u);
    return __builtin_convertvector(q, v4f) * vsplat(5.9604644775390625e-8f);
}

/* logf of the lanes in m (0 elsewhere). */
static v4f plog(v4f u, v4i m) {
    v4f r = vsplat(0.0f);
    for (int j = 0; j < 4; ++j)
        if (m[j]) r[j] = logf(u[j]);
    return r;
}

/* OctreeStepTraverser (KernelOctree.cpp:222-349): the Revelles walk of
 * process_subtree() with an explicit stack, one leaf per call, so a ray pays
 * only for the leaves it reaches. Same leaves, same intervals, same order. */
typedef struct { float t0[3], t1[3], tm[3]; uint32_t n, tn; } pframe;
typedef struct {
    pframe st[32];
    int sp;                   /* frames on the stack */
    uint32_t a;
    int root_leaf;            /* the root itself is the only leaf */
} pwalk;

static void pwalk_push(const onode *ob, pwalk *W, const float t0[3], const float t1[3], uint32_t n) {
    pframe *F = &W->st[W->sp++];
    for (int k = 0; k < 3; ++k) {
        F->t0[k] = t0[k];
        F->t1[k] = t1[k];
        F->tm[k] = 0.5f * (t0[k] + t1[k]);
    }
    F->n = n;
    F->tn = first_node(F->t0, F->tm);
    (void)ob;
}

/* OctreeInitialize (KernelOctree.cpp:78-110) + the root test. */
static void pwalk_init(const onode *ob, pwalk *W, const float ro[3], const float rd[3]) {
    const float *ce = ob[0].ce;
    float o[3], d[3];
    uint32_t a = 0;
    for (int k = 0; k < 3; ++k) {
        o[k] = (ro[k] - ce[k]) + ce[3];
        d[k] = rd[k];
    }
    if (d[0] < 0.0f) { o[0] = 2.0f * ce[3] - o[0]; d[0] = -d[0]; a |= 4; }
    if (d[1] < 0.0f) { o[1] = 2.0f * ce[3] - o[1]; d[1] = -d[1]; a |= 2; }
    if (d[2] < 0.0f) { o[2] = 2.0f * ce[3] - o[2]; d[2] = -d[2]; a |= 1; }
    float t0[3], t1[3], e2 = 2.0f * ce[3];
    for (int k = 0; k < 3; ++k) {
        t0[k] = (0.0f - o[k]) / d[k];
        t1[k] = (e2 - o[k]) / d[k];
    }
    W->sp = 0;
    W->a = a;
    W->root_leaf = 0;
    if (!(fmax3(t0) < fmin3(t1)) || t1[0] < 0.0f || t1[1] < 0.0f || t1[2] < 0.0f) return;
    if (is_leaf(&ob[0])) {
        W->root_leaf = 1;
        memcpy(W->st[0].t0, t0, 12);
        memcpy(W->st[0].t1, t1, 12);
        return;
    }
    pwalk_push(ob, W, t0, t1, 0);
}

/* The next leaf of the walk: 1 with (node, entry, exit), or 0 at the end. */
static int pwalk_next(const onode *ob, pwalk *W, uint32_t *node, float *ta, float *tb) {
    static const uint32_t nla[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    if (W->root_leaf) {
        W->root_leaf = 0;
        *node = 0;
        *ta = fmax3(W->st[0].t0);
        *tb = fmin3(W->st[0].t1);
        return 1;
    }
    while (W->sp > 0) {
        pframe *F = &W->st[W->sp - 1];
        if (F->tn >= 8) {
            W->sp--;
            continue;
        }
        const float *t0 = F->t0, *t1 = F->t1, *tm = F->tm;
        float a0[3], a1[3];
        uint32_t tn = F->tn;
        switch (tn) {
        case 0:
            memcpy(a0, t0, 12); memcpy(a1, tm, 12);
            F->tn = next_node(tm, 4, 2, 1);
            break;
        case 1:
            a0[0] = t0[0]; a0[1] = t0[1]; a0[2] = tm[2]; a1[0] = tm[0]; a1[1] = tm[1]; a1[2] = t1[2];
            F->tn = next_node(a1, 5, 3, 8);
            break;
        case 2:
            a0[0] = t0[0]; a0[1] = tm[1]; a0[2] = t0[2]; a1[0] = tm[0]; a1[1] = t1[1]; a1[2] = tm[2];
            F->tn = next_node(a1, 6, 8, 3);
            break;
        case 3:
            a0[0] = t0[0]; a0[1] = tm[1]; a0[2] = tm[2]; a1[0] = tm[0]; a1[1] = t1[1]; a1[2] = t1[2];
            F->tn = next_node(a1, 7, 8, 8);
            break;
        case 4:
            a0[0] = tm[0]; a0[1] = t0[1]; a0[2] = t0[2]; a1[0] = t1[0]; a1[1] = tm[1]; a1[2] = tm[2];
            F->tn = next_node(a1, 8, 6, 5);
            break;
        case 5:
            a0[0] = tm[0]; a0[1] = t0[1]; a0[2] = tm[2]; a1[0] = t1[0]; a1[1] = tm[1]; a1[2] = t1[2];
            F->tn = next_node(a1, 8, 7, 8);
            break;
        case 6:
            a0[0] = tm[0]; a0[1] = tm[1]; a0[2] = t0[2]; a1[0] = t1[0]; a1[1] = t1[1]; a1[2] = tm[2];
            F->tn = next_node(a1, 8, 8, 7);
            break;
        default:
            memcpy(a0, tm, 12); memcpy(a1, t1, 12);
            F->tn = 8;
            break;
        }
        uint32_t c = ob[F->n].chn[nla[tn ^ W->a]];
        if (a1[0] < 0.0f || a1[1] < 0.0f || a1[2] < 0.0f || c == 0) continue;
        if (is_leaf(&ob[c])) {
            *node = c;
            *ta = fmax3(a0);
            *tb = fmin3(a1);
            return 1;
        }
        if (W->sp == (int)(sizeof(W->st) / sizeof(W->st[0]))) abort();
        pwalk_push(ob, W, a0, a1, c);
    }
    return 0;
}

typedef struct {
    const oparams *P;
    olist L[4];               /* the camera rays' leaves (full traverser, kept for every sample) */
    pwalk W[4];               /* secondary rays (step traverser) */
    uint32_t key0[4], sample;
} pctx;

/* The leaf and Woodcock loops (kernel.cpp:197-317) for the lanes in gm;
 * track() for each lane. Returns the mask of the lanes that collided. */
static v4i ptrack(pctx *X, const v4f ro[3], const v4f rd[3], v4i gm, uint32_t ray_id, v4f depth, v4f *ptd) {
    const oparams *P = X->P;
    pstream rng;
    pstream_seed(&rng, X->key0, P->key1, X->sample, ray_id);
    /* Camera rays walk the leaf lists made once per packet; the others step
     * their traversal along, every lane that still has leaves advancing with
     * each round of the leaf loop (GetLeaf, kernel.cpp:206-218). */
    const int primary = ray_id == 0u;
    v4i lm = gm;
    if (!primary)
        for (int j = 0; j < 4; ++j)
            if (gm[j]) {
                float o[3] = {ro[0][j], ro[1][j], ro[2][j]}, d[3] = {rd[0][j], rd[1][j], rd[2][j]};
                pwalk_init(P->ob, &X->W[j], o, d);
            }
    const v4f sige = vsplat(P->sigma_a + P->sigma_s);
    v4f td = vsplat(0.0f);
    v4i rm = {-1, -1, -1, -1};
    for (uint32_t i = 0;; ++i) {
        uint32_t lnode[4] = {0, 0, 0, 0};
        float lt0[4] = {0, 0, 0, 0}, lt1[4] = {0, 0, 0, 0};
        for (int j = 0; j < 4; ++j) {
            if (!lm[j]) continue;
            if (primary) {
                if (i < X->L[j].n) {
                    lnode[j] = X->L[j].ls[i].node;
                    lt0[j] = X->L[j].ls[i].t0;
                    lt1[j] = X->L[j].ls[i].t1;
                } else {
                    lm[j] = 0;
                }
            } else if (!pwalk_next(P->ob, &X->W[j], &lnode[j], &lt0[j], &lt1[j])) {
                lm[j] = 0;
            }
        }
        v4i qm = gm & rm & lm & (td < depth);
        if (!vany(qm)) break;
        v4f tra = vsplat(0.0f), trb = vsplat(0.0f), smax = vsplat(1.0f);
        v4i skip = {0, 0, 0, 0};
        const onode *N[4] = {NULL, NULL, NULL, NULL};
        const float *fogb[4] = {NULL, NULL, NULL, NULL}, *sdfb[4] = {NULL, NULL, NULL, NULL};
        for (int j = 0; j < 4; ++j) {
            if (!qm[j]) continue;
            const onode *Nj = &P->ob[lnode[j]];
            N[j] = Nj;
            tra[j] = lt0[j];
            trb[j] = lt1[j];
            const int sdf = Nj->volx[0] != ONO;
            fogb[j] = Nj->volx[1] != ONO ? P->pvol_fog + (size_t)Nj->volx[1] * P->lvoxc3 : NULL;
            sdfb[j] = sdf ? P->pvol_sdf + (size_t)Nj->volx[0] * P->lvoxc3 : NULL;
            smax[j] = sdf ? 1.0f : Nj->qval[1];
            int sk = 1;
            if (sdf) {
                sk = 0;
                if (tra[j] > td[j]) {
                    float r0[3] = {ro[0][j] + rd[0][j] * tra[j], ro[1][j] + rd[1][j] * tra[j],
                                   ro[2][j] + rd[2][j] * tra[j]};
                    sk = or_voxel(sdfb[j], r0, Nj->ce, P->lvoxc) > 0.0f;
                }
            }
            skip[j] = sk ? -1 : 0;
        }
        const v4f lo = td;
        const v4f tr0 = tra - td, tr1 = trb - td;
        skip &= qm;
        td = vsel(skip, tra, td);
        const v4f s0 = vsel(skip, tr0, vsplat(0.0f));
        const v4f den = sige * smax;
        v4i sm = qm;
        v4f sr = -(plog(pstream_u(&rng, sm), sm) / den);
        for (;;) {
            sm &= rm;
            if (!vany(sm)) break;
            v4f sc = s0 + sr;
            td = vsel(sm, lo + sc, td);
            v4i inb = sc < tr1;
            td = vsel(sm & ~inb, trb, td);
            sm &= inb;
            rm = isel(sm, sc > tr0, rm);
            v4i act = sm & rm;
            if (!vany(act)) continue;
            v4f rc[3] = {ro[0] + rd[0] * td, ro[1] + rd[1] * td, ro[2] + rd[2] * td};
            v4f p = vsplat(-1.0f), d = vsplat(1.0f);
            for (int j = 0; j < 4; ++j) {
                if (!act[j]) continue;
                float x[3] = {rc[0][j], rc[1][j], rc[2][j]};
                if (sdfb[j]) {
                    d[j] = or_voxel(sdfb[j], x, N[j]->ce, P->lvoxc);
                    p[j] = (d[j] > 0.0f && fogb[j]) ? or_voxel(fogb[j], x, N[j]->ce, P->lvoxc) : -1.0f;
                } else {
                    p[j] = or_voxel(fogb[j], x, N[j]->ce, P->lvoxc);
                }
            }
            v4f q = pstream_u(&rng, act);
            rm = isel(act, (p / smax < q) & (d > 0.0f), rm);
            v4i cont = act & rm;
            if (vany(cont)) sr = vsel(cont, sr - plog(pstream_u(&rng, cont), cont) / den, sr);
        }
    }
    *ptd = td;
    return gm & ~rm;
}

/* SampleVolume (kernel.cpp:128-445) for a packet with lane mask gm, one
 * sample, depth r; sample_volume() for each lane of gm. */
static void psample_volume(pctx *X, const v4f ro[3], const v4f rd[3], v4i gm, uint32_t r, uint32_t ray_id,
                           v4f depth, v4f cl[4], v4f cs[4]) {
    const oparams *P = X->P;
    const v4f zero = vsplat(0.0f), one = vsplat(1.0f);
    v4f td;
    const v4i scat = ptrack(X, ro, rd, gm, ray_id, depth, &td);
    const v4i rm = ~scat;
    v4f lc[4] = {zero, zero, zero, zero}, le[4] = {zero, zero, zero, zero};
    const v4i esc = gm & rm;
    if (r > 0) {
        for (int j = 0; j < 4; ++j) {
            if (!esc[j]) continue;
            float d[3] = {rd[0][j], rd[1][j], rd[2][j]}, l3[3] = {0.0f, 0.0f, 0.0f}, e3[3];
            for (uint32_t l = 0; l < P->light_count; ++l)
                if (dot3(d, P->lights[l].dir) > P->lights[l].cosang)
                    for (int k = 0; k < 3; ++k) l3[k] += P->lights[l].color[k];
            env_eval(P, d, e3);
            for (int k = 0; k < 3; ++k) {
                lc[k][j] = l3[k];
                le[k][j] = e3[k];
            }
        }
        lc[3] = le[3] = one;
    } else {
        lc[3] = le[3] = vsel(rm, zero, one);
    }
    if (r < P->scattevs && vany(scat)) {
        v4f srd[3] = {zero, zero, one}, lrd[3] = {zero, zero, one}, p1[3] = {one, one, one}, p3[3] = {one, one, one};
        v4f p2 = zero, pd1 = zero;
        for (int j = 0; j < 4; ++j) {
            if (!scat[j]) continue;
            uint32_t c[4] = {X->sample, r, 0u, 1u}, v[4];
            philox(c, X->key0[j], P->key1, v);
            float u1 = u01(v[0]), u2 = u01(v[1]), u3 = u01(v[2]), u4 = u01(v[3]);
            float d[3] = {rd[0][j], rd[1][j], rd[2][j]}, s[3], l[3], q1[3], q3[3];
            if (P->phase == 0) or_mie_sample(P->mie, d, u1, u2, s);
            else or_hg_sample(P->hg_g, d, u1, u2, s);
            sun_sample(&P->lights[0], u3, u4, l);
            if (P->phase == 0) {
                or_mie_eval(P->mie, dot3(s, d), q1);
                or_mie_eval(P->mie, dot3(l, d), q3);
            } else {
                float h1 = or_hg_eval(P->hg_g, dot3(s, d)), h3 = or_hg_eval(P->hg_g, dot3(l, d));
                q1[0] = q1[1] = q1[2] = h1;
                q3[0] = q3[1] = q3[2] = h3;
            }
            p2[j] = sun_pdf(&P->lights[0], l);
            pd1[j] = sun_pdf(&P->lights[0], s);
            for (int k = 0; k < 3; ++k) {
                srd[k][j] = s[k];
                lrd[k][j] = l[k];
                p1[k][j] = q1[k];
                p3[k][j] = q3[k];
            }
        }
        v4f rc[3] = {ro[0] + rd[0] * td, ro[1] + rd[1] * td, ro[2] + rd[2] * td};
        v4f S1l[4], S1s[4], S2l[4], S2s[4];
        psample_volume(X, rc, srd, scat, r + 1, 2u * r + 2u, vsplat(FLT_MAX), S1l, S1s);
        psample_volume(X, rc, lrd, scat, P->scattevs, 2u * r + 1u, vsplat(FLT_MAX), S2l, S2s);
        const v4f sigs = vsplat(P->sigma_s), sige = vsplat(P->sigma_a + P->sigma_s);
        for (int k = 0; k < 3; ++k) {
            v4f w1 = p1[k] / (p1[k] + pd1);
            v4f w2 = p3[k] / (p3[k] + p2);
            v4f cl1 = ((S1l[k] * w1 + S2l[k] * w2) * sigs) / sige;
            v4f cs1 = ((S1s[k] * w1 + zero * w2) * sigs) / sige;
            cl[k] = zero + vsel(scat, cl1, vsel(rm, lc[k], zero));
            cs[k] = zero + vsel(scat, cs1, vsel(rm, le[k], zero));
        }
    } else {
        for (int k = 0; k < 3; ++k) {
            cl[k] = zero + vsel(rm, lc[k], zero);
            cs[k] = zero + vsel(rm, le[k], zero);
        }
    }
    cl[3] = zero + lc[3];
    cs[3] = zero + le[3];
}

/* One 2x2 packet (K_ParallelRender, kernel.cpp:453-476): lanes (0,0), (1,0),
 * (0,1), (1,1) of the packet at tile-local pixel (2*qx, 2*qy); lanes past the
 * tile's rendered extent are masked off like those outside the image. */
static void render_packet(ojob *J, pctx *X, uint32_t qx, uint32_t qy) {
    const oparams *P = J->P;
    v4f ro[3], rd[3], depth;
    v4i gm;
    size_t o[4];
    for (int j = 0; j < 4; ++j) {
        uint32_t lx = 2u * qx + (uint32_t)(j & 1), ly = 2u * qy + (uint32_t)(j >> 1);
        uint32_t px = J->x0 + lx, py = J->y0 + ly;
        float a[3], b[3];
        gm[j] = camera_ray(P, px, py, a, b) && lx < J->rx && ly < J->ry ? -1 : 0;
        for (int k = 0; k < 3; ++k) {
            ro[k][j] = a[k];
            rd[k][j] = b[k];
        }
        X->key0[j] = py * P->w + px;
        depth[j] = P->depth ? P->depth[X->key0[j]] : FLT_MAX;
        o[j] = ((size_t)ly * J->tilex + lx) * 4;
    }
    if (!vany(gm)) return;
    /* The camera rays' leaves, once for all samples (the reference's caching
     * full traverser, kernel.cpp:130-134). */
    for (int j = 0; j < 4; ++j) {
        X->L[j].n = 0;
        if (gm[j]) {
            float a[3] = {ro[0][j], ro[1][j], ro[2][j]}, b[3] = {rd[0][j], rd[1][j], rd[2][j]};
            traverse(&X->L[j], a, b);
        }
    }
    double acc[4][4][4] = {{{0}}};
    for (uint32_t s = 0; s < J->samples; ++s) {
        X->sample = J->sample_base + s;
        v4f cl[4], cs[4];
        psample_volume(X, ro, rd, gm, 0, 0u, depth, cl, cs);
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                acc[j][0][k] += cl[k][j];
                acc[j][1][k] += cs[k][j];
                acc[j][2][k] += k < 3 ? (double)cl[k][j] * cl[k][j] : cl[k][j];
                acc[j][3][k] += k < 3 ? (double)cs[k][j] * cs[k][j] : cs[k][j];
            }
    }
    for (int j = 0; j < 4; ++j) {
        if (!gm[j]) continue;
        for (int b = 0; b < 4; ++b)
            for (int k = 0; k < 4; ++k) J->out[b * J->plane + o[j] + k] = (float)acc[j][b][k];
    }
}

static void *render_packet_worker(void *arg) {
    ojob *J = (ojob *)arg;
    pctx X;
    memset(&X, 0, sizeof(X));
    X.P = J->P;
    for (int j = 0; j < 4; ++j) X.L[j].ob = J->P->ob;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        uint32_t row = J->next_row++;
        pthread_mutex_unlock(&J->mu);
        if (2u * row >= J->ry) break;
        for (uint32_t qx = 0; 2u * qx < J->rx; ++qx) render_packet(J, &X, qx, row);
    }
    for (int j = 0; j < 4; ++j) free(X.L[j].ls);
    return NULL;
}

/* or_render with the reference's 2x2-packet, 4-lane structure: same
 * arguments, same output bits; threads take rows of packets. */
int or_render_packets(const oparams *P, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
                      uint32_t sample_base, uint32_t threads, float *out) {
    ojob J;
    memset(&J, 0, sizeof(J));
    J.P = P;
    J.x0 = x0; J.y0 = y0;
    J.rx = tilex & ~1u; J.ry = tiley & ~1u;
    if (x0 + J.rx > P->w) J.rx = x0 < P->w ? P->w - x0 : 0;
    if (y0 + J.ry > P->h) J.ry = y0 < P->h ? P->h - y0 : 0;
    J.tilex = tilex;
    J.samples = samples;
    J.sample_base = sample_base;
    J.out = out;
    J.plane = (size_t)tilex * tiley * 4;
    memset(out, 0, sizeof(float) * 4 * J.plane);
    pthread_mutex_init(&J.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (uint32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, render_packet_worker, &J);
    for (uint32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J.mu);
    return 0;
}

/* ------------------------------------------------------------- adjoint -- */
/*
 * Derivative of the expected image with respect to the FOG brick values,
 * majorants fixed (no reference counterpart: the reference is not
 * differentiable). Differential delta tracking, written over the reference's
 * recursion: a tracked ray whose subtree has adjoint-weighted value W adds
 *   W * w_v(x)/rho(x)                      at its real collision x,
 *   -W * sigma_e * (b-a) * w_v(a+u(b-a))   per covered leaf interval [a,b]
 *                                          with a non-zero majorant,
 * w_v the trilinear weights of vertex v; u comes from the ray's adjoint
 * stream (Philox counter purpose 3), one draw per covered leaf.
 */
typedef struct {
    float *grad;             /* leafx * lvoxc^3 */
    float gL[3], gE[3], gA;
} oadj;

/* Corner indices and weights of or_voxel's lerps (d lerp/da = 1-t, /db = t). */
static void voxel_weights(const float p[3], const float ce[4], uint32_t lvoxc, uint32_t idx[8], float w[8]) {
    float nv = (float)lvoxc, nvm = nv - 1.0f, nl[3];
    uint32_t a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        float ni = ((-0.5f * ((ce[k] - ce[3]) - p[k])) * nvm) / ce[3];
        float nf = fmaxf(floorf(ni), 0.0f);
        float nc = fminf(ceilf(ni), nvm);
        nl[k] = ni - nf;
        nc = fmaxf(fminf(nc, (float)(lvoxc - 1)), 0.0f);
        nf = fmaxf(fminf(nf, (float)(lvoxc - 1)), 0.0f);
        a[k] = (uint32_t)nf;
        b[k] = (uint32_t)nc;
    }
    for (int v = 0; v < 8; ++v) {
        int bx = v & 1, by = (v >> 1) & 1, bz = (v >> 2) & 1;
        idx[v] = (bz ? b[2] : a[2]) * lvoxc * lvoxc + (by ? b[1] : a[1]) * lvoxc + (bx ? b[0] : a[0]);
        w[v] = (bx ? nl[0] : 1.0f - nl[0]) * (by ? nl[1] : 1.0f - nl[1]) * (bz ? nl[2] : 1.0f - nl[2]);
    }
}

static void add_grad(const oparams *P, oadj *A, const onode *N, const float x[3], float weight) {
    if (weight == 0.0f) return;
    uint32_t idx[8];
    float w[8];
    voxel_weights(x, N->ce, P->lvoxc, idx, w);
    float *g = A->grad + (size_t)N->volx[1] * P->lvoxc3;
    for (int v = 0; v < 8; ++v) g[idx[v]] += weight * w[v];
}

static void add_segment(const oparams *P, oadj *A, const onode *N, ostream *ar, const float ro[3], const float rd[3],
                        float a, float b, float W) {
    float u = stream_u(ar);
    if (!(b > a)) return;
    float t = a + u * (b - a);
    float x[3] = {ro[0] + rd[0] * t, ro[1] + rd[1] * t, ro[2] + rd[2] * t};
    add_grad(P, A, N, x, -W * (P->sigma_a + P->sigma_s) * (b - a));
}

/* Replays track() for ray `ray_id` and adds its derivative terms with weight W. */
static void track_adj(octx *X, oadj *A, const float ro[3], const float rd[3], uint32_t ray_id, float depth, float W) {
    const oparams *P = X->P;
    ostream rng, ar;
    stream_seed(&rng, X->key0, P->key1, X->sample, ray_id);
    {
        uint32_t c[4] = {X->sample, ray_id, 0u, 3u}, v[4];
        philox(c, X->key0, P->key1, v);
        ar.s[0] = v[0]; ar.s[1] = v[1]; ar.s[2] = v[2];
        ar.s[3] = v[3] | ((v[0] | v[1] | v[2]) == 0u ? 1u : 0u);
    }
    traverse(&X->L, ro, rd);
    const float sige = P->sigma_a + P->sigma_s;
    float td = 0.0f;
    for (uint32_t i = 0; i < X->L.n; ++i) {
        if (!(td < depth)) break;
        const oleaf *lf = &X->L.ls[i];
        const onode *N = &P->ob[lf->node];
        float tra = lf->t0, trb = lf->t1, lo = td, tr0 = tra - td, tr1 = trb - td, smax = N->qval[1];
        td = tra;
        const float *brick = P->pvol_fog + (size_t)N->volx[1] * P->lvoxc3;
        float sr = -(logf(stream_u(&rng)) / (sige * smax));
        for (;;) {
            float sc = tr0 + sr;
            td = lo + sc;
            if (!(sc < tr1)) {
                td = trb;
                if (sige * smax > 0.0f) add_segment(P, A, N, &ar, ro, rd, tra, trb, W);
                break;
            }
            if (!(sc > tr0)) return;          /* collision before the leaf entry: no density term */
            float rc[3] = {ro[0] + rd[0] * td, ro[1] + rd[1] * td, ro[2] + rd[2] * td};
            float p = or_voxel(brick, rc, N->ce, P->lvoxc);
            float q = stream_u(&rng);
            if (!(p / smax < q)) {
                add_segment(P, A, N, &ar, ro, rd, tra, td, W);
                add_grad(P, A, N, rc, W / p);
                return;
            }
            sr = sr - logf(stream_u(&rng)) / (sige * smax);
        }
    }
}

/* sample_volume with the adjoint: same value, plus this ray's and its
 * children's derivative terms. mL/mE: per-channel multipliers from the camera
 * sample down to this call. */
static void sample_volume_adj(octx *X, oadj *A, const float ro[3], const float rd[3], uint32_t r, uint32_t ray_id,
                              float depth, const float mL[3], const float mE[3], ores *out) {
    const oparams *P = X->P;
    memset(out, 0, sizeof(*out));
    float td;
    int scattered = track(X, ro, rd, ray_id, depth, &td);
    int rm = !scattered;
    float lc[4] = {0, 0, 0, 0}, le[4] = {0, 0, 0, 0};
    if (rm && r > 0) {
        for (uint32_t l = 0; l < P->light_count; ++l)
            if (dot3(rd, P->lights[l].dir) > P->lights[l].cosang)
                for (int k = 0; k < 3; ++k) lc[k] += P->lights[l].color[k];
        env_eval(P, rd, le);
        lc[3] = le[3] = 1.0f;
    } else {
        lc[3] = le[3] = rm ? 0.0f : 1.0f;
    }
    if (r < P->scattevs && !rm) {
        uint32_t c[4] = {X->sample, r, 0u, 1u}, v[4];
        philox(c, X->key0, P->key1, v);
        float u1 = u01(v[0]), u2 = u01(v[1]), u3 = u01(v[2]), u4 = u01(v[3]);
        float srd[3], lrd[3], p1[3], p3[3];
        if (P->phase == 0) or_mie_sample(P->mie, rd, u1, u2, srd);
        else or_hg_sample(P->hg_g, rd, u1, u2, srd);
        sun_sample(&P->lights[0], u3, u4, lrd);
        if (P->phase == 0) {
            or_mie_eval(P->mie, dot3(srd, rd), p1);
            or_mie_eval(P->mie, dot3(lrd, rd), p3);
        } else {
            float h1 = or_hg_eval(P->hg_g, dot3(srd, rd)), h3 = or_hg_eval(P->hg_g, dot3(lrd, rd));
            p1[0] = p1[1] = p1[2] = h1;
            p3[0] = p3[1] = p3[2] = h3;
        }
        float p2 = sun_pdf(&P->lights[0], lrd), pd1 = sun_pdf(&P->lights[0], srd);
        float sige = P->sigma_a + P->sigma_s;
        float w1[3], w2[3], m1L[3], m1E[3], m2L[3], m2E[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            w1[k] = p1[k] / (p1[k] + pd1);
            w2[k] = p3[k] / (p3[k] + p2);
            m1L[k] = mL[k] * w1[k] * P->sigma_s / sige;
            m1E[k] = mE[k] * w1[k] * P->sigma_s / sige;
            m2L[k] = mL[k] * w2[k] * P->sigma_s / sige;
        }
        float rc[3] = {ro[0] + rd[0] * td, ro[1] + rd[1] * td, ro[2] + rd[2] * td};
        ores S1, S2;
        sample_volume_adj(X, A, rc, srd, r + 1, 2u * r + 2u, FLT_MAX, m1L, m1E, &S1);
        sample_volume_adj(X, A, rc, lrd, P->scattevs, 2u * r + 1u, FLT_MAX, m2L, m2E, &S2);
        for (int k = 0; k < 3; ++k) {
            out->cl[k] += ((S1.cl[k] * w1[k] + S2.cl[k] * w2[k]) * P->sigma_s) / sige;
            out->cs[k] += ((S1.cs[k] * w1[k] + 0.0f * w2[k]) * P->sigma_s) / sige;
        }
        out->cl[3] += lc[3];
        out->cs[3] += le[3];
    } else {
        for (int k = 0; k < 3; ++k) {
            out->cl[k] += rm ? lc[k] : 0.0f;
            out->cs[k] += rm ? le[k] : 0.0f;
        }
        out->cl[3] += lc[3];
        out->cs[3] += le[3];
    }
    /* This ray's weight: the adjoint-weighted value of everything it leads to
     * (its escape radiance, or the whole subtree of its collision), as seen
     * from the camera sample. Its children added their own terms above. */
    /* Channels with a zero multiplier never reach the image (a shadow ray's
     * sky is dropped, sky2 = 0, and may be NaN near the sun): skip them. */
    float Wray = 0.0f;
    for (int k = 0; k < 3; ++k) {
        if (A->gL[k] != 0.0f && mL[k] != 0.0f) Wray += A->gL[k] * mL[k] * out->cl[k];
        if (A->gE[k] != 0.0f && mE[k] != 0.0f) Wray += A->gE[k] * mE[k] * out->cs[k];
    }
    if (r == 0) Wray += A->gA * out->cl[3];
    if (Wray != 0.0f) track_adj(X, A, ro, rd, ray_id, depth, Wray);
}

/*
 * Adjoint of or_render for a tile: adj holds dL/d(lights sums) then
 * dL/d(env sums), each tiley*tilex*4 floats; grad (leafx*lvoxc^3) is
 * overwritten with dL/d(FOG brick values).
 */
int or_adjoint(const oparams *P, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
               uint32_t sample_base, const float *adj, float *grad, size_t grad_n) {
    if (P->pvol_sdf) return -1;   /* SDF leaves are not differentiated */
    memset(grad, 0, sizeof(float) * grad_n);
    size_t plane = (size_t)tilex * tiley * 4;
    uint32_t rx = tilex & ~1u, ry = tiley & ~1u;
    if (x0 + rx > P->w) rx = x0 < P->w ? P->w - x0 : 0;
    if (y0 + ry > P->h) ry = y0 < P->h ? P->h - y0 : 0;
    octx X;
    memset(&X, 0, sizeof(X));
    X.P = P;
    X.L.ob = P->ob;
    oadj A;
    A.grad = grad;
    const float one[3] = {1.0f, 1.0f, 1.0f};
    for (uint32_t ly = 0; ly < ry; ++ly)
        for (uint32_t lx = 0; lx < rx; ++lx) {
            size_t o = ((size_t)ly * tilex + lx) * 4;
            for (int k = 0; k < 3; ++k) {
                A.gL[k] = adj[o + k];
                A.gE[k] = adj[plane + o + k];
            }
            A.gA = adj[o + 3] + adj[plane + o + 3];
            if (A.gL[0] == 0 && A.gL[1] == 0 && A.gL[2] == 0 && A.gE[0] == 0 && A.gE[1] == 0 && A.gE[2] == 0 && A.gA == 0)
                continue;
            uint32_t px = x0 + lx, py = y0 + ly;
            float ro[3], rd[3];
            if (!camera_ray(P, px, py, ro, rd)) continue;
            X.key0 = py * P->w + px;
            for (uint32_t s = 0; s < samples; ++s) {
                X.sample = sample_base + s;
                ores R;
                sample_volume_adj(&X, &A, ro, rd, 0, 0u, FLT_MAX, one, one, &R);
            }
        }
    free(X.L.ls);
    return 0;
}

/*
 * RenderKernel::Shadow (kernel.cpp:569-607): per pixel, the surface point at
 * the compositing depth (when depth < clip_end) sends `samples` sun-sampled
 * shadow rays, SampleVolume at r = scattevs; the lights' radiance of each
 * escaping ray is divided by lights[0]'s colour and summed into out buffer 0
 * (alpha: the sample count; the reference's alpha is 1/0 from a zero w). Lanes
 * without a surface are not traced and count as unshadowed.
 */
int or_shadow(const oparams *P, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
              uint32_t sample_base, float clip_end, float *out) {
    size_t plane = (size_t)tilex * tiley * 4;
    memset(out, 0, sizeof(float) * 4 * plane);
    uint32_t rx = tilex & ~1u, ry = tiley & ~1u;
    if (x0 + rx > P->w) rx = x0 < P->w ? P->w - x0 : 0;
    if (y0 + ry > P->h) ry = y0 < P->h ? P->h - y0 : 0;
    octx X;
    memset(&X, 0, sizeof(X));
    X.P = P;
    X.L.ob = P->ob;
    const olight *L0 = &P->lights[0];
    for (uint32_t ly = 0; ly < ry; ++ly)
        for (uint32_t lx = 0; lx < rx; ++lx) {
            uint32_t px = x0 + lx, py = y0 + ly;
            float ro[3], rd[3];
            int gm = camera_ray(P, px, py, ro, rd);
            if (!gm) continue;
            X.key0 = py * P->w + px;
            float depth = P->depth ? P->depth[X.key0] : FLT_MAX;
            int gm1 = gm && depth < clip_end;
            float ro1[3] = {ro[0] + rd[0] * depth, ro[1] + rd[1] * depth, ro[2] + rd[2] * depth};
            double acc[4] = {0, 0, 0, 0};
            for (uint32_t s = 0; s < samples; ++s) {
                X.sample = sample_base + s;
                uint32_t c[4] = {X.sample, 0u, 0u, 2u}, v[4];
                philox(c, X.key0, P->key1, v);
                float lrd[3];
                sun_sample(L0, u01(v[0]), u01(v[1]), lrd);
                float cl[3] = {0, 0, 0};
                if (gm1) {
                    ores R;
                    sample_volume(&X, ro1, lrd, P->scattevs, 1u, FLT_MAX, &R);
                    for (int k = 0; k < 3; ++k) cl[k] = R.cl[k];
                } else if (P->scattevs > 0) {
                    for (uint32_t l = 0; l < P->light_count; ++l)
                        if (dot3(lrd, P->lights[l].dir) > P->lights[l].cosang)
                            for (int k = 0; k < 3; ++k) cl[k] += P->lights[l].color[k];
                }
                for (int k = 0; k < 3; ++k) acc[k] += cl[k] / L0->color[k];
                acc[3] += 1.0;
            }
            size_t o = ((size_t)ly * tilex + lx) * 4;
            for (int k = 0; k < 4; ++k) out[o + k] = (float)acc[k];
        }
    free(X.L.ls);
    return 0;
}

uint32_t or_params_size(void) { return (uint32_t)sizeof(oparams); }
uint32_t or_mie_size(void) { return (uint32_t)sizeof(omie); }
