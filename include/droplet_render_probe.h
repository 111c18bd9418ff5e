/*
 * droplet_render_probe.h - test hooks of the MI355X path.
 *
 * No reference counterpart: each probe runs ONE device function of the
 * renderer on the GPU over a batch of inputs, so the parity tests can compare
 * the gfx950 code against the CPU oracle function by function (phase
 * functions, sun, sky, trilinear lookup, octree traversal, RNG streams).
 * Inputs and outputs are host arrays; each call is synchronous.
 * Return codes follow droplet_render.h.
 */
#ifndef DROPLET_RENDER_PROBE_H
#define DROPLET_RENDER_PROBE_H

#include <stdint.h>
#include "droplet_render.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox-4x32-10: out[4*i..] = philox(ctr[4*i..], key[2*i..]). */
int drk_probe_philox(const uint32_t *ctr, const uint32_t *key, uint32_t n, uint32_t *out);

/* First m uniforms of the tracking stream of (pixel key0, key1, sample, ray);
 * out is n*m floats. */
int drk_probe_ray_stream(const uint32_t *key0, const uint32_t *sample, const uint32_t *ray, uint32_t n,
                         uint32_t key1, uint32_t m, float *out);

/* MiePhase::EvaluateRGB at cos theta (out 3n) and MiePhase::Sample for
 * incident directions iv (3n) and uniforms u1,u2 (out 3n). */
int drk_probe_mie_eval(const float *ct, uint32_t n, float *out_rgb);
int drk_probe_mie_sample(const float *iv, const float *u1, const float *u2, uint32_t n, float *out_dir);

/* HGPhase::Evaluate / Sample with anisotropy g. */
int drk_probe_hg_eval(float g, const float *ct, uint32_t n, float *out);
int drk_probe_hg_sample(float g, const float *iv, const float *u1, const float *u2, uint32_t n, float *out_dir);

/* SunLight (lights[0]): Sample(u1,u2) directions, Pdf and Evaluate of dirs. */
int drk_probe_sun(const drk_sun *sun, const float *u1, const float *u2, const float *dirs, uint32_t n,
                  float *out_sample_dir, float *out_pdf, float *out_eval_rgb);

/* Sky radiance (Hosek-Wilkie, reference gamma and exposure) for directions rd
 * with the sun direction of `sun` and the given turbidity / albedo. */
int drk_probe_sky(const drk_sun *sun, float turbidity, float albedo, const float *rd, uint32_t n, float *out_rgb);

/* MapEnv::Evaluate then pow 2.2 for directions rd (3n): tex is w*h RGBA texels, row-major. */
int drk_probe_map_env(const float *tex, uint32_t w, uint32_t h, const float *rd, uint32_t n, float *out_rgb);

/* Hosek state in double, as the host computes it (configs 3x9, radiances 3). */
int drk_probe_hosek_state(double turbidity, double albedo, double elevation, double *cfg27, double *rad3);

/* SampleVoxelSpace: trilinear lookup of points p (3n) in one brick with node
 * ce; leaf_rcp != 0 divides by the extent through its correctly rounded
 * reciprocal and one Markstein correction (the render kernels' form for
 * scenes whose leaves share one extent), 0 by the division. */
int drk_probe_trilinear(const float *brick, uint32_t lvoxc, const float ce[4], const float *p, uint32_t n,
                        int leaf_rcp, float *out);

/* Leaf sequence of each ray through the scene's octree: up to max_leaves
 * (node, tr0, tr1) per ray; count[i] is the full sequence length (may exceed
 * max_leaves, in which case the list is truncated). */
int drk_probe_traverse(const drk_scene *scene, const float *ro, const float *rd, uint32_t n, uint32_t max_leaves,
                       uint32_t *nodes, float *tr0, float *tr1, uint32_t *count);

/* Camera rays of pixels (px, py) for an image w x h (origin 3n, dir 3n, mask n). */
int drk_probe_camera(const float viewi[16], const float proji[16], uint32_t w, uint32_t h, const uint32_t *px,
                     const uint32_t *py, uint32_t n, float *ro, float *rd, uint32_t *gm);

/* The tracker's free-flight logarithm against the device library's logf:
 * for u_i = (2i+1) * 2^-24, i < n (n <= 2^23: every uniform the RNG makes),
 * out_logf[i] = logf(u_i) and out_fast[i] = the kernels' log of u_i. */
int drk_probe_log_unit(uint32_t n, float *out_logf, float *out_fast);

/* Host check of the kernels' division by an invariant divisor (the work-unit
 * to pixel mapping): counts the n in {n0 + i*step mod 2^32 : i < count} and
 * the 64 largest 32-bit values whose quotient by d differs from n / d.
 * Runs on the CPU (no device needed). */
int drk_probe_udiv(uint32_t d, uint32_t n0, uint32_t step, uint64_t count, uint64_t *mismatches);

#ifdef __cplusplus
}
#endif

#endif
