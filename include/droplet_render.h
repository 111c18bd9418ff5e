/*
 * droplet_render.h - C ABI of the MI355X volumetric light-transport path.
 *
 * This is the drop-in boundary under droplet-render's C++ render entry points.
 * Each entry point names the reference interface it replaces (file:line in
 * jaelpark/droplet-render). Plain pointers and sizes only; no exceptions
 * cross this boundary. Every function returns 0 (DRK_OK) on success or a
 * negative DRK_E* code, with a message available from drk_last_error().
 *
 * Threading: calls on one drk_kernel are serialised by the caller (the
 * reference renders one tile at a time per process, src/main.cpp:650-667).
 * Each drk_kernel owns its own HIP stream and device buffers.
 */
#ifndef DROPLET_RENDER_H
#define DROPLET_RENDER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRK_OK        0
#define DRK_EINVAL   -1   /* invalid argument (null pointer, empty/oversized tile, bad scene) */
#define DRK_ENOMEM   -2   /* host or device allocation failed */
#define DRK_EDEVICE  -3   /* no HIP device, launch or runtime failure */
#define DRK_ELIMIT   -4   /* a bounded loop hit its iteration cap */

/* Volume buffer slots, src/scene.h:4-8 (VOLUME_BUFFER_SDF, VOLUME_BUFFER_FOG). */
#define DRK_VOLUME_SDF 0
#define DRK_VOLUME_FOG 1
#define DRK_NO_BRICK   0xffffffffu

/* Octree node record, bit-identical to OctreeStructure (src/scene.h:21-31):
 * 64 bytes, ce@0 chn@16 volx@48 qval@56. chn[] is indexed by the build slot
 * i = x + 2y + 4z (src/scene.cpp:284-285); 0 means "no child". */
typedef struct drk_octree_node {
    float    ce[4];      /* centre xyz, half-extent in w */
    uint32_t chn[8];
    uint32_t volx[2];    /* brick index per volume buffer, DRK_NO_BRICK if none */
    float    qval[2];    /* SDF: min distance; FOG: max density (majorant, >= 0.01) */
} drk_octree_node;

/* Host scene: the data Scene::Initialize leaves behind (src/scene.h:146-159). */
typedef struct drk_host_scene {
    drk_octree_node *ob;       /* node_count records, root at 0 */
    uint32_t node_count;
    float   *pvol[2];          /* bricks: leafx[b] * lvoxc^3 floats, vertex-centred, x fastest */
    uint32_t leafx[2];
    uint32_t lvoxc;            /* samples per brick edge (8 unless the depth cap doubled it) */
    uint32_t mlevel;           /* octree depth; leaves sit at level mlevel-1 */
    float    root_ce[4];
} drk_host_scene;

/*
 * Builds the octree and bricks from dense grids, replacing the OpenVDB part of
 * Scene::Initialize: leaf-AABB collection and root sizing
 * (src/scene.cpp:676-716), Octree::BuildPath (src/scene.cpp:267-305) and the
 * brick resampling with per-leaf majorants (src/scene.cpp:840-917).
 * Grids are x-fastest, voxel (i,j,k) at world (i,j,k)*voxel_size. A FOG
 * voxel is active when its value differs from the background 0. sdf (may be
 * NULL, as may fog, not both) is a signed distance in world units, negative
 * inside: it stands for the reference's level set, whose active voxels are
 * the narrow band |d| < 4*voxel_size (bvc, src/scene.cpp:844) and whose
 * values beyond it are +-that band (also the background outside the grid).
 * SDF leaves get SDF bricks and qval[SDF] = their minimum distance; FOG-only
 * leaves whose centre is inside the surface lose their FOG brick
 * (src/scene.cpp:874-883). maxd is the depth cap (src/scene.cpp:712-716).
 */
int drk_scene_build_dense(const float *fog, const float *sdf,
                          uint32_t nx, uint32_t ny, uint32_t nz,
                          float voxel_size, uint32_t maxd,
                          drk_host_scene **out);
void drk_host_scene_free(drk_host_scene *scene);

/*
 * On-disk scene blob (no reference counterpart; the reference caches the
 * OpenVDB grids, src/scene.cpp:420-477, and rebuilds the octree each time):
 * the node records and both brick buffers of a built scene, with a header
 * and an FNV-1a 64 checksum of the payload. drk_scene_load returns a host
 * scene (free it with drk_host_scene_free); a file with a bad magic, sizes,
 * checksum or out-of-range indices is rejected with DRK_EINVAL.
 */
int drk_scene_save(const drk_host_scene *scene, const char *path);
int drk_scene_load(const char *path, drk_host_scene **out);

/* Device-resident scene: uploads Scene::ob and Scene::pvol once
 * (the reference reads them in place, src/kernel.cpp:96-445). */
typedef struct drk_scene drk_scene;
int  drk_scene_create(const drk_octree_node *ob, uint32_t node_count,
                      const float *pvol_sdf, uint32_t leafx_sdf,
                      const float *pvol_fog, uint32_t leafx_fog,
                      uint32_t lvoxc, drk_scene **out);
/*
 * Octree walk of the kernels (no reference counterpart; the leaf sequence
 * and intervals are the Revelles traversal's, src/KernelOctree.cpp:9-349,
 * either way). When every leaf sits at one depth L <= 6 (the reference's
 * build), drk_scene_create also builds the leaf grid (2^L cells per axis ->
 * leaf node) and the kernels can step through it cell by cell instead of
 * climbing the tree; DRK_WALK_AUTO (the default) walks the grid when at
 * least 1/8 of its cells hold a leaf. drk_scene_walk_info reports the
 * grid's depth (0: none), the fraction of occupied cells and whether
 * launches walk it. Takes effect at the next launch on the scene.
 */
#define DRK_WALK_AUTO   0
#define DRK_WALK_OCTREE 1
#define DRK_WALK_GRID   2
int drk_scene_set_walk(drk_scene *scene, uint32_t walk);
int drk_scene_walk_info(const drk_scene *scene, uint32_t *grid_levels, float *occupancy, int *uses_grid);

/* Replaces the FOG bricks and per-leaf majorants in place (same leaf count);
 * used between optimisation steps. qval may be NULL to keep the majorants. */
int  drk_scene_update_fog(drk_scene *scene, const float *pvol_fog, const float *qval_fog);
/* Stream-ordered form for FOG bricks already in device memory (no reference
 * counterpart): an asynchronous copy on `stream`, ordered after the work
 * that produced d_pvol_fog and before later launches on that stream; the
 * majorants are kept. Launches on other streams must be ordered by the
 * caller. */
int  drk_scene_update_fog_device(drk_scene *scene, const float *d_pvol_fog, void *stream);
void drk_scene_destroy(drk_scene *scene);

/* KernelSampler objects (src/KernelSampler.h:6-97). */
#define DRK_PHASE_MIE 0      /* MiePhase::gmie, src/KernelSampler.cpp:206-287 */
#define DRK_PHASE_HG  1      /* HGPhase(g),     src/KernelSampler.cpp:165-204 */
#define DRK_ENV_HOSEK 0      /* Hosek-Wilkie sky, src/kernel.cpp:329-360 (USE_ARHOSEK_SKYMODEL) */
#define DRK_ENV_NULL  1      /* NullEnv, src/KernelSampler.cpp:355-367 */
#define DRK_ENV_MAP   2      /* MapEnv (equirect RGBA texels), src/KernelSampler.cpp:369-425 */
#define DRK_MAX_LIGHTS 8

/* SunLight(direction, color*intensity, angle), src/KernelSampler.cpp:305-345. */
typedef struct drk_sun {
    float direction[3];
    float color[3];
    float angle;
} drk_sun;

typedef struct drk_sampler_desc {
    uint32_t phase;            /* DRK_PHASE_* */
    float    hg_g;             /* HGPhase::g1 */
    uint32_t env;              /* DRK_ENV_* */
    float    sky_turbidity;    /* Hosek state parameters (reference hard-codes 7.0 / 0.15) */
    float    sky_albedo;
    const float *env_map;      /* env_w*env_h RGBA floats, row-major, for DRK_ENV_MAP */
    uint32_t env_w, env_h;
    const drk_sun *lights;     /* BaseLight::lights; lights[0] is the MIS light */
    uint32_t light_count;
} drk_sampler_desc;

/* RenderKernel flags, src/kernel.h:4. */
#define DRK_KERNEL_DEPTHCOMP 0x1

typedef struct drk_kernel drk_kernel;

/*
 * RenderKernel::Initialize (src/kernel.h:15, src/kernel.cpp:497-534).
 * viewi/proji are the inverted view and projection matrices as 4 rows of 4
 * floats, row-vector convention (src/main.cpp:112-115). pdepth (w*h floats,
 * NULL allowed) is the compositing depth. tilex*tiley sizes the two result
 * buffers phb[0] (lights) and phb[1] (environment).
 */
int drk_kernel_create(const drk_scene *scene, const float viewi[16], const float proji[16],
                      const drk_sampler_desc *samplers, const float *pdepth,
                      uint32_t scattevs, float msigmas, float msigmaa,
                      uint32_t tilex, uint32_t tiley, uint32_t w, uint32_t h,
                      uint32_t flags, drk_kernel **out);

/* Seeds the counter-based RNG. Every Render advances sample_base by its sample
 * count, so successive calls draw fresh samples (the reference's per-packet
 * RNG streams keep running between calls, src/kernel.cpp:23-84). */
int drk_kernel_set_seed(drk_kernel *k, uint64_t seed, uint32_t sample_base);

/*
 * RenderKernel::Render (src/kernel.h:16, src/kernel.cpp:536-567): renders the
 * tile (x0,y0,tilex,tiley) with `samples` paths per pixel and leaves per-pixel
 * SUMS (not means) in phb[0] (sun, alpha) and phb[1] (sky/env, alpha), tile
 * row-major with stride tilex. Synchronous, like the reference. Odd tile
 * sizes drop the last column/row as in src/kernel.cpp:453.
 */
int drk_kernel_render(drk_kernel *k, uint32_t x0, uint32_t y0,
                      uint32_t tilex, uint32_t tiley, uint32_t samples);

/* RenderKernel::Shadow (src/kernel.h:17, src/kernel.cpp:569-607). */
int drk_kernel_shadow(drk_kernel *k, uint32_t x0, uint32_t y0,
                      uint32_t tilex, uint32_t tiley, uint32_t samples);

/* RenderKernel::phb[buffer] (src/kernel.h:26): tilex*tiley RGBA floats,
 * valid until the next Render/Shadow or Destroy. buffer 2 and 3 hold the
 * per-pixel sums of squares of the lights and environment samples
 * (needed for Monte Carlo error bars; no reference counterpart). */
const float *drk_kernel_phb(const drk_kernel *k, uint32_t buffer);

/* RenderKernel::Destroy (src/kernel.h:18, src/kernel.cpp:609-615). */
void drk_kernel_destroy(drk_kernel *k);

/*
 * Device-buffer form of Render for callers that keep results in HBM
 * (benchmarks, the autograd wrapper, batched drivers). Output pointers are
 * device pointers of tilex*tiley float4 each; the sums are written (not
 * accumulated). sumsq pointers may be NULL. stream is a hipStream_t (NULL = the
 * default null stream, as in HIP). Asynchronous on that stream.
 */
int drk_kernel_render_device(drk_kernel *k, uint32_t x0, uint32_t y0,
                             uint32_t tilex, uint32_t tiley, uint32_t samples,
                             uint64_t seed, uint32_t sample_base,
                             float *d_lights, float *d_env,
                             float *d_sumsq_lights, float *d_sumsq_env,
                             void *stream);

/*
 * Allocation-free form of drk_kernel_render_device (no reference
 * counterpart): for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, work and results as
 * drk_kernel_render_device with the same chunking, but the launch's scratch
 * (its work-unit counters, launch parameters and per-chunk partial sums; the
 * camera rays are computed per path instead of once per pixel) lives in a fixed
 * 128 MiB device array that the library holds on every GPU it runs on until
 * the process ends, so the call allocates nothing. A call whose partial sums
 * exceed that array (pixels x sample chunks x 32 bytes, 64 with sums of
 * squares) is refused with DRK_EINVAL; a smaller tile or fewer chunks
 * (drk_kernel_set_tuning) fits. All calls on one GPU share the array: calls
 * that overlap in time on one GPU (other streams, other kernels, other
 * threads) race and must be serialised by the caller. Not for depth
 * compositing's adjoint or the path record (forward sums only).
 */
int drk_kernel_render_device_noalloc(drk_kernel *k, uint32_t x0, uint32_t y0,
                                     uint32_t tilex, uint32_t tiley, uint32_t samples,
                                     uint64_t seed, uint32_t sample_base,
                                     float *d_lights, float *d_env,
                                     float *d_sumsq_lights, float *d_sumsq_env,
                                     void *stream);

/*
 * Adjoint of Render (no reference counterpart; the reference is not
 * differentiable). Given dL/d(phb sums) for the lights and environment
 * buffers (tilex*tiley RGBA floats each; the alpha gradient counts once per
 * buffer), accumulates dL/d(FOG brick values) into grad_fog (leafx*lvoxc^3
 * floats, the layout of Scene::pvol[FOG]) for the render with the same tile,
 * samples, seed and sample_base. The majorants (qval) are held fixed. The
 * estimator is path replay with differential delta tracking: unbiased for the
 * derivative of the expected image, no stored path vertices.
 * _device: device pointers, accumulates (the caller zeroes grad), async on
 * stream. Host form: host pointers, grad_fog overwritten, synchronous.
 * Not supported with depth compositing or scattevs > 64.
 */
int drk_kernel_adjoint_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                              uint32_t samples, uint64_t seed, uint32_t sample_base, const float *d_adj_lights,
                              const float *d_adj_env, float *d_grad_fog, void *stream);
int drk_kernel_adjoint(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
                       uint64_t seed, uint32_t sample_base, const float *adj_lights, const float *adj_env,
                       float *grad_fog);

/*
 * Estimators of the adjoint's free-flight terms (no reference counterpart).
 * DRK_ADJ_SEGMENT (the default): differential delta tracking, one uniform
 * point per covered leaf interval, -W sigma_e (b-a) w_v(point).
 * DRK_ADJ_NULL_COLLISION: one term per null collision x of the replayed
 * ray, -W w_v(x) / (majorant - rho(x)): the exact derivative, path by path,
 * of drk_kernel_render_weighted_device at its reference densities, so
 * central differences of that forward with common random numbers equal it
 * up to O(h^2). Both are unbiased for the derivative of the expected image.
 */
#define DRK_ADJ_SEGMENT        0
#define DRK_ADJ_NULL_COLLISION 1
int drk_kernel_set_adjoint_estimator(drk_kernel *k, uint32_t estimator);

/*
 * Weighted delta tracking (no reference counterpart): a density-smooth
 * forward for finite-difference checks of the adjoint. The scene's
 * reference FOG bricks (drk_scene_set_reference_fog: leafx*lvoxc^3 floats,
 * host or device memory; NULL drops them) decide every tentative collision
 * with the reference's comparison, and each ray carries the ratio of that
 * decision's probabilities under the current and the reference densities
 * (rho/rho_ref when accepted, (q-rho)/(q-rho_ref) when rejected, q the
 * leaf's majorant). Its expected image equals the plain render's for the
 * current densities (at or below the majorants); where they equal the
 * reference the image is the plain render's bit for bit, and it is smooth
 * in the densities under common random numbers. Sums only (no alpha
 * counts: alpha is the weighted scatter count). FOG-only scenes.
 */
int drk_scene_set_reference_fog(drk_scene *scene, const float *pvol_ref);
int drk_kernel_render_weighted_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                      uint32_t samples, uint64_t seed, uint32_t sample_base,
                                      float *d_lights, float *d_env, void *stream);

/*
 * Path record (no reference counterpart): the forward render keeps, per
 * (pixel, sample) of the rendered extent, the path's lights and environment
 * radiance, its alpha, which of its shadow rays escaped and where its last
 * contribution arose (40 bytes each;
 * drk_kernel_record_bytes gives the size of a call's record). The adjoint of
 * the same call (same tile, samples, seed, sample_base and scene densities)
 * given that record skips the replay that recomputes those totals, so it
 * traces each path once instead of twice; the gradient is the same
 * (identical per-path contributions, summed in another order). d_record is a
 * 16-byte aligned device buffer of at least record_bytes bytes. Not supported
 * with scattevs > 64.
 */
/*
 * Gradient events (no reference counterpart). The adjoint writes each
 * gradient contribution as a 20-byte event (weight and point; the brick in
 * a separate plane), sorts the events by brick and sums them in on-chip
 * memory, which avoids per-corner float atomics in HBM. The event buffer
 * belongs to the kernel's scene and is shared by the scene's kernels, so
 * adjoint calls on kernels of one scene must not overlap in time. It holds
 * 40 bytes per slot (the event and its sorted copy), grows to 16 slots per
 * pixel sample of the largest call so far, and is limited to `bytes`
 * (UINT64_MAX, the default: a quarter of the device memory free when the
 * buffer grows, counting the buffer itself; 0: no events, every contribution
 * is added with atomics); events beyond it, or all of them when the buffer
 * cannot be allocated, are added with atomics. The gradient is the same
 * either way up to the order of float additions. Setting the budget frees
 * the current buffer (after a device synchronisation).
 */
int drk_kernel_set_event_budget(drk_kernel *k, uint64_t bytes);

/*
 * Gradient window (no reference counterpart): one event reduction for many
 * adjoint calls, e.g. the views of one optimisation step. After
 * drk_scene_events_begin, the _device adjoint entry points of the scene's
 * kernels append their events to the scene's event buffer (sized for
 * `samples` pixel samples in total: 16 slots each, within the budget above)
 * instead of reducing them at the end of each call; drk_scene_events_end
 * reduces all of them at once and adds the result to d_grad_fog. Events
 * beyond the buffer go to each call's own d_grad_fog with atomics, so the
 * gradient is the sum of what the calls and the end add (pass one gradient
 * to all of them to get it in one place). The calls and the end must run in
 * order (one stream, or synchronised by the caller). While the window is
 * open the host adjoint and drk_kernel_set_event_budget are refused.
 */
int drk_scene_events_begin(drk_scene *scene, uint64_t samples, void *stream);
int drk_scene_events_end(drk_scene *scene, float *d_grad_fog, void *stream);

/* Device memory held by a kernel (result, partial-sum and gradient-copy
 * buffers as currently allocated) and by a scene (node records, child links,
 * FOG bricks, the adjoint's event buffers), in bytes. No reference
 * counterpart. */
int drk_kernel_footprint(const drk_kernel *k, size_t *bytes);
int drk_scene_footprint(const drk_scene *s, size_t *bytes);

int drk_kernel_record_bytes(const drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                            uint32_t samples, size_t *bytes);
int drk_kernel_render_record_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                    uint32_t samples, uint64_t seed, uint32_t sample_base, float *d_lights,
                                    float *d_env, void *d_record, size_t record_bytes, void *stream);
int drk_kernel_adjoint_record_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                     uint32_t samples, uint64_t seed, uint32_t sample_base,
                                     const float *d_adj_lights, const float *d_adj_env, float *d_grad_fog,
                                     const void *d_record, size_t record_bytes, void *stream);

/*
 * Batched forward launch (no reference counterpart): many render calls --
 * scenes, views, tiles of uneven size, each a (kernel, tile) job writing its
 * own device sum buffers as drk_kernel_render_device would -- run in one
 * persistent grid. Each workgroup visits every job once, starting at a
 * different one; its waves drain a job's work units before moving on.
 * drk_batch_create uploads the job table (jobs may share kernels and
 * scenes; buffers must not overlap). drk_batch_launch enqueues the batch on
 * a hipStream_t (NULL: the default stream). drk_batch_launch_graph captures
 * the same launch sequence into a hipGraph on its first call and replays
 * the graph after that (a graph made before drk_batch_set_tuning or
 * drk_batch_autotune is rebuilt). drk_batch_set_tuning sets the grid
 * (workgroups; 0 = occupancy) and the work-unit granularity: unit_scale x 16
 * sample chunks per pixel of every job (default 1); drk_batch_autotune
 * times a grid x units sweep on the stream (synchronising it), keeps the
 * fastest and writes the table to report.
 * Per-job results equal those of separate launches with the same chunking.
 */
typedef struct drk_batch drk_batch;
typedef struct drk_batch_job {
    drk_kernel *kernel;
    uint32_t x0, y0, tilex, tiley, samples;
    uint64_t seed;
    uint32_t sample_base;
    float *d_lights, *d_env;               /* tilex*tiley float4 sums each */
    float *d_sumsq_lights, *d_sumsq_env;   /* NULL for every job, or set for every job */
} drk_batch_job;
int drk_batch_create(const drk_batch_job *jobs, uint32_t njobs, drk_batch **out);
int drk_batch_set_tuning(drk_batch *b, uint32_t grid_blocks, float unit_scale);
int drk_batch_launch(drk_batch *b, void *stream);
int drk_batch_launch_graph(drk_batch *b, void *stream);
int drk_batch_autotune(drk_batch *b, void *stream, uint32_t reps, char *report, uint32_t report_len);
void drk_batch_destroy(drk_batch *b);

/*
 * Views: the path-record forward and its adjoint for several render calls
 * (views, tiles) in one persistent launch each (no reference counterpart).
 * drk_views_render_record_device(jobs, n) gives every job the sums and
 * path record of drk_kernel_render_record_device with the job's arguments
 * (d_lights, d_env: the outputs), bit for bit; drk_views_adjoint_record_device
 * (jobs, n, grad) adds to d_grad_fog what drk_kernel_adjoint_record_device
 * would for every job (d_lights, d_env: the job's dL/d(sums), read; the
 * same gradient up to the order of float additions), with one event
 * reduction for all jobs (none inside the scene's gradient window); jobs
 * whose events (16 slots per pixel sample) would not fit the event buffer
 * within its budget run as consecutive groups that fit, one launch and one
 * reduction each. Each
 * kernel at most once per call; the adjoint's kernels share one scene and
 * adjoint estimator. A separate launch per call drains its last paths alone
 * at the end; a views launch moves each workgroup on to the next job
 * instead. Every job uses its kernel's buffers (counters, camera rays,
 * partial sums) and the first job's kernel holds the job tables, so calls
 * that share kernels must not overlap in time. The round limit applies per
 * job (drk_kernel_limit_hit of each job's kernel).
 */
typedef struct drk_view_job {
    drk_kernel *kernel;
    uint32_t x0, y0, tilex, tiley, samples;
    uint64_t seed;
    uint32_t sample_base;
    float *d_lights, *d_env;   /* tilex*tiley float4 each: the sums (render) or their gradients (adjoint) */
    void *d_record;            /* the job's path record, 16-byte aligned */
    size_t record_bytes;
} drk_view_job;
int drk_views_render_record_device(const drk_view_job *jobs, uint32_t njobs, void *stream);
int drk_views_adjoint_record_device(const drk_view_job *jobs, uint32_t njobs, float *d_grad_fog, void *stream);

/*
 * Extensions with no reference counterpart.
 *
 * drk_kernel_set_tuning: fixes the work-unit chunk (samples per unit, every
 * chunk that size; the built-in heuristic instead ends with chunks halving
 * down to one sample) and the persistent grid size; 0 keeps the heuristic.
 * Used by the autotuner, benchmarks and tests that need a given chunking.
 *
 * drk_kernel_count: runs the render once with per-lane counters and returns
 * counts[0] = octree node records read, counts[1] = brick lookups. The RNG is
 * counter-based, so the counts describe any run with the same arguments.
 */
int drk_kernel_set_tuning(drk_kernel *k, uint32_t chunk, uint32_t grid_blocks);

/* Live kernel timing: when enabled, HIP events are recorded on the launch
 * stream right around every path-tracing kernel launch. drk_kernel_timing
 * waits for them, returns the summed duration (ms) and launch count since the
 * last query, and resets. */
int drk_kernel_set_timing(drk_kernel *k, int enable);
int drk_kernel_timing(drk_kernel *k, double *total_ms, uint32_t *launches);
int drk_kernel_count(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                     uint32_t samples, uint64_t seed, uint32_t sample_base,
                     unsigned long long counts[2]);

/*
 * Synthetic procedural density grid (n^3, x fastest, values in [0,1]) used as
 * benchmark and test input in place of the reference's Blender/OpenVDB scene
 * construction. kind 0: fBm cloud, prm = {octaves, base frequency,
 * lacunarity, gain, threshold}; kind 1: fBm-displaced union of spheres,
 * prm = {spheres, centre min, centre max, radius min, radius max,
 * displacement amplitude, octaves, base frequency, band width}.
 */
int drk_synth_cloud(uint32_t kind, uint32_t n, uint64_t seed, const float *prm, uint32_t nprm, float *out);

/*
 * Watchdog (no reference counterpart; the reference's loops have no bound):
 * limits the rounds of its state machine a wave of the path tracer may run
 * in one launch of this kernel (0, the default: no limit). A lane takes at
 * most one step (a traversal batch, a Woodcock step, a ray end) per round,
 * so the limit bounds every lane's steps. A wave that reaches it stops and
 * the launch ends normally with incomplete sums: the synchronous calls
 * (Render, Shadow, count, the host adjoint) then return DRK_ELIMIT;
 * drk_kernel_limit_hit reports it for the kernel's last device-buffer launch
 * (call it after synchronising that launch's stream), drk_batch_limit_hit
 * for a batch's last launch (jobs take their kernel's limit).
 */
int drk_kernel_set_round_limit(drk_kernel *k, uint32_t rounds);
int drk_kernel_limit_hit(const drk_kernel *k, int *hit);
int drk_batch_limit_hit(const drk_batch *b, int *hit);

/*
 * Device-allocation gate (no reference counterpart; the reference checks
 * only RenderKernel::Initialize's buffer allocation, src/kernel.cpp:500-502):
 * when set, gate(bytes, user) is called before every device allocation the
 * library makes, and a non-zero return makes that allocation fail as if the
 * device were out of memory (the call returns DRK_ENOMEM and frees what it
 * had allocated; the adjoint's event buffer falls back to atomics). NULL
 * removes the gate. Process-wide; set it while no other call runs. Used to
 * test the failure paths and to cap the library's device memory.
 */
typedef int (*drk_alloc_gate_fn)(size_t bytes, void *user);
int drk_set_alloc_gate(drk_alloc_gate_fn gate, void *user);

/* Thread-local text of the last error. */
const char *drk_last_error(void);

/* Library/device information (for logs and benchmarks). */
int drk_device_info(char *name, uint32_t name_len, uint32_t *cu_count);

#ifdef __cplusplus
}
#endif

#endif
