// extern "C" boundary: argument checking, device memory, streams and the
// RenderKernel life cycle (src/kernel.cpp:497-615) around the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "drk_common.h"
#include "drk_internal.h"
#include "render_launch.h"

namespace drk {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

static int hip_fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? DRK_ENOMEM : DRK_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

// Device-allocation gate (drk_set_alloc_gate): consulted before every
// device allocation the library makes; a refusal reads as out of memory.
static drk_alloc_gate_fn g_gate = nullptr;
static void *g_gate_user = nullptr;

// Debugging aid: with DRK_POISON_ALLOC=1 in the environment every device
// buffer the library allocates starts filled with 0xff bytes (NaN floats,
// all-ones integers), so that a kernel reading memory it was never given a
// value for shows it in its results the same way every run. The fill is a
// synchronous hipMemset (not for callers that capture graphs).
static bool poison_alloc() {
    static const bool on = [] {
        const char *v = std::getenv("DRK_POISON_ALLOC");
        return v && v[0] == '1';
    }();
    return on;
}

template <class T>
static hipError_t dev_malloc(T **p, size_t bytes) {
    *p = nullptr;
    if (g_gate && g_gate(bytes, g_gate_user) != 0) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    if (e == hipSuccess && poison_alloc()) {
        e = hipMemset(*p, 0xff, bytes);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    return e;
}

#define HIP_TRY(expr, what)                      \
    do {                                         \
        hipError_t _e = (expr);                  \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

}  // namespace drk

using namespace drk;

// Default work-unit chunking of a launch: at most this many sample chunks per
// pixel, and at most this many bytes of per-chunk partial sums.
#ifndef DRK_MAX_AUTO_CHUNKS
#define DRK_MAX_AUTO_CHUNKS 64
#endif
constexpr uint64_t kMaxAutoChunks = DRK_MAX_AUTO_CHUNKS;
constexpr uint64_t kMaxPartialBytes = 4ull << 30;
// Batched launches: sample chunks per pixel of every job at unit scale 1.
constexpr double kBatchChunks = 16.0;

struct drk_scene {
    std::vector<drk_octree_node> h_ob;
    drk_octree_node *d_ob = nullptr;
    uint32_t *d_chx = nullptr;      // per node 8 child links: index | kChildLeaf if the child is a leaf
    float *d_bce = nullptr;         // per FOG brick: its leaf's ce (float4), for the adjoint's event reduction
    bool bce_ok = true;             // every FOG brick belongs to leaves of one ce (else no event reduction)
    float *d_fog = nullptr;
    float *d_sdf = nullptr;
    float *d_ref = nullptr;         // reference FOG bricks of weighted tracking (or null)
    // Leaf grid (every leaf at depth grid_levels): cell -> leaf node, and
    // whether the kernels walk it (DRK_WALK_*).
    uint32_t *d_cells = nullptr;
    uint32_t grid_levels = 0;
    float grid_occupancy = 0.0f;
    // RN(1 / e) when every leaf has the same half-extent e (the reference's
    // build puts all leaves at one depth, each level halving e exactly), for
    // the lookups' division by e; 0 otherwise.
    float leaf_rcp = 0.0f;
    uint32_t walk = DRK_WALK_AUTO;
    uint32_t node_count = 0;
    uint32_t leafx[2] = {0, 0};
    uint32_t lvoxc = 0;
    int device = 0;
    // Adjoint gradient events, shared by the scene's kernels (adjoint calls
    // on kernels of one scene run one at a time): budget (bytes; 0: direct
    // atomics only), slots allocated, and the buffers (events + sorted copy,
    // bin tables).
    size_t ev_budget = SIZE_MAX;    // SIZE_MAX: a quarter of device memory
    uint32_t ev_cap = 0;
    void *d_ev = nullptr;
    uint32_t *d_evbins = nullptr;   // count (2 words) + hist + bin_start + bin_fill + blk_start + bin_max
    // Gradient window (drk_scene_events_begin/end): adjoint calls append
    // their events to the buffer's first ev_window_cap slots and the
    // reduction runs once when the window ends.
    bool ev_window = false;
    uint32_t ev_window_cap = 0;
    uint64_t ev_window_samples = 0;
};

struct drk_kernel {
    drk_scene *scene = nullptr;     // shares its event buffers with the scene's other kernels
    RenderParams P;                 // launch template (tile fields filled per call)
    hipStream_t stream = nullptr;
    uint32_t tilex = 0, tiley = 0;
    uint64_t seed = 0x5eedULL;
    uint32_t sample_base = 0;
    float *d_out = nullptr;         // 4 buffers of tilex*tiley float4
    float *h_phb = nullptr;         // host copies, same layout
    float *d_partial = nullptr;
    size_t partial_bytes = 0;
    void *d_mail = nullptr;         // escape-service mailboxes, two float4 per lane of the grid
    size_t mail_bytes = 0;
    void *d_cam = nullptr;          // camera rays of the rendered extent, one float4 per pixel
    size_t cam_bytes = 0;
    uint32_t *d_counter = nullptr;  // launch block: counter block, then the launch's parameters
    unsigned long long *d_counts = nullptr;
    float *d_env = nullptr;
    KernelConsts C;                 // host copy of the per-kernel constants
    KernelConsts *d_kc = nullptr;   // device copy read by the kernels
    uint32_t bands = 1;             // work bands (one per XCD)
    uint32_t xcc = 1;               // XCDs of the device
    uint32_t grad_copies = 1;       // gradient copies of the adjoint (one per XCD by default)
    float *d_gradx = nullptr;       // grad_copies * (FOG gradient size) floats
    size_t gradx_bytes = 0;
    float *d_depth = nullptr;
    uint32_t cu_count = 0;
    uint32_t blocks_per_cu = 1;
    uint32_t round_limit = 0;       // watchdog: rounds per wave and launch (0: none)
    uint32_t adj_estimator = DRK_ADJ_SEGMENT;   // free-flight terms of the adjoint
    uint32_t *status = nullptr;     // counter block of the last launch (its status word at kStatusWord)
    void *d_jobs = nullptr;         // job tables of views calls led by this kernel (RenderParams, WorkParams)
    size_t jobs_bytes = 0;
    char *scratch = nullptr;        // this device's fixed scratch (allocation-free render), or null
    uint32_t chunk_override = 0;
    uint32_t max_chunks = (uint32_t)kMaxAutoChunks;   // default chunking: at most this many per pixel
    uint32_t grid_override = 0;
    bool depthcomp = false;
    // Live kernel timing: event pairs recorded around each k_render launch.
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free, ev_used;
};

struct drk_batch {
    std::vector<drk_batch_job> jobs;
    std::vector<RenderParams> hP;
    std::vector<WorkParams> hW;
    RenderParams *dP = nullptr;
    WorkParams *dW = nullptr;
    uint32_t *d_counters = nullptr;  // kCounterWords words per job
    float *d_partial = nullptr;
    size_t partial_bytes = 0;
    void *d_cam = nullptr;           // every job's camera rays (one float4 per pixel of its extent)
    size_t cam_bytes = 0;
    uint32_t cam_pix = 0;            // pixels of the largest extent
    uint32_t grid = 0;               // 0: occupancy-derived
    uint32_t grid_eff = 1;           // grid of the planned launch
    float unit_scale = 1.0f;         // sample chunks per pixel, relative to kBatchChunks
    uint32_t cu_count = 1;
    bool sumsq = false;
    bool planned = false;
    hipStream_t cap_stream = nullptr;  // stream the graph is captured on
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

const char *drk_last_error(void) { return g_err.c_str(); }

int drk_set_alloc_gate(drk_alloc_gate_fn fn, void *user) {
    g_gate = fn;
    g_gate_user = user;
    return DRK_OK;
}

int drk_device_info(char *name, uint32_t name_len, uint32_t *cu_count) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = (uint32_t)prop.multiProcessorCount;
    return DRK_OK;
}

int drk_scene_build_dense(const float *fog, const float *sdf, uint32_t nx, uint32_t ny, uint32_t nz, float voxel_size,
                          uint32_t maxd, drk_host_scene **out) {
    return scene_build_dense(fog, sdf, nx, ny, nz, voxel_size, maxd, out);
}

void drk_host_scene_free(drk_host_scene *scene) { host_scene_free(scene); }

int drk_scene_create(const drk_octree_node *ob, uint32_t node_count, const float *pvol_sdf, uint32_t leafx_sdf,
                     const float *pvol_fog, uint32_t leafx_fog, uint32_t lvoxc, drk_scene **out) {
    if (!out) return set_error(DRK_EINVAL, "drk_scene_create: out is NULL");
    *out = nullptr;
    if (!ob || node_count == 0 || node_count > kMaxNodes || lvoxc < 2 || lvoxc > 256)
        return set_error(DRK_EINVAL, "drk_scene_create: invalid octree (nodes=%u lvoxc=%u)", node_count, lvoxc);
    if (leafx_fog && !pvol_fog) return set_error(DRK_EINVAL, "drk_scene_create: FOG bricks missing");
    if (leafx_sdf && !pvol_sdf) return set_error(DRK_EINVAL, "drk_scene_create: SDF bricks missing");
    // Validate the structure the kernels index blindly.
    for (uint32_t i = 0; i < node_count; ++i) {
        const drk_octree_node &n = ob[i];
        for (int c = 0; c < 8; ++c)
            if (n.chn[c] >= node_count || (n.chn[c] != 0 && n.chn[c] <= i))
                return set_error(DRK_EINVAL, "drk_scene_create: node %u child %d out of range", i, c);
        if (n.volx[DRK_VOLUME_SDF] != DRK_NO_BRICK && n.volx[DRK_VOLUME_SDF] >= leafx_sdf)
            return set_error(DRK_EINVAL, "drk_scene_create: node %u SDF brick %u >= leafx %u", i, n.volx[DRK_VOLUME_SDF], leafx_sdf);
        if (n.volx[DRK_VOLUME_FOG] != DRK_NO_BRICK && n.volx[DRK_VOLUME_FOG] >= leafx_fog)
            return set_error(DRK_EINVAL, "drk_scene_create: node %u brick %u >= leafx %u", i, n.volx[DRK_VOLUME_FOG], leafx_fog);
        if (!(n.ce[3] > 0.0f)) return set_error(DRK_EINVAL, "drk_scene_create: node %u has non-positive extent", i);
    }
    // Interior nodes must sit above level kMaxDepth: the traversal keeps one
    // stack word per interior level. A node may be reached through several
    // parents (the reference's build never does this, but a hand-made or
    // loaded array can), so its level is the deepest over all of them.
    // Children have larger indices than their parents, so one forward pass
    // sees every parent of a node before the node itself.
    {
        constexpr uint16_t kUnseen = 0xffff;
        std::vector<uint16_t> lvl(node_count, kUnseen);
        lvl[0] = 0;
        for (uint32_t i = 0; i < node_count; ++i) {
            if (lvl[i] == kUnseen) continue;
            const drk_octree_node &n = ob[i];
            bool leaf = n.volx[0] != DRK_NO_BRICK || n.volx[1] != DRK_NO_BRICK;
            if (leaf) continue;
            if (lvl[i] >= kMaxDepth) return set_error(DRK_EINVAL, "drk_scene_create: octree deeper than %u levels", kMaxDepth);
            for (int c = 0; c < 8; ++c) {
                const uint32_t ch = n.chn[c];
                if (ch && (lvl[ch] == kUnseen || lvl[ch] < lvl[i] + 1)) lvl[ch] = (uint16_t)(lvl[i] + 1);
            }
        }
    }
    // Leaf grid: when every leaf reachable from the root sits at one depth L
    // (as the reference's build makes it), the cell of each leaf follows
    // from the child slots on its path (slot bit 0: x, 1: y, 2: z).
    std::vector<uint32_t> cells;
    uint32_t grid_levels = 0;
    float grid_occ = 0.0f;
    {
        struct Item { uint32_t node, level, x, y, z; };
        std::vector<Item> todo{{0u, 0u, 0u, 0u, 0u}};
        std::vector<Item> leaves;
        int depth = -1;
        bool ok = !(ob[0].volx[0] != DRK_NO_BRICK || ob[0].volx[1] != DRK_NO_BRICK);
        size_t visits = 0;
        while (ok && !todo.empty()) {
            const Item it = todo.back();
            todo.pop_back();
            if (++visits > (8u << 20)) { ok = false; break; }
            const drk_octree_node &n = ob[it.node];
            if (n.volx[0] != DRK_NO_BRICK || n.volx[1] != DRK_NO_BRICK) {
                if (depth < 0) depth = (int)it.level;
                ok = depth == (int)it.level;
                leaves.push_back(it);
                continue;
            }
            if (it.level >= kMaxGridLevels) { ok = false; break; }
            for (uint32_t c = 0; c < 8; ++c)
                if (n.chn[c]) todo.push_back({n.chn[c], it.level + 1, 2 * it.x + (c & 1), 2 * it.y + ((c >> 1) & 1),
                                              2 * it.z + (c >> 2)});
        }
        if (ok && depth >= 1) {
            grid_levels = (uint32_t)depth;
            const size_t g = (size_t)1 << grid_levels;
            cells.assign(g * g * g, 0u);
            for (const Item &it : leaves) cells[(it.z * g + it.y) * g + it.x] = it.node;
            size_t occ = 0;
            for (uint32_t v : cells) occ += v != 0;
            grid_occ = (float)occ / (float)cells.size();
        }
    }
    drk_scene *s = new (std::nothrow) drk_scene;
    if (!s) return set_error(DRK_ENOMEM, "drk_scene_create: host allocation");
    s->grid_levels = grid_levels;
    s->grid_occupancy = grid_occ;
    {
        float e = 0.0f;
        bool have = false, same = true;
        for (uint32_t i = 0; i < node_count && same; ++i) {
            const drk_octree_node &n = ob[i];
            if (n.volx[0] == DRK_NO_BRICK && n.volx[1] == DRK_NO_BRICK) continue;
            if (!have) e = n.ce[3];
            have = true;
            same = n.ce[3] == e;
        }
        // Markstein's correction needs a normal reciprocal and normal
        // quotients: the lookups use it for numerators of at least 2^-100,
        // whose quotients stay normal for extents up to 2^20.
        if (have && same && e >= 0x1p-60f && e <= 0x1p20f) s->leaf_rcp = 1.0f / e;
    }
    if (const char *w = std::getenv("DRK_WALK")) s->walk = (uint32_t)std::atoi(w);
    s->node_count = node_count;
    s->h_ob.assign(ob, ob + node_count);
    s->leafx[0] = leafx_sdf;
    s->leafx[1] = leafx_fog;
    s->lvoxc = lvoxc;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = dev_malloc(&s->d_ob, sizeof(drk_octree_node) * node_count);
    if (e == hipSuccess) e = hipMemcpy(s->d_ob, ob, sizeof(drk_octree_node) * node_count, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // Child links with the child's leaf flag folded in, so a traversal
        // step reads one word instead of the child index and then the child.
        std::vector<uint32_t> chx((size_t)node_count * 8, 0u);
        for (uint32_t i = 0; i < node_count; ++i)
            for (int c = 0; c < 8; ++c) {
                uint32_t ch = ob[i].chn[c];
                if (!ch) continue;
                const drk_octree_node &m = ob[ch];
                bool leaf = m.volx[0] != DRK_NO_BRICK || m.volx[1] != DRK_NO_BRICK;
                chx[(size_t)i * 8 + c] = ch | (leaf ? kChildLeaf : 0u);
            }
        e = dev_malloc(&s->d_chx, chx.size() * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemcpy(s->d_chx, chx.data(), chx.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && leafx_fog) {
        // Leaf box of every FOG brick: the adjoint's event reduction finds an
        // event's lookup cell from its point and this box. The reference's
        // build gives each brick to one leaf; a brick shared by leaves of
        // different boxes has no single cell, and the adjoint then adds its
        // gradient with atomics instead of events.
        std::vector<float> bce((size_t)leafx_fog * 4, 0.0f);
        std::vector<uint8_t> seen(leafx_fog, 0);
        for (uint32_t i = 0; i < node_count; ++i) {
            const uint32_t b = ob[i].volx[DRK_VOLUME_FOG];
            if (b == DRK_NO_BRICK) continue;
            if (seen[b] && std::memcmp(&bce[(size_t)b * 4], ob[i].ce, sizeof(ob[i].ce)) != 0) s->bce_ok = false;
            std::memcpy(&bce[(size_t)b * 4], ob[i].ce, sizeof(ob[i].ce));
            seen[b] = 1;
        }
        e = dev_malloc(&s->d_bce, bce.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(s->d_bce, bce.data(), bce.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !cells.empty()) {
        e = dev_malloc(&s->d_cells, cells.size() * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemcpy(s->d_cells, cells.data(), cells.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    size_t nb = (size_t)leafx_fog * lvoxc * lvoxc * lvoxc;
    if (e == hipSuccess && nb) e = dev_malloc(&s->d_fog, nb * sizeof(float));
    if (e == hipSuccess && nb) e = hipMemcpy(s->d_fog, pvol_fog, nb * sizeof(float), hipMemcpyHostToDevice);
    const size_t ns = (size_t)leafx_sdf * lvoxc * lvoxc * lvoxc;
    if (e == hipSuccess && ns) e = dev_malloc(&s->d_sdf, ns * sizeof(float));
    if (e == hipSuccess && ns) e = hipMemcpy(s->d_sdf, pvol_sdf, ns * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        drk_scene_destroy(s);
        return hip_fail(e, "drk_scene_create");
    }
    *out = s;
    return DRK_OK;
}

int drk_scene_update_fog(drk_scene *s, const float *pvol_fog, const float *qval_fog) {
    if (!s || !pvol_fog) return set_error(DRK_EINVAL, "drk_scene_update_fog: NULL argument");
    size_t nb = (size_t)s->leafx[1] * s->lvoxc * s->lvoxc * s->lvoxc;
    HIP_TRY(hipMemcpy(s->d_fog, pvol_fog, nb * sizeof(float), hipMemcpyDefault), "drk_scene_update_fog");
    if (qval_fog) {
        // qval_fog is indexed by node (host memory): patch every FOG majorant.
        for (uint32_t i = 0; i < s->node_count; ++i) s->h_ob[i].qval[DRK_VOLUME_FOG] = qval_fog[i];
        HIP_TRY(hipMemcpy(s->d_ob, s->h_ob.data(), sizeof(drk_octree_node) * s->node_count, hipMemcpyHostToDevice),
                "drk_scene_update_fog");
    }
    return DRK_OK;
}

int drk_scene_update_fog_device(drk_scene *s, const float *d_pvol_fog, void *stream) {
    if (!s || !d_pvol_fog) return set_error(DRK_EINVAL, "drk_scene_update_fog_device: NULL argument");
    const size_t nb = (size_t)s->leafx[1] * s->lvoxc * s->lvoxc * s->lvoxc;
    HIP_TRY(hipMemcpyAsync(s->d_fog, d_pvol_fog, nb * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)stream),
            "drk_scene_update_fog_device");
    return DRK_OK;
}

int drk_scene_set_reference_fog(drk_scene *s, const float *pvol_ref) {
    if (!s) return set_error(DRK_EINVAL, "drk_scene_set_reference_fog: NULL scene");
    if (!pvol_ref) {
        if (s->d_ref) {
            (void)hipDeviceSynchronize();   // a launch may still read them
            (void)hipFree(s->d_ref);
        }
        s->d_ref = nullptr;
        return DRK_OK;
    }
    const size_t nb = (size_t)s->leafx[1] * s->lvoxc * s->lvoxc * s->lvoxc;
    if (nb == 0) return set_error(DRK_EINVAL, "drk_scene_set_reference_fog: scene has no FOG bricks");
    if (!s->d_ref) HIP_TRY(dev_malloc(&s->d_ref, nb * sizeof(float)), "drk_scene_set_reference_fog");
    HIP_TRY(hipMemcpy(s->d_ref, pvol_ref, nb * sizeof(float), hipMemcpyDefault), "drk_scene_set_reference_fog");
    return DRK_OK;
}

// Whether launches on this scene walk the leaf grid: forced by
// drk_scene_set_walk, or (auto) when the grid exists and at least
// kGridMinOccupancy of its cells hold a leaf (in sparser grids most steps
// would cross empty cells that the octree walk skips a subtree at a time).
static bool scene_uses_grid(const drk_scene *s) {
    if (!s->d_cells || s->walk == DRK_WALK_OCTREE) return false;
    return s->walk == DRK_WALK_GRID || s->grid_occupancy >= kGridMinOccupancy;
}
static void set_walk_params(const drk_scene *s, RenderParams &P) {
    const bool g = scene_uses_grid(s);
    P.cells = g ? s->d_cells : nullptr;
    P.grid_levels = g ? s->grid_levels : 0u;
}

int drk_scene_set_walk(drk_scene *s, uint32_t walk) {
    if (!s) return set_error(DRK_EINVAL, "drk_scene_set_walk: NULL scene");
    if (walk > DRK_WALK_GRID) return set_error(DRK_EINVAL, "drk_scene_set_walk: unknown walk %u", walk);
    s->walk = walk;
    return DRK_OK;
}

int drk_scene_walk_info(const drk_scene *s, uint32_t *grid_levels, float *occupancy, int *uses_grid) {
    if (!s) return set_error(DRK_EINVAL, "drk_scene_walk_info: NULL scene");
    if (grid_levels) *grid_levels = s->d_cells ? s->grid_levels : 0u;
    if (occupancy) *occupancy = s->grid_occupancy;
    if (uses_grid) *uses_grid = scene_uses_grid(s) ? 1 : 0;
    return DRK_OK;
}

const drk_octree_node *drk_internal_scene_nodes(const drk_scene *s) { return s ? s->d_ob : nullptr; }
const uint32_t *drk_internal_scene_cells(const drk_scene *s) { return s && scene_uses_grid(s) ? s->d_cells : nullptr; }
uint32_t drk_internal_scene_grid_levels(const drk_scene *s) { return s && scene_uses_grid(s) ? s->grid_levels : 0u; }
const uint32_t *drk_internal_scene_links(const drk_scene *s) { return s ? s->d_chx : nullptr; }

void drk_scene_destroy(drk_scene *s) {
    if (!s) return;
    if (s->d_ob) (void)hipFree(s->d_ob);
    if (s->d_chx) (void)hipFree(s->d_chx);
    if (s->d_bce) (void)hipFree(s->d_bce);
    if (s->d_fog) (void)hipFree(s->d_fog);
    if (s->d_sdf) (void)hipFree(s->d_sdf);
    if (s->d_ref) (void)hipFree(s->d_ref);
    if (s->d_cells) (void)hipFree(s->d_cells);
    if (s->d_ev) (void)hipFree(s->d_ev);
    if (s->d_evbins) (void)hipFree(s->d_evbins);
    delete s;
}

static void sampling_basis_host(const float iv[3], float b1[3], float b2[3]) {
#pragma clang fp contract(off)
    float len = std::sqrt(iv[0] * iv[0] + 0.0f * 0.0f + iv[2] * iv[2]);
    b1[0] = -iv[2] / len;
    b1[1] = 0.0f / len;
    b1[2] = iv[0] / len;
    b2[0] = iv[1] * b1[2] - iv[2] * b1[1];
    b2[1] = iv[2] * b1[0] - iv[0] * b1[2];
    b2[2] = iv[0] * b1[1] - iv[1] * b1[0];
}

int drk_kernel_create(const drk_scene *scene, const float viewi[16], const float proji[16],
                      const drk_sampler_desc *smp, const float *pdepth, uint32_t scattevs, float msigmas,
                      float msigmaa, uint32_t tilex, uint32_t tiley, uint32_t w, uint32_t h, uint32_t flags,
                      drk_kernel **out) {
    if (!out) return set_error(DRK_EINVAL, "drk_kernel_create: out is NULL");
    *out = nullptr;
    if (!scene || !viewi || !proji || !smp) return set_error(DRK_EINVAL, "drk_kernel_create: NULL argument");
    if (tilex == 0 || tiley == 0 || w == 0 || h == 0 || tilex > 16384 || tiley > 16384)
        return set_error(DRK_EINVAL, "drk_kernel_create: bad tile/image size");
    if (smp->light_count == 0 || smp->light_count > DRK_MAX_LIGHTS || !smp->lights)
        return set_error(DRK_EINVAL, "drk_kernel_create: need 1..%d lights (lights[0] is sampled)", DRK_MAX_LIGHTS);
    if (smp->phase != DRK_PHASE_MIE && smp->phase != DRK_PHASE_HG)
        return set_error(DRK_EINVAL, "drk_kernel_create: unknown phase function %u", smp->phase);
    if (smp->phase == DRK_PHASE_HG && (smp->hg_g == 0.0f || !(std::fabs(smp->hg_g) < 1.0f)))
        return set_error(DRK_EINVAL, "drk_kernel_create: HG g must be in (-1,1) and non-zero");
    if (smp->env > DRK_ENV_MAP) return set_error(DRK_EINVAL, "drk_kernel_create: unknown environment %u", smp->env);
    if (smp->env == DRK_ENV_MAP && (!smp->env_map || smp->env_w < 2 || smp->env_h < 2))
        return set_error(DRK_EINVAL, "drk_kernel_create: env map needs >= 2x2 texels");
    if (smp->env == DRK_ENV_HOSEK && !(smp->sky_turbidity >= 1.0f && smp->sky_turbidity <= 10.0f &&
                                      smp->sky_albedo >= 0.0f && smp->sky_albedo <= 1.0f))
        return set_error(DRK_EINVAL, "drk_kernel_create: Hosek turbidity in [1,10], albedo in [0,1]");
    if (!(msigmas >= 0.0f) || !(msigmaa >= 0.0f) || !(msigmas + msigmaa > 0.0f))
        return set_error(DRK_EINVAL, "drk_kernel_create: cross sections must be >= 0 with a positive sum");
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(viewi[i]) || !std::isfinite(proji[i]))
            return set_error(DRK_EINVAL, "drk_kernel_create: non-finite camera matrix");

    drk_kernel *k = new (std::nothrow) drk_kernel;
    if (!k) return set_error(DRK_ENOMEM, "drk_kernel_create: host allocation");
    std::memset(&k->P, 0, sizeof(k->P));
    RenderParams &P = k->P;
    k->scene = const_cast<drk_scene *>(scene);
    k->tilex = tilex;
    k->tiley = tiley;
    P.ob = scene->d_ob;
    P.chx = scene->d_chx;
    P.pvol_fog = scene->d_fog;
    P.pvol_sdf = scene->d_sdf;
    P.lvoxc = scene->lvoxc;
    P.lvoxc3 = scene->lvoxc * scene->lvoxc * scene->lvoxc;
    P.lvoxm1 = (float)scene->lvoxc - 1.0f;
    P.leaf_rcp = scene->leaf_rcp;
    P.w = w;
    P.h = h;
    P.scattevs = scattevs;
    P.sigma_s = msigmas;
    P.sigma_e = msigmaa + msigmas;
    P.phase = smp->phase;
    P.hg_g = smp->hg_g;
    P.env = smp->env;
    KernelConsts &C = k->C;
    std::memset(&C, 0, sizeof(C));
    std::memcpy(C.viewi, viewi, sizeof(C.viewi));
    std::memcpy(C.proji, proji, sizeof(C.proji));
    C.light_count = smp->light_count;
    for (uint32_t l = 0; l < smp->light_count; ++l) {
        for (int c = 0; c < 3; ++c) {
            C.sun_dir[l][c] = smp->lights[l].direction[c];
            C.sun_color[l][c] = smp->lights[l].color[c];
        }
        C.sun_cos[l] = cosf(smp->lights[l].angle);   // SunLight ctor, KernelSampler.cpp:306
    }
    {
#pragma clang fp contract(off)
        float ctm = C.sun_cos[0];
        C.sun_pdf0 = 1.0f / (2.0f * kPi * (1.0f - ctm * ctm));
    }
    sampling_basis_host(C.sun_dir[0], C.sun_b1, C.sun_b2);
    if (smp->env == DRK_ENV_HOSEK) {
        // RenderKernel::Initialize, src/kernel.cpp:521-526, with the turbidity
        // and albedo made parameters.
        float th = acosf(C.sun_dir[0][2]);
        float se = 1.57079632679489661923f - th;
        double cfg[3][9], rad[3];
        hosek_state((double)smp->sky_turbidity, (double)smp->sky_albedo, (double)se, cfg, rad);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 9; ++j) C.sky.cfg[i][j] = (float)cfg[i][j];
            C.sky.rad[i] = (float)rad[i];
        }
        for (int c = 0; c < 3; ++c) C.sky.dir[c] = C.sun_dir[0][c];
    }
    {
        // Far plane: -z/w of (0,0,1,1)*proji (src/kernel.cpp:574-575).
        float zf_z = 0.0f * proji[0 * 4 + 2] + 0.0f * proji[1 * 4 + 2] + 1.0f * proji[2 * 4 + 2] + 1.0f * proji[3 * 4 + 2];
        float zf_w = 0.0f * proji[0 * 4 + 3] + 0.0f * proji[1 * 4 + 3] + 1.0f * proji[2 * 4 + 3] + 1.0f * proji[3 * 4 + 3];
        C.clip_end = -zf_z / zf_w;
    }

    hipError_t e = hipStreamCreateWithFlags(&k->stream, hipStreamNonBlocking);
    size_t tile_bytes = (size_t)tilex * tiley * 16;
    if (e == hipSuccess) e = dev_malloc(&k->d_out, 4 * tile_bytes);
    if (e == hipSuccess) e = dev_malloc(&k->d_counter, kLaunchBlockBytes);
    if (e == hipSuccess) e = dev_malloc(&k->d_counts, 64);
    if (e == hipSuccess && smp->env == DRK_ENV_MAP) {
        size_t eb = (size_t)smp->env_w * smp->env_h * 16;
        e = dev_malloc(&k->d_env, eb);
        if (e == hipSuccess) e = hipMemcpy(k->d_env, smp->env_map, eb, hipMemcpyHostToDevice);
        C.env_map = k->d_env;
        C.env_w = smp->env_w;
        C.env_h = smp->env_h;
    }
    if (e == hipSuccess) e = dev_malloc(&k->d_kc, sizeof(KernelConsts));
    if (e == hipSuccess) e = hipMemcpy(k->d_kc, &C, sizeof(KernelConsts), hipMemcpyHostToDevice);
    P.kc = k->d_kc;
    // The depth map serves compositing (flag) and the Shadow pass (always).
    k->depthcomp = (flags & DRK_KERNEL_DEPTHCOMP) != 0;
    if (e == hipSuccess && pdepth) {
        e = dev_malloc(&k->d_depth, (size_t)w * h * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(k->d_depth, pdepth, (size_t)w * h * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        drk_kernel_destroy(k);
        return hip_fail(e, "drk_kernel_create");
    }
    k->h_phb = static_cast<float *>(std::aligned_alloc(64, ((4 * tile_bytes + 63) / 64) * 64));
    if (!k->h_phb) {
        drk_kernel_destroy(k);
        return set_error(DRK_ENOMEM, "drk_kernel_create: host result buffers");
    }
    std::memset(k->h_phb, 0, 4 * tile_bytes);
    uint32_t cus = 0;
    if (drk_device_info(nullptr, 0, &cus) != DRK_OK || cus == 0) cus = 1;
    k->cu_count = cus;
    k->blocks_per_cu = (uint32_t)render_blocks_per_cu(kModeForwardSq);
    {
        // One adjoint gradient copy per XCD (each XCD has a private L2).
        int dev = 0, xcc = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || xcc < 1)
            xcc = 1;
        k->xcc = (uint32_t)xcc;
        k->grad_copies = (uint32_t)xcc;
        if (const char *g = std::getenv("DRK_GRAD_COPIES")) k->grad_copies = std::max(1, std::atoi(g));
        k->bands = (uint32_t)std::min<int>(xcc, (int)kMaxBands);
        if (const char *b = std::getenv("DRK_BANDS")) k->bands = std::max(1u, std::min<uint32_t>(std::atoi(b), kMaxBands));
    }
    {
        // The allocation-free render's fixed scratch: its device address,
        // looked up once per device (here, with the kernel's device current).
        void *p = nullptr;
        if (scratch_address(&p) == hipSuccess) k->scratch = static_cast<char *>(p);
        else (void)hipGetLastError();
    }
    k->status = k->d_counter;
    if (const char *c = std::getenv("DRK_CHUNK")) k->chunk_override = (uint32_t)std::atoi(c);
    if (const char *c = std::getenv("DRK_MAX_CHUNKS")) k->max_chunks = std::max(1, std::atoi(c));
    if (const char *g = std::getenv("DRK_GRID")) k->grid_override = (uint32_t)std::atoi(g);
    *out = k;
    return DRK_OK;
}

int drk_kernel_set_seed(drk_kernel *k, uint64_t seed, uint32_t sample_base) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_set_seed: NULL kernel");
    k->seed = seed;
    k->sample_base = sample_base;
    return DRK_OK;
}

int drk_kernel_set_tuning(drk_kernel *k, uint32_t chunk, uint32_t grid_blocks) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_set_tuning: NULL kernel");
    k->chunk_override = chunk;
    k->grid_override = grid_blocks;
    return DRK_OK;
}

// Adjoint-mode inputs of render_impl.
struct AdjointArgs {
    const float *adj_l = nullptr, *adj_e = nullptr;
    float *grad = nullptr;
};

// Rendered extent of a call: the tile rounded down to even sizes (the
// reference renders 2x2 packets, src/kernel.cpp:453) and clipped to the image.
static void rendered_extent(const drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                            uint32_t &rx, uint32_t &ry) {
    rx = x0 < k->P.w ? std::min(tilex & ~1u, k->P.w - x0) : 0u;
    ry = y0 < k->P.h ? std::min(tiley & ~1u, k->P.h - y0) : 0u;
}

// One job of a views call (several launches as one batched launch): what
// render_impl plans for it instead of launching, and what the call shares
// among its jobs (event slots, gradient copies).
struct ViewsPlan {
    uint32_t ev_cap = 0;          // event slots of the call (0: atomics only)
    float *gradx = nullptr;       // per-XCD gradient copies (copies > 1)
    uint32_t copies = 1;
    bool planned = false;         // false: the job renders nothing (empty extent)
    RenderParams P;
    WorkParams W;
};

// Event buffer of the scene's adjoint calls: 40 bytes per slot (event +
// sorted copy), sized to 16 events per sample of `samples` within the
// budget (a multiple of 64 slots, which keeps every plane 16-byte aligned);
// kept for later calls and reallocated when a call wants more. Returns the
// slots to use (0: no events; every contribution then goes to the gradient
// with atomics, the same gradient up to the order of float additions, so an
// allocation failure is not an error).
static uint32_t ensure_event_buffer(drk_scene *sc, uint64_t samples) {
    uint32_t shift = 0, nbins = 0;
    const bool binnable = event_bins(sc->leafx[1], sc->lvoxc * sc->lvoxc * sc->lvoxc, shift, nbins) && sc->bce_ok;
    if (!binnable) return 0;
    size_t budget = sc->ev_budget;
    if (budget == SIZE_MAX) {
        // A quarter of the memory that is free now (other allocators,
        // torch's caching one among them, may hold much of the device),
        // plus what the current buffer already holds.
        size_t fr = 0, tot = 0;
        const size_t held = sc->d_ev ? (size_t)sc->ev_cap * kEventSlotBytes : 0;
        budget = hipMemGetInfo(&fr, &tot) == hipSuccess ? (std::min(fr + held, tot)) / 4 : 0;
    }
    const uint64_t want = std::min<uint64_t>(samples * 16, 0xffffffffull);
    const uint64_t cap = std::min<uint64_t>(want, budget / kEventSlotBytes) & ~63ull;
    if (cap > sc->ev_cap) {
        if (sc->d_ev) (void)hipFree(sc->d_ev);
        if (sc->d_evbins) (void)hipFree(sc->d_evbins);
        sc->d_ev = nullptr;
        sc->d_evbins = nullptr;
        sc->ev_cap = 0;
        hipError_t e1 = dev_malloc((void **)&sc->d_ev, (size_t)cap * kEventSlotBytes);
        hipError_t e2 = e1 == hipSuccess ? dev_malloc((void **)&sc->d_evbins, kEvTableWords * sizeof(uint32_t)) : e1;
        if (e1 == hipSuccess && e2 == hipSuccess) {
            sc->ev_cap = (uint32_t)cap;
        } else {
            (void)hipGetLastError();
            if (sc->d_ev) (void)hipFree(sc->d_ev);
            sc->d_ev = nullptr;
            sc->d_evbins = nullptr;
        }
    }
    return cap > 0 && sc->ev_cap > 0 ? sc->ev_cap : 0u;
}

// Parameters of the event reduction over `cap` slots of the scene's buffer
// into grad, for calls that rendered `samples` pixel samples in total.
static void event_reduce_params(const drk_scene *sc, uint32_t cap, uint64_t samples, float *grad, EventReduce &E) {
    uint32_t shift = 0, nbins = 0;
    const uint32_t lvoxc3 = sc->lvoxc * sc->lvoxc * sc->lvoxc;
    (void)event_bins(sc->leafx[1], lvoxc3, shift, nbins);
    std::memset(&E, 0, sizeof(E));
    E.ev_base = static_cast<char *>(sc->d_ev);
    E.bins = sc->d_evbins;
    E.ev_cap = cap;
    E.nbins = nbins;
    E.shift = shift;
    // Accumulation blocks of 2^14 events, or 2^12 for calls of fewer
    // than 2^24 samples, whose few blocks would otherwise leave CUs idle
    // (256^3 frame: 2^14 3.38 ms, 2^12 3.72 ms; sixteen 256^2 fitting views
    // reduced one by one: step 42.1 ms at 2^14, 40.8 ms at 2^12).
    E.acc_log2 = samples < (1ull << 24) ? 12u : 14u;
    E.nbricks = sc->leafx[1];
    E.lvoxc = sc->lvoxc;
    E.lvoxc3 = lvoxc3;
    E.bce = reinterpret_cast<const float4 *>(sc->d_bce);
    E.grad = grad;
}

// Per-XCD gradient copies of an adjoint launch: one per XCD by default, while
// they stay small next to HBM (<= 4 GiB).
static uint32_t grad_copies_for(const drk_kernel *k, size_t gn) {
    uint32_t copies = k->grad_copies;
    while (copies > 1 && (double)copies * gn * sizeof(float) > 4.0 * (1ull << 30)) copies >>= 1;
    return copies;
}

// The kernel's per-XCD gradient copies, grown to copies * gn floats and
// cleared on stream s.
static hipError_t clear_grad_copies(drk_kernel *k, uint32_t copies, size_t gn, hipStream_t s, float **out) {
    const size_t need = (size_t)copies * gn * sizeof(float);
    if (need > k->gradx_bytes) {
        if (k->d_gradx) (void)hipFree(k->d_gradx);
        k->d_gradx = nullptr;
        k->gradx_bytes = 0;
        hipError_t e = dev_malloc(&k->d_gradx, need);
        if (e != hipSuccess) return e;
        k->gradx_bytes = need;
    }
    *out = k->d_gradx;
    return launch_fill(k->d_gradx, (size_t)copies * gn, 0.0f, s);
}

static int render_impl(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
                       uint64_t seed, uint32_t sample_base, float *d_l, float *d_e, float *d_sql, float *d_sqe,
                       hipStream_t s, bool count, bool shadow = false, const AdjointArgs *adj = nullptr,
                       float *rec = nullptr, bool noalloc = false, bool weighted = false,
                       ViewsPlan *vp = nullptr) {
    if (tilex == 0 || tiley == 0 || tilex > k->tilex || tiley > k->tiley)
        return set_error(DRK_EINVAL, "render: tile %ux%u exceeds the kernel's %ux%u", tilex, tiley, k->tilex, k->tiley);
    if (samples == 0) return set_error(DRK_EINVAL, "render: samples must be > 0");
    if ((uint64_t)x0 + tilex > 0xffffffffull || (uint64_t)y0 + tiley > 0xffffffffull)
        return set_error(DRK_EINVAL, "render: tile origin overflows");
    if ((uint64_t)sample_base + samples > 0xffffffffull) return set_error(DRK_EINVAL, "render: sample counter overflows");
    RenderParams P = k->P;
    set_walk_params(k->scene, P);
    P.x0 = x0;
    P.y0 = y0;
    P.rx = tilex & ~1u;   // K_ParallelRender iterates rx/2 packets of 2 (src/kernel.cpp:453)
    P.ry = tiley & ~1u;
    P.tilex = tilex;
    P.samples = samples;
    P.sample_base = sample_base;
    P.key1 = rng_key1(seed);
    P.shadow_mode = shadow ? 1u : 0u;
    P.depth = (shadow || k->depthcomp) ? k->d_depth : nullptr;
    const bool sumsq = d_sql && d_sqe;
    const int mode = adj ? (k->adj_estimator == DRK_ADJ_NULL_COLLISION ? kModeAdjointNull : kModeAdjoint)
                     : weighted ? kModeWeighted : count ? kModeCount : rec ? kModeRecord
                     : sumsq ? kModeForwardSq : kModeForward;
    P.pvol_ref = weighted ? k->scene->d_ref : nullptr;
    // Clip the rendered extent to the image (gm mask, src/kernel.cpp:472-474).
    uint32_t rx, ry;
    rendered_extent(k, x0, y0, tilex, tiley, rx, ry);
    P.rx = rx;
    P.ry = ry;
    // Pixels outside the rendered extent (odd tiles, or outside the image)
    // are zeroed; the reference leaves them untouched.
    if (!adj && (rx != tilex || ry != tiley)) {
        size_t tile_n = (size_t)tilex * tiley * 4;
        HIP_TRY(launch_fill(d_l, tile_n, 0.0f, s), "render: clear");
        HIP_TRY(launch_fill(d_e, tile_n, 0.0f, s), "render: clear");
        if (sumsq) {
            HIP_TRY(launch_fill(d_sql, tile_n, 0.0f, s), "render: clear");
            HIP_TRY(launch_fill(d_sqe, tile_n, 0.0f, s), "render: clear");
        }
    }
    if (rx == 0 || ry == 0) return DRK_OK;

    const uint32_t npix = rx * ry;
    const uint32_t bpc = mode == kModeForwardSq ? k->blocks_per_cu : (uint32_t)render_blocks_per_cu(mode);
    uint32_t grid = k->grid_override ? k->grid_override : k->cu_count * std::max(1u, bpc);
    uint32_t chunk = k->chunk_override;
    if (chunk == 0) {
        // Short units keep the drain at the end of a launch short (the last
        // units handed out are single samples of a pixel rather than
        // several samples of an expensive one), and the waves' unit reserves
        // make a refill cheap. Up to kMaxAutoChunks chunks (one sample each
        // for up to that many samples per pixel), fewer where the per-chunk
        // partial sums would exceed kMaxPartialBytes.
        const uint64_t bytes_per_chunk = (uint64_t)npix * 16 * partial_buffers(mode);
        uint64_t nch = std::min<uint64_t>(samples, k->max_chunks);
        nch = std::min<uint64_t>(nch, std::max<uint64_t>(1, kMaxPartialBytes / std::max<uint64_t>(bytes_per_chunk, 1)));
        nch = std::max<uint64_t>(1, nch);
        chunk = (uint32_t)((samples + nch - 1) / nch);
    }
    chunk = std::max(1u, std::min(chunk, samples));
    uint32_t nfull, nchunks;
    plan_chunks(samples, chunk, k->chunk_override == 0, nfull, nchunks);
    // Unit indices stay below 2^32 with room for the fetches that run past a
    // band's end (at most a wave's lanes plus its reserve per wave and band).
    if ((uint64_t)npix * nchunks > 0xf0000000ull) return set_error(DRK_EINVAL, "render: too many work units");
    WorkParams W;
    std::memset(&W, 0, sizeof(W));
    W.counter = noalloc ? reinterpret_cast<uint32_t *>(k->scratch) : k->d_counter;
    W.round_limit = k->round_limit;
    W.grad_copies = 1;
    W.grad_nxcc = k->xcc;
    W.nbands = k->bands;
    W.npix = npix;
    set_bands(W, rx);
    W.chunk = chunk;
    W.nfull = nfull;
    W.nchunks = nchunks;
    W.total_units = npix * nchunks;
    W.out_l = d_l;
    W.out_e = d_e;
    W.sq_l = sumsq ? d_sql : nullptr;
    W.sq_e = sumsq ? d_sqe : nullptr;
    W.counts = k->d_counts;
    W.rec = rec;
    uint32_t copies = 1;
    const size_t gn = (size_t)k->scene->leafx[1] * k->P.lvoxc3;
    EventReduce E;
    std::memset(&E, 0, sizeof(E));
    drk_scene *sc = k->scene;
    if (adj) {
        // Events: into the scene's buffer, reduced at the end of this call,
        // or, inside a gradient window, appended and reduced when the window
        // ends. Without room for events every contribution goes to the
        // gradient with atomics instead.
        const uint32_t cap = vp              ? vp->ev_cap
                             : sc->ev_window ? sc->ev_window_cap
                                             : ensure_event_buffer(sc, (uint64_t)npix * samples);
        if (cap > 0) {
            W.ev_base = static_cast<char *>(sc->d_ev);
            W.ev_count = reinterpret_cast<unsigned long long *>(sc->d_evbins);
            W.ev_cap = cap;
            if (!sc->ev_window && !vp) {
                event_reduce_params(sc, cap, (uint64_t)npix * samples, adj->grad, E);
                HIP_TRY(launch_fill(reinterpret_cast<float *>(sc->d_evbins), 2, 0.0f, s), "adjoint: clear");
            }
        }
        W.adj_l = adj->adj_l;
        W.adj_e = adj->adj_e;
        W.grad = adj->grad;
        // Per-XCD copies of the gradient (the atomics' targets), added to
        // grad after the launch.
        copies = vp ? vp->copies : grad_copies_for(k, gn);
        if (copies > 1) {
            float *gx = vp ? vp->gradx : nullptr;
            if (!vp) HIP_TRY(clear_grad_copies(k, copies, gn, s, &gx), "adjoint: gradient copies");
            W.grad = gx;
            W.grad_copies = copies;
            W.grad_stride = gn;
        }
    }
    if (nchunks > 1 && !adj && noalloc) {
        const size_t need = (size_t)partial_buffers(mode) * nchunks * npix * 16;
        if (need > kScratchBytes - kScratchPartialOffset)
            return set_error(DRK_EINVAL,
                             "render (no allocation): %u pixels x %u sample chunks need %zu bytes of partial sums; "
                             "the fixed scratch holds %zu (use a smaller tile or fewer chunks, drk_kernel_set_tuning)",
                             npix, nchunks, need, (size_t)(kScratchBytes - kScratchPartialOffset));
        W.partial = reinterpret_cast<float *>(k->scratch + kScratchPartialOffset);
    } else if (nchunks > 1 && !adj) {
        size_t need = (size_t)partial_buffers(mode) * nchunks * npix * 16;
        if (need > k->partial_bytes) {
            if (k->d_partial) (void)hipFree(k->d_partial);
            k->d_partial = nullptr;
            k->partial_bytes = 0;
            HIP_TRY(dev_malloc(&k->d_partial, need), "render: partial sums");
            k->partial_bytes = need;
        }
        W.partial = k->d_partial;
    }
    // Mailboxes of the escape service (forward modes; the allocation-free
    // entry computes escape radiance in place).
    if (!noalloc && !vp &&
        (mode == kModeForward || mode == kModeForwardSq || mode == kModeCount || mode == kModeWeighted ||
         mode == kModeRecord)) {
        const size_t need = (size_t)grid * render_block_size() * 32;
        if (need > k->mail_bytes) {
            if (k->d_mail) (void)hipFree(k->d_mail);
            k->d_mail = nullptr;
            k->mail_bytes = 0;
            HIP_TRY(dev_malloc(&k->d_mail, need), "render: escape mailboxes");
            k->mail_bytes = need;
        }
        W.esc_mail = static_cast<float4 *>(k->d_mail);
    }
    // Camera rays: every sample of a pixel reuses its one primary ray
    // (src/kernel.cpp:556), so a pass before the render computes them once
    // per pixel instead of once per path (the allocation-free entry computes
    // them per path).
    if (!noalloc && npix > 0) {
        const size_t need = (size_t)npix * 16;
        if (need > k->cam_bytes) {
            if (k->d_cam) (void)hipFree(k->d_cam);
            k->d_cam = nullptr;
            k->cam_bytes = 0;
            HIP_TRY(dev_malloc(&k->d_cam, need), "render: camera rays");
            k->cam_bytes = need;
        }
        W.cam = static_cast<float4 *>(k->d_cam);
    }
    if (vp) {
        // Planned for a batched launch by the caller.
        k->status = W.counter;
        vp->P = P;
        vp->W = W;
        vp->planned = true;
        return DRK_OK;
    }
    if (count) HIP_TRY(hipMemsetAsync(k->d_counts, 0, 16, s), "render: counters");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (k->timing && !count && !noalloc) {
        if (k->ev_free.empty()) {
            HIP_TRY(hipEventCreate(&e0), "render: event");
            HIP_TRY(hipEventCreate(&e1), "render: event");
        } else {
            e0 = k->ev_free.back().first;
            e1 = k->ev_free.back().second;
            k->ev_free.pop_back();
        }
        k->ev_used.emplace_back(e0, e1);
    }
    k->status = W.counter;
    HIP_TRY(launch_render(P, W, grid, mode, s, e0, e1), "render: launch");
    if (adj && copies > 1) HIP_TRY(launch_grad_sum(k->d_gradx, copies, gn, adj->grad, s), "adjoint: gradient sum");
    if (adj && E.ev_cap) HIP_TRY(launch_event_reduce(E, s), "adjoint: event reduction");
    if (adj && sc->ev_window) sc->ev_window_samples += (uint64_t)npix * samples;
    return DRK_OK;
}

// After a synchronous call: the status word of the kernel's last launch
// (the stream has been synchronised), DRK_ELIMIT if a wave hit the round limit.
static int check_limit(drk_kernel *k, hipStream_t s, const char *fn) {
    if (k->round_limit == 0) return DRK_OK;
    uint32_t st = 0;
    HIP_TRY(hipMemcpyAsync(&st, k->status + kStatusWord, sizeof(st), hipMemcpyDeviceToHost, s), fn);
    HIP_TRY(hipStreamSynchronize(s), fn);
    if (st) return set_error(DRK_ELIMIT, "%s: a wave reached the limit of %u rounds; the sums are incomplete", fn, k->round_limit);
    return DRK_OK;
}

int drk_kernel_render_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                             uint32_t samples, uint64_t seed, uint32_t sample_base, float *d_lights, float *d_env,
                             float *d_sumsq_lights, float *d_sumsq_env, void *stream) {
    if (!k || !d_lights || !d_env) return set_error(DRK_EINVAL, "drk_kernel_render_device: NULL argument");
    if ((d_sumsq_lights == nullptr) != (d_sumsq_env == nullptr))
        return set_error(DRK_EINVAL, "drk_kernel_render_device: pass both sum-of-squares buffers or neither");
    hipStream_t s = (hipStream_t)stream;   // NULL: the default (null) stream, as in HIP
    return render_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_lights, d_env, d_sumsq_lights,
                       d_sumsq_env, s, false);
}

int drk_kernel_render_device_noalloc(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                     uint32_t samples, uint64_t seed, uint32_t sample_base, float *d_lights,
                                     float *d_env, float *d_sumsq_lights, float *d_sumsq_env, void *stream) {
    if (!k || !d_lights || !d_env) return set_error(DRK_EINVAL, "drk_kernel_render_device_noalloc: NULL argument");
    if ((d_sumsq_lights == nullptr) != (d_sumsq_env == nullptr))
        return set_error(DRK_EINVAL, "drk_kernel_render_device_noalloc: pass both sum-of-squares buffers or neither");
    if (!k->scratch) return set_error(DRK_EDEVICE, "drk_kernel_render_device_noalloc: no device scratch address");
    hipStream_t s = (hipStream_t)stream;   // NULL: the default (null) stream, as in HIP
    return render_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_lights, d_env, d_sumsq_lights,
                       d_sumsq_env, s, false, false, nullptr, nullptr, true);
}

int drk_kernel_render_weighted_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                      uint32_t samples, uint64_t seed, uint32_t sample_base, float *d_lights,
                                      float *d_env, void *stream) {
    if (!k || !d_lights || !d_env) return set_error(DRK_EINVAL, "drk_kernel_render_weighted_device: NULL argument");
    if (!k->scene->d_ref)
        return set_error(DRK_EINVAL, "drk_kernel_render_weighted_device: the scene has no reference densities");
    if (k->scene->leafx[0] != 0)
        return set_error(DRK_EINVAL, "drk_kernel_render_weighted_device: scenes with SDF leaves are not supported");
    hipStream_t s = (hipStream_t)stream;   // NULL: the default (null) stream, as in HIP
    return render_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_lights, d_env, nullptr, nullptr, s,
                       false, false, nullptr, nullptr, false, true);
}

int drk_kernel_set_adjoint_estimator(drk_kernel *k, uint32_t estimator) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_set_adjoint_estimator: NULL kernel");
    if (estimator != DRK_ADJ_SEGMENT && estimator != DRK_ADJ_NULL_COLLISION)
        return set_error(DRK_EINVAL, "drk_kernel_set_adjoint_estimator: unknown estimator %u", estimator);
    k->adj_estimator = estimator;
    return DRK_OK;
}

int drk_kernel_record_bytes(const drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                            uint32_t samples, size_t *bytes) {
    if (!k || !bytes) return set_error(DRK_EINVAL, "drk_kernel_record_bytes: NULL argument");
    uint32_t rx, ry;
    rendered_extent(k, x0, y0, tilex, tiley, rx, ry);
    *bytes = (size_t)rx * ry * samples * kRecordBytes;
    return DRK_OK;
}

static int check_record(const drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                        uint32_t samples, const void *rec, size_t rec_bytes, const char *fn) {
    if (!rec) return set_error(DRK_EINVAL, "%s: NULL record", fn);
    if ((uintptr_t)rec % 16) return set_error(DRK_EINVAL, "%s: record must be 16-byte aligned", fn);
    size_t need = 0;
    drk_kernel_record_bytes(k, x0, y0, tilex, tiley, samples, &need);
    if (rec_bytes < need) return set_error(DRK_EINVAL, "%s: record holds %zu bytes, needs %zu", fn, rec_bytes, need);
    return DRK_OK;
}

int drk_kernel_render_record_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                    uint32_t samples, uint64_t seed, uint32_t sample_base, float *d_lights,
                                    float *d_env, void *d_record, size_t record_bytes, void *stream) {
    if (!k || !d_lights || !d_env) return set_error(DRK_EINVAL, "drk_kernel_render_record_device: NULL argument");
    if (k->P.scattevs > 64)
        return set_error(DRK_EINVAL, "drk_kernel_render_record_device: scattevs > 64 is not supported");
    int rc = check_record(k, x0, y0, tilex, tiley, samples, d_record, record_bytes, "drk_kernel_render_record_device");
    if (rc != DRK_OK) return rc;
    hipStream_t s = (hipStream_t)stream;   // NULL: the default (null) stream, as in HIP
    return render_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_lights, d_env, nullptr, nullptr, s,
                       false, false, nullptr, static_cast<float *>(d_record));
}

int drk_kernel_render(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_render: NULL kernel");
    size_t tb = (size_t)k->tilex * k->tiley * 16;
    char *base = reinterpret_cast<char *>(k->d_out);
    float *b[4] = {(float *)base, (float *)(base + tb), (float *)(base + 2 * tb), (float *)(base + 3 * tb)};
    int rc = render_impl(k, x0, y0, tilex, tiley, samples, k->seed, k->sample_base, b[0], b[1], b[2], b[3],
                         k->stream, false);
    if (rc != DRK_OK) return rc;
    HIP_TRY(hipMemcpyAsync(k->h_phb, k->d_out, 4 * tb, hipMemcpyDeviceToHost, k->stream), "drk_kernel_render: copy-out");
    HIP_TRY(hipStreamSynchronize(k->stream), "drk_kernel_render: sync");
    rc = check_limit(k, k->stream, "drk_kernel_render");
    if (rc != DRK_OK) return rc;
    k->sample_base += samples;
    return DRK_OK;
}

// Counting pass for the algorithmic-bytes model: same launch with per-lane
// counters of node-record reads and brick lookups (identical RNG keys, so the
// counts describe the timed run exactly).
int drk_kernel_count(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
                     uint64_t seed, uint32_t sample_base, unsigned long long counts[2]) {
    if (!k || !counts) return set_error(DRK_EINVAL, "drk_kernel_count: NULL argument");
    size_t tb = (size_t)k->tilex * k->tiley * 16;
    char *base = reinterpret_cast<char *>(k->d_out);
    int rc = render_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, (float *)base, (float *)(base + tb),
                         (float *)(base + 2 * tb), (float *)(base + 3 * tb), k->stream, true);
    if (rc != DRK_OK) return rc;
    HIP_TRY(hipMemcpyAsync(counts, k->d_counts, 16, hipMemcpyDeviceToHost, k->stream), "drk_kernel_count: copy");
    HIP_TRY(hipStreamSynchronize(k->stream), "drk_kernel_count: sync");
    return check_limit(k, k->stream, "drk_kernel_count");
}

static int adjoint_impl(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
                        uint64_t seed, uint32_t sample_base, const float *d_adj_lights, const float *d_adj_env,
                        float *d_grad_fog, const void *rec, void *stream, const char *fn) {
    if (!k || !d_adj_lights || !d_adj_env || !d_grad_fog) return set_error(DRK_EINVAL, "%s: NULL argument", fn);
    if (k->P.scattevs > 64) return set_error(DRK_EINVAL, "%s: scattevs > 64 is not supported", fn);
    if (k->depthcomp) return set_error(DRK_EINVAL, "%s: depth compositing is not differentiated", fn);
    if (k->scene->leafx[1] == 0) return set_error(DRK_EINVAL, "%s: scene has no FOG bricks", fn);
    if (k->scene->leafx[0] != 0)
        return set_error(DRK_EINVAL, "%s: scenes with SDF leaves are not differentiated", fn);
    AdjointArgs A;
    A.adj_l = d_adj_lights;
    A.adj_e = d_adj_env;
    A.grad = d_grad_fog;
    hipStream_t s = (hipStream_t)stream;   // NULL: the default (null) stream, as in HIP
    return render_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, nullptr, nullptr, nullptr, nullptr, s,
                       false, false, &A, const_cast<float *>(static_cast<const float *>(rec)));
}

int drk_kernel_adjoint_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                              uint32_t samples, uint64_t seed, uint32_t sample_base, const float *d_adj_lights,
                              const float *d_adj_env, float *d_grad_fog, void *stream) {
    return adjoint_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_adj_lights, d_adj_env, d_grad_fog,
                        nullptr, stream, "drk_kernel_adjoint_device");
}

int drk_kernel_adjoint_record_device(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley,
                                     uint32_t samples, uint64_t seed, uint32_t sample_base,
                                     const float *d_adj_lights, const float *d_adj_env, float *d_grad_fog,
                                     const void *d_record, size_t record_bytes, void *stream) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_adjoint_record_device: NULL kernel");
    int rc = check_record(k, x0, y0, tilex, tiley, samples, d_record, record_bytes, "drk_kernel_adjoint_record_device");
    if (rc != DRK_OK) return rc;
    return adjoint_impl(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_adj_lights, d_adj_env, d_grad_fog,
                        d_record, stream, "drk_kernel_adjoint_record_device");
}

// Several record-forward or adjoint calls as one batched launch: each job is
// planned by render_impl exactly as its own call would be (tile, chunks,
// record, buffers of its kernel), the plans go to a job table on the device,
// and one k_render_batch runs them all; then the per-job chunk reductions
// (forward) or the shared gradient sum and event reduction (adjoint).
static_assert(sizeof(RenderParams) % alignof(WorkParams) == 0, "the job table's WorkParams follow its RenderParams");
static int views_run(const drk_view_job *jobs, uint32_t nj, float *d_grad, uint64_t total, hipStream_t s);
static int views_impl(const drk_view_job *jobs, uint32_t nj, float *d_grad, void *stream, const char *fn) {
    const bool adjoint = d_grad != nullptr;
    if (!jobs || nj == 0) return set_error(DRK_EINVAL, "%s: no jobs", fn);
    drk_kernel *k0 = jobs[0].kernel;
    for (uint32_t j = 0; j < nj; ++j) {
        const drk_view_job &J = jobs[j];
        if (!J.kernel || !J.d_lights || !J.d_env) return set_error(DRK_EINVAL, "%s: job %u: NULL argument", fn, j);
        for (uint32_t i = 0; i < j; ++i)
            if (jobs[i].kernel == J.kernel) return set_error(DRK_EINVAL, "%s: jobs %u and %u share a kernel", fn, i, j);
        if (J.kernel->P.scattevs > 64) return set_error(DRK_EINVAL, "%s: job %u: scattevs > 64 is not supported", fn, j);
        int rc = check_record(J.kernel, J.x0, J.y0, J.tilex, J.tiley, J.samples, J.d_record, J.record_bytes, fn);
        if (rc != DRK_OK) return rc;
        if (adjoint) {
            const drk_kernel *k = J.kernel;
            if (k->scene != k0->scene) return set_error(DRK_EINVAL, "%s: job %u: the jobs' kernels need one scene", fn, j);
            if (k->adj_estimator != k0->adj_estimator)
                return set_error(DRK_EINVAL, "%s: job %u: the jobs' kernels need one adjoint estimator", fn, j);
            if (k->depthcomp) return set_error(DRK_EINVAL, "%s: depth compositing is not differentiated", fn);
        }
    }
    hipStream_t s = (hipStream_t)stream;   // NULL: the default (null) stream, as in HIP
    drk_scene *sc = k0->scene;
    if (adjoint) {
        if (sc->leafx[1] == 0) return set_error(DRK_EINVAL, "%s: scene has no FOG bricks", fn);
        if (sc->leafx[0] != 0) return set_error(DRK_EINVAL, "%s: scenes with SDF leaves are not differentiated", fn);
    }
    auto job_samples = [&](uint32_t j) {
        uint32_t rx, ry;
        rendered_extent(jobs[j].kernel, jobs[j].x0, jobs[j].y0, jobs[j].tilex, jobs[j].tiley, rx, ry);
        return (uint64_t)rx * ry * jobs[j].samples;
    };
    uint64_t total = 0;
    for (uint32_t j = 0; j < nj; ++j) total += job_samples(j);
    if (adjoint && !sc->ev_window && nj > 1 && total) {
        // Groups of consecutive jobs whose events (16 slots per sample, as a
        // single call sizes them) fit the event buffer, one launch and one
        // reduction per group: events beyond the buffer would go to float
        // atomics, many times slower (a 512^3 cloud's 8 full-HD views in one
        // launch with a quarter of the free memory for events: 557 ms; in
        // groups of 3: 50.5 ms).
        const uint64_t cap = ensure_event_buffer(sc, total);
        if (cap > 0 && cap < 16 * total) {
            for (uint32_t j0 = 0; j0 < nj;) {
                uint32_t j1 = j0 + 1;
                uint64_t g = job_samples(j0);
                while (j1 < nj && 16 * (g + job_samples(j1)) <= cap) g += job_samples(j1++);
                int rc = views_run(jobs + j0, j1 - j0, d_grad, g, s);
                if (rc != DRK_OK) return rc;
                j0 = j1;
            }
            return DRK_OK;
        }
    }
    return views_run(jobs, nj, d_grad, total, s);
}

// One batched launch of views jobs (validated by views_impl) that render
// `total` pixel samples, with its reductions.
static int views_run(const drk_view_job *jobs, uint32_t nj, float *d_grad, uint64_t total, hipStream_t s) {
    const bool adjoint = d_grad != nullptr;
    drk_kernel *k0 = jobs[0].kernel;
    drk_scene *sc = k0->scene;
    const int mode = !adjoint ? kModeRecord
                     : k0->adj_estimator == DRK_ADJ_NULL_COLLISION ? kModeAdjointNull : kModeAdjoint;
    ViewsPlan vp;
    EventReduce E;
    std::memset(&E, 0, sizeof(E));
    const size_t gn = (size_t)sc->leafx[1] * k0->P.lvoxc3;
    if (adjoint && total) {
        vp.ev_cap = sc->ev_window ? sc->ev_window_cap : ensure_event_buffer(sc, total);
        if (vp.ev_cap && !sc->ev_window) {
            event_reduce_params(sc, vp.ev_cap, total, d_grad, E);
            HIP_TRY(launch_fill(reinterpret_cast<float *>(sc->d_evbins), 2, 0.0f, s), "views: clear");
        }
        vp.copies = grad_copies_for(k0, gn);
        if (vp.copies > 1) HIP_TRY(clear_grad_copies(k0, vp.copies, gn, s, &vp.gradx), "views: gradient copies");
    }
    std::vector<RenderParams> hP;
    std::vector<WorkParams> hW;
    for (uint32_t j = 0; j < nj; ++j) {
        const drk_view_job &J = jobs[j];
        AdjointArgs A;
        A.adj_l = J.d_lights;
        A.adj_e = J.d_env;
        A.grad = d_grad;
        vp.planned = false;
        int rc = render_impl(J.kernel, J.x0, J.y0, J.tilex, J.tiley, J.samples, J.seed, J.sample_base,
                             adjoint ? nullptr : J.d_lights, adjoint ? nullptr : J.d_env, nullptr, nullptr, s, false,
                             false, adjoint ? &A : nullptr, static_cast<float *>(J.d_record), false, false, &vp);
        if (rc != DRK_OK) return rc;
        if (vp.planned) {
            hP.push_back(vp.P);
            hW.push_back(vp.W);
        }
    }
    const uint32_t n = (uint32_t)hP.size();
    if (n == 0) return DRK_OK;
    const size_t need = (size_t)n * (sizeof(RenderParams) + sizeof(WorkParams));
    if (need > k0->jobs_bytes) {
        if (k0->d_jobs) (void)hipFree(k0->d_jobs);
        k0->d_jobs = nullptr;
        k0->jobs_bytes = 0;
        HIP_TRY(dev_malloc(&k0->d_jobs, need), "views: job table");
        k0->jobs_bytes = need;
    }
    RenderParams *dP = static_cast<RenderParams *>(k0->d_jobs);
    WorkParams *dW = reinterpret_cast<WorkParams *>(dP + n);
    uint32_t cam_pix = 0;
    for (uint32_t j = 0; j < n; ++j) {
        HIP_TRY(launch_job_setup(hP[j], hW[j], dP + j, dW + j, s), "views: job setup");
        if (hW[j].cam) cam_pix = std::max(cam_pix, hW[j].npix);
    }
    const uint32_t grid = k0->grid_override ? k0->grid_override
                                            : k0->cu_count * (uint32_t)std::max(1, render_batch_blocks_per_cu(mode));
    HIP_TRY(launch_render_batch(dP, dW, n, nullptr, 0, grid, mode, cam_pix, s), "views: launch");
    if (!adjoint) {
        for (uint32_t j = 0; j < n; ++j) HIP_TRY(launch_reduce(hW[j], hP[j].rx, hP[j].tilex, 2u, s), "views: reduce");
        return DRK_OK;
    }
    if (vp.copies > 1) HIP_TRY(launch_grad_sum(vp.gradx, vp.copies, gn, d_grad, s), "views: gradient sum");
    if (E.ev_cap) HIP_TRY(launch_event_reduce(E, s), "views: event reduction");
    if (sc->ev_window) sc->ev_window_samples += total;
    return DRK_OK;
}

int drk_views_render_record_device(const drk_view_job *jobs, uint32_t njobs, void *stream) {
    return views_impl(jobs, njobs, nullptr, stream, "drk_views_render_record_device");
}

int drk_views_adjoint_record_device(const drk_view_job *jobs, uint32_t njobs, float *d_grad_fog, void *stream) {
    if (!d_grad_fog) return set_error(DRK_EINVAL, "drk_views_adjoint_record_device: NULL gradient");
    return views_impl(jobs, njobs, d_grad_fog, stream, "drk_views_adjoint_record_device");
}

int drk_kernel_adjoint(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples,
                       uint64_t seed, uint32_t sample_base, const float *adj_lights, const float *adj_env,
                       float *grad_fog) {
    if (!k || !adj_lights || !adj_env || !grad_fog) return set_error(DRK_EINVAL, "drk_kernel_adjoint: NULL argument");
    if (k->scene->ev_window)
        return set_error(DRK_EINVAL, "drk_kernel_adjoint: the scene's gradient window is open (use the _device forms)");
    if (tilex == 0 || tiley == 0 || tilex > k->tilex || tiley > k->tiley)
        return set_error(DRK_EINVAL, "drk_kernel_adjoint: tile %ux%u exceeds the kernel's %ux%u", tilex, tiley,
                         k->tilex, k->tiley);
    const size_t tb = (size_t)tilex * tiley * 16;
    const size_t gn = (size_t)k->scene->leafx[1] * k->P.lvoxc3;
    float *d_adj = nullptr, *d_grad = nullptr;
    hipError_t e = dev_malloc(&d_adj, 2 * tb);
    if (e == hipSuccess) e = dev_malloc(&d_grad, gn * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(d_adj, adj_lights, tb, hipMemcpyHostToDevice, k->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(reinterpret_cast<char *>(d_adj) + tb, adj_env, tb, hipMemcpyHostToDevice, k->stream);
    if (e == hipSuccess) e = launch_fill(d_grad, gn, 0.0f, k->stream);
    int rc = DRK_OK;
    if (e == hipSuccess) {
        rc = drk_kernel_adjoint_device(k, x0, y0, tilex, tiley, samples, seed, sample_base, d_adj,
                                       reinterpret_cast<const float *>(reinterpret_cast<char *>(d_adj) + tb), d_grad,
                                       k->stream);
        if (rc == DRK_OK) e = hipMemcpyAsync(grad_fog, d_grad, gn * sizeof(float), hipMemcpyDeviceToHost, k->stream);
        if (rc == DRK_OK && e == hipSuccess) e = hipStreamSynchronize(k->stream);
        if (rc == DRK_OK && e == hipSuccess) rc = check_limit(k, k->stream, "drk_kernel_adjoint");
    }
    if (d_adj) (void)hipFree(d_adj);
    if (d_grad) (void)hipFree(d_grad);
    if (rc != DRK_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "drk_kernel_adjoint");
    return DRK_OK;
}

int drk_kernel_shadow(drk_kernel *k, uint32_t x0, uint32_t y0, uint32_t tilex, uint32_t tiley, uint32_t samples) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_shadow: NULL kernel");
    size_t tb = (size_t)k->tilex * k->tiley * 16;
    char *base = reinterpret_cast<char *>(k->d_out);
    int rc = render_impl(k, x0, y0, tilex, tiley, samples, k->seed, k->sample_base, (float *)base,
                         (float *)(base + tb), (float *)(base + 2 * tb), (float *)(base + 3 * tb), k->stream, false, true);
    if (rc != DRK_OK) return rc;
    HIP_TRY(hipMemcpyAsync(k->h_phb, k->d_out, 4 * tb, hipMemcpyDeviceToHost, k->stream), "drk_kernel_shadow: copy-out");
    HIP_TRY(hipStreamSynchronize(k->stream), "drk_kernel_shadow: sync");
    rc = check_limit(k, k->stream, "drk_kernel_shadow");
    if (rc != DRK_OK) return rc;
    k->sample_base += samples;
    return DRK_OK;
}

int drk_kernel_footprint(const drk_kernel *k, size_t *bytes) {
    if (!k || !bytes) return set_error(DRK_EINVAL, "drk_kernel_footprint: NULL argument");
    size_t b = (size_t)4 * k->tilex * k->tiley * 16 + k->partial_bytes + k->mail_bytes + k->gradx_bytes + sizeof(KernelConsts);
    b += kCounterWords * sizeof(uint32_t) + 64;
    if (k->d_env) b += (size_t)k->C.env_w * k->C.env_h * 16;
    if (k->d_depth) b += (size_t)k->P.w * k->P.h * sizeof(float);
    *bytes = b;
    return DRK_OK;
}

int drk_scene_footprint(const drk_scene *s, size_t *bytes) {
    if (!s || !bytes) return set_error(DRK_EINVAL, "drk_scene_footprint: NULL argument");
    *bytes = (size_t)s->node_count * (sizeof(drk_octree_node) + 8 * sizeof(uint32_t)) +
             ((size_t)s->leafx[0] + s->leafx[1]) * s->lvoxc * s->lvoxc * s->lvoxc * sizeof(float) +
             (s->d_bce ? (size_t)s->leafx[1] * 4 * sizeof(float) : 0);
    if (s->d_ev) *bytes += (size_t)s->ev_cap * kEventSlotBytes + kEvTableWords * sizeof(uint32_t);
    if (s->d_ref) *bytes += (size_t)s->leafx[1] * s->lvoxc * s->lvoxc * s->lvoxc * sizeof(float);
    if (s->d_cells) *bytes += ((size_t)1 << (3 * s->grid_levels)) * sizeof(uint32_t);
    return DRK_OK;
}

int drk_kernel_set_event_budget(drk_kernel *k, uint64_t bytes) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_set_event_budget: NULL kernel");
    drk_scene *sc = k->scene;
    if (sc->ev_window)
        return set_error(DRK_EINVAL, "drk_kernel_set_event_budget: the scene's gradient window is open");
    sc->ev_budget = bytes == UINT64_MAX ? SIZE_MAX : (size_t)bytes;
    if (sc->d_ev || sc->d_evbins) (void)hipDeviceSynchronize();   // a launch may still use them
    if (sc->d_ev) (void)hipFree(sc->d_ev);
    if (sc->d_evbins) (void)hipFree(sc->d_evbins);
    sc->d_ev = nullptr;
    sc->d_evbins = nullptr;
    sc->ev_cap = 0;
    return DRK_OK;
}

int drk_scene_events_begin(drk_scene *s, uint64_t samples, void *stream) {
    if (!s) return set_error(DRK_EINVAL, "drk_scene_events_begin: NULL scene");
    if (s->ev_window) return set_error(DRK_EINVAL, "drk_scene_events_begin: the scene's gradient window is already open");
    if (samples == 0) return set_error(DRK_EINVAL, "drk_scene_events_begin: samples must be > 0");
    const uint32_t cap = ensure_event_buffer(s, samples);
    if (cap > 0)
        HIP_TRY(launch_fill(reinterpret_cast<float *>(s->d_evbins), 2, 0.0f, (hipStream_t)stream),
                "drk_scene_events_begin: clear");
    s->ev_window = true;
    s->ev_window_cap = cap;
    s->ev_window_samples = 0;
    return DRK_OK;
}

int drk_scene_events_end(drk_scene *s, float *d_grad_fog, void *stream) {
    if (!s || !d_grad_fog) return set_error(DRK_EINVAL, "drk_scene_events_end: NULL argument");
    if (!s->ev_window) return set_error(DRK_EINVAL, "drk_scene_events_end: no gradient window is open");
    s->ev_window = false;
    if (s->ev_window_cap == 0 || s->ev_window_samples == 0) return DRK_OK;
    EventReduce E;
    event_reduce_params(s, s->ev_window_cap, s->ev_window_samples, d_grad_fog, E);
    HIP_TRY(launch_event_reduce(E, (hipStream_t)stream), "drk_scene_events_end: event reduction");
    return DRK_OK;
}

int drk_kernel_set_round_limit(drk_kernel *k, uint32_t rounds) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_set_round_limit: NULL kernel");
    k->round_limit = rounds;
    return DRK_OK;
}

int drk_kernel_limit_hit(const drk_kernel *k, int *hit) {
    if (!k || !hit) return set_error(DRK_EINVAL, "drk_kernel_limit_hit: NULL argument");
    uint32_t st = 0;
    HIP_TRY(hipMemcpy(&st, k->status + kStatusWord, sizeof(st), hipMemcpyDeviceToHost), "drk_kernel_limit_hit");
    *hit = st != 0;
    return DRK_OK;
}

int drk_kernel_set_timing(drk_kernel *k, int enable) {
    if (!k) return set_error(DRK_EINVAL, "drk_kernel_set_timing: NULL kernel");
    k->timing = enable != 0;
    return DRK_OK;
}

int drk_kernel_timing(drk_kernel *k, double *total_ms, uint32_t *launches) {
    if (!k || !total_ms || !launches) return set_error(DRK_EINVAL, "drk_kernel_timing: NULL argument");
    double t = 0.0;
    for (auto &pr : k->ev_used) {
        HIP_TRY(hipEventSynchronize(pr.second), "drk_kernel_timing: sync");
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second), "drk_kernel_timing: elapsed");
        t += ms;
    }
    *total_ms = t;
    *launches = (uint32_t)k->ev_used.size();
    for (auto &pr : k->ev_used) k->ev_free.push_back(pr);
    k->ev_used.clear();
    return DRK_OK;
}

const float *drk_kernel_phb(const drk_kernel *k, uint32_t buffer) {
    if (!k || buffer > 3) {
        set_error(DRK_EINVAL, "drk_kernel_phb: bad kernel or buffer index");
        return nullptr;
    }
    return k->h_phb + (size_t)buffer * k->tilex * k->tiley * 4;
}

void drk_kernel_destroy(drk_kernel *k) {
    if (!k) return;
    if (k->stream) (void)hipStreamSynchronize(k->stream);
    if (k->d_out) (void)hipFree(k->d_out);
    if (k->d_partial) (void)hipFree(k->d_partial);
    if (k->d_mail) (void)hipFree(k->d_mail);
    if (k->d_cam) (void)hipFree(k->d_cam);
    if (k->d_counter) (void)hipFree(k->d_counter);
    if (k->d_counts) (void)hipFree(k->d_counts);
    if (k->d_env) (void)hipFree(k->d_env);
    if (k->d_kc) (void)hipFree(k->d_kc);
    if (k->d_gradx) (void)hipFree(k->d_gradx);
    if (k->d_depth) (void)hipFree(k->d_depth);
    if (k->d_jobs) (void)hipFree(k->d_jobs);
    for (auto &pr : k->ev_used) k->ev_free.push_back(pr);
    for (auto &pr : k->ev_free) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    if (k->stream) (void)hipStreamDestroy(k->stream);
    std::free(k->h_phb);
    delete k;
}

// ---------------------------------------------------------------------------
// Batched launches: many (kernel, tile) jobs in one persistent grid.
// ---------------------------------------------------------------------------
static void batch_drop_graph(drk_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Fills the per-job launch parameters (tile, chunking, outputs), the partial
// buffer and the device tables for the current grid and unit scale.
static int batch_plan(drk_batch *b) {
    const uint32_t nj = (uint32_t)b->jobs.size();
    const int mode = b->sumsq ? kModeForwardSq : kModeForward;
    const uint32_t grid = b->grid ? b->grid : b->cu_count * (uint32_t)std::max(1, render_batch_blocks_per_cu(mode));
    b->grid_eff = grid;
    size_t poff = 0;
    std::vector<size_t> pbase(nj, 0);
    for (uint32_t j = 0; j < nj; ++j) {
        const drk_batch_job &J = b->jobs[j];
        const drk_kernel *k = J.kernel;
        RenderParams &P = b->hP[j];
        WorkParams &W = b->hW[j];
        P = k->P;
        set_walk_params(k->scene, P);
        uint32_t rx, ry;
        rendered_extent(k, J.x0, J.y0, J.tilex, J.tiley, rx, ry);
        P.x0 = J.x0;
        P.y0 = J.y0;
        P.rx = rx;
        P.ry = ry;
        P.tilex = J.tilex;
        P.samples = J.samples;
        P.sample_base = J.sample_base;
        P.key1 = rng_key1(J.seed);
        P.shadow_mode = 0;
        P.depth = k->depthcomp ? k->d_depth : nullptr;
        std::memset(&W, 0, sizeof(W));
        const uint32_t npix = rx * ry;
        uint32_t chunk = J.samples;
        if (npix) {
            if (k->chunk_override) {
                chunk = k->chunk_override;
            } else {
                // kBatchChunks * unit_scale sample chunks per pixel for every
                // job (units of equal sample counts, short enough that the
                // drain at the end of the launch stays short), fewer where
                // the job's partial sums would exceed kMaxPartialBytes.
                const uint64_t bytes_per_chunk = (uint64_t)npix * 16 * (b->sumsq ? 4 : 2);
                uint64_t nch = (uint64_t)std::llround(kBatchChunks * b->unit_scale);
                nch = std::min<uint64_t>(nch, kMaxPartialBytes / std::max<uint64_t>(bytes_per_chunk, 1));
                nch = std::max<uint64_t>(1, std::min<uint64_t>(nch, J.samples));
                chunk = (uint32_t)((J.samples + nch - 1) / nch);
            }
            chunk = std::max(1u, std::min(chunk, J.samples));
        }
        uint32_t nfull, nchunks;
        plan_chunks(J.samples, chunk, k->chunk_override == 0, nfull, nchunks);
        if ((uint64_t)npix * nchunks > 0xf0000000ull) return set_error(DRK_EINVAL, "batch: job %u has too many work units", j);
        W.counter = b->d_counters + (size_t)j * kCounterWords;
        W.round_limit = k->round_limit;
        W.nbands = k->bands;
        W.npix = npix;
        set_bands(W, rx);
        W.chunk = chunk;
        W.nfull = nfull;
        W.nchunks = nchunks;
        W.total_units = npix * nchunks;
        W.out_l = J.d_lights;
        W.out_e = J.d_env;
        W.sq_l = b->sumsq ? J.d_sumsq_lights : nullptr;
        W.sq_e = b->sumsq ? J.d_sumsq_env : nullptr;
        W.counts = k->d_counts;
        W.grad_copies = 1;
        W.grad_nxcc = k->xcc;
        if (nchunks > 1 && npix) {
            pbase[j] = poff;
            poff += (size_t)(b->sumsq ? 4 : 2) * nchunks * npix * 16;
        }
    }
    if (poff > b->partial_bytes) {
        if (b->d_partial) (void)hipFree(b->d_partial);
        b->d_partial = nullptr;
        b->partial_bytes = 0;
        HIP_TRY(dev_malloc(&b->d_partial, poff), "batch: partial sums");
        b->partial_bytes = poff;
    }
    for (uint32_t j = 0; j < nj; ++j)
        if (b->hW[j].nchunks > 1 && b->hW[j].npix)
            b->hW[j].partial = reinterpret_cast<float *>(reinterpret_cast<char *>(b->d_partial) + pbase[j]);
    // Camera rays, once per pixel of every job (a pass before the render).
    size_t coff = 0;
    b->cam_pix = 0;
    for (uint32_t j = 0; j < nj; ++j) {
        coff += (size_t)b->hW[j].npix * 16;
        b->cam_pix = std::max(b->cam_pix, b->hW[j].npix);
    }
    if (coff > b->cam_bytes) {
        if (b->d_cam) (void)hipFree(b->d_cam);
        b->d_cam = nullptr;
        b->cam_bytes = 0;
        HIP_TRY(dev_malloc(&b->d_cam, coff), "batch: camera rays");
        b->cam_bytes = coff;
    }
    coff = 0;
    for (uint32_t j = 0; j < nj; ++j) {
        b->hW[j].cam = b->hW[j].npix ? reinterpret_cast<float4 *>(static_cast<char *>(b->d_cam) + coff) : nullptr;
        coff += (size_t)b->hW[j].npix * 16;
    }
    HIP_TRY(hipMemcpy(b->dP, b->hP.data(), sizeof(RenderParams) * nj, hipMemcpyHostToDevice), "batch: job table");
    HIP_TRY(hipMemcpy(b->dW, b->hW.data(), sizeof(WorkParams) * nj, hipMemcpyHostToDevice), "batch: job table");
    b->planned = true;
    batch_drop_graph(b);
    return DRK_OK;
}

// The launch sequence: clears of the tile pixels outside each rendered
// extent, the batched render, then the per-job chunk reductions. Kernels
// only (no memset nodes), so it can be captured into a graph.
static int batch_enqueue(drk_batch *b, hipStream_t s) {
    const uint32_t nj = (uint32_t)b->jobs.size();
    const int mode = b->sumsq ? kModeForwardSq : kModeForward;
    for (uint32_t j = 0; j < nj; ++j) {
        const drk_batch_job &J = b->jobs[j];
        const RenderParams &P = b->hP[j];
        if (P.rx != J.tilex || P.ry != J.tiley) {
            const size_t n = (size_t)J.tilex * J.tiley * 4;
            HIP_TRY(launch_fill(J.d_lights, n, 0.0f, s), "batch: clear");
            HIP_TRY(launch_fill(J.d_env, n, 0.0f, s), "batch: clear");
            if (b->sumsq) {
                HIP_TRY(launch_fill(J.d_sumsq_lights, n, 0.0f, s), "batch: clear");
                HIP_TRY(launch_fill(J.d_sumsq_env, n, 0.0f, s), "batch: clear");
            }
        }
    }
    HIP_TRY(launch_render_batch(b->dP, b->dW, nj, b->d_counters, nj * kCounterWords, b->grid_eff, mode, b->cam_pix, s),
            "batch: launch");
    for (uint32_t j = 0; j < nj; ++j)
        HIP_TRY(launch_reduce(b->hW[j], b->hP[j].rx, b->hP[j].tilex, b->sumsq ? 4u : 2u, s), "batch: reduce");
    return DRK_OK;
}

int drk_batch_create(const drk_batch_job *jobs, uint32_t njobs, drk_batch **out) {
    if (!out) return set_error(DRK_EINVAL, "drk_batch_create: out is NULL");
    *out = nullptr;
    if (!jobs || njobs == 0) return set_error(DRK_EINVAL, "drk_batch_create: no jobs");
    bool sumsq = jobs[0].d_sumsq_lights != nullptr;
    for (uint32_t j = 0; j < njobs; ++j) {
        const drk_batch_job &J = jobs[j];
        if (!J.kernel || !J.d_lights || !J.d_env) return set_error(DRK_EINVAL, "drk_batch_create: job %u: NULL argument", j);
        if ((J.d_sumsq_lights != nullptr) != sumsq || (J.d_sumsq_env != nullptr) != sumsq)
            return set_error(DRK_EINVAL, "drk_batch_create: pass sum-of-squares buffers for every job or for none");
        if (J.tilex == 0 || J.tiley == 0 || J.tilex > J.kernel->tilex || J.tiley > J.kernel->tiley)
            return set_error(DRK_EINVAL, "drk_batch_create: job %u: tile %ux%u exceeds its kernel's %ux%u", j, J.tilex,
                             J.tiley, J.kernel->tilex, J.kernel->tiley);
        if (J.samples == 0) return set_error(DRK_EINVAL, "drk_batch_create: job %u: samples must be > 0", j);
        if ((uint64_t)J.sample_base + J.samples > 0xffffffffull || (uint64_t)J.x0 + J.tilex > 0xffffffffull ||
            (uint64_t)J.y0 + J.tiley > 0xffffffffull)
            return set_error(DRK_EINVAL, "drk_batch_create: job %u: counter or origin overflows", j);
    }
    drk_batch *b = new (std::nothrow) drk_batch;
    if (!b) return set_error(DRK_ENOMEM, "drk_batch_create: host allocation");
    b->jobs.assign(jobs, jobs + njobs);
    b->hP.resize(njobs);
    b->hW.resize(njobs);
    b->sumsq = sumsq;
    b->cu_count = jobs[0].kernel->cu_count;
    hipError_t e = dev_malloc(&b->dP, sizeof(RenderParams) * njobs);
    if (e == hipSuccess) e = dev_malloc(&b->dW, sizeof(WorkParams) * njobs);
    if (e == hipSuccess) e = dev_malloc(&b->d_counters, sizeof(uint32_t) * njobs * kCounterWords);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        drk_batch_destroy(b);
        return hip_fail(e, "drk_batch_create");
    }
    int rc = batch_plan(b);
    if (rc != DRK_OK) {
        drk_batch_destroy(b);
        return rc;
    }
    *out = b;
    return DRK_OK;
}

int drk_batch_set_tuning(drk_batch *b, uint32_t grid_blocks, float unit_scale) {
    if (!b) return set_error(DRK_EINVAL, "drk_batch_set_tuning: NULL batch");
    if (!(unit_scale > 0.0f)) return set_error(DRK_EINVAL, "drk_batch_set_tuning: unit_scale must be > 0");
    b->grid = grid_blocks;
    b->unit_scale = unit_scale;
    return batch_plan(b);
}

int drk_batch_launch(drk_batch *b, void *stream) {
    if (!b) return set_error(DRK_EINVAL, "drk_batch_launch: NULL batch");
    return batch_enqueue(b, (hipStream_t)stream);
}

int drk_batch_launch_graph(drk_batch *b, void *stream) {
    if (!b) return set_error(DRK_EINVAL, "drk_batch_launch_graph: NULL batch");
    hipStream_t s = (hipStream_t)stream;
    if (!b->exec) {
        HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal), "batch: capture");
        int rc = batch_enqueue(b, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (rc != DRK_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIP_TRY(e, "batch: end capture");
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "batch: instantiate");
    }
    HIP_TRY(hipGraphLaunch(b->exec, s), "batch: graph launch");
    return DRK_OK;
}

int drk_batch_autotune(drk_batch *b, void *stream, uint32_t reps, char *report, uint32_t report_len) {
    if (!b) return set_error(DRK_EINVAL, "drk_batch_autotune: NULL batch");
    hipStream_t s = (hipStream_t)stream;
    reps = std::max(1u, reps);
    const int mode = b->sumsq ? kModeForwardSq : kModeForward;
    const uint32_t base_grid = b->cu_count * (uint32_t)std::max(1, render_batch_blocks_per_cu(mode));
    const uint32_t grids[2] = {base_grid, base_grid / 2 > 0 ? base_grid / 2 : 1};
    const float scales[5] = {0.25f, 0.5f, 1.0f, 2.0f, 4.0f};
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "autotune: event");
    HIP_TRY(hipEventCreate(&e1), "autotune: event");
    std::string log = "grid unit_scale ms\n";
    float best = 1e30f;
    uint32_t best_grid = base_grid;
    float best_scale = 1.0f;
    int rc = DRK_OK;
    for (uint32_t g : grids) {
        for (float sc : scales) {
            b->grid = g;
            b->unit_scale = sc;
            if ((rc = batch_plan(b)) != DRK_OK) break;
            if ((rc = batch_enqueue(b, s)) != DRK_OK) break;   // warm-up
            (void)hipEventRecord(e0, s);
            for (uint32_t r = 0; r < reps && rc == DRK_OK; ++r) rc = batch_enqueue(b, s);
            (void)hipEventRecord(e1, s);
            if (rc != DRK_OK) break;
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= reps;
            char line[96];
            std::snprintf(line, sizeof(line), "%u %.2f %.3f\n", g, sc, ms);
            log += line;
            if (ms < best) {
                best = ms;
                best_grid = g;
                best_scale = sc;
            }
        }
        if (rc != DRK_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != DRK_OK) return rc;
    b->grid = best_grid;
    b->unit_scale = best_scale;
    char line[96];
    std::snprintf(line, sizeof(line), "chosen %u %.2f %.3f\n", best_grid, best_scale, best);
    log += line;
    if (report && report_len) std::snprintf(report, report_len, "%s", log.c_str());
    return batch_plan(b);
}

int drk_batch_limit_hit(const drk_batch *b, int *hit) {
    if (!b || !hit) return set_error(DRK_EINVAL, "drk_batch_limit_hit: NULL argument");
    const uint32_t nj = (uint32_t)b->jobs.size();
    std::vector<uint32_t> words((size_t)nj * kCounterWords);
    HIP_TRY(hipMemcpy(words.data(), b->d_counters, words.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
            "drk_batch_limit_hit");
    *hit = 0;
    for (uint32_t j = 0; j < nj; ++j) *hit |= words[(size_t)j * kCounterWords + kStatusWord] != 0;
    return DRK_OK;
}

void drk_batch_destroy(drk_batch *b) {
    if (!b) return;
    if (b->cap_stream) (void)hipStreamSynchronize(b->cap_stream);
    (void)hipDeviceSynchronize();
    batch_drop_graph(b);
    if (b->dP) (void)hipFree(b->dP);
    if (b->dW) (void)hipFree(b->dW);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_partial) (void)hipFree(b->d_partial);
    if (b->d_cam) (void)hipFree(b->d_cam);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
}

}  // extern "C"
