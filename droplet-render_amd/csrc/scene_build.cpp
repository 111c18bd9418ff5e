// Host-side scene ingest: dense grids -> sparse octree (Scene::ob) and
// vertex-centred density bricks (Scene::pvol) with per-leaf majorants.
//
// Restates, without OpenVDB:
//   * leaf-AABB collection and root sizing   src/scene.cpp:676-716
//   * Octree::BuildPath                      src/scene.cpp:267-305 (+ Intersects :234-247)
//   * brick resampling and majorants         src/scene.cpp:840-917
// The float operation order of the reference is kept where it decides node
// geometry (centres, extents, overlap tests) so that node indices and ce are
// bit-identical for the same leaf AABB list.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "drk_common.h"
#include "drk_internal.h"

namespace drk {
namespace {

#pragma clang fp contract(off)

struct F4 { float x, y, z, w; };
static inline F4 f4(float x, float y, float z, float w) { return {x, y, z, w}; }
static inline F4 operator+(F4 a, F4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
static inline F4 operator-(F4 a, F4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
static inline F4 operator*(F4 a, F4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
static inline F4 operator*(float s, F4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
static inline F4 operator/(F4 a, F4 b) { return {a.x / b.x, a.y / b.y, a.z / b.z, a.w / b.w}; }
static inline F4 splat(float s) { return {s, s, s, s}; }

struct Aabb { float sc[3], se[3]; };

// One OpenVDB-like 8^3 leaf: its tree-order key and active bounding box.
struct VdbLeaf {
    uint64_t key;
    int mn[3], mx[3];
};

// Tree order of an OpenVDB FloatGrid (Tree4<float,5,4,3>): root children by
// coordinate, then the 32^3 internal node's child mask (x-major), then the
// 16^3 internal node's child mask, giving the leaf iteration order of
// cbeginLeaf() used at src/scene.cpp:677. Valid for coordinates in [0,2^21).
static uint64_t vdb_leaf_key(int x, int y, int z) {
    uint64_t k = 0;
    k = (k << 9) | (uint64_t)(x >> 12);
    k = (k << 9) | (uint64_t)(y >> 12);
    k = (k << 9) | (uint64_t)(z >> 12);
    k = (k << 5) | (uint64_t)((x >> 7) & 31);
    k = (k << 5) | (uint64_t)((y >> 7) & 31);
    k = (k << 5) | (uint64_t)((z >> 7) & 31);
    k = (k << 4) | (uint64_t)((x >> 3) & 15);
    k = (k << 4) | (uint64_t)((y >> 3) & 15);
    k = (k << 4) | (uint64_t)((z >> 3) & 15);
    return k;
}

// Active voxels: non-zero density for FOG; the narrow band |d| < band for a
// level set (OpenVDB keeps only those voxels active, src/scene.cpp:677-681).
static void collect_leaves(const float *g, uint32_t nx, uint32_t ny, uint32_t nz, float band,
                           std::vector<VdbLeaf> &out) {
    const uint32_t bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (nz + 7) / 8;
    for (uint32_t kz = 0; kz < bz; ++kz)
        for (uint32_t ky = 0; ky < by; ++ky)
            for (uint32_t kx = 0; kx < bx; ++kx) {
                VdbLeaf L;
                L.mn[0] = L.mn[1] = L.mn[2] = 1 << 30;
                L.mx[0] = L.mx[1] = L.mx[2] = -1;
                bool any = false;
                for (uint32_t z = kz * 8; z < std::min(nz, kz * 8 + 8); ++z)
                    for (uint32_t y = ky * 8; y < std::min(ny, ky * 8 + 8); ++y)
                        for (uint32_t x = kx * 8; x < std::min(nx, kx * 8 + 8); ++x) {
                            const float v = g[((size_t)z * ny + y) * nx + x];
                            if (band > 0.0f ? !(std::fabs(v) < band) : v == 0.0f) continue;
                            any = true;
                            int c[3] = {(int)x, (int)y, (int)z};
                            for (int a = 0; a < 3; ++a) {
                                L.mn[a] = std::min(L.mn[a], c[a]);
                                L.mx[a] = std::max(L.mx[a], c[a]);
                            }
                        }
                if (!any) continue;
                L.key = vdb_leaf_key((int)kx * 8, (int)ky * 8, (int)kz * 8);
                out.push_back(L);
            }
    std::sort(out.begin(), out.end(), [](const VdbLeaf &a, const VdbLeaf &b) { return a.key < b.key; });
}

// World AABB of a leaf's active voxels, src/scene.cpp:680-692:
// centre = indexToWorld(bbox.getCenter()) in double, extent = 0.5f*dim*s in float.
static Aabb leaf_aabb(const VdbLeaf &L, float s) {
    Aabb b;
    for (int a = 0; a < 3; ++a) {
        float dimi = (float)(L.mx[a] - L.mn[a] + 1);
        b.se[a] = 0.5f * dimi * s;
        double posi = ((double)L.mn[a] + (double)L.mx[a]) * 0.5;
        b.sc[a] = (float)(posi * (double)s);
    }
    return b;
}

// BoundingBox::Intersects(const BoundingBox&), src/scene.cpp:234-247.
static bool aabb_overlap(const float ca[3], const float ea[3], const float cb[3], const float eb[3]) {
    for (int a = 0; a < 3; ++a) {
        float mina = ca[a] - ea[a], maxa = ca[a] + ea[a];
        float minb = cb[a] - eb[a], maxb = cb[a] + eb[a];
        if (mina > maxb || minb > maxa) return false;
    }
    return true;
}

struct Builder {
    std::vector<drk_octree_node> ob;
    std::vector<uint32_t> scratch;   // unused; kept for symmetry with the node pool
    uint32_t index = 0;              // last assigned node index (root is 0)
    uint32_t leafx[2] = {0, 0};
    uint32_t mlevel = 0;

    uint32_t new_node() {
        uint32_t i = ++index;
        if (ob.size() < (size_t)i + 1) {
            drk_octree_node n;
            std::memset(&n, 0, sizeof(n));
            n.volx[0] = n.volx[1] = DRK_NO_BRICK;
            ob.resize(i + 1, n);
        }
        return i;
    }

    // Octree::BuildPath, src/scene.cpp:267-305. c/e: node centre and
    // (splatted) half-extent; c1/e1: the leaf AABB being inserted.
    void build_path(uint32_t x, F4 c, F4 e, const float c1[3], const float e1[3], uint32_t level, int bx) {
        drk_octree_node &n = ob[x];
        n.ce[0] = c.x; n.ce[1] = c.y; n.ce[2] = c.z; n.ce[3] = e.w;
        n.qval[0] = n.qval[1] = 0.0f;
        if (level >= mlevel - 1) {
            if (ob[x].volx[bx] == DRK_NO_BRICK) ob[x].volx[bx] = leafx[bx]++;
            return;
        }
        F4 ee = 0.5f * e;
        for (uint32_t i = 0; i < 8; ++i) {
            F4 sv = 2.0f * f4((float)(i % 2), (float)((i / 2) % 2), (float)(i / 4), 0.0f) - splat(1.0f);
            F4 cc = c + sv * ee;
            F4 ex = 1.1f * ee;
            float ccv[3] = {cc.x, cc.y, cc.z}, exv[3] = {ex.x, ex.y, ex.z};
            if (!aabb_overlap(ccv, exv, c1, e1)) continue;
            uint32_t ch = ob[x].chn[i];
            if (ch == 0) {
                ch = new_node();
                ob[x].chn[i] = ch;
            }
            build_path(ch, cc, ee, c1, e1, level + 1, bx);
        }
    }
};

// Trilinear sample of a dense grid, restating OpenVDB's BoxSampler (index =
// world/s in double, floor, interpolate z then y then x with float
// differences and double weights). Outside the grid the value is the
// background: 0 for FOG, +band for a level set, whose values are also held
// to [-band, band] (OpenVDB stores +-background beyond the narrow band).
// OpenVDB itself is not available; see the oracle for the independent
// restatement.
static inline float grid_at(const float *g, uint32_t nx, uint32_t ny, uint32_t nz, float band, long x, long y,
                            long z) {
    if (x < 0 || y < 0 || z < 0 || x >= (long)nx || y >= (long)ny || z >= (long)nz) return band;
    const float v = g[((size_t)z * ny + (size_t)y) * nx + (size_t)x];
    return band > 0.0f ? std::max(std::min(v, band), -band) : v;
}
static inline float lerp_vdb(float a, float b, double w) { return a + (float)((double)(b - a) * w); }

static float box_sample(const float *g, uint32_t nx, uint32_t ny, uint32_t nz, float band, double inv_s,
                        const float p[3]) {
    double q[3];
    long i0[3];
    double uvw[3];
    for (int a = 0; a < 3; ++a) {
        q[a] = (double)p[a] * inv_s;
        i0[a] = (long)std::floor(q[a]);
        uvw[a] = q[a] - (double)i0[a];
    }
    float d[2][2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) d[i][j][k] = grid_at(g, nx, ny, nz, band, i0[0] + i, i0[1] + j, i0[2] + k);
    return lerp_vdb(lerp_vdb(lerp_vdb(d[0][0][0], d[0][0][1], uvw[2]), lerp_vdb(d[0][1][0], d[0][1][1], uvw[2]), uvw[1]),
                    lerp_vdb(lerp_vdb(d[1][0][0], d[1][0][1], uvw[2]), lerp_vdb(d[1][1][0], d[1][1][1], uvw[2]), uvw[1]),
                    uvw[0]);
}

}  // namespace

int scene_build_dense(const float *fog, const float *sdf, uint32_t nx, uint32_t ny, uint32_t nz,
                      float s, uint32_t maxd, drk_host_scene **out) {
    if ((!fog && !sdf) || !out || nx == 0 || ny == 0 || nz == 0 || !(s > 0.0f) || maxd == 0 || maxd > kMaxDepth)
        return set_error(DRK_EINVAL, "drk_scene_build_dense: invalid argument");
    *out = nullptr;
    // Narrow band of the level set: bvc = 4 voxels (src/scene.cpp:844).
    const float band = 4.0f * s;

    // Leaf AABBs in buffer order SDF then FOG (src/scene.cpp:676-701).
    std::vector<Aabb> boxes[2];
    if (sdf) {
        std::vector<VdbLeaf> leaves;
        collect_leaves(sdf, nx, ny, nz, band, leaves);
        for (const VdbLeaf &L : leaves) boxes[DRK_VOLUME_SDF].push_back(leaf_aabb(L, s));
    }
    if (fog) {
        std::vector<VdbLeaf> leaves;
        collect_leaves(fog, nx, ny, nz, 0.0f, leaves);
        for (const VdbLeaf &L : leaves) boxes[DRK_VOLUME_FOG].push_back(leaf_aabb(L, s));
    }
    if (boxes[0].empty() && boxes[1].empty()) return set_error(DRK_EINVAL, "drk_scene_build_dense: grid has no active voxels");

    F4 mn = splat(FLT_MAX), mx = splat(-FLT_MAX);
    for (int b = 0; b < 2; ++b)
        for (const Aabb &a : boxes[b]) {
            F4 c = f4(a.sc[0], a.sc[1], a.sc[2], 0.0f), e = f4(a.se[0], a.se[1], a.se[2], 0.0f);
            F4 lo = c - e, hi = c + e;
            mn = f4(std::min(lo.x, mn.x), std::min(lo.y, mn.y), std::min(lo.z, mn.z), 0.0f);
            mx = f4(std::max(hi.x, mx.x), std::max(hi.y, mx.y), std::max(hi.z, mx.z), 0.0f);
        }
    // Root cube and depth, src/scene.cpp:703-716 (SM_LN2 = 0.69315f).
    F4 c = 0.5f * (mx + mn);
    F4 e = 0.5f * (mx - mn);
    float a = std::max(std::max(e.x, e.y), e.z);
    float lvc = 8.0f;
    float d = 2.0f * a;
    // A scene smaller than one leaf gives a depth <= 0 (the reference's
    // mlevel-1 then underflows); it is held at 1. The depth is computed as a
    // signed integer: a negative float converted to uint32_t is undefined.
    int32_t ml = (int32_t)ceilf(logf(d / (lvc * s)) / 0.69315f);
    if (ml > (int32_t)maxd) {
        lvc *= powf(2.0f, (float)(ml - (int32_t)maxd));
        ml = (int32_t)maxd;
    }
    const uint32_t mlevel = ml < 1 ? 1u : (uint32_t)ml;

    Builder B;
    B.mlevel = mlevel;
    try {
        B.ob.reserve(1024);
        drk_octree_node root;
        std::memset(&root, 0, sizeof(root));
        root.volx[0] = root.volx[1] = DRK_NO_BRICK;
        B.ob.push_back(root);
        for (int b = 0; b < 2; ++b)
            for (const Aabb &bx : boxes[b]) B.build_path(0, f4(c.x, c.y, c.z, a), splat(a), bx.sc, bx.se, 0, b);
    } catch (const std::bad_alloc &) {
        return set_error(DRK_ENOMEM, "drk_scene_build_dense: out of host memory");
    }

    const uint32_t lvoxc = (uint32_t)lvc;
    const uint32_t lvoxc3 = lvoxc * lvoxc * lvoxc;
    drk_host_scene *hs = (drk_host_scene *)std::calloc(1, sizeof(drk_host_scene));
    if (!hs) return set_error(DRK_ENOMEM, "drk_scene_build_dense: out of host memory");
    hs->node_count = (uint32_t)B.ob.size();
    hs->ob = (drk_octree_node *)std::malloc(sizeof(drk_octree_node) * hs->node_count);
    hs->leafx[0] = B.leafx[0];
    hs->leafx[1] = B.leafx[1];
    hs->lvoxc = lvoxc;
    hs->mlevel = mlevel;
    hs->root_ce[0] = c.x; hs->root_ce[1] = c.y; hs->root_ce[2] = c.z; hs->root_ce[3] = a;
    for (int b = 0; b < 2; ++b)
        hs->pvol[b] = B.leafx[b] ? (float *)std::calloc((size_t)B.leafx[b] * lvoxc3, sizeof(float)) : nullptr;
    if (!hs->ob || (B.leafx[1] && !hs->pvol[1]) || (B.leafx[0] && !hs->pvol[0])) {
        host_scene_free(hs);
        return set_error(DRK_ENOMEM, "drk_scene_build_dense: out of host memory");
    }
    std::memcpy(hs->ob, B.ob.data(), sizeof(drk_octree_node) * hs->node_count);
    // The node the resampling skips (the last one, below) keeps the bricks it
    // was given here: FOG 0 (its majorant 0 makes it transparent anyway) and
    // SDF +band, outside any surface. The reference leaves both uninitialised.
    if (hs->pvol[DRK_VOLUME_SDF])
        for (size_t j = 0; j < (size_t)B.leafx[DRK_VOLUME_SDF] * lvoxc3; ++j) hs->pvol[DRK_VOLUME_SDF][j] = band;

    // Resampling, src/scene.cpp:870-908. The reference iterates node indices
    // [0, index) where index is the LAST assigned node index, so the final
    // node (always a leaf) keeps qval = 0 and an unfilled brick; that leaf is
    // then skipped by the tracker (infinite free-flight). Reproduced here,
    // with the brick left at zero instead of uninitialised memory.
    const uint32_t nres = B.index;  // nodes [0, index)
    const double inv_s = 1.0 / (double)s;
    const float nv = (float)lvoxc;
    auto work = [&](uint32_t i0, uint32_t i1) {
        for (uint32_t i = i0; i < i1; ++i) {
            drk_octree_node &n = hs->ob[i];
            if (n.volx[DRK_VOLUME_SDF] == DRK_NO_BRICK) {
                if (n.volx[DRK_VOLUME_FOG] == DRK_NO_BRICK) continue;   // not a leaf
                // A fog leaf whose centre lies inside the surface (and that no
                // SDF leaf overlaps) is dropped (src/scene.cpp:874-883).
                // Without an SDF grid the distance is the positive background.
                if (sdf) {
                    float pc[3] = {n.ce[0], n.ce[1], n.ce[2]};
                    if (box_sample(sdf, nx, ny, nz, band, inv_s, pc) < 0.0f) {
                        n.volx[DRK_VOLUME_FOG] = DRK_NO_BRICK;
                        continue;
                    }
                }
            }
            F4 nc = f4(n.ce[0], n.ce[1], n.ce[2], n.ce[3]);
            F4 ne = splat(n.ce[3]);
            n.qval[DRK_VOLUME_SDF] = FLT_MAX;
            n.qval[DRK_VOLUME_FOG] = 0.01f;
            float *sbrick = n.volx[DRK_VOLUME_SDF] == DRK_NO_BRICK
                                ? nullptr
                                : hs->pvol[DRK_VOLUME_SDF] + (size_t)n.volx[DRK_VOLUME_SDF] * lvoxc3;
            float *fbrick = n.volx[DRK_VOLUME_FOG] == DRK_NO_BRICK
                                ? nullptr
                                : hs->pvol[DRK_VOLUME_FOG] + (size_t)n.volx[DRK_VOLUME_FOG] * lvoxc3;
            for (uint32_t j = 0; j < lvoxc3; ++j) {
                F4 idx = f4((float)(j % lvoxc), (float)((j / lvoxc) % lvoxc), (float)(j / (lvoxc * lvoxc)), 1.0f);
                F4 nn = (splat(nv) - 2.0f * idx - splat(1.0f)) / (splat(nv) - splat(1.0f));
                F4 nw = nc - ne * nn;
                float p[3] = {nw.x, nw.y, nw.z};
                if (sbrick) {
                    float v = box_sample(sdf, nx, ny, nz, band, inv_s, p);
                    sbrick[j] = v;
                    n.qval[DRK_VOLUME_SDF] = std::min(n.qval[DRK_VOLUME_SDF], v);
                }
                if (fbrick) {
                    float v = box_sample(fog, nx, ny, nz, 0.0f, inv_s, p);
                    v = std::max(std::min(v, 1.0f), 0.0f);
                    fbrick[j] = v;
                    n.qval[DRK_VOLUME_FOG] = std::max(n.qval[DRK_VOLUME_FOG], v);
                }
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nres < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        uint32_t i0 = (uint32_t)((uint64_t)nres * t / nt), i1 = (uint32_t)((uint64_t)nres * (t + 1) / nt);
        th.emplace_back(work, i0, i1);
    }
    for (auto &t : th) t.join();

    *out = hs;
    return DRK_OK;
}

void host_scene_free(drk_host_scene *hs) {
    if (!hs) return;
    std::free(hs->ob);
    std::free(hs->pvol[0]);
    std::free(hs->pvol[1]);
    std::free(hs);
}

}  // namespace drk
