// On-disk form of a built scene: the flattened node records and bricks
// (Scene::ob and Scene::pvol, src/scene.h:146-159) in one file, so a render
// can start from the octree the build produced without re-running it.
//
// Layout (little endian):
//   magic "DRKSCN\1\0" | u32 version (1) | u32 node_count | u32 leafx[2] |
//   u32 lvoxc | u32 mlevel | f32 root_ce[4] | u64 payload bytes |
//   u64 FNV-1a 64 of the header's first 56 bytes and then the payload |
//   payload: node_count 64-byte records, then leafx[0] SDF bricks and
//   leafx[1] FOG bricks of lvoxc^3 floats each.
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "drk_common.h"
#include "drk_internal.h"

namespace drk {

static const char kMagic[8] = {'D', 'R', 'K', 'S', 'C', 'N', 1, 0};
constexpr uint32_t kBlobVersion = 2;   // 2: the checksum covers the header fields too

struct BlobHeader {
    char magic[8];
    uint32_t version, node_count, leafx[2], lvoxc, mlevel;
    float root_ce[4];
    uint64_t payload, checksum;
};
static_assert(sizeof(BlobHeader) == 64, "blob header layout");

static uint64_t fnv1a(const void *p, size_t n, uint64_t h) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;

}  // namespace drk

using namespace drk;

extern "C" {

int drk_scene_save(const drk_host_scene *hs, const char *path) {
    if (!hs || !path || !hs->ob || hs->node_count == 0) return set_error(DRK_EINVAL, "drk_scene_save: invalid argument");
    const size_t l3 = (size_t)hs->lvoxc * hs->lvoxc * hs->lvoxc;
    const size_t nb = sizeof(drk_octree_node) * hs->node_count;
    const size_t sb[2] = {(size_t)hs->leafx[0] * l3 * sizeof(float), (size_t)hs->leafx[1] * l3 * sizeof(float)};
    for (int b = 0; b < 2; ++b)
        if (sb[b] && !hs->pvol[b]) return set_error(DRK_EINVAL, "drk_scene_save: missing bricks");
    BlobHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = kBlobVersion;
    h.node_count = hs->node_count;
    h.leafx[0] = hs->leafx[0];
    h.leafx[1] = hs->leafx[1];
    h.lvoxc = hs->lvoxc;
    h.mlevel = hs->mlevel;
    std::memcpy(h.root_ce, hs->root_ce, sizeof(h.root_ce));
    h.payload = nb + sb[0] + sb[1];
    uint64_t c = fnv1a(&h, offsetof(BlobHeader, checksum), kFnvBasis);
    c = fnv1a(hs->ob, nb, c);
    for (int b = 0; b < 2; ++b)
        if (sb[b]) c = fnv1a(hs->pvol[b], sb[b], c);
    h.checksum = c;
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(DRK_EINVAL, "drk_scene_save: cannot open %s", path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(hs->ob, 1, nb, f) == nb;
    for (int b = 0; b < 2 && ok; ++b)
        if (sb[b]) ok = std::fwrite(hs->pvol[b], 1, sb[b], f) == sb[b];
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(DRK_EINVAL, "drk_scene_save: write to %s failed", path);
    return DRK_OK;
}

int drk_scene_load(const char *path, drk_host_scene **out) {
    if (!out || !path) return set_error(DRK_EINVAL, "drk_scene_load: invalid argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_error(DRK_EINVAL, "drk_scene_load: cannot open %s", path);
    BlobHeader h;
    if (std::fread(&h, sizeof(h), 1, f) != 1) {
        std::fclose(f);
        return set_error(DRK_EINVAL, "drk_scene_load: %s: truncated header", path);
    }
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kBlobVersion) {
        std::fclose(f);
        return set_error(DRK_EINVAL, "drk_scene_load: %s: not a scene blob (magic/version)", path);
    }
    if (h.node_count == 0 || h.node_count > kMaxNodes || h.lvoxc < 2 || h.lvoxc > 256) {
        std::fclose(f);
        return set_error(DRK_EINVAL, "drk_scene_load: %s: bad header (nodes %u, lvoxc %u)", path, h.node_count, h.lvoxc);
    }
    const size_t l3 = (size_t)h.lvoxc * h.lvoxc * h.lvoxc;
    const size_t nb = sizeof(drk_octree_node) * h.node_count;
    const size_t sb[2] = {(size_t)h.leafx[0] * l3 * sizeof(float), (size_t)h.leafx[1] * l3 * sizeof(float)};
    if (h.payload != nb + sb[0] + sb[1]) {
        std::fclose(f);
        return set_error(DRK_EINVAL, "drk_scene_load: %s: payload size does not match the header", path);
    }
    drk_host_scene *hs = (drk_host_scene *)std::calloc(1, sizeof(drk_host_scene));
    if (!hs) {
        std::fclose(f);
        return set_error(DRK_ENOMEM, "drk_scene_load: host allocation");
    }
    hs->node_count = h.node_count;
    hs->leafx[0] = h.leafx[0];
    hs->leafx[1] = h.leafx[1];
    hs->lvoxc = h.lvoxc;
    hs->mlevel = h.mlevel;
    std::memcpy(hs->root_ce, h.root_ce, sizeof(h.root_ce));
    hs->ob = (drk_octree_node *)std::malloc(nb);
    for (int b = 0; b < 2; ++b) hs->pvol[b] = sb[b] ? (float *)std::malloc(sb[b]) : nullptr;
    bool ok = hs->ob && (!sb[0] || hs->pvol[0]) && (!sb[1] || hs->pvol[1]);
    if (!ok) {
        std::fclose(f);
        host_scene_free(hs);
        return set_error(DRK_ENOMEM, "drk_scene_load: host allocation");
    }
    ok = std::fread(hs->ob, 1, nb, f) == nb;
    for (int b = 0; b < 2 && ok; ++b)
        if (sb[b]) ok = std::fread(hs->pvol[b], 1, sb[b], f) == sb[b];
    const bool trailing = ok && std::fgetc(f) != EOF;
    std::fclose(f);
    if (!ok || trailing) {
        host_scene_free(hs);
        return set_error(DRK_EINVAL, "drk_scene_load: %s: %s", path, ok ? "trailing bytes" : "truncated payload");
    }
    uint64_t c = fnv1a(&h, offsetof(BlobHeader, checksum), kFnvBasis);
    c = fnv1a(hs->ob, nb, c);
    for (int b = 0; b < 2; ++b)
        if (sb[b]) c = fnv1a(hs->pvol[b], sb[b], c);
    if (c != h.checksum) {
        host_scene_free(hs);
        return set_error(DRK_EINVAL, "drk_scene_load: %s: checksum mismatch", path);
    }
    // Child links and brick indices are checked again by drk_scene_create;
    // reject what would index outside the arrays here already.
    for (uint32_t i = 0; i < hs->node_count; ++i) {
        const drk_octree_node &n = hs->ob[i];
        bool bad = false;
        for (int k = 0; k < 8; ++k) bad |= n.chn[k] >= hs->node_count || (n.chn[k] != 0 && n.chn[k] <= i);
        for (int b = 0; b < 2; ++b) bad |= n.volx[b] != DRK_NO_BRICK && n.volx[b] >= hs->leafx[b];
        if (bad) {
            host_scene_free(hs);
            return set_error(DRK_EINVAL, "drk_scene_load: %s: node %u indexes out of range", path, i);
        }
    }
    *out = hs;
    return DRK_OK;
}

}  // extern "C"
