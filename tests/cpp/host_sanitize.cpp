// Host-side driver for the AddressSanitizer/UndefinedBehaviorSanitizer build
// of the library's host code (octree build, scene blob, synthetic clouds,
// argument validation of the C ABI). Runs without a GPU: every call here
// either does host work only or fails before its first device call.
// Exits 0 after printing "host_sanitize ok"; a sanitizer report aborts it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "droplet_render.h"

static int g_checks = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        ++g_checks;                                                             \
        if (!(cond)) {                                                          \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            std::exit(2);                                                       \
        }                                                                       \
    } while (0)

static std::vector<float> cloud(uint32_t kind, uint32_t n, uint64_t seed) {
    std::vector<float> g((size_t)n * n * n);
    if (kind == 0) {
        const float prm[5] = {3, 3.0f, 2.0f, 0.5f, 0.2f};
        CHECK(drk_synth_cloud(0, n, seed, prm, 5, g.data()) == DRK_OK);
    } else {
        const float prm[9] = {6, 0.25f, 0.75f, 0.08f, 0.2f, 0.1f, 3, 4.0f, 2.0f / n};
        CHECK(drk_synth_cloud(1, n, seed, prm, 9, g.data()) == DRK_OK);
    }
    return g;
}

static std::vector<char> read_file(const std::string &p) {
    std::vector<char> b;
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return b;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    b.resize((size_t)n);
    if (n > 0 && std::fread(b.data(), 1, (size_t)n, f) != (size_t)n) b.clear();
    std::fclose(f);
    return b;
}

static void write_file(const std::string &p, const std::vector<char> &b) {
    FILE *f = std::fopen(p.c_str(), "wb");
    CHECK(f != nullptr);
    if (!b.empty()) CHECK(std::fwrite(b.data(), 1, b.size(), f) == b.size());
    std::fclose(f);
}

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    std::mt19937 rng(12345);

    // Synthetic clouds: argument checks and both kinds.
    {
        float g[8];
        const float prm[5] = {3, 3.0f, 2.0f, 0.5f, 0.2f};
        CHECK(drk_synth_cloud(7, 2, 1, prm, 5, g) == DRK_EINVAL);
        CHECK(drk_synth_cloud(0, 2, 1, nullptr, 0, g) == DRK_EINVAL);
        CHECK(drk_synth_cloud(0, 2, 1, prm, 5, nullptr) == DRK_EINVAL);
    }

    // Octree builds: FOG, SDF and both, cubic and not, with depth caps that
    // double the brick size.
    std::vector<drk_host_scene *> scenes;
    for (int t = 0; t < 12; ++t) {
        const uint32_t n = 16 + 8 * (t % 4);
        std::vector<float> fog = cloud(t % 2, n, 100 + t);
        std::vector<float> sdf;
        const bool with_sdf = t % 3 == 2;
        if (with_sdf) {
            sdf.resize(fog.size());
            const float s = 1.0f / n;
            for (uint32_t z = 0; z < n; ++z)
                for (uint32_t y = 0; y < n; ++y)
                    for (uint32_t x = 0; x < n; ++x) {
                        float dx = x * s - 0.5f, dy = y * s - 0.5f, dz = z * s - 0.5f;
                        sdf[((size_t)z * n + y) * n + x] = std::sqrt(dx * dx + dy * dy + dz * dz) - 0.3f;
                    }
        }
        const uint32_t maxd = t < 6 ? 12u : 1u + (uint32_t)(t % 3);
        drk_host_scene *hs = nullptr;
        const float *f = (t % 5 == 4 && with_sdf) ? nullptr : fog.data();
        int rc = drk_scene_build_dense(f, with_sdf ? sdf.data() : nullptr, n, n, n, 1.0f / n, maxd, &hs);
        CHECK(rc == DRK_OK && hs && hs->node_count > 0);
        scenes.push_back(hs);
    }
    {
        // a non-cubic grid and a grid with one active voxel
        const uint32_t nx = 24, ny = 16, nz = 40;
        std::vector<float> g((size_t)nx * ny * nz, 0.0f);
        for (size_t i = 0; i < g.size(); i += 7) g[i] = (float)(i % 11) / 10.0f;
        drk_host_scene *hs = nullptr;
        CHECK(drk_scene_build_dense(g.data(), nullptr, nx, ny, nz, 0.02f, 12, &hs) == DRK_OK);
        scenes.push_back(hs);
        std::fill(g.begin(), g.end(), 0.0f);
        g[(size_t)5 * nx * ny + 3 * nx + 9] = 1.0f;
        hs = nullptr;
        CHECK(drk_scene_build_dense(g.data(), nullptr, nx, ny, nz, 0.02f, 12, &hs) == DRK_OK);
        scenes.push_back(hs);
        std::fill(g.begin(), g.end(), 0.0f);
        hs = nullptr;
        CHECK(drk_scene_build_dense(g.data(), nullptr, nx, ny, nz, 0.02f, 12, &hs) == DRK_EINVAL);
        CHECK(drk_scene_build_dense(nullptr, nullptr, nx, ny, nz, 0.02f, 12, &hs) == DRK_EINVAL);
        CHECK(drk_scene_build_dense(g.data(), nullptr, nx, ny, nz, -1.0f, 12, &hs) == DRK_EINVAL);
    }

    // Scene blobs: round trips, then damaged copies (flipped bytes,
    // truncations, extensions) must be rejected without reading out of bounds.
    const std::string path = dir + "/host_sanitize_scene.bin";
    const std::string bad = dir + "/host_sanitize_bad.bin";
    for (size_t si = 0; si < scenes.size(); si += 3) {
        drk_host_scene *hs = scenes[si];
        CHECK(drk_scene_save(hs, path.c_str()) == DRK_OK);
        drk_host_scene *back = nullptr;
        CHECK(drk_scene_load(path.c_str(), &back) == DRK_OK && back);
        CHECK(back->node_count == hs->node_count && back->lvoxc == hs->lvoxc);
        CHECK(std::memcmp(back->ob, hs->ob, sizeof(drk_octree_node) * hs->node_count) == 0);
        drk_host_scene_free(back);
        std::vector<char> blob = read_file(path);
        CHECK(blob.size() > 64);
        for (int trial = 0; trial < 40; ++trial) {
            std::vector<char> b = blob;
            const int kind = trial % 4;
            if (kind == 0) {
                b[rng() % b.size()] ^= (char)(1 + rng() % 255);
            } else if (kind == 1) {
                b.resize(rng() % b.size());
            } else if (kind == 2) {
                b.resize(b.size() + 1 + rng() % 64, (char)rng());
            } else {
                b[8 + rng() % 56] ^= (char)(1 + rng() % 255);   // header fields
            }
            write_file(bad, b);
            drk_host_scene *x = nullptr;
            const int rc = drk_scene_load(bad.c_str(), &x);
            // every damage is caught: header and payload are checksummed,
            // sizes are checked, trailing bytes are refused
            CHECK(rc == DRK_EINVAL && x == nullptr);
        }
    }
    {
        drk_host_scene *x = nullptr;
        CHECK(drk_scene_load((dir + "/does_not_exist.bin").c_str(), &x) != DRK_OK);
        CHECK(drk_scene_save(nullptr, path.c_str()) == DRK_EINVAL);
    }

    // drk_scene_create's structural validation runs before any device call:
    // corrupted node arrays must be rejected (or pass to the device step,
    // which fails here without a GPU) without reading out of bounds.
    for (size_t si = 0; si < scenes.size(); ++si) {
        drk_host_scene *hs = scenes[si];
        std::vector<drk_octree_node> ob(hs->ob, hs->ob + hs->node_count);
        const size_t l3 = (size_t)hs->lvoxc * hs->lvoxc * hs->lvoxc;
        for (int trial = 0; trial < 20; ++trial) {
            std::vector<drk_octree_node> c = ob;
            drk_octree_node &n = c[rng() % c.size()];
            switch (trial % 5) {
            case 0: n.chn[rng() % 8] = (uint32_t)rng(); break;
            case 1: n.volx[rng() % 2] = (uint32_t)rng() % (hs->leafx[1] + 3); break;
            case 2: n.ce[3] = -std::fabs(n.ce[3]); break;
            case 3: n.chn[rng() % 8] = (uint32_t)(rng() % c.size()); break;
            default: n.ce[3] = std::nanf(""); break;
            }
            drk_scene *s = nullptr;
            int rc = drk_scene_create(c.data(), (uint32_t)c.size(), hs->pvol[0], hs->leafx[0], hs->pvol[1],
                                      hs->leafx[1], hs->lvoxc, &s);
            if (rc == DRK_OK) drk_scene_destroy(s);
            CHECK(rc == DRK_OK || rc == DRK_EINVAL || rc == DRK_EDEVICE || rc == DRK_ENOMEM);
            (void)l3;
        }
        drk_scene *s = nullptr;
        CHECK(drk_scene_create(ob.data(), 0, nullptr, 0, nullptr, 0, hs->lvoxc, &s) == DRK_EINVAL);
        CHECK(drk_scene_create(ob.data(), (uint32_t)ob.size(), nullptr, 0, nullptr, 0, 1, &s) == DRK_EINVAL);
    }

    // Kernel entry points with bad arguments.
    {
        float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        drk_kernel *k = nullptr;
        drk_sampler_desc d;
        std::memset(&d, 0, sizeof(d));
        CHECK(drk_kernel_create(nullptr, m, m, &d, nullptr, 4, 1.0f, 0.0f, 8, 8, 8, 8, 0, &k) == DRK_EINVAL);
        CHECK(drk_kernel_render(nullptr, 0, 0, 8, 8, 1) == DRK_EINVAL);
        CHECK(drk_kernel_shadow(nullptr, 0, 0, 8, 8, 1) == DRK_EINVAL);
        CHECK(drk_kernel_set_round_limit(nullptr, 1) == DRK_EINVAL);
        CHECK(drk_kernel_phb(nullptr, 0) == nullptr);
        CHECK(drk_batch_create(nullptr, 0, nullptr) == DRK_EINVAL);
        CHECK(std::strlen(drk_last_error()) > 0);
    }

    for (drk_host_scene *hs : scenes) drk_host_scene_free(hs);
    std::remove(path.c_str());
    std::remove(bad.c_str());
    std::printf("host_sanitize ok: %d checks\n", g_checks);
    return 0;
}
