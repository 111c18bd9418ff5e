// Renders one tile through the C++ RenderKernel facade (the reference's entry
// points) and writes phb[0], phb[1] to a file, so the GPU test can compare it
// with the same render through the Python mirror.
// usage: facade_render <in.bin> <out.bin>
//   in.bin: u32 n, w, h, tilex, tiley, x0, y0, spp, scattevs; f32 sigma_s, sigma_a,
//           viewi[16], proji[16], sun_dir[3], turbidity, albedo; u64 seed; f32 grid[n^3]
#include <cstdio>
#include <vector>

#include "RenderKernel.hpp"

using namespace droplet;

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    uint32_t u[9];
    float sig[2], vi[16], pj[16], sd[3], sky[2];
    uint64_t seed;
    if (std::fread(u, 4, 9, f) != 9 || std::fread(sig, 4, 2, f) != 2 || std::fread(vi, 4, 16, f) != 16 ||
        std::fread(pj, 4, 16, f) != 16 || std::fread(sd, 4, 3, f) != 3 || std::fread(sky, 4, 2, f) != 2 ||
        std::fread(&seed, 8, 1, f) != 1)
        return 4;
    const uint32_t n = u[0];
    std::vector<float> grid((size_t)n * n * n);
    if (std::fread(grid.data(), 4, grid.size(), f) != grid.size()) return 5;
    std::fclose(f);

    Scene scene;
    if (!scene.Initialize(grid.data(), n, n, n, 1.0f / n, 12)) return 6;
    dmatrix44 viewi, proji;
    std::memcpy(&viewi, vi, 64);
    std::memcpy(&proji, pj, 64);
    dfloat3 dir = {sd[0], sd[1], sd[2]}, col = {1.0f, 1.0f, 1.0f};
    new KernelSampler::SunLight(&dir, &col, 0.010f);      // registers itself in BaseLight::lights
    KernelSampler::HosekSky env(sky[0], sky[1]);
    RenderKernel k;
    if (!k.Initialize(&scene, nullptr, &viewi, &proji, &KernelSampler::MiePhase::gmie, &env, nullptr, u[8], sig[0],
                      sig[1], u[3], u[4], u[1], u[2], 0)) {
        std::fprintf(stderr, "Initialize failed: %s\n", drk_last_error());
        return 7;
    }
    drk_kernel_set_seed(k.handle(), seed, 0);
    k.Render(u[5], u[6], u[3], u[4], u[7]);
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 8;
    for (int b = 0; b < RenderKernel::BUFFER_COUNT; ++b) std::fwrite(k.phb[b], 16, (size_t)u[3] * u[4], o);
    std::fclose(o);
    k.Destroy();
    KernelSampler::BaseLight::DeleteAll();
    std::printf("facade render ok\n");
    return 0;
}
