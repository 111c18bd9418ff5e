// Writes Philox-4x32-10 known-answer vectors computed by rocRAND's
// independent implementation (rocrand_philox4x32_10.h, ROCm 7.2), used to pin
// the renderer's and the oracle's Philox. Build and run (host code only):
//   hipcc -O1 -o /tmp/philox_kat tests/golden/philox_kat_rocrand.cpp && /tmp/philox_kat > tests/golden/philox_kat.txt
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct Exposed : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

int main() {
    Exposed e;
    const unsigned cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 2u, 3u, 4u, 5u, 6u},
        {12345u, 7u, 0u, 1u, 589823u, 0xdeadbeefu},
    };
    for (auto &c : cases) {
        uint4 r = e.run(make_uint4(c[0], c[1], c[2], c[3]), make_uint2(c[4], c[5]));
        std::printf("%08x %08x %08x %08x %08x %08x -> %08x %08x %08x %08x\n", c[0], c[1], c[2], c[3], c[4], c[5],
                    r.x, r.y, r.z, r.w);
    }
    return 0;
}
