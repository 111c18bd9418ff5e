// Host check of the work-unit chunk schedule (plan_chunks / chunk_start):
// prints "samples chunk guided nfull nchunks start_0 ... start_nchunks" per case.
#include <cstdio>
#include "render_launch.h"

int main() {
    const unsigned samples[] = {1, 2, 3, 7, 16, 64, 65, 100, 1000, 4096};
    const unsigned chunks[] = {1, 2, 3, 11, 16, 64, 300};
    for (unsigned s : samples)
        for (unsigned c0 : chunks)
            for (int g = 0; g < 2; ++g) {
                const unsigned c = c0 < s ? c0 : s;
                uint32_t nfull, n;
                drk::plan_chunks(s, c, g != 0, nfull, n);
                std::printf("%u %u %d %u %u", s, c, g, nfull, n);
                for (uint32_t k = 0; k <= n; ++k) std::printf(" %u", drk::chunk_start(c, nfull, s, k));
                std::printf("\n");
            }
    return 0;
}
