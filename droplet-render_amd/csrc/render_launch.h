// Launch interface between the host library and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include "drk_common.h"

namespace drk {

// Work decomposition of one render launch: units are (pixel, sample chunk)
// pairs, unit u -> chunk u / npix, pixel u % npix (row-major in the tile).
struct WorkParams {
    uint32_t *counter;        // next unfetched unit (device, reset per launch)
    uint32_t total_units;     // npix * nchunks
    uint32_t npix;            // rx * ry
    uint32_t chunk;           // samples per unit
    uint32_t nchunks;
    float *partial;           // [buffers][nchunks][npix] float4 when nchunks > 1, else nullptr
    float *out_l, *out_e;     // tile float4 outputs (stride tilex)
    float *sq_l, *sq_e;       // sum-of-squares outputs or nullptr
    unsigned long long *counts;  // [node touches, brick lookups] for the counting pass
    const float *adj_l, *adj_e;  // adjoint mode: dL/d(phb sums), tile float4 (stride tilex)
    float *grad;                 // adjoint mode: dL/d(FOG brick values), accumulated
};

// Instantiations of the path tracer.
constexpr int kModeForward = 0;     // sums
constexpr int kModeForwardSq = 1;   // sums and sums of squares
constexpr int kModeCount = 2;       // sums, squares and work counters
constexpr int kModeAdjoint = 3;     // density gradients by path replay

int render_block_size();
int render_blocks_per_cu(int mode);
// ev0/ev1 (may be null) are recorded on s right before and after k_render.
hipError_t launch_render(const RenderParams &P, const WorkParams &W, uint32_t grid, int mode, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);

}  // namespace drk
