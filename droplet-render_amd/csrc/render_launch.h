// Launch interface between the host library and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include "drk_common.h"

namespace drk {

// Work decomposition of one render launch: units are (pixel, sample chunk)
// pairs. The pixels (row-major in the tile) are split into nbands contiguous
// bands [b*npix/nbands, (b+1)*npix/nbands) of bp pixels; unit u of band b is
// chunk u / bp, pixel b*npix/nbands + u % bp.
constexpr uint32_t kMaxBands = 16;
constexpr uint32_t kCounterStride = 16;   // u32 words: one 64-byte line per band counter
struct WorkParams {
    uint32_t *counter;        // per band: next unfetched unit, kCounterStride apart (reset per launch)
    uint32_t nbands;          // row bands of the tile's pixels (1..kMaxBands)
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
    size_t grad_stride;          // floats between gradient copies
    uint32_t grad_copies;        // copies the waves spread their atomics over (>= 1)
    uint32_t grad_nxcc;          // XCDs of the device
    // Per-sample path record, three planes over i = sample * npix + pixel
    // (samples of this launch, pixels of the rendered extent): float4
    // {lights rgb, alpha}, float4 {env rgb, 0}, uint2 shadow-escape bits.
    // Written by kModeRecord, read by kModeAdjoint (then pass 0 is skipped).
    float *rec;
};

// Instantiations of the path tracer.
constexpr int kModeForward = 0;     // sums
constexpr int kModeForwardSq = 1;   // sums and sums of squares
constexpr int kModeCount = 2;       // sums, squares and work counters
constexpr int kModeAdjoint = 3;     // density gradients by path replay
constexpr int kModeRecord = 4;      // sums plus the per-sample path record

// Bytes of one (pixel, sample) entry of the path record.
constexpr size_t kRecordBytes = 40;

int render_block_size();
int render_blocks_per_cu(int mode);
// ev0/ev1 (may be null) are recorded on s right before and after k_render.
hipError_t launch_render(const RenderParams &P, const WorkParams &W, uint32_t grid, int mode, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);
// out[i] += sum over c < copies of part[c * n + i].
hipError_t launch_grad_sum(const float *part, uint32_t copies, size_t n, float *out, hipStream_t s);

}  // namespace drk
