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

// Unsigned division by an invariant 32-bit divisor d >= 1 with one
// multiply-high (Granlund and Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, Fig. 4.1): with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n), the quotient
// floor(n / d) is (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) for every
// 32-bit n.
struct UDiv {
    uint32_t m, s1, s2;
};
inline UDiv udiv_make(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) ++l;
    UDiv r;
    r.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    r.s1 = l < 1 ? l : 1u;
    r.s2 = l > 0 ? l - 1 : 0u;
    return r;
}
__host__ __device__ inline uint32_t udiv(uint32_t n, const UDiv &D) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t t = __umulhi(D.m, n);
#else
    const uint32_t t = (uint32_t)(((uint64_t)D.m * n) >> 32);
#endif
    return (t + ((n - t) >> D.s1)) >> D.s2;
}

// First sample of chunk k, k <= nchunks. Chunks below nfull hold `chunk`
// samples (the last of them fewer when the samples run out first); each
// chunk after them takes half, rounded up, of the samples that remain, down
// to single samples. Units are handed out chunk by chunk, so the last units
// of a launch are short and its drain leaves few lanes idle.
__host__ __device__ inline uint32_t chunk_start(uint32_t chunk, uint32_t nfull, uint32_t samples, uint32_t k) {
    if (k <= nfull) return (uint32_t)((uint64_t)k * chunk < samples ? (uint64_t)k * chunk : samples);
    uint32_t s = nfull * chunk, rem = samples - s;
    for (uint32_t j = nfull; j < k; ++j) {
        const uint32_t z = (rem + 1) / 2;
        s += z;
        rem -= z;
    }
    return s;
}

// Chunking of `samples` with leading chunks of `chunk` samples: fixed-size
// chunks only (guided false), or full chunks while more than two chunks'
// worth remain and the halving tail after them.
inline void plan_chunks(uint32_t samples, uint32_t chunk, bool guided, uint32_t &nfull, uint32_t &nchunks) {
    if (!guided) {
        nfull = nchunks = (samples + chunk - 1) / chunk;
        return;
    }
    const uint64_t two = 2ull * chunk;
    nfull = samples > two ? (uint32_t)((samples - two + chunk - 1) / chunk) : 0u;
    nchunks = nfull;
    for (uint32_t rem = samples - nfull * chunk; rem > 0; rem -= (rem + 1) / 2) ++nchunks;
}
// Leaf grid: at most 2^6 cells per axis (1 MiB of cells); walked by
// default when at least this fraction of its cells hold a leaf.
constexpr uint32_t kMaxGridLevels = 6;
constexpr float kGridMinOccupancy = 0.125f;
constexpr uint32_t kCounterStride = 16;   // u32 words: one 64-byte line per band counter
// Counter block of a launch: kMaxBands band counters, then one line whose
// first word is the launch's status (1: a wave hit the round limit). The
// whole block is cleared before every launch.
constexpr uint32_t kStatusWord = kMaxBands * kCounterStride;
constexpr uint32_t kCounterWords = kStatusWord + kCounterStride;
struct WorkParams {
    uint32_t *counter;        // counter block: per band the next unfetched unit, kCounterStride apart,
                              // and the status word at kStatusWord (cleared per launch)
    uint32_t round_limit;     // rounds of the state machine a wave may run in this launch (0: no limit)
    uint32_t nbands;          // row bands of the tile's pixels (1..kMaxBands)
    uint32_t band_p0[kMaxBands + 1];   // first pixel of each band (band_p0[nbands] = npix)
    UDiv band_div[kMaxBands];          // division by each band's pixel count
    UDiv rx_div;                       // division by the rendered extent's width (rx)
    uint32_t total_units;     // npix * nchunks
    uint32_t npix;            // rx * ry
    uint32_t chunk;           // samples per unit in the leading chunks
    uint32_t nfull;           // leading chunks of `chunk` samples (see chunk_start)
    uint32_t nchunks;
    float *partial;           // [nchunks][npix][buffers] float4 when nchunks > 1, else nullptr: a unit's
                              // buffers side by side, written by one lane as one 32/64-byte span
    float4 *cam;              // per pixel of the rendered extent (pixel index as for units): the camera
                              // ray's direction and lane mask (w: 1 inside the image), written by a
                              // pass before the render; nullptr: each path computes its camera ray
    float4 *esc_mail;         // per lane of the grid two float4: the escape service's mailbox (nullptr:
                              // escape radiance computed in place)
    float *out_l, *out_e;     // tile float4 outputs (stride tilex)
    float *sq_l, *sq_e;       // sum-of-squares outputs or nullptr
    unsigned long long *counts;  // [node touches, brick lookups] for the counting pass
    const float *adj_l, *adj_e;  // adjoint mode: dL/d(phb sums), tile float4 (stride tilex)
    float *grad;                 // adjoint mode: dL/d(FOG brick values), accumulated
    // Adjoint gradient events (binned by brick and reduced in LDS after the
    // launch; once ev_cap is reached, lanes add to the bricks directly).
    // Both buffers are passed by their allocation base (the planes are
    // located in the kernels; see EventPlanes).
    unsigned long long *ev_count;  // events produced (may exceed ev_cap); base of the bin tables
    char *ev_base;                 // event buffer
    uint32_t ev_cap;               // event slots; 0: direct atomics only
    size_t grad_stride;          // floats between gradient copies
    uint32_t grad_copies;        // copies the waves spread their atomics over (>= 1)
    uint32_t grad_nxcc;          // XCDs of the device
    // Per-sample path record, three planes over i = sample * npix + pixel
    // (samples of this launch, pixels of the rendered extent): float4
    // {lights rgb, alpha}, float4 {env rgb, last contribution (int bits)},
    // uint2 shadow-escape bits.
    // Written by kModeRecord, read by kModeAdjoint (then pass 0 is skipped).
    float *rec;
};

// Instantiations of the path tracer.
constexpr int kModeForward = 0;     // sums
constexpr int kModeForwardSq = 1;   // sums and sums of squares
constexpr int kModeCount = 2;       // sums, squares and work counters
constexpr int kModeAdjoint = 3;     // density gradients by path replay
constexpr int kModeRecord = 4;      // sums plus the per-sample path record
constexpr int kModeWeighted = 5;    // sums of weighted delta tracking against reference densities
constexpr int kModeAdjointNull = 6; // density gradients by path replay, null-collision terms
// Float4 buffers a render mode sums per pixel: lights and env, plus their
// sums of squares in the modes that keep them.
inline uint32_t partial_buffers(int mode) { return (mode == kModeForwardSq || mode == kModeCount) ? 4u : 2u; }

// Bytes of one (pixel, sample) entry of the path record.
constexpr size_t kRecordBytes = 40;

// Fills W's band table and divisors for W.npix pixels in rows of rx and
// W.nbands bands.
inline void set_bands(WorkParams &W, uint32_t rx) {
    for (uint32_t b = 0; b <= W.nbands; ++b) W.band_p0[b] = (uint32_t)((uint64_t)b * W.npix / W.nbands);
    for (uint32_t b = 0; b < W.nbands; ++b) {
        const uint32_t bp = W.band_p0[b + 1] - W.band_p0[b];
        W.band_div[b] = udiv_make(bp ? bp : 1u);   // (an empty band is never divided by)
    }
    W.rx_div = udiv_make(rx ? rx : 1u);
}

// A launch's parameters in device memory, written by a setup kernel right
// before the launch (so the sequence stays kernels only and can be captured
// in a graph): the render kernels read them through a pointer with scalar
// loads where they are used. Passed by value, the argument block (148 words)
// was kept in scalar registers across the state machine's loop, and the
// registers it did not fit spilled into vector-register lanes whose
// read-backs took vector issue slots (256^3 frame forward 27.7 -> 26.7 ms).
struct LaunchParams {
    RenderParams P;
    WorkParams W;
};
// The launch block: the counter block, then the parameters at kParamOffset.
constexpr size_t kParamOffset = 1280;
constexpr size_t kLaunchBlockBytes = kParamOffset + sizeof(LaunchParams);
static_assert(kCounterWords * sizeof(uint32_t) <= kParamOffset, "counter block fits before the parameters");

int render_block_size();
int render_blocks_per_cu(int mode);
// W.counter is the launch block (kLaunchBlockBytes). ev0/ev1 (may be null)
// are recorded on s right before and after k_render.
hipError_t launch_render(const RenderParams &P, const WorkParams &W, uint32_t grid, int mode, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);
// Fixed device scratch of the allocation-free render (one per device, the
// size below): the launch block (counters and parameters) at offset 0,
// partial sums from kScratchPartialOffset. scratch_address gives the current device's copy
// (looked up once per device).
constexpr size_t kScratchBytes = 128ull << 20;
constexpr size_t kScratchPartialOffset = 4096;
constexpr uint32_t kScratchDevices = 64;
hipError_t scratch_address(void **ptr);
static_assert(kLaunchBlockBytes <= kScratchPartialOffset, "launch block fits before the partials");
// Batched launch over njobs (P, W) pairs in device memory (modes: forward,
// forward with squares, forward with the path record, the two adjoints);
// counters (ncounters words, every job's band counters; null: the caller
// cleared them) are cleared first.
int render_batch_blocks_per_cu(int mode);
// cam_pix: the largest job's pixel count when the jobs' camera rays are to be
// computed first (their W.cam), else 0.
hipError_t launch_render_batch(const RenderParams *dP, const WorkParams *dW, uint32_t njobs, uint32_t *counters,
                               uint32_t ncounters, uint32_t grid, int mode, uint32_t cam_pix, hipStream_t s);
// Stores one job's parameters at *dP and *dW and clears its counter block
// (W.counter), as kernels (no copies from host memory, nothing to keep alive).
hipError_t launch_job_setup(const RenderParams &P, const WorkParams &W, RenderParams *dP, WorkParams *dW,
                            hipStream_t s);
// Chunk-order sum of a launch's partials into its outputs (no-op without partials).
hipError_t launch_reduce(const WorkParams &W, uint32_t rx, uint32_t tilex, uint32_t buffers, hipStream_t s);
// out[i] += sum over c < copies of part[c * n + i].
hipError_t launch_grad_sum(const float *part, uint32_t copies, size_t n, float *out, hipStream_t s);

// Reduction of the adjoint's gradient events into the FOG gradient: a
// counting sort by bin (2^shift bricks per bin) followed by one LDS
// accumulation per (bin, chunk of events).
struct EventReduce {
    char *ev_base;         // event buffer (see EventPlanes)
    uint32_t *bins;        // bin tables: count (2 words), hist[nbins], bin_start[nbins+1],
                           // bin_fill[nbins], blk_start[nbins+1], bin_max[nbins]
    uint32_t ev_cap;
    uint32_t nbins, shift; // bin = brick >> shift
    uint32_t acc_log2;     // events per accumulation block: 2^acc_log2 (12..14)
    uint32_t nbricks, lvoxc, lvoxc3;
    const float4 *bce;     // per FOG brick: centre and half-size of its leaf
    float *grad;           // accumulated into
};

// Planes of the event buffer (cap slots each): events {weight, point x, y, z}
// float4 and their FOG brick, then their copies sorted by bin. The lookup
// cell and lerp fractions are not stored: the accumulation recomputes them
// from the point and the brick's leaf box (EventReduce::bce) with the same
// arithmetic as the lookup.
struct EventPlanes {
    float4 *wx;
    uint32_t *b;
    float4 *swx;
    uint32_t *sb;
};
__host__ __device__ inline EventPlanes event_planes(char *base, uint32_t cap) {
    EventPlanes e;
    e.wx = reinterpret_cast<float4 *>(base);
    e.b = reinterpret_cast<uint32_t *>(base + (size_t)cap * 16);
    e.swx = reinterpret_cast<float4 *>(base + (size_t)cap * 20);
    e.sb = reinterpret_cast<uint32_t *>(base + (size_t)cap * 36);
    return e;
}
constexpr uint32_t kEvMaxBins = 8192;          // LDS histogram of the sort
constexpr size_t kEvTableWords = 2 + 5 * (size_t)kEvMaxBins + 2;   // the bin tables (EventReduce::bins)
constexpr size_t kEventSlotBytes = 40;         // event (20 B) + its sorted copy (20 B)
constexpr uint32_t kEvChunk = 2048;            // event slots a wave reserves at a time
constexpr uint32_t kEvAccumLds = 65536;        // bytes of LDS gradient per accumulation block
// Largest shift (bricks per bin = 2^shift) and bin count for a scene, or
// false when a bin's bricks do not fit the accumulation block's LDS.
bool event_bins(uint32_t nbricks, uint32_t lvoxc3, uint32_t &shift, uint32_t &nbins);
hipError_t launch_event_reduce(const EventReduce &E, hipStream_t s);

}  // namespace drk
