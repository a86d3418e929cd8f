// svgf_internal.hpp — types shared by the host library and the gfx950 kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace svgf {

// A pitched 2D plane as the kernels see it.
struct PlaneDev {
    char* ptr;
    uint64_t pitch;
};

// Everything one frame's kernels need. The single-frame form passes it by value; the
// batched form keeps an array of them in device memory, indexed from the tile table.
struct FrameDesc {
    int32_t w, h;
    int32_t history_valid;
    int32_t tiles_x;        // tiles per row (batched form)
    PlaneDev rad;           // radiance (storage width)
    PlaneDev g;             // normal.xyz + depth, float4
    PlaneDev gprev;         // previous G-buffer, float4
    PlaneDev gsave;         // batched form: where this frame's G is kept for the next one
    PlaneDev mv;            // motion, float2
    PlaneDev hist_prev;     // colour history (rgb, length), storage width x4
    PlaneDev hist_curr;
    PlaneDev mom_prev;      // moments, storage width x2
    PlaneDev mom_curr;
    PlaneDev den;           // output, storage width x4
    PlaneDev wa, wb;        // wavelet ping-pong scratch, storage width x4
    PlaneDev ctmp;          // temporal colour + variance, storage width x4 (canonical)
    PlaneDev nlen;          // history length, uint8 (canonical)
};

// Pass parameters as plain floats (svgf_params after validation).
struct KParams {
    float phi_luma, phi_normal, phi_depth;
    float alpha_color, alpha_moments;
    float variance_eps, depth_eps;
    float depth_reject, normal_reject;
    int32_t history_cap;
    int32_t iterations;
    int32_t compensated;
};

// Tile = 64 x 4 pixels, one wave64 per row: coalesced 16 B (FP32) / 8 B (FP16) accesses.
constexpr int TILE_W = 64;
constexpr int TILE_H = 4;

// Batched form: the device-side table of a batch. Frame i owns tiles
// [starts[i], starts[i+1]) in row-major tile order; a block maps its tile index to a frame by
// binary search over `starts` (uniform across the block, so it runs on the scalar unit).
struct BatchTable {
    const uint32_t* hdr;       // hdr[0] = frames, hdr[1] = tiles
    const uint32_t* starts;    // first tile of each frame
    const FrameDesc* frames;
};
constexpr uint32_t MAX_BATCH_FRAMES = 1024;

enum class Storage : int { F32 = 0, F16 = 1 };

// Stage launchers (svgf_kernels.hip). All enqueue on `stream` and return hipGetLastError().
// `ev` (optional) receives stages + 1 events: ev[0] before the first stage, ev[i + 1] after
// stage i.
constexpr int MAX_STAGES = 12;
hipError_t launch_reference_frame(const FrameDesc& f, const KParams& p, Storage st, hipStream_t stream,
                                  hipEvent_t* ev = nullptr);
hipError_t launch_canonical_frame(const FrameDesc& f, const KParams& p, Storage st, hipStream_t stream,
                                  hipEvent_t* ev = nullptr);
// Batched: the table is already in device memory. Grid-stride over the tiles; launch shapes
// depend only on grid_blocks, so the sequence can be captured once and replayed.
hipError_t launch_reference_batch(const BatchTable& t, const KParams& p, Storage st, int grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev = nullptr);
hipError_t launch_canonical_batch(const BatchTable& t, const KParams& p, Storage st, int grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev = nullptr);

}  // namespace svgf
