// svgf_internal.hpp — types shared by the host library and the gfx950 kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace svgf {

// A pitched 2D plane as the kernels see it. Kernels address it through a raw buffer resource
// (base, pitch * height bytes), so every plane must stay below 2 GiB (checked by the host).
struct PlaneDev {
    char* ptr;
    uint64_t pitch;
};

// Everything one frame's kernels need. The single-frame form passes it by value; the
// batched form keeps an array of them in device memory.
struct FrameDesc {
    int32_t w, h;
    int32_t history_valid;
    int32_t reserved;
    PlaneDev rad;        // radiance (storage width x4)
    PlaneDev g;          // normal.xyz + depth, float4
    PlaneDev gprev;      // previous G-buffer, float4
    PlaneDev gsave;      // batched form: where this frame's G is kept for the next one
    PlaneDev mv;         // motion, float2
    PlaneDev hist_prev;  // colour history (rgb, length), storage width x4
    PlaneDev hist_curr;
    PlaneDev mom_prev;   // moments, storage width x2
    PlaneDev mom_curr;
    PlaneDev den;        // output, storage width x4
    PlaneDev wa, wb;     // wavelet ping-pong scratch, storage width x4
    PlaneDev nlen;       // history length, uint8 (canonical)
    PlaneDev grad;       // per-pixel depth gradient (gx, gy), float2 (canonical)
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

// Work decomposition. Every stage runs on 16x16 tiles of a stride-s lattice: for a-trous
// iteration k (s = 2^k) the pixels of one residue class (x mod s, y mod s) form a lattice
// whose 5x5 taps are the lattice neighbours, so a tile plus a 2-point halo (20x20 entries)
// holds every tap in LDS whatever the stride. Pointwise stages use s = 1 (plain pixel tiles).
// A frame at stride s has s^2 * ceil(ceil(w/s)/16) * ceil(ceil(h/s)/16) tiles, ordered tile-
// major with the residue fastest (the s^2 blocks of one 16s x 16s pixel region are adjacent).
constexpr int LT = 16;
constexpr int LHALO = 2;
constexpr int LREG = LT + 2 * LHALO;  // 20
constexpr int BLOCK = LT * LT;        // 256 threads

__host__ __device__ inline uint32_t tiles_at(uint32_t w, uint32_t h, uint32_t log2s)
{
    uint32_t s = 1u << log2s;
    uint32_t lx = (w + s - 1) >> log2s, ly = (h + s - 1) >> log2s;
    return (s * s) * ((lx + LT - 1) / LT) * ((ly + LT - 1) / LT);
}

// Stage list of a mode: REFERENCE = init + N a-trous; CANONICAL = temporal, variance,
// N a-trous. log2 stride of stage i:
__host__ __device__ inline uint32_t stage_log2s(int canonical, int stage)
{
    int first_atrous = canonical ? 2 : 1;
    return stage < first_atrous ? 0u : (uint32_t)(stage - first_atrous);
}
constexpr int MAX_STAGES = 12;
constexpr uint32_t MAX_BATCH_FRAMES = 1024;

// Batched form: the device-side table of a batch. hdr[0] = frames; for stage i, frame f owns
// tiles [starts[i * (MAX_BATCH_FRAMES + 1) + f], ... + f + 1); a block maps its tile index to
// a frame by binary search over that row (uniform across the block: scalar unit).
struct BatchTable {
    const uint32_t* hdr;
    const uint32_t* starts;
    const FrameDesc* frames;
};
constexpr uint32_t STARTS_ROW = MAX_BATCH_FRAMES + 1;

enum class Storage : int { F32 = 0, F16 = 1 };

// Stage launchers (svgf_kernels.hip). All enqueue on `stream` and return hipGetLastError().
// `ev` (optional) receives stages + 1 events: ev[0] before the first stage, ev[i + 1] after
// stage i.
hipError_t launch_reference_frame(const FrameDesc& f, const KParams& p, Storage st, hipStream_t stream,
                                  hipEvent_t* ev = nullptr);
hipError_t launch_canonical_frame(const FrameDesc& f, const KParams& p, Storage st, hipStream_t stream,
                                  hipEvent_t* ev = nullptr);
// Batched: the table is already in device memory; stage i launches grid_blocks[i] blocks, one
// per tile of that stage over all frames (a captured graph updates these per batch).
hipError_t launch_reference_batch(const BatchTable& t, const KParams& p, Storage st, const uint32_t* grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev = nullptr);
hipError_t launch_canonical_batch(const BatchTable& t, const KParams& p, Storage st, const uint32_t* grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev = nullptr);

}  // namespace svgf

// Read-only accessors of an svgf_context for other parts of the library (server).
struct svgf_context;
struct svgf_desc;
namespace svgf {
uint32_t context_storage(const svgf_context* c);
uint32_t context_mode(const svgf_context* c);
int context_device(const svgf_context* c);
const ::svgf_desc& context_desc(const svgf_context* c);
}  // namespace svgf
