// svgf_internal.hpp — types shared by the host library and the gfx950 kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "svgf_sched.hpp"
#include "svgf_state.hpp"

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
    int32_t wide;  // FP16 CANONICAL wide state (svgf_state.hpp)
    // Canonical a-trous edge-stop constants, folded on the host so the exponent of a tap is
    // two fused multiply-adds (the filter works in base 2):
    //   luma_k  = -log2(e) / phi_luma
    //   depth_c = ln(2) phi_depth,  depth_e = -ln(2) depth_eps (-inf when phi_depth is 0)
    // so that -log2(e) / (phi_depth |t| + depth_eps) = 1 / (depth_e - depth_c |t|), which is
    // -0 (depth term off) when phi_depth is 0.
    float luma_k, depth_c, depth_e;
};

// Batched form: the device-side table of a batch (header and starts rows as in
// svgf_sched.hpp); a block maps its tile index to a frame by binary search over the stage's
// row (uniform across the block: scalar unit).
struct BatchTable {
    const uint32_t* hdr;
    const uint32_t* starts;
    const FrameDesc* frames;
};

enum class Storage : int { F32 = 0, F16 = 1 };

// Form of the CANONICAL temporal state: raw moments (mu1, mu2), compensated (mean, sigma), or
// compensated with the wide FP16 history and moment words of svgf_state.hpp.
enum StateForm : int { SF_RAW = 0, SF_COMP = 1, SF_WIDE = 2 };
inline int state_form(const KParams& p) { return p.wide ? SF_WIDE : (p.compensated ? SF_COMP : SF_RAW); }

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

// Persistent batched form: one launch of `blocks` workgroups runs every stage of the batch,
// taking tickets from *ctl (protocol in svgf_sched.hpp). *ctl must be zero before the first
// launch and is zero again after each one. *err (host-mapped) becomes 1 when a dependency wait
// gave up after max_polls polls.
hipError_t launch_batch_persistent(int canonical, const BatchTable& t, const KParams& p, Storage st, PersistCtl* ctl,
                                   uint32_t* err, uint64_t max_polls, uint32_t blocks, hipStream_t stream);
// Resident workgroups per CU of the persistent kernel (0 if the query failed).
int persistent_blocks_per_cu(int canonical, Storage st, int state_form);

// Makes `dev` current for the scope of the object and restores the caller's device after.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
    DeviceGuard(const DeviceGuard&) = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;
};

inline bool aligned_to(const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

// Per-call descriptor tables of the one-launch-per-call kernels (wire encoder, composite):
// RING slots of pinned host staging and a device copy. A call fills the next slot's staging,
// copies it to the device table on the caller's stream and records the slot's event; a slot is
// written again RING calls later, after its event (the copy that read the staging) completed.
struct TableRing {
    static constexpr int RING = 3;
    struct Slot {
        void* staging = nullptr;  // pinned
        void* table = nullptr;    // device
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    Slot slots[RING];
    size_t bytes = 0;
    int next = 0;

    bool create(size_t table_bytes)
    {
        bytes = table_bytes;
        for (auto& s : slots) {
            if (hipHostMalloc(&s.staging, bytes, hipHostMallocDefault) != hipSuccess ||
                hipMalloc(&s.table, bytes) != hipSuccess ||
                hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
                return false;
        }
        return true;
    }
    void destroy()
    {
        for (auto& s : slots) {
            if (s.pending) (void)hipEventSynchronize(s.done);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.staging) (void)hipHostFree(s.staging);
            if (s.table) (void)hipFree(s.table);
            s = Slot{};
        }
    }
    // The next slot, with its staging free to write (nullptr on a HIP error).
    Slot* acquire()
    {
        Slot& s = slots[next];
        next = (next + 1) % RING;
        if (s.pending) {
            if (hipEventSynchronize(s.done) != hipSuccess) return nullptr;
            s.pending = false;
        }
        return &s;
    }
    // Enqueue the copy of the first `used` bytes of the slot's staging into its device table.
    static bool upload(Slot& s, size_t used, hipStream_t stream)
    {
        return hipMemcpyAsync(s.table, s.staging, used, hipMemcpyHostToDevice, stream) == hipSuccess;
    }
    // After the launch that reads the table: the slot is free once this point of the stream is.
    static bool release(Slot& s, hipStream_t stream)
    {
        if (hipEventRecord(s.done, stream) != hipSuccess) return false;
        s.pending = true;
        return true;
    }
};

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
