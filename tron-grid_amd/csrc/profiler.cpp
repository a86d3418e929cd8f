// profiler.cpp — TronGrid's per-pass GPU profiler on hipEvents (host code).
//
// The reference (src/main.cpp:145-180, 1180-1207, 2902-2953) writes a start / end timestamp
// pair per pass into a query pool slot per frame in flight; when a slot comes round again,
// after the wait on that slot's fence, it reads the pairs back, turns each into nanoseconds,
// folds them into per-pass exponential moving averages (the first readback initialises them)
// and every GPU_PROFILER_LOG_INTERVAL_FRAMES frames logs "GPU times (ms): name=ms ...". Here a
// slot holds one start / end hipEvent pair per pass; the readback waits on the slot's end
// events (the fence), and hipEventElapsedTime gives the duration. The bookkeeping (slot
// ring, pending flags, frame counter, when the average starts and when a line is due) follows
// the reference loop step for step.
#include "../../include/svgf.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

constexpr uint32_t MAX_PASSES = 32;
constexpr uint32_t MAX_SLOTS = 8;
constexpr size_t NAME_MAX_CHARS = 31;

struct Slot {
    std::vector<hipEvent_t> start, end;  // one pair per pass (device profilers)
    std::vector<float> host_ns;          // host profilers: the frame's samples
    uint32_t recorded_begin = 0, recorded_end = 0;  // bit p: pass p's start / end recorded
    bool pending = false;                // holds a frame not yet read back
};

}  // namespace

struct svgf_profiler {
    uint32_t passes = 0, slots = 0, interval = 0;
    float alpha = 0.0f;
    uint64_t counter = 0;
    int device = -1;
    std::vector<std::string> names;
    std::vector<Slot> ring;
    uint32_t current = 0;
    bool in_frame = false;
    std::vector<float> ema;
    bool ema_init = false;
    std::string line;
};

namespace {

void destroy_events(svgf_profiler* p)
{
    for (auto& s : p->ring) {
        for (hipEvent_t e : s.start)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : s.end)
            if (e) (void)hipEventDestroy(e);
        s.start.clear();
        s.end.clear();
    }
}

// Read back slot s: per-pass durations in nanoseconds (0 for a pass not bracketed that frame).
svgf_status read_slot(svgf_profiler* p, Slot& s, std::vector<float>& ns)
{
    ns.assign(p->passes, 0.0f);
    if (p->device < 0) {
        ns = s.host_ns;
        return SVGF_OK;
    }
    for (uint32_t k = 0; k < p->passes; ++k) {
        const uint32_t bit = 1u << k;
        if (!(s.recorded_begin & bit) || !(s.recorded_end & bit)) continue;
        if (hipEventSynchronize(s.end[k]) != hipSuccess) return SVGF_ERR_HIP;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s.start[k], s.end[k]) != hipSuccess) return SVGF_ERR_HIP;
        ns[k] = ms * 1.0e6f;
    }
    return SVGF_OK;
}

struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev)
    {
        if (dev < 0) return;
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceScope()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

svgf_status svgf_profiler_default_desc(svgf_profiler_desc* out)
{
    if (!out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = svgf_profiler_desc{};
    out->passes = 1;
    out->slots = 2;            // MAX_FRAMES_IN_FLIGHT, src/main.cpp:74
    out->ema_alpha = 0.05f;    // GPU_PROFILER_EMA_ALPHA, src/main.cpp:170
    out->log_interval = 60;    // GPU_PROFILER_LOG_INTERVAL_FRAMES, src/main.cpp:175
    out->first_frame = 16;     // WARMUP_ITERATIONS, src/main.cpp:2611-2612
    out->device = 0;
    out->names = nullptr;
    return SVGF_OK;
}

svgf_status svgf_profiler_create(const svgf_profiler_desc* d, svgf_profiler** out)
{
    if (!d || !out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (d->passes < 1 || d->passes > MAX_PASSES || d->slots < 1 || d->slots > MAX_SLOTS || d->log_interval < 1 ||
        !(d->ema_alpha >= 0.0f && d->ema_alpha <= 1.0f) || d->device < -1)
        return SVGF_ERR_INVALID_ARGUMENT;
    if (d->names) {
        for (uint32_t k = 0; k < d->passes; ++k)
            if (!d->names[k] || std::strlen(d->names[k]) > NAME_MAX_CHARS) return SVGF_ERR_INVALID_ARGUMENT;
    }
    if (d->device >= 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || d->device >= ndev) return SVGF_ERR_NO_DEVICE;
    }
    svgf_profiler* p = new (std::nothrow) svgf_profiler();
    if (!p) return SVGF_ERR_OUT_OF_MEMORY;
    p->passes = d->passes;
    p->slots = d->slots;
    p->interval = d->log_interval;
    p->alpha = d->ema_alpha;
    p->counter = d->first_frame;
    p->device = d->device;
    for (uint32_t k = 0; k < d->passes; ++k)
        p->names.push_back(d->names ? std::string(d->names[k]) : "pass" + std::to_string(k));
    p->ring.resize(d->slots);
    p->ema.assign(d->passes, 0.0f);
    if (p->device >= 0) {
        DeviceScope scope(p->device);
        for (auto& s : p->ring) {
            s.start.assign(p->passes, nullptr);
            s.end.assign(p->passes, nullptr);
            for (uint32_t k = 0; k < p->passes; ++k) {
                if (hipEventCreate(&s.start[k]) != hipSuccess || hipEventCreate(&s.end[k]) != hipSuccess) {
                    destroy_events(p);
                    delete p;
                    return SVGF_ERR_HIP;
                }
            }
        }
    } else {
        for (auto& s : p->ring) s.host_ns.assign(p->passes, 0.0f);
    }
    *out = p;
    return SVGF_OK;
}

svgf_status svgf_profiler_destroy(svgf_profiler* p)
{
    if (!p) return SVGF_ERR_INVALID_ARGUMENT;
    {
        DeviceScope scope(p->device);
        destroy_events(p);
    }
    delete p;
    return SVGF_OK;
}

svgf_status svgf_profiler_begin_frame(svgf_profiler* p, int32_t* logged)
{
    if (!p || p->in_frame) return SVGF_ERR_INVALID_ARGUMENT;
    if (logged) *logged = 0;
    Slot& s = p->ring[p->current];
    if (s.pending) {  // src/main.cpp:2910: the slot's previous frame has finished (its fence)
        DeviceScope scope(p->device);
        std::vector<float> ns;
        const svgf_status st = read_slot(p, s, ns);
        if (st != SVGF_OK) return st;
        if (!p->ema_init) {
            p->ema = ns;
            p->ema_init = true;
        } else {
            for (uint32_t k = 0; k < p->passes; ++k) p->ema[k] = (p->ema[k] * (1.0f - p->alpha)) + (ns[k] * p->alpha);
        }
        if (p->counter % p->interval == 0) {
            std::string line = "GPU times (ms):";
            for (uint32_t k = 0; k < p->passes; ++k) {
                const float ms = p->ema[k] / 1.0e6f;
                char buf[64];
                std::snprintf(buf, sizeof(buf), "%f", (double)ms);
                line += " " + p->names[k] + "=" + buf;
            }
            p->line = line;
            if (logged) *logged = 1;
        }
    }
    ++p->counter;
    s.pending = false;
    s.recorded_begin = s.recorded_end = 0;
    if (p->device < 0) s.host_ns.assign(p->passes, 0.0f);
    p->in_frame = true;
    return SVGF_OK;
}

svgf_status svgf_profiler_pass_begin(svgf_profiler* p, uint32_t pass, void* hip_stream)
{
    if (!p || !p->in_frame || pass >= p->passes || p->device < 0) return SVGF_ERR_INVALID_ARGUMENT;
    Slot& s = p->ring[p->current];
    DeviceScope scope(p->device);
    if (hipEventRecord(s.start[pass], static_cast<hipStream_t>(hip_stream)) != hipSuccess) return SVGF_ERR_HIP;
    s.recorded_begin |= 1u << pass;
    return SVGF_OK;
}

svgf_status svgf_profiler_pass_end(svgf_profiler* p, uint32_t pass, void* hip_stream)
{
    if (!p || !p->in_frame || pass >= p->passes || p->device < 0) return SVGF_ERR_INVALID_ARGUMENT;
    Slot& s = p->ring[p->current];
    if (!(s.recorded_begin & (1u << pass))) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceScope scope(p->device);
    if (hipEventRecord(s.end[pass], static_cast<hipStream_t>(hip_stream)) != hipSuccess) return SVGF_ERR_HIP;
    s.recorded_end |= 1u << pass;
    return SVGF_OK;
}

svgf_status svgf_profiler_sample(svgf_profiler* p, uint32_t pass, float ns)
{
    if (!p || !p->in_frame || pass >= p->passes || p->device >= 0) return SVGF_ERR_INVALID_ARGUMENT;
    p->ring[p->current].host_ns[pass] = ns;
    return SVGF_OK;
}

svgf_status svgf_profiler_end_frame(svgf_profiler* p)
{
    if (!p || !p->in_frame) return SVGF_ERR_INVALID_ARGUMENT;
    p->ring[p->current].pending = true;  // src/main.cpp:3096
    p->current = (p->current + 1) % p->slots;  // src/main.cpp:3148
    p->in_frame = false;
    return SVGF_OK;
}

svgf_status svgf_profiler_ema(svgf_profiler* p, float* ema_ns, uint32_t capacity, uint32_t* count)
{
    if (!p || !count || (!ema_ns && capacity)) return SVGF_ERR_INVALID_ARGUMENT;
    *count = p->passes;
    if (!p->ema_init) return SVGF_ERR_NOT_READY;
    for (uint32_t k = 0; k < p->passes && k < capacity; ++k) ema_ns[k] = p->ema[k];
    return SVGF_OK;
}

svgf_status svgf_profiler_log_line(svgf_profiler* p, char* buf, uint32_t capacity)
{
    if (!p || !buf || capacity < 1) return SVGF_ERR_INVALID_ARGUMENT;
    const size_t n = std::min<size_t>(p->line.size(), capacity - 1);
    std::memcpy(buf, p->line.data(), n);
    buf[n] = '\0';
    return SVGF_OK;
}

}  // extern "C"
