// svgf_agent_consumer.cpp — a C++20 host of the agent-frame entry points of include/svgf.h in
// the tick the integration guide describes: per tick one svgf_denoise_batch over every agent
// stream, svgf_composite_run over the denoised output, svgf_wire_encode and one
// svgf_readback_push into the pinned ring. It runs the same ticks twice: in sequence on one
// stream, and pipelined (the wire buffers alternate with the tick; once the next tick's denoise
// and composite are enqueued, the host waits for a tick's encode and pushes its buffer on a copy
// stream; tick k + 2's encode waits on the device for that push). It compares the ring's bytes
// tick by tick. Inputs come from svgf_synth_frame; every plane is FP32 and tightly packed.
//
// usage: svgf_agent_consumer <agents> <ticks>
// exit status 0 and "ok ..." on stdout: every tick's bytes equal in both schedules, and
// consecutive ticks differ.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "svgf.h"

namespace {

void check(svgf_status s, const char* what)
{
    if (s != SVGF_OK) {
        std::fprintf(stderr, "svgf_agent_consumer: %s: %s\n", what, svgf_status_str(s));
        std::exit(1);
    }
}
void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "svgf_agent_consumer: %s failed (%d)\n", what, (int)e);
        std::exit(1);
    }
}

struct Buffer {
    void* ptr = nullptr;
    explicit Buffer(size_t bytes) { check(hipMalloc(&ptr, bytes), "hipMalloc"); }
    ~Buffer() { (void)hipFree(ptr); }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
};

svgf_plane plane(const Buffer& b, uint32_t w, uint32_t bytes_px, uint32_t fmt)
{
    return svgf_plane{b.ptr, (uint64_t)w * bytes_px, fmt, 0};
}

struct Agent {
    uint32_t w, h;
    std::vector<Buffer*> owned;
    Buffer* radiance[2];
    Buffer* gbuffer[2];
    Buffer* motion[2];
    Buffer *denoised, *emissive, *fresnel, *reflection, *refl_colour, *refl_base, *out;
    Buffer* make(size_t bytes)
    {
        owned.push_back(new Buffer(bytes));
        return owned.back();
    }
};

std::vector<uint8_t> run(std::vector<Agent>& agents, int ticks, bool pipelined, size_t& wire_bytes)
{
    const uint32_t n = (uint32_t)agents.size();
    uint64_t npx = 0;
    for (const Agent& a : agents) npx += (uint64_t)a.w * a.h;
    std::vector<size_t> rgb_off(n), dep_off(n);
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        rgb_off[i] = off;
        off += ((size_t)3 * agents[i].w * agents[i].h + 15) / 16 * 16;
    }
    for (uint32_t i = 0; i < n; ++i) {
        dep_off[i] = off;
        off += (size_t)4 * agents[i].w * agents[i].h;
    }
    wire_bytes = off;
    Buffer wire_buf0(off), wire_buf1(off);
    Buffer* wire_buf[2] = {&wire_buf0, &wire_buf1};

    svgf_desc d{};
    d.mode = SVGF_MODE_CANONICAL;
    d.storage = SVGF_STORAGE_FP32;
    check(svgf_default_params(d.mode, d.storage, &d.params), "svgf_default_params");
    d.max_width = 256;
    d.max_height = 256;
    d.max_streams = n;
    d.max_batch_frames = n;
    d.max_batch_pixels = npx;
    svgf_context* ctx = nullptr;
    check(svgf_create(&d, &ctx), "svgf_create");
    svgf_composite* comp = nullptr;
    check(svgf_composite_create(0, n, &comp), "svgf_composite_create");
    svgf_wire* wire = nullptr;
    check(svgf_wire_create(0, n, &wire), "svgf_wire_create");
    svgf_readback* rb = nullptr;
    check(svgf_readback_create(0, off, 3, &rb), "svgf_readback_create");
    svgf_composite_params cp;
    check(svgf_composite_default_params(&cp), "svgf_composite_default_params");
    svgf_wire_params wp;
    check(svgf_wire_default_params(&wp), "svgf_wire_default_params");

    std::vector<uint32_t> sid(n);
    for (uint32_t i = 0; i < n; ++i) check(svgf_stream_open(ctx, agents[i].w, agents[i].h, &sid[i]), "svgf_stream_open");
    // requests per frame parity, composite frames, wire frames per (frame parity, wire buffer)
    std::vector<svgf_request> req[2];
    std::vector<svgf_composite_frame> cf(n);
    std::vector<svgf_wire_frame> wf[2][2];
    for (int f = 0; f < 2; ++f) {
        req[f].resize(n);
        for (int b = 0; b < 2; ++b) wf[f][b].resize(n);
    }
    for (uint32_t i = 0; i < n; ++i) {
        Agent& a = agents[i];
        for (int f = 0; f < 2; ++f) {
            svgf_request& r = req[f][i];
            r.stream_id = sid[i];
            r.flags = SVGF_REQUEST_KEEP_GBUFFER;  // each frame's G-buffer stays until the next
            r.radiance = plane(*a.radiance[f], a.w, 16, SVGF_FMT_RGBA32F);
            r.gbuffer = plane(*a.gbuffer[f], a.w, 16, SVGF_FMT_RGBA32F);
            r.motion = plane(*a.motion[f], a.w, 8, SVGF_FMT_RG32F);
            r.denoised = plane(*a.denoised, a.w, 16, SVGF_FMT_RGBA32F);
            for (int b = 0; b < 2; ++b) {
                svgf_wire_frame& q = wf[f][b][i];
                q.width = a.w;
                q.height = a.h;
                q.colour = plane(*a.out, a.w, 16, SVGF_FMT_RGBA32F);
                q.gbuffer = r.gbuffer;
                q.rgb = static_cast<char*>(wire_buf[b]->ptr) + rgb_off[i];
                q.depth = static_cast<char*>(wire_buf[b]->ptr) + dep_off[i];
            }
        }
        svgf_composite_frame& c = cf[i];
        c.width = a.w;
        c.height = a.h;
        for (int k = 0; k < 16; ++k) c.prev_view_projection[k] = (k % 5 == 0) ? 1.0f : 0.0f;
        c.history = plane(*a.denoised, a.w, 16, SVGF_FMT_RGBA32F);
        c.emissive = plane(*a.emissive, a.w, 16, SVGF_FMT_RGBA32F);
        c.fresnel = plane(*a.fresnel, a.w, 16, SVGF_FMT_RGBA32F);
        c.reflection = plane(*a.reflection, a.w, 16, SVGF_FMT_RGBA32F);
        c.reflection_colour = plane(*a.refl_colour, a.w, 16, SVGF_FMT_RGBA32F);
        c.reflection_base = plane(*a.refl_base, a.w, 16, SVGF_FMT_RGBA32F);
        c.out = plane(*a.out, a.w, 16, SVGF_FMT_RGBA32F);
    }

    hipStream_t main_s, copy_s;
    check(hipStreamCreate(&main_s), "hipStreamCreate");
    check(hipStreamCreate(&copy_s), "hipStreamCreate");
    hipEvent_t encoded[2], pushed[2];
    for (int b = 0; b < 2; ++b) {
        check(hipEventCreateWithFlags(&encoded[b], hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&pushed[b], hipEventDisableTiming), "hipEventCreate");
    }
    std::vector<uint8_t> bytes((size_t)ticks * off);
    std::vector<uint64_t> seq_of(ticks, 0);
    auto read_tick = [&](int k) {
        const void* host = nullptr;
        uint64_t got = 0;
        check(svgf_readback_view(rb, seq_of[k], 1, &host, &got), "svgf_readback_view");
        if (got != off) std::exit(1);
        std::memcpy(bytes.data() + (size_t)k * off, host, off);
    };
    auto push = [&](int k) {  // the host has seen tick k's encode finish: the copy has its input
        const int b = k % 2;
        check(hipEventSynchronize(encoded[b]), "hipEventSynchronize");
        check(svgf_readback_push(rb, wire_buf[b]->ptr, off, copy_s, &seq_of[k]), "svgf_readback_push");
        check(hipEventRecord(pushed[b], copy_s), "hipEventRecord");
        if (k >= 1) read_tick(k - 1);  // the ring holds 3 slots: read tick k - 1 before it is reused
    };
    for (int k = 0; k < ticks; ++k) {
        const int f = k % 2, b = pipelined ? k % 2 : 0;
        check(svgf_denoise_batch(ctx, req[f].data(), n, main_s), "svgf_denoise_batch");
        check(svgf_composite_run(comp, cf.data(), n, &cp, main_s), "svgf_composite_run");
        if (!pipelined) {
            check(svgf_wire_encode(wire, wf[f][b].data(), n, &wp, main_s), "svgf_wire_encode");
            check(svgf_readback_push(rb, wire_buf[b]->ptr, off, main_s, &seq_of[k]), "svgf_readback_push");
            read_tick(k);
            continue;
        }
        if (k >= 1) push(k - 1);
        if (k >= 2) check(hipStreamWaitEvent(main_s, pushed[b], 0), "hipStreamWaitEvent");
        check(svgf_wire_encode(wire, wf[f][b].data(), n, &wp, main_s), "svgf_wire_encode");
        check(hipEventRecord(encoded[b], main_s), "hipEventRecord");
    }
    if (pipelined) {
        push(ticks - 1);
        read_tick(ticks - 1);
    }
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    for (int b = 0; b < 2; ++b) {
        (void)hipEventDestroy(encoded[b]);
        (void)hipEventDestroy(pushed[b]);
    }
    (void)hipStreamDestroy(main_s);
    (void)hipStreamDestroy(copy_s);
    check(svgf_readback_destroy(rb), "svgf_readback_destroy");
    check(svgf_wire_destroy(wire), "svgf_wire_destroy");
    check(svgf_composite_destroy(comp), "svgf_composite_destroy");
    check(svgf_destroy(ctx), "svgf_destroy");
    return bytes;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: svgf_agent_consumer <agents> <ticks>\n");
        return 2;
    }
    const int n = std::atoi(argv[1]), ticks = std::atoi(argv[2]);
    if (n < 1 || ticks < 2) return 2;
    const uint32_t eye[5] = {8, 32, 64, 128, 256};  // docs/AI_INTERFACE.md:354-360
    std::vector<Agent> agents(n);
    for (int i = 0; i < n; ++i) {
        Agent& a = agents[i];
        a.w = a.h = eye[i % 5];
        const size_t px = (size_t)a.w * a.h;
        std::vector<float> rad(px * 4), g(px * 4), mv(px * 2), con(px * 4);
        for (int f = 0; f < 2; ++f) {
            svgf_synth_desc sd{};
            sd.width = a.w;
            sd.height = a.h;
            sd.seed = 31 * (uint64_t)i + (uint64_t)f;
            sd.stream = (uint64_t)i;
            sd.frame = (uint32_t)f;
            sd.pan_px = 0.5f;
            sd.disocclusion = 0.05f;
            sd.threads = 1;
            check(svgf_synth_frame(&sd, rad.data(), g.data(), mv.data()), "svgf_synth_frame");
            a.radiance[f] = a.make(px * 16);
            a.gbuffer[f] = a.make(px * 16);
            a.motion[f] = a.make(px * 8);
            check(hipMemcpy(a.radiance[f]->ptr, rad.data(), px * 16, hipMemcpyHostToDevice), "hipMemcpy");
            check(hipMemcpy(a.gbuffer[f]->ptr, g.data(), px * 16, hipMemcpyHostToDevice), "hipMemcpy");
            check(hipMemcpy(a.motion[f]->ptr, mv.data(), px * 8, hipMemcpyHostToDevice), "hipMemcpy");
        }
        a.denoised = a.make(px * 16);
        a.out = a.make(px * 16);
        // a plain surface: emissive 0.05, F_view 0.9, no reflection (kind 0)
        auto fill = [&](Buffer*& dst, float v, float alpha) {
            for (size_t p = 0; p < px; ++p) {
                con[4 * p] = con[4 * p + 1] = con[4 * p + 2] = v;
                con[4 * p + 3] = alpha;
            }
            dst = a.make(px * 16);
            check(hipMemcpy(dst->ptr, con.data(), px * 16, hipMemcpyHostToDevice), "hipMemcpy");
        };
        fill(a.emissive, 0.05f, 1.0f);
        fill(a.fresnel, 0.9f, 1.0f);
        fill(a.reflection, 0.0f, 0.0f);
        fill(a.refl_colour, 0.0f, 1.0f);
        fill(a.refl_base, 0.0f, 1.0f);
    }
    size_t off = 0;
    const std::vector<uint8_t> seq = run(agents, ticks, false, off);
    const std::vector<uint8_t> pipe = run(agents, ticks, true, off);
    for (int k = 0; k < ticks; ++k) {
        if (std::memcmp(seq.data() + (size_t)k * off, pipe.data() + (size_t)k * off, off) != 0) {
            std::printf("tick %d differs between the schedules\n", k);
            return 1;
        }
        if (k > 0 && std::memcmp(seq.data() + (size_t)k * off, seq.data() + (size_t)(k - 1) * off, off) == 0) {
            std::printf("ticks %d and %d are equal\n", k - 1, k);
            return 1;
        }
    }
    for (Agent& a : agents)
        for (Buffer* b : a.owned) delete b;
    std::printf("ok %d agents, %d ticks, %zu wire bytes per tick\n", n, ticks, off);
    return 0;
}
