// svgf_batch_consumer.cpp — a C++20 host of the batched SVGF C ABI (include/svgf.h): a serving
// process that keeps one resident CANONICAL stream per agent and denoises one frame of every
// stream per batch. It opens the streams (svgf_stream_open), and per frame submits one
// svgf_denoise_batch over all of them, then reads that batch's own outcome (svgf_batch_last,
// svgf_batch_status). Each stream's G-buffer is kept per frame in flight, two planes per stream
// alternating with the frame, so every request carries SVGF_REQUEST_KEEP_GBUFFER: the next frame
// reads the previous G-buffer in place. After the last frame it reads each stream's resident
// state back through svgf_stream_state_get. Planes are pitched hipMallocPitch buffers.
//
// usage: svgf_batch_consumer <storage 0|1> <frames> <dir> <w0> <h0> [<w1> <h1> ...]
//   reads  <dir>/s<i>_radiance_<f>.bin (w h 4 values of the storage type, rows packed),
//          <dir>/s<i>_gbuffer_<f>.bin (w h 4 float32), <dir>/s<i>_motion_<f>.bin (w h 2 float32)
//   writes <dir>/s<i>_denoised_<f>.bin (w h 4 values of the storage type) and, after the last
//          frame, <dir>/s<i>_history.bin and <dir>/s<i>_moments.bin (the stream's state, rows packed)
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "svgf.h"

namespace {

[[noreturn]] void fail(const char* what, int code)
{
    std::fprintf(stderr, "svgf_batch_consumer: %s failed (%d)\n", what, code);
    std::exit(1);
}
void check(svgf_status s, const char* what)
{
    if (s != SVGF_OK) {
        std::fprintf(stderr, "svgf_batch_consumer: %s: %s\n", what, svgf_status_str(s));
        std::exit(1);
    }
}
void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(what, (int)e);
}

std::vector<char> read_file(const std::string& path, size_t bytes)
{
    std::vector<char> buf(bytes);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(buf.data(), 1, bytes, f) != bytes) fail(path.c_str(), 0);
    std::fclose(f);
    return buf;
}
void write_file(const std::string& path, const std::vector<char>& buf)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) fail(path.c_str(), 0);
    std::fclose(f);
}

// One caller-owned pitched device plane.
struct Plane {
    void* ptr = nullptr;
    size_t pitch = 0;
    uint32_t format = 0;
    size_t row = 0;  // bytes of one packed row
    uint32_t h = 0;
    Plane(uint32_t w, uint32_t rows, uint32_t fmt, uint32_t bytes_px) : format(fmt), row((size_t)w * bytes_px), h(rows)
    {
        check(hipMallocPitch(&ptr, &pitch, row, h), "hipMallocPitch");
    }
    ~Plane() { (void)hipFree(ptr); }
    Plane(const Plane&) = delete;
    Plane& operator=(const Plane&) = delete;
    svgf_plane abi() const { return svgf_plane{ptr, (uint64_t)pitch, format, 0}; }
    void upload(const std::string& path, hipStream_t s)
    {
        const std::vector<char> buf = read_file(path, row * h);
        check(hipMemcpy2DAsync(ptr, pitch, buf.data(), row, row, h, hipMemcpyHostToDevice, s), "hipMemcpy2DAsync");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
};

// A plane of the library's resident state (svgf_stream_state_get), downloaded rows packed.
void download_state(const svgf_plane& p, size_t row, uint32_t h, const std::string& path)
{
    std::vector<char> buf(row * h);
    check(hipMemcpy2D(buf.data(), row, p.ptr, (size_t)p.pitch_bytes, row, h, hipMemcpyDeviceToHost), "hipMemcpy2D");
    write_file(path, buf);
}

// One agent's caller-owned planes: radiance, motion and output, and its G-buffer kept per frame
// in flight (two planes, alternating with the frame).
struct Agent {
    uint32_t w, h, id = 0;
    Plane radiance, motion, denoised;
    Plane gbuffer[2];
    Agent(uint32_t w_, uint32_t h_, uint32_t f4, uint32_t b4)
        : w(w_), h(h_), radiance(w_, h_, f4, b4), motion(w_, h_, SVGF_FMT_RG32F, 8), denoised(w_, h_, f4, b4),
          gbuffer{Plane(w_, h_, SVGF_FMT_RGBA32F, 16), Plane(w_, h_, SVGF_FMT_RGBA32F, 16)}
    {
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 6 || (argc - 4) % 2 != 0) {
        std::fprintf(stderr, "usage: %s <storage> <frames> <dir> <w0> <h0> [<w1> <h1> ...]\n", argv[0]);
        return 2;
    }
    const uint32_t storage = (uint32_t)std::atoi(argv[1]);
    const int frames = std::atoi(argv[2]);
    const std::string dir = argv[3];
    const uint32_t n = (uint32_t)(argc - 4) / 2;
    const bool fp16 = storage == SVGF_STORAGE_FP16;
    const uint32_t f4 = fp16 ? SVGF_FMT_RGBA16F : SVGF_FMT_RGBA32F, b4 = fp16 ? 8 : 16;

    std::vector<std::unique_ptr<Agent>> agents;
    uint32_t max_w = 1, max_h = 1;
    uint64_t pixels = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t w = (uint32_t)std::atoi(argv[4 + 2 * i]), h = (uint32_t)std::atoi(argv[5 + 2 * i]);
        agents.push_back(std::make_unique<Agent>(w, h, f4, b4));
        max_w = w > max_w ? w : max_w;
        max_h = h > max_h ? h : max_h;
        pixels += (uint64_t)w * h;
    }

    // one context for every agent stream, sized for one frame of each per batch
    svgf_desc desc{};
    desc.mode = SVGF_MODE_CANONICAL;
    desc.storage = storage;
    check(svgf_default_params(desc.mode, storage, &desc.params), "svgf_default_params");
    desc.max_width = max_w;
    desc.max_height = max_h;
    desc.max_streams = n;
    desc.max_batch_frames = n;
    desc.max_batch_pixels = pixels;
    desc.device = 0;
    svgf_context* svgf = nullptr;
    check(svgf_create(&desc, &svgf), "svgf_create");
    hipStream_t stream = nullptr;
    check(hipStreamCreate(&stream), "hipStreamCreate");
    for (auto& a : agents) check(svgf_stream_open(svgf, a->w, a->h, &a->id), "svgf_stream_open");

    std::vector<svgf_request> req(n);
    for (int f = 0; f < frames; ++f) {
        const std::string fs = std::to_string(f);
        for (uint32_t i = 0; i < n; ++i) {
            Agent& a = *agents[i];
            const std::string p = dir + "/s" + std::to_string(i) + "_";
            a.radiance.upload(p + "radiance_" + fs + ".bin", stream);
            a.gbuffer[f & 1].upload(p + "gbuffer_" + fs + ".bin", stream);  // the other one is the kept G_prev
            a.motion.upload(p + "motion_" + fs + ".bin", stream);
            svgf_request& r = req[i];
            r = svgf_request{};
            r.stream_id = a.id;
            r.flags = SVGF_REQUEST_KEEP_GBUFFER;
            r.radiance = a.radiance.abi();
            r.gbuffer = a.gbuffer[f & 1].abi();
            r.motion = a.motion.abi();
            r.denoised = a.denoised.abi();
        }
        check(svgf_denoise_batch(svgf, req.data(), n, stream), "svgf_denoise_batch");
        uint64_t batch = 0;
        check(svgf_batch_last(svgf, &batch), "svgf_batch_last");
        check(svgf_batch_status(svgf, batch, 1), "svgf_batch_status");
        for (uint32_t i = 0; i < n; ++i) {
            Agent& a = *agents[i];
            std::vector<char> buf(a.denoised.row * a.h);
            check(hipMemcpy2D(buf.data(), a.denoised.row, a.denoised.ptr, a.denoised.pitch, a.denoised.row, a.h,
                              hipMemcpyDeviceToHost),
                  "hipMemcpy2D");
            write_file(dir + "/s" + std::to_string(i) + "_denoised_" + fs + ".bin", buf);
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        const Agent& a = *agents[i];
        svgf_stream_state st{};
        check(svgf_stream_state_get(svgf, a.id, &st), "svgf_stream_state_get");
        const size_t bm = st.moments.format == SVGF_FMT_RG32F ? 8 : (st.moments.format == SVGF_FMT_RG16F ? 4 : 8);
        const std::string p = dir + "/s" + std::to_string(i) + "_";
        download_state(st.history, (size_t)a.w * b4, a.h, p + "history.bin");
        download_state(st.moments, (size_t)a.w * bm, a.h, p + "moments.bin");
        check(svgf_stream_close(svgf, a.id), "svgf_stream_close");
    }
    check(svgf_destroy(svgf), "svgf_destroy");
    check(hipStreamDestroy(stream), "hipStreamDestroy");
    std::printf("svgf_batch_consumer (C++20): %d batches of %u streams, storage %u\n", frames, n, storage);
    return 0;
}
