// svgf_consumer.cpp — a C++20 host of the SVGF C ABI (include/svgf.h), written the way TronGrid's
// render thread records the pass: the context is created once at start-up, in place of the SVGF
// pipeline and descriptor sets (src/main.cpp:2288-2446); each frame makes one svgf_denoise call
// in place of the four dispatches of recordFrame (main.cpp:907-947), reading moments_prev from
// image 1 - current_frame and writing moments_curr and denoised to image current_frame (the
// descriptor-set ping-pong of main.cpp:2401-2446); current_frame flips after the frame
// (main.cpp:3148). Planes are pitched hipMallocPitch buffers.
//
// usage: svgf_consumer <mode 0|1> <storage 0|1> <width> <height> <frames> <dir>
//   reads  <dir>/radiance_<f>.bin  (w h 4 values of the storage type, rows packed)
//          and, CANONICAL, <dir>/gbuffer_<f>.bin (w h 4 float32), <dir>/motion_<f>.bin (w h 2 float32)
//   writes <dir>/denoised_<f>.bin (w h 4 values of the storage type) and, after the last frame,
//          <dir>/moments.bin and (CANONICAL) <dir>/history.bin: that frame's state planes, rows packed
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "svgf.h"

namespace {

[[noreturn]] void fail(const char* what, int code)
{
    std::fprintf(stderr, "svgf_consumer: %s failed (%d)\n", what, code);
    std::exit(1);
}
void check(svgf_status s, const char* what)
{
    if (s != SVGF_OK) {
        std::fprintf(stderr, "svgf_consumer: %s: %s\n", what, svgf_status_str(s));
        std::exit(1);
    }
}
void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(what, (int)e);
}

// One pitched device plane.
struct Plane {
    void* ptr = nullptr;
    size_t pitch = 0;
    uint32_t format = 0;
    size_t row = 0;  // bytes of one packed row
    uint32_t h = 0;
    Plane(uint32_t w, uint32_t rows, uint32_t fmt, uint32_t bytes_px) : format(fmt), row((size_t)w * bytes_px), h(rows)
    {
        check(hipMallocPitch(&ptr, &pitch, row, h), "hipMallocPitch");
        check(hipMemset2D(ptr, pitch, 0, row, h), "hipMemset2D");
    }
    ~Plane() { (void)hipFree(ptr); }
    Plane(const Plane&) = delete;
    Plane& operator=(const Plane&) = delete;
    svgf_plane abi() const { return svgf_plane{ptr, (uint64_t)pitch, format, 0}; }
    void upload(const std::string& path)
    {
        std::vector<char> buf(row * h);
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f || std::fread(buf.data(), 1, buf.size(), f) != buf.size()) fail(path.c_str(), 0);
        std::fclose(f);
        check(hipMemcpy2D(ptr, pitch, buf.data(), row, row, h, hipMemcpyHostToDevice), "hipMemcpy2D");
    }
    void download(const std::string& path) const
    {
        std::vector<char> buf(row * h);
        check(hipMemcpy2D(buf.data(), row, ptr, pitch, row, h, hipMemcpyDeviceToHost), "hipMemcpy2D");
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f || std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) fail(path.c_str(), 0);
        std::fclose(f);
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s <mode> <storage> <width> <height> <frames> <dir>\n", argv[0]);
        return 2;
    }
    const uint32_t mode = (uint32_t)std::atoi(argv[1]), storage = (uint32_t)std::atoi(argv[2]);
    const uint32_t w = (uint32_t)std::atoi(argv[3]), h = (uint32_t)std::atoi(argv[4]);
    const int frames = std::atoi(argv[5]);
    const std::string dir = argv[6];
    const bool canonical = mode == SVGF_MODE_CANONICAL, fp16 = storage == SVGF_STORAGE_FP16;

    // start-up: the pipeline and descriptor sets of main.cpp:2288-2446
    svgf_desc desc{};
    desc.mode = mode;
    desc.storage = storage;
    check(svgf_default_params(mode, storage, &desc.params), "svgf_default_params");
    desc.max_width = w;
    desc.max_height = h;
    desc.device = 0;
    svgf_context* svgf = nullptr;
    check(svgf_create(&desc, &svgf), "svgf_create");
    hipStream_t stream = nullptr;
    check(hipStreamCreate(&stream), "hipStreamCreate");

    // the images of main.cpp:1333-1370 (two of each per-frame image: the ping-pong)
    const uint32_t f4 = fp16 ? SVGF_FMT_RGBA16F : SVGF_FMT_RGBA32F, b4 = fp16 ? 8 : 16;
    const bool wide = (desc.params.flags & SVGF_FLAG_WIDE_STATE) != 0;
    const uint32_t fm = wide ? SVGF_FMT_RGBA16F : (fp16 ? SVGF_FMT_RG16F : SVGF_FMT_RG32F), bm = wide ? 8 : (fp16 ? 4 : 8);
    Plane radiance(w, h, f4, b4);
    Plane denoised[2] = {Plane(w, h, f4, b4), Plane(w, h, f4, b4)};
    Plane moments[2] = {Plane(w, h, fm, bm), Plane(w, h, fm, bm)};
    Plane history[2] = {Plane(w, h, f4, b4), Plane(w, h, f4, b4)};
    Plane gbuffer[2] = {Plane(w, h, SVGF_FMT_RGBA32F, 16), Plane(w, h, SVGF_FMT_RGBA32F, 16)};
    Plane motion(w, h, SVGF_FMT_RG32F, 8);

    uint32_t current_frame = 0;
    for (int f = 0; f < frames; ++f) {
        const std::string n = std::to_string(f);
        radiance.upload(dir + "/radiance_" + n + ".bin");
        if (canonical) {
            gbuffer[current_frame].upload(dir + "/gbuffer_" + n + ".bin");
            motion.upload(dir + "/motion_" + n + ".bin");
        }
        // recordFrame's SVGF block: one call, enqueued on the frame's stream
        svgf_frame fr{};
        fr.width = w;
        fr.height = h;
        fr.history_valid = f > 0 ? 1u : 0u;  // zero-cleared images before the first frame
        fr.radiance = radiance.abi();
        fr.moments_prev = moments[1 - current_frame].abi();
        fr.moments_curr = moments[current_frame].abi();
        fr.denoised = denoised[current_frame].abi();
        if (canonical) {
            fr.gbuffer = gbuffer[current_frame].abi();
            fr.gbuffer_prev = gbuffer[1 - current_frame].abi();
            fr.motion = motion.abi();
            fr.history_prev = history[1 - current_frame].abi();
            fr.history_curr = history[current_frame].abi();
        }
        check(svgf_denoise(svgf, &fr, stream), "svgf_denoise");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        denoised[current_frame].download(dir + "/denoised_" + n + ".bin");
        if (f == frames - 1) {
            moments[current_frame].download(dir + "/moments.bin");
            if (canonical) history[current_frame].download(dir + "/history.bin");
        }
        current_frame ^= 1u;  // main.cpp:3148
    }
    check(svgf_destroy(svgf), "svgf_destroy");
    check(hipStreamDestroy(stream), "hipStreamDestroy");
    std::printf("svgf_consumer (C++20): %d frames of %ux%u, mode %u, storage %u\n", frames, w, h, mode, storage);
    return 0;
}
