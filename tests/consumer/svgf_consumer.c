/* svgf_consumer.c — the same host as svgf_consumer.cpp in plain C11 (gcc): one context created
 * at start-up (src/main.cpp:2288-2446), one svgf_denoise per frame in place of recordFrame's
 * SVGF block (main.cpp:907-947), the moments / denoised / history images chosen by current_frame
 * (main.cpp:2401-2446), flipped after each frame (main.cpp:3148). Same arguments and files as
 * svgf_consumer.cpp. */
#include <hip/hip_runtime_api.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "svgf.h"

typedef struct plane {
    void* ptr;
    size_t pitch, row;
    uint32_t format, h;
} plane;

static void die(const char* what, const char* why)
{
    fprintf(stderr, "svgf_consumer_c: %s: %s\n", what, why);
    exit(1);
}
static void check_s(svgf_status s, const char* what)
{
    if (s != SVGF_OK) die(what, svgf_status_str(s));
}
static void check_h(hipError_t e, const char* what)
{
    if (e != hipSuccess) die(what, hipGetErrorString(e));
}

static plane plane_make(uint32_t w, uint32_t h, uint32_t format, uint32_t bytes_px)
{
    plane p = {NULL, 0, (size_t)w * bytes_px, format, h};
    check_h(hipMallocPitch(&p.ptr, &p.pitch, p.row, h), "hipMallocPitch");
    check_h(hipMemset2D(p.ptr, p.pitch, 0, p.row, h), "hipMemset2D");
    return p;
}
static svgf_plane plane_abi(const plane* p)
{
    svgf_plane a = {p->ptr, (uint64_t)p->pitch, p->format, 0};
    return a;
}
static void plane_upload(plane* p, const char* dir, const char* name, int f)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%s_%d.bin", dir, name, f);
    size_t n = p->row * p->h;
    char* buf = malloc(n);
    FILE* fp = fopen(path, "rb");
    if (!buf || !fp || fread(buf, 1, n, fp) != n) die(path, "read");
    fclose(fp);
    check_h(hipMemcpy2D(p->ptr, p->pitch, buf, p->row, p->row, p->h, hipMemcpyHostToDevice), "hipMemcpy2D");
    free(buf);
}
static void plane_download(const plane* p, const char* path)
{
    size_t n = p->row * p->h;
    char* buf = malloc(n);
    if (!buf) die(path, "malloc");
    check_h(hipMemcpy2D(buf, p->row, p->ptr, p->pitch, p->row, p->h, hipMemcpyDeviceToHost), "hipMemcpy2D");
    FILE* fp = fopen(path, "wb");
    if (!fp || fwrite(buf, 1, n, fp) != n) die(path, "write");
    fclose(fp);
    free(buf);
}

int main(int argc, char** argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s <mode> <storage> <width> <height> <frames> <dir>\n", argv[0]);
        return 2;
    }
    const uint32_t mode = (uint32_t)atoi(argv[1]), storage = (uint32_t)atoi(argv[2]);
    const uint32_t w = (uint32_t)atoi(argv[3]), h = (uint32_t)atoi(argv[4]);
    const int frames = atoi(argv[5]);
    const char* dir = argv[6];
    const int canonical = mode == SVGF_MODE_CANONICAL, fp16 = storage == SVGF_STORAGE_FP16;

    svgf_desc desc;
    memset(&desc, 0, sizeof desc);
    desc.mode = mode;
    desc.storage = storage;
    check_s(svgf_default_params(mode, storage, &desc.params), "svgf_default_params");
    desc.max_width = w;
    desc.max_height = h;
    svgf_context* svgf = NULL;
    check_s(svgf_create(&desc, &svgf), "svgf_create");
    hipStream_t stream = NULL;
    check_h(hipStreamCreate(&stream), "hipStreamCreate");

    const uint32_t f4 = fp16 ? SVGF_FMT_RGBA16F : SVGF_FMT_RGBA32F, b4 = fp16 ? 8 : 16;
    const int wide = (desc.params.flags & SVGF_FLAG_WIDE_STATE) != 0;
    const uint32_t fm = wide ? SVGF_FMT_RGBA16F : (fp16 ? SVGF_FMT_RG16F : SVGF_FMT_RG32F), bm = wide ? 8 : (fp16 ? 4 : 8);
    plane radiance = plane_make(w, h, f4, b4), motion = plane_make(w, h, SVGF_FMT_RG32F, 8);
    plane denoised[2], moments[2], history[2], gbuffer[2];
    for (int i = 0; i < 2; ++i) {
        denoised[i] = plane_make(w, h, f4, b4);
        moments[i] = plane_make(w, h, fm, bm);
        history[i] = plane_make(w, h, f4, b4);
        gbuffer[i] = plane_make(w, h, SVGF_FMT_RGBA32F, 16);
    }

    uint32_t current_frame = 0;
    for (int f = 0; f < frames; ++f) {
        plane_upload(&radiance, dir, "radiance", f);
        if (canonical) {
            plane_upload(&gbuffer[current_frame], dir, "gbuffer", f);
            plane_upload(&motion, dir, "motion", f);
        }
        svgf_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.width = w;
        fr.height = h;
        fr.history_valid = f > 0;
        fr.radiance = plane_abi(&radiance);
        fr.moments_prev = plane_abi(&moments[1 - current_frame]);
        fr.moments_curr = plane_abi(&moments[current_frame]);
        fr.denoised = plane_abi(&denoised[current_frame]);
        if (canonical) {
            fr.gbuffer = plane_abi(&gbuffer[current_frame]);
            fr.gbuffer_prev = plane_abi(&gbuffer[1 - current_frame]);
            fr.motion = plane_abi(&motion);
            fr.history_prev = plane_abi(&history[1 - current_frame]);
            fr.history_curr = plane_abi(&history[current_frame]);
        }
        check_s(svgf_denoise(svgf, &fr, stream), "svgf_denoise");
        check_h(hipStreamSynchronize(stream), "hipStreamSynchronize");
        char path[1024];
        snprintf(path, sizeof path, "%s/denoised_%d.bin", dir, f);
        plane_download(&denoised[current_frame], path);
        if (f == frames - 1) {
            snprintf(path, sizeof path, "%s/moments.bin", dir);
            plane_download(&moments[current_frame], path);
            if (canonical) {
                snprintf(path, sizeof path, "%s/history.bin", dir);
                plane_download(&history[current_frame], path);
            }
        }
        current_frame ^= 1u;
    }
    check_s(svgf_destroy(svgf), "svgf_destroy");
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(denoised[i].ptr);
        (void)hipFree(moments[i].ptr);
        (void)hipFree(history[i].ptr);
        (void)hipFree(gbuffer[i].ptr);
    }
    (void)hipFree(radiance.ptr);
    (void)hipFree(motion.ptr);
    check_h(hipStreamDestroy(stream), "hipStreamDestroy");
    printf("svgf_consumer (C11): %d frames of %ux%u, mode %u, storage %u\n", frames, w, h, mode, storage);
    return 0;
}
