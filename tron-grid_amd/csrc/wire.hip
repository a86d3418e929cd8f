// wire.hip — agent frames on the wire: denoised HDR colour + depth -> RGB888 + normalised depth,
// and the pinned host ring an agent reads them from.
//
// Layout (docs/AI_INTERFACE.md:346-383, TgVisionData): per frame, RGB888 row-major from the top
// row down with no row padding (3 w h bytes), and float32 depth in [0, 1] (4 w h bytes). Colour
// goes through TronGrid's display tonemap (src/postprocess.slang:47-103, :153-160): the fitted
// ACES of Hill / MJP BakingLab (sRGB -> AP1 matrix, RRT+ODT rational fit per channel,
// AP1 -> sRGB matrix, saturate) and the IEC 61966-2-1 sRGB transfer; the display path's
// screen-space effects (bloom, chromatic aberration, grade, vignette) are for a human viewer and
// are not applied. The 8-bit code of a channel is floor(255 srgb(x) + 1/2) with srgb in double,
// evaluated exactly as the count of 255 float decision thresholds not above x (the smallest
// float reaching each half code; computed once on the host), so the bytes do not depend on a
// pow() implementation. Depth: (z - near) / (far - near) clamped to [0, 1]; z not > 0 or not
// finite (no geometry) reads 1.
//
// Readback (docs/AI_INTERFACE.md:385-400): a ring of pinned host slots. Each push copies one
// device region into the next slot with hipMemcpyAsync and records an event; a slot is reused
// `slots` pushes later, so with 3 slots an agent reads frame n while frames n + 1 and n + 2 are
// being produced, as the reference's triple-buffered staging does.
#include "../../include/svgf.h"
#include "svgf_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace svgf {
namespace {

constexpr int WIRE_BLOCK = 256;
constexpr int WIRE_PX_PER_THREAD = 4;
constexpr int WIRE_PX_PER_BLOCK = WIRE_BLOCK * WIRE_PX_PER_THREAD;  // 1024
constexpr uint32_t WIRE_MAX_FRAMES = 4096;
constexpr int BUF_WORD3 = 0x00020000;

struct WireDesc {
    uint32_t w, h;
    uint32_t colour_fp16;
    uint32_t reserved;
    const char* colour;
    uint64_t colour_pitch;
    const char* gbuffer;
    uint64_t gbuffer_pitch;
    char* rgb;
    char* depth;
};

struct WireArgs {
    const uint32_t* starts;  // n + 1 block prefix
    const WireDesc* frames;
    const float* thresholds;  // 255 ascending
    uint32_t n;
    float exposure, near_z, far_z;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, BUF_WORD3);
}

// sRGB -> AP1 and AP1 -> sRGB, row-major (postprocess.slang:49-60)
__constant__ float ACES_IN[9] = {0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f,
                                 0.01566f, 0.02840f, 0.13383f, 0.83777f};
__constant__ float ACES_OUT[9] = {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f,
                                  -0.00605f, -0.00327f, -0.07276f, 1.07602f};

__device__ __forceinline__ void mat3(const float* m, const float v[3], float o[3])
{
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = __builtin_fmaf(m[3 * r], v[0], __builtin_fmaf(m[3 * r + 1], v[1], m[3 * r + 2] * v[2]));
}

// Code of x in [0, 1]: the count of thresholds <= x, with t[k] the k-th threshold (k = 1..255)
// and t[0] = -inf, t[256] = +inf in LDS. The sRGB curve in float (hardware log2 / exp2) puts the
// estimate e = floor(255 srgb(x) + 1/2) within one code of the count: its error, ~1e-5 of the
// curve's range, is far below one code step (1/255). One comparison on each side of e then gives
// the count exactly: two independent LDS reads instead of a dependent 8-step search.
__device__ __forceinline__ uint32_t code_of(const float* t, float x)
{
    const float s = x <= 0.0031308f ? x * 12.92f
                                    : __builtin_fmaf(1.055f, __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 2.4f)),
                                                     -0.055f);
    const int e = min(max((int)__builtin_fmaf(s, 255.0f, 0.5f), 0), 255);
    const float below = t[e], above = t[e + 1];  // thresholds e and e + 1
    return (uint32_t)(above <= x ? e + 1 : (below > x ? e - 1 : e));
}

__device__ __forceinline__ uint32_t encode_rgb(const float* t, float r, float g, float b, float exposure)
{
#pragma clang fp contract(off)
    float c[3] = {r * exposure, g * exposure, b * exposure};
    float a[3];
    mat3(ACES_IN, c, a);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float v = a[i];
        const float num = __builtin_fmaf(v, v + 0.0245786f, -0.000090537f);
        const float den = __builtin_fmaf(v, __builtin_fmaf(0.983729f, v, 0.4329510f), 0.238081f);
        a[i] = num / den;  // correctly rounded division
    }
    mat3(ACES_OUT, a, c);
    uint32_t out = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float s = c[i] > 0.0f ? (c[i] < 1.0f ? c[i] : 1.0f) : 0.0f;
        out |= code_of(t, s) << (8 * i);
    }
    return out;  // bytes R, G, B in the low 24 bits
}

__device__ __forceinline__ float encode_depth(float z, float near_z, float far_z)
{
#pragma clang fp contract(off)
    if (!(z > 0.0f) || !__builtin_isfinite(z)) return 1.0f;
    const float d = (z - near_z) / (far_z - near_z);
    return d > 0.0f ? (d < 1.0f ? d : 1.0f) : 0.0f;
}

// One launch over the frames of a request: block b covers pixels [1024 k, 1024 k + 1024) of the
// frame whose block range holds b (binary search over the prefix, uniform per block); thread t
// encodes 4 consecutive pixels of the frame's row-major pixel order and stores their 12 RGB
// bytes as 3 dwords and their 4 depths as one 16-byte store.
template <bool FP16>
__global__ __launch_bounds__(WIRE_BLOCK) void k_wire(WireArgs a)
{
    __shared__ float sT[257];
    const int tid = (int)threadIdx.x;
    if (tid < 255) sT[tid + 1] = a.thresholds[tid];
    if (tid == 255) {
        sT[0] = -__builtin_inff();
        sT[256] = __builtin_inff();
    }
    const uint32_t b = blockIdx.x;
    if (b >= a.starts[a.n]) return;  // uniform per block (before any barrier)
    __syncthreads();
    const uint32_t fi = frame_of(a.starts, a.n, b);
    WireDesc f;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.frames + fi);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&f);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(WireDesc) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    }
    const uint64_t npx = (uint64_t)f.w * f.h;
    const uint64_t p0 = (uint64_t)(b - a.starts[fi]) * WIRE_PX_PER_BLOCK + (uint64_t)tid * WIRE_PX_PER_THREAD;
    if (p0 >= npx) return;
    const __amdgpu_buffer_rsrc_t cr = rsrc(f.colour, f.colour_pitch * f.h);
    const __amdgpu_buffer_rsrc_t gr = rsrc(f.gbuffer, f.gbuffer_pitch * f.h);
    uint32_t y = (uint32_t)(p0 / f.w), x = (uint32_t)(p0 - (uint64_t)y * f.w);
    uint32_t rgb[4] = {0, 0, 0, 0};
    float dep[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    const int cnt = (int)min<uint64_t>(WIRE_PX_PER_THREAD, npx - p0);
    // The colour and depth of all four pixels are loaded before the first is encoded (one
    // memory round trip). A pixel past the frame's last (the tail of the last block) lies past
    // the planes, so the buffer resources answer its loads with 0; it is neither encoded nor
    // stored.
    float cr_[WIRE_PX_PER_THREAD], cg_[WIRE_PX_PER_THREAD], cb_[WIRE_PX_PER_THREAD], z_[WIRE_PX_PER_THREAD];
#pragma unroll
    for (int j = 0; j < WIRE_PX_PER_THREAD; ++j) {
        if constexpr (FP16) {
            typedef _Float16 h4 __attribute__((ext_vector_type(4)));
            const uint32_t off = y * (uint32_t)f.colour_pitch + x * 8;
            const h4 v = __builtin_bit_cast(h4, __builtin_amdgcn_raw_buffer_load_b64(cr, off, 0, 0));
            cr_[j] = (float)v.x;
            cg_[j] = (float)v.y;
            cb_[j] = (float)v.z;
        } else {
            const uint32_t off = y * (uint32_t)f.colour_pitch + x * 16;
            typedef float f4 __attribute__((ext_vector_type(4)));
            const f4 v = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(cr, off, 0, 0));
            cr_[j] = v.x;
            cg_[j] = v.y;
            cb_[j] = v.z;
        }
        z_[j] = __builtin_bit_cast(
            float, __builtin_amdgcn_raw_buffer_load_b32(gr, y * (uint32_t)f.gbuffer_pitch + x * 16 + 12, 0, 0));
        if (++x == f.w) {
            x = 0;
            ++y;
        }
    }
#pragma unroll
    for (int j = 0; j < WIRE_PX_PER_THREAD; ++j) {
        if (j < cnt) {
            rgb[j] = encode_rgb(sT, cr_[j], cg_[j], cb_[j], a.exposure);
            dep[j] = encode_depth(z_[j], a.near_z, a.far_z);
        }
    }
    const __amdgpu_buffer_rsrc_t rr = rsrc(f.rgb, npx * 3);
    const __amdgpu_buffer_rsrc_t dr = rsrc(f.depth, npx * 4);
    const uint32_t rgb_off = (uint32_t)(p0 * 3), dep_off = (uint32_t)(p0 * 4);
    if (cnt == WIRE_PX_PER_THREAD) {
        typedef uint32_t u3 __attribute__((ext_vector_type(3)));
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const u3 packed = {rgb[0] | (rgb[1] << 24), (rgb[1] >> 8) | (rgb[2] << 16), (rgb[2] >> 16) | (rgb[3] << 8)};
        __builtin_amdgcn_raw_buffer_store_b96(__builtin_bit_cast(u3, packed), rr, rgb_off, 0, 0);
        const u4 d = {__builtin_bit_cast(uint32_t, dep[0]), __builtin_bit_cast(uint32_t, dep[1]),
                      __builtin_bit_cast(uint32_t, dep[2]), __builtin_bit_cast(uint32_t, dep[3])};
        __builtin_amdgcn_raw_buffer_store_b128(d, dr, dep_off, 0, 0);
    } else {
        for (int j = 0; j < cnt; ++j) {
            for (int c = 0; c < 3; ++c)
                __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(rgb[j] >> (8 * c)), rr, rgb_off + 3 * j + c, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, dep[j]), dr, dep_off + 4 * j, 0, 0);
        }
    }
}

double srgb_encode(double x) { return x <= 0.0031308 ? x * 12.92 : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055; }
double srgb_decode(double v) { return v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4); }

// smallest float whose sRGB value (in double) reaches (k - 1/2) / 255, k = 1..255
void make_thresholds(float* t)
{
    for (int k = 1; k <= 255; ++k) {
        const double target = (k - 0.5) / 255.0;
        float x = (float)srgb_decode(target);
        while (x > 0.0f && srgb_encode((double)std::nextafter(x, 0.0f)) >= target) x = std::nextafter(x, 0.0f);
        while (srgb_encode((double)x) < target) x = std::nextafter(x, 2.0f);
        t[k - 1] = x;
    }
}

}  // namespace
}  // namespace svgf

using namespace svgf;

struct svgf_wire {
    int device = 0;
    uint32_t max_frames = 0;
    size_t frames_off = 0, table_bytes = 0;
    float* thresholds = nullptr;  // device, 255
    TableRing ring;                // per call: starts (n + 1), then the descriptors
};

struct svgf_readback {
    int device = 0;
    uint64_t slot_bytes = 0;
    uint32_t slots = 0;
    std::vector<void*> host;        // pinned
    std::vector<hipEvent_t> done;
    std::vector<uint64_t> seq_of;   // sequence number held by each slot (0 = none)
    std::vector<uint64_t> bytes_of;
    uint64_t next_seq = 1;
};

namespace {

void free_wire(svgf_wire* w)
{
    w->ring.destroy();
    if (w->thresholds) (void)hipFree(w->thresholds);
    delete w;
}

void free_readback(svgf_readback* r)
{
    for (size_t i = 0; i < r->done.size(); ++i) {
        if (r->done[i]) {
            (void)hipEventSynchronize(r->done[i]);
            (void)hipEventDestroy(r->done[i]);
        }
    }
    for (void* p : r->host)
        if (p) (void)hipHostFree(p);
    delete r;
}

}  // namespace

extern "C" {

svgf_status svgf_wire_default_params(svgf_wire_params* out)
{
    if (!out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = svgf_wire_params{};
    out->exposure = 1.0f;
    out->depth_near = 0.1f;  // camera planes, src/camera.hpp:99-100
    out->depth_far = 100.0f;
    return SVGF_OK;
}

svgf_status svgf_wire_create(int32_t device, uint32_t max_frames, svgf_wire** out)
{
    if (!out || max_frames < 1 || max_frames > WIRE_MAX_FRAMES) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SVGF_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return SVGF_ERR_NO_DEVICE;
    DeviceGuard guard(device);
    svgf_wire* w = new (std::nothrow) svgf_wire();
    if (!w) return SVGF_ERR_OUT_OF_MEMORY;
    w->device = device;
    w->max_frames = max_frames;
    w->frames_off = (((size_t)max_frames + 1) * 4 + 255) / 256 * 256;
    w->table_bytes = w->frames_off + sizeof(WireDesc) * max_frames;
    float t[255];
    make_thresholds(t);
    bool ok = hipMalloc(reinterpret_cast<void**>(&w->thresholds), sizeof(t)) == hipSuccess &&
              hipMemcpy(w->thresholds, t, sizeof(t), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && w->ring.create(w->table_bytes);
    if (!ok) {
        free_wire(w);
        return SVGF_ERR_OUT_OF_MEMORY;
    }
    *out = w;
    return SVGF_OK;
}

svgf_status svgf_wire_destroy(svgf_wire* w)
{
    if (!w) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(w->device);
    free_wire(w);
    return SVGF_OK;
}

svgf_status svgf_wire_encode(svgf_wire* w, const svgf_wire_frame* frames, uint32_t n, const svgf_wire_params* p,
                             void* hip_stream)
{
    if (!w || (!frames && n) || !p) return SVGF_ERR_INVALID_ARGUMENT;
    if (n == 0) return SVGF_OK;
    if (n > w->max_frames) return SVGF_ERR_CAPACITY;
    if (!(p->depth_far > p->depth_near) || !std::isfinite(p->depth_near) || !std::isfinite(p->depth_far) ||
        !(p->exposure >= 0.0f) || !std::isfinite(p->exposure))
        return SVGF_ERR_INVALID_ARGUMENT;
    int fp16 = -1;
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const svgf_wire_frame& f = frames[i];
        if (f.width < 1 || f.height < 1 || !f.colour.ptr || !f.gbuffer.ptr || !f.rgb || !f.depth) return SVGF_ERR_INVALID_ARGUMENT;
        const int h16 = f.colour.format == SVGF_FMT_RGBA16F ? 1 : f.colour.format == SVGF_FMT_RGBA32F ? 0 : -1;
        if (h16 < 0 || (fp16 >= 0 && h16 != fp16)) return SVGF_ERR_INVALID_ARGUMENT;  // one colour format per call
        fp16 = h16;
        const uint64_t cb = h16 ? 8 : 16;
        if (f.gbuffer.format != SVGF_FMT_RGBA32F || f.colour.pitch_bytes < f.width * cb || f.colour.pitch_bytes % cb ||
            f.gbuffer.pitch_bytes < (uint64_t)f.width * 16 || f.gbuffer.pitch_bytes % 16 || !aligned_to(f.colour.ptr, cb) ||
            !aligned_to(f.gbuffer.ptr, 16) || !aligned_to(f.rgb, 4) || !aligned_to(f.depth, 16))
            return SVGF_ERR_INVALID_ARGUMENT;
        const uint64_t px = (uint64_t)f.width * f.height;
        if (f.colour.pitch_bytes * f.height >= (1ull << 31) || f.gbuffer.pitch_bytes * f.height >= (1ull << 31) ||
            px * 4 >= (1ull << 31))
            return SVGF_ERR_CAPACITY;
        blocks += (px + WIRE_PX_PER_BLOCK - 1) / WIRE_PX_PER_BLOCK;
    }
    if (blocks >= (1ull << 31)) return SVGF_ERR_CAPACITY;
    DeviceGuard guard(w->device);
    TableRing::Slot* slot = w->ring.acquire();
    if (!slot) return SVGF_ERR_HIP;
    TableRing::Slot& s = *slot;
    uint32_t* starts = static_cast<uint32_t*>(s.staging);
    WireDesc* descs = reinterpret_cast<WireDesc*>(static_cast<char*>(s.staging) + w->frames_off);
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const svgf_wire_frame& f = frames[i];
        starts[i] = acc;
        acc += (uint32_t)(((uint64_t)f.width * f.height + WIRE_PX_PER_BLOCK - 1) / WIRE_PX_PER_BLOCK);
        WireDesc d{};
        d.w = f.width;
        d.h = f.height;
        d.colour_fp16 = (uint32_t)fp16;
        d.colour = static_cast<const char*>(f.colour.ptr);
        d.colour_pitch = f.colour.pitch_bytes;
        d.gbuffer = static_cast<const char*>(f.gbuffer.ptr);
        d.gbuffer_pitch = f.gbuffer.pitch_bytes;
        d.rgb = static_cast<char*>(f.rgb);
        d.depth = static_cast<char*>(f.depth);
        descs[i] = d;
    }
    starts[n] = acc;
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    const size_t copy = w->frames_off + sizeof(WireDesc) * n;
    if (!TableRing::upload(s, copy, stream)) return SVGF_ERR_HIP;
    WireArgs a{};
    a.starts = static_cast<const uint32_t*>(s.table);
    a.frames = reinterpret_cast<const WireDesc*>(static_cast<const char*>(s.table) + w->frames_off);
    a.thresholds = w->thresholds;
    a.n = n;
    a.exposure = p->exposure;
    a.near_z = p->depth_near;
    a.far_z = p->depth_far;
    if (fp16) hipLaunchKernelGGL(k_wire<true>, dim3(acc), dim3(WIRE_BLOCK), 0, stream, a);
    else hipLaunchKernelGGL(k_wire<false>, dim3(acc), dim3(WIRE_BLOCK), 0, stream, a);
    if (hipGetLastError() != hipSuccess) return SVGF_ERR_HIP;
    if (!TableRing::release(s, stream)) return SVGF_ERR_HIP;
    return SVGF_OK;
}

svgf_status svgf_readback_create(int32_t device, uint64_t slot_bytes, uint32_t slots, svgf_readback** out)
{
    if (!out || slot_bytes < 1 || slots < 2 || slots > 16) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SVGF_ERR_NO_DEVICE;
    DeviceGuard guard(device);
    svgf_readback* r = new (std::nothrow) svgf_readback();
    if (!r) return SVGF_ERR_OUT_OF_MEMORY;
    r->device = device;
    r->slot_bytes = slot_bytes;
    r->slots = slots;
    r->host.assign(slots, nullptr);
    r->done.assign(slots, nullptr);
    r->seq_of.assign(slots, 0);
    r->bytes_of.assign(slots, 0);
    for (uint32_t i = 0; i < slots; ++i) {
        if (hipHostMalloc(&r->host[i], slot_bytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&r->done[i], hipEventDisableTiming) != hipSuccess) {
            free_readback(r);
            return SVGF_ERR_OUT_OF_MEMORY;
        }
    }
    *out = r;
    return SVGF_OK;
}

svgf_status svgf_readback_destroy(svgf_readback* r)
{
    if (!r) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(r->device);
    free_readback(r);
    return SVGF_OK;
}

svgf_status svgf_readback_push(svgf_readback* r, const void* device_src, uint64_t bytes, void* hip_stream, uint64_t* seq)
{
    if (!r || !device_src || !seq || bytes < 1) return SVGF_ERR_INVALID_ARGUMENT;
    if (bytes > r->slot_bytes) return SVGF_ERR_CAPACITY;
    DeviceGuard guard(r->device);
    const uint64_t q = r->next_seq;
    const uint32_t i = (uint32_t)(q % r->slots);
    if (r->seq_of[i] && hipEventSynchronize(r->done[i]) != hipSuccess) return SVGF_ERR_HIP;
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    if (hipMemcpyAsync(r->host[i], device_src, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess) return SVGF_ERR_HIP;
    if (hipEventRecord(r->done[i], stream) != hipSuccess) return SVGF_ERR_HIP;
    r->seq_of[i] = q;
    r->bytes_of[i] = bytes;
    r->next_seq = q + 1;
    *seq = q;
    return SVGF_OK;
}

svgf_status svgf_readback_view(svgf_readback* r, uint64_t seq, int32_t wait, const void** host, uint64_t* bytes)
{
    if (!r || !host || !bytes || seq == 0) return SVGF_ERR_INVALID_ARGUMENT;
    const uint32_t i = (uint32_t)(seq % r->slots);
    if (r->seq_of[i] != seq) return SVGF_ERR_INVALID_ARGUMENT;  // never pushed, or already overwritten
    DeviceGuard guard(r->device);
    if (wait) {
        if (hipEventSynchronize(r->done[i]) != hipSuccess) return SVGF_ERR_HIP;
    } else {
        const hipError_t q = hipEventQuery(r->done[i]);
        if (q == hipErrorNotReady) return SVGF_ERR_NOT_READY;
        if (q != hipSuccess) return SVGF_ERR_HIP;
    }
    *host = r->host[i];
    *bytes = r->bytes_of[i];
    return SVGF_OK;
}

}  // extern "C"
