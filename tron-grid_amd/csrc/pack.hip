// pack.hip — the producer side of the denoiser input: radiance, G-buffer and motion planes
// built from the mesh pass's per-pixel terms in one pass that writes every pixel (the
// per-frame zero-clear of the radiance plane folded in), for many frames in one launch.
//
// TronGrid clears its SVGF input image every frame and lets the mesh fragment shader write it
// where opaque geometry covers a pixel (src/main.cpp:595-646, src/mesh.slang:995-998):
//   covered:   radiance = ((indirect + direct_diffuse) * ao + direct_specular, 1)
//   otherwise: radiance = 0
// Here one kernel reads the terms and writes all three planes the denoiser reads:
//   gbuffer = (unpackOctahedralNormal(normal_oct), (VP (p, 1)).w)   (mesh.slang:341-352)
//   motion  = previous position - pixel centre, previous position from clip = PVP (p, 1):
//             (clip.xy / max(clip.w, 1e-6) (0.5, -0.5) + 0.5) (W, H)   (mesh.slang:589-591)
//             (-1e6, -1e6) for clip.w <= 0, uncovered pixels, or |motion| > 1e6 px
// Uncovered pixels get radiance 0 and gbuffer 0 and do not read the lighting planes.
//
// Arithmetic: float, one rounding per operation (contraction off), matrix rows summed left to
// right, correctly rounded division and square root. Launch shape as the composite: 256
// threads, 1024 consecutive pixels per block, thread t takes pixels t + 256 j.
#include "../../include/svgf.h"
#include "svgf_internal.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>

namespace svgf {
namespace {

constexpr int PK_BLOCK = 256;
constexpr int PK_PX_PER_THREAD = 4;
constexpr int PK_PX_PER_BLOCK = PK_BLOCK * PK_PX_PER_THREAD;  // 1024
constexpr uint32_t PK_MAX_FRAMES = 4096;
constexpr int BUF_WORD3 = 0x00020000;
constexpr float NO_MOTION = -1.0e6f;

enum : int { Q_POS = 0, Q_NRM, Q_IND, Q_DIF, Q_SPC, Q_RAD, Q_G, Q_MV, Q_COUNT };

struct PackDesc {
    uint32_t w, h;
    uint32_t half;  // bit k: plane k is RGBA16F (lighting planes and radiance only)
    uint32_t reserved;
    float vp[16];
    float pvp[16];
    const char* plane[Q_COUNT];
    uint32_t pitch[Q_COUNT];
};
static_assert(sizeof(PackDesc) % 8 == 0, "descriptor array alignment");

struct PackArgs {
    const uint32_t* starts;
    const PackDesc* frames;
    uint32_t n;
};

constexpr uint32_t PK_INLINE = 10;
struct PackInlineArgs {
    uint32_t starts[PK_INLINE + 1];
    uint32_t n;
    PackDesc frames[PK_INLINE];
};
static_assert(sizeof(PackInlineArgs) <= 4096 - 64, "kernel argument segment");

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, BUF_WORD3);
}

__device__ __forceinline__ f4 load4(__amdgpu_buffer_rsrc_t r, uint32_t off_row, uint32_t x, bool half)
{
    if (half) {
        const h4 v = __builtin_bit_cast(h4, __builtin_amdgcn_raw_buffer_load_b64(r, off_row + x * 8, 0, 0));
        return f4{(float)v.x, (float)v.y, (float)v.z, (float)v.w};
    }
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, off_row + x * 16, 0, 0));
}

__device__ __forceinline__ float row4(const float* m, int r, float x, float y, float z)
{
#pragma clang fp contract(off)
    return ((m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z) + m[4 * r + 3];
}

// mesh.slang:341-352, normalised with a correctly rounded sqrt and division
__device__ __forceinline__ void unpack_octahedral(uint32_t packed, float& x, float& y, float& z)
{
#pragma clang fp contract(off)
    const float qx = (float)(packed & 0xFFFFu) / 65535.0f;
    const float qy = (float)((packed >> 16) & 0xFFFFu) / 65535.0f;
    const float px = qx * 2.0f - 1.0f;
    const float py = qy * 2.0f - 1.0f;
    float nx = px, ny = py;
    const float nz = (1.0f - __builtin_fabsf(px)) - __builtin_fabsf(py);
    if (nz < 0.0f) {
        nx = (1.0f - __builtin_fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        ny = (1.0f - __builtin_fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    }
    const float inv = 1.0f / __builtin_sqrtf((nx * nx + ny * ny) + nz * nz);
    x = nx * inv;
    y = ny * inv;
    z = nz * inv;
}

// Planes of a frame that may be RGBA16F: the three lighting planes and the radiance.
constexpr uint32_t PK_HALF_PLANES = (1u << Q_IND) | (1u << Q_DIF) | (1u << Q_SPC) | (1u << Q_RAD);
// Plane formats as the block body sees them: all four RGBA16F, all RGBA32F (compile-time, so
// the lighting loads of all four pixels issue before the first conversion), or mixed (per
// plane at run time: each lighting load then waits in a branch of its own).
enum Formats : int { FM_HALF = 0, FM_FLOAT = 1, FM_MIXED = 2 };

template <int FM>
__device__ __forceinline__ bool half_of(uint32_t mask, int k)
{
    if constexpr (FM == FM_HALF) return true;
    else if constexpr (FM == FM_FLOAT) return false;
    else return (mask >> k) & 1u;
}

// one pixel's raw words (binary16: 8 bytes, binary32: 16 bytes) and their rgba as float
template <bool H>
__device__ __forceinline__ std::conditional_t<H, u2, u4> load_raw(__amdgpu_buffer_rsrc_t r, uint32_t off_row, uint32_t x)
{
    if constexpr (H) return __builtin_amdgcn_raw_buffer_load_b64(r, off_row + x * 8, 0, 0);
    else return __builtin_amdgcn_raw_buffer_load_b128(r, off_row + x * 16, 0, 0);
}
template <bool H>
__device__ __forceinline__ f4 raw_f4(std::conditional_t<H, u2, u4> v)
{
    if constexpr (H) {
        const h4 q = __builtin_bit_cast(h4, v);
        return f4{(float)q.x, (float)q.y, (float)q.z, (float)q.w};
    } else {
        return __builtin_bit_cast(f4, v);
    }
}

// The pack of a block's pixels: position and normal of all four pixels of the thread, then the
// lighting terms of the covered ones, then the arithmetic and the stores.
template <int FM>
__device__ __forceinline__ void pack_pixels(const PackDesc& f, const __amdgpu_buffer_rsrc_t* r, uint32_t p0)
{
#pragma clang fp contract(off)
    const bool hi = half_of<FM>(f.half, Q_IND), hd = half_of<FM>(f.half, Q_DIF), hs = half_of<FM>(f.half, Q_SPC);
    const bool hr = half_of<FM>(f.half, Q_RAD);
    const uint32_t npx = f.w * f.h;

    uint32_t xs[PK_PX_PER_THREAD], ys[PK_PX_PER_THREAD], oct[PK_PX_PER_THREAD];
    f4 pos[PK_PX_PER_THREAD];
    // the four pixels' coordinates first, then their loads back to back (otherwise the
    // optimiser reads one pixel's coverage while the next pixel's address is still being
    // computed, and that read waits for the loads already issued)
#pragma unroll
    for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
        const uint32_t p = min(p0 + j * PK_BLOCK, npx - 1);  // clamped tail lanes load a valid pixel
        ys[j] = p / f.w;
        xs[j] = p - ys[j] * f.w;
    }
#pragma unroll
    for (int j = 0; j < PK_PX_PER_THREAD; ++j) asm volatile("" ::"v"(ys[j]), "v"(xs[j]));
#pragma unroll
    for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
        pos[j] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r[Q_POS], ys[j] * f.pitch[Q_POS] + xs[j] * 16, 0, 0));
        oct[j] = __builtin_amdgcn_raw_buffer_load_b32(r[Q_NRM], ys[j] * f.pitch[Q_NRM] + xs[j] * 4, 0, 0);
    }
    f4 ind[PK_PX_PER_THREAD], dif[PK_PX_PER_THREAD], spc[PK_PX_PER_THREAD];
    if constexpr (FM == FM_MIXED) {
#pragma unroll
        for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
            ind[j] = dif[j] = spc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            if (pos[j].w > 0.0f) {
                ind[j] = load4(r[Q_IND], ys[j] * f.pitch[Q_IND], xs[j], hi);
                dif[j] = load4(r[Q_DIF], ys[j] * f.pitch[Q_DIF], xs[j], hd);
                spc[j] = load4(r[Q_SPC], ys[j] * f.pitch[Q_SPC], xs[j], hs);
            }
        }
    } else {
        // No branch: an uncovered pixel's lighting loads take an offset past the plane, which
        // the buffer resource answers with 0 and no memory access (the ABI keeps planes under
        // 2 GiB). Straight-line loads of all four pixels are all in flight before the first
        // conversion, where a load in a branch would be waited for at the branch's end.
        constexpr bool H = FM == FM_HALF;
        using Raw = std::conditional_t<H, u2, u4>;
        constexpr uint32_t OOB = 0x80000000u;
        // every pixel's coverage is read before the first lighting load is issued: a load
        // issued in between would have to complete too (loads retire in order)
        uint32_t cy[PK_PX_PER_THREAD], cx[PK_PX_PER_THREAD];
#pragma unroll
        for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
            const bool cov = pos[j].w > 0.0f;
            cy[j] = cov ? ys[j] : OOB;  // row index: the planes' row offsets exceed 2^31 past it
            cx[j] = cov ? xs[j] : 0u;
        }
#pragma unroll
        for (int j = 0; j < PK_PX_PER_THREAD; ++j) asm volatile("" ::"v"(cy[j]), "v"(cx[j]));
        Raw ri[PK_PX_PER_THREAD], rd[PK_PX_PER_THREAD], rs[PK_PX_PER_THREAD];
#pragma unroll
        for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
            ri[j] = load_raw<H>(r[Q_IND], cy[j] == OOB ? OOB : cy[j] * f.pitch[Q_IND], cx[j]);
            rd[j] = load_raw<H>(r[Q_DIF], cy[j] == OOB ? OOB : cy[j] * f.pitch[Q_DIF], cx[j]);
            rs[j] = load_raw<H>(r[Q_SPC], cy[j] == OOB ? OOB : cy[j] * f.pitch[Q_SPC], cx[j]);
        }
#pragma unroll
        for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
            ind[j] = raw_f4<H>(ri[j]);
            dif[j] = raw_f4<H>(rd[j]);
            spc[j] = raw_f4<H>(rs[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < PK_PX_PER_THREAD; ++j) {
        if (p0 + j * PK_BLOCK >= npx) break;
        const uint32_t x = xs[j], y = ys[j];
        f4 rad = {0.0f, 0.0f, 0.0f, 0.0f}, g = {0.0f, 0.0f, 0.0f, 0.0f};
        f2 mv = {NO_MOTION, NO_MOTION};
        if (pos[j].w > 0.0f) {
            rad = f4{(ind[j].x + dif[j].x) * dif[j].w + spc[j].x, (ind[j].y + dif[j].y) * dif[j].w + spc[j].y,
                     (ind[j].z + dif[j].z) * dif[j].w + spc[j].z, 1.0f};
            float ox, oy, oz;
            unpack_octahedral(oct[j], ox, oy, oz);
            g.x = ox;
            g.y = oy;
            g.z = oz;
            const float px = pos[j].x, py = pos[j].y, pz = pos[j].z;
            g.w = row4(f.vp, 3, px, py, pz);
            const float cw = row4(f.pvp, 3, px, py, pz);
            if (cw > 0.0f) {
                const float d = cw > 1e-6f ? cw : 1e-6f;
                const float nx = row4(f.pvp, 0, px, py, pz) / d;
                const float ny = row4(f.pvp, 1, px, py, pz) / d;
                const float mx = (nx * 0.5f + 0.5f) * (float)f.w - ((float)x + 0.5f);
                const float my = (ny * -0.5f + 0.5f) * (float)f.h - ((float)y + 0.5f);
                if (__builtin_fabsf(mx) <= 1.0e6f && __builtin_fabsf(my) <= 1.0e6f) mv = f2{mx, my};
            }
        }
        if (hr) {
            const h4 q = {(_Float16)rad.x, (_Float16)rad.y, (_Float16)rad.z, (_Float16)rad.w};
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, q), r[Q_RAD], y * f.pitch[Q_RAD] + x * 8, 0, 0);
        } else {
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, rad), r[Q_RAD], y * f.pitch[Q_RAD] + x * 16, 0, 0);
        }
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, g), r[Q_G], y * f.pitch[Q_G] + x * 16, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, mv), r[Q_MV], y * f.pitch[Q_MV] + x * 8, 0, 0);
    }
}

__device__ __forceinline__ void pack_block(const uint32_t* starts, const PackDesc* frames, uint32_t n)
{
    const uint32_t b = blockIdx.x;
    if (b >= starts[n]) return;
    const uint32_t fi = frame_of(starts, n, b);
    PackDesc f;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(frames + fi);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&f);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(PackDesc) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    }
    __amdgpu_buffer_rsrc_t r[Q_COUNT];
#pragma unroll
    for (int k = 0; k < Q_COUNT; ++k) r[k] = rsrc(f.plane[k], f.pitch[k] * f.h);
    const uint32_t p0 = (b - starts[fi]) * PK_PX_PER_BLOCK + threadIdx.x;
    const uint32_t mask = f.half & PK_HALF_PLANES;
    if (mask == PK_HALF_PLANES) pack_pixels<FM_HALF>(f, r, p0);
    else if (mask == 0u) pack_pixels<FM_FLOAT>(f, r, p0);
    else pack_pixels<FM_MIXED>(f, r, p0);
}

__global__ __launch_bounds__(PK_BLOCK) void k_pack(PackArgs a) { pack_block(a.starts, a.frames, a.n); }

__global__ __launch_bounds__(PK_BLOCK) void k_pack_inline(PackInlineArgs a) { pack_block(a.starts, a.frames, a.n); }

// bytes per pixel of an accepted plane format, 0 if the slot does not take it
uint32_t slot_bytes(int slot, uint32_t fmt)
{
    switch (slot) {
    case Q_POS:
    case Q_G: return fmt == SVGF_FMT_RGBA32F ? 16 : 0;
    case Q_NRM: return fmt == SVGF_FMT_R32UI ? 4 : 0;
    case Q_MV: return fmt == SVGF_FMT_RG32F ? 8 : 0;
    default: return fmt == SVGF_FMT_RGBA32F ? 16 : fmt == SVGF_FMT_RGBA16F ? 8 : 0;
    }
}

}  // namespace
}  // namespace svgf

using namespace svgf;

struct svgf_pack {
    int device = 0;
    uint32_t max_frames = 0;
    size_t frames_off = 0, table_bytes = 0;
    TableRing ring;
};

extern "C" {

svgf_status svgf_pack_create(int32_t device, uint32_t max_frames, svgf_pack** out)
{
    if (!out || max_frames < 1 || max_frames > PK_MAX_FRAMES) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SVGF_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return SVGF_ERR_NO_DEVICE;
    DeviceGuard guard(device);
    svgf_pack* c = new (std::nothrow) svgf_pack();
    if (!c) return SVGF_ERR_OUT_OF_MEMORY;
    c->device = device;
    c->max_frames = max_frames;
    c->frames_off = (((size_t)max_frames + 1) * 4 + 255) / 256 * 256;
    c->table_bytes = c->frames_off + sizeof(PackDesc) * max_frames;
    if (!c->ring.create(c->table_bytes)) {
        c->ring.destroy();
        delete c;
        return SVGF_ERR_OUT_OF_MEMORY;
    }
    *out = c;
    return SVGF_OK;
}

svgf_status svgf_pack_destroy(svgf_pack* c)
{
    if (!c) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->device);
    c->ring.destroy();
    delete c;
    return SVGF_OK;
}

svgf_status svgf_pack_run(svgf_pack* c, const svgf_pack_frame* frames, uint32_t n, void* hip_stream)
{
    if (!c || (!frames && n)) return SVGF_ERR_INVALID_ARGUMENT;
    if (n == 0) return SVGF_OK;
    if (n > c->max_frames) return SVGF_ERR_CAPACITY;
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const svgf_pack_frame& f = frames[i];
        if (f.width < 1 || f.height < 1) return SVGF_ERR_INVALID_ARGUMENT;
        for (int k = 0; k < 16; ++k)
            if (!std::isfinite(f.view_projection[k]) || !std::isfinite(f.prev_view_projection[k]))
                return SVGF_ERR_INVALID_ARGUMENT;
        const svgf_plane* pl[Q_COUNT] = {&f.position, &f.normal_oct, &f.indirect, &f.direct_diffuse,
                                         &f.direct_specular, &f.radiance, &f.gbuffer, &f.motion};
        for (int k = 0; k < Q_COUNT; ++k) {
            const uint64_t px = slot_bytes(k, pl[k]->format);
            if (!pl[k]->ptr || !px) return SVGF_ERR_INVALID_ARGUMENT;
            if (pl[k]->pitch_bytes < (uint64_t)f.width * px || pl[k]->pitch_bytes % px || !aligned_to(pl[k]->ptr, px))
                return SVGF_ERR_INVALID_ARGUMENT;
            if (pl[k]->pitch_bytes * f.height >= (1ull << 31)) return SVGF_ERR_CAPACITY;
        }
        // outputs may not overlap each other or any input of the frame
        for (int o = Q_RAD; o < Q_COUNT; ++o) {
            const uintptr_t o0 = reinterpret_cast<uintptr_t>(pl[o]->ptr), o1 = o0 + pl[o]->pitch_bytes * f.height;
            for (int k = 0; k < Q_COUNT; ++k) {
                if (k == o) continue;
                const uintptr_t k0 = reinterpret_cast<uintptr_t>(pl[k]->ptr), k1 = k0 + pl[k]->pitch_bytes * f.height;
                if (o0 < k1 && k0 < o1) return SVGF_ERR_INVALID_ARGUMENT;
            }
        }
        blocks += ((uint64_t)f.width * f.height + PK_PX_PER_BLOCK - 1) / PK_PX_PER_BLOCK;
    }
    if (blocks >= (1ull << 31)) return SVGF_ERR_CAPACITY;
    DeviceGuard guard(c->device);
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    const bool inl = n <= PK_INLINE;
    PackInlineArgs ia{};
    TableRing::Slot* slot = nullptr;
    uint32_t* starts = ia.starts;
    PackDesc* descs = ia.frames;
    if (!inl) {
        slot = c->ring.acquire();
        if (!slot) return SVGF_ERR_HIP;
        starts = static_cast<uint32_t*>(slot->staging);
        descs = reinterpret_cast<PackDesc*>(static_cast<char*>(slot->staging) + c->frames_off);
    }
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const svgf_pack_frame& f = frames[i];
        starts[i] = acc;
        acc += (uint32_t)(((uint64_t)f.width * f.height + PK_PX_PER_BLOCK - 1) / PK_PX_PER_BLOCK);
        const svgf_plane* pl[Q_COUNT] = {&f.position, &f.normal_oct, &f.indirect, &f.direct_diffuse,
                                         &f.direct_specular, &f.radiance, &f.gbuffer, &f.motion};
        PackDesc d{};
        d.w = f.width;
        d.h = f.height;
        std::memcpy(d.vp, f.view_projection, sizeof(d.vp));
        std::memcpy(d.pvp, f.prev_view_projection, sizeof(d.pvp));
        for (int k = 0; k < Q_COUNT; ++k) {
            if (pl[k]->format == SVGF_FMT_RGBA16F) d.half |= 1u << k;
            d.plane[k] = static_cast<const char*>(pl[k]->ptr);
            d.pitch[k] = (uint32_t)pl[k]->pitch_bytes;
        }
        descs[i] = d;
    }
    starts[n] = acc;
    if (inl) {
        ia.n = n;
        hipLaunchKernelGGL(k_pack_inline, dim3(acc), dim3(PK_BLOCK), 0, stream, ia);
        return hipGetLastError() == hipSuccess ? SVGF_OK : SVGF_ERR_HIP;
    }
    if (!TableRing::upload(*slot, c->frames_off + sizeof(PackDesc) * n, stream)) return SVGF_ERR_HIP;
    PackArgs a{};
    a.starts = static_cast<const uint32_t*>(slot->table);
    a.frames = reinterpret_cast<const PackDesc*>(static_cast<const char*>(slot->table) + c->frames_off);
    a.n = n;
    hipLaunchKernelGGL(k_pack, dim3(acc), dim3(PK_BLOCK), 0, stream, a);
    if (hipGetLastError() != hipSuccess) return SVGF_ERR_HIP;
    if (!TableRing::release(*slot, stream)) return SVGF_ERR_HIP;
    return SVGF_OK;
}

}  // extern "C"
