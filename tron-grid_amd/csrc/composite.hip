// composite.hip — the reader of the denoised output: one-frame-late history at the same pixel,
// reflection hits reprojected into the history, and the final composite, for many frames in
// one launch.
//
// TronGrid does this inside its mesh fragment shader (src/mesh.slang:1007, 1161-1174, 1191):
//   indirect  = history[pixel].rgb
//   clip      = prev_view_projection (hit, 1)
//   if clip.w > 0: px = int2((clip.xy / clip.w (0.5, -0.5) + 0.5) (W, H));
//                  inside: hit_indirect = history[px].rgb
//   otherwise      hit_indirect = hit_base * AMBIENT
//   reflected = hit_emissive + hit_indirect
//   colour    = (indirect + emissive) + reflected * F_view, alpha 1
// Here the shading terms come in as planes (emissive, F_view, the reflection's kind and hit
// position, its emissive or ready colour, its base colour), so the composite is a streaming
// pass over HBM plus one gather per reprojected reflection.
//
// Arithmetic is float with one rounding per operation (contraction off); matrix rows are summed
// left to right, ((m0 x + m1 y) + m2 z) + m3; the division is correctly rounded. int2()
// truncates toward zero, so a product in (-1, 0) lands on column/row 0; NaN and out-of-range
// products are off screen.
//
// Launch shape: 256 threads per block, 1024 consecutive pixels of one frame's row-major order
// per block; thread t takes pixels t, t + 256, t + 512, t + 768 of the block, so every load and
// store instruction of a wave covers 64 consecutive pixels. Blocks map to frames by a binary
// search over the per-call prefix of block counts (uniform per block).
#include "../../include/svgf.h"
#include "svgf_internal.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>

namespace svgf {
namespace {

constexpr int COMP_BLOCK = 256;
constexpr int COMP_PX_PER_THREAD = 4;
constexpr int COMP_PX_PER_BLOCK = COMP_BLOCK * COMP_PX_PER_THREAD;  // 1024
constexpr uint32_t COMP_MAX_FRAMES = 4096;
constexpr int BUF_WORD3 = 0x00020000;

// plane slots of a frame descriptor
enum : int { P_HIST = 0, P_EMIS, P_FRES, P_REFL, P_RCOL, P_RBASE, P_OUT, P_COUNT };

struct CompDesc {
    uint32_t w, h;
    uint32_t half;      // bit k: plane k is RGBA16F (else RGBA32F)
    uint32_t reserved;
    float m[16];        // row-major previous view-projection
    const char* plane[P_COUNT];
    uint32_t pitch[P_COUNT];
    uint32_t pad;
};
static_assert(sizeof(CompDesc) % 8 == 0, "descriptor array alignment");

struct CompArgs {
    const uint32_t* starts;  // n + 1 block prefix
    const CompDesc* frames;
    uint32_t n;
    float ambient;
};

// Calls of up to COMP_INLINE frames pass prefix and descriptors in the kernel arguments
// instead of a device table: no host-to-device copy ahead of the kernel.
constexpr uint32_t COMP_INLINE = 12;
struct CompInlineArgs {
    uint32_t starts[COMP_INLINE + 1];
    uint32_t n;
    float ambient;
    uint32_t pad;
    CompDesc frames[COMP_INLINE];
};
static_assert(sizeof(CompInlineArgs) <= 4096 - 64, "kernel argument segment");

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, BUF_WORD3);
}

// rgb of one pixel of an RGBA16F (half) or RGBA32F plane; `half` is uniform
__device__ __forceinline__ f4 load_px(__amdgpu_buffer_rsrc_t r, uint32_t off_row, uint32_t x, bool half)
{
    if (half) {
        const h4 v = __builtin_bit_cast(h4, __builtin_amdgcn_raw_buffer_load_b64(r, off_row + x * 8, 0, 0));
        return f4{(float)v.x, (float)v.y, (float)v.z, (float)v.w};
    }
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, off_row + x * 16, 0, 0));
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

__device__ __forceinline__ void store_px(__amdgpu_buffer_rsrc_t r, uint32_t off, f4 v, bool half)
{
    if (half) {
        const h4 q = {(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, q), r, off, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, off, 0, 0);
    }
}

// row r of m times (x, y, z, 1), summed left to right
__device__ __forceinline__ float row4(const float* m, int r, float x, float y, float z)
{
#pragma clang fp contract(off)
    return ((m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z) + m[4 * r + 3];
}

// mesh.slang:1162-1170: the history pixel a world position reprojects to; false when it is
// behind the camera or off screen
__device__ __forceinline__ bool reproject(const float* m, uint32_t w, uint32_t h, float x, float y, float z,
                                          uint32_t& px, uint32_t& py)
{
#pragma clang fp contract(off)
    const float cw = row4(m, 3, x, y, z);
    if (!(cw > 0.0f)) return false;
    const float nx = row4(m, 0, x, y, z) / cw;
    const float ny = row4(m, 1, x, y, z) / cw;
    const float fx = (nx * 0.5f + 0.5f) * (float)w;
    const float fy = (ny * -0.5f + 0.5f) * (float)h;
    // trunc(f) lies in [0, n) exactly when -1 < f < n; NaN fails both tests
    if (!(fx > -1.0f && fx < (float)w && fy > -1.0f && fy < (float)h)) return false;
    px = (uint32_t)(int)fx;
    py = (uint32_t)(int)fy;
    return true;
}

// Planes of a frame that may be RGBA16F (every one but the reflection's kind / hit position).
constexpr uint32_t HALF_PLANES = ((1u << P_COUNT) - 1u) & ~(1u << P_REFL);
// Plane formats of a frame as the block body sees them: every colour plane RGBA16F, every one
// RGBA32F (both compile-time, so all independent loads issue before the first conversion), or
// mixed (per plane at run time: each load then sits in a branch of its own and waits there).
enum Formats : int { FM_HALF = 0, FM_FLOAT = 1, FM_MIXED = 2 };

template <int FM>
__device__ __forceinline__ bool half_of(uint32_t mask, int k)
{
    if constexpr (FM == FM_HALF) return k != P_REFL;
    else if constexpr (FM == FM_FLOAT) return false;
    else return (mask >> k) & 1u;
}

// The composite of a block's pixels: the direct loads of all four pixels of the thread first,
// then the reflection loads that depend on them (the colour where a reflection has one, the
// history pixel its hit reprojects to or the hit's base colour), then the arithmetic.
template <int FM>
__device__ __forceinline__ void composite_pixels(const CompDesc& f, const __amdgpu_buffer_rsrc_t* r, uint32_t p0,
                                                 float ambient)
{
#pragma clang fp contract(off)
    const bool hh = half_of<FM>(f.half, P_HIST), he = half_of<FM>(f.half, P_EMIS), hf = half_of<FM>(f.half, P_FRES);
    const bool hc = half_of<FM>(f.half, P_RCOL), hb = half_of<FM>(f.half, P_RBASE), ho = half_of<FM>(f.half, P_OUT);
    const uint32_t npx = f.w * f.h;
    uint32_t xs[COMP_PX_PER_THREAD], ys[COMP_PX_PER_THREAD];
    f4 hist[COMP_PX_PER_THREAD], emis[COMP_PX_PER_THREAD], fres[COMP_PX_PER_THREAD], refl[COMP_PX_PER_THREAD];
#pragma unroll
    for (int j = 0; j < COMP_PX_PER_THREAD; ++j) {
        const uint32_t p = min(p0 + j * COMP_BLOCK, npx - 1);  // clamped tail lanes load a valid pixel
        ys[j] = p / f.w;
        xs[j] = p - ys[j] * f.w;
        hist[j] = load_px(r[P_HIST], ys[j] * f.pitch[P_HIST], xs[j], hh);
        emis[j] = load_px(r[P_EMIS], ys[j] * f.pitch[P_EMIS], xs[j], he);
        fres[j] = load_px(r[P_FRES], ys[j] * f.pitch[P_FRES], xs[j], hf);
        refl[j] = load_px(r[P_REFL], ys[j] * f.pitch[P_REFL], xs[j], false);
    }
    f4 col[COMP_PX_PER_THREAD], ind[COMP_PX_PER_THREAD];
    bool hit[COMP_PX_PER_THREAD];
    if constexpr (FM == FM_MIXED) {
#pragma unroll
        for (int j = 0; j < COMP_PX_PER_THREAD; ++j) {
            const float kind = refl[j].w;
            col[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            ind[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            hit[j] = false;
            if (kind == 1.0f || kind == 2.0f) col[j] = load_px(r[P_RCOL], ys[j] * f.pitch[P_RCOL], xs[j], hc);
            if (kind == 2.0f) {
                uint32_t px, py;
                hit[j] = reproject(f.m, f.w, f.h, refl[j].x, refl[j].y, refl[j].z, px, py);
                ind[j] = hit[j] ? load_px(r[P_HIST], py * f.pitch[P_HIST], px, hh)
                                : load_px(r[P_RBASE], ys[j] * f.pitch[P_RBASE], xs[j], hb);
            }
        }
    } else {
        // the raw words are loaded in the branches and converted after all of them are issued
        // (a conversion inside a branch would wait there for its load)
        constexpr bool H = FM == FM_HALF;
        using Raw = std::conditional_t<H, u2, u4>;
        // every pixel's reprojection first: computed after one wait for the direct loads, where
        // a gather issued in between would have to complete too (loads retire in order)
        uint32_t gx[COMP_PX_PER_THREAD], gy[COMP_PX_PER_THREAD];
#pragma unroll
        for (int j = 0; j < COMP_PX_PER_THREAD; ++j) {
            gx[j] = xs[j];
            gy[j] = ys[j];
            hit[j] = refl[j].w == 2.0f && reproject(f.m, f.w, f.h, refl[j].x, refl[j].y, refl[j].z, gx[j], gy[j]);
        }
        // No branch: a load a pixel does not need takes an offset past its plane, which the
        // buffer resource answers with 0 and no memory access (the ABI keeps planes under
        // 2 GiB). The loads of all four pixels are then in flight before the first conversion.
        constexpr uint32_t OOB = 0x80000000u;
        Raw rcol[COMP_PX_PER_THREAD], rhist[COMP_PX_PER_THREAD], rbase[COMP_PX_PER_THREAD];
#pragma unroll
        for (int j = 0; j < COMP_PX_PER_THREAD; ++j) {
            const float kind = refl[j].w;
            const bool has_col = kind == 1.0f || kind == 2.0f, miss = kind == 2.0f && !hit[j];
            rcol[j] = load_raw<H>(r[P_RCOL], has_col ? ys[j] * f.pitch[P_RCOL] : OOB, has_col ? xs[j] : 0u);
            rhist[j] = load_raw<H>(r[P_HIST], hit[j] ? gy[j] * f.pitch[P_HIST] : OOB, hit[j] ? gx[j] : 0u);
            rbase[j] = load_raw<H>(r[P_RBASE], miss ? ys[j] * f.pitch[P_RBASE] : OOB, miss ? xs[j] : 0u);
        }
#pragma unroll
        for (int j = 0; j < COMP_PX_PER_THREAD; ++j) {
            col[j] = raw_f4<H>(rcol[j]);
            ind[j] = hit[j] ? raw_f4<H>(rhist[j]) : raw_f4<H>(rbase[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < COMP_PX_PER_THREAD; ++j) {
        if (p0 + j * COMP_BLOCK >= npx) break;
        const float kind = refl[j].w;
        f4 rc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (kind == 2.0f) {
            const f4 in = hit[j] ? ind[j] : f4{ind[j].x * ambient, ind[j].y * ambient, ind[j].z * ambient, 0.0f};
            rc = f4{col[j].x + in.x, col[j].y + in.y, col[j].z + in.z, 0.0f};
        } else if (kind == 1.0f) {
            rc = col[j];
        }
        const f4 o = {(hist[j].x + emis[j].x) + rc.x * fres[j].x, (hist[j].y + emis[j].y) + rc.y * fres[j].y,
                      (hist[j].z + emis[j].z) + rc.z * fres[j].z, 1.0f};
        store_px(r[P_OUT], ys[j] * f.pitch[P_OUT] + xs[j] * (ho ? 8u : 16u), o, ho);
    }
}

__device__ __forceinline__ void composite_block(const uint32_t* starts, const CompDesc* frames, uint32_t n,
                                                float ambient)
{
    const uint32_t b = blockIdx.x;
    if (b >= starts[n]) return;
    const uint32_t fi = frame_of(starts, n, b);
    CompDesc f;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(frames + fi);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&f);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(CompDesc) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    }
    __amdgpu_buffer_rsrc_t r[P_COUNT];
#pragma unroll
    for (int k = 0; k < P_COUNT; ++k) r[k] = rsrc(f.plane[k], f.pitch[k] * f.h);
    const uint32_t p0 = (b - starts[fi]) * COMP_PX_PER_BLOCK + threadIdx.x;
    const uint32_t mask = f.half & HALF_PLANES;
    if (mask == HALF_PLANES) composite_pixels<FM_HALF>(f, r, p0, ambient);
    else if (mask == 0u) composite_pixels<FM_FLOAT>(f, r, p0, ambient);
    else composite_pixels<FM_MIXED>(f, r, p0, ambient);
}

__global__ __launch_bounds__(COMP_BLOCK) void k_composite(CompArgs a)
{
    composite_block(a.starts, a.frames, a.n, a.ambient);
}

__global__ __launch_bounds__(COMP_BLOCK) void k_composite_inline(CompInlineArgs a)
{
    composite_block(a.starts, a.frames, a.n, a.ambient);
}

bool plane_half(const svgf_plane& p, bool allow_half, bool& half)
{
    if (p.format == SVGF_FMT_RGBA32F) half = false;
    else if (p.format == SVGF_FMT_RGBA16F && allow_half) half = true;
    else return false;
    return true;
}

}  // namespace
}  // namespace svgf

using namespace svgf;

struct svgf_composite {
    int device = 0;
    uint32_t max_frames = 0;
    size_t frames_off = 0, table_bytes = 0;
    TableRing ring;  // per call: starts (n + 1), then the descriptors
};

extern "C" {

svgf_status svgf_composite_default_params(svgf_composite_params* out)
{
    if (!out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = svgf_composite_params{};
    out->ambient = 0.03f;  // AMBIENT, src/mesh.slang:276
    return SVGF_OK;
}

svgf_status svgf_composite_create(int32_t device, uint32_t max_frames, svgf_composite** out)
{
    if (!out || max_frames < 1 || max_frames > COMP_MAX_FRAMES) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SVGF_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return SVGF_ERR_NO_DEVICE;
    DeviceGuard guard(device);
    svgf_composite* c = new (std::nothrow) svgf_composite();
    if (!c) return SVGF_ERR_OUT_OF_MEMORY;
    c->device = device;
    c->max_frames = max_frames;
    c->frames_off = (((size_t)max_frames + 1) * 4 + 255) / 256 * 256;
    c->table_bytes = c->frames_off + sizeof(CompDesc) * max_frames;
    if (!c->ring.create(c->table_bytes)) {
        c->ring.destroy();
        delete c;
        return SVGF_ERR_OUT_OF_MEMORY;
    }
    *out = c;
    return SVGF_OK;
}

svgf_status svgf_composite_destroy(svgf_composite* c)
{
    if (!c) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->device);
    c->ring.destroy();
    delete c;
    return SVGF_OK;
}

svgf_status svgf_composite_run(svgf_composite* c, const svgf_composite_frame* frames, uint32_t n,
                               const svgf_composite_params* params, void* hip_stream)
{
    if (!c || (!frames && n) || !params) return SVGF_ERR_INVALID_ARGUMENT;
    if (n == 0) return SVGF_OK;
    if (n > c->max_frames) return SVGF_ERR_CAPACITY;
    if (!std::isfinite(params->ambient)) return SVGF_ERR_INVALID_ARGUMENT;
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const svgf_composite_frame& f = frames[i];
        if (f.width < 1 || f.height < 1) return SVGF_ERR_INVALID_ARGUMENT;
        for (float v : f.prev_view_projection)
            if (!std::isfinite(v)) return SVGF_ERR_INVALID_ARGUMENT;
        const svgf_plane* pl[P_COUNT] = {&f.history, &f.emissive, &f.fresnel, &f.reflection, &f.reflection_colour,
                                         &f.reflection_base, &f.out};
        for (int k = 0; k < P_COUNT; ++k) {
            bool half = false;
            if (!pl[k]->ptr || !plane_half(*pl[k], k != P_REFL, half)) return SVGF_ERR_INVALID_ARGUMENT;
            const uint64_t px = half ? 8 : 16;
            if (pl[k]->pitch_bytes < (uint64_t)f.width * px || pl[k]->pitch_bytes % px || !aligned_to(pl[k]->ptr, px))
                return SVGF_ERR_INVALID_ARGUMENT;
            if (pl[k]->pitch_bytes * f.height >= (1ull << 31)) return SVGF_ERR_CAPACITY;
        }
        // the output may not overlap the history other pixels gather from
        const uintptr_t o0 = reinterpret_cast<uintptr_t>(f.out.ptr), o1 = o0 + f.out.pitch_bytes * f.height;
        const uintptr_t h0 = reinterpret_cast<uintptr_t>(f.history.ptr), h1 = h0 + f.history.pitch_bytes * f.height;
        if (o0 < h1 && h0 < o1) return SVGF_ERR_INVALID_ARGUMENT;
        blocks += ((uint64_t)f.width * f.height + COMP_PX_PER_BLOCK - 1) / COMP_PX_PER_BLOCK;
    }
    if (blocks >= (1ull << 31)) return SVGF_ERR_CAPACITY;
    DeviceGuard guard(c->device);
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    const bool inl = n <= COMP_INLINE;
    CompInlineArgs ia{};
    TableRing::Slot* slot = nullptr;
    uint32_t* starts = ia.starts;
    CompDesc* descs = ia.frames;
    if (!inl) {
        slot = c->ring.acquire();
        if (!slot) return SVGF_ERR_HIP;
        starts = static_cast<uint32_t*>(slot->staging);
        descs = reinterpret_cast<CompDesc*>(static_cast<char*>(slot->staging) + c->frames_off);
    }
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const svgf_composite_frame& f = frames[i];
        starts[i] = acc;
        acc += (uint32_t)(((uint64_t)f.width * f.height + COMP_PX_PER_BLOCK - 1) / COMP_PX_PER_BLOCK);
        const svgf_plane* pl[P_COUNT] = {&f.history, &f.emissive, &f.fresnel, &f.reflection, &f.reflection_colour,
                                         &f.reflection_base, &f.out};
        CompDesc d{};
        d.w = f.width;
        d.h = f.height;
        std::memcpy(d.m, f.prev_view_projection, sizeof(d.m));
        for (int k = 0; k < P_COUNT; ++k) {
            if (pl[k]->format == SVGF_FMT_RGBA16F) d.half |= 1u << k;
            d.plane[k] = static_cast<const char*>(pl[k]->ptr);
            d.pitch[k] = (uint32_t)pl[k]->pitch_bytes;
        }
        descs[i] = d;
    }
    starts[n] = acc;
    if (inl) {
        ia.n = n;
        ia.ambient = params->ambient;
        hipLaunchKernelGGL(k_composite_inline, dim3(acc), dim3(COMP_BLOCK), 0, stream, ia);
        return hipGetLastError() == hipSuccess ? SVGF_OK : SVGF_ERR_HIP;
    }
    if (!TableRing::upload(*slot, c->frames_off + sizeof(CompDesc) * n, stream)) return SVGF_ERR_HIP;
    CompArgs a{};
    a.starts = static_cast<const uint32_t*>(slot->table);
    a.frames = reinterpret_cast<const CompDesc*>(static_cast<const char*>(slot->table) + c->frames_off);
    a.n = n;
    a.ambient = params->ambient;
    hipLaunchKernelGGL(k_composite, dim3(acc), dim3(COMP_BLOCK), 0, stream, a);
    if (hipGetLastError() != hipSuccess) return SVGF_ERR_HIP;
    if (!TableRing::release(*slot, stream)) return SVGF_ERR_HIP;
    return SVGF_OK;
}

}  // extern "C"
