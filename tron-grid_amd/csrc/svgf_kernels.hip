// svgf_kernels.hip — SVGF stages as gfx950 kernels.
//
// Every stage is a block-level __device__ function over one 16x16 lattice tile (see
// svgf_internal.hpp). The single-frame kernels launch one block per tile of one frame; the
// batched kernels grid-stride over the tiles of every frame of a batch. Both call the same
// function, so a frame gives bit-identical results alone and inside any batch.
//
// Memory: planes are read and written through raw buffer resources (SGPR descriptor + 32-bit
// offset): no 64-bit address arithmetic per access, and an access past the end of a plane
// reads 0 / writes nothing instead of faulting. A-trous taps come from LDS: each block stages
// its tile plus a 2-point lattice halo once (converted to float, luminance precomputed), so
// the 25 taps of every output cost only LDS reads.
//
// REFERENCE stages follow src/svgf_denoise.slang:152-275 operation by operation (no
// contraction where the shader's rounding matters: luminance, lerp, variance). CANONICAL
// stages follow Schied et al. 2017 with FMA contraction (checked against the double oracle).
#include "svgf_internal.hpp"

namespace svgf {
namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// gfx9 raw buffer descriptor word 3 (32-bit data format, untyped access)
constexpr int BUF_WORD3 = 0x00020000;

struct Buf {
    __amdgpu_buffer_rsrc_t r;
    uint32_t pitch;
};

__device__ __forceinline__ Buf bind(const PlaneDev& p, int h)
{
    return Buf{__builtin_amdgcn_make_buffer_rsrc(p.ptr, (short)0, (int)(p.pitch * (uint64_t)h), BUF_WORD3),
               (uint32_t)p.pitch};
}

__device__ __forceinline__ float4 ld_f4(const Buf& b, uint32_t off)
{
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(b.r, off, 0, 0));
}
__device__ __forceinline__ void st_f4(const Buf& b, uint32_t off, float4 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), b.r, off, 0, 0);
}
__device__ __forceinline__ float2 ld_f2(const Buf& b, uint32_t off)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.r, off, 0, 0));
}
__device__ __forceinline__ void st_f2(const Buf& b, uint32_t off, float2 v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), b.r, off, 0, 0);
}
__device__ __forceinline__ float ld_f1(const Buf& b, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, off, 0, 0));
}
__device__ __forceinline__ void st_f1(const Buf& b, uint32_t off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), b.r, off, 0, 0);
}
__device__ __forceinline__ uint32_t ld_u8(const Buf& b, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b8(b.r, off, 0, 0);
}
__device__ __forceinline__ void st_u8(const Buf& b, uint32_t off, uint32_t v)
{
    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v, b.r, off, 0, 0);
}

// Storage traits: FP32 planes (float4 / float2) or FP16 planes (half4 / half2), float math.
struct F32S {
    static constexpr uint32_t B4 = 16, B2 = 8;
    __device__ static __forceinline__ float4 ld4(const Buf& b, int x, int y) { return ld_f4(b, y * b.pitch + x * B4); }
    __device__ static __forceinline__ void st4(const Buf& b, int x, int y, float4 v) { st_f4(b, y * b.pitch + x * B4, v); }
    __device__ static __forceinline__ float2 ld2(const Buf& b, int x, int y) { return ld_f2(b, y * b.pitch + x * B2); }
    __device__ static __forceinline__ void st2(const Buf& b, int x, int y, float2 v) { st_f2(b, y * b.pitch + x * B2, v); }
    // pixels x and x + 1 of row y, v1 = v0 when `first` (a read past the plane's end returns
    // zeros)
    __device__ static __forceinline__ void ld4x2(const Buf& b, int x, int y, bool first, float4& v0, float4& v1)
    {
        v0 = ld_f4(b, y * b.pitch + x * B4);
        v1 = ld_f4(b, y * b.pitch + x * B4 + (first ? 0 : B4));
    }
    __device__ static __forceinline__ void ld2x2(const Buf& b, int x, int y, bool first, float2& v0, float2& v1)
    {
        const float4 v = ld_f4(b, y * b.pitch + x * B2);
        v0 = make_float2(v.x, v.y);
        v1 = first ? v0 : make_float2(v.z, v.w);
    }
    // alpha channel alone
    __device__ static __forceinline__ float lda(const Buf& b, int x, int y) { return ld_f1(b, y * b.pitch + x * B4 + 12); }
    __device__ static __forceinline__ void sta(const Buf& b, int x, int y, float v) { st_f1(b, y * b.pitch + x * B4 + 12, v); }
};

struct F16S {
    static constexpr uint32_t B4 = 8, B2 = 4;
    __device__ static __forceinline__ float4 ld4(const Buf& b, int x, int y)
    {
        half4_t h = __builtin_bit_cast(half4_t, __builtin_amdgcn_raw_buffer_load_b64(b.r, y * b.pitch + x * B4, 0, 0));
        return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
    }
    __device__ static __forceinline__ void st4(const Buf& b, int x, int y, float4 v)
    {
        half4_t h = {(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, h), b.r, y * b.pitch + x * B4, 0, 0);
    }
    __device__ static __forceinline__ float2 ld2(const Buf& b, int x, int y)
    {
        half2_t h = __builtin_bit_cast(half2_t, __builtin_amdgcn_raw_buffer_load_b32(b.r, y * b.pitch + x * B2, 0, 0));
        return make_float2((float)h.x, (float)h.y);
    }
    __device__ static __forceinline__ void st2(const Buf& b, int x, int y, float2 v)
    {
        half2_t h = {(_Float16)v.x, (_Float16)v.y};
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, h), b.r, y * b.pitch + x * B2, 0, 0);
    }
    // pixels x and x + 1 of row y in one load, v1 = v0 when `first` (selected before the
    // conversion; a read past the plane's end returns zeros)
    __device__ static __forceinline__ void ld4x2(const Buf& b, int x, int y, bool first, float4& v0, float4& v1)
    {
        const u32x4_t r = __builtin_amdgcn_raw_buffer_load_b128(b.r, y * b.pitch + x * B4, 0, 0);
        const half4_t h0 = __builtin_bit_cast(half4_t, u32x2_t{r.x, r.y});
        const half4_t h1 = __builtin_bit_cast(half4_t, first ? u32x2_t{r.x, r.y} : u32x2_t{r.z, r.w});
        v0 = make_float4((float)h0.x, (float)h0.y, (float)h0.z, (float)h0.w);
        v1 = make_float4((float)h1.x, (float)h1.y, (float)h1.z, (float)h1.w);
    }
    __device__ static __forceinline__ void ld2x2(const Buf& b, int x, int y, bool first, float2& v0, float2& v1)
    {
        const u32x2_t r = __builtin_amdgcn_raw_buffer_load_b64(b.r, y * b.pitch + x * B2, 0, 0);
        const half2_t h0 = __builtin_bit_cast(half2_t, r.x), h1 = __builtin_bit_cast(half2_t, first ? r.x : r.y);
        v0 = make_float2((float)h0.x, (float)h0.y);
        v1 = make_float2((float)h1.x, (float)h1.y);
    }
    // raw 64-bit words of pixels x and x + 1 (r1 = r0 when `first`)
    __device__ static __forceinline__ void ld_raw4x2(const Buf& b, int x, int y, bool first, u32x2_t& r0, u32x2_t& r1)
    {
        const u32x4_t r = __builtin_amdgcn_raw_buffer_load_b128(b.r, y * b.pitch + x * B4, 0, 0);
        r0 = u32x2_t{r.x, r.y};
        r1 = first ? r0 : u32x2_t{r.z, r.w};
    }
    __device__ static __forceinline__ void st_raw4(const Buf& b, int x, int y, half4_t h)
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, h), b.r, y * b.pitch + x * B4, 0, 0);
    }
    __device__ static __forceinline__ float lda(const Buf& b, int x, int y)
    {
        return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(b.r, y * b.pitch + x * B4 + 6, 0, 0));
    }
    __device__ static __forceinline__ void sta(const Buf& b, int x, int y, float v)
    {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), b.r, y * b.pitch + x * B4 + 6,
                                              0, 0);
    }
};

// float4 G-buffer, float2 motion (FP32 in both storage forms)
__device__ __forceinline__ float4 ldg4(const Buf& b, int x, int y) { return ld_f4(b, y * b.pitch + x * 16); }
__device__ __forceinline__ void stg4(const Buf& b, int x, int y, float4 v) { st_f4(b, y * b.pitch + x * 16, v); }
__device__ __forceinline__ float2 ldg2(const Buf& b, int x, int y) { return ld_f2(b, y * b.pitch + x * 8); }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Wide state of the FP16 CANONICAL form (word layout in svgf_state.hpp).
// History pixel from its raw 64-bit word: colour decoded with its residual codes, length n.
__device__ __forceinline__ float4 wide_history(u32x2_t r, uint32_t& n)
{
    const half4_t h = __builtin_bit_cast(half4_t, r);
    const uint16_t word = (uint16_t)(r.y >> 16);
    n = wide_history_length(word);
    return make_float4(wide_decode_channel(h.x, wide_history_code(word, 0)),
                       wide_decode_channel(h.y, wide_history_code(word, 1)),
                       wide_decode_channel(h.z, wide_history_code(word, 2)), 0.0f);
}
__device__ __forceinline__ void st_wide_history(const Buf& b, int x, int y, float4 o, uint32_t n)
{
    uint32_t q0, q1, q2;
    const _Float16 h0 = wide_encode_channel(o.x, q0), h1 = wide_encode_channel(o.y, q1), h2 = wide_encode_channel(o.z, q2);
    const half4_t h = {h0, h1, h2, __builtin_bit_cast(_Float16, wide_history_word(n, q0, q1, q2))};
    F16S::st_raw4(b, x, y, h);
}
// Moments pixel (mean, sigma) from its raw (hi, hi, lo, lo) binary16 words.
__device__ __forceinline__ float2 wide_moments(u32x2_t r)
{
    const half4_t h = __builtin_bit_cast(half4_t, r);
    return make_float2(wide_decode_moment(h.x, h.z), wide_decode_moment(h.y, h.w));
}
__device__ __forceinline__ void st_wide_moments(const Buf& b, int x, int y, float mu, float sd)
{
    _Float16 mh, ml, sh, sl;
    wide_encode_moment(mu, mh, ml);
    wide_encode_moment(sd, sh, sl);
    F16S::st_raw4(b, x, y, half4_t{mh, sh, ml, sl});
}

// Completes this wave's outstanding LDS operations (s_waitcnt lgkmcnt(0)). A value one thread
// writes to LDS for the other waves must have landed before that thread's wave reaches the
// barrier that publishes it; the compiler does not always place this wait before a barrier
// reached through a loop back edge, and a wave that passes the barrier early reads the old value.
__device__ __forceinline__ void lds_writes_done() { __builtin_amdgcn_s_waitcnt(0xc07f); }

// The thread index as a value the optimiser cannot see through. Stage functions derive their
// LDS addresses from it; in the persistent kernel's ticket loop those addresses would
// otherwise be hoisted out of the loop one per tap, each holding a VGPR for the whole kernel.
__device__ __forceinline__ int lane_tid()
{
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// computeLuminance (svgf_denoise.slang:109-112) rounded as the shader's dot: three products,
// two sums, left to right, no fused multiply-add.
__device__ __forceinline__ float luma_ref(float r, float g, float b)
{
#pragma clang fp contract(off)
    return r * 0.2126f + g * 0.7152f + b * 0.0722f;
}
// Same weights with fused multiply-adds (canonical stages).
__device__ __forceinline__ float luma_fma(float r, float g, float b)
{
    return fmaf(b, 0.0722f, fmaf(g, 0.7152f, r * 0.2126f));
}

constexpr float LOG2E = 1.44269504f;

// exp(x) as the hardware exp2 of a prescaled argument (v_exp_f32); relative error about
// |x| * 6e-8 from the prescale rounding.
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * LOG2E); }

// Exact decision of dot(a, b) >= t for float inputs: each product is split into its rounded
// value and exact error (FMA), the sum is carried with two-sum error terms, so the sign
// matches the real-number comparison except within ~2^-48 of the threshold.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ bool dot_at_least(float4 a, float4 b, float t)
{
#pragma clang fp contract(off)
    float p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
    float e0 = __builtin_fmaf(a.x, b.x, -p0);
    float e1 = __builtin_fmaf(a.y, b.y, -p1);
    float e2 = __builtin_fmaf(a.z, b.z, -p2);
    float s1, c1, s2, c2, s3, c3;
    two_sum(p0, p1, s1, c1);
    two_sum(s1, p2, s2, c2);
    two_sum(s2, -t, s3, c3);
    float tail = ((c1 + c2) + c3) + ((e0 + e1) + e2);
    return s3 + tail >= 0.0f;
}

constexpr float K5[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
// log2(K[i] * K[j]) for the weights folded into one exp2
__host__ __device__ constexpr float log2k(int i, int j)
{
    constexpr float L[5] = {-4.0f, -2.0f, -1.41503751f, -2.0f, -4.0f};  // log2(1/16), log2(4/16), log2(6/16)
    return L[i] + L[j];
}

// ----------------------------------------------------------------------------------------
// Tiles (decoded by svgf_sched.hpp)
// ----------------------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2). Renumber so
// that XCD x runs one contiguous range of logical tiles: neighbouring tiles, which share
// halo rows and (for s > 1) cache lines of other residues, then meet in the same L2.
__device__ __forceinline__ uint32_t xcd_logical(uint32_t b, uint32_t n)
{
    uint32_t q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

// ========================================================================================
// LDS of the stage tiles
// ========================================================================================
// Bytes a stage tile stages in LDS: the a-trous tile (colour + luminance, G, variance per entry
// of NR staged regions, the young list and its count), the REFERENCE a-trous halo, the
// temporal stage's depth tile.
__host__ __device__ constexpr uint32_t atrous_lds_bytes(bool first, int nr, bool tall)
{
    const uint32_t h = first ? LHALO + 1 : LHALO, th = tall ? 2 * LT : LT;
    const uint32_t entries = (uint32_t)nr * (LT + 2 * h) * (th + 2 * h);
    const uint32_t list = first ? (LT + 2 * LHALO) * (th + 2 * LHALO) * 2u : 0u;
    return ((entries * 36u + list + 15u) & ~15u) + 16u;
}
constexpr uint32_t REF_ATROUS_LDS = LREG * LREG * 16;
constexpr uint32_t cmax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// A kernel that runs one stage gives its tile function arrays of its own (OwnLds). The
// persistent kernel runs every stage in turn: its tile functions share one arena sized for the
// largest (ArenaLds), so its LDS is one stage's rather than the sum of all of them.
struct OwnLds {
    template <uint32_t BYTES>
    __device__ static __forceinline__ char* get()
    {
        __shared__ __align__(16) char m[BYTES];
        return m;
    }
};
constexpr uint32_t PERSIST_LDS_CAN =
    cmax(cmax(atrous_lds_bytes(true, 1, true), atrous_lds_bytes(false, 2, false)),
         atrous_lds_bytes(false, 1, false));
__shared__ __align__(16) char g_persist_lds_can[PERSIST_LDS_CAN];
__shared__ __align__(16) char g_persist_lds_ref[REF_ATROUS_LDS];
template <bool CANON>
struct ArenaLds {
    template <uint32_t BYTES>
    __device__ static __forceinline__ char* get()
    {
        static_assert(BYTES <= (CANON ? PERSIST_LDS_CAN : REF_ATROUS_LDS), "persistent LDS arena too small");
        if constexpr (CANON) return g_persist_lds_can;
        else return g_persist_lds_ref;
    }
};

// ========================================================================================
// REFERENCE stages (svgf_denoise.slang)
// ========================================================================================

// Pass 0 (svgf_denoise.slang:176-191): moments EMA, variance, sqrt into wavelet_a alpha.
template <class S>
__device__ __forceinline__ void ref_init_tile(const FrameDesc& f, const KParams& p, const Tile& T)
{
#pragma clang fp contract(off)
    const int tid = lane_tid();
    const int x = T.i0 + (tid & (LT - 1)), y = T.j0 + (tid >> 4);
    if (x >= f.w || y >= f.h) return;
    const Buf rad = bind(f.rad, f.h), mc = bind(f.mom_curr, f.h), wa = bind(f.wa, f.h);
    float4 L = S::ld4(rad, x, y);
    float l = luma_ref(L.x, L.y, L.z);
    float m1 = 0.0f, m2 = 0.0f;
    if (f.history_valid) {
        float2 m = S::ld2(bind(f.mom_prev, f.h), x, y);
        m1 = m.x;
        m2 = m.y;
    }
    float mu1 = m1 + p.alpha_moments * (l - m1);
    float mu2 = m2 + p.alpha_moments * (l * l - m2);
    float var = fmaxf(mu2 - mu1 * mu1, 0.0f);
    S::st2(mc, x, y, make_float2(mu1, mu2));
    S::st4(wa, x, y, make_float4(L.x, L.y, L.z, __builtin_sqrtf(var + p.variance_eps)));
}

// Passes 1..N (svgf_denoise.slang:209-274): 5x5 B3-spline taps at `stride`, clamp-to-edge,
// weight K[dx] K[dy] exp(-|l_q - l_c| / max(phi * sigma_c, eps)), dy outer / dx inner.
template <class S, class L>
__device__ __forceinline__ void ref_atrous_tile(const FrameDesc& f, const KParams& p, const Tile& T, int pass)
{
    float4* const sA = reinterpret_cast<float4*>(L::template get<REF_ATROUS_LDS>());  // (r, g, b, luminance) of the tile + halo
    const bool source_is_a = (pass & 1) == 1;
    const Buf src = bind(source_is_a ? f.wa : f.wb, f.h);
    const int max_x = f.w - 1, max_y = f.h - 1;
    const int tid = lane_tid();
    const int ti = tid & (LT - 1), tj = tid >> 4;
    const int x = T.px(T.i0 + ti), y = T.py(T.j0 + tj);
    // The centre's sigma (alpha) and every entry of the region are loaded before the first LDS
    // write: one memory round trip. The last pass's surplus lanes load the final entry again.
    const float cw = S::lda(src, min(x, max_x), min(y, max_y));
    constexpr int NE = LREG * LREG, NIT = (NE + BLOCK - 1) / BLOCK;
    float4 v[NIT];
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
        const int e = min(tid + j * BLOCK, NE - 1);
        const int li = e % LREG, lj = e / LREG;
        const int qx = clampi(T.px(T.i0 - LHALO + li), 0, max_x), qy = clampi(T.py(T.j0 - LHALO + lj), 0, max_y);
        v[j] = S::ld4(src, qx, qy);
    }
#pragma unroll
    for (int j = 0; j < NIT; ++j) asm volatile("" ::"v"(v[j].x), "v"(v[j].y), "v"(v[j].z));
#pragma unroll
    for (int j = 0; j < NIT; ++j)
        if (tid + j * BLOCK < NE) sA[tid + j * BLOCK] = make_float4(v[j].x, v[j].y, v[j].z, luma_ref(v[j].x, v[j].y, v[j].z));
    __syncthreads();
    if (x <= max_x && y <= max_y) {
        const int c0 = (tj + LHALO) * LREG + ti + LHALO;
        float4 c = sA[c0];
        float lc = c.w;
        float sigma = fmaxf(p.phi_luma * cw, p.variance_eps);
        float nrs = -1.0f / sigma;
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, ws = 0.0f;
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
            for (int dx = -2; dx <= 2; ++dx) {
                float4 v = sA[c0 + dy * LREG + dx];
                float tw = (K5[dx + 2] * K5[dy + 2]) * fast_exp(fabsf(v.w - lc) * nrs);
                s0 = fmaf(v.x, tw, s0);
                s1 = fmaf(v.y, tw, s1);
                s2 = fmaf(v.z, tw, s2);
                ws += tw;
            }
            // the row's sums are complete here (keeps later rows' weights from staying live)
            asm volatile("" : "+v"(s0), "+v"(s1), "+v"(s2), "+v"(ws));
        }
        float4 o;
        if (ws > 1e-6f) o = make_float4(s0 / ws, s1 / ws, s2 / ws, cw);
        else o = make_float4(c.x, c.y, c.z, cw);
        if (pass == p.iterations) {
            o.w = 1.0f;
            S::st4(bind(f.den, f.h), x, y, o);
        } else {
            S::st4(bind(source_is_a ? f.wb : f.wa, f.h), x, y, o);
        }
    }
    __syncthreads();  // the next tile of a grid-stride loop reuses sA
}

// ========================================================================================
// CANONICAL stages (Schied et al. 2017)
// ========================================================================================

// The wavelet intermediates of the canonical chain (W0..W4: colour, sigma) stay FP32 in both
// storage forms: the a-trous iterations are bound by arithmetic, not by these bytes, and an
// FP16 round trip per iteration would add five half-ulp errors to the output.
using WavS = F32S;

// Canonical stages use the hardware reciprocal / square root / reciprocal square root
// (v_rcp_f32, v_sqrt_f32, v_rsq_f32: about 1 ulp) instead of the correctly rounded
// multi-instruction sequences; they are checked against the double oracle like the rest.
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }

// Temporal reprojection + colour / moments accumulation. Writes W0 = (colour, sqrt(temporal
// variance)) into wavelet_a, the history length (uint8) and the moments; the batched form
// also keeps G for the next frame. Standard deviations rather than variances are stored so
// that heavy-tailed radiance stays inside the FP16 range.
// Pixel (x, y) of the frame, with its G, radiance R and motion m (unused without history).
template <class S, int SF>
__device__ __forceinline__ void can_temporal_pixel(const FrameDesc& f, const KParams& p, int x, int y, float4 G,
                                                   float4 R, float2 m)
{
    constexpr bool COMP = SF != SF_RAW, WIDE = SF == SF_WIDE;
    static_assert(!WIDE || S::B4 == 8, "the wide state is a binary16 form");
    float l = luma_fma(R.x, R.y, R.z);
    if (f.gsave.ptr) stg4(bind(f.gsave, f.h), x, y, G);

    float sumb = 0.0f, h0 = 0.0f, h1 = 0.0f, h2 = 0.0f, hm1 = 0.0f, hm2 = 0.0f;
    // Compensated form: per-tap (weight, mean, second moment), weight 0 for unused taps so
    // the loops below stay fully unrolled (no dynamically indexed private arrays).
    float tb[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tmu[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tm2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int nmax = 0;
    if (f.history_valid) {
        const Buf gp = bind(f.gprev, f.h), hp = bind(f.hist_prev, f.h), mp = bind(f.mom_prev, f.h);
        float flx = floorf(m.x), fly = floorf(m.y);
        float fx = m.x - flx, fy = m.y - fly;
        int x0 = x + (int)flx, y0 = y + (int)fly;
        // Issue the G / history / moment loads of all four taps at once (addresses clamped into
        // the frame; taps outside it are discarded below) so the reprojection costs one
        // dependent memory round trip instead of three. History and moments come as the pixel
        // pair (px, px + 1) of each row, px = clamp(x0): when x0 = -1 the right tap (x = 0) is
        // the pair's first pixel; every other in-frame tap is the pixel it names.
        // (wide form: the history words carry the residual codes and the length; the moments
        // are hi / lo pairs, two RGBA16F pixels per load)
        float4 Gq[4], H[4];
        float2 M[4];
        uint32_t NQ[4];
        const int px = clampi(x0, 0, f.w - 1);
        const bool right_first = x0 < 0;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            const int cy = clampi(y0 + jj, 0, f.h - 1);
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) Gq[jj * 2 + ii] = ldg4(gp, clampi(x0 + ii, 0, f.w - 1), cy);
            if constexpr (WIDE) {
                u32x2_t r0, r1;
                F16S::ld_raw4x2(hp, px, cy, right_first, r0, r1);
                H[jj * 2] = wide_history(r0, NQ[jj * 2]);
                H[jj * 2 + 1] = wide_history(r1, NQ[jj * 2 + 1]);
                F16S::ld_raw4x2(mp, px, cy, right_first, r0, r1);
                M[jj * 2] = wide_moments(r0);
                M[jj * 2 + 1] = wide_moments(r1);
            } else {
                S::ld4x2(hp, px, cy, right_first, H[jj * 2], H[jj * 2 + 1]);
                S::ld2x2(mp, px, cy, right_first, M[jj * 2], M[jj * 2 + 1]);
                NQ[jj * 2] = (uint32_t)(int)H[jj * 2].w;
                NQ[jj * 2 + 1] = (uint32_t)(int)H[jj * 2 + 1].w;
            }
        }
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                const int t = jj * 2 + ii;
                float b = (ii ? fx : 1.0f - fx) * (jj ? fy : 1.0f - fy);
                int qx = x0 + ii, qy = y0 + jj;
                if (!(b > 0.0f) || qx < 0 || qx >= f.w || qy < 0 || qy >= f.h) continue;
                float dz = fabsf(Gq[t].w - G.w);
                if (__builtin_fmaf(-p.depth_reject, G.w, dz) > 0.0f) continue;
                if (!dot_at_least(Gq[t], G, p.normal_reject)) continue;
                const int nq = (int)NQ[t];
                if (nq < 1) continue;
                sumb += b;
                h0 = fmaf(b, H[t].x, h0);
                h1 = fmaf(b, H[t].y, h1);
                h2 = fmaf(b, H[t].z, h2);
                nmax = max(nmax, nq);
                if (COMP) {
                    tb[t] = b;
                    tmu[t] = M[t].x;
                    tm2[t] = M[t].y * M[t].y;
                } else {
                    hm1 = fmaf(b, M[t].x, hm1);
                    hm2 = fmaf(b, M[t].y, hm2);
                }
            }
        }
    }
    int n_new = 1;
    float hmu = 0.0f, hvar = 0.0f;
    if (sumb > 0.0f) {
        n_new = min(nmax + 1, p.history_cap);
        float inv = rcp_hw(sumb);
        h0 *= inv;
        h1 *= inv;
        h2 *= inv;
        if (COMP) {
            float mu = 0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t) mu = fmaf(tb[t], tmu[t], mu);
            mu *= inv;
            float m2 = 0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                float d = tmu[t] - mu;
                m2 = fmaf(tb[t], fmaf(d, d, tm2[t]), m2);
            }
            hmu = mu;
            hvar = m2 * inv;
        } else {
            hm1 *= inv;
            hm2 *= inv;
        }
    } else {
        h0 = h1 = h2 = 0.0f;
        hm1 = hm2 = 0.0f;
    }
    float rn = rcp_hw((float)n_new);
    float ac = fmaxf(p.alpha_color, rn), am = fmaxf(p.alpha_moments, rn);
    float c0 = fmaf(ac, R.x - h0, h0), c1 = fmaf(ac, R.y - h1, h1), c2 = fmaf(ac, R.z - h2, h2);
    float var;
    const Buf mc = bind(f.mom_curr, f.h);
    if (COMP) {
        float d = l - hmu;
        float mu = fmaf(am, d, hmu);
        float m2 = (1.0f - am) * fmaf(am * d, d, hvar);
        if constexpr (WIDE) st_wide_moments(mc, x, y, mu, sqrt_hw(m2));
        else S::st2(mc, x, y, make_float2(mu, sqrt_hw(m2)));
        var = m2;
    } else {
        float mu1 = fmaf(am, l - hm1, hm1);
        float mu2 = fmaf(am, fmaf(l, l, -hm2), hm2);
        S::st2(mc, x, y, make_float2(mu1, mu2));
        var = fmaxf(mu2 - mu1 * mu1, 0.0f);
    }
    WavS::st4(bind(f.wa, f.h), x, y, make_float4(c0, c1, c2, sqrt_hw(var)));
    const Buf nl = bind(f.nlen, f.h);
    st_u8(nl, y * nl.pitch + x, (uint32_t)n_new);
}

template <class S, int SF, class L>
__device__ __forceinline__ void can_temporal_tile(const FrameDesc& f, const KParams& p, const Tile& T)
{
    const int tid = lane_tid();
    const int ti = tid & (LT - 1), tj = tid >> 4;
    const int x = T.i0 + ti, y = T.j0 + tj;
    if (x < f.w && y < f.h) {
        // G, radiance and motion in one memory round trip, the reprojection gathers in a
        // second. The depth gradient the strided a-trous iterations use is stored by iteration
        // 0, which has the neighbours' depths staged (the same floats).
        const float4 G = ldg4(bind(f.g, f.h), x, y);
        const float4 R = S::ld4(bind(f.rad, f.h), x, y);
        const float2 m = f.history_valid ? ldg2(bind(f.mv, f.h), x, y) : make_float2(0.0f, 0.0f);
        can_temporal_pixel<S, SF>(f, p, x, y, G, R, m);
    }
    // the tile functions of the persistent kernel's ticket loop all end at a barrier
    __syncthreads();
}

// Spatial variance of a young pixel (history length n < 4): the edge-aware variance of the
// temporal luminance over its 3x3 window (clamped at the frame border). A tap at offset o is
// weighted by 2^e with
//   e = -|z_o - z| / (phi_z |g.o| + eps) log2(e) + phi_n log2(n.n_o)   (n.n_o <= 0: weight 0)
// where g is the centre's depth gradient; each term only when its phi is positive. The
// denominator of o and -o is the same float, so five reciprocals serve the nine taps. The
// code is branch-free per tap (selects), so the taps' LDS reads overlap. tap(dx, dy) gives
// the LDS entry holding the window pixel.
template <class Tap>
__device__ __forceinline__ float young_variance(const KParams& p, const float4* sA, const float4* sB, float4 G,
                                                float gx, float gy, Tap tap)
{
    // o = (0,0), (1,0), (0,1), (1,1), (-1,1); a mirrored offset uses its partner's entry
    float nrd[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const float ox = i == 1 || i == 3 ? 1.0f : (i == 4 ? -1.0f : 0.0f), oy = i >= 2 ? 1.0f : 0.0f;
        nrd[i] = __builtin_amdgcn_rcpf(fmaf(p.phi_depth, fabsf(fmaf(gx, ox, gy * oy)), p.depth_eps));
    }
    const bool use_z = p.phi_depth > 0.0f, use_n = p.phi_normal > 0.0f;
    float tw[9], tl[9], sw = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const int t = (dy + 1) * 3 + dx + 1;
            const int mi = dy == 0 ? (dx == 0 ? 0 : 1) : (dx == 0 ? 2 : (dx == dy ? 3 : 4));
            const float4 Gq = sB[tap(dx, dy)];
            const float a = use_z ? -fabsf(Gq.w - G.w) * nrd[mi] * LOG2E : 0.0f;
            const float d = fmaf(G.x, Gq.x, fmaf(G.y, Gq.y, G.z * Gq.z));
            const float lg = __builtin_amdgcn_logf(d);
            const float e = use_n ? (d > 0.0f ? fmaf(p.phi_normal, lg, a) : -INFINITY) : a;
            const float w = __builtin_amdgcn_exp2f(e);
            tw[t] = w;
            tl[t] = sA[tap(dx, dy)].w;
            sw += w;
            s1 = fmaf(w, tl[t], s1);
        }
    }
    float var = 0.0f;
    if (sw > 0.0f) {
        float inv = rcp_hw(sw);
        float mean = s1 * inv, s2 = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            float d = tl[t] - mean;
            s2 = fmaf(tw[t] * d, d, s2);
        }
        var = s2 * inv;
    }
    return var;
}

// Per-pixel terms of an a-trous tap weight: luminance and depth of the centre, the luminance
// scale -log2(e) / sigma_l, the normal (zero and nb = 1 when the normal stop is off, so every
// tap's clamped dot product is 1) and the depth scales of the 12 offsets after the centre.
struct Centre {
    float lc, zc, nls, nx, ny, nz, nb;
    float nrd[12];
};

// C: centre (colour, luminance), Gc: centre (normal, depth), gv: prefiltered variance,
// (gx, gy): depth gradient, s: stride.
__device__ __forceinline__ Centre centre_of(const KParams& p, float4 C, float4 Gc, float gv, float gx, float gy, int s)
{
    Centre c;
    c.lc = C.w;
    c.zc = Gc.w;
    c.nls = p.luma_k * rsq_hw(fmaxf(gv, 0.0f) + p.variance_eps);
    const bool nrm_on = p.phi_normal > 0.0f;
    c.nx = nrm_on ? Gc.x : 0.0f;
    c.ny = nrm_on ? Gc.y : 0.0f;
    c.nz = nrm_on ? Gc.z : 0.0f;
    c.nb = nrm_on ? 0.0f : 1.0f;
    // -log2(e) / (phi_z |g.o| + eps) = 1 / (depth_e - |dx U + dy V|) with (U, V) = ln(2) phi_z s g,
    // for the 12 offsets after the centre in row-major order, (1,0) (2,0) (-2,1) (-1,1) (0,1)
    // (1,1) (2,1) (-2,2) (-1,2) (0,2) (1,2) (2,2); the mirrored offset has the bitwise-same
    // denominator. |dx U + dy V| = tk |tx| (the factors 2 are exact). All -0 when phi_z is 0.
    const float cs = p.depth_c * (float)s, U = gx * cs, V = gy * cs;
    const float VmU = V - U, UpV = U + V;
    const float tx[12] = {U,   U, fmaf(-2.0f, U, V), VmU, V, UpV, fmaf(2.0f, U, V),
                          VmU, fmaf(2.0f, V, -U),    V,   fmaf(2.0f, V, U), UpV};
    constexpr float tk[12] = {1.0f, 2.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 2.0f, 1.0f, 2.0f, 1.0f, 2.0f};
#pragma unroll
    for (int i = 0; i < 12; ++i) c.nrd[i] = __builtin_amdgcn_rcpf(fmaf(-tk[i], fabsf(tx[i]), p.depth_e));
    return c;
}

// Weight of the tap at offset (dx, dy) (compile-time after unrolling) with entry colour A and
// (normal, depth) B: K[dx] K[dy] 2^e with e the sum of the depth, luminance and normal terms.
__device__ __forceinline__ float tap_weight(const KParams& p, const Centre& c, int dx, int dy, float4 A, float4 B)
{
    const int o = (dy + 2) * 5 + dx + 2;  // 0..24, centre 12
    float e = log2k(dx + 2, dy + 2);
    if (o != 12) e = fmaf(fabsf(B.w - c.zc), c.nrd[o > 12 ? o - 13 : 11 - o], e);
    // the centre's own luminance difference is 0 (or NaN, and then so is every tap's, whose
    // weights all turn NaN and the result falls back to the centre as it would with it)
    if (o != 12) e = fmaf(fabsf(A.w - c.lc), c.nls, e);
    // max(0, n.n_q)^phi_n as phi_n log2(d) in the exponent, d clamped to [0, 1] by the dot
    // product's clamp bit: d <= 0 gives log2 = -inf, weight 0 (unit normals exceed 1 only by
    // rounding)
    const float d = __builtin_amdgcn_fmed3f(fmaf(c.nx, B.x, fmaf(c.ny, B.y, fmaf(c.nz, B.z, c.nb))), 0.0f, 1.0f);
    e = fmaf(p.phi_normal, __builtin_amdgcn_logf(d), e);
    return __builtin_amdgcn_exp2f(e);
}

// Weighted sums of a pixel: colour (3), weight and, unless VAR is false (the last iteration,
// whose output carries no variance), the squared-weight variance.
template <bool VAR>
struct Sums {
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sw = 0.0f, sv = 0.0f;
    __device__ __forceinline__ void add(float w, float4 A, float V)
    {
        s0 = fmaf(w, A.x, s0);
        s1 = fmaf(w, A.y, s1);
        s2 = fmaf(w, A.z, s2);
        sw += w;
        if constexpr (VAR) sv = fmaf(w * w, V, sv);
    }
    // The row's sums are complete here. Without this the optimiser sinks the weighted sums
    // below later rows' exponents and keeps those taps' weights and colours live (77 VGPRs
    // instead of 58 in the one-pixel filter, 6 waves per SIMD instead of 8, a quarter slower).
    __device__ __forceinline__ void pin()
    {
        if constexpr (VAR) asm volatile("" : "+v"(s0), "+v"(s1), "+v"(s2), "+v"(sw), "+v"(sv));
        else asm volatile("" : "+v"(s0), "+v"(s1), "+v"(s2), "+v"(sw));
    }
    // (colour, sigma) of the filtered pixel; the centre's own when the weights vanish (sigma 0
    // without VAR)
    __device__ __forceinline__ float4 result(float4 C, float Vc) const
    {
        if (sw > 1e-6f) {
            const float inv = rcp_hw(sw);
            return make_float4(s0 * inv, s1 * inv, s2 * inv, VAR ? sqrt_hw(sv) * inv : 0.0f);
        }
        return make_float4(C.x, C.y, C.z, VAR ? sqrt_hw(Vc) : 0.0f);
    }
};

// A-trous iteration k >= 1 (stride s = 2^k) on a direct tile: pixels T.i0 .. T.i0 + 255 of the
// frame (row-major), one per thread, each tap (and the variance prefilter's 3x3) loaded from
// memory at clamped coordinates, the values a pair tile stages for the same pixel; the same
// arithmetic in the same order as the staged filter below, so the outputs are the same floats.
// Used for frames whose stride-s lattice is tiny (svgf_sched.hpp, direct_lattice), where the
// lanes of neighbouring pixels load neighbouring taps.
template <class S, bool LAST>
__device__ __forceinline__ void can_atrous_direct(const FrameDesc& f, const KParams& p, const Tile& T, int k)
{
    const int s = 1 << k;
    const Buf src = bind((k & 1) ? f.wb : f.wa, f.h), gb = bind(f.g, f.h), gr = bind(f.grad, f.h);
    const int max_x = f.w - 1, max_y = f.h - 1;
    const int pix = T.i0 + lane_tid();
    if (pix < f.w * f.h) {
        const int y = pix / f.w, x = pix - y * f.w;
        auto qx = [&](int dx) { return clampi(x + dx * s, 0, max_x); };
        auto qy = [&](int dy) { return clampi(y + dy * s, 0, max_y); };
        // 3x3 binomial of the variance over the lattice neighbours
        float gv = 0.0f;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const float sd = WavS::ld4(src, qx(dx), qy(dy)).w;
                gv = fmaf((dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f), sd * sd, gv);
            }
        float4 C = WavS::ld4(src, x, y);
        const float Vc = C.w * C.w;
        C.w = luma_fma(C.x, C.y, C.z);
        const float2 g = ld_f2(gr, y * gr.pitch + x * 8);
        const Centre c = centre_of(p, C, ldg4(gb, x, y), gv, g.x, g.y, s);
        Sums<!LAST> a;
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
            for (int dx = -2; dx <= 2; ++dx) {
                float4 A = WavS::ld4(src, qx(dx), qy(dy));
                const float V = LAST ? 0.0f : A.w * A.w;
                A.w = luma_fma(A.x, A.y, A.z);
                a.add(tap_weight(p, c, dx, dy, A, ldg4(gb, qx(dx), qy(dy))), A, V);
            }
            a.pin();
        }
        const float4 o = a.result(C, Vc);
        if (LAST || k == p.iterations - 1) S::st4(bind(f.den, f.h), x, y, make_float4(o.x, o.y, o.z, 1.0f));
        else WavS::st4(bind((k & 1) ? f.wa : f.wb, f.h), x, y, o);
    }
    __syncthreads();  // the tile functions of the persistent kernel's ticket loop all end at a barrier
}

// A-trous iteration k (stride s = 2^k) on one lattice tile: luminance, depth and normal
// edge-stops, variance filtered by squared weights. Iteration 0 also writes the colour
// history, the last one the output.
// FIRST (iteration 0, stride 1) runs on a tall 16x32 pixel tile (TALL; a 16x16 tile for single
// small frames) and first replaces the temporal variance of young pixels (n < 4) by their
// spatial variance. The taps and the variance prefilter read the variance of every entry of
// the 20x36 (20x20) tap region, and a young entry's 3x3 window reaches one pixel further, so
// this iteration stages a 22x38 (22x22) region; the young entries of the tap region are listed
// and the block's threads compute their windows from LDS before the filter runs.
// NR = 2: the tile holds residues (rx, ry) and (rx + 1, ry) (strides >= 2): the region of each
// is staged (the pair's two pixels of an entry are loaded by neighbouring lanes, one cache line);
// each half of the block filters one residue.
// TALL and NR = 2 (blocked): a thread filters two vertically adjacent lattice points, whose 5x5
// taps share 4 of their 5 rows: 6 rows of LDS entries serve both. Otherwise (the 16x16 tiles
// of single small frames) a thread filters one lattice point.
// LAST: the caller's k is iterations - 1 >= 1; the output carries no variance, so the taps
// skip the squared-weight sum and its LDS reads. SF: the moment / history form (the wide form
// encodes the colour history written by FIRST).
template <class S, int SF, bool FIRST, bool LAST, int NR, bool TALL, bool EARLY, class L>
__device__ __forceinline__ void can_atrous_tile(const FrameDesc& f, const KParams& p, const Tile& T, int k)
{
    static_assert(!FIRST || NR == 1, "the first iteration runs at stride 1");
    static_assert(!FIRST || EARLY, "the first iteration loads its history lengths with the staged region");
    static_assert(!(FIRST && LAST), "the first iteration keeps its runtime last-iteration test");
    static_assert(!TALL || FIRST, "tall tiles are the first iteration's");
    constexpr bool BLK = TALL || NR == 2;
    constexpr int H = FIRST ? LHALO + 1 : LHALO;          // staged halo
    constexpr int TH = TALL ? 2 * LT : LT;                // tile height (lattice points)
    constexpr int RW = LT + 2 * H, RH = TH + 2 * H;       // staged region
    constexpr int RR = RW * RH;
    constexpr int QW = LT + 2 * LHALO, QH = TH + 2 * LHALO;  // tap region
    constexpr uint32_t NA = NR * RR, LIST = FIRST ? QW * QH : 0;
    static_assert(((NA * 36 + LIST * 2 + 15) & ~15u) + 16 == atrous_lds_bytes(FIRST, NR, TALL), "LDS layout");
    if constexpr (NR == 2) {
        if (T.rx < 0) {  // a direct tile (uniform across the block)
            can_atrous_direct<S, LAST>(f, p, T, k);
            return;
        }
    }
    char* const lds = L::template get<atrous_lds_bytes(FIRST, NR, TALL)>();
    float4* const sA = reinterpret_cast<float4*>(lds);                // (r, g, b, luminance), residue-major
    float4* const sB = reinterpret_cast<float4*>(lds + NA * 16);      // (normal, depth)
    float* const sV = reinterpret_cast<float*>(lds + NA * 32);        // variance
    uint16_t* const sList = reinterpret_cast<uint16_t*>(lds + NA * 36);  // young entries
    int& sCount = *reinterpret_cast<int*>(lds + ((NA * 36 + LIST * 2 + 15) & ~15u));
    const int s = 1 << k;
    const Buf src = bind((k & 1) ? f.wb : f.wa, f.h), gb = bind(f.g, f.h);
    const int max_x = f.w - 1, max_y = f.h - 1;
    const int tid = lane_tid();
    // FIRST: history lengths of the tap region's entries tid + 256 j, loaded together with the
    // staged region (no extra round trip)
    constexpr int NQ = FIRST ? (QW * QH + BLOCK - 1) / BLOCK : 0;
    uint32_t nq[NQ > 0 ? NQ : 1];
    // The thread's lattice points: (x, y0) and, in blocked tiles, (x, y1) with y1 = y0 + s.
    // Their depth gradients (strided iterations: the plane iteration 0 stores; iteration 0
    // computes them from the staged depths) or history lengths (iteration 0) are loaded at coordinates
    // clamped into the frame (the values of in-frame points). EARLY: together with the staged
    // region, whose loads all complete before its first LDS write, one memory round trip for
    // the whole tile. Otherwise after the barrier: at strides of 8 and more every lane's
    // gradient is a cache line of its own, and loading those lines beside the staging loads
    // (which there touch a line per entry) delays the staging more than the round trip costs
    // (measured; at strides 1 to 4 the early form is faster).
    const Buf gr = bind(f.grad, f.h), nl = bind(f.nlen, f.h);
    // residue res (NR = 2: tid / 128) and lattice point (ti, tj) of the tile
    const int res = NR == 2 ? tid >> 7 : 0, lt = NR == 2 ? tid & 127 : tid, ti = lt & (LT - 1);
    const int tj = BLK ? 2 * (lt >> 4) : lt >> 4;
    const int x = T.px(T.i0 + ti) + res, y0 = T.py(T.j0 + tj), y1 = y0 + s;
    const int cx = min(x, max_x), cy0 = min(y0, max_y), cy1 = min(y1, max_y);
    float2 g0 = make_float2(0.0f, 0.0f), g1 = make_float2(0.0f, 0.0f);
    uint32_t n0 = 0, n1 = 0;
    auto load_point_terms = [&] {
        if constexpr (FIRST) {
            n0 = ld_u8(nl, cy0 * nl.pitch + cx);
            if constexpr (BLK) n1 = ld_u8(nl, cy1 * nl.pitch + cx);
        } else {
            g0 = ld_f2(gr, cy0 * gr.pitch + cx * 8);
            if constexpr (BLK) g1 = ld_f2(gr, cy1 * gr.pitch + cx * 8);  // unused past the frame
        }
    };
    if constexpr (EARLY) load_point_terms();
    if constexpr (FIRST) {
        if (tid == 0) sCount = 0;
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int e = tid + j * BLOCK;
            const int qx = clampi(T.i0 - LHALO + e % QW, 0, max_x), qy = clampi(T.j0 - LHALO + e / QW, 0, max_y);
            nq[j] = e < QW * QH ? ld_u8(nl, qy * nl.pitch + qx) : 4u;
        }
    }
    {
        // every load of the region is issued before the first LDS write (one memory round trip);
        // the last pass's surplus lanes load the final entry again and store nothing
        constexpr uint32_t NE = NR * RR, NIT = (NE + BLOCK - 1) / BLOCK;
        float4 va[NIT], vb[NIT];
        uint32_t slot[NIT];  // LDS index of the entry
        const int bx = T.i0 - H, by = T.j0 - H;
#pragma unroll
        for (uint32_t j = 0; j < NIT; ++j) {
            // unsigned index arithmetic: divisions by the constant widths are a multiply-high
            const uint32_t e2 = j + 1 < NIT ? (uint32_t)tid + j * BLOCK : min((uint32_t)tid + j * BLOCK, NE - 1);
            const uint32_t r = e2 % NR, e = e2 / NR;  // neighbouring lanes: the two residues of one entry
            const uint32_t li = e % RW, lj = e / RW;
            slot[j] = r * RR + e;
            const int qx = clampi(T.rx + ((bx + (int)li) << T.log2s) + (int)r, 0, max_x);
            const int qy = clampi(T.ry + ((by + (int)lj) << T.log2s), 0, max_y);
            va[j] = WavS::ld4(src, qx, qy);
            vb[j] = ldg4(gb, qx, qy);
        }
        // EARLY: the loaded values are needed here; otherwise the optimiser sinks the last
        // pass's loads into its conditional store, a second memory round trip
        if constexpr (EARLY) {
#pragma unroll
            for (uint32_t j = 0; j < NIT; ++j)
                asm volatile("" ::"v"(va[j].x), "v"(va[j].y), "v"(va[j].z), "v"(va[j].w), "v"(vb[j].x),
                             "v"(vb[j].y), "v"(vb[j].z), "v"(vb[j].w));
        }
#pragma unroll
        for (uint32_t j = 0; j < NIT; ++j) {
            if (j + 1 < NIT || (uint32_t)tid + j * BLOCK < NE) {
                // the variance first, then the luminance into the loaded value's own alpha
                // register (a fresh float4 would copy r, g, b: three moves per entry)
                float4& v = va[j];
                sV[slot[j]] = v.w * v.w;
                v.w = luma_fma(v.x, v.y, v.z);
                sA[slot[j]] = v;
                sB[slot[j]] = vb[j];
            }
        }
    }
    if constexpr (FIRST) {
        lds_writes_done();
        __syncthreads();  // the staged region and sCount = 0
        // list the young entries: one LDS atomic per wave (its count), lanes placed by their
        // rank among the wave's young lanes
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const bool yng = nq[j] < 4u;
            const uint64_t m = __ballot(yng);
            int base = 0;
            if ((tid & 63) == 0 && m) base = atomicAdd(&sCount, (int)__popcll(m));
            base = __builtin_amdgcn_readfirstlane(base);
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (yng) sList[base + rank] = (uint16_t)(tid + j * BLOCK);
        }
        lds_writes_done();
        __syncthreads();
        const int count = sCount;
        for (int i = tid; i < count; i += BLOCK) {
            const int e = sList[i];
            const int qx = clampi(T.i0 - LHALO + e % QW, 0, max_x), qy = clampi(T.j0 - LHALO + e / QW, 0, max_y);
            // staged entry (li, lj) holds pixel (clamp(i0 - 3 + li), clamp(j0 - 3 + lj)), so
            // pixel (clamp(qx + dx), clamp(qy + dy)) sits at its offset from (i0 - 3, j0 - 3)
            const int ec = (qy - (T.j0 - H)) * RW + (qx - (T.i0 - H));
            auto tap = [&](int dx, int dy) {
                return (clampi(qy + dy, 0, max_y) - (T.j0 - H)) * RW + (clampi(qx + dx, 0, max_x) - (T.i0 - H));
            };
            // the depth gradient of the entry from the staged depths of its clamped neighbours
            // (the temporal stage's central differences, the same floats)
            const float gx = 0.5f * (sB[tap(1, 0)].w - sB[tap(-1, 0)].w), gy = 0.5f * (sB[tap(0, 1)].w - sB[tap(0, -1)].w);
            const float var = young_variance(p, sA, sB, sB[ec], gx, gy, tap);
            const float sd = sqrt_hw(var);  // the value W0 would hold
            sV[(e / QW + H - LHALO) * RW + e % QW + H - LHALO] = sd * sd;
        }
        lds_writes_done();
    }
    __syncthreads();
    // 3x3 binomial of the variance over the lattice neighbours (offsets +-s, in LDS)
    auto prefilter = [&](int c) {
        float gv = 0.0f;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
                gv = fmaf((dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f), sV[c + dy * RW + dx], gv);
        return gv;
    };
    if constexpr (!EARLY) load_point_terms();
    // FIRST also writes the colour history: the output with the pixel's history length
    auto store = [&](int x, int y, float4 o, uint32_t n) {
        if constexpr (FIRST) {
            if constexpr (SF == SF_WIDE) st_wide_history(bind(f.hist_curr, f.h), x, y, o, n);
            else S::st4(bind(f.hist_curr, f.h), x, y, make_float4(o.x, o.y, o.z, (float)n));
        }
        if (LAST || k == p.iterations - 1) S::st4(bind(f.den, f.h), x, y, make_float4(o.x, o.y, o.z, 1.0f));
        else WavS::st4(bind((k & 1) ? f.wa : f.wb, f.h), x, y, o);
    };
    if constexpr (!BLK) {
        const int y = y0;
        if (x <= max_x && y <= max_y) {
            const int c0 = (tj + H) * RW + ti + H;
            const float4 C = sA[c0];
            const float2 g = FIRST ? make_float2(0.5f * (sB[c0 + 1].w - sB[c0 - 1].w), 0.5f * (sB[c0 + RW].w - sB[c0 - RW].w))
                                   : g0;
            // iteration 0 keeps the gradient for the strided iterations
            if (FIRST && p.iterations > 1) st_f2(gr, y * gr.pitch + x * 8, g);
            const uint32_t n = n0;
            const Centre c = centre_of(p, C, sB[c0], prefilter(c0), g.x, g.y, s);
            Sums<!LAST> a;
#pragma unroll
            for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
                for (int dx = -2; dx <= 2; ++dx) {
                    const int q = c0 + dy * RW + dx;
                    const float4 A = sA[q];
                    a.add(tap_weight(p, c, dx, dy, A, sB[q]), A, LAST ? 0.0f : sV[q]);
                }
                a.pin();
            }
            store(x, y, a.result(C, sV[c0]), n);
        }
    } else {
        // lattice points (ti, tj) and (ti, tj + 1) of residue res
        if (x <= max_x && y0 <= max_y) {
            const int c0 = res * RR + (tj + H) * RW + ti + H, c1 = c0 + RW;
            const float4 C0 = sA[c0], C1 = sA[c1];
            if constexpr (FIRST) {
                // stride 1: the clamped neighbours are staged entries (the temporal stage's
                // central differences, the same floats)
                g0 = make_float2(0.5f * (sB[c0 + 1].w - sB[c0 - 1].w), 0.5f * (sB[c1].w - sB[c0 - RW].w));
                g1 = make_float2(0.5f * (sB[c1 + 1].w - sB[c1 - 1].w), 0.5f * (sB[c1 + RW].w - sB[c0].w));
                // kept for the strided iterations
                if (p.iterations > 1) {
                    st_f2(gr, y0 * gr.pitch + x * 8, g0);
                    if (y1 <= max_y) st_f2(gr, y1 * gr.pitch + x * 8, g1);
                }
            }
            const Centre e0 = centre_of(p, C0, sB[c0], prefilter(c0), g0.x, g0.y, s);
            const Centre e1 = centre_of(p, C1, sB[c1], prefilter(c1), g1.x, g1.y, s);
            Sums<!LAST> a0, a1;
#pragma unroll
            for (int j = -2; j <= 3; ++j) {
#pragma unroll
                for (int dx = -2; dx <= 2; ++dx) {
                    const int q = c0 + j * RW + dx;
                    const float4 A = sA[q], B = sB[q];
                    const float V = LAST ? 0.0f : sV[q];
                    if (j <= 2) a0.add(tap_weight(p, e0, dx, j, A, B), A, V);
                    if (j >= -1) a1.add(tap_weight(p, e1, dx, j - 1, A, B), A, V);
                }
                a0.pin();
                a1.pin();
            }
            store(x, y0, a0.result(C0, sV[c0]), n0);
            if (y1 <= max_y) store(x, y1, a1.result(C1, sV[c1]), n1);
        }
    }
    __syncthreads();  // the next tile of a grid-stride loop reuses the LDS arrays
}

// ========================================================================================
// Kernels
// ========================================================================================
enum StageId : int {
    ST_REF_INIT = 0, ST_REF_ATROUS = 1, ST_CAN_TEMPORAL = 2, ST_CAN_ATROUS0 = 3, ST_CAN_ATROUS = 4, ST_CAN_ATROUS_PAIR = 5,
    ST_CAN_ATROUS0_SQ = 6,  // the first a-trous iteration on 16x16 tiles (single small frames)
    // the last iteration (k = iterations - 1 >= 1): writes the output, no filtered variance
    ST_CAN_ATROUS_L = 7, ST_CAN_ATROUS_PAIR_L = 8,
    // residue pairs at strides 2 and 4 (not the last iteration): the per-point loads travel
    // with the staged region
    ST_CAN_ATROUS_PAIR_E = 9
};
constexpr bool is_pair_stage(int st)
{
    return st == ST_CAN_ATROUS_PAIR || st == ST_CAN_ATROUS_PAIR_L || st == ST_CAN_ATROUS_PAIR_E;
}

template <class S, int STAGE, int SF, class L = OwnLds>
__device__ __forceinline__ void run_stage(const FrameDesc& f, const KParams& p, const Tile& T, int arg)
{
    if constexpr (STAGE == ST_REF_INIT) ref_init_tile<S>(f, p, T);
    else if constexpr (STAGE == ST_REF_ATROUS) ref_atrous_tile<S, L>(f, p, T, arg);
    else if constexpr (STAGE == ST_CAN_TEMPORAL) can_temporal_tile<S, SF, L>(f, p, T);
    else if constexpr (STAGE == ST_CAN_ATROUS0) can_atrous_tile<S, SF, true, false, 1, true, true, L>(f, p, T, 0);
    else if constexpr (STAGE == ST_CAN_ATROUS0_SQ) can_atrous_tile<S, SF, true, false, 1, false, true, L>(f, p, T, 0);
    else if constexpr (STAGE == ST_CAN_ATROUS) can_atrous_tile<S, SF, false, false, 1, false, false, L>(f, p, T, arg);
    else if constexpr (STAGE == ST_CAN_ATROUS_L) can_atrous_tile<S, SF, false, true, 1, false, false, L>(f, p, T, arg);
    else if constexpr (STAGE == ST_CAN_ATROUS_PAIR_L) can_atrous_tile<S, SF, false, true, 2, false, false, L>(f, p, T, arg);
    else if constexpr (STAGE == ST_CAN_ATROUS_PAIR_E) can_atrous_tile<S, SF, false, false, 2, false, true, L>(f, p, T, arg);
    else can_atrous_tile<S, SF, false, false, 2, false, false, L>(f, p, T, arg);
}

// Frame descriptor of a batch entry into scalar registers (the index is uniform across the block).
__device__ __forceinline__ FrameDesc load_frame(const FrameDesc* frames, uint32_t i)
{
    FrameDesc f;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(frames + i);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&f);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(FrameDesc) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    return f;
}

// The tile of stage STAGE that logical workgroup t of a frame runs.
template <int STAGE>
__device__ __forceinline__ Tile stage_tile(int w, int h, uint32_t log2s, uint32_t t)
{
    if constexpr (is_pair_stage(STAGE)) return decode_pair_tile(w, h, log2s, t);
    else if constexpr (STAGE == ST_CAN_ATROUS0) return decode_tall_tile(w, t);
    else return decode_tile(w, log2s, t);
}

// arg: REFERENCE a-trous pass (1..N) or CANONICAL a-trous iteration (0..N-1); log2s: stride.
template <class S, int STAGE, int SF>
__global__ __launch_bounds__(BLOCK) void k_frame(FrameDesc f, KParams p, int arg, uint32_t log2s)
{
    Tile T = stage_tile<STAGE>(f.w, f.h, log2s, xcd_logical(blockIdx.x, gridDim.x));
    run_stage<S, STAGE, SF>(f, p, T, arg);
}

// One block per tile of the batch (the host sizes the grid to the stage's tile count; a
// grid-stride loop here would keep a second tile's state live and double the VGPR count).
template <class S, int STAGE, int SF>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchTable bt, KParams p, int arg, uint32_t log2s, int stage)
{
    const uint32_t nf = bt.hdr[HDR_FRAMES];
    const uint32_t* starts = bt.starts + (uint32_t)stage * STARTS_ROW;
    const uint32_t t = xcd_logical(blockIdx.x, gridDim.x);
    if (t >= starts[nf]) return;
    const uint32_t fi = frame_of(starts, nf, t);
    const FrameDesc f = load_frame(bt.frames, fi);
    Tile T = stage_tile<STAGE>(f.w, f.h, log2s, t - starts[fi]);
    run_stage<S, STAGE, SF>(f, p, T, arg);
}

// One tile of a persistent launch. Each stage's branch loads the frame descriptor itself: one
// load ahead of the branches would keep every descriptor field live across all of them and raise
// the register count. The stage tiles share one LDS arena (ArenaLds).
template <class S, int STAGE, int SF>
__device__ __forceinline__ void persist_tile(const FrameDesc* frames, uint32_t fi, const KParams& p, uint32_t log2s,
                                          uint32_t tile, int arg)
{
    const FrameDesc f = load_frame(frames, fi);
    constexpr bool CANON = STAGE != ST_REF_INIT && STAGE != ST_REF_ATROUS;
    run_stage<S, STAGE, SF, ArenaLds<CANON>>(f, p, stage_tile<STAGE>(f.w, f.h, log2s, tile), arg);
}

// Persistent form: every stage of a batch in one launch. Each workgroup takes chunks of
// consecutive tickets from ctl->ticket until none is left (order: svgf_sched.hpp; a chunk's
// tiles are neighbours, which then share one XCD's L2). Stage i > 0 may start once every
// tile of stage i - 1 (all frames) is done; a workgroup waits for that once per stage, then
// acquires (every wave invalidates its caches' stale lines of the previous stage's outputs).
// Finished tiles are published lazily: a workgroup counts them and publishes the count (every
// wave makes its stores visible at device scope, then one atomic add) when its next ticket
// belongs to another stage or none is left - always before it waits. A workgroup that waits
// holds no unpublished tile and only waits for tickets smaller than its own, so the holder of
// the smallest unfinished ticket can always run: no deadlock, whether or not all workgroups
// are resident at once. Per-stage (not per-frame) dependencies also mean that when stage i
// starts no read of a stage < i - 1 is still in flight, so nothing refills a cache with stale
// lines after the acquire. A wait of more than max_polls polls (about 1.3 us each) gives up,
// sets *err (host memory) and ctl->abort, and skips its tile; once ctl->abort is set every
// wait gives up at once, so the kernel always ends soon after. The last workgroup to find
// the tickets exhausted clears the counters, leaving ctl zeroed for the next launch.
template <class S, bool CANON, int SF>
__global__ __launch_bounds__(BLOCK) void k_persist(BatchTable bt, KParams p, PersistCtl* ctl, uint32_t* err,
                                                   uint64_t max_polls)
{
    // Thread 0 writes the next chunk index into the slot the other waves are not reading; the
    // barrier at the top of the loop orders both uses of a slot.
    __shared__ uint32_t sChunk[2];
    __shared__ int sFlag;
    const int tid = (int)threadIdx.x;
    const uint32_t* hdr = bt.hdr;
    const uint32_t nstages = hdr[HDR_STAGES];
    const uint32_t total = hdr[HDR_BASE + nstages];
    const uint32_t chunk = hdr[HDR_CHUNK];
    // Thread 0 takes a chunk in one block at the end of the loop body (the first one before
    // the loop); a `tid == 0` block on each side of the back edge would be threaded into one
    // and give thread 0 a loop path of its own, so the other lanes would reach the next barrier
    // without it.
    if (tid == 0) {
        sChunk[0] = __hip_atomic_fetch_add(&ctl->ticket[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lds_writes_done();
    }
    uint32_t slot = 0, t = 0, end = 0;        // tickets [t, end) of the current chunk
    uint32_t pending = 0, pending_stage = 0;  // finished, unpublished tiles (all of one stage)
    uint32_t ready_stage = 0;                 // highest stage whose inputs this workgroup has acquired
    for (;;) {
        __syncthreads();
        if (t == end) {
            const uint32_t c = __builtin_amdgcn_readfirstlane(sChunk[slot]);
            t = c < total / chunk + 1 ? c * chunk : total;  // (no overflow of c * chunk)
            end = min(t + chunk, total);
        }
        const bool more = t < total;
        Ticket k{0, 0, 0};
        if (more) k = decode_ticket(hdr, bt.starts, t);
        if (pending != 0 && (!more || k.stage != pending_stage)) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (tid == 0)
                __hip_atomic_fetch_add(&ctl->done[pending_stage * CTL_LINE], pending, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
            pending = 0;
        }
        if (!more) break;
        bool run = true;
        if (k.stage > ready_stage) {
            if (tid == 0) {
                const uint32_t need = hdr[HDR_BASE + k.stage] - hdr[HDR_BASE + k.stage - 1];
                uint32_t* c = &ctl->done[(k.stage - 1) * CTL_LINE];
                int ok = 1;
                // bounded by a number of polls, not by a clock: a wave that is swapped out
                // (queue preemption) does not poll, so the bound cannot expire while it is away
                for (uint64_t polls = 0; __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need; ++polls) {
                    if (polls >= max_polls ||
                        ((polls & 7) == 7 && __hip_atomic_load(&ctl->abort[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(16);
                }
                if (!ok) {
                    __hip_atomic_store(&ctl->abort[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    // a plain store (no read-modify-write over the host link): only 1 is written
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
                sFlag = ok;
                lds_writes_done();
            }
            __syncthreads();
            // every wave: later loads must not hit lines cached before the producers' stores
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            run = __builtin_amdgcn_readfirstlane(sFlag) != 0;
            if (run) ready_stage = k.stage;
        }
        if (run) {
            if constexpr (CANON) {
                // stage i >= 2 is a-trous iteration i - 1 (residue pairs from stride 2^PAIR_LOG2S);
                // the last one writes the output without the filtered variance
                const bool single = PAIR_LOG2S > 1 && k.stage - 1 < PAIR_LOG2S, last = k.stage == nstages - 1;
                if (k.stage == 0) persist_tile<S, ST_CAN_TEMPORAL, SF>(bt.frames, k.frame, p, 0, k.tile, 0);
                else if (k.stage == 1) persist_tile<S, ST_CAN_ATROUS0, SF>(bt.frames, k.frame, p, 0, k.tile, 0);
                else if (single && !last)
                    persist_tile<S, ST_CAN_ATROUS, SF>(bt.frames, k.frame, p, k.stage - 1, k.tile, (int)k.stage - 1);
                else if (single)
                    persist_tile<S, ST_CAN_ATROUS_L, SF>(bt.frames, k.frame, p, k.stage - 1, k.tile, (int)k.stage - 1);
                else if (!last)
                    persist_tile<S, ST_CAN_ATROUS_PAIR, SF>(bt.frames, k.frame, p, k.stage - 1, k.tile, (int)k.stage - 1);
                else persist_tile<S, ST_CAN_ATROUS_PAIR_L, SF>(bt.frames, k.frame, p, k.stage - 1, k.tile, (int)k.stage - 1);
            } else {
                if (k.stage == 0) persist_tile<S, ST_REF_INIT, SF_RAW>(bt.frames, k.frame, p, 0, k.tile, 0);
                else persist_tile<S, ST_REF_ATROUS, SF_RAW>(bt.frames, k.frame, p, k.stage - 1, k.tile, (int)k.stage);
            }
        }
        ++pending;
        pending_stage = k.stage;
        if (++t == end) {
            slot ^= 1u;
            if (tid == 0) {
                sChunk[slot] = __hip_atomic_fetch_add(&ctl->ticket[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                lds_writes_done();
            }
        }
    }
    if (tid == 0) {
        sFlag = __hip_atomic_fetch_add(&ctl->exited[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
        lds_writes_done();
    }
    __syncthreads();
    if (sFlag) {
        // every other workgroup has exited: nothing reads the counters any more
        if (tid < (int)nstages) __hip_atomic_store(&ctl->done[tid * CTL_LINE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tid == 0) {
            __hip_atomic_store(&ctl->ticket[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->exited[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->abort[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}
template <class S, int STAGE, int SF>
void launch_frame(const FrameDesc& f, const KParams& p, int arg, uint32_t log2s, hipStream_t stream)
{
    const uint32_t w = (uint32_t)f.w, h = (uint32_t)f.h;
    const uint32_t n = is_pair_stage(STAGE)          ? pair_tiles_at(w, h, log2s)
                       : STAGE == ST_CAN_ATROUS0   ? tall_tiles_at(w, h)
                                                   : tiles_at(w, h, log2s);
    hipLaunchKernelGGL((k_frame<S, STAGE, SF>), dim3(n), dim3(BLOCK), 0, stream, f, p, arg, log2s);
}

template <class S, int STAGE, int SF>
void launch_batch(const BatchTable& bt, const KParams& p, int arg, uint32_t log2s, int stage, const uint32_t* grids,
                  hipStream_t stream)
{
    hipLaunchKernelGGL((k_batch<S, STAGE, SF>), dim3(grids[stage]), dim3(BLOCK), 0, stream, bt, p, arg, log2s, stage);
}

inline void mark(hipEvent_t* ev, int i, hipStream_t s)
{
    if (ev) (void)hipEventRecord(ev[i], s);
}

template <class S>
void reference_frame(const FrameDesc& f, const KParams& p, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_frame<S, ST_REF_INIT, SF_RAW>(f, p, 0, 0, s);
    mark(ev, 1, s);
    for (int pass = 1; pass <= p.iterations; ++pass) {
        launch_frame<S, ST_REF_ATROUS, SF_RAW>(f, p, pass, (uint32_t)(pass - 1), s);
        mark(ev, pass + 1, s);
    }
}

template <class S, int SF>
void canonical_frame(const FrameDesc& f, const KParams& p, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_frame<S, ST_CAN_TEMPORAL, SF>(f, p, 0, 0, s);
    mark(ev, 1, s);
    // A single small frame has few tiles: halving them (tall tiles, residue pairs) would leave
    // the device emptier, so both are used from 2048 16x16 single-residue tiles on.
    if (tiles_at((uint32_t)f.w, (uint32_t)f.h, 0) < 2048u) launch_frame<S, ST_CAN_ATROUS0_SQ, SF>(f, p, 0, 0, s);
    else launch_frame<S, ST_CAN_ATROUS0, SF>(f, p, 0, 0, s);
    mark(ev, 2, s);
    for (int k = 1; k < p.iterations; ++k) {
        const bool last = k == p.iterations - 1;
        if (k < (int)PAIR_LOG2S || tiles_at((uint32_t)f.w, (uint32_t)f.h, (uint32_t)k) < 2048u) {
            if (last) launch_frame<S, ST_CAN_ATROUS_L, SF>(f, p, k, (uint32_t)k, s);
            else launch_frame<S, ST_CAN_ATROUS, SF>(f, p, k, (uint32_t)k, s);
        } else {
            if (last) launch_frame<S, ST_CAN_ATROUS_PAIR_L, SF>(f, p, k, (uint32_t)k, s);
            else if (k <= 2) launch_frame<S, ST_CAN_ATROUS_PAIR_E, SF>(f, p, k, (uint32_t)k, s);
            else launch_frame<S, ST_CAN_ATROUS_PAIR, SF>(f, p, k, (uint32_t)k, s);
        }
        mark(ev, k + 2, s);
    }
}

template <class S>
void reference_batch(const BatchTable& bt, const KParams& p, const uint32_t* g, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_batch<S, ST_REF_INIT, SF_RAW>(bt, p, 0, 0, 0, g, s);
    mark(ev, 1, s);
    for (int pass = 1; pass <= p.iterations; ++pass) {
        launch_batch<S, ST_REF_ATROUS, SF_RAW>(bt, p, pass, (uint32_t)(pass - 1), pass, g, s);
        mark(ev, pass + 1, s);
    }
}

template <class S, int SF>
void canonical_batch(const BatchTable& bt, const KParams& p, const uint32_t* g, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_batch<S, ST_CAN_TEMPORAL, SF>(bt, p, 0, 0, 0, g, s);
    mark(ev, 1, s);
    launch_batch<S, ST_CAN_ATROUS0, SF>(bt, p, 0, 0, 1, g, s);
    mark(ev, 2, s);
    for (int k = 1; k < p.iterations; ++k) {
        const bool last = k == p.iterations - 1;
        if (k < (int)PAIR_LOG2S) {
            if (last) launch_batch<S, ST_CAN_ATROUS_L, SF>(bt, p, k, (uint32_t)k, k + 1, g, s);
            else launch_batch<S, ST_CAN_ATROUS, SF>(bt, p, k, (uint32_t)k, k + 1, g, s);
        } else {
            if (last) launch_batch<S, ST_CAN_ATROUS_PAIR_L, SF>(bt, p, k, (uint32_t)k, k + 1, g, s);
            else if (k <= 2) launch_batch<S, ST_CAN_ATROUS_PAIR_E, SF>(bt, p, k, (uint32_t)k, k + 1, g, s);
            else launch_batch<S, ST_CAN_ATROUS_PAIR, SF>(bt, p, k, (uint32_t)k, k + 1, g, s);
        }
        mark(ev, k + 2, s);
    }
}

template <class S, bool CANON, int SF>
hipError_t persist_launch(const BatchTable& bt, const KParams& p, PersistCtl* ctl, uint32_t* err, uint64_t max_polls,
                          uint32_t blocks, hipStream_t s)
{
    hipLaunchKernelGGL((k_persist<S, CANON, SF>), dim3(blocks), dim3(BLOCK), 0, s, bt, p, ctl, err, max_polls);
    return hipGetLastError();
}

template <class S, bool CANON, int SF>
int persist_occupancy()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persist<S, CANON, SF>, BLOCK, 0) != hipSuccess) return 0;
    return n;
}

}  // namespace

hipError_t launch_reference_frame(const FrameDesc& f, const KParams& p, Storage st, hipStream_t stream, hipEvent_t* ev)
{
    if (st == Storage::F32) reference_frame<F32S>(f, p, stream, ev);
    else reference_frame<F16S>(f, p, stream, ev);
    return hipGetLastError();
}

hipError_t launch_canonical_frame(const FrameDesc& f, const KParams& p, Storage st, hipStream_t stream, hipEvent_t* ev)
{
    const int sf = state_form(p);
    if (st == Storage::F32) {
        if (sf == SF_COMP) canonical_frame<F32S, SF_COMP>(f, p, stream, ev);
        else canonical_frame<F32S, SF_RAW>(f, p, stream, ev);
    } else {
        if (sf == SF_WIDE) canonical_frame<F16S, SF_WIDE>(f, p, stream, ev);
        else if (sf == SF_COMP) canonical_frame<F16S, SF_COMP>(f, p, stream, ev);
        else canonical_frame<F16S, SF_RAW>(f, p, stream, ev);
    }
    return hipGetLastError();
}

hipError_t launch_reference_batch(const BatchTable& bt, const KParams& p, Storage st, const uint32_t* grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev)
{
    if (st == Storage::F32) reference_batch<F32S>(bt, p, grid_blocks, stream, ev);
    else reference_batch<F16S>(bt, p, grid_blocks, stream, ev);
    return hipGetLastError();
}

hipError_t launch_canonical_batch(const BatchTable& bt, const KParams& p, Storage st, const uint32_t* grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev)
{
    const int sf = state_form(p);
    if (st == Storage::F32) {
        if (sf == SF_COMP) canonical_batch<F32S, SF_COMP>(bt, p, grid_blocks, stream, ev);
        else canonical_batch<F32S, SF_RAW>(bt, p, grid_blocks, stream, ev);
    } else {
        if (sf == SF_WIDE) canonical_batch<F16S, SF_WIDE>(bt, p, grid_blocks, stream, ev);
        else if (sf == SF_COMP) canonical_batch<F16S, SF_COMP>(bt, p, grid_blocks, stream, ev);
        else canonical_batch<F16S, SF_RAW>(bt, p, grid_blocks, stream, ev);
    }
    return hipGetLastError();
}

hipError_t launch_batch_persistent(int canonical, const BatchTable& bt, const KParams& p, Storage st, PersistCtl* ctl,
                                   uint32_t* err, uint64_t max_polls, uint32_t blocks, hipStream_t stream)
{
    if (!canonical) {
        if (st == Storage::F32) return persist_launch<F32S, false, SF_RAW>(bt, p, ctl, err, max_polls, blocks, stream);
        return persist_launch<F16S, false, SF_RAW>(bt, p, ctl, err, max_polls, blocks, stream);
    }
    const int sf = state_form(p);
    if (st == Storage::F32) {
        if (sf == SF_COMP) return persist_launch<F32S, true, SF_COMP>(bt, p, ctl, err, max_polls, blocks, stream);
        return persist_launch<F32S, true, SF_RAW>(bt, p, ctl, err, max_polls, blocks, stream);
    }
    if (sf == SF_WIDE) return persist_launch<F16S, true, SF_WIDE>(bt, p, ctl, err, max_polls, blocks, stream);
    if (sf == SF_COMP) return persist_launch<F16S, true, SF_COMP>(bt, p, ctl, err, max_polls, blocks, stream);
    return persist_launch<F16S, true, SF_RAW>(bt, p, ctl, err, max_polls, blocks, stream);
}

int persistent_blocks_per_cu(int canonical, Storage st, int sf)
{
    if (!canonical) return st == Storage::F32 ? persist_occupancy<F32S, false, SF_RAW>() : persist_occupancy<F16S, false, SF_RAW>();
    if (st == Storage::F32) return sf == SF_COMP ? persist_occupancy<F32S, true, SF_COMP>() : persist_occupancy<F32S, true, SF_RAW>();
    if (sf == SF_WIDE) return persist_occupancy<F16S, true, SF_WIDE>();
    return sf == SF_COMP ? persist_occupancy<F16S, true, SF_COMP>() : persist_occupancy<F16S, true, SF_RAW>();
}

}  // namespace svgf
