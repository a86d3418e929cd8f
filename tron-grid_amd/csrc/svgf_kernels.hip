// svgf_kernels.hip — SVGF stages as gfx950 kernels.
//
// Every stage is a per-pixel __device__ function; the single-frame kernels (2D grid over
// one frame's 64x4 tiles) and the batched kernels (grid-stride loop over a ragged tile table
// spanning many frames) call the same function, so a frame gives bit-identical results
// alone and inside a batch.
//
// REFERENCE stages follow src/svgf_denoise.slang:152-275 operation by operation (no
// contraction where the shader's rounding matters: luminance, lerp, variance). CANONICAL
// stages follow Schied et al. 2017 with FMA contraction (they are checked against the
// double-precision oracle, not against a float restatement).
#include "svgf_internal.hpp"

namespace svgf {
namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ char* row_of(const PlaneDev& p, int y) { return p.ptr + (size_t)y * p.pitch; }

// Storage traits: FP32 planes (float4 / float2) or FP16 planes (half4 / half2), float math.
struct F32S {
    __device__ static __forceinline__ float4 ld4(const PlaneDev& p, int x, int y)
    {
        return reinterpret_cast<const float4*>(row_of(p, y))[x];
    }
    __device__ static __forceinline__ void st4(const PlaneDev& p, int x, int y, float4 v)
    {
        reinterpret_cast<float4*>(row_of(p, y))[x] = v;
    }
    __device__ static __forceinline__ float2 ld2(const PlaneDev& p, int x, int y)
    {
        return reinterpret_cast<const float2*>(row_of(p, y))[x];
    }
    __device__ static __forceinline__ void st2(const PlaneDev& p, int x, int y, float2 v)
    {
        reinterpret_cast<float2*>(row_of(p, y))[x] = v;
    }
    __device__ static __forceinline__ float lda(const PlaneDev& p, int x, int y)
    {
        return reinterpret_cast<const float*>(row_of(p, y))[4 * x + 3];
    }
};

struct F16S {
    __device__ static __forceinline__ float4 ld4(const PlaneDev& p, int x, int y)
    {
        half4_t h = reinterpret_cast<const half4_t*>(row_of(p, y))[x];
        return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
    }
    __device__ static __forceinline__ void st4(const PlaneDev& p, int x, int y, float4 v)
    {
        half4_t h = {(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
        reinterpret_cast<half4_t*>(row_of(p, y))[x] = h;
    }
    __device__ static __forceinline__ float2 ld2(const PlaneDev& p, int x, int y)
    {
        half2_t h = reinterpret_cast<const half2_t*>(row_of(p, y))[x];
        return make_float2((float)h.x, (float)h.y);
    }
    __device__ static __forceinline__ void st2(const PlaneDev& p, int x, int y, float2 v)
    {
        half2_t h = {(_Float16)v.x, (_Float16)v.y};
        reinterpret_cast<half2_t*>(row_of(p, y))[x] = h;
    }
    __device__ static __forceinline__ float lda(const PlaneDev& p, int x, int y)
    {
        return (float)reinterpret_cast<const _Float16*>(row_of(p, y))[4 * x + 3];
    }
};

__device__ __forceinline__ float4 ldg4(const PlaneDev& p, int x, int y)
{
    return reinterpret_cast<const float4*>(row_of(p, y))[x];
}
__device__ __forceinline__ float ldgz(const PlaneDev& p, int x, int y)
{
    return reinterpret_cast<const float*>(row_of(p, y))[4 * x + 3];
}
__device__ __forceinline__ void stg4(const PlaneDev& p, int x, int y, float4 v)
{
    reinterpret_cast<float4*>(row_of(p, y))[x] = v;
}
__device__ __forceinline__ float2 ldg2(const PlaneDev& p, int x, int y)
{
    return reinterpret_cast<const float2*>(row_of(p, y))[x];
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

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

// exp(x) as the hardware exp2 of a prescaled argument (v_exp_f32); relative error about
// |x| * 6e-8 from the prescale rounding.
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

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

// ========================================================================================
// REFERENCE stages (svgf_denoise.slang)
// ========================================================================================

// Pass 0 (svgf_denoise.slang:176-191): moments EMA, variance, sqrt into wavelet_a alpha.
template <class S>
__device__ __forceinline__ void ref_init_px(const FrameDesc& f, const KParams& p, int x, int y)
{
#pragma clang fp contract(off)
    float4 L = S::ld4(f.rad, x, y);
    float l = luma_ref(L.x, L.y, L.z);
    float m1 = 0.0f, m2 = 0.0f;
    if (f.history_valid) {
        float2 m = S::ld2(f.mom_prev, x, y);
        m1 = m.x;
        m2 = m.y;
    }
    float mu1 = m1 + p.alpha_moments * (l - m1);
    float mu2 = m2 + p.alpha_moments * (l * l - m2);
    float var = fmaxf(mu2 - mu1 * mu1, 0.0f);
    S::st2(f.mom_curr, x, y, make_float2(mu1, mu2));
    S::st4(f.wa, x, y, make_float4(L.x, L.y, L.z, __builtin_sqrtf(var + p.variance_eps)));
}

// Passes 1..N (svgf_denoise.slang:209-274): 5x5 B3-spline taps at `stride`, clamp-to-edge,
// weight K[dx] K[dy] exp(-|l_q - l_c| / max(phi * sigma_c, eps)), dy outer / dx inner.
template <class S>
__device__ __forceinline__ void ref_atrous_px(const FrameDesc& f, const KParams& p, int x, int y, int pass)
{
    const int stride = 1 << (pass - 1);
    const bool source_is_a = (pass & 1) == 1;
    const PlaneDev& src = source_is_a ? f.wa : f.wb;
    const bool final_pass = pass == p.iterations;
    const int max_x = f.w - 1, max_y = f.h - 1;

    float4 c = S::ld4(src, x, y);
    float lc = luma_ref(c.x, c.y, c.z);
    float sigma = fmaxf(p.phi_luma * c.w, p.variance_eps);
    float nrs = -1.0f / sigma;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, ws = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        int sy = clampi(y + dy * stride, 0, max_y);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            int sx = clampi(x + dx * stride, 0, max_x);
            float4 v = S::ld4(src, sx, sy);
            float l = luma_ref(v.x, v.y, v.z);
            float tw = (K5[dx + 2] * K5[dy + 2]) * fast_exp(fabsf(l - lc) * nrs);
            s0 = fmaf(v.x, tw, s0);
            s1 = fmaf(v.y, tw, s1);
            s2 = fmaf(v.z, tw, s2);
            ws += tw;
        }
    }
    float4 o;
    if (ws > 1e-6f) {
        o = make_float4(s0 / ws, s1 / ws, s2 / ws, c.w);
    } else {
        o = make_float4(c.x, c.y, c.z, c.w);
    }
    if (final_pass) {
        o.w = 1.0f;
        S::st4(f.den, x, y, o);
    } else {
        S::st4(source_is_a ? f.wb : f.wa, x, y, o);
    }
}

// ========================================================================================
// CANONICAL stages (Schied et al. 2017)
// ========================================================================================

// Temporal reprojection + colour / moments accumulation. Writes ctmp = (colour, temporal
// variance), nlen = history length, mom_curr; the batched form also keeps G for next frame.
template <class S, bool COMP>
__device__ __forceinline__ void can_temporal_px(const FrameDesc& f, const KParams& p, int x, int y)
{
    float4 R = S::ld4(f.rad, x, y);
    float l = luma_fma(R.x, R.y, R.z);
    float4 G = ldg4(f.g, x, y);
    if (f.gsave.ptr) stg4(f.gsave, x, y, G);

    float sumb = 0.0f, h0 = 0.0f, h1 = 0.0f, h2 = 0.0f, hm1 = 0.0f, hm2 = 0.0f;
    // Compensated form: per-tap (weight, mean, second moment), weight 0 for unused taps so
    // the loops below stay fully unrolled (no dynamically indexed private arrays).
    float tb[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tmu[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tm2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int nmax = 0;
    if (f.history_valid) {
        float2 m = ldg2(f.mv, x, y);
        float flx = floorf(m.x), fly = floorf(m.y);
        float fx = m.x - flx, fy = m.y - fly;
        int x0 = x + (int)flx, y0 = y + (int)fly;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                float b = (ii ? fx : 1.0f - fx) * (jj ? fy : 1.0f - fy);
                int qx = x0 + ii, qy = y0 + jj;
                if (!(b > 0.0f) || qx < 0 || qx >= f.w || qy < 0 || qy >= f.h) continue;
                float4 Gq = ldg4(f.gprev, qx, qy);
                float dz = fabsf(Gq.w - G.w);
                if (__builtin_fmaf(-p.depth_reject, G.w, dz) > 0.0f) continue;
                if (!dot_at_least(Gq, G, p.normal_reject)) continue;
                float4 H = S::ld4(f.hist_prev, qx, qy);
                int nq = (int)H.w;
                if (nq < 1) continue;
                sumb += b;
                h0 = fmaf(b, H.x, h0);
                h1 = fmaf(b, H.y, h1);
                h2 = fmaf(b, H.z, h2);
                nmax = max(nmax, nq);
                float2 M = S::ld2(f.mom_prev, qx, qy);
                if (COMP) {
                    tb[jj * 2 + ii] = b;
                    tmu[jj * 2 + ii] = M.x;
                    tm2[jj * 2 + ii] = M.y * M.y;
                } else {
                    hm1 = fmaf(b, M.x, hm1);
                    hm2 = fmaf(b, M.y, hm2);
                }
            }
        }
    }
    int n_new = 1;
    float hmu = 0.0f, hvar = 0.0f;
    if (sumb > 0.0f) {
        n_new = min(nmax + 1, p.history_cap);
        float inv = 1.0f / sumb;
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
    float rn = 1.0f / (float)n_new;
    float ac = fmaxf(p.alpha_color, rn), am = fmaxf(p.alpha_moments, rn);
    float c0 = fmaf(ac, R.x - h0, h0), c1 = fmaf(ac, R.y - h1, h1), c2 = fmaf(ac, R.z - h2, h2);
    float var;
    if (COMP) {
        float d = l - hmu;
        float mu = fmaf(am, d, hmu);
        float m2 = (1.0f - am) * fmaf(am * d, d, hvar);
        S::st2(f.mom_curr, x, y, make_float2(mu, __builtin_sqrtf(m2)));
        var = m2;
    } else {
        float mu1 = fmaf(am, l - hm1, hm1);
        float mu2 = fmaf(am, fmaf(l, l, -hm2), hm2);
        S::st2(f.mom_curr, x, y, make_float2(mu1, mu2));
        var = fmaxf(mu2 - mu1 * mu1, 0.0f);
    }
    S::st4(f.ctmp, x, y, make_float4(c0, c1, c2, var));
    reinterpret_cast<uint8_t*>(row_of(f.nlen, y))[x] = (uint8_t)n_new;
}

// Depth-gradient (central differences of clamped neighbours, halved).
__device__ __forceinline__ void depth_gradient(const FrameDesc& f, int x, int y, float& gx, float& gy)
{
    float zl = ldgz(f.g, clampi(x - 1, 0, f.w - 1), y), zr = ldgz(f.g, clampi(x + 1, 0, f.w - 1), y);
    float zu = ldgz(f.g, x, clampi(y - 1, 0, f.h - 1)), zd = ldgz(f.g, x, clampi(y + 1, 0, f.h - 1));
    gx = 0.5f * (zr - zl);
    gy = 0.5f * (zd - zu);
}

// log2 of the depth / normal edge-stop factor: -|dz| / (phi_z |g.o| + eps) * log2(e)
// + phi_n * log2(max(0, n.n')). Returns -inf when the normal term is zero.
__device__ __forceinline__ float edge_log2(const KParams& p, float4 Gc, float4 Gq, float gx, float gy, float ox, float oy)
{
    float a = 0.0f;
    if (p.phi_depth > 0.0f) {
        float den = fmaf(p.phi_depth, fabsf(fmaf(gx, ox, gy * oy)), p.depth_eps);
        a = -fabsf(Gq.w - Gc.w) * __builtin_amdgcn_rcpf(den) * 1.44269504f;
    }
    if (p.phi_normal > 0.0f) {
        float d = fmaf(Gc.x, Gq.x, fmaf(Gc.y, Gq.y, Gc.z * Gq.z));
        a = d > 0.0f ? fmaf(p.phi_normal, __builtin_amdgcn_logf(d), a) : -INFINITY;
    }
    return a;
}

// Spatial variance where the history is young (n < 4), else the temporal variance.
template <class S, bool COMP>
__device__ __forceinline__ void can_variance_px(const FrameDesc& f, const KParams& p, int x, int y)
{
    float4 C = S::ld4(f.ctmp, x, y);
    int n = reinterpret_cast<const uint8_t*>(row_of(f.nlen, y))[x];
    float var = C.w;
    if (n < 4) {
        float4 G = ldg4(f.g, x, y);
        float gx, gy;
        depth_gradient(f, x, y, gx, gy);
        float tw[9], tl[9], sw = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            int qy = clampi(y + dy, 0, f.h - 1);
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                int qx = clampi(x + dx, 0, f.w - 1);
                int t = (dy + 1) * 3 + dx + 1;
                float4 Gq = ldg4(f.g, qx, qy);
                float e = edge_log2(p, G, Gq, gx, gy, (float)dx, (float)dy);
                float w = __builtin_amdgcn_exp2f(e);
                float4 Cq = S::ld4(f.ctmp, qx, qy);
                tw[t] = w;
                tl[t] = luma_fma(Cq.x, Cq.y, Cq.z);
                sw += w;
                s1 = fmaf(w, tl[t], s1);
            }
        }
        if (sw > 0.0f) {
            float inv = 1.0f / sw;
            float mean = s1 * inv, s2 = 0.0f;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                float d = tl[t] - mean;
                s2 = fmaf(tw[t] * d, d, s2);
            }
            var = s2 * inv;
        } else {
            var = 0.0f;
        }
    }
    S::st4(f.wa, x, y, make_float4(C.x, C.y, C.z, var));
}

// A-trous iteration k (stride 2^k) with luminance, depth and normal edge-stops; variance
// filtered by squared weights. Iteration 0 also writes the colour history.
template <class S>
__device__ __forceinline__ void can_atrous_px(const FrameDesc& f, const KParams& p, int x, int y, int k)
{
    const int s = 1 << k;
    const PlaneDev& src = (k & 1) ? f.wb : f.wa;
    const bool fin = k == p.iterations - 1;
    const int max_x = f.w - 1, max_y = f.h - 1;

    float4 c = S::ld4(src, x, y);
    float lc = luma_fma(c.x, c.y, c.z);
    float gv = 0.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
        int qy = clampi(y + dy, 0, max_y);
        float ky = dy == 0 ? 0.5f : 0.25f;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            int qx = clampi(x + dx, 0, max_x);
            float kx = dx == 0 ? 0.5f : 0.25f;
            gv = fmaf(kx * ky, S::lda(src, qx, qy), gv);
        }
    }
    float sigma = p.phi_luma * __builtin_sqrtf(fmaxf(gv, 0.0f) + p.variance_eps);
    float nls = -1.44269504f / sigma;
    float4 G = ldg4(f.g, x, y);
    float gx, gy;
    depth_gradient(f, x, y, gx, gy);

    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sw = 0.0f, sv = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        int qy = clampi(y + dy * s, 0, max_y);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            int qx = clampi(x + dx * s, 0, max_x);
            float4 v = S::ld4(src, qx, qy);
            float4 Gq = ldg4(f.g, qx, qy);
            float e = edge_log2(p, G, Gq, gx, gy, (float)(dx * s), (float)(dy * s));
            e = fmaf(fabsf(luma_fma(v.x, v.y, v.z) - lc), nls, e);
            float w = (K5[dx + 2] * K5[dy + 2]) * __builtin_amdgcn_exp2f(e);
            s0 = fmaf(w, v.x, s0);
            s1 = fmaf(w, v.y, s1);
            s2 = fmaf(w, v.z, s2);
            sw += w;
            sv = fmaf(w * w, v.w, sv);
        }
    }
    float4 o;
    if (sw > 1e-6f) {
        float inv = 1.0f / sw;
        o = make_float4(s0 * inv, s1 * inv, s2 * inv, sv * inv * inv);
    } else {
        o = c;
    }
    if (k == 0) {
        float n = (float)reinterpret_cast<const uint8_t*>(row_of(f.nlen, y))[x];
        S::st4(f.hist_curr, x, y, make_float4(o.x, o.y, o.z, n));
    }
    if (fin) {
        S::st4(f.den, x, y, make_float4(o.x, o.y, o.z, 1.0f));
    } else {
        S::st4((k & 1) ? f.wa : f.wb, x, y, o);
    }
}

// ========================================================================================
// Kernels
// ========================================================================================
enum StageId : int { ST_REF_INIT = 0, ST_REF_ATROUS = 1, ST_CAN_TEMPORAL = 2, ST_CAN_VARIANCE = 3, ST_CAN_ATROUS = 4 };

template <class S, int STAGE, bool COMP>
__device__ __forceinline__ void run_stage_px(const FrameDesc& f, const KParams& p, int x, int y, int arg)
{
    if constexpr (STAGE == ST_REF_INIT) ref_init_px<S>(f, p, x, y);
    else if constexpr (STAGE == ST_REF_ATROUS) ref_atrous_px<S>(f, p, x, y, arg);
    else if constexpr (STAGE == ST_CAN_TEMPORAL) can_temporal_px<S, COMP>(f, p, x, y);
    else if constexpr (STAGE == ST_CAN_VARIANCE) can_variance_px<S, COMP>(f, p, x, y);
    else can_atrous_px<S>(f, p, x, y, arg);
}

template <class S, int STAGE, bool COMP>
__global__ __launch_bounds__(TILE_W* TILE_H) void k_frame(FrameDesc f, KParams p, int arg)
{
    int x = blockIdx.x * TILE_W + threadIdx.x;
    int y = blockIdx.y * TILE_H + threadIdx.y;
    if (x >= f.w || y >= f.h) return;
    run_stage_px<S, STAGE, COMP>(f, p, x, y, arg);
}

template <class S, int STAGE, bool COMP>
__global__ __launch_bounds__(TILE_W* TILE_H) void k_batch(BatchTable bt, KParams p, int arg)
{
    const uint32_t nf = bt.hdr[0], nt = bt.hdr[1];
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        // largest frame index i with starts[i] <= t
        uint32_t lo = 0, hi = nf - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (bt.starts[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const FrameDesc& f = bt.frames[lo];
        uint32_t local = t - bt.starts[lo];
        uint32_t ty = local / (uint32_t)f.tiles_x, tx = local - ty * (uint32_t)f.tiles_x;
        int x = (int)tx * TILE_W + threadIdx.x;
        int y = (int)ty * TILE_H + threadIdx.y;
        if (x < f.w && y < f.h) run_stage_px<S, STAGE, COMP>(f, p, x, y, arg);
    }
}

template <class S, int STAGE, bool COMP>
void launch_frame(const FrameDesc& f, const KParams& p, int arg, hipStream_t stream)
{
    dim3 grid((f.w + TILE_W - 1) / TILE_W, (f.h + TILE_H - 1) / TILE_H);
    dim3 block(TILE_W, TILE_H);
    hipLaunchKernelGGL((k_frame<S, STAGE, COMP>), grid, block, 0, stream, f, p, arg);
}

template <class S, int STAGE, bool COMP>
void launch_batch(const BatchTable& bt, const KParams& p, int arg, int grid_blocks, hipStream_t stream)
{
    dim3 block(TILE_W, TILE_H);
    hipLaunchKernelGGL((k_batch<S, STAGE, COMP>), dim3(grid_blocks), block, 0, stream, bt, p, arg);
}

inline void mark(hipEvent_t* ev, int i, hipStream_t s)
{
    if (ev) (void)hipEventRecord(ev[i], s);
}

template <class S>
void reference_frame(const FrameDesc& f, const KParams& p, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_frame<S, ST_REF_INIT, false>(f, p, 0, s);
    mark(ev, 1, s);
    for (int pass = 1; pass <= p.iterations; ++pass) {
        launch_frame<S, ST_REF_ATROUS, false>(f, p, pass, s);
        mark(ev, pass + 1, s);
    }
}

template <class S, bool COMP>
void canonical_frame(const FrameDesc& f, const KParams& p, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_frame<S, ST_CAN_TEMPORAL, COMP>(f, p, 0, s);
    mark(ev, 1, s);
    launch_frame<S, ST_CAN_VARIANCE, COMP>(f, p, 0, s);
    mark(ev, 2, s);
    for (int k = 0; k < p.iterations; ++k) {
        launch_frame<S, ST_CAN_ATROUS, COMP>(f, p, k, s);
        mark(ev, k + 3, s);
    }
}

template <class S>
void reference_batch(const BatchTable& bt, const KParams& p, int g, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_batch<S, ST_REF_INIT, false>(bt, p, 0, g, s);
    mark(ev, 1, s);
    for (int pass = 1; pass <= p.iterations; ++pass) {
        launch_batch<S, ST_REF_ATROUS, false>(bt, p, pass, g, s);
        mark(ev, pass + 1, s);
    }
}

template <class S, bool COMP>
void canonical_batch(const BatchTable& bt, const KParams& p, int g, hipStream_t s, hipEvent_t* ev)
{
    mark(ev, 0, s);
    launch_batch<S, ST_CAN_TEMPORAL, COMP>(bt, p, 0, g, s);
    mark(ev, 1, s);
    launch_batch<S, ST_CAN_VARIANCE, COMP>(bt, p, 0, g, s);
    mark(ev, 2, s);
    for (int k = 0; k < p.iterations; ++k) {
        launch_batch<S, ST_CAN_ATROUS, COMP>(bt, p, k, g, s);
        mark(ev, k + 3, s);
    }
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
    if (st == Storage::F32) {
        if (p.compensated) canonical_frame<F32S, true>(f, p, stream, ev);
        else canonical_frame<F32S, false>(f, p, stream, ev);
    } else {
        if (p.compensated) canonical_frame<F16S, true>(f, p, stream, ev);
        else canonical_frame<F16S, false>(f, p, stream, ev);
    }
    return hipGetLastError();
}

hipError_t launch_reference_batch(const BatchTable& bt, const KParams& p, Storage st, int grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev)
{
    if (st == Storage::F32) reference_batch<F32S>(bt, p, grid_blocks, stream, ev);
    else reference_batch<F16S>(bt, p, grid_blocks, stream, ev);
    return hipGetLastError();
}

hipError_t launch_canonical_batch(const BatchTable& bt, const KParams& p, Storage st, int grid_blocks,
                                  hipStream_t stream, hipEvent_t* ev)
{
    if (st == Storage::F32) {
        if (p.compensated) canonical_batch<F32S, true>(bt, p, grid_blocks, stream, ev);
        else canonical_batch<F32S, false>(bt, p, grid_blocks, stream, ev);
    } else {
        if (p.compensated) canonical_batch<F16S, true>(bt, p, grid_blocks, stream, ev);
        else canonical_batch<F16S, false>(bt, p, grid_blocks, stream, ev);
    }
    return hipGetLastError();
}

}  // namespace svgf
