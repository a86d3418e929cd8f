/*
 * svgf_oracle.c — CPU oracle for the SVGF denoiser. TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 *
 * Build: oracle/Makefile (gcc, -O2, no -ffast-math, no FMA contraction) so that float
 * arithmetic is IEEE binary32 with one rounding per operation, as in Slang's CPU target.
 */
#include "svgf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------------------
 * binary16 conversion (round to nearest even). The Vulkan path of the reference stores every
 * image as rgba16f / rg16f (src/main.cpp:1333-1370); this models that store.
 * ------------------------------------------------------------------------------------- */
uint16_t oracle_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) { /* inf / nan */
        if (ax == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x1ffu));
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 -> inf */
    if (ax >= 0x38800000u) {                                   /* normal half */
        uint32_t mant = ax & 0x7fffffu;
        uint32_t e = (ax >> 23) - 112u;
        uint32_t hb = (e << 10) | (mant >> 13);
        uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (hb & 1u))) hb++;
        return (uint16_t)(sign | hb);
    }
    if (ax < 0x33000000u) return (uint16_t)sign; /* below half of the smallest subnormal */
    {
        uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        uint32_t e = ax >> 23;       /* 102..112 */
        uint32_t shift = 126u - e;   /* 14..24 */
        uint32_t hb = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (hb & 1u))) hb++;
        return (uint16_t)(sign | hb);
    }
}

float oracle_f16_to_f32(uint16_t hb)
{
    uint32_t sign = ((uint32_t)hb & 0x8000u) << 16;
    uint32_t e = ((uint32_t)hb >> 10) & 0x1fu;
    uint32_t m = (uint32_t)hb & 0x3ffu;
    uint32_t bits;
    if (e == 0) {
        float v = ldexpf((float)m, -24);
        return sign ? -v : v;
    }
    if (e == 31) bits = sign | 0x7f800000u | (m << 13);
    else bits = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

void oracle_round_f16_array(float* data, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) data[i] = oracle_f16_to_f32(oracle_f32_to_f16(data[i]));
}

static inline float store_q(float v, int store)
{
    return store == ORACLE_STORE_F16 ? oracle_f16_to_f32(oracle_f32_to_f16(v)) : v;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Rec.709 luminance, computeLuminance (svgf_denoise.slang:109-112): dot(c, (0.2126, 0.7152,
 * 0.0722)) with the constants as float literals, summed r, g, b. */
static const float LUMA_R = 0.2126f, LUMA_G = 0.7152f, LUMA_B = 0.0722f;

/* 5-tap B3-spline, WAVELET_KERNEL (svgf_denoise.slang:118). */
static const float KERNEL5[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};

/* ---------------------------------------------------------------------------------------
 * REFERENCE mode, float storage/arithmetic (optionally half-rounded stores).
 * Pass structure: main.cpp:929-944 (pass 0 init, passes 1..N a-trous at stride 2^(k-1));
 * source = wavelet_a on odd passes (svgf_denoise.slang:209), final pass -> denoised.
 * ------------------------------------------------------------------------------------- */
typedef struct r_job_f32 {
    int w, h, pass, iterations, store, history_valid;
    const oracle_params* p;
    const float *lighting, *moments_prev;
    float *moments_curr, *denoised, *wa, *wb;
    int y0, y1;
} r_job_f32;

/* Pass 0 (svgf_denoise.slang:176-191). lerp(a, b, t) = a + t * (b - a). */
static void r_init_rows_f32(const r_job_f32* j)
{
    const float alpha = j->p->alpha_moments, eps = j->p->variance_eps;
    for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < j->w; ++x) {
            size_t i = (size_t)y * j->w + x;
            float r = store_q(j->lighting[4 * i + 0], j->store);
            float g = store_q(j->lighting[4 * i + 1], j->store);
            float b = store_q(j->lighting[4 * i + 2], j->store);
            float luma = r * LUMA_R + g * LUMA_G + b * LUMA_B;
            float m1 = 0.0f, m2 = 0.0f;
            if (j->history_valid) {
                m1 = store_q(j->moments_prev[2 * i + 0], j->store);
                m2 = store_q(j->moments_prev[2 * i + 1], j->store);
            }
            float mu1 = m1 + alpha * (luma - m1);
            float mu2 = m2 + alpha * (luma * luma - m2);
            float var = fmaxf(mu2 - mu1 * mu1, 0.0f);
            j->moments_curr[2 * i + 0] = store_q(mu1, j->store);
            j->moments_curr[2 * i + 1] = store_q(mu2, j->store);
            j->wa[4 * i + 0] = store_q(r, j->store);
            j->wa[4 * i + 1] = store_q(g, j->store);
            j->wa[4 * i + 2] = store_q(b, j->store);
            j->wa[4 * i + 3] = store_q(sqrtf(var + eps), j->store);
        }
    }
}

/* Passes 1..N (svgf_denoise.slang:209-274). */
static void r_atrous_rows_f32(const r_job_f32* j)
{
    const int pass = j->pass;
    const int stride = 1 << (pass - 1);
    const int source_is_a = (pass % 2) == 1;
    const float* src = source_is_a ? j->wa : j->wb;
    float* dst = source_is_a ? j->wb : j->wa;
    const int final_pass = (pass == j->iterations);
    const int max_x = j->w - 1, max_y = j->h - 1;
    const float phi = j->p->phi_luma, eps = j->p->variance_eps;
    for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < j->w; ++x) {
            size_t i = (size_t)y * j->w + x;
            const float* c = src + 4 * i;
            float centre_luma = c[0] * LUMA_R + c[1] * LUMA_G + c[2] * LUMA_B;
            float sigma = fmaxf(phi * c[3], eps);
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, ws = 0.0f;
            for (int dy = -2; dy <= 2; ++dy) {
                for (int dx = -2; dx <= 2; ++dx) {
                    int sx = clampi(x + dx * stride, 0, max_x);
                    int sy = clampi(y + dy * stride, 0, max_y);
                    const float* v = src + 4 * ((size_t)sy * j->w + sx);
                    float l = v[0] * LUMA_R + v[1] * LUMA_G + v[2] * LUMA_B;
                    float kw = KERNEL5[dx + 2] * KERNEL5[dy + 2];
                    float d = fabsf(l - centre_luma);
                    float ew = expf(-d / sigma);
                    float tw = kw * ew;
                    s0 += v[0] * tw;
                    s1 += v[1] * tw;
                    s2 += v[2] * tw;
                    ws += tw;
                }
            }
            float o0, o1, o2;
            if (ws > 1e-6f) {
                o0 = s0 / ws;
                o1 = s1 / ws;
                o2 = s2 / ws;
            } else {
                o0 = c[0];
                o1 = c[1];
                o2 = c[2];
            }
            float* out = final_pass ? j->denoised + 4 * i : dst + 4 * i;
            out[0] = store_q(o0, j->store);
            out[1] = store_q(o1, j->store);
            out[2] = store_q(o2, j->store);
            out[3] = final_pass ? 1.0f : c[3];
        }
    }
}

static void* r_job_entry(void* arg)
{
    r_job_f32* j = (r_job_f32*)arg;
    if (j->pass == 0) r_init_rows_f32(j);
    else r_atrous_rows_f32(j);
    return NULL;
}

int oracle_r_frame_f32(int w, int h, const oracle_params* p, int store, int history_valid,
                       const float* lighting, const float* moments_prev, float* moments_curr,
                       float* denoised, float* wavelet_a, float* wavelet_b, int threads)
{
    if (w <= 0 || h <= 0 || !p || p->iterations < 1 || p->iterations > 16) return 1;
    if (!lighting || !moments_curr || !denoised || !wavelet_a || !wavelet_b) return 1;
    if (history_valid && !moments_prev) return 1;
    if (threads < 1) threads = 1;
    if (threads > h) threads = h;
    if (threads > 256) threads = 256;
    r_job_f32 jobs[256];
    pthread_t tid[256];
    for (int pass = 0; pass <= (int)p->iterations; ++pass) {
        for (int t = 0; t < threads; ++t) {
            r_job_f32* j = &jobs[t];
            j->w = w;
            j->h = h;
            j->pass = pass;
            j->iterations = (int)p->iterations;
            j->store = store;
            j->history_valid = history_valid;
            j->p = p;
            j->lighting = lighting;
            j->moments_prev = moments_prev;
            j->moments_curr = moments_curr;
            j->denoised = denoised;
            j->wa = wavelet_a;
            j->wb = wavelet_b;
            j->y0 = (int)((int64_t)h * t / threads);
            j->y1 = (int)((int64_t)h * (t + 1) / threads);
        }
        if (threads == 1) {
            r_job_entry(&jobs[0]);
        } else {
            for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, r_job_entry, &jobs[t]);
            for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * REFERENCE mode in double: same steps with every operation in binary64. The constants are
 * the float values the shader holds (its literals are float).
 * ------------------------------------------------------------------------------------- */
int oracle_r_frame_f64(int w, int h, const oracle_params* p, int history_valid,
                       const double* lighting, const double* moments_prev, double* moments_curr,
                       double* denoised, double* wa, double* wb)
{
    if (w <= 0 || h <= 0 || !p || p->iterations < 1 || p->iterations > 16) return 1;
    if (!lighting || !moments_curr || !denoised || !wa || !wb) return 1;
    if (history_valid && !moments_prev) return 1;
    const double alpha = p->alpha_moments, eps = p->variance_eps, phi = p->phi_luma;
    const double lr = LUMA_R, lg = LUMA_G, lb = LUMA_B;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            const double* L = lighting + 4 * i;
            double luma = L[0] * lr + L[1] * lg + L[2] * lb;
            double m1 = history_valid ? moments_prev[2 * i] : 0.0;
            double m2 = history_valid ? moments_prev[2 * i + 1] : 0.0;
            double mu1 = m1 + alpha * (luma - m1);
            double mu2 = m2 + alpha * (luma * luma - m2);
            double var = fmax(mu2 - mu1 * mu1, 0.0);
            moments_curr[2 * i] = mu1;
            moments_curr[2 * i + 1] = mu2;
            wa[4 * i + 0] = L[0];
            wa[4 * i + 1] = L[1];
            wa[4 * i + 2] = L[2];
            wa[4 * i + 3] = sqrt(var + eps);
        }
    }
    for (int pass = 1; pass <= (int)p->iterations; ++pass) {
        int stride = 1 << (pass - 1);
        const double* src = (pass % 2 == 1) ? wa : wb;
        double* dst = (pass % 2 == 1) ? wb : wa;
        int fin = pass == (int)p->iterations;
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                size_t i = (size_t)y * w + x;
                const double* c = src + 4 * i;
                double lc = c[0] * lr + c[1] * lg + c[2] * lb;
                double sigma = fmax(phi * c[3], eps);
                double s0 = 0, s1 = 0, s2 = 0, ws = 0;
                for (int dy = -2; dy <= 2; ++dy) {
                    for (int dx = -2; dx <= 2; ++dx) {
                        int sx = clampi(x + dx * stride, 0, w - 1);
                        int sy = clampi(y + dy * stride, 0, h - 1);
                        const double* v = src + 4 * ((size_t)sy * w + sx);
                        double l = v[0] * lr + v[1] * lg + v[2] * lb;
                        double tw = (double)KERNEL5[dx + 2] * (double)KERNEL5[dy + 2] * exp(-fabs(l - lc) / sigma);
                        s0 += v[0] * tw;
                        s1 += v[1] * tw;
                        s2 += v[2] * tw;
                        ws += tw;
                    }
                }
                double* out = fin ? denoised + 4 * i : dst + 4 * i;
                if (ws > 1e-6) {
                    out[0] = s0 / ws;
                    out[1] = s1 / ws;
                    out[2] = s2 / ws;
                } else {
                    out[0] = c[0];
                    out[1] = c[1];
                    out[2] = c[2];
                }
                out[3] = fin ? 1.0 : c[3];
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * CANONICAL mode: the oracle in binary64 (oracle_c_frame_f64*), and the same code in binary32
 * (oracle_c_frame_f32_mt) as the CPU baseline port of the FP32 canonical form.
 *
 * Temporal stage (per pixel p = (x, y)):
 *   l = luma(radiance); reprojected position = (x + floor(m.x) + fx, y + floor(m.y) + fy),
 *   fx = m.x - floor(m.x). Bilinear 2x2 history taps, row-major (j outer, i inner); a tap is
 *   used when its weight b > 0, it lies inside the image, |z' - z| <= depth_reject * z,
 *   n.n' >= normal_reject and its history length >= 1 (the test of
 *   volumetric_filter.slang:171-206 applied per tap). With sum(b) > 0 the history is the
 *   b-weighted mean of the used taps and n = min(1 + max used length, history_cap);
 *   otherwise n = 1 (disocclusion). a_c = max(alpha_color, 1/n), a_m = max(alpha_moments, 1/n).
 *   colour = h + a_c (R - h)   (the EMA of volumetric_filter.slang:223)
 *   raw moments:         mu1 = h1 + a_m (l - h1); mu2 = h2 + a_m (l^2 - h2); var = max(mu2 - mu1^2, 0)
 *   compensated moments: d = l - mu; mu' = mu + a_m d; m2' = (1 - a_m)(m2 + a_m d^2); var = m2'
 *     with the interpolated state mu = sum b mu_q / B, m2 = sum b (m2_q + (mu_q - mu)^2) / B.
 * Variance stage: n < 4 -> 3x3 clamp-to-edge window of temporal luminances weighted by the
 *   depth and normal edge-stops (stride 1): var = sum w (l_q - mean)^2 / sum w; else the
 *   temporal variance. Output W0 = (colour, var).
 * A-trous iteration k (stride s = 2^k), taps as svgf_denoise.slang:225-254:
 *   gv = 3x3 binomial (1,2,1)/4 x (1,2,1)/4 of the variance at p + (i s, j s), i, j in
 *        {-1, 0, 1} (clamped): the prefilter runs on the iteration's own lattice
 *   sigma_l = phi_luma * sqrt(max(gv, 0) + variance_eps)
 *   (gx, gy) = central differences of depth at p (clamped neighbours, halved)
 *   w = K[dx] K[dy] exp(-|l_q - l_p| / sigma_l - |z_q - z_p| / (phi_depth |gx dx s + gy dy s| + depth_eps))
 *       * max(0, n_p.n_q)^phi_normal        (depth / normal terms off when their phi is 0)
 *   colour = sum w c_q / sum w; variance = sum w^2 var_q / (sum w)^2 (centre when sum w <= 1e-6)
 *   Iteration 0's colour (with n in alpha) is the next frame's history; the last iteration
 *   writes denoised = (colour, 1).
 * ------------------------------------------------------------------------------------- */

#define CREAL double
#define CSUF(name) name##_f64
#define C_FRAME oracle_c_frame_f64
#define C_FRAME_MT oracle_c_frame_f64_mt
#define C_LIT(x) ((double)(x))
#define C_EXP exp
#define C_POW pow
#define C_SQRT sqrt
#define C_FABS fabs
#define C_FLOOR floor
#define C_FMAX fmax
#include "svgf_oracle_canon.inc"
#undef CREAL
#undef CSUF
#undef C_FRAME
#undef C_FRAME_MT
#undef C_LIT
#undef C_EXP
#undef C_POW
#undef C_SQRT
#undef C_FABS
#undef C_FLOOR
#undef C_FMAX

#define CREAL float
#define CSUF(name) name##_f32
#define C_FRAME oracle_c_frame_f32
#define C_FRAME_MT oracle_c_frame_f32_mt
#define C_LIT(x) ((float)(x))
#define C_EXP expf
#define C_POW powf
#define C_SQRT sqrtf
#define C_FABS fabsf
#define C_FLOOR floorf
#define C_FMAX fmaxf
#include "svgf_oracle_canon.inc"
