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
 * CANONICAL mode, binary64.
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
static inline double luma_d(const double* c) { return c[0] * (double)LUMA_R + c[1] * (double)LUMA_G + c[2] * (double)LUMA_B; }

static double depth_normal_weight(const oracle_params* p, const float* gp, const float* gq, double gx,
                                  double gy, double ox, double oy, double* arg_out)
{
    double arg = 0.0;
    if (p->phi_depth > 0.0f) {
        double den = (double)p->phi_depth * fabs(gx * ox + gy * oy) + (double)p->depth_eps;
        arg = fabs((double)gq[3] - (double)gp[3]) / den;
    }
    double wn = 1.0;
    if (p->phi_normal > 0.0f) {
        double d = (double)gp[0] * gq[0] + (double)gp[1] * gq[1] + (double)gp[2] * gq[2];
        wn = d > 0.0 ? pow(d, (double)p->phi_normal) : 0.0;
    }
    *arg_out = arg;
    return wn;
}

static void depth_gradient(const float* g, int w, int h, int x, int y, double* gx, double* gy)
{
    int xl = clampi(x - 1, 0, w - 1), xr = clampi(x + 1, 0, w - 1);
    int yu = clampi(y - 1, 0, h - 1), yd = clampi(y + 1, 0, h - 1);
    *gx = 0.5 * ((double)g[4 * ((size_t)y * w + xr) + 3] - (double)g[4 * ((size_t)y * w + xl) + 3]);
    *gy = 0.5 * ((double)g[4 * ((size_t)yd * w + x) + 3] - (double)g[4 * ((size_t)yu * w + x) + 3]);
}

typedef struct c_job {
    int w, h, history_valid, as_ref, comp, k;
    const oracle_params* p;
    const float *radiance, *gbuffer, *gbuffer_prev, *motion;
    const double *history_prev, *moments_prev;
    double *history_curr, *moments_curr, *denoised;
    double *C, *WA, *WB;
    const double* src;
    double* dst;
    int* N;
    int y0, y1;
} c_job;

static void c_temporal_rows(const c_job* J)
{
    const int w = J->w, h = J->h, history_valid = J->history_valid, as_ref = J->as_ref, comp = J->comp;
    const oracle_params* p = J->p;
    const float *radiance = J->radiance, *gbuffer = J->gbuffer, *gbuffer_prev = J->gbuffer_prev, *motion = J->motion;
    const double *history_prev = J->history_prev, *moments_prev = J->moments_prev;
    double *moments_curr = J->moments_curr, *C = J->C;
    int* N = J->N;
    const double alpha_m = p->alpha_moments, alpha_c = p->alpha_color;
    for (int y = J->y0; y < J->y1; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            const float* R = radiance + 4 * i;
            double Rd[3] = {R[0], R[1], R[2]};
            double l = luma_d(Rd);
            if (as_ref) {
                /* same-pixel moments, colour = radiance (svgf_denoise.slang:176-191) */
                double m1 = history_valid ? moments_prev[2 * i] : 0.0;
                double m2 = history_valid ? moments_prev[2 * i + 1] : 0.0;
                double mu1 = m1 + alpha_m * (l - m1), mu2 = m2 + alpha_m * (l * l - m2);
                moments_curr[2 * i] = mu1;
                moments_curr[2 * i + 1] = mu2;
                C[4 * i + 0] = Rd[0];
                C[4 * i + 1] = Rd[1];
                C[4 * i + 2] = Rd[2];
                C[4 * i + 3] = sqrt(fmax(mu2 - mu1 * mu1, 0.0) + (double)p->variance_eps); /* sigma */
                N[i] = 1000000;
                continue;
            }
            const float* G = gbuffer + 4 * i;
            double z = G[3];
            double sumb = 0.0, hc[3] = {0, 0, 0}, hm1 = 0.0, hm2 = 0.0;
            double tb[4], tmu[4], tm2[4];
            int used = 0, nmax = 0;
            if (history_valid) {
                double mx = motion[2 * i], my = motion[2 * i + 1];
                double flx = floor(mx), fly = floor(my);
                double fx = mx - flx, fy = my - fly;
                int x0 = x + (int)flx, y0 = y + (int)fly;
                for (int jj = 0; jj < 2; ++jj) {
                    for (int ii = 0; ii < 2; ++ii) {
                        double b = (ii ? fx : 1.0 - fx) * (jj ? fy : 1.0 - fy);
                        int qx = x0 + ii, qy = y0 + jj;
                        if (!(b > 0.0) || qx < 0 || qx >= w || qy < 0 || qy >= h) continue;
                        size_t q = (size_t)qy * w + qx;
                        const float* Gq = gbuffer_prev + 4 * q;
                        double dz = fabs((double)Gq[3] - z);
                        if (dz > (double)p->depth_reject * z) continue;
                        double nd = (double)Gq[0] * G[0] + (double)Gq[1] * G[1] + (double)Gq[2] * G[2];
                        if (nd < (double)p->normal_reject) continue;
                        const double* Hq = history_prev + 4 * q;
                        int nq = (int)Hq[3];
                        if (nq < 1) continue;
                        sumb += b;
                        hc[0] += b * Hq[0];
                        hc[1] += b * Hq[1];
                        hc[2] += b * Hq[2];
                        if (nq > nmax) nmax = nq;
                        if (comp) {
                            tb[used] = b;
                            tmu[used] = moments_prev[2 * q];
                            tm2[used] = moments_prev[2 * q + 1] * moments_prev[2 * q + 1];
                        } else {
                            hm1 += b * moments_prev[2 * q];
                            hm2 += b * moments_prev[2 * q + 1];
                        }
                        used++;
                    }
                }
            }
            int n_new = 1;
            double hmu = 0.0, hvar = 0.0;
            if (sumb > 0.0) {
                n_new = nmax + 1;
                if (n_new > (int)p->history_cap) n_new = (int)p->history_cap;
                hc[0] /= sumb;
                hc[1] /= sumb;
                hc[2] /= sumb;
                if (comp) {
                    double mu = 0.0;
                    for (int t = 0; t < used; ++t) mu += tb[t] * tmu[t];
                    mu /= sumb;
                    double m2 = 0.0;
                    for (int t = 0; t < used; ++t) m2 += tb[t] * (tm2[t] + (tmu[t] - mu) * (tmu[t] - mu));
                    hmu = mu;
                    hvar = m2 / sumb;
                } else {
                    hm1 /= sumb;
                    hm2 /= sumb;
                }
            } else {
                hc[0] = hc[1] = hc[2] = 0.0;
                hm1 = hm2 = 0.0;
            }
            double ac = fmax(alpha_c, 1.0 / n_new), am = fmax(alpha_m, 1.0 / n_new);
            double col[3];
            for (int c = 0; c < 3; ++c) col[c] = hc[c] + ac * (Rd[c] - hc[c]);
            double var;
            if (comp) {
                double d = l - hmu;
                double mu = hmu + am * d;
                double m2 = (1.0 - am) * (hvar + am * d * d);
                moments_curr[2 * i] = mu;
                moments_curr[2 * i + 1] = sqrt(m2);
                var = m2;
            } else {
                double mu1 = hm1 + am * (l - hm1), mu2 = hm2 + am * (l * l - hm2);
                moments_curr[2 * i] = mu1;
                moments_curr[2 * i + 1] = mu2;
                var = fmax(mu2 - mu1 * mu1, 0.0);
            }
            C[4 * i + 0] = col[0];
            C[4 * i + 1] = col[1];
            C[4 * i + 2] = col[2];
            C[4 * i + 3] = var;
            N[i] = n_new;
        }
    }
}

static void c_variance_rows(const c_job* J)
{
    const int w = J->w, h = J->h, as_ref = J->as_ref;
    const oracle_params* p = J->p;
    const float* gbuffer = J->gbuffer;
    const double* C = J->C;
    double* WA = J->WA;
    const int* N = J->N;
    for (int y = J->y0; y < J->y1; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            double var = C[4 * i + 3];
            if (!as_ref && N[i] < 4) {
                const float* Gp = gbuffer + 4 * i;
                double gx, gy;
                depth_gradient(gbuffer, w, h, x, y, &gx, &gy);
                double tw[9], tl[9], sw = 0.0, s1 = 0.0;
                int t = 0;
                for (int dy = -1; dy <= 1; ++dy) {
                    for (int dx = -1; dx <= 1; ++dx, ++t) {
                        int qx = clampi(x + dx, 0, w - 1), qy = clampi(y + dy, 0, h - 1);
                        size_t q = (size_t)qy * w + qx;
                        double arg;
                        double wn = depth_normal_weight(p, Gp, gbuffer + 4 * q, gx, gy, dx, dy, &arg);
                        tw[t] = exp(-arg) * wn;
                        tl[t] = luma_d(C + 4 * q);
                        sw += tw[t];
                        s1 += tw[t] * tl[t];
                    }
                }
                if (sw > 0.0) {
                    double mean = s1 / sw, s2 = 0.0;
                    for (t = 0; t < 9; ++t) s2 += tw[t] * (tl[t] - mean) * (tl[t] - mean);
                    var = s2 / sw;
                } else {
                    var = 0.0;
                }
            }
            WA[4 * i + 0] = C[4 * i + 0];
            WA[4 * i + 1] = C[4 * i + 1];
            WA[4 * i + 2] = C[4 * i + 2];
            WA[4 * i + 3] = var;
        }
    }
}

static void c_atrous_rows(const c_job* J)
{
    const int w = J->w, h = J->h, as_ref = J->as_ref, k = J->k;
    const oracle_params* p = J->p;
    const float* gbuffer = J->gbuffer;
    const double* src = J->src;
    double *dst = J->dst, *history_curr = J->history_curr, *denoised = J->denoised;
    const int* N = J->N;
    const int s = 1 << k;
    const int fin = k == (int)p->iterations - 1;
    for (int y = J->y0; y < J->y1; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            const double* c = src + 4 * i;
            double lc = luma_d(c);
            double sigma;
            double gx = 0.0, gy = 0.0;
            const float* Gp = NULL;
            if (as_ref) {
                sigma = fmax((double)p->phi_luma * c[3], (double)p->variance_eps);
            } else {
                double gv = 0.0;
                static const double K3[3] = {0.25, 0.5, 0.25};
                const int ps = (p->flags & ORACLE_FLAG_PREFILTER_UNIT) ? 1 : s;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int qx = clampi(x + dx * ps, 0, w - 1), qy = clampi(y + dy * ps, 0, h - 1);
                        gv += K3[dx + 1] * K3[dy + 1] * src[4 * ((size_t)qy * w + qx) + 3];
                    }
                if (p->flags & ORACLE_FLAG_EPS_OUTSIDE)
                    sigma = (double)p->phi_luma * sqrt(fmax(gv, 0.0)) + (double)p->variance_eps;
                else
                    sigma = (double)p->phi_luma * sqrt(fmax(gv, 0.0) + (double)p->variance_eps);
                Gp = gbuffer + 4 * i;
                depth_gradient(gbuffer, w, h, x, y, &gx, &gy);
            }
            double s0 = 0, s1 = 0, s2 = 0, sw = 0, sv = 0;
            for (int dy = -2; dy <= 2; ++dy) {
                for (int dx = -2; dx <= 2; ++dx) {
                    int qx = clampi(x + dx * s, 0, w - 1), qy = clampi(y + dy * s, 0, h - 1);
                    size_t q = (size_t)qy * w + qx;
                    const double* v = src + 4 * q;
                    double arg = fabs(luma_d(v) - lc) / sigma;
                    double wn = 1.0;
                    if (!as_ref) {
                        double az;
                        wn = depth_normal_weight(p, Gp, gbuffer + 4 * q, gx, gy, (double)dx * s, (double)dy * s, &az);
                        arg += az;
                    }
                    double wt = (double)KERNEL5[dx + 2] * (double)KERNEL5[dy + 2] * exp(-arg) * wn;
                    s0 += wt * v[0];
                    s1 += wt * v[1];
                    s2 += wt * v[2];
                    sw += wt;
                    sv += wt * wt * v[3];
                }
            }
            double* o = dst + 4 * i;
            if (sw > 1e-6) {
                o[0] = s0 / sw;
                o[1] = s1 / sw;
                o[2] = s2 / sw;
                o[3] = as_ref ? c[3] : sv / (sw * sw);
            } else {
                o[0] = c[0];
                o[1] = c[1];
                o[2] = c[2];
                o[3] = c[3];
            }
            if (k == 0 && !as_ref) {
                history_curr[4 * i + 0] = o[0];
                history_curr[4 * i + 1] = o[1];
                history_curr[4 * i + 2] = o[2];
                history_curr[4 * i + 3] = (double)N[i];
            }
            if (fin) {
                denoised[4 * i + 0] = o[0];
                denoised[4 * i + 1] = o[1];
                denoised[4 * i + 2] = o[2];
                denoised[4 * i + 3] = 1.0;
            }
        }
    }
}

typedef void (*c_rows_fn)(const c_job*);
typedef struct c_task {
    c_rows_fn fn;
    c_job job;
} c_task;

static void* c_task_entry(void* arg)
{
    c_task* t = (c_task*)arg;
    t->fn(&t->job);
    return NULL;
}

/* Runs fn over all rows, split statically over `threads` pthreads (rows are independent
 * within a stage). */
static void c_run_rows(c_rows_fn fn, const c_job* base, int threads)
{
    if (threads <= 1) {
        c_job j = *base;
        j.y0 = 0;
        j.y1 = base->h;
        fn(&j);
        return;
    }
    c_task tasks[256];
    pthread_t tid[256];
    for (int t = 0; t < threads; ++t) {
        tasks[t].fn = fn;
        tasks[t].job = *base;
        tasks[t].job.y0 = (int)((int64_t)base->h * t / threads);
        tasks[t].job.y1 = (int)((int64_t)base->h * (t + 1) / threads);
        pthread_create(&tid[t], NULL, c_task_entry, &tasks[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

int oracle_c_frame_f64_mt(int w, int h, const oracle_params* p, int history_valid,
                          const float* radiance, const float* gbuffer, const float* gbuffer_prev,
                          const float* motion, const double* history_prev, const double* moments_prev,
                          double* history_curr, double* moments_curr, double* denoised,
                          double* dbg_temporal, double* dbg_variance, int threads)
{
    if (w <= 0 || h <= 0 || !p || p->iterations < 1 || p->iterations > 16) return 1;
    if (!radiance || !history_curr || !moments_curr || !denoised) return 1;
    const int as_ref = (p->flags & ORACLE_FLAG_AS_REFERENCE) != 0;
    const int comp = (p->flags & ORACLE_FLAG_COMPENSATED) != 0 && !as_ref;
    if (!as_ref && (!gbuffer || !motion)) return 1;
    if (history_valid && (!moments_prev || (!as_ref && (!history_prev || !gbuffer_prev)))) return 1;
    if (threads < 1) threads = 1;
    if (threads > h) threads = h;
    if (threads > 256) threads = 256;
    const size_t np = (size_t)w * h;
    double* C = (double*)malloc(np * 4 * sizeof(double));
    int* N = (int*)malloc(np * sizeof(int));
    double* WA = (double*)malloc(np * 4 * sizeof(double));
    double* WB = (double*)malloc(np * 4 * sizeof(double));
    if (!C || !N || !WA || !WB) {
        free(C);
        free(N);
        free(WA);
        free(WB);
        return 2;
    }
    c_job J;
    memset(&J, 0, sizeof J);
    J.w = w;
    J.h = h;
    J.history_valid = history_valid;
    J.as_ref = as_ref;
    J.comp = comp;
    J.p = p;
    J.radiance = radiance;
    J.gbuffer = gbuffer;
    J.gbuffer_prev = gbuffer_prev;
    J.motion = motion;
    J.history_prev = history_prev;
    J.moments_prev = moments_prev;
    J.history_curr = history_curr;
    J.moments_curr = moments_curr;
    J.denoised = denoised;
    J.C = C;
    J.WA = WA;
    J.WB = WB;
    J.N = N;

    /* ---- temporal ---- */
    c_run_rows(c_temporal_rows, &J, threads);
    if (dbg_temporal) {
        for (size_t i = 0; i < np; ++i) {
            dbg_temporal[4 * i + 0] = C[4 * i + 0];
            dbg_temporal[4 * i + 1] = C[4 * i + 1];
            dbg_temporal[4 * i + 2] = C[4 * i + 2];
            dbg_temporal[4 * i + 3] = as_ref ? C[4 * i + 3] : (double)N[i];
        }
    }

    /* ---- variance ---- */
    c_run_rows(c_variance_rows, &J, threads);
    if (dbg_variance) memcpy(dbg_variance, WA, np * 4 * sizeof(double));

    /* ---- a-trous ---- */
    J.src = WA;
    J.dst = WB;
    for (int k = 0; k < (int)p->iterations; ++k) {
        J.k = k;
        c_run_rows(c_atrous_rows, &J, threads);
        const double* t = J.src;
        J.src = J.dst;
        J.dst = (double*)t;
    }
    free(C);
    free(N);
    free(WA);
    free(WB);
    return 0;
}

int oracle_c_frame_f64(int w, int h, const oracle_params* p, int history_valid,
                       const float* radiance, const float* gbuffer, const float* gbuffer_prev,
                       const float* motion, const double* history_prev, const double* moments_prev,
                       double* history_curr, double* moments_curr, double* denoised,
                       double* dbg_temporal, double* dbg_variance)
{
    return oracle_c_frame_f64_mt(w, h, p, history_valid, radiance, gbuffer, gbuffer_prev, motion, history_prev,
                                 moments_prev, history_curr, moments_curr, denoised, dbg_temporal, dbg_variance, 1);
}
