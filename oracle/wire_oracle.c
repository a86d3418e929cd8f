/* wire_oracle.c — CPU restatement of the agent-frame wire encoder (test infrastructure only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's CPU-baseline leg as the checker;
 * the product path never uses it.
 *
 * The wire format is the offscreen-frame layout of docs/AI_INTERFACE.md:346-383 (TgVisionData):
 * RGB888, row-major, top to bottom, no row padding, and float32 depth normalised to [0, 1].
 * Colour follows TronGrid's display tonemap, src/postprocess.slang:47-103 and :153-160:
 * the Hill / MJP BakingLab fitted ACES (sRGB -> AP1 matrix, RRT+ODT rational fit per channel,
 * AP1 -> sRGB matrix, saturate), then the IEC 61966-2-1 sRGB transfer function. The display
 * path's cinematic terms (bloom, chromatic aberration, colour grade, vignette,
 * postprocess.slang:120-152, 162-176) are screen-space effects for a human viewer and are not
 * part of an agent's eye.
 *
 * Quantisation: code k of a channel is floor(255 * srgb(x) + 1/2) with srgb evaluated in double.
 * It is computed as the number of decision thresholds T_1..T_255 not above x, where T_k is the
 * smallest float whose double sRGB value reaches (k - 1/2) / 255, so the result does not depend
 * on how a platform rounds pow(). Depth: (z - near) / (far - near) clamped to [0, 1]; pixels
 * without geometry (z not > 0 or not finite) read 1 (far), as an empty depth buffer does.
 *
 * Arithmetic: float, one rounding per operation, no contraction (the Makefile builds with
 * -ffp-contract=off); the fused steps are written with fmaf, as the GPU kernel computes them.
 */
#include "svgf_oracle.h"

#include <math.h>
#include <stdint.h>

/* sRGB -> AP1 and AP1 -> sRGB, row-major (postprocess.slang:49-60) */
static const float ACES_IN[9] = {0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f,
                                 0.01566f, 0.02840f, 0.13383f, 0.83777f};
static const float ACES_OUT[9] = {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f,
                                  -0.00605f, -0.00327f, -0.07276f, 1.07602f};

static double srgb_encode_d(double x)
{
    if (x <= 0.0031308) return x * 12.92;
    return 1.055 * pow(x, 1.0 / 2.4) - 0.055;
}

static double srgb_decode_d(double v)
{
    if (v <= 0.04045) return v / 12.92;
    return pow((v + 0.055) / 1.055, 2.4);
}

void oracle_wire_thresholds(float* t)
{
    for (int k = 1; k <= 255; ++k) {
        const double target = (k - 0.5) / 255.0;
        float x = (float)srgb_decode_d(target);
        /* smallest float with srgb(x) >= target */
        while (x > 0.0f && srgb_encode_d((double)nextafterf(x, 0.0f)) >= target) x = nextafterf(x, 0.0f);
        while (srgb_encode_d((double)x) < target) x = nextafterf(x, 2.0f);
        t[k - 1] = x;
    }
}

/* rows of m times v, each row as fma(m0, v0, fma(m1, v1, m2 * v2)) */
static void mat3(const float* m, const float* v, float* o)
{
    for (int r = 0; r < 3; ++r) o[r] = fmaf(m[3 * r], v[0], fmaf(m[3 * r + 1], v[1], m[3 * r + 2] * v[2]));
}

void oracle_wire_tonemap(const float* rgb_in, float exposure, float* out)
{
    float c[3] = {rgb_in[0] * exposure, rgb_in[1] * exposure, rgb_in[2] * exposure};
    float a[3];
    mat3(ACES_IN, c, a);
    for (int i = 0; i < 3; ++i) {
        const float v = a[i];
        const float num = fmaf(v, v + 0.0245786f, -0.000090537f);
        const float den = fmaf(v, fmaf(0.983729f, v, 0.4329510f), 0.238081f);
        a[i] = num / den;
    }
    mat3(ACES_OUT, a, c);
    for (int i = 0; i < 3; ++i) out[i] = c[i] > 0.0f ? (c[i] < 1.0f ? c[i] : 1.0f) : 0.0f;
}

static uint8_t code_of(const float* t, float x)
{
    /* number of thresholds <= x (thresholds ascending); NaN reads 0 */
    int lo = 0, hi = 255;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return (uint8_t)lo;
}

void oracle_wire_codes(const float* x, int64_t n, uint8_t* out)
{
    float t[255];
    oracle_wire_thresholds(t);
    for (int64_t i = 0; i < n; ++i) out[i] = code_of(t, x[i]);
}

float oracle_wire_depth(float z, float near_z, float far_z)
{
    if (!(z > 0.0f) || !isfinite(z)) return 1.0f;
    const float d = (z - near_z) / (far_z - near_z);
    return d > 0.0f ? (d < 1.0f ? d : 1.0f) : 0.0f;
}

int oracle_wire_encode(int w, int h, float exposure, float near_z, float far_z, const float* colour,
                       const float* gbuffer, uint8_t* rgb, float* depth)
{
    if (w < 1 || h < 1 || !colour || !gbuffer || !rgb || !depth || !(far_z > near_z)) return 1;
    float t[255];
    oracle_wire_thresholds(t);
    for (int64_t i = 0; i < (int64_t)w * h; ++i) {
        float tm[3];
        oracle_wire_tonemap(colour + 4 * i, exposure, tm);
        rgb[3 * i + 0] = code_of(t, tm[0]);
        rgb[3 * i + 1] = code_of(t, tm[1]);
        rgb[3 * i + 2] = code_of(t, tm[2]);
        depth[i] = oracle_wire_depth(gbuffer[4 * i + 3], near_z, far_z);
    }
    return 0;
}
