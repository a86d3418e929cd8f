/*
 * svgf_oracle.h — CPU oracle for the SVGF denoiser. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library; it is the checker, never the product path.
 *
 * Two restatements live here:
 *   oracle_r_*  the reference shader src/svgf_denoise.slang:152-275 as written (REFERENCE
 *               mode), with storage = float (Slang CPU-target semantics), float rounded to
 *               half at every image store (the Vulkan rgba16f/rg16f images of
 *               src/main.cpp:1333-1370), or double.
 *   oracle_c_*  canonical SVGF (Schied et al. 2017, cited at svgf_denoise.slang:18) in
 *               double precision (CANONICAL mode), with the reference's clamp-to-edge taps
 *               and tap order, and the history-rejection pattern of
 *               src/volumetric_filter.slang:171-223.
 *
 * Planes are tightly packed row-major arrays with interleaved channels (4 for colour-like
 * planes, 2 for moments / motion).
 */
#ifndef SVGF_ORACLE_H
#define SVGF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as svgf_params in include/svgf.h. */
typedef struct oracle_params {
    uint32_t iterations;
    float phi_luma;
    float phi_normal;
    float phi_depth;
    float alpha_color;
    float alpha_moments;
    float variance_eps;
    float depth_eps;
    float depth_reject;
    float normal_reject;
    uint32_t history_cap;
    uint32_t flags;
} oracle_params;

#define ORACLE_FLAG_COMPENSATED 0x1u /* moments as (mean, sqrt(centred second moment)) */
#define ORACLE_FLAG_AS_REFERENCE 0x100u /* oracle_c_* with every canonical feature switched
                                           off: must reproduce oracle_r_f64 */
/* Schied et al. 2017's forms where CANONICAL deviates (for measuring the deviations only):
 * the 3x3 variance prefilter over the pixel's direct neighbours (+-1) instead of its lattice
 * neighbours at +-s, and the luminance scale phi_l sqrt(g(var)) + eps instead of
 * phi_l sqrt(g(var) + eps). */
#define ORACLE_FLAG_PREFILTER_UNIT 0x200u
#define ORACLE_FLAG_EPS_OUTSIDE 0x400u

enum { ORACLE_STORE_F32 = 0, ORACLE_STORE_F16 = 1 };

/* IEEE binary16 round-to-nearest-even (overflow to inf, subnormals kept). */
uint16_t oracle_f32_to_f16(float f);
float oracle_f16_to_f32(uint16_t h);
void oracle_round_f16_array(float* data, int64_t n);

/* One REFERENCE-mode frame. lighting: w*h*4; moments_prev/curr: w*h*2; denoised,
 * wavelet_a, wavelet_b: w*h*4. history_valid = 0 reads moments_prev as zero.
 * threads > 1 splits rows statically over pthreads (CPU baseline timing). */
int oracle_r_frame_f32(int w, int h, const oracle_params* p, int store, int history_valid,
                       const float* lighting, const float* moments_prev, float* moments_curr,
                       float* denoised, float* wavelet_a, float* wavelet_b, int threads);
int oracle_r_frame_f64(int w, int h, const oracle_params* p, int history_valid,
                       const double* lighting, const double* moments_prev, double* moments_curr,
                       double* denoised, double* wavelet_a, double* wavelet_b);

/* One CANONICAL frame in double precision.
 * radiance, gbuffer, gbuffer_prev: w*h*4 float; motion: w*h*2 float (inputs as the GPU sees
 * them). history_prev/curr: w*h*4 (rgb, history length); moments_prev/curr: w*h*2 (raw
 * (mu1, mu2), or (mean, sigma) with ORACLE_FLAG_COMPENSATED); denoised: w*h*4.
 * Optional debug outputs (may be NULL): temporal w*h*4 (colour, history length),
 * variance w*h*4 (colour, variance) = the input of the first a-trous iteration. */
int oracle_c_frame_f64(int w, int h, const oracle_params* p, int history_valid,
                       const float* radiance, const float* gbuffer, const float* gbuffer_prev,
                       const float* motion, const double* history_prev, const double* moments_prev,
                       double* history_curr, double* moments_curr, double* denoised,
                       double* dbg_temporal, double* dbg_variance);
/* Same, each stage split by rows over `threads` pthreads (results identical). */
int oracle_c_frame_f64_mt(int w, int h, const oracle_params* p, int history_valid,
                          const float* radiance, const float* gbuffer, const float* gbuffer_prev,
                          const float* motion, const double* history_prev, const double* moments_prev,
                          double* history_curr, double* moments_curr, double* denoised,
                          double* dbg_temporal, double* dbg_variance, int threads);
/* The same CANONICAL code in binary32 (float state planes, float arithmetic): the CPU baseline
 * port of the FP32 canonical form. Not an oracle; pinned within 2e-4 of the binary64 one. */
int oracle_c_frame_f32(int w, int h, const oracle_params* p, int history_valid,
                       const float* radiance, const float* gbuffer, const float* gbuffer_prev,
                       const float* motion, const float* history_prev, const float* moments_prev,
                       float* history_curr, float* moments_curr, float* denoised,
                       float* dbg_temporal, float* dbg_variance);
int oracle_c_frame_f32_mt(int w, int h, const oracle_params* p, int history_valid,
                          const float* radiance, const float* gbuffer, const float* gbuffer_prev,
                          const float* motion, const float* history_prev, const float* moments_prev,
                          float* history_curr, float* moments_curr, float* denoised,
                          float* dbg_temporal, float* dbg_variance, int threads);

/* Agent-frame wire encoder (wire_oracle.c): docs/AI_INTERFACE.md:346-383 layout, colour through
 * postprocess.slang's fitted ACES + sRGB, quantised by decision thresholds. */
void oracle_wire_thresholds(float* t255);
void oracle_wire_tonemap(const float* rgb, float exposure, float* out3);
void oracle_wire_codes(const float* x, int64_t n, uint8_t* out); /* 8-bit sRGB code of linear x */
float oracle_wire_depth(float z, float near_z, float far_z);
/* colour, gbuffer: w*h*4 float; rgb: w*h*3 bytes; depth: w*h float. Returns 0 on success. */
int oracle_wire_encode(int w, int h, float exposure, float near_z, float far_z, const float* colour,
                       const float* gbuffer, uint8_t* rgb, float* depth);

/* Consumer of the denoised output (composite_oracle.c): mesh.slang:1007, 1161-1174, 1191.
 * m: row-major 4x4 previous view-projection (clip = m (x, y, z, 1)). Lookup: the history pixel
 * a world position reprojects to; 0 and *px, *py set, or -1 when behind the camera / off
 * screen. Composite: every plane w*h*4 float; reflection.w is the kind (2 reproject,
 * 1 ready colour, else none); store = 1 rounds the output to half. Returns 0 on success. */
int oracle_composite_lookup(const float* m, int w, int h, float x, float y, float z, int* px, int* py);
int oracle_composite(int w, int h, const float* m, float ambient, const float* history, const float* emissive,
                     const float* fresnel, const float* reflection, const float* refl_colour, const float* refl_base,
                     int store, float* out);

/* Producer side of the denoiser input (pack_oracle.c): main.cpp:595-646 (zero-clear),
 * mesh.slang:995-998 (lighting), :341-352 (octahedral normal), :589-591 (previous position).
 * vp, pvp: row-major 4x4 current / previous view-projection. position, indirect,
 * direct_diffuse (ao in .w), direct_specular: w*h*4 float; normal_oct: w*h uint32. Outputs:
 * radiance w*h*4 (store = 1 rounds to half), gbuffer w*h*4 (normal, depth), motion w*h*2. */
void oracle_unpack_octahedral(uint32_t packed, float* n3);
int oracle_pack(int w, int h, const float* vp, const float* pvp, const float* position, const uint32_t* normal_oct,
                const float* indirect, const float* direct_diffuse, const float* direct_specular, int store,
                float* radiance, float* gbuffer, float* motion);

/* The reference's GPU pass profiler loop (profiler_oracle.c; main.cpp:2910-2951, 3043, 3096,
 * 3148). samples, ema: frames*passes floats (ns); logged: frames bytes. Returns 0 on success. */
int oracle_profiler_run(int passes, int slots, int frames, const float* samples, float alpha, uint32_t interval,
                        uint32_t first, float* ema, uint8_t* logged);

#ifdef __cplusplus
}
#endif

#endif
