/* composite_oracle.c — CPU restatement of the consumer of the denoised output (test
 * infrastructure only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's CPU-baseline leg as the checker;
 * the product path never uses it.
 *
 * TronGrid reads the SVGF output one frame late, in the mesh pass (src/mesh.slang):
 *   - :1007      indirect = history[pixel].rgb, the previous frame's denoised lighting at the
 *                same pixel;
 *   - :1161-1174 a reflection ray that hit a non-emissive surface reprojects the hit position
 *                through the previous frame's view-projection and reads the history there:
 *                  clip = PVP (hit, 1); if clip.w > 0: ndc = clip.xy / clip.w,
 *                  uv = ndc (0.5, -0.5) + 0.5, px = int2(uv (W, H));
 *                  px inside [0, W) x [0, H): hit_indirect = history[px].rgb,
 *                  otherwise hit_indirect = hit_base * AMBIENT (AMBIENT = 0.03, :276);
 *                reflected = hit_emissive + hit_indirect;
 *   - :1191      colour = (indirect + emissive) + reflected * F_view, alpha 1.
 * Per pixel the caller supplies what the mesh pass has at that point besides the history:
 * emissive, F_view, and the reflection as a kind (reflection.w): 2 = reproject (reflection.xyz
 * = hit position, refl_colour = hit emissive, refl_base = hit base colour), 1 = a reflected
 * colour that needs no lookup (sky miss or emissive hit, :1066-1131) in refl_colour, anything
 * else = no reflection ray (reflected = 0; refl_colour and refl_base are not read).
 *
 * int2() of the reference truncates toward zero, so uv (W, H) in (-1, 0) lands on column/row 0
 * and is inside; a product that is NaN or out of int range is off screen here (the shader
 * leaves that conversion undefined).
 *
 * Arithmetic: float, one rounding per operation, no contraction (the Makefile builds with
 * -ffp-contract=off). Matrix rows are dot products from left to right,
 * ((m0 x + m1 y) + m2 z) + m3, the order of the Slang CPU target's mul(). A half output
 * (store = 1) is rounded to nearest even at the store.
 */
#include "svgf_oracle.h"

#include <stdint.h>

/* row r of the row-major matrix m times (x, y, z, 1) */
static float row4(const float* m, int r, float x, float y, float z)
{
    return ((m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z) + m[4 * r + 3];
}

int oracle_composite_lookup(const float* m, int w, int h, float x, float y, float z, int* px, int* py)
{
    const float cw = row4(m, 3, x, y, z);
    if (!(cw > 0.0f)) return -1;
    const float nx = row4(m, 0, x, y, z) / cw;
    const float ny = row4(m, 1, x, y, z) / cw;
    const float u = nx * 0.5f + 0.5f;
    const float v = ny * -0.5f + 0.5f;
    const float fx = u * (float)w, fy = v * (float)h;
    /* trunc(f) lies in [0, n) exactly when -1 < f < n; NaN fails both tests */
    if (!(fx > -1.0f && fx < (float)w && fy > -1.0f && fy < (float)h)) return -1;
    *px = (int)fx;
    *py = (int)fy;
    return 0;
}

int oracle_composite(int w, int h, const float* m, float ambient, const float* history, const float* emissive,
                     const float* fresnel, const float* reflection, const float* refl_colour, const float* refl_base,
                     int store, float* out)
{
    if (w < 1 || h < 1 || !m || !history || !emissive || !fresnel || !reflection || !refl_colour || !refl_base || !out)
        return 1;
    for (int64_t i = 0; i < (int64_t)w * h; ++i) {
        const float kind = reflection[4 * i + 3];
        float refl[3] = {0.0f, 0.0f, 0.0f};
        if (kind == 2.0f) {
            int px, py;
            const float* r = reflection + 4 * i;
            if (oracle_composite_lookup(m, w, h, r[0], r[1], r[2], &px, &py) == 0) {
                const float* hh = history + 4 * ((int64_t)py * w + px);
                for (int c = 0; c < 3; ++c) refl[c] = refl_colour[4 * i + c] + hh[c];
            } else {
                for (int c = 0; c < 3; ++c) refl[c] = refl_colour[4 * i + c] + refl_base[4 * i + c] * ambient;
            }
        } else if (kind == 1.0f) {
            for (int c = 0; c < 3; ++c) refl[c] = refl_colour[4 * i + c];
        }
        for (int c = 0; c < 3; ++c) {
            const float base = history[4 * i + c] + emissive[4 * i + c];
            const float v = base + refl[c] * fresnel[4 * i + c];
            out[4 * i + c] = store ? oracle_f16_to_f32(oracle_f32_to_f16(v)) : v;
        }
        out[4 * i + 3] = 1.0f;
    }
    return 0;
}
