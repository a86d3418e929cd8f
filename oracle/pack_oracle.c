/* pack_oracle.c — CPU restatement of the producer side of the denoiser input (test
 * infrastructure only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's CPU-baseline leg as the checker;
 * the product path never uses it.
 *
 * TronGrid's mesh pass writes the denoiser input only where opaque geometry covers a pixel,
 * and the frame clears it to zero first so that sky and transparent-only pixels enter the
 * denoiser as 0 (src/main.cpp:595-646):
 *   covered:   radiance = ((indirect + direct_diffuse) * ao + direct_specular, 1)
 *              (src/mesh.slang:995-998)
 *   otherwise: radiance = (0, 0, 0, 0)
 * The pack writes every pixel once, so the clear is folded into the same pass. It also builds
 * the G-buffer and motion planes the canonical denoiser reads:
 *   normal  = unpackOctahedralNormal(normal_oct) (src/mesh.slang:341-352): the reference's
 *             16 + 16-bit octahedral encoding, normalised with a correctly rounded square root
 *             and division;
 *   depth   = (VP (p, 1)).w, the view-space distance along the camera axis of a perspective
 *             projection;
 *   motion  = previous screen position - pixel centre, in pixels, where the previous position
 *             follows src/mesh.slang:589-591: clip = PVP (p, 1), ndc = clip.xy / max(clip.w,
 *             1e-6), uv = ndc (0.5, -0.5) + 0.5, position = uv (W, H). A point with
 *             clip.w <= 0 (behind the previous camera, the reference's prev_in_screen test at
 *             :594-596), an uncovered pixel, or a motion beyond +-1e6 px (or not finite) gets
 *             the motion (-1e6, -1e6): no previous position inside any frame.
 *   uncovered pixels: gbuffer = (0, 0, 0, 0).
 * Coverage: position.w > 0.
 *
 * Arithmetic: float, one rounding per operation, no contraction (the Makefile builds with
 * -ffp-contract=off); matrix rows are summed left to right, ((m0 x + m1 y) + m2 z) + m3.
 */
#include "svgf_oracle.h"

#include <math.h>
#include <stdint.h>

#define PACK_NO_MOTION (-1.0e6f)

static float row4(const float* m, int r, float x, float y, float z)
{
    return ((m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z) + m[4 * r + 3];
}

void oracle_unpack_octahedral(uint32_t packed, float* n3)
{
    const float qx = (float)(packed & 0xFFFFu) / 65535.0f;
    const float qy = (float)((packed >> 16) & 0xFFFFu) / 65535.0f;
    const float px = qx * 2.0f - 1.0f;
    const float py = qy * 2.0f - 1.0f;
    float nx = px, ny = py;
    const float nz = (1.0f - fabsf(px)) - fabsf(py);
    if (nz < 0.0f) {
        nx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        ny = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    }
    const float inv = 1.0f / sqrtf((nx * nx + ny * ny) + nz * nz);
    n3[0] = nx * inv;
    n3[1] = ny * inv;
    n3[2] = nz * inv;
}

int oracle_pack(int w, int h, const float* vp, const float* pvp, const float* position, const uint32_t* normal_oct,
                const float* indirect, const float* direct_diffuse, const float* direct_specular, int store,
                float* radiance, float* gbuffer, float* motion)
{
    if (w < 1 || h < 1 || !vp || !pvp || !position || !normal_oct || !indirect || !direct_diffuse ||
        !direct_specular || !radiance || !gbuffer || !motion)
        return 1;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int64_t i = (int64_t)y * w + x;
            const float* p = position + 4 * i;
            float* rad = radiance + 4 * i;
            float* g = gbuffer + 4 * i;
            float* mv = motion + 2 * i;
            if (!(p[3] > 0.0f)) {
                rad[0] = rad[1] = rad[2] = rad[3] = 0.0f;
                g[0] = g[1] = g[2] = g[3] = 0.0f;
                mv[0] = mv[1] = PACK_NO_MOTION;
                continue;
            }
            const float ao = direct_diffuse[4 * i + 3];
            for (int c = 0; c < 3; ++c) {
                const float v = (indirect[4 * i + c] + direct_diffuse[4 * i + c]) * ao + direct_specular[4 * i + c];
                rad[c] = store ? oracle_f16_to_f32(oracle_f32_to_f16(v)) : v;
            }
            rad[3] = 1.0f;
            oracle_unpack_octahedral(normal_oct[i], g);
            g[3] = row4(vp, 3, p[0], p[1], p[2]);
            const float cw = row4(pvp, 3, p[0], p[1], p[2]);
            float mx = PACK_NO_MOTION, my = PACK_NO_MOTION;
            if (cw > 0.0f) {
                const float d = cw > 1e-6f ? cw : 1e-6f;
                const float nx = row4(pvp, 0, p[0], p[1], p[2]) / d;
                const float ny = row4(pvp, 1, p[0], p[1], p[2]) / d;
                mx = (nx * 0.5f + 0.5f) * (float)w - ((float)x + 0.5f);
                my = (ny * -0.5f + 0.5f) * (float)h - ((float)y + 0.5f);
                if (!(fabsf(mx) <= 1.0e6f && fabsf(my) <= 1.0e6f)) mx = my = PACK_NO_MOTION;
            }
            mv[0] = mx;
            mv[1] = my;
        }
    }
    return 0;
}
