/* The wire encoder's code estimate (tron-grid_amd/csrc/wire.hip, code_of): for x in [0, 1] the
 * estimate e = floor(255 srgb_f(x) + 1/2), with the sRGB curve evaluated in float, must lie within
 * one code of C = the count of the oracle's 255 thresholds not above x; the kernel then corrects
 * it exactly with one comparison on each side. srgb_f is perturbed by a relative +-1e-4, far
 * beyond the error of the GPU's log2 / exp2 approximations. Checked on every 16th float of
 * [0, 1] and on every float within 64 ulps of each threshold. Exit status 0: no estimate off by
 * more than one code. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

void oracle_wire_thresholds(float* t255);

static float T[255];
static uint64_t checked, bad;

static int count_of(float x)
{
    int lo = 0, hi = 255;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (T[mid - 1] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

static void check(uint32_t bits)
{
    if (bits > 0x3f800000u) return; /* [0, 1] only */
    float x;
    memcpy(&x, &bits, 4);
    const int c = count_of(x);
    const float s0 = x <= 0.0031308f ? x * 12.92f : fmaf(1.055f, exp2f(log2f(x) * (1.0f / 2.4f)), -0.055f);
    for (int p = -1; p <= 1; ++p) {
        int e = (int)fmaf(s0 * (1.0f + 1e-4f * (float)p), 255.0f, 0.5f);
        e = e < 0 ? 0 : (e > 255 ? 255 : e);
        if (e - c > 1 || c - e > 1) {
            if (bad < 5) printf("x %.9g count %d estimate %d\n", x, c, e);
            ++bad;
        }
    }
    ++checked;
}

int main(void)
{
    oracle_wire_thresholds(T);
    for (uint32_t b = 0; b <= 0x3f800000u; b += 16) check(b);
    for (int k = 0; k < 255; ++k) {
        uint32_t tb;
        memcpy(&tb, &T[k], 4);
        for (int d = -64; d <= 64; ++d) check((uint32_t)((int64_t)tb + d));
    }
    printf("%s %llu values, %llu estimates off by more than one code\n", bad ? "FAIL" : "ok",
           (unsigned long long)checked, (unsigned long long)bad);
    return bad != 0;
}
