// Host build of the wide FP16 state codec (tron-grid_amd/csrc/svgf_state.hpp, the functions the
// kernels use), checked against bounds computed here independently of it:
//   colour channel: |v - decode(encode(v))| <= ulp(h) / 8, with ulp(h) = 2^(E - 10) for a normal
//     binary16 h of exponent E and 2^-24 for a subnormal (from frexp, not from the bit fields);
//     16 / 16 codes of each colour word survive the packing with every history length;
//   moment: |v - decode| <= 2^-22 |v| + 2^-36 (binary16 rounding of the residual in units of
//     ulp(h), at most 1/2);
//   infinities and NaN stay infinite / NaN with code 0 and lo 0.
// Writes (v, h, code) of every colour sample to argv[1] for the numpy decoder's comparison.
#include "../tron-grid_amd/csrc/svgf_state.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace svgf;

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static double unit()
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

static double ulp_of_half(float h)
{
    const double a = std::fabs((double)h);
    if (a < std::ldexp(1.0, -14)) return std::ldexp(1.0, -24);  // subnormal (and zero)
    int e;
    std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1): binary16 exponent E = e - 1
    return std::ldexp(1.0, e - 1 - 10);
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    FILE* dump = std::fopen(argv[1], "wb");
    if (!dump) return 2;
    int bad = 0;
    long n = 0;
    double worst_c = 0.0, worst_m = 0.0;
    for (int i = 0; i < 2000000; ++i) {
        // log-uniform magnitudes from below the subnormal range to just under the overflow, both signs
        const double mag = std::exp2(-30.0 + 46.0 * unit());
        float v = (float)(unit() < 0.1 ? -mag : mag);
        if (i < 64) v = (float)std::ldexp(1.0 + i / 64.0, i % 2 ? -14 : 0);  // binade edges
        uint32_t code;
        const _Float16 h = wide_encode_channel(v, code);
        const float d = wide_decode_channel(h, code);
        const double err = std::fabs((double)v - (double)d);
        if (std::isfinite((double)(float)h)) {
            const double bound = ulp_of_half((float)h) / 8.0;
            if (err > bound * (1.0 + 1e-12)) {
                if (bad++ < 10) std::printf("colour %.9g -> %.9g (h %.9g, code %u) err %.3g > %.3g\n", v, d, (float)h, code, err, bound);
            }
            if (err / std::fabs((double)v) > worst_c && std::fabs(v) > 1e-4) worst_c = err / std::fabs((double)v);
        }
        float rec[3] = {v, (float)h, (float)code};
        std::fwrite(rec, sizeof rec, 1, dump);
        ++n;
        // history words: every length with this code in every channel slot
        for (uint32_t len = 0; len <= WIDE_MAX_LEN; len += 21) {
            const uint16_t w = wide_history_word(len, code, (code + 3) & 7u, (code + 5) & 7u);
            if (wide_history_length(w) != len || wide_history_code(w, 0) != code || wide_history_code(w, 1) != ((code + 3) & 7u) ||
                wide_history_code(w, 2) != ((code + 5) & 7u) || (w >> 15) != 0) {
                if (bad++ < 10) std::printf("word %u/%u -> %04x\n", len, code, w);
            }
        }
        _Float16 hi, lo;
        wide_encode_moment(v, hi, lo);
        const double em = std::fabs((double)v - (double)wide_decode_moment(hi, lo));
        const double bm = std::ldexp(std::fabs((double)v), -22) + std::ldexp(1.0, -36);
        if (std::isfinite((double)(float)hi)) {
            if (em > bm && bad++ < 10) std::printf("moment %.9g err %.3g > %.3g\n", v, em, bm);
            if (std::fabs(v) > 1e-4 && em / std::fabs((double)v) > worst_m) worst_m = em / std::fabs((double)v);
        }
    }
    const float specials[4] = {INFINITY, -INFINITY, NAN, 70000.0f};
    for (float v : specials) {
        uint32_t code;
        const _Float16 h = wide_encode_channel(v, code);
        const float d = wide_decode_channel(h, code);
        _Float16 hi, lo;
        wide_encode_moment(v, hi, lo);
        const float m = wide_decode_moment(hi, lo);
        const bool ok = code == 0 && (float)lo == 0.0f &&
                        (std::isnan(v) ? std::isnan(d) && std::isnan(m) : std::isinf(d) && std::isinf(m) && (d > 0) == (v > 0));
        if (!ok && bad++ < 10) std::printf("special %g -> %g / %g (code %u)\n", v, d, m, code);
    }
    std::fclose(dump);
    std::printf("%ld samples, worst relative error above 1e-4: colour %.3g, moment %.3g, %d failures\n", n, worst_c, worst_m, bad);
    return bad ? 1 : 0;
}
