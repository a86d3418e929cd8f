// svgf_state.hpp — binary16 words of the wide state of the CANONICAL FP16 form.
//
// The colour history and the moments are fed back into the next frame (colour through the
// a = max(alpha, 1/n) blend, moments through the same recursion), so a binary16 rounding of
// each stored value contracts only by (1 - a) per frame and adds up to about u / a = 5
// half-ulps. The wide state keeps binary16 planes but carries the rounding residual with the
// value, so the feedback loop runs at 13 (colour) and ~21 (moments) significant bits:
//
//   history plane, RGBA16F:  r, g, b = binary16 (round to nearest even) of the colour;
//     the alpha slot is a 16-bit word, not a half:
//       bits 0..5    history length n (0..63)
//       bits 6..8    residual code q of r  (two's complement, -4..3)
//       bits 9..11   residual code of g
//       bits 12..14  residual code of b
//       bit 15       0
//     A channel with binary16 value h (biased exponent e, m = max(e, 1)) decodes as
//       c = h + q * 2^(m - 28)        (2^(m - 28) is an eighth of h's unit in the last place),
//     q = clamp(round(8 (c - h) / ulp(h)), -4, 3), so |c - decoded| <= ulp(h) / 8; q = 0 when
//     h is infinite or NaN.
//   moments plane, RGBA16F:  (mean_hi, sigma_hi, mean_lo, sigma_lo) with hi = binary16 of the
//     value and lo = binary16 of (value - hi) / ulp(hi), ulp(hi) = 2^(m - 25) (lo = 0 when hi is
//     not finite); a value decodes as hi + lo * ulp(hi), exact in float, within
//     ulp(hi) * 2^-12 <= 2^-22 |hi| of the value (2^-36 for subnormal hi).
//
// Every step is exact in float except the roundings named above, so the host build of this
// header (the CPU test of the codec) and the kernels produce the same words.
#pragma once

#include <cstdint>

#if defined(__HIP__)
#define SVGF_HD __host__ __device__
#else
#define SVGF_HD
#endif

namespace svgf {

constexpr uint32_t WIDE_LEN_BITS = 6;
constexpr uint32_t WIDE_LEN_MASK = (1u << WIDE_LEN_BITS) - 1u;  // history lengths up to 63
constexpr uint32_t WIDE_MAX_LEN = WIDE_LEN_MASK;

SVGF_HD inline float f32_from_bits(uint32_t b) { return __builtin_bit_cast(float, b); }

// m = max(biased exponent, 1) of a binary16 bit pattern (a 32-bit field extract: the 16-bit
// shift and mask take an instruction more on the GPU)
SVGF_HD inline uint32_t half_scale_exp(uint16_t hb)
{
    const uint32_t e = ((uint32_t)hb >> 10) & 31u;
    return e > 1u ? e : 1u;
}

// Colour channel -> (binary16, 3-bit residual code)
SVGF_HD inline _Float16 wide_encode_channel(float v, uint32_t& code)
{
    const _Float16 h = (_Float16)v;
    const uint16_t hb = __builtin_bit_cast(uint16_t, h);
    float q = 0.0f;
    if ((((uint32_t)hb >> 10) & 31u) != 31u) {
        const float inv_step = f32_from_bits((155u - half_scale_exp(hb)) << 23);  // 2^(28 - m)
        // h is finite here, so v is and q is: min / max clamp it as the comparisons would
        // (one v_med3_f32 on the GPU)
        q = __builtin_fminf(__builtin_fmaxf(__builtin_rintf((v - (float)h) * inv_step), -4.0f), 3.0f);
    }
    code = (uint32_t)(int)q & 7u;
    return h;
}

SVGF_HD inline float wide_decode_channel(_Float16 h, uint32_t code)
{
    const uint16_t hb = __builtin_bit_cast(uint16_t, h);
    const float step = f32_from_bits((half_scale_exp(hb) + 99u) << 23);  // 2^(m - 28)
    const int q = (int)(code << 29) >> 29;
    return (float)h + (float)q * step;
}

// History word of (n, codes r, g, b)
SVGF_HD inline uint16_t wide_history_word(uint32_t n, uint32_t qr, uint32_t qg, uint32_t qb)
{
    return (uint16_t)((n & WIDE_LEN_MASK) | (qr << 6) | (qg << 9) | (qb << 12));
}
SVGF_HD inline uint32_t wide_history_length(uint16_t word) { return word & WIDE_LEN_MASK; }
SVGF_HD inline uint32_t wide_history_code(uint16_t word, int channel) { return (word >> (6 + 3 * channel)) & 7u; }

// Moment value -> (hi, lo): lo is the residual in units of hi's ulp, in [-1/2, 1/2]
SVGF_HD inline void wide_encode_moment(float v, _Float16& hi, _Float16& lo)
{
    hi = (_Float16)v;
    const uint16_t hb = __builtin_bit_cast(uint16_t, hi);
    const float inv_ulp = f32_from_bits((152u - half_scale_exp(hb)) << 23);  // 2^(25 - m)
    lo = ((hb >> 10) & 31u) != 31u ? (_Float16)((v - (float)hi) * inv_ulp) : (_Float16)0.0f;
}
SVGF_HD inline float wide_decode_moment(_Float16 hi, _Float16 lo)
{
    const float ulp = f32_from_bits((half_scale_exp(__builtin_bit_cast(uint16_t, hi)) + 102u) << 23);  // 2^(m - 25)
    return (float)hi + (float)lo * ulp;
}

}  // namespace svgf
