// synth.cpp — seeded synthetic G-buffer sequences for the denoiser (host code, no GPU).
//
// Stands in for TronGrid's producer, the mesh-shader pass that writes the noisy combined
// lighting into lighting_raw (src/mesh.slang:995-998), and for the G-buffer a canonical SVGF
// needs. Every value is a pure function of (seed, stream, frame, pixel), so generation order
// and thread count never change the data.
//
// Scene: the camera pans `pan_px` pixels per frame, so frame f samples a static world at
// u = x + f * pan_px, v = y, and the previous position of pixel x is x + pan_px (motion =
// (+pan_px, 0)). Depth is a smooth field in [0.1, 100] m, normals come from a procedural
// heightfield, radiance is a smooth 0-10 gradient (tinted) times per-pixel noise with mean 1:
// exponential (1 spp) or lognormal (sigma 1.5, mu -1.125), or none (the noise-free radiance
// both have as their mean: the ground truth of a denoised sequence). A `disocclusion` fraction of the
// pixels gets its depth scaled by 1.5, which the history depth test rejects.
#include "../../include/svgf.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t stream, uint64_t frame, uint64_t pixel, uint64_t lane)
{
    uint64_t h = splitmix64(seed ^ 0x5356474600000000ull);
    h = splitmix64(h ^ stream);
    h = splitmix64(h ^ frame);
    h = splitmix64(h ^ pixel);
    return splitmix64(h ^ lane);
}

// Uniform in (0, 1): 53 random bits, never 0 or 1.
inline double u01(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

constexpr double PI = 3.14159265358979323846;

struct Gen {
    const svgf_synth_desc* d;
    double W, H;

    double depth(double u, double v) const
    {
        double t = 0.5 + 0.32 * std::sin(2.0 * PI * (1.3 * u / W) + 0.7) * std::cos(2.0 * PI * (0.8 * v / H)) +
                   0.12 * std::sin(2.0 * PI * (u + 0.5 * v) / 97.0);
        t = std::min(1.0, std::max(0.0, t));
        return 0.1 * std::pow(1000.0, t);
    }
    // heightfield h = A (sin(k1 u) cos(k2 v) + 0.5 sin(k3 u + k4 v)); normal = normalize(-dh/du, -dh/dv, 1)
    void normal(double u, double v, double n[3]) const
    {
        const double A = 6.0, k1 = 2.0 * PI / 41.0, k2 = 2.0 * PI / 67.0, k3 = 2.0 * PI / 23.0, k4 = 2.0 * PI / 31.0;
        double dhu = A * (k1 * std::cos(k1 * u) * std::cos(k2 * v) + 0.5 * k3 * std::cos(k3 * u + k4 * v));
        double dhv = A * (-k2 * std::sin(k1 * u) * std::sin(k2 * v) + 0.5 * k4 * std::cos(k3 * u + k4 * v));
        double l = std::sqrt(dhu * dhu + dhv * dhv + 1.0);
        n[0] = -dhu / l;
        n[1] = -dhv / l;
        n[2] = 1.0 / l;
    }
    void base(double u, double v, double c[3]) const
    {
        double g = 10.0 * (0.62 * u / W + 0.38 * v / H);
        g = std::min(10.0, std::max(0.0, g));
        c[0] = g;
        c[1] = g * (0.85 + 0.1 * std::sin(u / 50.0));
        c[2] = g * (0.7 + 0.2 * std::cos(v / 40.0));
    }
    double noise(uint64_t pixel) const
    {
        uint64_t h1 = key(d->seed, d->stream, d->frame, pixel, 1);
        if (d->noise == 2) return 1.0;
        if (d->noise == 0) return -std::log(u01(h1));
        uint64_t h2 = key(d->seed, d->stream, d->frame, pixel, 2);
        double nrm = std::sqrt(-2.0 * std::log(u01(h1))) * std::cos(2.0 * PI * u01(h2));
        return std::exp(-1.125 + 1.5 * nrm);
    }
};

void rows(const Gen& g, uint32_t y0, uint32_t y1, float* rad, float* gb, float* mv)
{
    const svgf_synth_desc* d = g.d;
    for (uint32_t y = y0; y < y1; ++y) {
        for (uint32_t x = 0; x < d->width; ++x) {
            uint64_t pix = (uint64_t)y * d->width + x;
            double u = x + (double)d->frame * d->pan_px, v = y;
            bool dis = d->disocclusion > 0.0f && u01(key(d->seed, d->stream, d->frame, pix, 3)) < d->disocclusion;
            if (rad) {
                double c[3];
                g.base(u, v, c);
                double nz = g.noise(pix);
                rad[4 * pix + 0] = (float)(c[0] * nz);
                rad[4 * pix + 1] = (float)(c[1] * nz);
                rad[4 * pix + 2] = (float)(c[2] * nz);
                rad[4 * pix + 3] = 1.0f;
            }
            if (gb) {
                double n[3];
                g.normal(u, v, n);
                double z = g.depth(u, v);
                if (dis) z *= 1.5;
                gb[4 * pix + 0] = (float)n[0];
                gb[4 * pix + 1] = (float)n[1];
                gb[4 * pix + 2] = (float)n[2];
                gb[4 * pix + 3] = (float)z;
            }
            if (mv) {
                mv[2 * pix + 0] = d->pan_px;
                mv[2 * pix + 1] = 0.0f;
            }
        }
    }
}

}  // namespace

extern "C" svgf_status svgf_synth_frame(const svgf_synth_desc* d, float* radiance, float* gbuffer, float* motion)
{
    if (!d || d->width < 1 || d->height < 1 || d->noise > 2 || !(d->disocclusion >= 0.0f && d->disocclusion <= 1.0f))
        return SVGF_ERR_INVALID_ARGUMENT;
    Gen g{d, (double)d->width, (double)d->height};
    uint32_t nt = d->threads ? d->threads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<uint32_t>(nt, d->height);
    nt = std::min<uint32_t>(nt, 64);
    if (nt <= 1) {
        rows(g, 0, d->height, radiance, gbuffer, motion);
        return SVGF_OK;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        uint32_t y0 = (uint32_t)((uint64_t)d->height * t / nt), y1 = (uint32_t)((uint64_t)d->height * (t + 1) / nt);
        th.emplace_back(rows, std::cref(g), y0, y1, radiance, gbuffer, motion);
    }
    for (auto& t : th) t.join();
    return SVGF_OK;
}
