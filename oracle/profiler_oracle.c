/* profiler_oracle.c — CPU restatement of the reference's GPU pass profiler loop (test
 * infrastructure only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker; the product path never uses it.
 *
 * src/main.cpp, per iteration i of the render loop (frame counter F = first + i, slot
 * c = i mod slots):
 *   :2910-2951  if slot c holds an earlier frame (it does from iteration `slots` on), read its
 *               samples back: the first readback sets ema = sample, later ones
 *               ema = ema * (1 - alpha) + sample * alpha (float); if F % interval == 0 the log
 *               line is composed;
 *   :3043       ++F;
 *   :3096       slot c now holds this iteration's samples;
 *   :3148       c = (c + 1) mod slots.
 * samples: frames x passes durations (ns) of each iteration's passes. Outputs per iteration:
 * ema (frames x passes, after that iteration's readback; 0 before the first) and logged
 * (1 where a line is composed).
 */
#include "svgf_oracle.h"

#include <stdint.h>

int oracle_profiler_run(int passes, int slots, int frames, const float* samples, float alpha, uint32_t interval,
                        uint32_t first, float* ema, uint8_t* logged)
{
    if (passes < 1 || slots < 1 || frames < 0 || !samples || !ema || !logged || interval < 1) return 1;
    int init = 0;
    float cur[64];
    if (passes > 64) return 1;
    for (int k = 0; k < passes; ++k) cur[k] = 0.0f;
    uint64_t counter = first;
    for (int i = 0; i < frames; ++i) {
        logged[i] = 0;
        if (i >= slots) {
            const float* s = samples + (int64_t)(i - slots) * passes;
            if (!init) {
                for (int k = 0; k < passes; ++k) cur[k] = s[k];
                init = 1;
            } else {
                for (int k = 0; k < passes; ++k) cur[k] = (cur[k] * (1.0f - alpha)) + (s[k] * alpha);
            }
            if (counter % interval == 0) logged[i] = 1;
        }
        for (int k = 0; k < passes; ++k) ema[(int64_t)i * passes + k] = cur[k];
        ++counter;
    }
    return 0;
}
