// Exact cover of every stage's tiles (svgf_sched.hpp, the functions the kernels use): for each
// frame size and stage, the tiles that stage_tiles counts, decoded as the kernels decode them,
// hold every pixel of the frame exactly once. A tile covers lattice points (i0 + i, j0 + j),
// i < 16, j < its height (16, or 32 for the tall tiles of the first CANONICAL a-trous
// iteration), of its residue (and of residue rx + 1 for a pair tile); a direct tile (rx = -1)
// covers the 256 pixels from row-major index i0.
#include "../tron-grid_amd/csrc/svgf_sched.hpp"

#include <cstdio>
#include <vector>

using namespace svgf;

static int check(int canonical, int stage, uint32_t w, uint32_t h)
{
    const uint32_t log2s = stage_log2s(canonical, stage);
    const bool tall = canonical && stage == 1;
    const bool pair = !tall && paired_stage(canonical, stage, log2s);
    const int th = tall ? 2 * LT : LT, nr = pair ? 2 : 1;
    std::vector<int> hits((size_t)w * h, 0);
    const uint32_t n = stage_tiles(canonical, stage, w, h);
    for (uint32_t t = 0; t < n; ++t) {
        const Tile T = tall ? decode_tall_tile((int)w, t) : pair ? decode_pair_tile((int)w, (int)h, log2s, t) : decode_tile((int)w, log2s, t);
        if (pair && T.rx < 0) {  // direct tile: pixels i0 .. i0 + 255, row-major
            if (!direct_lattice(w, h, log2s) || T.i0 != (int)(t * BLOCK)) {
                std::printf("bad direct tile: canonical %d stage %d %ux%u tile %u\n", canonical, stage, w, h, t);
                return 1;
            }
            for (int i = 0; i < BLOCK; ++i)
                if ((size_t)T.i0 + i < hits.size()) ++hits[(size_t)T.i0 + i];
            continue;
        }
        if (T.log2s != (int)log2s || T.rx < 0 || T.ry < 0 || T.rx + nr - 1 >= (1 << log2s) || T.ry >= (1 << log2s)) {
            std::printf("bad residue: canonical %d stage %d %ux%u tile %u\n", canonical, stage, w, h, t);
            return 1;
        }
        for (int r = 0; r < nr; ++r)
            for (int j = 0; j < th; ++j)
                for (int i = 0; i < LT; ++i) {
                    const int x = T.px(T.i0 + i) + r, y = T.py(T.j0 + j);
                    if (x < (int)w && y < (int)h) ++hits[(size_t)y * w + x];
                }
    }
    for (size_t k = 0; k < hits.size(); ++k)
        if (hits[k] != 1) {
            std::printf("pixel (%zu, %zu) covered %d times: canonical %d stage %d %ux%u (%u tiles)\n", k % w, k / w,
                        hits[k], canonical, stage, w, h, n);
            return 1;
        }
    return 0;
}

int main()
{
    const uint32_t sizes[][2] = {{1, 1},   {2, 9},     {9, 2},     {17, 33},   {31, 47},    {96, 96},   {128, 4},
                                 {130, 70}, {320, 180}, {257, 255}, {640, 480}, {1280, 720}, {8, 8},     {32, 32},
                                 {64, 64},  {128, 128}, {256, 256}, {33, 200}};
    int bad = 0, cases = 0;
    for (const auto& sz : sizes)
        for (int canonical = 0; canonical < 2; ++canonical)
            for (int stage = 0; stage < 9; ++stage, ++cases) bad += check(canonical, stage, sz[0], sz[1]);
    if (bad) return 1;
    std::printf("ok %d cases\n", cases);
    return 0;
}
