// svgf_sched.hpp — tile decomposition and the work order of the batched kernels.
//
// Plain integer code with no HIP dependency: the kernels use it on the device, and a host
// simulation of the persistent kernel's ticket protocol (tests/sched_sim.cpp) compiles the
// same functions with g++.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define SVGF_HD __host__ __device__
#else
#define SVGF_HD
#endif

namespace svgf {

// Stages run on tiles of a stride-s lattice: for a-trous iteration k (s = 2^k) the pixels of one
// residue class (x mod s, y mod s) form a lattice whose 5x5 taps are the lattice neighbours, so a
// tile plus a 2-point halo holds every tap in LDS whatever the stride. Pointwise stages use s = 1
// (plain pixel tiles). With 16x16 tiles a frame at stride s has
// s^2 * ceil(ceil(w/s)/16) * ceil(ceil(h/s)/16) tiles, ordered tile-major with the residue
// fastest (the s^2 blocks of one 16s x 16s pixel region are adjacent); the CANONICAL a-trous
// iterations use the residue-pair and tall variants below.
constexpr int LT = 16;
constexpr int LHALO = 2;
constexpr int LREG = LT + 2 * LHALO;  // 20
constexpr int BLOCK = LT * LT;        // 256 threads
// Blocks of one launch: its work-items (blocks x BLOCK) must fit in 32 bits.
constexpr uint32_t MAX_LAUNCH_BLOCKS = 0xFFFFFFFFu / BLOCK;

SVGF_HD inline uint32_t tiles_at(uint32_t w, uint32_t h, uint32_t log2s)
{
    uint32_t s = 1u << log2s;
    uint32_t lx = (w + s - 1) >> log2s, ly = (h + s - 1) >> log2s;
    return (s * s) * ((lx + LT - 1) / LT) * ((ly + LT - 1) / LT);
}

// Stage list of a mode: REFERENCE = init + N a-trous; CANONICAL = temporal + N a-trous (the
// spatial variance of young pixels runs inside the first a-trous iteration). log2 stride of
// stage i:
SVGF_HD inline uint32_t stage_log2s(int /*canonical*/, int stage) { return stage < 1 ? 0u : (uint32_t)(stage - 1); }
// CANONICAL a-trous iterations at strides s >= 2 run residue pairs: one workgroup filters the
// lattices of residues (rx, ry) and (rx + 1, ry) (rx even), whose entries are adjacent pixels.
// A halo entry then brings two pixels of one cache line instead of one, which halves the L2 ->
// L1 requests of the large strides (each lattice row touches one line per entry), and the
// workgroup's 256 threads each filter two lattice points of one residue that share most of
// their taps. Such a stage has s^2 / 2 * ceil(ceil(w/s)/16) * ceil(ceil(h/s)/16) tiles.
constexpr uint32_t PAIR_LOG2S = 1;

SVGF_HD inline bool paired_stage(int canonical, int stage, uint32_t log2s)
{
    return canonical && stage >= 2 && log2s >= PAIR_LOG2S;
}

// A frame whose stride-s lattice has at most DIRECT_MAX_POINTS points (ceil(w/s) ceil(h/s): an
// 8x8 agent eye from stride 2 on, a 64x64 one at stride 16) would leave most of every pair tile's
// 2 x 16 x 16 slots idle, each tile staging its whole region for a few points. Such a frame
// runs direct tiles instead: 256 consecutive pixels (row-major) per tile, one per thread, every
// tap loaded from memory (adjacent lanes' taps are adjacent pixels), ceil(w h / 256) tiles.
constexpr uint32_t DIRECT_MAX_POINTS = 40;

SVGF_HD inline bool direct_lattice(uint32_t w, uint32_t h, uint32_t log2s)
{
    const uint32_t s = 1u << log2s;
    return ((w + s - 1) >> log2s) * ((h + s - 1) >> log2s) <= DIRECT_MAX_POINTS;
}

SVGF_HD inline uint32_t pair_tiles_at(uint32_t w, uint32_t h, uint32_t log2s)
{
    if (direct_lattice(w, h, log2s)) return (w * h + BLOCK - 1) / BLOCK;
    return tiles_at(w, h, log2s) / 2;
}

// The first CANONICAL a-trous iteration (stride 1, with the spatial variance of young pixels)
// runs on tall 16 x 32 pixel tiles, whose 256 threads each filter two vertically adjacent
// pixels: ceil(w/16) * ceil(h/32) tiles.
SVGF_HD inline uint32_t tall_tiles_at(uint32_t w, uint32_t h)
{
    return ((w + LT - 1) / LT) * ((h + 2 * LT - 1) / (2 * LT));
}

// Tiles (workgroups) of stage `stage` for a w x h frame.
SVGF_HD inline uint32_t stage_tiles(int canonical, int stage, uint32_t w, uint32_t h)
{
    const uint32_t log2s = stage_log2s(canonical, stage);
    if (canonical && stage == 1) return tall_tiles_at(w, h);
    return paired_stage(canonical, stage, log2s) ? pair_tiles_at(w, h, log2s) : tiles_at(w, h, log2s);
}

// A tile: lattice points (i0 + i, j0 + j) of residue (rx, ry) at stride 2^log2s (pixel
// (rx + (i0 + i) 2^log2s, ry + (j0 + j) 2^log2s)); a pair tile also holds residue rx + 1.
struct Tile {
    int log2s, rx, ry, i0, j0;
    SVGF_HD int px(int i) const { return rx + (i << log2s); }
    SVGF_HD int py(int j) const { return ry + (j << log2s); }
};

SVGF_HD inline Tile decode_tile(int w, uint32_t log2s, uint32_t t)
{
    uint32_t s = 1u << log2s;
    uint32_t lx = ((uint32_t)w + s - 1) >> log2s;
    uint32_t ntx = (lx + LT - 1) / LT;
    uint32_t res = t & (s * s - 1);
    uint32_t tile = t >> (2 * log2s);
    uint32_t ty = tile / ntx, tx = tile - ty * ntx;
    return Tile{(int)log2s, (int)(res & (s - 1)), (int)(res >> log2s), (int)(tx * LT), (int)(ty * LT)};
}

// Tile t of a residue-pair stage (stride s >= 2): residues (rx, ry) and (rx + 1, ry), rx even,
// ordered as decode_tile orders single residues; or, for a frame of direct tiles, rx = -1 and
// i0 = its first pixel (row-major index).
SVGF_HD inline Tile decode_pair_tile(int w, int h, uint32_t log2s, uint32_t t)
{
    if (direct_lattice((uint32_t)w, (uint32_t)h, log2s)) return Tile{(int)log2s, -1, 0, (int)(t * BLOCK), 0};
    const uint32_t s = 1u << log2s;
    const uint32_t lx = ((uint32_t)w + s - 1) >> log2s;
    const uint32_t ntx = (lx + LT - 1) / LT;
    const uint32_t res = t & (s * s / 2 - 1);
    const uint32_t tile = t >> (2 * log2s - 1);
    const uint32_t ty = tile / ntx, tx = tile - ty * ntx;
    return Tile{(int)log2s, (int)(2 * (res & (s / 2 - 1))), (int)(res >> (log2s - 1)), (int)(tx * LT), (int)(ty * LT)};
}

// Tile t of the first CANONICAL a-trous iteration (stride 1): 16 x 32 pixels, row-major.
SVGF_HD inline Tile decode_tall_tile(int w, uint32_t t)
{
    const uint32_t ntx = ((uint32_t)w + LT - 1) / LT;
    const uint32_t ty = t / ntx, tx = t - ty * ntx;
    return Tile{0, 0, 0, (int)(tx * LT), (int)(ty * 2 * LT)};
}

constexpr int MAX_STAGES = 12;
constexpr uint32_t MAX_BATCH_FRAMES = 1024;
constexpr uint32_t STARTS_ROW = MAX_BATCH_FRAMES + 1;

// Header words of a batch table (16 words):
//   hdr[HDR_FRAMES]          frames in the batch (>= 1)
//   hdr[HDR_STAGES]          stages per frame (<= MAX_STAGES)
//   hdr[HDR_BASE + i]        first ticket of stage i (i <= stages; the last entry is the total)
// starts[i * STARTS_ROW + f] is the first tile of frame f within stage i (prefix sums of the
// frames' tile counts; starts[i * STARTS_ROW + frames] = tiles of stage i).
//   hdr[HDR_CHUNK]           tiles per ticket-counter increment of the persistent kernel
constexpr int HDR_FRAMES = 0;
constexpr int HDR_STAGES = 1;
constexpr int HDR_BASE = 2;
constexpr int HDR_CHUNK = 15;
constexpr int HDR_WORDS = 16;
static_assert(HDR_BASE + MAX_STAGES + 1 <= HDR_CHUNK, "batch header too small");

// Largest frame index f with row[f] <= t (row is non-decreasing, row[0] = 0 <= t).
SVGF_HD inline uint32_t frame_of(const uint32_t* row, uint32_t frames, uint32_t t)
{
    uint32_t lo = 0, hi = frames - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (row[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The persistent batched kernel hands out tickets 0 .. total-1 from one atomic counter, in
// chunks of hdr[HDR_CHUNK] consecutive tickets per increment. Ticket order is stage-major (every tile of stage i before any tile of stage i + 1), frame-major
// inside a stage. A tile of stage i > 0 may start once every tile of stage i - 1 has finished;
// those tiles all hold smaller tickets, so this order is a topological order of the
// dependencies.
struct Ticket {
    uint32_t stage, frame, tile;
};

SVGF_HD inline Ticket decode_ticket(const uint32_t* hdr, const uint32_t* starts, uint32_t t)
{
    const uint32_t ns = hdr[HDR_STAGES];
    uint32_t st = 0;
    while (st + 1 < ns && hdr[HDR_BASE + st + 1] <= t) ++st;
    const uint32_t local = t - hdr[HDR_BASE + st];
    const uint32_t* row = starts + (uint64_t)st * STARTS_ROW;
    const uint32_t f = frame_of(row, hdr[HDR_FRAMES], local);
    return Ticket{st, f, local - row[f]};
}

// Fills the header and the starts rows of a batch of n frames (1 <= n <= MAX_BATCH_FRAMES)
// for a mode with `stages` stages; size(i, w, h) gives frame i's size. Returns the total
// number of tickets (tiles over all stages); the header's ticket fields are valid only when
// that total is below 2^32.
template <class SizeOf>
inline uint64_t build_schedule(uint32_t n, SizeOf size, int canonical, int stages, uint32_t* hdr, uint32_t* starts)
{
    uint64_t ticket = 0;
    for (int st = 0; st < stages; ++st) {
        uint32_t* row = starts + (uint64_t)st * STARTS_ROW;
        uint32_t tile = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t w = 0, h = 0;
            size(i, w, h);
            row[i] = tile;
            tile += stage_tiles(canonical, st, w, h);
        }
        row[n] = tile;
        hdr[HDR_BASE + st] = (uint32_t)ticket;
        ticket += tile;
    }
    hdr[HDR_BASE + stages] = (uint32_t)ticket;
    hdr[HDR_FRAMES] = n;
    hdr[HDR_STAGES] = (uint32_t)stages;
    return ticket;
}

// Tiles of frame f in stage st.
SVGF_HD inline uint32_t frame_tiles(const uint32_t* starts, uint32_t st, uint32_t f)
{
    const uint32_t* row = starts + (uint64_t)st * STARTS_ROW;
    return row[f + 1] - row[f];
}

// Counters of the persistent kernel, each on a 256-byte line of its own (workgroups poll the
// stage counters while others add to the ticket counter). The kernel zeroes what it used
// before it ends.
constexpr int CTL_LINE = 64;  // words
struct PersistCtl {
    uint32_t ticket[CTL_LINE];                // [0]: next chunk of tickets
    uint32_t exited[CTL_LINE];                // [0]: workgroups that found no ticket left
    uint32_t abort[CTL_LINE];                 // [0]: set by the first wait that gives up
    uint32_t done[MAX_STAGES * CTL_LINE];     // [i * CTL_LINE]: published tiles of stage i
};

}  // namespace svgf
