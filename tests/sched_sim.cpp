// Host model of the persistent batched kernel's ticket protocol (k_persist in
// tron-grid_amd/csrc/svgf_kernels.hip): workgroups are threads, device atomics are std::atomic,
// and the work order comes from the same svgf_sched.hpp functions the kernel and the library
// use (build_schedule, decode_ticket, frame_tiles).
//
// Over random ragged batches (REFERENCE and CANONICAL stage lists, 1..N workers, random delays)
// it checks that
//   - decode_ticket enumerates (stage, frame, tile) stage-major, frame-major, tile-major;
//   - every ticket runs exactly once;
//   - no tile starts before every tile of the previous stage has finished (the lazily
//     published counts never run ahead of the work);
//   - no wait gives up, also with a single worker or more workers than tickets;
//   - the counters are zero after each batch, as the next batch on the same slot requires.
// Prints "ok <batches> <tickets>" and exits 0, or prints the first violation and exits 1.
#include "../tron-grid_amd/csrc/svgf_sched.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <thread>
#include <vector>

using namespace svgf;

namespace {

struct Ctl {
    std::atomic<uint32_t> ticket{0};
    std::atomic<uint32_t> exited{0};
    std::atomic<uint32_t> done[MAX_STAGES];
    Ctl()
    {
        for (auto& d : done) d = 0;
    }
};

std::atomic<int> g_fail{0};

void fail(const char* what, uint32_t a, uint32_t b, uint32_t c)
{
    if (g_fail.exchange(1) == 0) std::printf("FAIL %s (%u, %u, %u)\n", what, a, b, c);
}

struct Batch {
    uint32_t hdr[HDR_WORDS] = {};
    std::vector<uint32_t> starts = std::vector<uint32_t>((size_t)MAX_STAGES * STARTS_ROW, 0);
    uint32_t total = 0;
    // shadow state of the checker (not part of the protocol)
    std::unique_ptr<std::atomic<uint32_t>[]> runs;      // per ticket: times run
    std::unique_ptr<std::atomic<uint32_t>[]> finished;  // per stage: tiles finished
};

// One workgroup of k_persist: take tickets until none is left; publish the finished tiles of a
// stage (one add) when the next ticket belongs to another stage or none is left, before any
// wait; before the first tile of a higher stage wait until every tile of the stage below is
// published; on exit the last workgroup clears the counters.
void worker(Ctl& ctl, Batch& B, uint32_t grid, uint32_t seed, int delay_ppm)
{
    std::mt19937 rng(seed);
    const uint32_t* hdr = B.hdr;
    const uint32_t* starts = B.starts.data();
    const uint32_t nstages = hdr[HDR_STAGES];
    const uint32_t total = hdr[HDR_BASE + nstages];
    const uint32_t chunk = hdr[HDR_CHUNK];
    uint32_t t = 0, end = 0;
    uint32_t pending = 0, pending_stage = 0, ready_stage = 0;
    for (;;) {
        if (t == end) {
            const uint32_t c = ctl.ticket.fetch_add(1, std::memory_order_relaxed);
            t = c < total / chunk + 1 ? c * chunk : total;
            end = std::min(t + chunk, total);
        }
        const bool more = t < total;
        Ticket k{0, 0, 0};
        if (more) k = decode_ticket(hdr, starts, t);
        if (pending != 0 && (!more || k.stage != pending_stage)) {
            ctl.done[pending_stage].fetch_add(pending, std::memory_order_release);
            pending = 0;
        }
        if (!more) break;
        if (k.stage > ready_stage) {
            const uint32_t need = hdr[HDR_BASE + k.stage] - hdr[HDR_BASE + k.stage - 1];
            std::atomic<uint32_t>& c = ctl.done[k.stage - 1];
            const auto t0 = std::chrono::steady_clock::now();
            while (c.load(std::memory_order_acquire) < need) {
                if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10)) {
                    fail("wait gave up", t, k.stage, k.frame);
                    return;
                }
                std::this_thread::yield();
            }
            ready_stage = k.stage;
        }
        if (k.stage > 0) {
            const uint32_t need = hdr[HDR_BASE + k.stage] - hdr[HDR_BASE + k.stage - 1];
            if (B.finished[k.stage - 1].load() != need) fail("tile started before the previous stage finished", t, k.stage, k.frame);
        }
        if (B.runs[t].fetch_add(1) != 0) fail("ticket ran twice", t, k.stage, k.frame);
        if ((int)(rng() % 1000000) < delay_ppm) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
        B.finished[k.stage].fetch_add(1);
        ++pending;
        pending_stage = k.stage;
        ++t;
    }
    if (ctl.exited.fetch_add(1, std::memory_order_acq_rel) == grid - 1) {
        for (uint32_t st = 0; st < nstages; ++st) ctl.done[st].store(0, std::memory_order_relaxed);
        ctl.ticket.store(0, std::memory_order_relaxed);
        ctl.exited.store(0, std::memory_order_relaxed);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    const uint32_t seed = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 1u;
    const int batches = argc > 2 ? std::atoi(argv[2]) : 40;
    const int max_workers = argc > 3 ? std::atoi(argv[3]) : 12;
    std::mt19937 rng(seed);
    static const uint32_t sizes[][2] = {{1, 1},     {2, 9},     {9, 2},     {17, 3},    {65, 5},     {96, 96},
                                        {128, 128}, {130, 70},  {256, 256}, {320, 180}, {640, 480}, {1280, 720}};
    const int nsizes = (int)(sizeof(sizes) / sizeof(sizes[0]));
    Ctl ctl;
    uint64_t tickets = 0;
    for (int b = 0; b < batches && !g_fail.load(); ++b) {
        Batch B;
        const uint32_t n = b == 0 ? 1u : 1u + rng() % (b % 5 == 0 ? 200u : 24u);
        const int canonical = (int)(rng() & 1);
        const int iterations = 1 + (int)(rng() % 5);
        const int stages = iterations + 1;
        std::vector<uint32_t> w(n), h(n);
        for (uint32_t i = 0; i < n; ++i) {
            // mostly small frames, so that the simulation stays fast; 1280x720 now and then
            int s = (int)(rng() % (uint32_t)nsizes);
            if (s == nsizes - 1 && rng() % 4 != 0) s = (int)(rng() % 6);
            w[i] = sizes[s][0];
            h[i] = sizes[s][1];
        }
        const uint64_t total = build_schedule(
            n, [&](uint32_t i, uint32_t& ww, uint32_t& hh) { ww = w[i]; hh = h[i]; }, canonical, stages, B.hdr,
            B.starts.data());
        B.total = (uint32_t)total;
        // chunks of 1..5 tickets mostly, and up to 64 now and then
        B.hdr[HDR_CHUNK] = b % 3 == 2 ? 1u + rng() % 64u : 1u + rng() % 5u;
        // decode_ticket against a direct enumeration
        {
            uint32_t t = 0;
            for (int st = 0; st < stages; ++st)
                for (uint32_t f = 0; f < n; ++f) {
                    const uint32_t nt = stage_tiles(canonical, st, w[f], h[f]);
                    if (frame_tiles(B.starts.data(), (uint32_t)st, f) != nt) fail("frame_tiles", (uint32_t)st, f, nt);
                    for (uint32_t i = 0; i < nt; ++i, ++t) {
                        const Ticket k = decode_ticket(B.hdr, B.starts.data(), t);
                        if (k.stage != (uint32_t)st || k.frame != f || k.tile != i) fail("decode_ticket", t, k.stage, k.frame);
                    }
                }
            if (t != B.total) fail("ticket total", t, B.total, 0);
        }
        B.runs.reset(new std::atomic<uint32_t>[B.total]);
        for (uint32_t i = 0; i < B.total; ++i) B.runs[i] = 0;
        B.finished.reset(new std::atomic<uint32_t>[MAX_STAGES]);
        for (uint32_t i = 0; i < (uint32_t)MAX_STAGES; ++i) B.finished[i] = 0;

        // workers: 1, more than tickets now and then, otherwise 2..max_workers
        uint32_t grid = b % 7 == 1 ? 1u : (b % 7 == 2 ? B.total + 3u : 2u + rng() % (uint32_t)std::max(1, max_workers - 1));
        grid = std::min<uint32_t>(grid, 64);
        const int delay_ppm = (int)(rng() % 3) * 2000;
        std::vector<std::thread> th;
        for (uint32_t g = 0; g < grid; ++g) th.emplace_back(worker, std::ref(ctl), std::ref(B), grid, (uint32_t)rng(), delay_ppm);
        for (auto& x : th) x.join();
        for (uint32_t i = 0; i < B.total && !g_fail.load(); ++i)
            if (B.runs[i].load() != 1) fail("ticket did not run once", i, B.runs[i].load(), 0);
        if (ctl.ticket.load() != 0 || ctl.exited.load() != 0) fail("ticket/exited not reset", ctl.ticket.load(), ctl.exited.load(), 0);
        for (uint32_t i = 0; i < (uint32_t)MAX_STAGES; ++i)
            if (ctl.done[i].load() != 0) fail("completion counter not reset", i, ctl.done[i].load(), 0);
        tickets += B.total;
    }
    if (g_fail.load()) return 1;
    std::printf("ok %d %llu\n", batches, (unsigned long long)tickets);
    return 0;
}
