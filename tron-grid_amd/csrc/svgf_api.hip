// svgf_api.hip — the C ABI of include/svgf.h over HIP.
//
// Replaces TronGrid's SVGF host glue: pipeline / descriptor creation (src/main.cpp:2288-2446)
// becomes svgf_create; the per-frame dispatch block of recordFrame (src/main.cpp:907-947)
// becomes svgf_denoise; image allocation + zero-init + resize reset (src/main.cpp:1326-1496,
// 1570-1603) become caller-owned planes (single frame) or library-resident streams (batched).
#include "../../include/svgf.h"
#include "svgf_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

using namespace svgf;

namespace {

constexpr int BATCH_SLOTS = 3;
// Outcomes of completed batches kept for svgf_batch_status after their slot is reused (a
// caller with several batches in flight, such as the serving loop's lanes, may read a batch's
// outcome only after later batches have cycled through the slots).
constexpr int BATCH_OUTCOMES = 64;
constexpr size_t ALIGN = 256;

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

uint32_t bytes_of(uint32_t fmt)
{
    switch (fmt) {
    case SVGF_FMT_RGBA32F: return 16;
    case SVGF_FMT_RGBA16F: return 8;
    case SVGF_FMT_RG32F: return 8;
    case SVGF_FMT_RG16F: return 4;
    default: return 0;
    }
}

struct StreamSlot {
    bool open = false;
    uint32_t w = 0, h = 0;
    uint32_t frames = 0;
    int parity = 0;  // state[parity] holds what the next frame reads
    void* block = nullptr;
    uint64_t bytes = 0;
    PlaneDev hist[2]{}, mom[2]{}, gbuf[2]{};
    // The last frame was requested with SVGF_REQUEST_KEEP_GBUFFER: the next one reads that
    // request's G-buffer (gkept) instead of gbuf[parity].
    bool g_kept = false;
    PlaneDev gkept{};
};

struct BatchSlot {
    void* staging = nullptr;  // pinned host copy of the table
    void* table = nullptr;    // device table
    void* arena = nullptr;    // per-frame scratch planes
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
    std::vector<uint32_t> ids;  // streams the pending batch touches
    std::vector<hipGraphNode_t> knodes;      // captured stage kernels, launch order
    std::vector<hipKernelNodeParams> kparams;
    PersistCtl* ctl = nullptr;  // persistent form: ticket and completion counters (zero between launches)
    uint32_t* err_host = nullptr;  // persistent form: host-mapped timeout flag
    uint32_t* err_dev = nullptr;
    hipEvent_t sev[MAX_STAGES + 1]{};  // SVGF_BATCH_TIMED_STAGES: events around the batch's stage launches
    int sev_n = 0;                      // stages those events bracket in the slot's last batch
    uint64_t seq = 0;        // id of the batch the slot last ran (0: none)
    hipStream_t stream = nullptr;  // the hip_stream it was enqueued on
    bool checked = false;    // its outcome has been read (after completion)
    bool failed = false;     // a dependency wait of its persistent launch gave up
};

}  // namespace

struct svgf_context {
    svgf_desc desc{};
    KParams kp{};
    Storage st = Storage::F32;
    int mode = SVGF_MODE_REFERENCE;
    uint32_t s4 = 16, s2 = 8;  // storage bytes of colour-like and moment planes
    uint32_t w4 = 16;          // bytes of a wavelet-scratch pixel (FP32 in CANONICAL mode)

    void* scratch = nullptr;
    hipEvent_t scratch_ev = nullptr;     // recorded after every single-frame call
    hipStream_t scratch_stream = nullptr;  // the stream of the last single-frame call
    bool scratch_used = false;

    std::vector<StreamSlot> streams;
    uint64_t budget_used = 0;
    // Freed stream-state blocks kept for reuse: closing and reopening streams (churn) then
    // needs no hipFree, which would synchronise the whole device.
    std::multimap<uint64_t, void*> pool;
    uint64_t pool_bytes = 0;

    BatchSlot slots[BATCH_SLOTS];
    int next_slot = 0;
    uint64_t last_seq = 0;  // id of the last batch enqueued
    struct Outcome {
        uint64_t seq = 0;
        bool failed = false;
    } outcomes[BATCH_OUTCOMES];  // [seq % BATCH_OUTCOMES]: checked batches
    size_t table_bytes = 0, starts_off = 0, frames_off = 0, arena_bytes = 0;
    int grid_blocks = 0;
    uint32_t persist_blocks = 0;  // resident workgroups of the persistent kernel on the device
    int cus = 1;                  // compute units of the device
    uint64_t max_polls = 0;       // persistent form: longest dependency wait, in polls
    hipStream_t capture_stream = nullptr;

    // SVGF_PROFILE_STAGES
    hipEvent_t ev[MAX_STAGES + 1]{};
    int ev_stages = 0;
    // SVGF_BATCH_TIMED_STAGES: per-stage GPU time summed over completed batches
    double stage_sum[MAX_STAGES]{};
    int stage_n = 0;
    uint64_t stage_batches = 0;
};

namespace {

svgf_status validate_params(const svgf_params& p, uint32_t mode, uint32_t storage)
{
    if (p.iterations < 1 || p.iterations > 10) return SVGF_ERR_INVALID_ARGUMENT;
    auto finite_nonneg = [](float v) { return v >= 0.0f && v < 3.0e38f; };
    if (!finite_nonneg(p.phi_luma) || !finite_nonneg(p.alpha_moments) || p.alpha_moments > 1.0f) return SVGF_ERR_INVALID_ARGUMENT;
    if (!(p.variance_eps > 0.0f) || !finite_nonneg(p.variance_eps)) return SVGF_ERR_INVALID_ARGUMENT;
    if (mode == SVGF_MODE_CANONICAL) {
        if (!(p.phi_luma > 0.0f) || !finite_nonneg(p.phi_normal) || !finite_nonneg(p.phi_depth) || !finite_nonneg(p.alpha_color) ||
            p.alpha_color > 1.0f || !(p.depth_eps > 0.0f) || !finite_nonneg(p.depth_reject) ||
            !(p.normal_reject >= -1.0f && p.normal_reject <= 1.0f) || p.history_cap < 1 || p.history_cap > 255)
            return SVGF_ERR_INVALID_ARGUMENT;
    }
    if (p.flags & ~(SVGF_FLAG_COMPENSATED_MOMENTS | SVGF_FLAG_WIDE_STATE)) return SVGF_ERR_INVALID_ARGUMENT;
    if (mode == SVGF_MODE_REFERENCE && p.flags) return SVGF_ERR_UNSUPPORTED;
    if (p.flags & SVGF_FLAG_WIDE_STATE) {
        // binary16 words of compensated moments; the length field has 6 bits
        if (storage != SVGF_STORAGE_FP16 || !(p.flags & SVGF_FLAG_COMPENSATED_MOMENTS)) return SVGF_ERR_UNSUPPORTED;
        if (p.history_cap > WIDE_MAX_LEN) return SVGF_ERR_INVALID_ARGUMENT;
    }
    return SVGF_OK;
}

KParams to_kparams(const svgf_params& p)
{
    KParams k{};
    k.phi_luma = p.phi_luma;
    k.phi_normal = p.phi_normal;
    k.phi_depth = p.phi_depth;
    k.alpha_color = p.alpha_color;
    k.alpha_moments = p.alpha_moments;
    k.variance_eps = p.variance_eps;
    k.depth_eps = p.depth_eps;
    k.depth_reject = p.depth_reject;
    k.normal_reject = p.normal_reject;
    k.history_cap = (int32_t)p.history_cap;
    k.iterations = (int32_t)p.iterations;
    k.compensated = (p.flags & SVGF_FLAG_COMPENSATED_MOMENTS) ? 1 : 0;
    k.wide = (p.flags & SVGF_FLAG_WIDE_STATE) ? 1 : 0;
    constexpr float LOG2E = 1.44269504f, LN2 = 0.693147181f;
    k.luma_k = p.phi_luma > 0.0f ? -LOG2E / p.phi_luma : 0.0f;
    k.depth_c = LN2 * p.phi_depth;
    k.depth_e = p.phi_depth > 0.0f ? -LN2 * p.depth_eps : -INFINITY;
    return k;
}

svgf_status check_plane(const svgf_plane& pl, uint32_t want_fmt, uint32_t w)
{
    if (!pl.ptr) return SVGF_ERR_INVALID_ARGUMENT;
    if (pl.format != want_fmt) return SVGF_ERR_INVALID_ARGUMENT;
    uint32_t b = bytes_of(want_fmt);
    if ((reinterpret_cast<uintptr_t>(pl.ptr) % b) != 0) return SVGF_ERR_INVALID_ARGUMENT;
    if (pl.pitch_bytes < (uint64_t)w * b || (pl.pitch_bytes % b) != 0) return SVGF_ERR_INVALID_ARGUMENT;
    return SVGF_OK;
}

PlaneDev dev_plane(const svgf_plane& p) { return PlaneDev{static_cast<char*>(p.ptr), p.pitch_bytes}; }

uint32_t fmt4(Storage st) { return st == Storage::F32 ? SVGF_FMT_RGBA32F : SVGF_FMT_RGBA16F; }
// moments: two channels at the storage width, or the wide state's four binary16 words
uint32_t fmt2(Storage st, const KParams& kp)
{
    if (kp.wide) return SVGF_FMT_RGBA16F;
    return st == Storage::F32 ? SVGF_FMT_RG32F : SVGF_FMT_RG16F;
}
uint32_t fmt4(const svgf_context* c) { return fmt4(c->st); }
uint32_t fmt2(const svgf_context* c) { return fmt2(c->st, c->kp); }
// bytes of a moments pixel
uint32_t moment_bytes(uint32_t storage, uint32_t flags)
{
    if (flags & SVGF_FLAG_WIDE_STATE) return 8;
    return storage == SVGF_STORAGE_FP32 ? 8 : 4;
}

uint64_t stream_state_bytes(uint32_t mode, uint32_t s4, uint32_t s2, uint32_t w, uint32_t h)
{
    uint64_t px = (uint64_t)w * h;
    uint64_t b = 2 * align_up(px * s2, ALIGN);
    if (mode == SVGF_MODE_CANONICAL) b += 2 * align_up(px * s4, ALIGN) + 2 * align_up(px * 16, ALIGN);
    return b;
}

// Scratch of one frame in a batch or of the single-frame form: the wavelet ping-pong
// (storage x4), the uint8 history length and the float2 depth gradient.
constexpr uint64_t frame_scratch_bytes(uint32_t s4, uint64_t px)
{
    return 2 * align_up(px * s4, ALIGN) + align_up(px, ALIGN) + align_up(px * 8, ALIGN);
}

// wavelet ping-pong, history length (u8) and depth gradient (float2) of one frame
void carve_scratch(char* base, uint32_t s4, uint32_t w, uint32_t h, FrameDesc& f)
{
    uint64_t px = (uint64_t)w * h;
    size_t p4 = align_up(px * s4, ALIGN), p1 = align_up(px, ALIGN);
    f.wa = PlaneDev{base, (uint64_t)w * s4};
    f.wb = PlaneDev{base + p4, (uint64_t)w * s4};
    f.nlen = PlaneDev{base + 2 * p4, (uint64_t)w};
    f.grad = PlaneDev{base + 2 * p4 + p1, (uint64_t)w * 8};
}

// Kernels address planes with 32-bit offsets through buffer resources.
bool plane_fits(const svgf_plane& pl, uint32_t h) { return pl.pitch_bytes * (uint64_t)h < (1ull << 31); }

void free_slot(BatchSlot& s)
{
    for (auto& e : s.sev)
        if (e) (void)hipEventDestroy(e);
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.staging) (void)hipHostFree(s.staging);
    if (s.err_host) (void)hipHostFree(s.err_host);
    if (s.ctl) (void)hipFree(s.ctl);
    if (s.table) (void)hipFree(s.table);
    if (s.arena) (void)hipFree(s.arena);
    s = BatchSlot{};
}

void drop_graphs(svgf_context* c)
{
    for (auto& s : c->slots) {
        if (s.pending) {
            (void)hipEventSynchronize(s.done);
            s.pending = false;
        }
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
        if (s.graph) (void)hipGraphDestroy(s.graph);
        s.exec = nullptr;
        s.graph = nullptr;
    }
}

// Reads the outcome of a slot's batch once it has completed (the caller has seen its event):
// when a dependency wait of its persistent launch gave up, its outputs and the state it wrote
// are wrong, so its streams restart with cold history and never reproject that state.
void check_slot(svgf_context* c, BatchSlot& s)
{
    if (s.pending || s.checked || !s.seq) return;
    s.checked = true;
    auto record = [&] { c->outcomes[s.seq % BATCH_OUTCOMES] = {s.seq, s.failed}; };
    if (s.sev_n > 0) {
        // stage times of a batch recorded with SVGF_BATCH_TIMED_STAGES
        float t[MAX_STAGES];
        bool ok = true;
        for (int i = 0; i < s.sev_n && ok; ++i) ok = hipEventElapsedTime(&t[i], s.sev[i], s.sev[i + 1]) == hipSuccess;
        if (ok) {
            if (c->stage_n != s.sev_n) {
                std::fill(std::begin(c->stage_sum), std::end(c->stage_sum), 0.0);
                c->stage_batches = 0;
                c->stage_n = s.sev_n;
            }
            for (int i = 0; i < s.sev_n; ++i) c->stage_sum[i] += t[i];
            c->stage_batches++;
        }
    }
    if (!s.err_host || *s.err_host == 0) return record();
    *s.err_host = 0;
    s.failed = true;
    for (uint32_t id : s.ids)
        if (id < c->streams.size() && c->streams[id].open) c->streams[id].frames = 0;
    record();
}

// Marks completed slots (non-blocking) and reads their outcome.
svgf_status poll_slots(svgf_context* c)
{
    for (auto& s : c->slots) {
        if (s.pending) {
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return SVGF_ERR_HIP;
            s.pending = false;
        }
        check_slot(c, s);
    }
    return SVGF_OK;
}

void destroy_context(svgf_context* c)
{
    (void)hipDeviceSynchronize();
    for (auto& s : c->slots) free_slot(s);
    for (auto& s : c->streams)
        if (s.block) (void)hipFree(s.block);
    for (auto& kv : c->pool) (void)hipFree(kv.second);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->scratch_ev) (void)hipEventDestroy(c->scratch_ev);
    if (c->capture_stream) (void)hipStreamDestroy(c->capture_stream);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    delete c;
}

int stage_count(const svgf_context* c) { return c->kp.iterations + 1; }

// Resident workgroups of the persistent kernel variant the parameters select.
void size_persistent(svgf_context* c)
{
    const int per_cu = persistent_blocks_per_cu(c->mode == SVGF_MODE_CANONICAL, c->st, state_form(c->kp));
    c->persist_blocks = (uint32_t)std::max(1, per_cu * c->cus);
}

// One persistent launch per batch (stage profiling needs one launch per stage).
bool persistent(const svgf_context* c)
{
    return (c->desc.batch_flags & SVGF_BATCH_PERSISTENT) && !(c->desc.batch_flags & SVGF_PROFILE_STAGES);
}

// Kernel launches per batch and the grid of launch i for a batch whose stages have
// `tiles[i]` tiles: one block per tile, or for the persistent kernel at most as many
// workgroups as the device holds at once (more would only queue behind them).
int launches_per_batch(const svgf_context* c) { return persistent(c) ? 1 : stage_count(c); }

// Tiles per ticket-counter increment of the persistent kernel: fewer same-address atomics and
// neighbouring tiles on one workgroup (one L2), against a longer tail at each stage boundary.
constexpr uint32_t PERSIST_CHUNK = 2;

uint32_t launch_grid(const svgf_context* c, const uint32_t* tiles, int i)
{
    if (!persistent(c)) return tiles[i];
    uint64_t total = 0;
    for (int st = 0; st < stage_count(c); ++st) total += tiles[st];
    const uint64_t chunks = (total + PERSIST_CHUNK - 1) / PERSIST_CHUNK;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, c->persist_blocks));
}

hipEvent_t* profile_events(svgf_context* c)
{
    if (!(c->desc.batch_flags & SVGF_PROFILE_STAGES)) return nullptr;
    c->ev_stages = stage_count(c);
    return c->ev;
}

bool timed_stages(const svgf_context* c)
{
    return (c->desc.batch_flags & SVGF_BATCH_TIMED_STAGES) && !(c->desc.batch_flags & SVGF_PROFILE_STAGES);
}

// Events around the stage launches of a batch in slot s launched directly: the context's
// (SVGF_PROFILE_STAGES) or the slot's own (SVGF_BATCH_TIMED_STAGES). A captured batch gets the
// slot's events as graph nodes instead (add_stage_events): stream capture does not turn
// hipEventRecord into a graph node.
hipEvent_t* batch_events(svgf_context* c, BatchSlot& s, bool capturing)
{
    if (hipEvent_t* ev = profile_events(c)) return ev;
    if (!timed_stages(c) || capturing) return nullptr;
    s.sev_n = launches_per_batch(c);
    return s.sev;
}

// Inserts an event-record node before every kernel node of a captured chain (table copy ->
// kernel 0 -> kernel 1 -> ...) and one after the last: ev[i] is recorded before kernel i and
// ev[n] after kernel n - 1, as direct launches record them. Returns the number of kernels.
hipError_t add_stage_events(hipGraph_t g, hipEvent_t* ev, int max_kernels, int& kernels)
{
    std::vector<hipGraphNode_t> chain;
    hipGraphNode_t node = nullptr;
    size_t n = 1;
    hipError_t e = hipGraphGetRootNodes(g, &node, &n);
    while (e == hipSuccess && n == 1 && node) {
        chain.push_back(node);
        size_t nd = 0;
        e = hipGraphNodeGetDependentNodes(node, nullptr, &nd);
        if (e != hipSuccess || nd == 0) break;
        if (nd != 1) return hipErrorUnknown;
        hipGraphNode_t next = nullptr;
        e = hipGraphNodeGetDependentNodes(node, &next, &nd);
        node = next;
    }
    kernels = 0;
    hipGraphNode_t last = nullptr;
    for (size_t i = 0; i < chain.size() && e == hipSuccess; ++i) {
        hipGraphNodeType t;
        if ((e = hipGraphNodeGetType(chain[i], &t)) != hipSuccess) break;
        if (t != hipGraphNodeTypeKernel) continue;
        if (kernels >= max_kernels) return hipErrorUnknown;
        hipGraphNode_t en = nullptr;
        if (i > 0) {
            if ((e = hipGraphRemoveDependencies(g, &chain[i - 1], &chain[i], 1)) != hipSuccess) break;
            e = hipGraphAddEventRecordNode(&en, g, &chain[i - 1], 1, ev[kernels]);
        } else {
            e = hipGraphAddEventRecordNode(&en, g, nullptr, 0, ev[kernels]);
        }
        if (e == hipSuccess) e = hipGraphAddDependencies(g, &en, &chain[i], 1);
        ++kernels;
        last = chain[i];
    }
    if (e == hipSuccess && last) {
        hipGraphNode_t en = nullptr;
        e = hipGraphAddEventRecordNode(&en, g, &last, 1, ev[kernels]);
    }
    return e;
}

hipError_t enqueue_batch(svgf_context* c, BatchSlot& s, const uint32_t* tiles, hipStream_t stream, bool capturing = false)
{
    hipError_t e = hipMemcpyAsync(s.table, s.staging, c->table_bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    char* t = static_cast<char*>(s.table);
    BatchTable bt{reinterpret_cast<const uint32_t*>(t), reinterpret_cast<const uint32_t*>(t + c->starts_off),
                  reinterpret_cast<const FrameDesc*>(t + c->frames_off)};
    hipEvent_t* ev = batch_events(c, s, capturing);
    if (persistent(c)) {
        if (ev && (e = hipEventRecord(ev[0], stream)) != hipSuccess) return e;
        e = launch_batch_persistent(c->mode == SVGF_MODE_CANONICAL, bt, c->kp, c->st, s.ctl, s.err_dev, c->max_polls,
                                    launch_grid(c, tiles, 0), stream);
        if (e == hipSuccess && ev) e = hipEventRecord(ev[1], stream);
        return e;
    }
    const uint32_t* grids = tiles;
    if (c->mode == SVGF_MODE_REFERENCE) return launch_reference_batch(bt, c->kp, c->st, grids, stream, ev);
    return launch_canonical_batch(bt, c->kp, c->st, grids, stream, ev);
}

// Captures table copy + kernel launches of a slot once; later batches only change the grid
// size of each kernel node, so the graph is never re-instantiated.
hipError_t capture_slot(svgf_context* c, BatchSlot& s, const uint32_t* tiles)
{
    hipError_t e = hipStreamBeginCapture(c->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = enqueue_batch(c, s, tiles, c->capture_stream, true);
    e = hipStreamEndCapture(c->capture_stream, &s.graph);
    if (le != hipSuccess) e = le;
    if (e == hipSuccess && timed_stages(c)) {
        int kernels = 0;
        e = add_stage_events(s.graph, s.sev, MAX_STAGES, kernels);
        if (e == hipSuccess && kernels != launches_per_batch(c)) e = hipErrorUnknown;
        s.sev_n = e == hipSuccess ? kernels : 0;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0);
    // kernel nodes in launch order: follow the dependency chain from the root (the copy)
    s.knodes.clear();
    s.kparams.clear();
    hipGraphNode_t node = nullptr;
    size_t n = 1;
    if (e == hipSuccess) e = hipGraphGetRootNodes(s.graph, &node, &n);
    while (e == hipSuccess && n == 1 && node) {
        hipGraphNodeType type;
        e = hipGraphNodeGetType(node, &type);
        if (e == hipSuccess && type == hipGraphNodeTypeKernel) {
            hipKernelNodeParams kp{};
            e = hipGraphKernelNodeGetParams(node, &kp);
            s.knodes.push_back(node);
            s.kparams.push_back(kp);
        }
        hipGraphNode_t next = nullptr;
        size_t nd = 0;
        if (e == hipSuccess) e = hipGraphNodeGetDependentNodes(node, nullptr, &nd);
        if (e != hipSuccess || nd == 0) break;
        if (nd != 1) {
            e = hipErrorUnknown;
            break;
        }
        e = hipGraphNodeGetDependentNodes(node, &next, &nd);
        node = next;
    }
    if (e == hipSuccess && (int)s.knodes.size() != launches_per_batch(c)) e = hipErrorUnknown;
    if (e != hipSuccess) {
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
        if (s.graph) (void)hipGraphDestroy(s.graph);
        s.exec = nullptr;
        s.graph = nullptr;
        s.knodes.clear();
        s.kparams.clear();
    }
    return e;
}

// Checks the planes of one single-frame call and fills its frame descriptor (the scratch
// planes are the caller's to carve).
svgf_status single_frame_desc(uint32_t mode, Storage st, const KParams& kp, const svgf_frame* fr, FrameDesc& f)
{
    const uint32_t w = fr->width, h = fr->height;
    if (w < 1 || h < 1) return SVGF_ERR_INVALID_ARGUMENT;
    svgf_status s;
    const bool hv = fr->history_valid != 0;
    if ((s = check_plane(fr->radiance, fmt4(st), w)) != SVGF_OK) return s;
    if ((s = check_plane(fr->moments_curr, fmt2(st, kp), w)) != SVGF_OK) return s;
    if ((s = check_plane(fr->denoised, fmt4(st), w)) != SVGF_OK) return s;
    if (hv && (s = check_plane(fr->moments_prev, fmt2(st, kp), w)) != SVGF_OK) return s;
    if (mode == SVGF_MODE_CANONICAL) {
        if ((s = check_plane(fr->gbuffer, SVGF_FMT_RGBA32F, w)) != SVGF_OK) return s;
        if ((s = check_plane(fr->history_curr, fmt4(st), w)) != SVGF_OK) return s;
        if (hv) {
            if ((s = check_plane(fr->gbuffer_prev, SVGF_FMT_RGBA32F, w)) != SVGF_OK) return s;
            if ((s = check_plane(fr->motion, SVGF_FMT_RG32F, w)) != SVGF_OK) return s;
            if ((s = check_plane(fr->history_prev, fmt4(st), w)) != SVGF_OK) return s;
        }
    }
    for (const svgf_plane* pl : {&fr->radiance, &fr->gbuffer, &fr->gbuffer_prev, &fr->motion, &fr->history_prev,
                                 &fr->history_curr, &fr->moments_prev, &fr->moments_curr, &fr->denoised})
        if (pl->ptr && !plane_fits(*pl, h)) return SVGF_ERR_CAPACITY;
    f = FrameDesc{};
    f.w = (int32_t)w;
    f.h = (int32_t)h;
    f.history_valid = hv ? 1 : 0;
    f.rad = dev_plane(fr->radiance);
    f.mom_curr = dev_plane(fr->moments_curr);
    f.den = dev_plane(fr->denoised);
    if (hv) f.mom_prev = dev_plane(fr->moments_prev);
    if (mode == SVGF_MODE_CANONICAL) {
        f.g = dev_plane(fr->gbuffer);
        f.hist_curr = dev_plane(fr->history_curr);
        if (hv) {
            f.gprev = dev_plane(fr->gbuffer_prev);
            f.mv = dev_plane(fr->motion);
            f.hist_prev = dev_plane(fr->history_prev);
        }
    }
    return SVGF_OK;
}

hipError_t launch_single_frame(uint32_t mode, Storage st, const KParams& kp, const FrameDesc& f, hipStream_t stream,
                               hipEvent_t* ev)
{
    return mode == SVGF_MODE_REFERENCE ? launch_reference_frame(f, kp, st, stream, ev)
                                       : launch_canonical_frame(f, kp, st, stream, ev);
}

}  // namespace

// Scratch of svgf_denoise_static: the single-frame scratch (wavelet ping-pong at FP32 width,
// history length, depth gradient) of a frame of up to SVGF_STATIC_MAX_PIXELS pixels, in a
// fixed device array of the library, so that the call allocates nothing. Each GPU the library
// runs on holds its own copy at its own address until the process ends. Host code never uses
// the array's name as an address: the per-device address comes from hipGetSymbolAddress.
constexpr uint64_t STATIC_SCRATCH_BYTES = frame_scratch_bytes(16, SVGF_STATIC_MAX_PIXELS);
static __device__ __align__(256) char scratch[STATIC_SCRATCH_BYTES];

namespace {

constexpr int STATIC_MAX_DEVICES = 64;
struct StaticDevice {
    void* ptr = nullptr;  // address of `scratch` on the device
    int state = 0;        // 0: not set up, 1: ready, < 0: svgf_status of the failed set-up
};
std::mutex g_static_mu;
StaticDevice g_static[STATIC_MAX_DEVICES];

// Set-up of the current device, once: gfx950 check and the array's device address.
svgf_status static_device(int dev, void** ptr)
{
    if (dev < 0 || dev >= STATIC_MAX_DEVICES) return SVGF_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_static_mu);
    StaticDevice& d = g_static[dev];
    if (d.state == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            d.state = -(int)SVGF_ERR_NO_DEVICE;
        } else if (hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch)) != hipSuccess || !d.ptr) {
            d.ptr = nullptr;
            d.state = -(int)SVGF_ERR_HIP;
        } else {
            d.state = 1;
        }
    }
    if (d.state < 0) return (svgf_status)(-d.state);
    *ptr = d.ptr;
    return SVGF_OK;
}

}  // namespace

extern "C" {

uint32_t svgf_abi_version(void) { return SVGF_ABI_VERSION; }

const char* svgf_status_str(svgf_status s)
{
    switch (s) {
    case SVGF_OK: return "ok";
    case SVGF_ERR_INVALID_ARGUMENT: return "invalid argument";
    case SVGF_ERR_UNSUPPORTED: return "unsupported";
    case SVGF_ERR_OUT_OF_MEMORY: return "out of device memory";
    case SVGF_ERR_OVER_BUDGET: return "over HBM budget";
    case SVGF_ERR_NO_STREAM: return "no such stream";
    case SVGF_ERR_CAPACITY: return "over context capacity";
    case SVGF_ERR_HIP: return "HIP runtime error";
    case SVGF_ERR_NO_DEVICE: return "no gfx950 device";
    case SVGF_ERR_TIMEOUT: return "timeout";
    case SVGF_ERR_NOT_READY: return "not ready";
    }
    return "unknown status";
}

svgf_status svgf_default_params(uint32_t mode, uint32_t storage, svgf_params* out)
{
    if (!out || mode > SVGF_MODE_CANONICAL || storage > SVGF_STORAGE_FP16) return SVGF_ERR_INVALID_ARGUMENT;
    svgf_params p{};
    if (mode == SVGF_MODE_REFERENCE) {
        // svgf_denoise.slang:118-150 and main.cpp:929
        p.iterations = 3;
        p.phi_luma = 2.0f;
        p.alpha_moments = 0.2f;
        p.variance_eps = 1e-4f;
        p.alpha_color = 1.0f;
        p.history_cap = 1;
        p.depth_eps = 1e-3f;
    } else {
        p.iterations = 5;
        p.phi_luma = 4.0f;
        p.phi_normal = 128.0f;
        p.phi_depth = 1.0f;
        p.alpha_color = 0.2f;
        p.alpha_moments = 0.2f;
        p.variance_eps = 1e-4f;
        p.depth_eps = 1e-3f;
        p.depth_reject = 0.1f;
        p.normal_reject = 0.9f;
        p.history_cap = 32;
        // FP16: compensated moments, and the wide state so that binary16 roundings of the
        // fed-back history and moments do not compound over frames
        p.flags = storage == SVGF_STORAGE_FP16 ? SVGF_FLAG_COMPENSATED_MOMENTS | SVGF_FLAG_WIDE_STATE : 0u;
    }
    *out = p;
    return SVGF_OK;
}

svgf_status svgf_query_bytes(const svgf_desc* d, uint32_t w, uint32_t h, uint64_t* context_bytes, uint64_t* stream_bytes)
{
    if (!d || d->mode > SVGF_MODE_CANONICAL || d->storage > SVGF_STORAGE_FP16) return SVGF_ERR_INVALID_ARGUMENT;
    uint32_t s4 = d->storage == SVGF_STORAGE_FP32 ? 16 : 8, s2 = moment_bytes(d->storage, d->params.flags);
    const uint32_t w4 = d->mode == SVGF_MODE_CANONICAL ? 16 : s4;
    if (context_bytes) {
        uint64_t b = frame_scratch_bytes(w4, (uint64_t)d->max_width * d->max_height);
        if (d->max_batch_frames) {
            uint64_t table = align_up(64 + 4ull * MAX_STAGES * STARTS_ROW, ALIGN) +
                             sizeof(FrameDesc) * (uint64_t)d->max_batch_frames;
            uint64_t arena = d->max_batch_pixels * (2ull * w4 + 9) + (uint64_t)d->max_batch_frames * 4 * ALIGN;
            b += BATCH_SLOTS * (table + arena);
        }
        *context_bytes = b;
    }
    if (stream_bytes) *stream_bytes = stream_state_bytes(d->mode, s4, s2, w, h);
    return SVGF_OK;
}

svgf_status svgf_create(const svgf_desc* d, svgf_context** out)
{
    if (!d || !out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (d->mode > SVGF_MODE_CANONICAL || d->storage > SVGF_STORAGE_FP16) return SVGF_ERR_INVALID_ARGUMENT;
    if (d->max_width < 1 || d->max_height < 1 || d->max_width > 16384 || d->max_height > 16384)
        return SVGF_ERR_INVALID_ARGUMENT;
    // every plane, the float4 G-buffer included, must stay addressable with 32-bit offsets
    if ((uint64_t)d->max_width * d->max_height * 16 >= (1ull << 31)) return SVGF_ERR_CAPACITY;
    if (d->max_batch_frames > MAX_BATCH_FRAMES) return SVGF_ERR_INVALID_ARGUMENT;
    if (d->max_batch_frames && d->max_batch_pixels < 1) return SVGF_ERR_INVALID_ARGUMENT;
    if (d->batch_flags & ~(SVGF_BATCH_NO_GRAPH | SVGF_PROFILE_STAGES | SVGF_BATCH_PERSISTENT | SVGF_BATCH_TIMED_STAGES))
        return SVGF_ERR_INVALID_ARGUMENT;
    svgf_status ps = validate_params(d->params, d->mode, d->storage);
    if (ps != SVGF_OK) return ps;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || d->device < 0 || d->device >= ndev) return SVGF_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d->device) != hipSuccess) return SVGF_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SVGF_ERR_NO_DEVICE;
    DeviceGuard guard(d->device);

    svgf_context* c = new (std::nothrow) svgf_context();
    if (!c) return SVGF_ERR_OUT_OF_MEMORY;
    c->desc = *d;
    c->mode = (int)d->mode;
    c->st = d->storage == SVGF_STORAGE_FP32 ? Storage::F32 : Storage::F16;
    c->s4 = c->st == Storage::F32 ? 16 : 8;
    c->s2 = moment_bytes(d->storage, d->params.flags);
    c->w4 = c->mode == SVGF_MODE_CANONICAL ? 16 : c->s4;
    c->kp = to_kparams(d->params);
    c->grid_blocks = std::max(1, prop.multiProcessorCount) * 8;
    c->cus = std::max(1, prop.multiProcessorCount);
    {
        size_persistent(c);
        // dependency-wait bound of the persistent kernel, in polls of about 1.3 us (a 1024-cycle
        // sleep plus one load): 250 ms
        uint64_t us = 250000;
        // diagnostic override in microseconds (0 makes every wait that finds its dependency
        // unfinished give up at once: the error path without a stuck kernel)
        if (const char* env = std::getenv("SVGF_PERSISTENT_WAIT_US")) us = std::strtoull(env, nullptr, 10);
        c->max_polls = us * 10 / 13;
    }

    uint64_t sb = frame_scratch_bytes(c->w4,(uint64_t)d->max_width * d->max_height);
    if (hipMalloc(&c->scratch, sb) != hipSuccess) {
        c->scratch = nullptr;
        destroy_context(c);
        return SVGF_ERR_OUT_OF_MEMORY;
    }
    if (hipEventCreateWithFlags(&c->scratch_ev, hipEventDisableTiming) != hipSuccess) {
        c->scratch_ev = nullptr;
        destroy_context(c);
        return SVGF_ERR_HIP;
    }
    if (d->batch_flags & SVGF_PROFILE_STAGES) {
        for (auto& e : c->ev)
            if (hipEventCreate(&e) != hipSuccess) {
                e = nullptr;
                destroy_context(c);
                return SVGF_ERR_HIP;
            }
    }

    if (d->max_streams) c->streams.resize(d->max_streams);
    if (d->max_batch_frames) {
        c->starts_off = 64;
        c->frames_off = align_up(c->starts_off + 4ull * MAX_STAGES * STARTS_ROW, ALIGN);
        c->table_bytes = c->frames_off + sizeof(FrameDesc) * (size_t)d->max_batch_frames;
        c->arena_bytes = d->max_batch_pixels * (2ull * c->w4 + 9) + (uint64_t)d->max_batch_frames * 4 * ALIGN;
        for (auto& s : c->slots) {
            if (hipHostMalloc(&s.staging, c->table_bytes, hipHostMallocDefault) != hipSuccess ||
                hipMalloc(&s.table, c->table_bytes) != hipSuccess || hipMalloc(&s.arena, c->arena_bytes) != hipSuccess ||
                hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
                destroy_context(c);
                return SVGF_ERR_OUT_OF_MEMORY;
            }
            std::memset(s.staging, 0, c->table_bytes);
            // persistent form: counters start (and are left by every launch) at zero; the
            // timeout flag lives in host-mapped memory so the host reads it without a copy
            if (hipMalloc(reinterpret_cast<void**>(&s.ctl), sizeof(PersistCtl)) != hipSuccess ||
                hipMemset(s.ctl, 0, sizeof(PersistCtl)) != hipSuccess ||
                hipHostMalloc(reinterpret_cast<void**>(&s.err_host), sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
                hipHostGetDevicePointer(reinterpret_cast<void**>(&s.err_dev), s.err_host, 0) != hipSuccess) {
                destroy_context(c);
                return SVGF_ERR_OUT_OF_MEMORY;
            }
            *s.err_host = 0;
            if (d->batch_flags & SVGF_BATCH_TIMED_STAGES)
                for (auto& e : s.sev)
                    if (hipEventCreate(&e) != hipSuccess) {
                        e = nullptr;
                        destroy_context(c);
                        return SVGF_ERR_HIP;
                    }
        }
        if (hipDeviceSynchronize() != hipSuccess) {
            destroy_context(c);
            return SVGF_ERR_HIP;
        }
        if (hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking) != hipSuccess) {
            destroy_context(c);
            return SVGF_ERR_HIP;
        }
    }
    *out = c;
    return SVGF_OK;
}

svgf_status svgf_destroy(svgf_context* c)
{
    if (!c) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->desc.device);
    destroy_context(c);
    return SVGF_OK;
}

svgf_status svgf_set_params(svgf_context* c, const svgf_params* p)
{
    if (!c || !p) return SVGF_ERR_INVALID_ARGUMENT;
    svgf_status s = validate_params(*p, c->desc.mode, c->desc.storage);
    if (s != SVGF_OK) return s;
    // the flags select how resident streams store their moments and history: open streams
    // would read their state in the other representation
    if (p->flags != c->desc.params.flags &&
        std::any_of(c->streams.begin(), c->streams.end(), [](const StreamSlot& t) { return t.open; }))
        return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->desc.device);
    drop_graphs(c);  // captured kernels hold the old parameters by value
    c->desc.params = *p;
    c->kp = to_kparams(*p);
    c->s2 = moment_bytes(c->desc.storage, p->flags);
    size_persistent(c);
    return SVGF_OK;
}

svgf_status svgf_denoise(svgf_context* c, const svgf_frame* fr, void* hip_stream)
{
    if (!c || !fr) return SVGF_ERR_INVALID_ARGUMENT;
    const uint32_t w = fr->width, h = fr->height;
    if (w < 1 || h < 1) return SVGF_ERR_INVALID_ARGUMENT;
    if (w > c->desc.max_width || h > c->desc.max_height) return SVGF_ERR_CAPACITY;
    FrameDesc f;
    svgf_status s = single_frame_desc(c->desc.mode, c->st, c->kp, fr, f);
    if (s != SVGF_OK) return s;
    DeviceGuard guard(c->desc.device);
    // Scratch planes are packed at this frame's width.
    carve_scratch(static_cast<char*>(c->scratch), c->w4, w, h, f);
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    // Every call uses the context's one scratch: a call on another stream than the previous
    // call's first waits (on the device) for that call, so calls never overlap through it.
    if (c->scratch_used && st != c->scratch_stream && hipStreamWaitEvent(st, c->scratch_ev, 0) != hipSuccess)
        return SVGF_ERR_HIP;
    hipError_t e = launch_single_frame(c->desc.mode, c->st, c->kp, f, st, profile_events(c));
    if (e == hipSuccess) e = hipEventRecord(c->scratch_ev, st);
    if (e != hipSuccess) return SVGF_ERR_HIP;
    c->scratch_used = true;
    c->scratch_stream = st;
    return SVGF_OK;
}

svgf_status svgf_denoise_static(uint32_t mode, uint32_t storage, const svgf_params* params, const svgf_frame* fr,
                                void* hip_stream)
{
    if (!params || !fr || mode > SVGF_MODE_CANONICAL || storage > SVGF_STORAGE_FP16) return SVGF_ERR_INVALID_ARGUMENT;
    svgf_status s = validate_params(*params, mode, storage);
    if (s != SVGF_OK) return s;
    const uint32_t w = fr->width, h = fr->height;
    if (w < 1 || h < 1) return SVGF_ERR_INVALID_ARGUMENT;
    if ((uint64_t)w * h > SVGF_STATIC_MAX_PIXELS) return SVGF_ERR_CAPACITY;
    const Storage st = storage == SVGF_STORAGE_FP32 ? Storage::F32 : Storage::F16;
    const KParams kp = to_kparams(*params);
    FrameDesc f;
    if ((s = single_frame_desc(mode, st, kp, fr, f)) != SVGF_OK) return s;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return SVGF_ERR_NO_DEVICE;
    void* base = nullptr;
    if ((s = static_device(dev, &base)) != SVGF_OK) return s;
    // every plane of the scratch is written (temporal / init stage) before it is read, so
    // nothing needs clearing; the scratch planes are packed at this frame's width
    carve_scratch(static_cast<char*>(base), mode == SVGF_MODE_CANONICAL ? 16u : (st == Storage::F32 ? 16u : 8u), w, h, f);
    const hipError_t e = launch_single_frame(mode, st, kp, f, static_cast<hipStream_t>(hip_stream), nullptr);
    return e == hipSuccess ? SVGF_OK : SVGF_ERR_HIP;
}

svgf_status svgf_stream_open(svgf_context* c, uint32_t w, uint32_t h, uint32_t* id)
{
    if (!c || !id || w < 1 || h < 1) return SVGF_ERR_INVALID_ARGUMENT;
    if (w > c->desc.max_width || h > c->desc.max_height) return SVGF_ERR_CAPACITY;
    auto it = std::find_if(c->streams.begin(), c->streams.end(), [](const StreamSlot& s) { return !s.open; });
    if (it == c->streams.end()) return SVGF_ERR_CAPACITY;
    uint64_t bytes = stream_state_bytes(c->desc.mode, c->s4, c->s2, w, h);
    const uint64_t budget = c->desc.hbm_budget_bytes;
    if (budget && c->budget_used + bytes > budget) return SVGF_ERR_OVER_BUDGET;
    DeviceGuard guard(c->desc.device);
    void* block = nullptr;
    auto it_pool = c->pool.lower_bound(bytes);
    if (it_pool != c->pool.end() && it_pool->first <= 2 * bytes &&
        (!budget || c->budget_used + it_pool->first <= budget)) {
        block = it_pool->second;
        bytes = it_pool->first;
        c->pool_bytes -= it_pool->first;
        c->pool.erase(it_pool);
    } else {
        // keep open streams + cached blocks within the budget
        while (budget && !c->pool.empty() && c->budget_used + c->pool_bytes + bytes > budget) {
            auto big = std::prev(c->pool.end());
            (void)hipFree(big->second);
            c->pool_bytes -= big->first;
            c->pool.erase(big);
        }
        if (hipMalloc(&block, bytes) != hipSuccess) return SVGF_ERR_OUT_OF_MEMORY;
    }
    StreamSlot& s = *it;
    s = StreamSlot{};
    s.open = true;
    s.w = w;
    s.h = h;
    s.block = block;
    s.bytes = bytes;
    uint64_t px = (uint64_t)w * h;
    char* b = static_cast<char*>(block);
    for (int i = 0; i < 2; ++i) {
        s.mom[i] = PlaneDev{b, (uint64_t)w * c->s2};
        b += align_up(px * c->s2, ALIGN);
    }
    if (c->desc.mode == SVGF_MODE_CANONICAL) {
        for (int i = 0; i < 2; ++i) {
            s.hist[i] = PlaneDev{b, (uint64_t)w * c->s4};
            b += align_up(px * c->s4, ALIGN);
        }
        for (int i = 0; i < 2; ++i) {
            s.gbuf[i] = PlaneDev{b, (uint64_t)w * 16};
            b += align_up(px * 16, ALIGN);
        }
    }
    c->budget_used += bytes;
    *id = (uint32_t)(it - c->streams.begin());
    return SVGF_OK;
}

static StreamSlot* find_stream(svgf_context* c, uint32_t id)
{
    if (!c || id >= c->streams.size() || !c->streams[id].open) return nullptr;
    return &c->streams[id];
}

svgf_status svgf_stream_close(svgf_context* c, uint32_t id)
{
    StreamSlot* s = find_stream(c, id);
    if (!s) return c ? SVGF_ERR_NO_STREAM : SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->desc.device);
    // Batches in flight that touch this stream may still read its state.
    for (auto& sl : c->slots)
        if (sl.pending && std::find(sl.ids.begin(), sl.ids.end(), id) != sl.ids.end()) {
            (void)hipEventSynchronize(sl.done);
            sl.pending = false;
            check_slot(c, sl);  // before the id can name another stream
        }
    c->pool.emplace(s->bytes, s->block);
    c->pool_bytes += s->bytes;
    c->budget_used -= s->bytes;
    *s = StreamSlot{};
    return SVGF_OK;
}

svgf_status svgf_stream_reset(svgf_context* c, uint32_t id, void* hip_stream)
{
    (void)hip_stream;
    StreamSlot* s = find_stream(c, id);
    if (!s) return c ? SVGF_ERR_NO_STREAM : SVGF_ERR_INVALID_ARGUMENT;
    // Cold history: the next frame runs with history_valid = 0 and never reads the old state
    // (the zero-cleared images of main.cpp:1449-1455 / 1570-1603).
    s->frames = 0;
    s->g_kept = false;
    return SVGF_OK;
}

svgf_status svgf_stream_state_get(svgf_context* c, uint32_t id, svgf_stream_state* out)
{
    StreamSlot* s = find_stream(c, id);
    if (!s) return c ? SVGF_ERR_NO_STREAM : SVGF_ERR_INVALID_ARGUMENT;
    if (!out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = svgf_stream_state{};
    out->width = s->w;
    out->height = s->h;
    out->frames = s->frames;
    int p = s->parity;
    out->moments = svgf_plane{s->mom[p].ptr, s->mom[p].pitch, fmt2(c), 0};
    if (c->desc.mode == SVGF_MODE_CANONICAL) {
        out->history = svgf_plane{s->hist[p].ptr, s->hist[p].pitch, fmt4(c), 0};
        const PlaneDev& g = s->g_kept ? s->gkept : s->gbuf[p];
        out->gbuffer = svgf_plane{g.ptr, g.pitch, SVGF_FMT_RGBA32F, 0};
    }
    return SVGF_OK;
}

svgf_status svgf_budget_used(svgf_context* c, uint64_t* bytes)
{
    if (!c || !bytes) return SVGF_ERR_INVALID_ARGUMENT;
    *bytes = c->budget_used;
    return SVGF_OK;
}

svgf_status svgf_denoise_batch(svgf_context* c, const svgf_request* req, uint32_t n, void* hip_stream)
{
    if (!c || (!req && n)) return SVGF_ERR_INVALID_ARGUMENT;
    if (n == 0) return SVGF_OK;
    if (!c->desc.max_batch_frames) return SVGF_ERR_UNSUPPORTED;
    if (n > c->desc.max_batch_frames) return SVGF_ERR_CAPACITY;
    uint64_t total_px = 0, arena_need = 0;
    for (uint32_t i = 0; i < n; ++i) {
        StreamSlot* s = find_stream(c, req[i].stream_id);
        if (!s) return SVGF_ERR_NO_STREAM;
        for (uint32_t j = 0; j < i; ++j)
            if (req[j].stream_id == req[i].stream_id) return SVGF_ERR_INVALID_ARGUMENT;
        if (req[i].flags & ~SVGF_REQUEST_KEEP_GBUFFER) return SVGF_ERR_INVALID_ARGUMENT;
        svgf_status st;
        if ((st = check_plane(req[i].radiance, fmt4(c), s->w)) != SVGF_OK) return st;
        if ((st = check_plane(req[i].denoised, fmt4(c), s->w)) != SVGF_OK) return st;
        if (c->mode == SVGF_MODE_CANONICAL) {
            if ((st = check_plane(req[i].gbuffer, SVGF_FMT_RGBA32F, s->w)) != SVGF_OK) return st;
            if (s->frames > 0 && (st = check_plane(req[i].motion, SVGF_FMT_RG32F, s->w)) != SVGF_OK) return st;
            if (!plane_fits(req[i].gbuffer, s->h) || (s->frames > 0 && !plane_fits(req[i].motion, s->h)))
                return SVGF_ERR_CAPACITY;
        }
        if (!plane_fits(req[i].radiance, s->h) || !plane_fits(req[i].denoised, s->h)) return SVGF_ERR_CAPACITY;
        total_px += (uint64_t)s->w * s->h;
        arena_need += frame_scratch_bytes(c->w4,(uint64_t)s->w * s->h);
    }
    if (total_px > c->desc.max_batch_pixels || arena_need > c->arena_bytes) return SVGF_ERR_CAPACITY;
    if (!persistent(c)) {
        // one block of BLOCK threads per tile of a stage, and a launch holds at most 2^32 - 1
        // work-items; the tiles of a stride-s stage grow with s^2 (a frame has at least s^2 / 2
        // tiles at s = 512), so many frames at 9 or 10 iterations can exceed that
        for (int st = 0; st < stage_count(c); ++st) {
            uint64_t t = 0;
            for (uint32_t i = 0; i < n; ++i) {
                const StreamSlot& s = c->streams[req[i].stream_id];
                t += stage_tiles(c->mode == SVGF_MODE_CANONICAL, st, s.w, s.h);
            }
            if (t > MAX_LAUNCH_BLOCKS) return SVGF_ERR_CAPACITY;
        }
    }

    DeviceGuard guard(c->desc.device);
    if (persistent(c)) {
        // A batch that reads the state a batch in flight writes waits for it, so that a failed
        // launch (its streams then restart cold) is found before its state is read.
        for (auto& sl : c->slots) {
            if (!sl.pending) continue;
            bool reads = false;
            for (uint32_t i = 0; i < n && !reads; ++i)
                reads = std::find(sl.ids.begin(), sl.ids.end(), req[i].stream_id) != sl.ids.end();
            if (!reads) continue;
            if (hipEventSynchronize(sl.done) != hipSuccess) return SVGF_ERR_HIP;
            sl.pending = false;
        }
    }
    if (poll_slots(c) != SVGF_OK) return SVGF_ERR_HIP;
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    // A batch in flight on another hip_stream that names one of these streams writes the state
    // this batch reads (and reads the parity this batch writes): this stream waits for it on
    // the device. Batches on one hip_stream are ordered by the stream itself.
    for (auto& sl : c->slots) {
        if (!sl.pending || sl.stream == stream) continue;
        bool shares = false;
        for (uint32_t i = 0; i < n && !shares; ++i)
            shares = std::find(sl.ids.begin(), sl.ids.end(), req[i].stream_id) != sl.ids.end();
        if (shares && hipStreamWaitEvent(stream, sl.done, 0) != hipSuccess) return SVGF_ERR_HIP;
    }
    BatchSlot& slot = c->slots[c->next_slot];
    c->next_slot = (c->next_slot + 1) % BATCH_SLOTS;
    if (slot.pending) {
        if (hipEventSynchronize(slot.done) != hipSuccess) return SVGF_ERR_HIP;
        slot.pending = false;
        check_slot(c, slot);
    }
    char* host = static_cast<char*>(slot.staging);
    uint32_t* hdr = reinterpret_cast<uint32_t*>(host);
    uint32_t* starts = reinterpret_cast<uint32_t*>(host + c->starts_off);
    FrameDesc* frames = reinterpret_cast<FrameDesc*>(host + c->frames_off);
    char* arena = static_cast<char*>(slot.arena);
    for (uint32_t i = 0; i < n; ++i) {
        StreamSlot& s = c->streams[req[i].stream_id];
        FrameDesc f{};
        f.w = (int32_t)s.w;
        f.h = (int32_t)s.h;
        f.history_valid = s.frames > 0 ? 1 : 0;
        const int rd = s.parity, wr = s.parity ^ 1;
        f.rad = dev_plane(req[i].radiance);
        f.den = dev_plane(req[i].denoised);
        f.mom_prev = s.mom[rd];
        f.mom_curr = s.mom[wr];
        if (c->mode == SVGF_MODE_CANONICAL) {
            f.g = dev_plane(req[i].gbuffer);
            if (f.history_valid) f.mv = dev_plane(req[i].motion);
            f.gprev = s.g_kept ? s.gkept : s.gbuf[rd];
            // kept by the caller: the next frame reads it in place
            if (!(req[i].flags & SVGF_REQUEST_KEEP_GBUFFER)) f.gsave = s.gbuf[wr];
            f.hist_prev = s.hist[rd];
            f.hist_curr = s.hist[wr];
        }
        carve_scratch(arena, c->w4, s.w, s.h, f);
        arena += frame_scratch_bytes(c->w4,(uint64_t)s.w * s.h);
        frames[i] = f;
    }
    // per stage: first tile of each frame (stage strides differ, so do the tile counts), and
    // the first ticket of each stage for the persistent form
    const int nstages = stage_count(c);
    const uint64_t total = build_schedule(
        n, [frames](uint32_t i, uint32_t& w, uint32_t& h) { w = (uint32_t)frames[i].w; h = (uint32_t)frames[i].h; },
        c->mode == SVGF_MODE_CANONICAL, nstages, hdr, starts);
    if (total >= (1ull << 32) - (1ull << 24)) return SVGF_ERR_CAPACITY;  // tickets are 32-bit
    hdr[HDR_CHUNK] = PERSIST_CHUNK;
    uint32_t tiles[MAX_STAGES];
    for (int st = 0; st < nstages; ++st) tiles[st] = starts[(size_t)st * STARTS_ROW + n];

    hipError_t e = hipSuccess;
    if (c->desc.batch_flags & (SVGF_BATCH_NO_GRAPH | SVGF_PROFILE_STAGES)) {
        e = enqueue_batch(c, slot, tiles, stream);
    } else {
        if (!slot.exec) {
            if (capture_slot(c, slot, tiles) != hipSuccess) return SVGF_ERR_HIP;
        } else {
            for (int i = 0; i < launches_per_batch(c) && e == hipSuccess; ++i) {
                hipKernelNodeParams kp = slot.kparams[i];
                kp.gridDim = dim3(launch_grid(c, tiles, i), 1, 1);
                e = hipGraphExecKernelNodeSetParams(slot.exec, slot.knodes[i], &kp);
            }
            if (e != hipSuccess) return SVGF_ERR_HIP;
        }
        e = hipGraphLaunch(slot.exec, stream);
    }
    if (e != hipSuccess) return SVGF_ERR_HIP;
    if (hipEventRecord(slot.done, stream) != hipSuccess) return SVGF_ERR_HIP;
    slot.pending = true;
    slot.stream = stream;
    slot.seq = ++c->last_seq;
    slot.checked = false;
    slot.failed = false;
    slot.ids.clear();
    for (uint32_t i = 0; i < n; ++i) slot.ids.push_back(req[i].stream_id);
    for (uint32_t i = 0; i < n; ++i) {
        StreamSlot& s = c->streams[req[i].stream_id];
        s.frames++;
        s.parity ^= 1;
        s.g_kept = c->mode == SVGF_MODE_CANONICAL && (req[i].flags & SVGF_REQUEST_KEEP_GBUFFER);
        s.gkept = s.g_kept ? frames[i].g : PlaneDev{};
    }
    return SVGF_OK;
}

svgf_status svgf_batch_last(svgf_context* c, uint64_t* batch)
{
    if (!c || !batch) return SVGF_ERR_INVALID_ARGUMENT;
    *batch = c->last_seq;
    return SVGF_OK;
}

svgf_status svgf_batch_status(svgf_context* c, uint64_t batch, int32_t wait)
{
    if (!c || batch == 0) return SVGF_ERR_INVALID_ARGUMENT;
    BatchSlot* s = nullptr;
    for (auto& sl : c->slots)
        if (sl.seq == batch) s = &sl;
    if (!s) {
        // its slot was reused since (the reuse waited for it and recorded its outcome)
        const auto& o = c->outcomes[batch % BATCH_OUTCOMES];
        if (o.seq != batch) return SVGF_ERR_INVALID_ARGUMENT;  // never enqueued, or too old
        return o.failed ? SVGF_ERR_TIMEOUT : SVGF_OK;
    }
    DeviceGuard guard(c->desc.device);
    if (s->pending) {
        const hipError_t q = wait ? hipEventSynchronize(s->done) : hipEventQuery(s->done);
        if (q == hipErrorNotReady) return SVGF_ERR_NOT_READY;
        if (q != hipSuccess) return SVGF_ERR_HIP;
        s->pending = false;
    }
    check_slot(c, *s);
    return s->failed ? SVGF_ERR_TIMEOUT : SVGF_OK;
}

svgf_status svgf_stage_times(svgf_context* c, float* ms, uint32_t capacity, uint32_t* count)
{
    if (!c || !count || (!ms && capacity)) return SVGF_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(c->desc.device);
    if (!(c->desc.batch_flags & SVGF_PROFILE_STAGES) && (c->desc.batch_flags & SVGF_BATCH_TIMED_STAGES)) {
        // mean over the batches completed since the previous call (waits for those in flight)
        for (auto& s : c->slots) {
            if (s.pending) {
                if (hipEventSynchronize(s.done) != hipSuccess) return SVGF_ERR_HIP;
                s.pending = false;
            }
            check_slot(c, s);
        }
        if (c->stage_batches == 0) return SVGF_ERR_NOT_READY;
        *count = (uint32_t)c->stage_n;
        for (int i = 0; i < c->stage_n && (uint32_t)i < capacity; ++i)
            ms[i] = (float)(c->stage_sum[i] / (double)c->stage_batches);
        std::fill(std::begin(c->stage_sum), std::end(c->stage_sum), 0.0);
        c->stage_batches = 0;
        return SVGF_OK;
    }
    if (!(c->desc.batch_flags & SVGF_PROFILE_STAGES) || c->ev_stages == 0) return SVGF_ERR_UNSUPPORTED;
    *count = (uint32_t)c->ev_stages;
    if (hipEventSynchronize(c->ev[c->ev_stages]) != hipSuccess) return SVGF_ERR_HIP;
    for (int i = 0; i < c->ev_stages && (uint32_t)i < capacity; ++i)
        if (hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]) != hipSuccess) return SVGF_ERR_HIP;
    return SVGF_OK;
}

}  // extern "C"

namespace svgf {
uint32_t context_storage(const svgf_context* c) { return c->desc.storage; }
uint32_t context_mode(const svgf_context* c) { return c->desc.mode; }
int context_device(const svgf_context* c) { return c->desc.device; }
const svgf_desc& context_desc(const svgf_context* c) { return c->desc; }
}  // namespace svgf
