// svgf_server.cpp — serving form: continuous batching of denoise requests over resident streams.
//
// TronGrid denoises one frame per render-thread iteration (src/main.cpp:907-947). Here many
// agent streams (docs/AI_INTERFACE.md:346-383 describes the brain-requested offscreen frames)
// submit frames of uneven sizes at random times. A scheduler thread keeps up to two batches in
// flight on their own HIP streams; whenever one finishes, the queued requests (FIFO, one per
// stream, streams with a frame still on the GPU wait their turn) form the next batch. A
// request's result never depends on its batch mates: every stage is a per-pixel function of its
// own frame's planes and state.
#include "../../include/svgf.h"
#include "svgf_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <random>
#include <thread>
#include <vector>

namespace {

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Pending {
    svgf_request req;
    uint64_t tag;
    uint64_t submit_ns;
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    uint64_t launch_ns = 0;
    uint64_t batch = 0;  // number of the batch the lane carries (svgf_batch_last)
    svgf_status status = SVGF_OK;
    std::vector<Pending> items;
};

constexpr int MAX_LANES = 2;

}  // namespace

struct svgf_server {
    svgf_context* ctx = nullptr;
    svgf_server_desc d{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> queue;
    std::vector<uint32_t> busy;  // per stream id: frames queued or in flight
    std::vector<svgf_completion> done;
    Lane lanes[MAX_LANES];
    uint64_t outstanding = 0;
    bool stop = false;
    std::thread th;
};

namespace {

void complete_lane(svgf_server* s, Lane& l, uint64_t t)
{
    for (auto& it : l.items) {
        svgf_completion c{};
        c.tag = it.tag;
        c.stream_id = it.req.stream_id;
        c.status = (int32_t)l.status;
        c.submit_ns = it.submit_ns;
        c.launch_ns = l.launch_ns;
        c.done_ns = t;
        c.batch_frames = (uint32_t)l.items.size();
        s->done.push_back(c);
        if (it.req.stream_id < s->busy.size() && s->busy[it.req.stream_id]) s->busy[it.req.stream_id]--;
        s->outstanding--;
    }
    l.items.clear();
    l.busy = false;
}

// Packs the next batch: FIFO over the queue, first request of each stream only, skipping
// streams whose previous frame is in flight, within the frame / pixel limits.
void form_batch(svgf_server* s, std::vector<Pending>& out, const std::vector<uint8_t>& inflight)
{
    out.clear();
    uint64_t px = 0;
    std::vector<uint32_t> seen;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        uint32_t id = it->req.stream_id;
        bool blocked = (id < inflight.size() && inflight[id]) || std::find(seen.begin(), seen.end(), id) != seen.end();
        if (!blocked) {
            svgf_stream_state st{};
            uint64_t fpx = 0;
            if (svgf_stream_state_get(s->ctx, id, &st) == SVGF_OK) fpx = (uint64_t)st.width * st.height;
            if (out.size() + 1 > s->d.max_batch_frames || px + fpx > s->d.max_batch_pixels) break;
            px += fpx;
            out.push_back(*it);
            seen.push_back(id);  // later frames of this stream wait for the next batch
            it = s->queue.erase(it);
            continue;
        }
        seen.push_back(id);
        ++it;
    }
}

void scheduler(svgf_server* s)
{
    (void)hipSetDevice(s->device);
    std::vector<Pending> batch;
    std::vector<svgf_request> reqs;
    std::vector<uint8_t> inflight;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool any_busy = false;
        for (auto& l : s->lanes) {
            if (!l.busy) continue;
            hipError_t q = hipEventQuery(l.ev);
            if (q == hipErrorNotReady) {
                any_busy = true;
                continue;
            }
            // the batch's own outcome (a persistent launch whose dependency wait gave up
            // reports SVGF_ERR_TIMEOUT; its streams restart cold)
            l.status = q != hipSuccess ? SVGF_ERR_HIP : svgf_batch_status(s->ctx, l.batch, 1);
            complete_lane(s, l, now_ns());
            s->cv_done.notify_all();
        }
        for (uint32_t li = 0; li < s->d.lanes && !s->queue.empty(); ++li) {
            Lane& l = s->lanes[li];
            if (l.busy) continue;
            if (s->d.max_wait_us && any_busy && s->queue.size() < s->d.max_batch_frames &&
                now_ns() - s->queue.front().submit_ns < (uint64_t)s->d.max_wait_us * 1000)
                break;  // let a partial batch fill while the GPU is busy anyway
            inflight.assign(s->busy.size(), 0);
            for (auto& o : s->lanes)
                for (auto& it : o.items)
                    if (it.req.stream_id < inflight.size()) inflight[it.req.stream_id] = 1;
            form_batch(s, batch, inflight);
            if (batch.empty()) break;
            reqs.resize(batch.size());
            for (size_t i = 0; i < batch.size(); ++i) reqs[i] = batch[i].req;
            l.items = batch;
            l.launch_ns = now_ns();
            l.status = svgf_denoise_batch(s->ctx, reqs.data(), (uint32_t)reqs.size(), l.stream);
            if (l.status == SVGF_OK) l.status = svgf_batch_last(s->ctx, &l.batch);
            if (l.status == SVGF_OK && hipEventRecord(l.ev, l.stream) != hipSuccess) l.status = SVGF_ERR_HIP;
            if (l.status != SVGF_OK) {
                complete_lane(s, l, now_ns());
                s->cv_done.notify_all();
                continue;
            }
            l.busy = true;
            any_busy = true;
        }
        if (s->stop && s->queue.empty() && !any_busy) break;
        if (any_busy) {
            // poll the in-flight batches without holding the lock
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else if (s->queue.empty()) {
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        } else {
            // queued requests all blocked behind in-flight frames of their stream: cannot
            // happen without a busy lane; wait briefly for safety
            s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

}  // namespace

extern "C" {

svgf_status svgf_server_create(svgf_context* ctx, const svgf_server_desc* d, svgf_server** out)
{
    if (!ctx || !d || !out) return SVGF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const svgf_desc& cd = svgf::context_desc(ctx);
    if (d->lanes < 1 || d->lanes > MAX_LANES || d->max_batch_frames < 1 || d->max_batch_pixels < 1)
        return SVGF_ERR_INVALID_ARGUMENT;
    if (!cd.max_streams || !cd.max_batch_frames) return SVGF_ERR_UNSUPPORTED;
    if (d->max_batch_frames > cd.max_batch_frames || d->max_batch_pixels > cd.max_batch_pixels)
        return SVGF_ERR_CAPACITY;
    svgf_server* s = new (std::nothrow) svgf_server();
    if (!s) return SVGF_ERR_OUT_OF_MEMORY;
    s->ctx = ctx;
    s->d = *d;
    s->device = cd.device;
    s->busy.assign(cd.max_streams, 0);
    (void)hipSetDevice(s->device);
    for (uint32_t i = 0; i < d->lanes; ++i) {
        if (hipStreamCreateWithFlags(&s->lanes[i].stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->lanes[i].ev, hipEventDisableTiming) != hipSuccess) {
            for (auto& l : s->lanes) {
                if (l.ev) (void)hipEventDestroy(l.ev);
                if (l.stream) (void)hipStreamDestroy(l.stream);
            }
            delete s;
            return SVGF_ERR_HIP;
        }
    }
    s->th = std::thread(scheduler, s);
    *out = s;
    return SVGF_OK;
}

svgf_status svgf_server_destroy(svgf_server* s)
{
    if (!s) return SVGF_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->th.join();
    for (auto& l : s->lanes) {
        if (l.stream) (void)hipStreamSynchronize(l.stream);
        if (l.ev) (void)hipEventDestroy(l.ev);
        if (l.stream) (void)hipStreamDestroy(l.stream);
    }
    delete s;
    return SVGF_OK;
}

svgf_status svgf_server_stream_open(svgf_server* s, uint32_t w, uint32_t h, uint32_t* id)
{
    if (!s) return SVGF_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    svgf_status st = svgf_stream_open(s->ctx, w, h, id);
    if (st == SVGF_OK) s->busy[*id] = 0;
    return st;
}

svgf_status svgf_server_stream_close(svgf_server* s, uint32_t id)
{
    if (!s) return SVGF_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    if (id >= s->busy.size()) return SVGF_ERR_NO_STREAM;
    // wait until the stream has nothing queued or in flight
    s->cv_done.wait(lk, [&] { return s->busy[id] == 0; });
    return svgf_stream_close(s->ctx, id);
}

svgf_status svgf_server_submit(svgf_server* s, const svgf_request* req, uint64_t tag)
{
    if (!s || !req || (req->flags & ~SVGF_REQUEST_KEEP_GBUFFER)) return SVGF_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return SVGF_ERR_INVALID_ARGUMENT;
    svgf_stream_state st{};
    svgf_status e = svgf_stream_state_get(s->ctx, req->stream_id, &st);
    if (e != SVGF_OK) return e;
    if ((uint64_t)st.width * st.height > s->d.max_batch_pixels) return SVGF_ERR_CAPACITY;
    s->queue.push_back(Pending{*req, tag, now_ns()});
    s->busy[req->stream_id]++;
    s->outstanding++;
    s->cv_work.notify_one();
    return SVGF_OK;
}

svgf_status svgf_server_poll(svgf_server* s, svgf_completion* out, uint32_t cap, uint32_t* count, uint32_t wait_us)
{
    if (!s || !count || (!out && cap)) return SVGF_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_us)
        s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !s->done.empty(); });
    uint32_t n = (uint32_t)std::min<size_t>(cap, s->done.size());
    std::copy(s->done.begin(), s->done.begin() + n, out);
    s->done.erase(s->done.begin(), s->done.begin() + n);
    *count = n;
    return SVGF_OK;
}

svgf_status svgf_server_drain(svgf_server* s)
{
    if (!s) return SVGF_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    return SVGF_OK;
}

svgf_status svgf_server_simulate(svgf_server* s, const svgf_sim_desc* d, svgf_sim_result* out)
{
    if (!s || !d || !out || d->streams < 1 || d->n_classes < 1 || !d->class_width || !d->class_height ||
        !d->class_inputs || d->inputs_per_class < 1 || !(d->rate_hz > 0.0) || !(d->duration_s > 0.0) ||
        d->churn_per_s < 0.0)
        return SVGF_ERR_INVALID_ARGUMENT;
    const svgf_desc& cd = svgf::context_desc(s->ctx);
    if (d->streams > cd.max_streams) return SVGF_ERR_CAPACITY;
    *out = svgf_sim_result{};
    const bool fp16 = cd.storage == SVGF_STORAGE_FP16;
    const uint32_t pxb = fp16 ? 8 : 16;
    const uint32_t fmt = fp16 ? SVGF_FMT_RGBA16F : SVGF_FMT_RGBA32F;
    uint32_t maxw = 0, maxh = 0;
    for (uint32_t c = 0; c < d->n_classes; ++c) {
        maxw = std::max(maxw, d->class_width[c]);
        maxh = std::max(maxh, d->class_height[c]);
    }
    (void)hipSetDevice(s->device);
    // one output buffer per agent slot at the largest class size (no allocation during the run)
    void* outbuf = nullptr;
    if (hipMalloc(&outbuf, (size_t)d->streams * maxw * maxh * pxb) != hipSuccess) return SVGF_ERR_OUT_OF_MEMORY;

    struct Agent {
        uint32_t id = 0, cls = 0, frame = 0;
        bool open = false;
    };
    std::mt19937_64 rng(d->seed);
    std::uniform_int_distribution<uint32_t> pick_cls(0, d->n_classes - 1), pick_agent(0, d->streams - 1);
    std::exponential_distribution<double> gap(d->rate_hz);
    std::vector<Agent> agents(d->streams);
    svgf_sim_result r{};
    svgf_status status = SVGF_OK;
    auto open_agent = [&](Agent& a) {
        a.cls = pick_cls(rng);
        a.frame = 0;
        svgf_status e = svgf_server_stream_open(s, d->class_width[a.cls], d->class_height[a.cls], &a.id);
        a.open = e == SVGF_OK;
        if (a.open) r.streams_opened++;
        return e;
    };
    for (auto& a : agents) {
        svgf_status e = open_agent(a);
        if (e != SVGF_OK && e != SVGF_ERR_OVER_BUDGET) status = e;
    }
    std::vector<double> lat_ms;
    lat_ms.reserve((size_t)(d->rate_hz * d->duration_s * 1.2) + 16);
    std::vector<svgf_completion> comp(256);
    std::vector<uint64_t> tag_px;  // tag -> pixels
    uint64_t first_ns = 0, last_done = 0;
    double batch_sum = 0.0;  // sum over completions of 1 / (frames in its batch) = batches
    auto drain_completions = [&](uint32_t wait_us) {
        uint32_t n = 0;
        svgf_server_poll(s, comp.data(), (uint32_t)comp.size(), &n, wait_us);
        for (uint32_t i = 0; i < n; ++i) {
            const auto& c = comp[i];
            if (c.status == SVGF_OK) {
                r.completed++;
                batch_sum += 1.0 / (double)std::max<uint32_t>(1, c.batch_frames);
                r.pixels += tag_px[c.tag];
                lat_ms.push_back((double)(c.done_ns - c.submit_ns) * 1e-6);
                last_done = std::max(last_done, c.done_ns);
            } else if (status == SVGF_OK) {
                status = (svgf_status)c.status;
            }
        }
        return n;
    };

    const uint64_t t0 = now_ns();
    const uint64_t t_end = t0 + (uint64_t)(d->duration_s * 1e9);
    double t_arr = 0.0;
    double next_churn = d->churn_per_s > 0.0
                            ? std::exponential_distribution<double>(d->churn_per_s * d->streams)(rng)
                            : 1e300;
    std::exponential_distribution<double> churn_gap(d->churn_per_s > 0.0 ? d->churn_per_s * d->streams : 1.0);
    uint64_t offered_px = 0;
    while (status == SVGF_OK) {
        t_arr += gap(rng);
        uint64_t t_next = t0 + (uint64_t)(t_arr * 1e9);
        if (t_next >= t_end) break;
        while (next_churn <= t_arr) {  // replace an agent: close, reopen cold with a new size
            Agent& a = agents[pick_agent(rng)];
            if (a.open) {
                svgf_server_stream_close(s, a.id);
                r.streams_closed++;
                a.open = false;
            }
            svgf_status e = open_agent(a);
            if (e != SVGF_OK && e != SVGF_ERR_OVER_BUDGET) status = e;
            next_churn += churn_gap(rng);
        }
        while (now_ns() < t_next) {
            if (!drain_completions(0)) std::this_thread::yield();
        }
        uint32_t ai = pick_agent(rng);
        Agent& a = agents[ai];
        if (!a.open) {
            r.rejected++;
            continue;
        }
        const svgf_request& in = d->class_inputs[a.cls * d->inputs_per_class + a.frame % d->inputs_per_class];
        svgf_request q = in;
        q.stream_id = a.id;
        uint32_t w = d->class_width[a.cls], h = d->class_height[a.cls];
        q.denoised = svgf_plane{static_cast<char*>(outbuf) + (size_t)ai * maxw * maxh * pxb, (uint64_t)w * pxb, fmt, 0};
        tag_px.push_back((uint64_t)w * h);
        if (!first_ns) first_ns = now_ns();
        svgf_status e = svgf_server_submit(s, &q, tag_px.size() - 1);
        if (e != SVGF_OK) {
            status = e;
            break;
        }
        offered_px += (uint64_t)w * h;
        r.submitted++;
        a.frame++;
    }
    svgf_server_drain(s);
    while (drain_completions(0)) {
    }
    for (auto& a : agents)
        if (a.open) {
            svgf_server_stream_close(s, a.id);
            r.streams_closed++;
        }
    (void)hipFree(outbuf);
    r.batches = (uint64_t)std::llround(batch_sum);
    r.mean_batch_frames = batch_sum > 0 ? (double)r.completed / batch_sum : 0.0;
    r.wall_s = last_done > first_ns ? (double)(last_done - first_ns) * 1e-9 : 0.0;
    r.mpx_per_s = r.wall_s > 0 ? (double)r.pixels / r.wall_s * 1e-6 : 0.0;
    r.offered_mpx_per_s = (double)offered_px / d->duration_s * 1e-6;
    if (!lat_ms.empty()) {
        std::vector<double> v = lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            double pos = p * (double)(v.size() - 1);
            size_t lo = (size_t)std::floor(pos), hi = std::min(v.size() - 1, lo + 1);
            return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
        };
        r.latency_p50_ms = pct(0.50);
        r.latency_p99_ms = pct(0.99);
        r.latency_max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        r.latency_mean_ms = sum / (double)v.size();
    }
    *out = r;
    return status;
}

}  // extern "C"
