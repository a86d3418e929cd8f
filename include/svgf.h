/*
 * svgf.h — C ABI of the MI355X-native SVGF denoiser (gfx950, HIP).
 *
 * Drop-in replacement for the SVGF compute pass of TronGrid:
 *   - the shader          src/svgf_denoise.slang:61-275 (bindings, push block, svgfMain)
 *   - its host recording  src/main.cpp:907-947 (4 dispatches + barriers inside recordFrame)
 *   - its resources       src/main.cpp:1326-1496 (7 images, zero-init = cold history)
 *   - its descriptor sets src/main.cpp:2288-2446 (moments / denoised ping-pong)
 *
 * Conventions
 *   - Every function returns an svgf_status; nothing throws or aborts across the ABI.
 *   - Planes are plain device pointers with a row pitch in bytes. The caller owns every
 *     per-frame plane; the library owns only its scratch and (batched form) per-stream
 *     history, exactly as TronGrid's render thread owns its images (main.cpp:1333-1370).
 *   - `hip_stream` is a hipStream_t passed as void* (NULL = the default stream). All work is
 *     enqueued on it; calls return before the GPU finishes, like vkCmdDispatch.
 *   - A context is externally synchronised: one host thread uses it at a time.
 */
#ifndef SVGF_H
#define SVGF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVGF_ABI_VERSION 1u

#if defined(__GNUC__) || defined(__clang__)
#define SVGF_API __attribute__((visibility("default")))
#else
#define SVGF_API
#endif

typedef enum svgf_status {
    SVGF_OK = 0,
    SVGF_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, pitch, format or parameter */
    SVGF_ERR_UNSUPPORTED = 2,      /* valid request this build does not implement */
    SVGF_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed */
    SVGF_ERR_OVER_BUDGET = 4,      /* stream admission would exceed the HBM budget */
    SVGF_ERR_NO_STREAM = 5,        /* unknown or closed stream id */
    SVGF_ERR_CAPACITY = 6,         /* more streams / pixels than the context was created for */
    SVGF_ERR_HIP = 7,              /* a HIP runtime call failed */
    SVGF_ERR_NO_DEVICE = 8,        /* no usable gfx950 device */
    SVGF_ERR_TIMEOUT = 9,          /* a bounded wait gave up (batched scheduler) */
    SVGF_ERR_NOT_READY = 10        /* a non-blocking query found the work still in flight */
} svgf_status;

/* Algorithm selection.
 * REFERENCE: the exact semantics of svgf_denoise.slang (same-pixel luminance moments, variance
 *            from temporal moments only, luminance-only edge stop, sigma passed through the
 *            alpha channel, output alpha = 1). Planes used: radiance, moments_prev,
 *            moments_curr, denoised.
 * CANONICAL: Schied et al. 2017 SVGF (cited at svgf_denoise.slang:18): motion-vector
 *            reprojection with depth/normal history rejection, colour history, 3x3 spatial
 *            variance for young history, depth/normal/luminance edge stops, variance
 *            filtered by squared weights. Uses every plane of svgf_frame. */
typedef enum svgf_mode { SVGF_MODE_REFERENCE = 0, SVGF_MODE_CANONICAL = 1 } svgf_mode;

/* Storage width of radiance, history colour, moments, wavelet scratch and output.
 * FP32: every plane float (the Slang CPU-target semantics of the reference).
 * FP16: those planes half, all arithmetic float (the Vulkan rgba16f / rg16f formats of
 *       main.cpp:1333-1370). G-buffer and motion stay FP32 in both. */
typedef enum svgf_storage { SVGF_STORAGE_FP32 = 0, SVGF_STORAGE_FP16 = 1 } svgf_storage;

typedef enum svgf_format {
    SVGF_FMT_RGBA32F = 1, /* 16 B/px */
    SVGF_FMT_RGBA16F = 2, /*  8 B/px */
    SVGF_FMT_RG32F = 3,   /*  8 B/px */
    SVGF_FMT_RG16F = 4,   /*  4 B/px */
    SVGF_FMT_R32UI = 5    /*  4 B/px: packed codes (octahedral normals of svgf_pack) */
} svgf_format;

/* Flags of svgf_params.flags */
#define SVGF_FLAG_COMPENSATED_MOMENTS 0x1u /* moments kept as (mean, sqrt(centred 2nd moment)),
                                              updated by the West/Finch recursion; never forms
                                              E[x^2]. Default for FP16 storage. */
#define SVGF_FLAG_WIDE_STATE 0x2u /* CANONICAL FP16 with compensated moments only (default
    there): the fed-back state carries its binary16 rounding residual, so the rounding does not
    compound from frame to frame (history_cap <= 63).
      history planes, RGBA16F: rgb = binary16 of the colour; the alpha slot is a 16-bit word, not
        a half: bits 0-5 history length, bits 6-8 / 9-11 / 12-14 residual code q of r / g / b
        (two's complement -4..3); a channel with binary16 value h of biased exponent e decodes as
        h + q 2^(max(e, 1) - 28).
      moments planes, RGBA16F (not RG16F): (mean_hi, sigma_hi, mean_lo, sigma_lo); a value
        decodes as hi + lo 2^(max(e, 1) - 25), e the biased exponent of hi.
    Without the flag the FP16 history alpha is the length as a half and the moments are RG16F. */

typedef struct svgf_params {
    uint32_t iterations;  /* a-trous iterations, strides 1,2,4,...; reference 3 (main.cpp:929) */
    float phi_luma;       /* luminance edge-stop scale; reference SIGMA_COLOR_BASE = 2
                             (svgf_denoise.slang:137) */
    float phi_normal;     /* CANONICAL: exponent of max(0, n.n') for unit normals (n.n' is
                             taken as at most 1); 0 disables */
    float phi_depth;      /* CANONICAL: depth edge-stop scale; 0 disables */
    float alpha_color;    /* CANONICAL: colour history blend */
    float alpha_moments;  /* moments blend; reference ALPHA_MOMENTS = 0.2 (svgf_denoise.slang:144) */
    float variance_eps;   /* reference VARIANCE_EPSILON = 1e-4 (svgf_denoise.slang:150) */
    float depth_eps;      /* CANONICAL: additive term of the depth edge-stop denominator */
    float depth_reject;   /* CANONICAL: history tap rejected when |z'-z| > depth_reject * z */
    float normal_reject;  /* CANONICAL: history tap rejected when n.n' < normal_reject */
    uint32_t history_cap; /* CANONICAL: cap of the history length */
    uint32_t flags;       /* SVGF_FLAG_* */
} svgf_params;

typedef struct svgf_desc {
    uint32_t mode;             /* svgf_mode */
    uint32_t storage;          /* svgf_storage */
    svgf_params params;
    uint32_t max_width;        /* largest frame svgf_denoise will see (scratch sizing) */
    uint32_t max_height;
    uint32_t max_streams;      /* batched form: resident streams; 0 = single-frame use only */
    uint32_t max_batch_frames; /* batched form: requests per svgf_denoise_batch call */
    uint64_t max_batch_pixels; /* batched form: total pixels per svgf_denoise_batch call */
    uint64_t hbm_budget_bytes; /* batched form: admission budget for stream state; 0 = none */
    int32_t device;            /* HIP device ordinal */
    uint32_t batch_flags;      /* SVGF_BATCH_* */
} svgf_desc;

/* Flags of svgf_desc.batch_flags */
#define SVGF_BATCH_NO_GRAPH 0x1u /* launch the batched stages directly instead of replaying a
                                    captured hipGraph */
#define SVGF_PROFILE_STAGES 0x2u /* record a hipEvent around every stage launch (no graph,
                                    one launch per stage); read back with svgf_stage_times */
#define SVGF_BATCH_PERSISTENT 0x4u /* run all stages of a batch in one persistent launch: the
                                      workgroups pull (stage, frame, tile) tickets in dependency
                                      order and a tile waits until every tile of the previous
                                      stage is done, instead of one launch per stage */
#define SVGF_BATCH_TIMED_STAGES 0x8u /* batched form: record a hipEvent before and after every
                                        stage launch of each batch, inside the captured graph
                                        (event nodes), so that per-stage GPU times come from the
                                        batches themselves; svgf_stage_times then returns their
                                        mean over the batches completed since its previous call
                                        (the persistent form: one stage, the whole launch) */

typedef struct svgf_plane {
    void* ptr;            /* device pointer to row 0, pixel 0 */
    uint64_t pitch_bytes; /* bytes between rows; >= width * pixel size, multiple of pixel size */
    uint32_t format;      /* svgf_format */
    uint32_t reserved;
} svgf_plane;

/* One frame of the single-frame form (caller-owned planes).
 * Reference binding map (svgf_denoise.slang:61-94, main.cpp:2314-2342):
 *   radiance = binding 0 lighting_raw, denoised = binding 3, moments_prev = binding 4,
 *   moments_curr = binding 5. wavelet_a / wavelet_b (bindings 1, 2) are library scratch.
 * CANONICAL adds gbuffer (unit normal.xyz, linear depth; RGBA32F), gbuffer_prev, motion (RG32F,
 *   pixels, previous position = current position + motion) and the colour history
 *   (rgb, history length in alpha). */
typedef struct svgf_frame {
    uint32_t width;
    uint32_t height;
    uint32_t history_valid; /* 0 = cold history: *_prev planes are ignored (REFERENCE reads
                               them as zero, as after main.cpp:1449-1455); 1 = use them */
    uint32_t reserved;
    svgf_plane radiance;
    svgf_plane gbuffer;
    svgf_plane gbuffer_prev;
    svgf_plane motion;
    svgf_plane history_prev;
    svgf_plane history_curr;
    svgf_plane moments_prev;
    svgf_plane moments_curr;
    svgf_plane denoised;
} svgf_frame;

/* One request of the batched form: per-frame inputs and output; history is library-resident
 * per stream (opened with svgf_stream_open). The stream's next frame reprojects against this
 * frame's G-buffer: by default the temporal stage saves a copy of it in the stream's state
 * (16 B per pixel written). */
/* svgf_request.flags (CANONICAL; other bits are SVGF_ERR_INVALID_ARGUMENT):
 * the caller leaves this request's gbuffer plane unchanged and allocated until the stream's
 * next frame has completed on the GPU (a G-buffer kept per frame in flight, as the renderer
 * keeps its per-frame resources by main.cpp:3148's current_frame flip, and written on the
 * stream that runs the batches meets this). The next frame then reads it in place as its
 * previous G-buffer, and no copy is saved. Results are bit-identical to the copying form. */
#define SVGF_REQUEST_KEEP_GBUFFER 0x1u
typedef struct svgf_request {
    uint32_t stream_id;
    uint32_t flags; /* SVGF_REQUEST_* */
    svgf_plane radiance;
    svgf_plane gbuffer; /* CANONICAL only */
    svgf_plane motion;  /* CANONICAL only */
    svgf_plane denoised;
} svgf_request;

/* Device pointers of a stream's resident state for the frame the NEXT request will read
 * (history colour, moments, previous G-buffer). For export / inspection. */
typedef struct svgf_stream_state {
    uint32_t width;
    uint32_t height;
    uint32_t frames;        /* frames denoised since open/reset */
    uint32_t reserved;
    svgf_plane history;     /* CANONICAL */
    svgf_plane moments;
    svgf_plane gbuffer;     /* CANONICAL */
} svgf_stream_state;

typedef struct svgf_context svgf_context;

/* --- library ---------------------------------------------------------------------------- */
SVGF_API uint32_t svgf_abi_version(void);
SVGF_API const char* svgf_status_str(svgf_status status);
/* Defaults: REFERENCE = the shader's constants (3 iterations, phi 2, alpha 0.2, eps 1e-4);
 * CANONICAL = 5 iterations, phi_luma 4, phi_normal 128, phi_depth 1, alpha 0.2/0.2, history
 * cap 32; with FP16 storage also SVGF_FLAG_COMPENSATED_MOMENTS | SVGF_FLAG_WIDE_STATE. */
SVGF_API svgf_status svgf_default_params(uint32_t mode, uint32_t storage, svgf_params* out);
/* Bytes the library allocates: scratch for a context (at max_width x max_height and the
 * batch limits) and the resident state of one stream of w x h. */
SVGF_API svgf_status svgf_query_bytes(const svgf_desc* desc, uint32_t w, uint32_t h,
                             uint64_t* context_bytes, uint64_t* stream_bytes);

/* --- context (replaces pipeline + descriptor-set creation, main.cpp:2288-2446) ---------- */
SVGF_API svgf_status svgf_create(const svgf_desc* desc, svgf_context** out);
SVGF_API svgf_status svgf_destroy(svgf_context* ctx);
SVGF_API svgf_status svgf_set_params(svgf_context* ctx, const svgf_params* params);

/* --- single frame (replaces the SVGF block of recordFrame, main.cpp:907-947) ------------
 * The calls of one context share its scratch: a call on another hip_stream than the previous
 * call first waits, on the device, for that call, so calls on different streams run one after
 * the other on the GPU (they never race through the scratch). */
SVGF_API svgf_status svgf_denoise(svgf_context* ctx, const svgf_frame* frame, void* hip_stream);

/* --- single frame without allocation --------------------------------------------------------
 * No counterpart in the reference, whose pass always runs inside an allocated pipeline. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): svgf_denoise without a context, on the current HIP device,
 * with mode, storage and parameters given per call. It calls no hipMalloc, hipFree or other
 * allocating routine: its scratch is a fixed device array of the library holding one frame of
 * up to SVGF_STATIC_MAX_PIXELS pixels (41 B per pixel, about 38 MB, on each GPU the library runs
 * on, until the process ends); a larger frame is SVGF_ERR_CAPACITY. Results equal svgf_denoise's
 * bit for bit. Every call on a GPU uses that GPU's one array and nothing in it needs clearing:
 * calls that overlap on one GPU (other streams, other threads) race and give wrong results, so
 * order them on one stream. */
#define SVGF_STATIC_MAX_PIXELS (1280u * 720u)
SVGF_API svgf_status svgf_denoise_static(uint32_t mode, uint32_t storage, const svgf_params* params,
                                         const svgf_frame* frame, void* hip_stream);

/* --- resident streams (history reset = main.cpp:1570-1603 resize semantics) ------------- */
SVGF_API svgf_status svgf_stream_open(svgf_context* ctx, uint32_t width, uint32_t height, uint32_t* stream_id);
SVGF_API svgf_status svgf_stream_close(svgf_context* ctx, uint32_t stream_id);
SVGF_API svgf_status svgf_stream_reset(svgf_context* ctx, uint32_t stream_id, void* hip_stream);
SVGF_API svgf_status svgf_stream_state_get(svgf_context* ctx, uint32_t stream_id, svgf_stream_state* out);
/* Bytes of the HBM budget in use by open streams. */
SVGF_API svgf_status svgf_budget_used(svgf_context* ctx, uint64_t* bytes);

/* --- batched form: n requests of uneven sizes in one ragged-tile launch sequence -------- *
 * Each call that returns SVGF_OK enqueues one batch, numbered 1, 2, ... (svgf_batch_last).
 * With SVGF_BATCH_PERSISTENT, a batch is one kernel launch; its dependency waits are bounded
 * (250 ms each; the environment variable SVGF_PERSISTENT_WAIT_US, read by svgf_create, sets
 * another bound). A batch whose wait gave up has wrong outputs; once its completion is
 * observed (by svgf_batch_status, or by a later svgf_denoise_batch / svgf_stream_close call)
 * its streams restart with cold history, so no later frame reprojects the wrong state. In the
 * persistent form a call whose requests name a stream of a batch still in flight first waits
 * for that batch on the host; in the per-stage form, when that batch was enqueued on another
 * hip_stream, this call's hip_stream waits for it on the device (hipStreamWaitEvent), so the
 * frames of a stream run in call order whatever streams the calls use. SVGF_ERR_CAPACITY: more requests or pixels than svgf_desc allows,
 * or (without SVGF_BATCH_PERSISTENT) a stage with more than 16777215 tiles, which only strides
 * of 256 and 512 reach. A REFERENCE frame has at least 65536 tiles at 9 passes and 262144 at 10,
 * so keep such batches under 256 and 64 frames. A CANONICAL frame whose lattice at the stride has
 * at most 40 points has one tile per 256 pixels; a wider one has at least s^2 / 2 (32768 at
 * stride 256, 131072 at 512). */
SVGF_API svgf_status svgf_denoise_batch(svgf_context* ctx, const svgf_request* requests, uint32_t n,
                               void* hip_stream);
/* Number of the batch the last successful svgf_denoise_batch call enqueued (0: none yet). */
SVGF_API svgf_status svgf_batch_last(svgf_context* ctx, uint64_t* batch);
/* Outcome of a batch: SVGF_OK when it completed correctly, SVGF_ERR_TIMEOUT when a dependency
 * wait of its persistent launch gave up (its outputs are wrong; its streams have been reset to
 * cold history), SVGF_ERR_NOT_READY while it runs (wait = 0; wait = 1 blocks until it has
 * completed). The outcomes of the last 64 batches are kept; an older or unknown number is
 * SVGF_ERR_INVALID_ARGUMENT. */
SVGF_API svgf_status svgf_batch_status(svgf_context* ctx, uint64_t batch, int32_t wait);

/* --- per-stage GPU time ---------------------------------------------------------------
 * With SVGF_PROFILE_STAGES in svgf_desc.batch_flags: of the last svgf_denoise /
 * svgf_denoise_batch call (hipEvents around each stage launch, the counterpart of the
 * reference's timestamp pair around the SVGF chain, main.cpp:907, 947); waits for that call to
 * finish. With SVGF_BATCH_TIMED_STAGES (and not SVGF_PROFILE_STAGES): the mean over the batches
 * completed since the previous call, waiting for those in flight (SVGF_ERR_NOT_READY if none).
 * Stage order: REFERENCE init, a-trous 0..N-1; CANONICAL temporal, a-trous 0..N-1 (a-trous 0
 * also computes the spatial variance of young pixels). *count receives the number of stages. */
SVGF_API svgf_status svgf_stage_times(svgf_context* ctx, float* ms, uint32_t capacity, uint32_t* count);

/* --- serving form: continuous batching over resident streams -----------------------------
 * A server owns a context (created with max_streams / max_batch_* set) and a scheduler thread.
 * Requests are queued by svgf_server_submit from any thread. Whenever one of `lanes` in-flight
 * slots is free, the scheduler packs queued requests (FIFO, at most one per stream, a stream
 * whose previous frame is still on the GPU waits) into one svgf_denoise_batch on that lane's
 * HIP stream, so batches overlap on the device and new arrivals join the next batch. Results
 * of a request do not depend on what it was batched with. While a server runs, use the
 * svgf_server_* stream calls instead of svgf_stream_* on its context. */
typedef struct svgf_server svgf_server;

typedef struct svgf_server_desc {
    uint32_t lanes;            /* batches in flight at once (1..2) */
    uint32_t max_batch_frames; /* <= context max_batch_frames */
    uint64_t max_batch_pixels; /* <= context max_batch_pixels */
    uint32_t max_wait_us;      /* hold a partial batch this long for more requests (0 = none) */
    uint32_t reserved;
} svgf_server_desc;

typedef struct svgf_completion {
    uint64_t tag;          /* caller's tag from svgf_server_submit */
    uint32_t stream_id;
    int32_t status;        /* svgf_status of the batch that carried it */
    uint64_t submit_ns;    /* steady-clock times: queued, launched, observed complete */
    uint64_t launch_ns;
    uint64_t done_ns;
    uint32_t batch_frames; /* frames in the batch that carried it */
    uint32_t reserved;
} svgf_completion;

SVGF_API svgf_status svgf_server_create(svgf_context* ctx, const svgf_server_desc* desc, svgf_server** out);
SVGF_API svgf_status svgf_server_destroy(svgf_server* srv);
SVGF_API svgf_status svgf_server_stream_open(svgf_server* srv, uint32_t width, uint32_t height, uint32_t* stream_id);
SVGF_API svgf_status svgf_server_stream_close(svgf_server* srv, uint32_t stream_id);
/* Thread-safe. The request's planes must stay valid until its completion is polled. */
SVGF_API svgf_status svgf_server_submit(svgf_server* srv, const svgf_request* req, uint64_t tag);
/* Returns up to `capacity` completions, waiting up to wait_us for the first one. */
SVGF_API svgf_status svgf_server_poll(svgf_server* srv, svgf_completion* out, uint32_t capacity,
                                      uint32_t* count, uint32_t wait_us);
/* Blocks until every submitted request has completed (completions stay pollable). */
SVGF_API svgf_status svgf_server_drain(svgf_server* srv);

/* Open-loop load test of a server (the serving benchmark): Poisson arrivals at rate_hz from
 * `streams` agent streams whose resolutions are drawn uniformly from the class list, streams
 * replaced (closed, reopened cold with a new resolution) at churn_per_s of the population per
 * second. Inputs come from `inputs_per_class` device frames per class (radiance, gbuffer,
 * motion of class_inputs[class * inputs_per_class + k]; their `denoised` field is ignored);
 * outputs go to per-stream buffers the simulation allocates. */
typedef struct svgf_sim_desc {
    uint32_t streams;
    uint32_t n_classes;
    const uint32_t* class_width;
    const uint32_t* class_height;
    const svgf_request* class_inputs;
    uint32_t inputs_per_class;
    uint32_t reserved;
    double rate_hz;
    double duration_s;
    double churn_per_s;
    uint64_t seed;
} svgf_sim_desc;

typedef struct svgf_sim_result {
    uint64_t submitted, completed, rejected; /* rejected: admission over budget */
    uint64_t batches, streams_opened, streams_closed;
    uint64_t pixels;             /* completed pixels */
    double wall_s;               /* first arrival to last completion */
    double mpx_per_s;
    double latency_p50_ms, latency_p99_ms, latency_max_ms, latency_mean_ms;
    double mean_batch_frames;
    double offered_mpx_per_s;
} svgf_sim_result;

SVGF_API svgf_status svgf_server_simulate(svgf_server* srv, const svgf_sim_desc* desc, svgf_sim_result* out);

/* --- seeded synthetic G-buffer sequences (host memory; no GPU needed) -------------------- */
typedef struct svgf_synth_desc {
    uint32_t width;
    uint32_t height;
    uint64_t seed;          /* per-frame noise seed */
    uint64_t stream;        /* agent-stream key */
    uint32_t frame;         /* frame index within the sequence (camera position) */
    uint32_t noise;         /* 0 = exponential (mean 1), 1 = lognormal (sigma 1.5, mean 1),
                               2 = none (the noise-free radiance, the mean of both) */
    float pan_px;           /* camera pan per frame in pixels (motion vectors) */
    float disocclusion;     /* fraction of pixels whose G-buffer is forced inconsistent */
    uint32_t threads;       /* host threads; 0 = hardware concurrency */
    uint32_t reserved;
} svgf_synth_desc;
/* Fills tightly packed host arrays (any may be NULL): radiance RGBA32F, gbuffer RGBA32F
 * (normal.xyz, depth), motion RG32F. */
SVGF_API svgf_status svgf_synth_frame(const svgf_synth_desc* desc, float* radiance, float* gbuffer,
                             float* motion);

/* --- agent frames on the wire ------------------------------------------------------------
 * The offscreen-frame layout an agent ("brain") reads (docs/AI_INTERFACE.md:346-383,
 * TgVisionData): per frame RGB888, row-major from the top row, no row padding (3 w h bytes),
 * and float32 depth normalised to [0, 1] (4 w h bytes). Colour is the denoised HDR output
 * through TronGrid's display tonemap (src/postprocess.slang:47-103, 153-160: fitted ACES, then
 * sRGB), each channel quantised to floor(255 srgb + 1/2); the display-only screen effects
 * (bloom, chromatic aberration, grade, vignette) are not applied. Depth is
 * (z - near) / (far - near) clamped to [0, 1]; pixels without geometry (z <= 0, not finite) read
 * 1. One call encodes the frames of many agents in a single launch. */
typedef struct svgf_wire svgf_wire;

typedef struct svgf_wire_params {
    float exposure;   /* linear scale before the tonemap (1: none) */
    float depth_near; /* depth mapped to 0 (camera near plane, src/camera.hpp:99) */
    float depth_far;  /* depth mapped to 1 (camera far plane, src/camera.hpp:100) */
    uint32_t reserved;
} svgf_wire_params;

typedef struct svgf_wire_frame {
    uint32_t width, height;
    svgf_plane colour;  /* denoised output, RGBA16F or RGBA32F (one format per call) */
    svgf_plane gbuffer; /* RGBA32F normal.xyz + linear depth (depth read from .w) */
    void* rgb;          /* device, 3 w h bytes, 4-byte aligned */
    void* depth;        /* device, w h floats, 16-byte aligned */
} svgf_wire_frame;

SVGF_API svgf_status svgf_wire_default_params(svgf_wire_params* out);
SVGF_API svgf_status svgf_wire_create(int32_t device, uint32_t max_frames, svgf_wire** out);
SVGF_API svgf_status svgf_wire_destroy(svgf_wire* wire);
/* Enqueues the encoding of n frames on hip_stream (asynchronous). */
SVGF_API svgf_status svgf_wire_encode(svgf_wire* wire, const svgf_wire_frame* frames, uint32_t n,
                                      const svgf_wire_params* params, void* hip_stream);

/* --- readback ring ----------------------------------------------------------------------------
 * The agent-side copy of wire frames (docs/AI_INTERFACE.md:385-400): `slots` pinned host buffers
 * of slot_bytes each. svgf_readback_push enqueues a device -> host copy into the next slot on
 * hip_stream and returns its sequence number (1, 2, ...); slot (seq mod slots) is overwritten
 * `slots` pushes later (push waits, on the host, for that slot's previous copy), so with 3 slots
 * a reader may hold frame n while frames n + 1 and n + 2 are produced. svgf_readback_view gives
 * the host address of a held sequence number: wait = 1 blocks until its copy is done, wait = 0
 * returns SVGF_ERR_NOT_READY while it is in flight; an overwritten or unknown sequence number is
 * SVGF_ERR_INVALID_ARGUMENT. To overlap a push with the next frame's kernels, push on a second
 * stream once its source is complete on the host's side (an event synchronised after the next
 * frame's denoise is enqueued): the HIP runtime then copies on the DMA engine, whereas a copy
 * ordered behind another stream's event by hipStreamWaitEvent is run as a copy kernel whose
 * writes across PCIe slow the memory-bound kernels beside it several times over. */
typedef struct svgf_readback svgf_readback;

SVGF_API svgf_status svgf_readback_create(int32_t device, uint64_t slot_bytes, uint32_t slots, svgf_readback** out);
SVGF_API svgf_status svgf_readback_destroy(svgf_readback* rb);
SVGF_API svgf_status svgf_readback_push(svgf_readback* rb, const void* device_src, uint64_t bytes, void* hip_stream,
                                        uint64_t* seq);
SVGF_API svgf_status svgf_readback_view(svgf_readback* rb, uint64_t seq, int32_t wait, const void** host,
                                        uint64_t* bytes);

/* --- consumer composite -----------------------------------------------------------------------
 * The reader of the denoised output: TronGrid's mesh pass shows the previous frame's denoised
 * lighting at the same pixel (src/mesh.slang:1007) and, for a reflection ray that hit a
 * non-emissive surface, at the pixel the hit position reprojects to through the previous
 * frame's view-projection (mesh.slang:1161-1174), then composites (mesh.slang:1191):
 *   clip = M (hit, 1); clip.w > 0 and px = int2((clip.xy / clip.w (0.5, -0.5) + 0.5) (W, H))
 *     inside the frame: hit_indirect = history[px].rgb, else hit_base * ambient;
 *   reflected = hit_emissive + hit_indirect;
 *   out = (history[pixel].rgb + emissive) + reflected * F_view, alpha 1.
 * int2() truncates toward zero (a product in (-1, 0) lands on column/row 0); a NaN or
 * out-of-range product is off screen. Arithmetic is float, one rounding per operation, matrix
 * rows summed left to right: ((m0 x + m1 y) + m2 z) + m3.
 * Per pixel the caller passes what the mesh pass holds besides the history: emissive, F_view and
 * the reflection, whose kind is reflection.w: 2 = reproject (reflection.xyz = hit position in
 * world space, reflection_colour = hit emissive, reflection_base = hit base colour); 1 = a
 * reflected colour that needs no lookup (sky miss, emissive hit; mesh.slang:1066-1131) in
 * reflection_colour; any other value = no reflection ray (reflected = 0; reflection_colour and
 * reflection_base are not read). One call composites many frames in one launch. */
typedef struct svgf_composite svgf_composite;

typedef struct svgf_composite_frame {
    uint32_t width, height;          /* render target = history size (camera.screen_width/height) */
    uint32_t reserved[2];
    float prev_view_projection[16];  /* row-major, element (r, c) at 4 r + c: clip = M (x, y, z, 1)
                                        (CameraData.prev_view_projection, mesh.slang:29) */
    svgf_plane history;              /* previous frame's denoised output: RGBA16F or RGBA32F */
    svgf_plane emissive;             /* rgb used: RGBA16F or RGBA32F */
    svgf_plane fresnel;              /* rgb used (F_view): RGBA16F or RGBA32F */
    svgf_plane reflection;           /* RGBA32F: hit position xyz, kind in w */
    svgf_plane reflection_colour;    /* rgb used: RGBA16F or RGBA32F */
    svgf_plane reflection_base;      /* rgb used: RGBA16F or RGBA32F */
    svgf_plane out;                  /* RGBA16F or RGBA32F; may not alias history */
} svgf_composite_frame;

typedef struct svgf_composite_params {
    float ambient; /* scale of the hit base colour when the lookup misses; AMBIENT = 0.03
                      (mesh.slang:276) */
    uint32_t reserved[3];
} svgf_composite_params;

SVGF_API svgf_status svgf_composite_default_params(svgf_composite_params* out);
SVGF_API svgf_status svgf_composite_create(int32_t device, uint32_t max_frames, svgf_composite** out);
SVGF_API svgf_status svgf_composite_destroy(svgf_composite* comp);
/* Enqueues the composite of n frames on hip_stream (asynchronous). Each frame's planes may have
 * their own formats and pitches; every plane stays below 2 GiB. */
SVGF_API svgf_status svgf_composite_run(svgf_composite* comp, const svgf_composite_frame* frames, uint32_t n,
                                        const svgf_composite_params* params, void* hip_stream);

/* --- producer pack --------------------------------------------------------------------------
 * Builds the denoiser's input planes from what TronGrid's mesh pass holds per pixel, writing
 * every pixel once (the per-frame zero-clear of src/main.cpp:595-646 folded into the pass):
 *   covered (position.w > 0):
 *     radiance = ((indirect + direct_diffuse) * ao + direct_specular, 1)  (mesh.slang:995-998)
 *     gbuffer  = (unpackOctahedralNormal(normal_oct), (VP (p, 1)).w)      (mesh.slang:341-352)
 *     motion   = previous position - pixel centre, in pixels, previous position from
 *                clip = PVP (p, 1): (clip.xy / max(clip.w, 1e-6) (0.5, -0.5) + 0.5) (W, H)
 *                (mesh.slang:589-591); (-1e6, -1e6) when clip.w <= 0 or |motion| > 1e6 px
 *   not covered (sky, transparent-only): radiance = 0, gbuffer = 0, motion = (-1e6, -1e6).
 * The depth (VP (p, 1)).w is the view-space distance along the camera axis of a perspective
 * projection. Normals are normalised with a correctly rounded square root and division.
 * Arithmetic is float, one rounding per operation, matrix rows summed left to right. */
typedef struct svgf_pack svgf_pack;

typedef struct svgf_pack_frame {
    uint32_t width, height;
    uint32_t reserved[2];
    float view_projection[16];      /* row-major, this frame (CameraData.view x projection) */
    float prev_view_projection[16]; /* row-major, previous frame (mesh.slang:29) */
    svgf_plane position;            /* RGBA32F: world position xyz, coverage in w */
    svgf_plane normal_oct;          /* R32UI: octahedral code (mesh.slang:327-339) */
    svgf_plane indirect;            /* rgb used: RGBA16F or RGBA32F */
    svgf_plane direct_diffuse;      /* rgb + ao_factor in w: RGBA16F or RGBA32F */
    svgf_plane direct_specular;     /* rgb used: RGBA16F or RGBA32F */
    svgf_plane radiance;            /* out: RGBA16F or RGBA32F (the denoiser's radiance plane) */
    svgf_plane gbuffer;             /* out: RGBA32F normal.xyz + linear depth */
    svgf_plane motion;              /* out: RG32F, pixels */
} svgf_pack_frame;

SVGF_API svgf_status svgf_pack_create(int32_t device, uint32_t max_frames, svgf_pack** out);
SVGF_API svgf_status svgf_pack_destroy(svgf_pack* pack);
/* Enqueues the pack of n frames on hip_stream (asynchronous). Outputs may not overlap inputs. */
SVGF_API svgf_status svgf_pack_run(svgf_pack* pack, const svgf_pack_frame* frames, uint32_t n, void* hip_stream);

/* --- frame profiler ---------------------------------------------------------------------------
 * TronGrid's per-pass GPU profiler (src/main.cpp:145-180, 1180-1207, 2902-2953) on hipEvents:
 * every pass of a frame is bracketed by a start and an end event recorded into the frame's slot
 * (`slots` frames in flight, MAX_FRAMES_IN_FLIGHT = 2); when a slot comes round again its
 * events are read back (waiting for them, as the reference waits for the slot's fence) and each
 * pass's duration updates an exponential moving average,
 *   first readback: ema = sample;  later: ema = ema * (1 - alpha) + sample * alpha   (float)
 * and when the frame counter at that readback is a multiple of log_interval the line
 *   "GPU times (ms): <name>=<ema / 1e6 as %f> ..." is composed. The frame counter starts at
 * first_frame (WARMUP_ITERATIONS = 16) and advances once per svgf_profiler_begin_frame.
 * Per frame: svgf_profiler_begin_frame, then pass_begin / pass_end pairs around the work, then
 * svgf_profiler_end_frame. A pass without both events in a frame samples 0.
 * device = -1 makes a host-only profiler for replaying recorded durations: samples are given
 * with svgf_profiler_sample instead of events. */
typedef struct svgf_profiler svgf_profiler;

typedef struct svgf_profiler_desc {
    uint32_t passes;          /* 1..32 */
    uint32_t slots;           /* frames in flight, 1..8; reference 2 (main.cpp:74) */
    float ema_alpha;          /* reference GPU_PROFILER_EMA_ALPHA = 0.05 (main.cpp:170) */
    uint32_t log_interval;    /* frames; reference GPU_PROFILER_LOG_INTERVAL_FRAMES = 60 (main.cpp:175) */
    uint32_t first_frame;     /* initial frame counter; reference WARMUP_ITERATIONS = 16 (main.cpp:2611) */
    int32_t device;           /* HIP device ordinal, or -1 for host samples only */
    const char* const* names; /* `passes` names (copied, at most 31 characters each); NULL: pass0, pass1, ... */
} svgf_profiler_desc;

SVGF_API svgf_status svgf_profiler_default_desc(svgf_profiler_desc* out);
SVGF_API svgf_status svgf_profiler_create(const svgf_profiler_desc* desc, svgf_profiler** out);
SVGF_API svgf_status svgf_profiler_destroy(svgf_profiler* prof);
/* Reads back the slot this frame reuses (if it holds a frame), updates the averages and composes
 * the log line when due (*logged = 1, else 0; logged may be NULL); then advances the counter. */
SVGF_API svgf_status svgf_profiler_begin_frame(svgf_profiler* prof, int32_t* logged);
SVGF_API svgf_status svgf_profiler_pass_begin(svgf_profiler* prof, uint32_t pass, void* hip_stream);
SVGF_API svgf_status svgf_profiler_pass_end(svgf_profiler* prof, uint32_t pass, void* hip_stream);
/* Host-only profiler: the current frame's duration of `pass`, in nanoseconds. */
SVGF_API svgf_status svgf_profiler_sample(svgf_profiler* prof, uint32_t pass, float ns);
SVGF_API svgf_status svgf_profiler_end_frame(svgf_profiler* prof);
/* Averages in nanoseconds (count = passes; SVGF_ERR_NOT_READY before the first readback). */
SVGF_API svgf_status svgf_profiler_ema(svgf_profiler* prof, float* ema_ns, uint32_t capacity, uint32_t* count);
/* The last composed log line (empty before the first), NUL-terminated, truncated to capacity. */
SVGF_API svgf_status svgf_profiler_log_line(svgf_profiler* prof, char* buf, uint32_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* SVGF_H */
