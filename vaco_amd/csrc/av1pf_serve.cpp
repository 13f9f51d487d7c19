// av1pf_serve.cpp - the serving loop: admission by HBM budget, continuous
// batching over many streams, per-request completions with timestamps.
//
// One worker thread owns the context. Submitters hand requests over through
// a mutex-protected queue; the worker admits them FIFO while the device
// bytes in flight stay within the budget, launches batches through the
// batched path (av1pf_enqueue_batch), retires every batch whose event has
// completed, settles each frame's status and publishes completions. Frames
// in one batch run the same kernels as a frame alone, so results do not
// depend on what they were batched with.
//
// Each batch in flight has a stream and an intermediate-image workspace of
// its own (max_inflight of each), so batches overlap on the device when the
// runtime gives each stream a hardware queue (its pool holds
// GPU_MAX_HW_QUEUES queues per process; streams beyond that share them and
// then run one after the other). The
// workspaces and the context's staging slots are sized at server_create;
// the worker closes a batch early rather than let it outgrow either, so no
// device or pinned memory is allocated after creation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/vaco_av1pf.h"
#include "av1pf_device.h"
#include "av1pf_host.h"
#include "av1pf_internal.h"

using namespace av1pf;

extern "C" uint64_t vaco_av1pf_now_ns(void)
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

namespace {

struct Request {
    vaco_av1pf_frame_desc desc;
    vaco_av1pf_picture in, out;
    uint64_t tag = 0, t_submit = 0, t_launch = 0;
    size_t bytes = 0;    // device bytes in flight (admission)
    size_t stage = 0;    // staging-slot bytes
    size_t scratch = 0;  // workspace samples
    bool replay = false;  // parameters can fail: settle with the reference's checks first
    uint64_t gpu_start = 0, gpu_end = 0;  // device clock of its batch
};

struct Batch {
    std::vector<Request> reqs;
    const uint32_t *status = nullptr;
    hipEvent_t done = nullptr;
    int stream = -1;  // index into the server's streams and workspaces
    int slot = -1;    // the context's staging slot holding the status words
    uint64_t gpu_start = 0, gpu_end = 0;  // device clock (publish)
};

constexpr int kMaxStreams = 4;

}  // namespace

struct vaco_av1pf_server {
    vaco_av1pf_ctx *ctx = nullptr;
    vaco_av1pf_server_config cfg{};
    hipStream_t streams[kMaxStreams] = {};
    // Device-clock markers: t0 when the server started, and a start / end
    // pair per stream around the batch in flight on it.
    hipEvent_t t0 = nullptr;
    hipEvent_t mark[kMaxStreams][2] = {};
    uint16_t *ws_mem = nullptr;  // max_inflight workspaces of ws_cap samples
    size_t ws_cap = 0;
    size_t stage_cap = 0;  // staging bytes of one batch
    bool stream_busy[kMaxStreams] = {};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> incoming;
    std::vector<vaco_av1pf_completion> completions;
    std::atomic<uint64_t> submitted{0}, completed{0};
    std::atomic<bool> stop{false};
    std::thread worker;
    // worker-owned
    std::deque<Request> waiting;
    std::vector<Batch> inflight;
    size_t bytes_in_flight = 0;
};

static void publish(vaco_av1pf_server *s, const Request &r, int32_t status, uint32_t batch,
                    uint64_t t_done)
{
    vaco_av1pf_completion c;
    c.tag = r.tag;
    c.status = status;
    c.batch_size = batch;
    c.t_submit_ns = r.t_submit;
    c.t_launch_ns = r.t_launch;
    c.t_done_ns = t_done;
    c.t_gpu_start_ns = r.gpu_start;
    c.t_gpu_end_ns = r.gpu_end;
    std::lock_guard<std::mutex> lk(s->mu);
    s->completions.push_back(c);
    s->completed.fetch_add(1);
    s->cv.notify_all();
}

// Publish a finished batch. `device_ok` false: the device reported an error
// for the batch's event, and every frame of it fails with DEVICE_ERROR.
// The marker's device time, ns since t0 (0 if it cannot be read).
static uint64_t since_t0_ns(vaco_av1pf_server *s, hipEvent_t e)
{
    float ms = 0.f;
    if (hipEventSynchronize(e) != hipSuccess || hipEventElapsedTime(&ms, s->t0, e) != hipSuccess || ms <= 0.f)
        return 0;
    return (uint64_t)((double)ms * 1e6);
}

static void retire(vaco_av1pf_server *s, Batch &b, bool device_ok)
{
    const uint64_t t = vaco_av1pf_now_ns();
    if (device_ok) {
        const uint64_t g0 = since_t0_ns(s, s->mark[b.stream][0]), g1 = since_t0_ns(s, s->mark[b.stream][1]);
        for (auto &r : b.reqs) {
            r.gpu_start = g0;
            r.gpu_end = g1;
        }
    }
    for (size_t i = 0; i < b.reqs.size(); i++) {
        int32_t st = device_ok ? VACO_AV1PF_OK : VACO_AV1PF_DEVICE_ERROR;
        if (device_ok && (b.status[i] & kFlagSampleRange)) {
            const char *msg = "";
            st = av1pf_replay_device_frame(s->ctx, b.reqs[i].desc, b.reqs[i].in, &msg);
        }
        s->bytes_in_flight -= b.reqs[i].bytes;
        publish(s, b.reqs[i], st, (uint32_t)b.reqs.size(), t);
    }
    // The batch's event has completed: the context must not wait on it (or
    // its stream) again, since the stream goes away with the server.
    av1pf_settle_slot(s->ctx, b.slot);
    s->stream_busy[b.stream] = false;
}

// Retire every batch that has completed; with `wait_slot` >= 0, first wait
// for the batch whose status words are in that staging slot (the next call
// takes it and clears them).
static void retire_done(vaco_av1pf_server *s, int wait_slot = -1)
{
    for (size_t i = 0; i < s->inflight.size();) {
        Batch &b = s->inflight[i];
        hipError_t q = b.slot == wait_slot ? hipEventSynchronize(b.done) : hipEventQuery(b.done);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        retire(s, b, q == hipSuccess);
        s->inflight.erase(s->inflight.begin() + (long)i);
    }
}

static void launch(vaco_av1pf_server *s)
{
    Batch b;
    for (int i = 0; i < (int)s->cfg.max_inflight; i++)
        if (!s->stream_busy[i]) {
            b.stream = i;
            break;
        }
    if (b.stream < 0)
        return;
    retire_done(s, av1pf_next_slot(s->ctx));
    size_t budget = s->bytes_in_flight, stage = kBatchStageBase, scratch = 0;
    while (!s->waiting.empty() && b.reqs.size() < s->cfg.max_batch) {
        Request &r = s->waiting.front();
        if (r.replay) {  // parameter errors are reported without running
            const char *msg = "";
            const vaco_av1pf_status st = av1pf_replay_device_frame(s->ctx, r.desc, r.in, &msg);
            if (st != VACO_AV1PF_OK) {
                r.t_launch = vaco_av1pf_now_ns();
                publish(s, r, st, 0, r.t_launch);
                s->waiting.pop_front();
                continue;
            }
            r.replay = false;
        }
        if (budget + r.bytes > s->cfg.hbm_budget_bytes)
            break;  // FIFO: wait for memory to free up
        if (!b.reqs.empty() && (stage + r.stage > s->stage_cap || scratch + r.scratch > s->ws_cap))
            break;  // the rest goes into the next batch
        budget += r.bytes;
        stage += r.stage;
        scratch += r.scratch;
        b.reqs.push_back(r);
        s->waiting.pop_front();
    }
    if (b.reqs.empty())
        return;
    const size_t n = b.reqs.size();
    std::vector<vaco_av1pf_frame_desc> descs(n);
    std::vector<vaco_av1pf_picture> ins(n), outs(n);
    const uint64_t t = vaco_av1pf_now_ns();
    for (size_t i = 0; i < n; i++) {
        descs[i] = b.reqs[i].desc;
        ins[i] = b.reqs[i].in;
        outs[i] = b.reqs[i].out;
        b.reqs[i].t_launch = t;
    }
    const av1pf_workspace ws{s->ws_mem + (size_t)b.stream * s->ws_cap, s->ws_cap};
    hipStream_t q = s->streams[b.stream];
    vaco_av1pf_status st = hipEventRecord(s->mark[b.stream][0], q) == hipSuccess ? VACO_AV1PF_OK
                                                                                : VACO_AV1PF_DEVICE_ERROR;
    if (st == VACO_AV1PF_OK)
        st = av1pf_enqueue_batch(s->ctx, (uint32_t)n, descs.data(), ins.data(), outs.data(), q, &b.status,
                                 &b.done, &ws, &b.slot);
    if (st == VACO_AV1PF_OK && hipEventRecord(s->mark[b.stream][1], q) != hipSuccess)
        st = VACO_AV1PF_DEVICE_ERROR;
    if (st != VACO_AV1PF_OK) {
        for (auto &r : b.reqs)
            publish(s, r, st, (uint32_t)n, vaco_av1pf_now_ns());
        return;
    }
    s->bytes_in_flight = budget;
    s->stream_busy[b.stream] = true;
    s->inflight.push_back(std::move(b));
}

static void work(vaco_av1pf_server *s)
{
    (void)hipSetDevice(av1pf_ctx_device(s->ctx));  // a failure shows in the calls after it
    const uint64_t window = (uint64_t)s->cfg.window_us * 1000;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->incoming.empty() && s->waiting.empty() && s->inflight.empty()) {
                if (s->stop.load())
                    return;
                s->cv.wait_for(lk, std::chrono::microseconds(200));
            }
            while (!s->incoming.empty()) {
                s->waiting.push_back(std::move(s->incoming.front()));
                s->incoming.pop_front();
            }
        }
        // Retire finished batches, in whatever order they finish.
        retire_done(s);
        // Launch: continuous batching with a window while the GPU is busy.
        if (!s->waiting.empty() && s->inflight.size() < s->cfg.max_inflight) {
            const uint64_t now = vaco_av1pf_now_ns();
            const bool go = s->inflight.empty() || s->waiting.size() >= s->cfg.max_batch ||
                            now - s->waiting.front().t_submit >= window;
            if (go)
                launch(s);
        }
        if (!s->inflight.empty())
            std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" vaco_av1pf_status vaco_av1pf_server_create(vaco_av1pf_ctx *ctx,
                                                      const vaco_av1pf_server_config *cfg,
                                                      vaco_av1pf_server **out)
{
    if (!ctx || !cfg || !out || cfg->max_batch == 0 || cfg->max_batch > 256 ||
        cfg->max_inflight == 0 || cfg->max_inflight > 4 || cfg->hbm_budget_bytes == 0)
        return VACO_AV1PF_INVALID_ARGUMENT;
    vaco_av1pf_server *s = new (std::nothrow) vaco_av1pf_server();
    if (!s)
        return VACO_AV1PF_DEVICE_ERROR;
    s->ctx = ctx;
    s->cfg = *cfg;
    // Every buffer the worker uses is sized here. A workspace holds the
    // intermediates of the largest frame the context accepts and otherwise
    // up to 2 GiB (or the budget's share of one batch); a staging slot the
    // largest frame's side information and otherwise up to 16 MiB.
    const size_t worst_ws = av1pf_worst_frame_scratch_samples(ctx);
    const size_t share = std::min<uint64_t>(2ull << 30, cfg->hbm_budget_bytes / cfg->max_inflight) / 2;
    s->ws_cap = (std::max(worst_ws, share) + 127) & ~size_t(127);
    const size_t worst_stage = kBatchStageBase + av1pf_worst_frame_stage_bytes(ctx);
    const size_t want_stage =
        std::max(worst_stage, std::min<size_t>(16u << 20, kBatchStageBase + cfg->max_batch *
                                                           av1pf_worst_frame_stage_bytes(ctx)));
    vaco_av1pf_status st = VACO_AV1PF_OK;
    if (hipSetDevice(av1pf_ctx_device(ctx)) != hipSuccess)
        st = VACO_AV1PF_DEVICE_ERROR;
    if (st == VACO_AV1PF_OK)
        st = av1pf_reserve_staging(ctx, want_stage);
    s->stage_cap = av1pf_staging_capacity(ctx);
    for (uint32_t i = 0; st == VACO_AV1PF_OK && i < cfg->max_inflight; i++)
        if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s->mark[i][0]) != hipSuccess || hipEventCreate(&s->mark[i][1]) != hipSuccess)
            st = VACO_AV1PF_DEVICE_ERROR;
    if (st == VACO_AV1PF_OK && (hipEventCreate(&s->t0) != hipSuccess ||
                                hipEventRecord(s->t0, s->streams[0]) != hipSuccess ||
                                hipEventSynchronize(s->t0) != hipSuccess))
        st = VACO_AV1PF_DEVICE_ERROR;
    if (st == VACO_AV1PF_OK &&
        hipMalloc((void **)&s->ws_mem, s->ws_cap * cfg->max_inflight * sizeof(uint16_t)) != hipSuccess)
        st = VACO_AV1PF_DEVICE_ERROR;
    if (st != VACO_AV1PF_OK) {
        for (int i = 0; i < kMaxStreams; i++) {
            if (s->streams[i])
                (void)hipStreamDestroy(s->streams[i]);
            for (auto &e : s->mark[i])
                if (e)
                    (void)hipEventDestroy(e);
        }
        if (s->t0)
            (void)hipEventDestroy(s->t0);
        if (s->ws_mem)
            (void)hipFree(s->ws_mem);
        delete s;
        return st;
    }
    s->worker = std::thread(work, s);
    *out = s;
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_server_submit(vaco_av1pf_server *s,
                                                      const vaco_av1pf_frame_desc *desc,
                                                      const vaco_av1pf_picture *in,
                                                      vaco_av1pf_picture *out, uint64_t tag)
{
    if (!s || !desc || !in || !out)
        return VACO_AV1PF_INVALID_ARGUMENT;
    Request r;
    r.desc = *desc;
    r.in = *in;
    r.out = *out;
    r.tag = tag;
    r.bytes = av1pf_frame_device_bytes(*desc);
    r.stage = av1pf_frame_stage_bytes(*desc);
    r.scratch = av1pf_frame_scratch_samples(*desc);
    if (r.bytes > s->cfg.hbm_budget_bytes || kBatchStageBase + r.stage > s->stage_cap ||
        r.scratch > s->ws_cap)
        return VACO_AV1PF_LIMIT_EXCEEDED;
    FrameGeo g;
    frame_geometry(*desc, &g);
    for (int p = 0; p < g.nplanes; p++)
        if (!in->plane[p].data || !out->plane[p].data || in->plane[p].width != g.in_w[p] ||
            in->plane[p].height != g.in_h[p] || out->plane[p].width != g.out_w[p] ||
            out->plane[p].height != g.out_h[p] || in->plane[p].stride < g.in_w[p] ||
            out->plane[p].stride < g.out_w[p])
            return VACO_AV1PF_INVALID_ARGUMENT;
    if (g.cdef && (!desc->cdef_idx || !desc->skip8 || desc->cdef_stride < (g.mi_cols + 15) / 16))
        return VACO_AV1PF_INVALID_ARGUMENT;
    r.replay = !params_clean(*desc, g);
    r.t_submit = vaco_av1pf_now_ns();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->incoming.push_back(r);
        s->submitted.fetch_add(1);
    }
    s->cv.notify_all();
    return VACO_AV1PF_OK;
}

extern "C" uint32_t vaco_av1pf_server_poll(vaco_av1pf_server *s, vaco_av1pf_completion *out,
                                           uint32_t max)
{
    if (!s || !out || max == 0)
        return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    const uint32_t n = (uint32_t)std::min<size_t>(max, s->completions.size());
    std::copy(s->completions.begin(), s->completions.begin() + n, out);
    s->completions.erase(s->completions.begin(), s->completions.begin() + n);
    return n;
}

extern "C" uint64_t vaco_av1pf_server_frame_bytes(const vaco_av1pf_frame_desc *desc)
{
    return desc ? (uint64_t)av1pf_frame_device_bytes(*desc) : 0;
}

extern "C" vaco_av1pf_status vaco_av1pf_server_drain(vaco_av1pf_server *s)
{
    if (!s)
        return VACO_AV1PF_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->completed.load() >= s->submitted.load(); });
    return VACO_AV1PF_OK;
}

extern "C" void vaco_av1pf_server_destroy(vaco_av1pf_server *s)
{
    if (!s)
        return;
    vaco_av1pf_server_drain(s);
    s->stop.store(true);
    s->cv.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    for (int i = 0; i < kMaxStreams; i++) {
        if (s->streams[i]) {
            (void)hipStreamSynchronize(s->streams[i]);
            (void)hipStreamDestroy(s->streams[i]);
        }
        for (auto &e : s->mark[i])
            if (e)
                (void)hipEventDestroy(e);
    }
    if (s->t0)
        (void)hipEventDestroy(s->t0);
    if (s->ws_mem)
        (void)hipFree(s->ws_mem);
    delete s;
}
