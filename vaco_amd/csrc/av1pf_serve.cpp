// av1pf_serve.cpp - the serving loop: admission by HBM budget, continuous
// batching over many streams, per-request completions with timestamps.
//
// One worker thread owns the context. Submitters hand requests over through
// a mutex-protected queue; the worker admits them FIFO while the device
// bytes in flight stay within the budget, launches batches through the
// batched path (av1pf_enqueue_batch) on the context's stream, retires
// batches whose event has completed, settles each frame's status and
// publishes completions. Frames in one batch run the same kernels as a
// frame alone, so results do not depend on what they were batched with.
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
    size_t bytes = 0;
    bool replay = false;  // parameters can fail: settle with the reference's checks first
};

struct Batch {
    std::vector<Request> reqs;
    const uint32_t *status = nullptr;
    hipEvent_t done = nullptr;
};

}  // namespace

struct vaco_av1pf_server {
    vaco_av1pf_ctx *ctx = nullptr;
    vaco_av1pf_server_config cfg{};
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> incoming;
    std::vector<vaco_av1pf_completion> completions;
    std::atomic<uint64_t> submitted{0}, completed{0};
    std::atomic<bool> stop{false};
    std::thread worker;
    // worker-owned
    std::deque<Request> waiting;
    std::deque<Batch> inflight;
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
    std::lock_guard<std::mutex> lk(s->mu);
    s->completions.push_back(c);
    s->completed.fetch_add(1);
    s->cv.notify_all();
}

// Publish a finished batch. `device_ok` false: the device reported an error
// for the batch's event, and every frame of it fails with DEVICE_ERROR.
static void retire(vaco_av1pf_server *s, Batch &b, bool device_ok)
{
    const uint64_t t = vaco_av1pf_now_ns();
    for (size_t i = 0; i < b.reqs.size(); i++) {
        int32_t st = device_ok ? VACO_AV1PF_OK : VACO_AV1PF_DEVICE_ERROR;
        if (device_ok && (b.status[i] & kFlagSampleRange)) {
            const char *msg = "";
            st = av1pf_replay_device_frame(s->ctx, b.reqs[i].desc, b.reqs[i].in, &msg);
        }
        s->bytes_in_flight -= b.reqs[i].bytes;
        publish(s, b.reqs[i], st, (uint32_t)b.reqs.size(), t);
    }
}

static void launch(vaco_av1pf_server *s)
{
    Batch b;
    size_t budget = s->bytes_in_flight;
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
        budget += r.bytes;
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
    const vaco_av1pf_status st = av1pf_enqueue_batch(s->ctx, (uint32_t)n, descs.data(), ins.data(),
                                                     outs.data(), s->stream, &b.status, &b.done);
    if (st != VACO_AV1PF_OK) {
        for (auto &r : b.reqs)
            publish(s, r, st, (uint32_t)n, vaco_av1pf_now_ns());
        return;
    }
    s->bytes_in_flight = budget;
    s->inflight.push_back(std::move(b));
}

static void work(vaco_av1pf_server *s)
{
    hipSetDevice(av1pf_ctx_device(s->ctx));
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
        // Retire finished batches (in launch order).
        while (!s->inflight.empty()) {
            const hipError_t q = hipEventQuery(s->inflight.front().done);
            if (q == hipErrorNotReady)
                break;
            retire(s, s->inflight.front(), q == hipSuccess);
            s->inflight.pop_front();
        }
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
    s->stream = nullptr;  // the context's stream
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
    if (r.bytes > s->cfg.hbm_budget_bytes)
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
    delete s;
}
