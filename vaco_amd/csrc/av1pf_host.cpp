// av1pf_host.cpp - host runtime and C ABI (include/vaco_av1pf.h) of the
// MI355X AV1 post-reconstruction filter stage.
//
// The per-frame entry replaces decode.rs:554-559 of matteopolak/vaco
// (crates/codec/vaco-codec-av1/src/): pre-CDEF copy, apply_cdef,
// apply_superres, apply_restoration. Device buffers are sized once at
// vaco_av1pf_create; a frame call only fills a pinned staging arena with
// work descriptors and side information, uploads it with one copy and
// launches at most three kernels (CDEF, superres, restoration).
//
// Errors: the fast path validates parameters on the host and lets the
// kernels flag any sample above the bit depth. Whenever a parameter is out
// of range, fuel is limited, or a kernel raised the flag, the host replays
// the reference's checks in the reference's order (av1pf_replay.cpp) so the
// status and message are exactly the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <new>
#include <vector>

#include "../../include/vaco_av1pf.h"
#include "av1pf_device.h"
#include "av1pf_host.h"
#include "av1pf_internal.h"

namespace av1pf {
hipError_t launch_cdef(const CdefFrame *frames, int n, const CdefFrame *single, uint32_t blocks, int mode,
                       bool xcd, int waves, const void *up_src, void *up_dst, size_t up_bytes, hipStream_t s,
                       hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_superres(const SrJob *jobs, int n, uint32_t blocks, int rows, hipStream_t s,
                           hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_chain(const CdefFrame &f, int mode, bool xcd, const LrPlane *planes, uint32_t n_items,
                        const uint64_t *list, uint32_t n_tiles, bool lr_xcd, const void *up_src, void *up_dst,
                        size_t up_bytes,
                        const ChainSync &cs, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);
uint32_t chain_copiers(uint32_t n_sb, size_t up_bytes);
hipError_t launch_lr(const LrPlane *planes, uint32_t n_items, const uint64_t *list, uint32_t n, bool xcd, int rows,
                     hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_upload(void *dst, const void *src, size_t bytes, hipStream_t s,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_to_bytes(const uint16_t *src, uint32_t src_stride, uint8_t *dst,
                           uint32_t dst_stride, uint32_t w, uint32_t h, hipStream_t s);
hipError_t launch_cdef_cases(const uint16_t *inputs, const void *params, uint8_t *dirs,
                             uint32_t *vars, uint16_t *outs, uint32_t n, hipStream_t s);
}  // namespace av1pf

using namespace av1pf;

// ------------------------------------------------------------ staging --

namespace {

constexpr int kSlots = 8;
constexpr uint32_t kMaxBatch = 256;  // frames per batched call
constexpr uint32_t kMaxSuperresWidth = 65536;  // k_superres's 32-bit source positions
// Launch variants of a new context (vaco_av1pf_autotune may change them).
constexpr vaco_av1pf_tuning kDefaultTuning = {8, 1, 0, 64, 16, 0};
constexpr const char *kChainStallMsg =
    "a restoration tile of a chained launch gave up waiting for its CDEF superblocks";
// Polls (each ~0.1-0.3 us) a chained launch's restoration tile makes before
// it gives up on its superblocks: about a second, far beyond any frame.
constexpr uint32_t kChainSpinLimit = 1u << 22;
// Staging bytes of the static-scratch entry (the rest of its scratch holds
// intermediate images).
constexpr size_t kStaticStageBytes = 2u << 20;

// One in-flight frame: its staging area (pinned host + device mirror), its
// status word (coherent pinned host memory, raised by the kernels), and
// what the deferred error replay needs (the caller keeps desc's arrays and
// the input planes valid until vaco_av1pf_finish()).
struct Arena {
    uint8_t *h = nullptr;   // pinned, coherent host memory
    uint8_t *hd = nullptr;  // the same memory as the device addresses it
    uint8_t *d = nullptr;   // device mirror
    size_t cap = 0, used = 0;
    // The frames that used this slot have finished: `done` without the
    // system-scope fence (enough to reuse the slot and read its status words,
    // which are coherent host memory written by system-scope atomics),
    // `done_sys` with it (a completion the host hands on, which must make the
    // outputs visible to any agent). `done_ev` is the one last recorded.
    hipEvent_t done = nullptr, done_sys = nullptr, done_ev = nullptr;
    hipStream_t stream = nullptr;  // the stream done_ev was recorded on
    bool pending = false;
    // h is ordinary (pageable) host memory: the upload is a runtime copy, not
    // k_upload's direct read of coherent pinned memory
    bool pageable = false;
    uint32_t *status = nullptr;  // kMaxBatch words, one per frame
    // frames whose status must be checked before reuse
    std::vector<vaco_av1pf_frame_desc> descs;
    std::vector<vaco_av1pf_picture> ins;
    // Instantiated kernel chains of calls that used this slot, keyed by
    // their launch parameters (vaco_av1pf_use_graphs). Only this slot's calls
    // launch them, so an entry is idle whenever the slot is taken.
    struct Graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;

    size_t tail = 0;  // bytes taken from the end of the slot (take_host)

    template <class T>
    T *take(size_t n, T **dev)
    {
        size_t a = (used + 255) & ~size_t(255);
        size_t bytes = n * sizeof(T);
        if (a + bytes + tail > cap)
            return nullptr;
        used = a + bytes;
        *dev = reinterpret_cast<T *>(d + a);
        return reinterpret_cast<T *>(h + a);
    }

    // Bytes that only the CDEF kernel reads (cdef_idx, skip flags): taken
    // from the end of the slot and never uploaded; *dev is their address in
    // the slot's coherent host memory as the device sees it, which the CDEF
    // kernel reads in place (a few hundred bytes per superblock, issued with
    // its sample loads). A slot of ordinary host memory has no such view:
    // there they are uploaded like the rest.
    template <class T>
    T *take_host(size_t n, T **dev)
    {
        if (pageable || !hd)
            return take(n, dev);
        const size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
        if (((used + 255) & ~size_t(255)) + tail + bytes > cap)
            return nullptr;
        tail += bytes;
        *dev = reinterpret_cast<T *>(hd + cap - tail);
        return reinterpret_cast<T *>(h + cap - tail);
    }
};

}  // namespace

namespace {

// A buffer that calls of a context share (a set of per-frame intermediates,
// the batch workspace, the chained launch's counters): the stream of its last
// user and the event that ends that use. A call on another stream waits for
// that event on the device before using the buffer; calls on one stream are
// ordered by the stream.
struct Use {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

constexpr int kSets = 2;                 // per-frame intermediate sets
constexpr int kSetPlanes = 6;            // CDEF image + upscaled CDEF image, 3 planes each
constexpr int kPoolPlanes = 6 + kSets * kSetPlanes;

}  // namespace

struct vaco_av1pf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t max_w = 0, max_h = 0;
    size_t plane_cap = 0;          // samples per plane buffer
    uint16_t *pool = nullptr;      // kPoolPlanes planes
    uint16_t *img_in[3] = {}, *img_out[3] = {};  // the host entry's input and output
    // Intermediates of per-frame calls (CDEF image, upscaled CDEF image):
    // two sets, so that calls on two streams overlap on the device.
    uint16_t *set_img[kSets] = {};
    Use set_use[kSets];
    int next_set = 0;
    Arena slot[kSlots];
    int next_slot = 0;
    uint32_t *h_status = nullptr;  // coherent pinned status words, one per slot
    const char *err = "";
    // first error found by a deferred check (device entry), reported by finish
    vaco_av1pf_status deferred = VACO_AV1PF_OK;
    const char *deferred_msg = "";
    // device intermediates of batched calls (grown on demand)
    uint16_t *ws = nullptr;
    size_t ws_cap = 0;  // samples
    // The batch workspace's and the chained launches' last use (a chained
    // launch's counters assume the previous one has completed).
    Use ws_use, chain_use;
    // Chained launches (tuning.chain): per-superblock completion flags and
    // the staged-bytes copy counter (device memory, zeroed at creation,
    // never reset); the counter's value once every chained launch so far has
    // completed (calls on a context are ordered on the device), and the
    // epoch of the last launch.
    uint32_t *chain_flags = nullptr;
    unsigned long long *chain_copied = nullptr;
    size_t chain_sbs = 0;
    uint64_t chain_copied_total = 0;
    uint32_t chain_epoch = 0;
    uint32_t chain_spin_limit = kChainSpinLimit;
    // kernel timing (vaco_av1pf_profile)
    struct ProfRec {
        int kernel;
        uint32_t a, b;
        uint64_t bytes;
    };
    vaco_av1pf_tuning tune = kDefaultTuning;  // launch variants (vaco_av1pf_autotune)
    bool use_graphs = false;                // replay kernel chains as hipGraphs
    bool prof = false;
    std::vector<hipEvent_t> prof_pool;
    uint32_t prof_next = 0;
    std::vector<ProfRec> prof_recs;
    double prof_ms[VACO_AV1PF_KERNEL_COUNT] = {};
    uint64_t prof_n[VACO_AV1PF_KERNEL_COUNT] = {};
    uint64_t prof_bytes[VACO_AV1PF_KERNEL_COUNT] = {};
};

static hipError_t profile_collect(vaco_av1pf_ctx *ctx);
template <class F>
static hipError_t timed_launch(vaco_av1pf_ctx *ctx, int kernel, uint64_t bytes, hipStream_t s,
                               F launch);

static vaco_av1pf_status fail(vaco_av1pf_ctx *ctx, vaco_av1pf_status st, const char *msg)
{
    if (ctx)
        ctx->err = msg;
    return st;
}

static vaco_av1pf_status hip_fail(vaco_av1pf_ctx *ctx, hipError_t e)
{
    return fail(ctx, VACO_AV1PF_DEVICE_ERROR, hipGetErrorString(e));
}

#define HIPCHECK(expr)                       \
    do {                                     \
        hipError_t e_ = (expr);              \
        if (e_ != hipSuccess)                \
            return hip_fail(ctx, e_);        \
    } while (0)

static vaco_av1pf_status resolve_slot(vaco_av1pf_ctx *ctx, Arena &a);
static size_t arena_need(const vaco_av1pf_frame_desc &d, const FrameGeo &g);

// (Re)allocate a slot's staging buffers at `bytes` (a multiple of 256). The
// host side is coherent pinned memory that k_upload reads directly.
static hipError_t alloc_slot(Arena &a, size_t bytes)
{
    if (a.h)
        (void)hipHostFree(a.h);
    if (a.d)
        (void)hipFree(a.d);
    a.h = a.hd = a.d = nullptr;
    a.cap = 0;
    a.used = a.tail = 0;
    hipError_t e = hipHostMalloc((void **)&a.h, bytes, hipHostMallocCoherent);
    if (e == hipSuccess)
        e = hipHostGetDevicePointer((void **)&a.hd, a.h, 0);
    if (e == hipSuccess)
        e = hipMalloc((void **)&a.d, bytes);
    if (e == hipSuccess)
        a.cap = bytes;
    return e;
}

// Next staging slot. Waits for the frame that last used it and settles that
// frame's deferred status check first.
static vaco_av1pf_status take_slot(vaco_av1pf_ctx *ctx, Arena **out)
{
    Arena &a = ctx->slot[ctx->next_slot];
    ctx->next_slot = (ctx->next_slot + 1) % kSlots;
    if (a.pending) {
        HIPCHECK(hipEventSynchronize(a.done_ev));
        a.pending = false;
    }
    vaco_av1pf_status st = resolve_slot(ctx, a);
    if (st != VACO_AV1PF_OK)
        return st;
    a.used = 0;
    a.tail = 0;
    memset(a.status, 0, kMaxBatch * sizeof(uint32_t));
    *out = &a;
    return VACO_AV1PF_OK;
}

// Grow a slot's staging buffers (its previous frames have finished). Every
// other slot whose work has finished grows with it, so a workload settles
// after one growth instead of one per slot.
static vaco_av1pf_status grow_slot(vaco_av1pf_ctx *ctx, Arena *a, size_t bytes)
{
    if (bytes <= a->cap)
        return VACO_AV1PF_OK;
    bytes = (std::max(bytes, a->cap * 2) + 255) & ~size_t(255);
    HIPCHECK(alloc_slot(*a, bytes));
    for (int i = 0; i < kSlots; i++) {
        Arena &b = ctx->slot[i];
        if (&b == a || b.cap >= bytes)
            continue;
        if (b.pending) {
            const hipError_t q = hipEventQuery(b.done_ev);
            if (q == hipErrorNotReady)
                continue;
            HIPCHECK(q);
        }
        HIPCHECK(alloc_slot(b, bytes));
    }
    return VACO_AV1PF_OK;
}

template <class F>
static hipError_t timed_launch(vaco_av1pf_ctx *ctx, int kernel, uint64_t bytes, hipStream_t s,
                               F launch);

// Copy the slot's staged bytes (descriptors; the side information the
// kernels read in place stays in host memory) to its device mirror, on the
// call's stream s ahead of its kernels: tens of microseconds at most for a
// large batch, against a cross-stream event hand-off, which cost as much and
// needed a hardware queue of its own. The slot's previous frame has finished
// (take_slot), so its mirror is free.
static vaco_av1pf_status upload_slot(vaco_av1pf_ctx *ctx, Arena *a, hipStream_t s)
{
    if (a->used == 0)
        return VACO_AV1PF_OK;
    if (a->pageable) {
        HIPCHECK(hipMemcpyAsync(a->d, a->h, a->used, hipMemcpyHostToDevice, s));
        return VACO_AV1PF_OK;
    }
    HIPCHECK(timed_launch(ctx, VACO_AV1PF_KERNEL_UPLOAD, (uint64_t)a->used * 2, s,
                          [&](hipEvent_t e0, hipEvent_t e1) { return launch_upload(a->d, a->hd, a->used, s, e0, e1); }));
    return VACO_AV1PF_OK;
}

// Marks the end of the slot's frames on stream s; host_visible: the event is
// handed on as the frames' completion (the server), so it carries the
// system-scope fence. u0, u1 (may be null): the shared buffers the call used,
// whose last use it now is.
static vaco_av1pf_status close_slot(vaco_av1pf_ctx *ctx, Arena *a, hipStream_t s, bool host_visible = false,
                                    Use *u0 = nullptr, Use *u1 = nullptr)
{
    a->done_ev = host_visible ? a->done_sys : a->done;
    a->stream = s;
    HIPCHECK(hipEventRecord(a->done_ev, s));
    a->pending = true;
    for (Use *u : {u0, u1})
        if (u) {
            u->stream = s;
            u->done = a->done_ev;
        }
    return VACO_AV1PF_OK;
}

// A call about to use shared buffer u on stream s: if its last use was on
// another stream, s waits for that use's end. (A slot's event is recorded
// again only once its previous work has completed, take_slot, so a newer
// record never hides an older use.)
static vaco_av1pf_status acquire(vaco_av1pf_ctx *ctx, const Use &u, hipStream_t s)
{
    if (u.done && s != u.stream)
        HIPCHECK(hipStreamWaitEvent(s, u.done, 0));
    return VACO_AV1PF_OK;
}

// The intermediate set for a per-frame call on stream s: one last used on s
// (no wait), else an unused one, else the sets in turn.
static int pick_set(vaco_av1pf_ctx *ctx, hipStream_t s)
{
    for (int k = 0; k < kSets; k++)
        if (ctx->set_use[k].stream == s)
            return k;
    for (int k = 0; k < kSets; k++)
        if (!ctx->set_use[k].done)
            return k;
    const int k = ctx->next_set;
    ctx->next_set = (k + 1) % kSets;
    return k;
}

// ------------------------------------------------------------ create --

extern "C" uint32_t vaco_av1pf_abi_version(void) { return VACO_AV1PF_ABI_VERSION; }

extern "C" const char *vaco_av1pf_last_error(const vaco_av1pf_ctx *ctx)
{
    return ctx ? ctx->err : "null context";
}

extern "C" vaco_av1pf_status vaco_av1pf_create(int device, uint32_t max_width, uint32_t max_height,
                                               vaco_av1pf_ctx **out)
{
    if (!out || max_width == 0 || max_height == 0 || max_width > 65536 || max_height > 65536)
        return VACO_AV1PF_INVALID_ARGUMENT;
    *out = nullptr;
    vaco_av1pf_ctx *ctx = new (std::nothrow) vaco_av1pf_ctx();
    if (!ctx)
        return VACO_AV1PF_DEVICE_ERROR;
    auto bail = [&](hipError_t e) {
        vaco_av1pf_destroy(ctx);
        (void)e;
        return VACO_AV1PF_DEVICE_ERROR;
    };
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess)
        return bail(e);
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return bail(e);
    // Mi-padded (multiple of 8) maximum geometry.
    ctx->max_w = (max_width + 7) & ~7u;
    ctx->max_h = (max_height + 7) & ~7u;
    ctx->plane_cap = (size_t)ctx->max_w * ctx->max_h;
    e = hipMalloc(&ctx->pool, ctx->plane_cap * kPoolPlanes * sizeof(uint16_t));
    if (e != hipSuccess)
        return bail(e);
    for (int p = 0; p < 3; p++) {
        ctx->img_in[p] = ctx->pool + ctx->plane_cap * (size_t)p;
        ctx->img_out[p] = ctx->pool + ctx->plane_cap * (size_t)(3 + p);
    }
    for (int k = 0; k < kSets; k++)
        ctx->set_img[k] = ctx->pool + ctx->plane_cap * (size_t)(6 + k * kSetPlanes);
    const size_t blocks8 = (size_t)(ctx->max_w / 8 + 2) * (ctx->max_h / 8 + 2);
    const size_t units = (size_t)(ctx->max_w / 32 + 2) * (ctx->max_h / 32 + 2) * 3;
    const size_t cap =
        ((1u << 20) + blocks8 * 2 + units * sizeof(LrUnit) + ctx->max_h * 64 + 255) & ~size_t(255);
    for (int i = 0; i < kSlots; i++) {
        Arena &a = ctx->slot[i];
        e = alloc_slot(a, cap);
        if (e != hipSuccess)
            return bail(e);
        // Without the system-scope fence an event that ends a frame costs
        // ~2 us less on the per-frame path (no cache writeback for the host).
        e = hipEventCreateWithFlags(&a.done, hipEventDisableTiming | hipEventDisableSystemFence);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&a.done_sys, hipEventDisableTiming);
        if (e != hipSuccess)
            return bail(e);
    }
    // Chained-launch state: one flag per superblock of the largest frame,
    // then the copy counter.
    ctx->chain_sbs = (size_t)((ctx->max_w + 63) / 64) * ((ctx->max_h + 63) / 64);
    const size_t flag_bytes = (ctx->chain_sbs * sizeof(uint32_t) + 255) & ~size_t(255);
    e = hipMalloc((void **)&ctx->chain_flags, flag_bytes + 256);
    if (e == hipSuccess)
        e = hipMemset(ctx->chain_flags, 0, flag_bytes + 256);
    if (e != hipSuccess)
        return bail(e);
    ctx->chain_copied = reinterpret_cast<unsigned long long *>(reinterpret_cast<uint8_t *>(ctx->chain_flags) + flag_bytes);
    // Status words: coherent so kernel atomics are visible to the host after
    // the frame's event completes; one 64-byte line per slot.
    e = hipHostMalloc((void **)&ctx->h_status, kSlots * kMaxBatch * sizeof(uint32_t),
                      hipHostMallocCoherent);
    if (e != hipSuccess)
        return bail(e);
    for (int i = 0; i < kSlots; i++) {
        ctx->slot[i].status = ctx->h_status + i * kMaxBatch;
        memset(ctx->slot[i].status, 0, kMaxBatch * sizeof(uint32_t));
    }
    *out = ctx;
    return VACO_AV1PF_OK;
}

extern "C" void vaco_av1pf_destroy(vaco_av1pf_ctx *ctx)
{
    if (!ctx)
        return;
    // Teardown has no caller to report to: every step runs whatever the
    // previous one returned.
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < kSlots; i++)
        if (ctx->slot[i].pending)
            (void)hipEventSynchronize(ctx->slot[i].done_ev);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < kSlots; i++) {
        if (ctx->slot[i].h)
            (void)hipHostFree(ctx->slot[i].h);
        if (ctx->slot[i].d)
            (void)hipFree(ctx->slot[i].d);
        if (ctx->slot[i].done)
            (void)hipEventDestroy(ctx->slot[i].done);
        if (ctx->slot[i].done_sys)
            (void)hipEventDestroy(ctx->slot[i].done_sys);
        for (auto &gr : ctx->slot[i].graphs)
            (void)hipGraphExecDestroy(gr.exec);
    }
    if (ctx->h_status)
        (void)hipHostFree(ctx->h_status);
    for (auto ev : ctx->prof_pool)
        (void)hipEventDestroy(ev);
    if (ctx->pool)
        (void)hipFree(ctx->pool);
    if (ctx->ws)
        (void)hipFree(ctx->ws);
    if (ctx->chain_flags)
        (void)hipFree(ctx->chain_flags);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// ------------------------------------------------------- host helpers --

extern "C" uint16_t vaco_av1pf_adjust_strength(uint16_t primary, uint32_t variance)
{
    return host_adjust_strength(primary, variance);
}

extern "C" uint8_t vaco_av1pf_chroma_direction(uint8_t direction, uint8_t sub_x, uint8_t sub_y)
{
    return host_chroma_direction(direction, sub_x, sub_y);
}

extern "C" void vaco_av1pf_lr_unit_counts(const vaco_av1pf_lr_plane_config *cfg, uint32_t *cols,
                                          uint32_t *rows)
{
    uint32_t c = 1, r = 1;
    if (cfg)
        lr_unit_counts(cfg->width, cfg->height, cfg->unit_size, &c, &r);
    if (cols)
        *cols = c;
    if (rows)
        *rows = r;
}

extern "C" uint32_t vaco_av1pf_lr_unit_index(const vaco_av1pf_lr_plane_config *cfg, uint32_t x,
                                             uint32_t y)
{
    if (!cfg)
        return 0;
    uint32_t cols, rows;
    lr_unit_counts(cfg->width, cfg->height, cfg->unit_size, &cols, &rows);
    const uint32_t u = cfg->unit_size;
    uint32_t row = u ? (y + (8u >> (cfg->subsampling_y ? 1 : 0))) / u : 0;
    uint32_t col = u ? x / u : 0;
    row = std::min(row, rows - 1);
    col = std::min(col, cols - 1);
    return row * cols + col;
}

extern "C" void vaco_av1pf_skip8_from_mi(const uint8_t *mi_skip, uint32_t mi_rows, uint32_t mi_cols,
                                         uint8_t *skip8)
{
    if (!mi_skip || !skip8)
        return;
    for (uint32_t r = 0; r + 1 < mi_rows + 1 && r < mi_rows; r += 2)
        for (uint32_t c = 0; c < mi_cols; c += 2) {
            // decode.rs:892-896: all four cells present and skip (mi_at is
            // None outside the grid, which counts as not skipped).
            bool all = true;
            for (uint32_t dr = 0; dr < 2; dr++)
                for (uint32_t dc = 0; dc < 2; dc++) {
                    const uint32_t rr = r + dr, cc = c + dc;
                    if (rr >= mi_rows || cc >= mi_cols || !mi_skip[(size_t)rr * mi_cols + cc])
                        all = false;
                }
            skip8[(size_t)(r / 2) * (mi_cols / 2 ? mi_cols / 2 : 1) + c / 2] = all ? 1 : 0;
        }
}

extern "C" vaco_av1pf_status vaco_av1pf_frame_input_geometry(const vaco_av1pf_frame_desc *d,
                                                             uint32_t width[3], uint32_t height[3])
{
    if (!d || !width || !height)
        return VACO_AV1PF_INVALID_ARGUMENT;
    FrameGeo g;
    frame_geometry(*d, &g);
    for (int p = 0; p < 3; p++) {
        width[p] = g.in_w[p];
        height[p] = g.in_h[p];
    }
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_frame_output_geometry(const vaco_av1pf_frame_desc *d,
                                                              uint32_t width[3],
                                                              uint32_t height[3])
{
    if (!d || !width || !height)
        return VACO_AV1PF_INVALID_ARGUMENT;
    FrameGeo g;
    frame_geometry(*d, &g);
    for (int p = 0; p < 3; p++) {
        width[p] = g.out_w[p];
        height[p] = g.out_h[p];
    }
    return VACO_AV1PF_OK;
}

// ------------------------------------------------- per-frame planning --

namespace {

// A device plane. Frame outputs (vaco_av1pf_frame_device_to_frame) set u8
// for one byte per sample (stride then counts bytes) and crop_w/crop_h to
// the frame plane's size: the final stage stores only that region.
struct DevPlane {
    uint16_t *data;
    uint32_t stride;
    uint8_t u8 = 0;
    uint32_t crop_w = 0, crop_h = 0;  // 0: no crop
};

// Row segments: intersections of restoration stripes (restoration.rs:218-231)
// with unit rows (restoration.rs:124-133), clipped to the plane.
void lr_row_segments(uint32_t height, uint32_t unit, bool ss_y, std::vector<LrRowSeg> &out)
{
    out.clear();
    uint32_t ucols_dummy, urows;
    lr_unit_counts(1, height, unit, &ucols_dummy, &urows);
    const int sh = 64 >> (ss_y ? 1 : 0), off = 8 >> (ss_y ? 1 : 0);
    std::vector<int> cuts;
    cuts.push_back(0);
    for (int y = sh - off; y < (int)height; y += sh)
        cuts.push_back(y);
    for (uint32_t r = 1; r < urows; r++)
        cuts.push_back((int)(r * unit) - off);
    cuts.push_back((int)height);
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    for (size_t i = 0; i + 1 < cuts.size(); i++) {
        const int y0 = cuts[i], y1 = cuts[i + 1];
        if (y0 < 0 || y1 <= y0)
            continue;
        LrRowSeg s;
        s.y0 = y0;
        s.h = y1 - y0;
        const int stripe = (y0 + off) / sh;
        s.stripe_start = stripe * sh - off;
        s.stripe_end = s.stripe_start + sh - 1;
        s.unit_row = std::min<uint32_t>(urows - 1, (uint32_t)(y0 + off) / unit);
        out.push_back(s);
    }
}

void lr_col_segments(uint32_t width, uint32_t unit, std::vector<LrColSeg> &out)
{
    out.clear();
    uint32_t ucols, urows_dummy;
    lr_unit_counts(width, 1, unit, &ucols, &urows_dummy);
    for (uint32_t c = 0; c < ucols; c++) {
        const uint32_t x0 = c * unit, x1 = c + 1 == ucols ? width : std::min(width, (c + 1) * unit);
        for (uint32_t x = x0; x < x1; x += 64) {
            LrColSeg s;
            s.x0 = (int)x;
            s.w = (int)std::min<uint32_t>(64, x1 - x);
            s.unit_col = c;
            out.push_back(s);
        }
    }
}

// Row segments cut into tiles of at most `rows` rows; each piece keeps its
// segment's stripe and unit row.
void lr_cut_rows(const std::vector<LrRowSeg> &segs, int rows, std::vector<LrRowSeg> &out)
{
    out.clear();
    for (const LrRowSeg &r : segs)
        for (int y = 0; y < r.h; y += rows) {
            LrRowSeg c = r;
            c.y0 = r.y0 + y;
            c.h = std::min(r.h - y, rows);
            out.push_back(c);
        }
}

// One k_lr list entry: the tile's row segment ri and column segment ci of
// its item (LrPlane index), and its kind (0 copy, 1 Wiener, 2 self-guided).
// Segment counts stay below 2^16 (a plane of at most 65536 samples a side
// has at most 2050 row and 1025 column segments).
static uint64_t lr_entry(uint32_t ri, uint32_t ci, uint32_t item, int kind)
{
    return (uint64_t)ri << 48 | (uint64_t)ci << 32 | (uint64_t)item << 2 | (uint64_t)kind;
}

// k_lr's tile list: self-guided tiles, then Wiener, then copies (by_kind), or
// the one plane-order list (remapped in the kernel, xcd_remap), written to
// `out`. In a by-kind list, each kind's tiles (in plane order) are dealt to
// the XCDs round-robin in chunks of kLrChunk consecutive tiles (workgroup i
// runs on XCD i % 8, each XCD with its own L2): neighbours in a chunk share
// halo lines in one L2, while every XCD still gets a share of each kind's
// long tiles.
constexpr size_t kLrChunk = 4;
constexpr size_t kCus = 256;  // compute units of an MI355X (8 XCDs of 32)
static void lr_tile_list(const std::vector<uint64_t> (&tiles)[3], bool by_kind, uint64_t *out)
{
    size_t a = 0;
    for (int k = 2; k >= 0; k--) {
        const std::vector<uint64_t> &t = tiles[k];
        const size_t n = t.size(), b = a + n;
        if (!by_kind) {
            std::copy(t.begin(), t.end(), out + a);
        } else {
            // the next free position of each XCD within [a, b)
            size_t next[8];
            for (size_t x = 0; x < 8; x++)
                next[x] = a + (x + 8 - a % 8) % 8;
            size_t x = 0;
            for (size_t i = 0; i < n; i++) {
                if (i % kLrChunk == 0 && i)
                    x = (x + 1) % 8;
                while (next[x] >= b)  // this XCD's positions are used up
                    x = (x + 1) % 8;
                out[next[x]] = t[i];
                next[x] += 8;
            }
        }
        a = b;
    }
    // Workgroup i of a launch lands on CU i % kCus, so CU c takes entries c,
    // c + kCus, c + 2 kCus ...: reversing every other block of kCus entries
    // pairs the CUs that took the longest tiles of one block with the
    // shortest of the next (kind order is roughly longest first). A block
    // reversed keeps its XCDs (kCus is a multiple of 8), so chunks stay
    // together.
    if (by_kind)
        for (size_t r = kCus; r < a; r += 2 * kCus)
            std::reverse(out + r, out + std::min(a, r + kCus));
}

}  // namespace

// Template instance of the CDEF kernel for a frame's chroma layout.
static int cdef_mode(const vaco_av1pf_frame_desc &d)
{
    if (d.monochrome)
        return 1;
    if (d.subsampling_x && d.subsampling_y)
        return 0;
    if (!d.subsampling_x && !d.subsampling_y)
        return 1;
    return d.subsampling_x ? 2 : 3;
}

// Device intermediates (CDEF image, upscaled images) of the frames of one
// launch set, carved from one allocation.
struct Scratch {
    uint16_t *base = nullptr;
    size_t cap = 0, used = 0;  // in samples
    uint16_t *take(size_t n)
    {
        const size_t a = (used + 127) & ~size_t(127);
        if (a + n > cap)
            return nullptr;
        used = a + n;
        return base + a;
    }
};

// Samples of intermediates a frame needs (see plan_frame): the CDEF image
// when a stage follows CDEF, and with superres and restoration, the upscaled
// CDEF image of each plane that restoration only copies (planes it filters
// upscale inside restoration's staging).
static size_t scratch_need(const vaco_av1pf_frame_desc &d, const FrameGeo &g)
{
    size_t n = 0;
    for (int p = 0; p < g.nplanes; p++) {
        if (g.cdef && (g.superres || g.lr))
            n += (size_t)g.in_w[p] * g.in_h[p] + 128;
        if (g.superres && g.lr && d.lr_type[p] == VACO_AV1PF_RESTORE_NONE)
            n += (size_t)g.sr_out_w[p] * g.sr_h[p] + 128;
    }
    return n;
}

// An upper bound of scratch_need over every parameter choice of frames of
// geometry g: both images in every plane.
static size_t scratch_bound(const FrameGeo &g)
{
    size_t n = 0;
    for (int p = 0; p < g.nplanes; p++)
        n += (size_t)g.in_w[p] * g.in_h[p] + (size_t)g.sr_out_w[p] * g.sr_h[p] + 256;
    return n;
}

// Work descriptors of one launch set: every frame's CDEF superblocks, superres
// rows and restoration tiles, each kernel's items with running block offsets.
struct Plan {
    std::vector<CdefFrame> cdef[4];  // by chroma layout (cdef_mode)
    uint32_t cdef_blocks[4] = {};
    std::vector<SrJob> sr;
    uint32_t sr_blocks = 0;
    std::vector<LrPlane> lr;
    uint32_t lr_blocks = 0;
    // restoration tiles by kernel: 0 copy, 1 Wiener, 2 self-guided (lr_entry)
    std::vector<uint64_t> lr_tiles[3];
    struct Copy {
        DevPlane src, dst;
        uint32_t w, h;
    };  // w x h samples; dst.u8: to bytes
    std::vector<Copy> copies;  // frames in which no stage runs
    bool superres = false;  // some frame is upscaled (restoration then upscales inside its staging)
    // Restoration units stay in the slot's host memory (read in place by the
    // restoration tiles, as CDEF reads its side information): batches, whose
    // deep grids hide that latency, so that their upload carries descriptors
    // only.
    bool host_units = false;
    vaco_av1pf_tuning tune = kDefaultTuning;  // launch variants this plan is built for
    // algorithmic bytes per kernel id (samples read once + written once, 2 B)
    uint64_t bytes[VACO_AV1PF_KERNEL_COUNT] = {};
    uint64_t cdef_bytes[4] = {};  // CDEF by chroma layout
};

#define TAKE(ptr, dptr, type, n)                                                    \
    type *dptr;                                                                     \
    type *ptr = a->take<type>((n), &dptr);                                          \
    if (!ptr)                                                                       \
        return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "staging arena exhausted");

// Append one frame's work (decode.rs:554-559) to `plan`: side information
// goes to the staging arena `a`, intermediates come from `ws`.
static vaco_av1pf_status plan_frame(vaco_av1pf_ctx *ctx, const vaco_av1pf_frame_desc &d,
                                    const FrameGeo &g, const DevPlane in[3], const DevPlane out[3],
                                    Arena *a, uint32_t *status, Scratch &ws, Plan &plan)
{
    const int np = g.nplanes;
    DevPlane cdef_out[3], cur[3], pre[3];
    for (int p = 0; p < np; p++)
        cdef_out[p] = cur[p] = pre[p] = in[p];
    if (!g.cdef && !g.superres && !g.lr) {
        // Nothing runs: the stage's output is its input (decode.rs:741-748).
        for (int p = 0; p < np; p++) {
            uint32_t w = g.in_w[p], h = g.in_h[p];
            if (out[p].crop_w) {
                w = std::min(w, out[p].crop_w);
                h = std::min(h, out[p].crop_h);
            }
            plan.copies.push_back({in[p], out[p], w, h});
        }
        return VACO_AV1PF_OK;
    }

    // ---- CDEF (decode.rs:739-861)
    if (g.cdef) {
        const size_t nidx = (size_t)g.idx_rows * d.cdef_stride;
        const size_t nskip = (size_t)(g.mi_rows / 2) * (g.mi_cols / 2);
        int8_t *didx;
        uint8_t *dskip;
        int8_t *hidx = a->take_host<int8_t>(nidx, &didx);
        uint8_t *hskip = a->take_host<uint8_t>(nskip, &dskip);
        if (!hidx || !hskip)
            return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "staging arena exhausted");
        memcpy(hidx, d.cdef_idx, nidx);
        memcpy(hskip, d.skip8, nskip);
        CdefFrame f;
        memset(&f, 0, sizeof f);
        for (int p = 0; p < np; p++) {
            if (!g.superres && !g.lr) {
                cdef_out[p] = out[p];
            } else {
                cdef_out[p] = DevPlane{ws.take((size_t)g.in_w[p] * g.in_h[p]), g.in_w[p]};
                if (!cdef_out[p].data)
                    return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "workspace exhausted");
            }
            f.in[p] = in[p].data;
            f.out[p] = cdef_out[p].data;
            f.width[p] = g.in_w[p];
            f.height[p] = g.in_h[p];
            f.in_stride[p] = in[p].stride;
            f.out_stride[p] = cdef_out[p].stride;
            f.crop_w[p] = g.in_w[p];
            f.crop_h[p] = g.in_h[p];
            if (cdef_out[p].crop_w) {
                f.crop_w[p] = std::min(f.crop_w[p], cdef_out[p].crop_w);
                f.crop_h[p] = std::min(f.crop_h[p], cdef_out[p].crop_h);
            }
        }
        f.out_u8 = cdef_out[0].u8;
        // Superres reads the CDEF output from the L2 right after; restoration
        // and the caller gain from non-temporal stores (nothing dirty is left
        // for the release at the launch's end to write back).
        f.out_nt = !g.superres;
        f.cdef_idx = didx;
        f.skip8 = dskip;
        f.idx_stride = d.cdef_stride;
        f.skip_stride = g.mi_cols / 2;
        f.sb_cols = (g.in_w[0] + 63) / 64;
        f.sb_rows = (g.in_h[0] + 63) / 64;
        f.status = status;
        f.bit_depth = d.bit_depth;
        f.ss_x = d.subsampling_x;
        f.ss_y = d.subsampling_y;
        f.nplanes = (uint8_t)np;
        f.damping = d.cdef.damping;
        f.vec_ok = 1;
        for (int p = 0; p < np; p++)
            if (((uintptr_t)in[p].data & 7) || (in[p].stride & 3) || (g.in_w[p] & 3))
                f.vec_ok = 0;
        for (int i = 0; i < 8; i++) {
            f.y_pri[i] = d.cdef.y[i].primary;
            f.y_sec[i] = d.cdef.y[i].secondary;
            f.uv_pri[i] = d.cdef.uv[i].primary;
            f.uv_sec[i] = d.cdef.uv[i].secondary;
        }
        const int mode = cdef_mode(d);
        for (int p = 0; p < np; p++)
            plan.cdef_bytes[mode] += (uint64_t)g.in_w[p] * g.in_h[p] * 4;
        f.block_begin = plan.cdef_blocks[mode];
        plan.cdef_blocks[mode] += f.sb_cols * f.sb_rows;
        plan.cdef[mode].push_back(f);
        for (int p = 0; p < np; p++)
            cur[p] = cdef_out[p];
    }

    // ---- superres (decode.rs:865-890, superres.rs:226-274)
    std::vector<LrRowSeg> lr_rows[3];
    std::vector<LrColSeg> lr_cols[3];
    if (g.lr) {
        for (int p = 0; p < np; p++) {
            if (d.lr_type[p] == VACO_AV1PF_RESTORE_NONE)
                continue;
            lr_row_segments(g.lr_h[p], d.lr_unit_size[p], p && d.subsampling_y, lr_rows[p]);
            lr_col_segments(g.lr_w[p], d.lr_unit_size[p], lr_cols[p]);
        }
    }
    // A plane that restoration filters is upscaled inside restoration's
    // staging (LrPlane::sr_step): no superres job, no upscaled images.
    plan.superres |= g.superres;
    auto folded = [&](int p) { return g.superres && g.lr && d.lr_type[p] != VACO_AV1PF_RESTORE_NONE; };
    if (g.superres) {
        for (int p = 0; p < np; p++) {
            // the CDEF image (the frame's output, or with restoration the
            // image a plane without restoration is copied from)
            if (folded(p))
                continue;
            DevPlane src = cur[p];
            DevPlane dst = out[p];  // the frame plane equals the upscaled plane
            if (g.lr) {
                dst = DevPlane{ws.take((size_t)g.sr_out_w[p] * g.sr_h[p]), g.sr_out_w[p]};
                if (!dst.data)
                    return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "workspace exhausted");
            }
            SrJob j;
            memset(&j, 0, sizeof j);
            j.in = src.data;
            j.out = dst.data;
            j.status = status;
            j.rows = nullptr;  // every row
            j.in_width = g.in_w[p];
            j.in_stride = src.stride;
            j.out_stride = dst.stride;
            j.out_width = g.sr_out_w[p];
            j.n_rows = g.sr_h[p];
            j.segs = (g.sr_out_w[p] + kSrSeg - 1) / kSrSeg;
            j.step = g.sr_step[p];
            j.start = g.sr_start[p];
            j.block_begin = plan.sr_blocks;
            j.max_sample = (uint16_t)((1u << d.bit_depth) - 1);
            j.out_u8 = dst.u8;
            j.vec_ok = !((uintptr_t)j.in & 7) && !(j.in_stride & 3);
            plan.sr_blocks += sr_job_blocks(j.n_rows, j.segs, plan.tune.superres_rows);
            plan.bytes[VACO_AV1PF_KERNEL_SUPERRES] += (uint64_t)j.n_rows * (g.sr_vis[p] + g.sr_out_w[p]) * 2;
            plan.sr.push_back(j);
            cur[p] = dst;
        }
    }

    // ---- loop restoration (decode.rs:683-737, restoration.rs:178-207)
    if (g.lr) {
        for (int p = 0; p < np; p++) {
            LrPlane L;
            memset(&L, 0, sizeof L);
            const bool none = d.lr_type[p] == VACO_AV1PF_RESTORE_NONE;
            std::vector<LrRowSeg> rows_none;
            std::vector<LrColSeg> cols_none;
            // Tiles: row segments cut to the tuned height (the stripe and unit
            // of each piece are its segment's).
            std::vector<LrRowSeg> rows_cut;
            lr_cut_rows(lr_rows[p], plan.tune.lr_tile_rows, rows_cut);
            const std::vector<LrRowSeg> *rows = &rows_cut;
            const std::vector<LrColSeg> *cols = &lr_cols[p];
            if (none) {
                // copy_visible_plane (decode.rs:673-681): 64x64 copy tiles
                for (uint32_t y = 0; y < g.lr_h[p]; y += 64)
                    rows_none.push_back({(int)y, (int)std::min<uint32_t>(64, g.lr_h[p] - y), 0, 0, 0});
                for (uint32_t x = 0; x < g.lr_w[p]; x += 64)
                    cols_none.push_back({(int)x, (int)std::min<uint32_t>(64, g.lr_w[p] - x), 0});
                rows = &rows_none;
                cols = &cols_none;
            }
            TAKE(hrs, drs, LrRowSeg, rows->size());
            TAKE(hcs, dcs, LrColSeg, cols->size());
            memcpy(hrs, rows->data(), rows->size() * sizeof(LrRowSeg));
            memcpy(hcs, cols->data(), cols->size() * sizeof(LrColSeg));
            if (!none) {
                static_assert(sizeof(LrUnit) == sizeof(vaco_av1pf_lr_unit), "unit layout");
                LrUnit *du;
                LrUnit *hu = plan.host_units ? a->take_host<LrUnit>(d.lr_unit_count[p], &du)
                                             : a->take<LrUnit>(d.lr_unit_count[p], &du);
                if (!hu)
                    return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "staging arena exhausted");
                memcpy(hu, d.lr_units[p], sizeof(LrUnit) * d.lr_unit_count[p]);
                L.units = du;
                uint32_t ucols, urows;
                lr_unit_counts(g.lr_w[p], g.lr_h[p], d.lr_unit_size[p], &ucols, &urows);
                L.unit_cols = ucols;
            }
            L.before = pre[p].data;
            L.after = cur[p].data;
            if (folded(p)) {  // before and after are the images before upscaling
                L.sr_step = g.sr_step[p];
                L.sr_start = g.sr_start[p];
                L.sr_in_width = g.in_w[p];
            }
            L.out = out[p].data;
            L.rows = drs;
            L.cols = dcs;
            L.status = status;
            L.before_stride = pre[p].stride;
            L.after_stride = cur[p].stride;
            L.out_stride = out[p].stride;
            L.width = g.lr_w[p];
            L.height = g.lr_h[p];
            L.n_rows = (uint32_t)rows->size();
            L.n_cols = (uint32_t)cols->size();
            L.block_begin = plan.lr_blocks;
            L.bit_depth = d.bit_depth;
            L.ss_y = (uint8_t)(p && d.subsampling_y);
            L.vec_ok = !((uintptr_t)L.before & 7) && !((uintptr_t)L.after & 7) &&
                       !(L.before_stride & 3) && !(L.after_stride & 3);
            L.out_u8 = out[p].u8;  // restored planes lie inside the frame planes
            L.out_vec = !((uintptr_t)L.out & 7) && !(L.out_stride & 3);
            plan.lr_blocks += L.n_rows * L.n_cols;
            const uint32_t item = (uint32_t)plan.lr.size();
            uint64_t rows_h = 0, cols_w = 0;
            for (const LrRowSeg &r : *rows)
                rows_h += (uint64_t)r.h;
            for (const LrColSeg &c : *cols)
                cols_w += (uint64_t)c.w;
            plan.bytes[VACO_AV1PF_KERNEL_RESTORATION] += rows_h * cols_w * 4;
            // Tiles in plane order (lr_order 1) need no list: the kernel finds
            // each workgroup's item and tile (fetch_tile). By kind: three lists.
            if (plan.tune.lr_order == 0)
                for (uint32_t ri = 0; ri < L.n_rows; ri++)
                    for (uint32_t ci = 0; ci < L.n_cols; ci++) {
                        int kind = 0;
                        if (!none) {
                            const uint8_t k =
                                d.lr_units[p][(*rows)[ri].unit_row * L.unit_cols + (*cols)[ci].unit_col].kind;
                            kind = k == VACO_AV1PF_UNIT_WIENER ? 1 : k == VACO_AV1PF_UNIT_SELF_GUIDED ? 2 : 0;
                        }
                        plan.lr_tiles[kind].push_back(lr_entry(ri, ci, item, kind));
                    }
            plan.lr.push_back(L);
        }
    }
    return VACO_AV1PF_OK;
}

// Upload the plan's descriptors with the arena (one copy) and launch each
// kernel once over all of its items.
//
// A single frame whose chain starts with CDEF (the per-frame path) has no
// separate upload: the CDEF launch reads its descriptor, cdef_idx and skip
// flags in place from the slot's coherent host memory (a few hundred bytes
// per superblock, read with the samples) and copies the staged bytes to the
// device mirror for the kernels after it. That saves a launch and a host
// round trip per frame.
static vaco_av1pf_status launch_plan(vaco_av1pf_ctx *ctx, const Plan &plan, Arena *a, hipStream_t s,
                                     bool single_frame = false)
{
    int cdef_modes = 0;
    for (int m = 0; m < 4; m++)
        cdef_modes += !plan.cdef[m].empty();
    // (Not when the chain is replayed as a graph: a graph keeps the kernel
    // arguments it was captured with, and this descriptor changes per frame.)
    const bool graphs = ctx && ctx->use_graphs && !ctx->prof;
    const bool fused_upload = single_frame && cdef_modes == 1 && !a->pageable && a->hd && !graphs;
    const CdefFrame *dc[4] = {};
    CdefFrame single;  // the frame's descriptor, passed as a kernel argument
    memset(&single, 0, sizeof single);
    for (int m = 0; m < 4; m++) {
        if (plan.cdef[m].empty())
            continue;
        if (fused_upload) {
            single = plan.cdef[m][0];  // its side information is in the slot's host memory
            dc[m] = &single;  // marks the mode; the kernel takes `single`
            continue;
        }
        TAKE(h, d, CdefFrame, plan.cdef[m].size());
        memcpy(h, plan.cdef[m].data(), plan.cdef[m].size() * sizeof(CdefFrame));
        dc[m] = d;
    }
    const SrJob *dj = nullptr;
    if (!plan.sr.empty()) {
        TAKE(h, d, SrJob, plan.sr.size());
        memcpy(h, plan.sr.data(), plan.sr.size() * sizeof(SrJob));
        dj = d;
    }
    const LrPlane *dl = nullptr;
    const uint64_t *dlist = nullptr;
    uint32_t nlist[3] = {};
    if (!plan.lr.empty()) {
        TAKE(h, d, LrPlane, plan.lr.size());
        memcpy(h, plan.lr.data(), plan.lr.size() * sizeof(LrPlane));
        dl = d;
        if (plan.tune.lr_order == 0) {  // by kind: the list
            size_t n = 0;
            for (int k = 0; k < 3; k++)
                n += nlist[k] = (uint32_t)plan.lr_tiles[k].size();
            TAKE(hl, dlk, uint64_t, n);
            lr_tile_list(plan.lr_tiles, true, hl);
            dlist = dlk;
        } else {  // plane order: no list
            nlist[0] = plan.lr_blocks;
        }
    }
    const uint32_t n_lr_items = (uint32_t)plan.lr.size();
    if (!fused_upload) {
        const vaco_av1pf_status st = upload_slot(ctx, a, s);
        if (st != VACO_AV1PF_OK)
            return st;
    }
    const size_t up_bytes = fused_upload ? a->used : 0;
    for (const auto &c : plan.copies) {
        if (c.dst.u8)
            HIPCHECK(launch_to_bytes(c.src.data, c.src.stride, (uint8_t *)c.dst.data, c.dst.stride,
                                     c.w, c.h, s));
        else
            HIPCHECK(hipMemcpy2DAsync(c.dst.data, c.dst.stride * 2, c.src.data, c.src.stride * 2,
                                      c.w * 2, c.h, hipMemcpyDeviceToDevice, s));
    }
    const uint32_t n_tiles = nlist[0] + nlist[1] + nlist[2];
    // A single frame with CDEF and restoration but no superres: both in one
    // launch (k_chain), if the tuning says so. (With superres a restoration
    // tile reads CDEF columns through the upscale, which the chain's
    // superblock dependencies do not follow.)
    int chain_mode = -1;
    if (fused_upload && ctx && ctx->tune.chain && !plan.superres && dj == nullptr && dl && n_tiles &&
        plan.copies.empty())
        for (int m = 0; m < 4; m++)
            if (dc[m] && plan.cdef[m].size() == 1 &&
                (size_t)plan.cdef[m][0].sb_cols * plan.cdef[m][0].sb_rows <= ctx->chain_sbs)
                chain_mode = m;
    if (chain_mode >= 0) {
        const int m = chain_mode;
        ChainSync cs;
        cs.flags = ctx->chain_flags;
        cs.copied = ctx->chain_copied;
        if (++ctx->chain_epoch == 0) {  // after 2^32 launches: clear the flags, start again
            HIPCHECK(hipMemsetAsync(ctx->chain_flags, 0, ctx->chain_sbs * sizeof(uint32_t), s));
            ctx->chain_epoch = 1;
        }
        cs.epoch = ctx->chain_epoch;
        cs.copied_target = ctx->chain_copied_total + chain_copiers(single.sb_cols * single.sb_rows, up_bytes);
        ctx->chain_copied_total = cs.copied_target;
        cs.spin_limit = ctx->chain_spin_limit;
        HIPCHECK(timed_launch(ctx, VACO_AV1PF_KERNEL_CHAIN,
                              plan.cdef_bytes[m] / 2 + plan.bytes[VACO_AV1PF_KERNEL_RESTORATION] / 2, s,
                              [&](hipEvent_t e0, hipEvent_t e1) {
                                  return launch_chain(single, m, plan.tune.cdef_xcd_order != 0, dl, n_lr_items,
                                                      dlist, n_tiles,
                                                      plan.tune.lr_order != 0, a->hd, a->d, up_bytes, cs, s, e0,
                                                      e1);
                              }));
        return VACO_AV1PF_OK;
    }
    auto launch_all = [&]() -> vaco_av1pf_status {
        for (int m = 0; m < 4; m++)
            if (dc[m])
                HIPCHECK(timed_launch(ctx, VACO_AV1PF_KERNEL_CDEF, plan.cdef_bytes[m], s,
                                      [&](hipEvent_t e0, hipEvent_t e1) {
                                          return launch_cdef(dc[m], (int)plan.cdef[m].size(),
                                                             fused_upload ? &single : nullptr,
                                                             plan.cdef_blocks[m], m,
                                                             plan.tune.cdef_xcd_order != 0,
                                                             plan.tune.cdef_waves, a->hd, a->d, up_bytes,
                                                             s, e0, e1);
                                      }));
        if (dj)
            HIPCHECK(timed_launch(ctx, VACO_AV1PF_KERNEL_SUPERRES,
                                  plan.bytes[VACO_AV1PF_KERNEL_SUPERRES], s,
                                  [&](hipEvent_t e0, hipEvent_t e1) {
                                      return launch_superres(dj, (int)plan.sr.size(), plan.sr_blocks,
                                                             plan.tune.superres_rows, s, e0, e1);
                                  }));
        if (dl)
            HIPCHECK(timed_launch(ctx, VACO_AV1PF_KERNEL_RESTORATION,
                                  plan.bytes[VACO_AV1PF_KERNEL_RESTORATION], s,
                                  [&](hipEvent_t e0, hipEvent_t e1) {
                                      return launch_lr(dl, n_lr_items, dlist, n_tiles, plan.tune.lr_order != 0,
                                                       plan.tune.lr_tile_rows, s, e0, e1);
                                  }));
        return VACO_AV1PF_OK;
    };
    if (!ctx || !ctx->use_graphs || ctx->prof)
        return launch_all();

    // Graph replay: the chain's launch parameters (descriptor addresses in
    // this slot, grid sizes, variants) identify its graph. Frames of one
    // geometry in one slot reuse it; the descriptors it reads are rewritten
    // by each call's upload.
    std::vector<uint64_t> key;
    for (int m = 0; m < 4; m++)
        key.insert(key.end(), {(uint64_t)(uintptr_t)dc[m], plan.cdef_blocks[m], plan.cdef[m].size()});
    key.insert(key.end(), {(uint64_t)(uintptr_t)dj, plan.sr_blocks, plan.sr.size(),
                           (uint64_t)(uintptr_t)dl, (uint64_t)(uintptr_t)dlist, n_tiles, n_lr_items,
                           plan.tune.superres_rows, plan.tune.cdef_xcd_order, plan.tune.lr_order,
                           plan.tune.cdef_waves, plan.tune.lr_tile_rows});
    hipGraphExec_t exec = nullptr;
    for (auto &gr : a->graphs)
        if (gr.key == key)
            exec = gr.exec;
    if (!exec) {
        HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const vaco_av1pf_status st2 = launch_all();
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(s, &g);
        if (st2 != VACO_AV1PF_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return st2;
        }
        if (ee != hipSuccess)
            return hip_fail(ctx, ee);
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess)
            return hip_fail(ctx, ei);
        if (a->graphs.size() >= 4) {
            HIPCHECK(hipGraphExecDestroy(a->graphs.front().exec));
            a->graphs.erase(a->graphs.begin());
        }
        a->graphs.push_back({key, exec});
    }
    HIPCHECK(hipGraphLaunch(exec, s));
    return VACO_AV1PF_OK;
}

#undef TAKE

// One frame on the per-frame path: intermediates from the context's images.
static vaco_av1pf_status enqueue_frame(vaco_av1pf_ctx *ctx, const vaco_av1pf_frame_desc &d,
                                       const FrameGeo &g, const DevPlane in[3],
                                       const DevPlane out[3], Arena *a, uint32_t *status,
                                       hipStream_t s, int set)
{
    Scratch ws;
    ws.base = ctx->set_img[set];
    ws.cap = ctx->plane_cap * kSetPlanes;
    Plan plan;
    plan.tune = ctx->tune;
    vaco_av1pf_status st = plan_frame(ctx, d, g, in, out, a, status, ws, plan);
    if (st != VACO_AV1PF_OK)
        return st;
    return launch_plan(ctx, plan, a, s, true);
}

// ---------------------------------------------------------- profiling --

static hipError_t profile_collect(vaco_av1pf_ctx *ctx)
{
    if (ctx->prof_recs.empty())
        return hipSuccess;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess)
        return e;
    for (const auto &r : ctx->prof_recs) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, ctx->prof_pool[r.a], ctx->prof_pool[r.b]);
        if (e != hipSuccess)
            return e;
        ctx->prof_ms[r.kernel] += ms;
        ctx->prof_n[r.kernel] += 1;
        ctx->prof_bytes[r.kernel] += r.bytes;
    }
    ctx->prof_recs.clear();
    ctx->prof_next = 0;
    return hipSuccess;
}

// launch(ev0, ev1) dispatches one kernel; while profiling, ev0 and ev1 are
// set by the dispatch itself at its start and end (hipExtLaunchKernel), the
// same interval a kernel trace reports.
template <class F>
static hipError_t timed_launch(vaco_av1pf_ctx *ctx, int kernel, uint64_t bytes, hipStream_t s,
                               F launch)
{
    (void)s;
    if (!ctx || !ctx->prof)
        return launch(nullptr, nullptr);
    if (ctx->prof_next + 2 > ctx->prof_pool.size()) {
        hipError_t e = profile_collect(ctx);
        if (e != hipSuccess)
            return e;
    }
    const uint32_t a = ctx->prof_next++, b = ctx->prof_next++;
    hipError_t e = launch(ctx->prof_pool[a], ctx->prof_pool[b]);
    if (e == hipSuccess)
        ctx->prof_recs.push_back({kernel, a, b, bytes});
    return e;
}

extern "C" vaco_av1pf_status vaco_av1pf_profile(vaco_av1pf_ctx *ctx, int enable)
{
    if (!ctx)
        return VACO_AV1PF_INVALID_ARGUMENT;
    HIPCHECK(hipSetDevice(ctx->device));
    if (enable && ctx->prof_pool.empty()) {
        ctx->prof_pool.resize(2048);
        for (auto &ev : ctx->prof_pool)
            HIPCHECK(hipEventCreate(&ev));
    }
    if (!enable)
        HIPCHECK(profile_collect(ctx));
    ctx->prof = enable != 0;
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_profile_read(vaco_av1pf_ctx *ctx, double ms[VACO_AV1PF_KERNEL_COUNT],
                                                     uint64_t launches[VACO_AV1PF_KERNEL_COUNT],
                                                     uint64_t bytes[VACO_AV1PF_KERNEL_COUNT])
{
    if (!ctx)
        return VACO_AV1PF_INVALID_ARGUMENT;
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(profile_collect(ctx));
    for (int k = 0; k < VACO_AV1PF_KERNEL_COUNT; k++) {
        if (ms)
            ms[k] = ctx->prof_ms[k];
        if (launches)
            launches[k] = ctx->prof_n[k];
        if (bytes)
            bytes[k] = ctx->prof_bytes[k];
        ctx->prof_ms[k] = 0;
        ctx->prof_n[k] = 0;
        ctx->prof_bytes[k] = 0;
    }
    return VACO_AV1PF_OK;
}

// Validate the caller's planes against the geometry the reference implies.
static vaco_av1pf_status check_geometry(const FrameGeo &g, uint32_t cdef_stride, const vaco_av1pf_picture *in,
                                        const vaco_av1pf_picture *out, const char **msg)
{
    for (int p = 0; p < g.nplanes; p++) {
        const vaco_av1pf_plane &i = in->plane[p], &o = out->plane[p];
        *msg = "missing plane";
        if (!i.data || !o.data)
            return VACO_AV1PF_INVALID_ARGUMENT;
        *msg = "input plane geometry differs from the Mi-padded frame";
        if (i.width != g.in_w[p] || i.height != g.in_h[p] || i.stride < i.width)
            return VACO_AV1PF_INVALID_ARGUMENT;
        *msg = "output plane geometry differs from the stage output";
        if (o.width != g.out_w[p] || o.height != g.out_h[p] || o.stride < o.width)
            return VACO_AV1PF_INVALID_ARGUMENT;
    }
    *msg = "CDEF side information missing";
    if (g.cdef && (!g.has_idx || !g.has_skip || cdef_stride < (g.mi_cols + 15) / 16))
        return VACO_AV1PF_INVALID_ARGUMENT;
    *msg = "";
    return VACO_AV1PF_OK;
}

// check_geometry plus the context's capacity.
static vaco_av1pf_status check_planes(vaco_av1pf_ctx *ctx, const FrameGeo &g, uint32_t cdef_stride,
                                      const vaco_av1pf_picture *in, const vaco_av1pf_picture *out)
{
    const char *msg = "";
    const vaco_av1pf_status st = check_geometry(g, cdef_stride, in, out, &msg);
    if (st != VACO_AV1PF_OK)
        return fail(ctx, st, msg);
    for (int p = 0; p < g.nplanes; p++)
        if ((size_t)g.in_w[p] > ctx->max_w || (size_t)g.in_h[p] > ctx->max_h ||
            (size_t)g.out_w[p] > ctx->max_w || (size_t)g.out_h[p] > ctx->max_h)
            return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "frame exceeds the context's capacity");
    return VACO_AV1PF_OK;
}

static void host_planes_of(const vaco_av1pf_picture *pic, int np, HostPlane hp[3])
{
    for (int p = 0; p < 3; p++) {
        hp[p].data = p < np ? pic->plane[p].data : nullptr;
        hp[p].width = p < np ? pic->plane[p].width : 0;
        hp[p].height = p < np ? pic->plane[p].height : 0;
        hp[p].stride = p < np ? pic->plane[p].stride : 0;
    }
}

extern "C" vaco_av1pf_status vaco_av1pf_frame(vaco_av1pf_ctx *ctx, const vaco_av1pf_frame_desc *desc,
                                              const vaco_av1pf_picture *in, vaco_av1pf_picture *out,
                                              uint64_t *fuel)
{
    if (!ctx || !desc || !in || !out)
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    ctx->err = "";
    FrameGeo g;
    frame_geometry(*desc, &g);
    vaco_av1pf_status st = check_planes(ctx, g, desc->cdef_stride, in, out);
    if (st != VACO_AV1PF_OK)
        return st;
    HIPCHECK(hipSetDevice(ctx->device));
    HostPlane hin[3];
    host_planes_of(in, g.nplanes, hin);
    // Slow path first when the reference's checks could fail on parameters
    // or fuel alone: replay them in the reference's order.
    const bool fuel_short = fuel && required_fuel(*desc, g) > *fuel;
    if (!params_clean(*desc, g) || fuel_short) {
        const char *msg = "";
        vaco_av1pf_status r = replay_checks(*desc, g, hin, fuel, &msg);
        if (r != VACO_AV1PF_OK)
            return fail(ctx, r, msg);
        fuel = nullptr;  // replay already charged it
    }
    Arena *a;
    if ((st = take_slot(ctx, &a)) != VACO_AV1PF_OK)
        return st;
    if ((st = grow_slot(ctx, a, arena_need(*desc, g))) != VACO_AV1PF_OK)
        return st;
    const int set = pick_set(ctx, ctx->stream);
    Use *chain = ctx->tune.chain ? &ctx->chain_use : nullptr;
    if ((st = acquire(ctx, ctx->set_use[set], ctx->stream)) != VACO_AV1PF_OK ||
        (chain && (st = acquire(ctx, *chain, ctx->stream)) != VACO_AV1PF_OK))
        return st;
    DevPlane din[3], dout[3];
    for (int p = 0; p < g.nplanes; p++) {
        din[p] = DevPlane{ctx->img_in[p], g.in_w[p]};
        dout[p] = DevPlane{ctx->img_out[p], g.out_w[p]};
        HIPCHECK(hipMemcpy2DAsync(din[p].data, din[p].stride * 2, in->plane[p].data,
                                  in->plane[p].stride * 2, g.in_w[p] * 2, g.in_h[p],
                                  hipMemcpyHostToDevice, ctx->stream));
    }
    st = enqueue_frame(ctx, *desc, g, din, dout, a, a->status, ctx->stream, set);
    if (st != VACO_AV1PF_OK)
        return st;
    for (int p = 0; p < g.nplanes; p++)
        HIPCHECK(hipMemcpy2DAsync(out->plane[p].data, out->plane[p].stride * 2, dout[p].data,
                                  dout[p].stride * 2, g.out_w[p] * 2, g.out_h[p],
                                  hipMemcpyDeviceToHost, ctx->stream));
    if ((st = close_slot(ctx, a, ctx->stream, false, &ctx->set_use[set], chain)) != VACO_AV1PF_OK)
        return st;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (*a->status & kFlagChainStall)
        return fail(ctx, VACO_AV1PF_DEVICE_ERROR, kChainStallMsg);
    if (*a->status & kFlagSampleRange) {
        const char *msg = "";
        vaco_av1pf_status r = replay_checks(*desc, g, hin, fuel, &msg);
        if (r != VACO_AV1PF_OK)
            return fail(ctx, r, msg);
    } else if (fuel) {
        *fuel -= required_fuel(*desc, g);
    }
    return VACO_AV1PF_OK;
}

// The device path of one frame into output planes `dout` (u16 picture
// planes, or frame planes of pic_to_frame's layout).
static vaco_av1pf_status frame_device_to(vaco_av1pf_ctx *ctx, const vaco_av1pf_frame_desc *desc,
                                         const FrameGeo &g, const vaco_av1pf_picture *in,
                                         const DevPlane dout[3], void *stream)
{
    vaco_av1pf_status st;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (!params_clean(*desc, g)) {
        // Parameter errors need the samples to order them exactly.
        std::vector<uint16_t> host[3];
        vaco_av1pf_picture hpic = *in;
        HIPCHECK(hipStreamSynchronize(s));
        for (int p = 0; p < g.nplanes; p++) {
            host[p].resize((size_t)g.in_w[p] * g.in_h[p]);
            HIPCHECK(hipMemcpy2D(host[p].data(), g.in_w[p] * 2, in->plane[p].data,
                                 in->plane[p].stride * 2, g.in_w[p] * 2, g.in_h[p],
                                 hipMemcpyDeviceToHost));
            hpic.plane[p].data = host[p].data();
            hpic.plane[p].stride = g.in_w[p];
        }
        HostPlane hin[3];
        host_planes_of(&hpic, g.nplanes, hin);
        const char *msg = "";
        vaco_av1pf_status r = replay_checks(*desc, g, hin, nullptr, &msg);
        if (r != VACO_AV1PF_OK)
            return fail(ctx, r, msg);
    }
    Arena *a;
    if ((st = take_slot(ctx, &a)) != VACO_AV1PF_OK)
        return st;
    if ((st = grow_slot(ctx, a, arena_need(*desc, g))) != VACO_AV1PF_OK)
        return st;
    // Intermediates from a set last used on this stream where there is one;
    // the chained launch's counters are ordered across streams (a chain
    // assumes the previous chain has completed).
    const int set = pick_set(ctx, s);
    if ((st = acquire(ctx, ctx->set_use[set], s)) != VACO_AV1PF_OK)
        return st;
    Use *chain = ctx->tune.chain ? &ctx->chain_use : nullptr;
    if (chain && (st = acquire(ctx, *chain, s)) != VACO_AV1PF_OK)
        return st;
    DevPlane din[3];
    for (int p = 0; p < g.nplanes; p++)
        din[p] = DevPlane{in->plane[p].data, in->plane[p].stride};
    st = enqueue_frame(ctx, *desc, g, din, dout, a, a->status, s, set);
    if (st != VACO_AV1PF_OK)
        return st;
    if ((st = close_slot(ctx, a, s, false, &ctx->set_use[set], chain)) != VACO_AV1PF_OK)
        return st;
    // The status is settled when the slot is reused or at finish().
    a->descs.assign(1, *desc);
    a->ins.assign(1, *in);
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_frame_device(vaco_av1pf_ctx *ctx,
                                                     const vaco_av1pf_frame_desc *desc,
                                                     const vaco_av1pf_picture *in,
                                                     vaco_av1pf_picture *out, void *stream)
{
    if (!ctx || !desc || !in || !out)
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    ctx->err = "";
    FrameGeo g;
    frame_geometry(*desc, &g);
    vaco_av1pf_status st = check_planes(ctx, g, desc->cdef_stride, in, out);
    if (st != VACO_AV1PF_OK)
        return st;
    DevPlane dout[3];
    for (int p = 0; p < g.nplanes; p++)
        dout[p] = DevPlane{out->plane[p].data, out->plane[p].stride};
    return frame_device_to(ctx, desc, g, in, dout, stream);
}

extern "C" vaco_av1pf_status vaco_av1pf_frame_layout(const vaco_av1pf_frame_desc *d,
                                                     uint32_t width[3], uint32_t height[3],
                                                     uint32_t *bytes_per_sample)
{
    if (!d || !width || !height || !bytes_per_sample)
        return VACO_AV1PF_INVALID_ARGUMENT;
    // pic_to_frame, decode.rs:3821-3829: upscaled_width x coded_height, chroma
    // rounded up.
    const uint32_t w = d->upscaled_width, h = d->coded_height;
    const uint32_t np = d->monochrome ? 1 : 3;
    for (uint32_t p = 0; p < 3; p++) {
        const bool chroma = p > 0;
        width[p] = p < np ? (chroma ? (w + d->subsampling_x) >> d->subsampling_x : w) : 0;
        height[p] = p < np ? (chroma ? (h + d->subsampling_y) >> d->subsampling_y : h) : 0;
    }
    *bytes_per_sample = d->bit_depth > 8 ? 2 : 1;
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_frame_device_to_frame(vaco_av1pf_ctx *ctx,
                                                              const vaco_av1pf_frame_desc *desc,
                                                              const vaco_av1pf_picture *in,
                                                              const vaco_av1pf_frame_planes *out,
                                                              void *stream)
{
    if (!ctx || !desc || !in || !out)
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    ctx->err = "";
    FrameGeo g;
    frame_geometry(*desc, &g);
    // Input planes are checked as for the picture output; the output is the
    // frame layout instead of the final picture.
    vaco_av1pf_picture probe;
    memset(&probe, 0, sizeof probe);
    for (int p = 0; p < g.nplanes; p++) {
        probe.plane[p].data = (uint16_t *)out->data[p];
        probe.plane[p].width = g.out_w[p];
        probe.plane[p].height = g.out_h[p];
        probe.plane[p].stride = g.out_w[p];
    }
    vaco_av1pf_status st = check_planes(ctx, g, desc->cdef_stride, in, &probe);
    if (st != VACO_AV1PF_OK)
        return st;
    uint32_t fw[3], fh[3], bps;
    vaco_av1pf_frame_layout(desc, fw, fh, &bps);
    DevPlane dout[3];
    for (int p = 0; p < g.nplanes; p++) {
        if (!out->data[p] || out->stride_bytes[p] < fw[p] * bps || out->stride_bytes[p] % bps)
            return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "frame plane geometry");
        dout[p].data = (uint16_t *)out->data[p];
        dout[p].stride = out->stride_bytes[p] / bps;
        dout[p].u8 = bps == 1;
        dout[p].crop_w = fw[p];
        dout[p].crop_h = fh[p];
    }
    return frame_device_to(ctx, desc, g, in, dout, stream);
}

// Settle a slot's deferred check: if its frame raised the sample-range flag,
// replay the reference's checks over a host copy of the device input.
static vaco_av1pf_status resolve_slot(vaco_av1pf_ctx *ctx, Arena &a)
{
    for (size_t i = 0; i < a.descs.size(); i++) {
        if ((a.status[i] & kFlagChainStall) && ctx->deferred == VACO_AV1PF_OK) {
            ctx->deferred = VACO_AV1PF_DEVICE_ERROR;
            ctx->deferred_msg = kChainStallMsg;
            continue;
        }
        if (!(a.status[i] & kFlagSampleRange) || ctx->deferred != VACO_AV1PF_OK)
            continue;
        const vaco_av1pf_frame_desc &d = a.descs[i];
        const vaco_av1pf_picture &in = a.ins[i];
        FrameGeo g;
        frame_geometry(d, &g);
        std::vector<uint16_t> host[3];
        vaco_av1pf_picture hpic = in;
        for (int p = 0; p < g.nplanes; p++) {
            host[p].resize((size_t)g.in_w[p] * g.in_h[p]);
            HIPCHECK(hipMemcpy2D(host[p].data(), g.in_w[p] * 2, in.plane[p].data,
                                 in.plane[p].stride * 2, g.in_w[p] * 2, g.in_h[p],
                                 hipMemcpyDeviceToHost));
            hpic.plane[p].data = host[p].data();
            hpic.plane[p].stride = g.in_w[p];
        }
        HostPlane hin[3];
        host_planes_of(&hpic, g.nplanes, hin);
        const char *msg = "";
        vaco_av1pf_status r = replay_checks(d, g, hin, nullptr, &msg);
        if (r != VACO_AV1PF_OK) {
            ctx->deferred = r;
            ctx->deferred_msg = msg;
        }
    }
    a.descs.clear();
    a.ins.clear();
    return VACO_AV1PF_OK;
}

// Staging bytes one frame can need (side information, segment lists,
// descriptors), an upper bound. Every TAKE rounds up to 256 bytes.
static size_t arena_need(const vaco_av1pf_frame_desc &d, const FrameGeo &g)
{
    size_t n = 16384;  // descriptors (CdefFrame, SrJob, LrPlane) and rounding
    if (g.cdef)
        n += (size_t)g.idx_rows * d.cdef_stride + (size_t)(g.mi_rows / 2) * (g.mi_cols / 2) + 1024;
    for (int p = 0; p < g.nplanes; p++) {
        // row pieces: at most one cut per 16 rows plus stripe and unit cuts
        // (every 32 rows at least); columns: 64-wide pieces plus unit cuts
        const size_t rows = (size_t)g.lr_h[p] / 8 + 8, cols = (size_t)g.lr_w[p] / 32 + 4;
        n += (size_t)d.lr_unit_count[p] * sizeof(LrUnit) + 512;
        n += rows * sizeof(LrRowSeg) + cols * sizeof(LrColSeg) + 512;
        n += rows * cols * sizeof(uint64_t) + 256;            // tile list
    }
    return n;
}

size_t av1pf_frame_device_bytes(const vaco_av1pf_frame_desc &d)
{
    FrameGeo g;
    frame_geometry(d, &g);
    size_t n = scratch_need(d, g);
    for (int p = 0; p < g.nplanes; p++)
        n += (size_t)g.in_w[p] * g.in_h[p] + (size_t)g.out_w[p] * g.out_h[p];
    return n * sizeof(uint16_t);
}

int av1pf_ctx_device(const vaco_av1pf_ctx *ctx) { return ctx->device; }

size_t av1pf_frame_stage_bytes(const vaco_av1pf_frame_desc &d)
{
    FrameGeo g;
    frame_geometry(d, &g);
    return arena_need(d, g);
}

size_t av1pf_frame_scratch_samples(const vaco_av1pf_frame_desc &d)
{
    FrameGeo g;
    frame_geometry(d, &g);
    return scratch_need(d, g);
}

// The largest frame a context accepts: 4:4:4 at its maximum geometry, with
// CDEF, superres and 32-sample restoration units in every plane.
static void worst_frame(const vaco_av1pf_ctx *ctx, vaco_av1pf_frame_desc &d, FrameGeo &g)
{
    static const int8_t idx = 0;
    static const uint8_t skip = 0;
    memset(&d, 0, sizeof d);
    d.coded_width = ctx->max_w - 8;
    d.upscaled_width = ctx->max_w;
    d.coded_height = ctx->max_h;
    d.use_superres = 1;
    d.bit_depth = 12;
    d.cdef.enabled = 1;
    d.cdef.y[0].primary = 1;
    d.cdef_idx = &idx;
    d.skip8 = &skip;
    d.cdef_stride = (ctx->max_w / 4 + 15) / 16;
    for (int p = 0; p < 3; p++) {
        d.lr_type[p] = VACO_AV1PF_RESTORE_SWITCHABLE;
        d.lr_unit_size[p] = 32;
        uint32_t c, r;
        lr_unit_counts(ctx->max_w, ctx->max_h, 32, &c, &r);
        d.lr_unit_count[p] = c * r;
    }
    frame_geometry(d, &g);
    // the Mi-padded input of a frame coded max_w wide (no superres)
    g.in_w[0] = g.in_w[1] = g.in_w[2] = ctx->max_w;
}

size_t av1pf_worst_frame_stage_bytes(const vaco_av1pf_ctx *ctx)
{
    vaco_av1pf_frame_desc d;
    FrameGeo g;
    worst_frame(ctx, d, g);
    return arena_need(d, g);
}

size_t av1pf_worst_frame_scratch_samples(const vaco_av1pf_ctx *ctx)
{
    vaco_av1pf_frame_desc d;
    FrameGeo g;
    worst_frame(ctx, d, g);
    return scratch_bound(g);
}

vaco_av1pf_status av1pf_reserve_staging(vaco_av1pf_ctx *ctx, size_t bytes)
{
    bytes = (bytes + 255) & ~size_t(255);
    HIPCHECK(hipSetDevice(ctx->device));
    for (int i = 0; i < kSlots; i++) {
        Arena &a = ctx->slot[i];
        if (a.cap >= bytes)
            continue;
        if (a.pending)
            HIPCHECK(hipEventSynchronize(a.done_ev));
        HIPCHECK(alloc_slot(a, bytes));
    }
    return VACO_AV1PF_OK;
}

size_t av1pf_staging_capacity(const vaco_av1pf_ctx *ctx)
{
    size_t c = ctx->slot[0].cap;
    for (int i = 1; i < kSlots; i++)
        c = std::min(c, ctx->slot[i].cap);
    return c;
}

int av1pf_next_slot(const vaco_av1pf_ctx *ctx) { return ctx->next_slot; }

void av1pf_settle_slot(vaco_av1pf_ctx *ctx, int slot)
{
    if (slot < 0 || slot >= kSlots)
        return;
    Arena &a = ctx->slot[slot];
    a.pending = false;
    a.stream = nullptr;
}

// Settle one frame's sample-range flag: replay the reference's checks over a
// host copy of its device input. Returns the reference's outcome.
vaco_av1pf_status av1pf_replay_device_frame(vaco_av1pf_ctx *ctx, const vaco_av1pf_frame_desc &d,
                                            const vaco_av1pf_picture &in, const char **msg)
{
    FrameGeo g;
    frame_geometry(d, &g);
    std::vector<uint16_t> host[3];
    vaco_av1pf_picture hpic = in;
    for (int p = 0; p < g.nplanes; p++) {
        host[p].resize((size_t)g.in_w[p] * g.in_h[p]);
        HIPCHECK(hipMemcpy2D(host[p].data(), g.in_w[p] * 2, in.plane[p].data, in.plane[p].stride * 2,
                             g.in_w[p] * 2, g.in_h[p], hipMemcpyDeviceToHost));
        hpic.plane[p].data = host[p].data();
        hpic.plane[p].stride = g.in_w[p];
    }
    HostPlane hin[3];
    host_planes_of(&hpic, g.nplanes, hin);
    *msg = "";
    return replay_checks(d, g, hin, nullptr, msg);
}

extern "C" vaco_av1pf_status vaco_av1pf_batch_device(vaco_av1pf_ctx *ctx, uint32_t n,
                                                     const vaco_av1pf_frame_desc *descs,
                                                     const vaco_av1pf_picture *ins,
                                                     vaco_av1pf_picture *outs, void *stream)
{
    if (!ctx)
        return VACO_AV1PF_INVALID_ARGUMENT;
    return av1pf_enqueue_batch(ctx, n, descs, ins, outs,
                               stream ? (hipStream_t)stream : ctx->stream, nullptr, nullptr, nullptr,
                               nullptr);
}

// The batched call. With `statuses` set, the caller takes over the frames'
// status checks: *statuses receives the slot's per-frame status words
// (valid until the slot is reused, kSlots batched calls later) and
// *done the event that completes the batch; otherwise the checks are
// deferred to vaco_av1pf_finish().
vaco_av1pf_status av1pf_enqueue_batch(vaco_av1pf_ctx *ctx, uint32_t n,
                                      const vaco_av1pf_frame_desc *descs,
                                      const vaco_av1pf_picture *ins, vaco_av1pf_picture *outs,
                                      hipStream_t s, const uint32_t **statuses, hipEvent_t *done,
                                      const av1pf_workspace *own_ws, int *slot_index)
{
    if (!ctx || (n && (!descs || !ins || !outs)))
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    if (n > kMaxBatch)
        return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "batch larger than 256 frames");
    ctx->err = "";
    if (n == 0)
        return VACO_AV1PF_OK;
    HIPCHECK(hipSetDevice(ctx->device));
    if (!s)
        s = ctx->stream;
    std::vector<FrameGeo> geo(n);
    size_t ws_need = 0, arena = kBatchStageBase;
    for (uint32_t i = 0; i < n; i++) {
        frame_geometry(descs[i], &geo[i]);
        vaco_av1pf_status st = check_planes(ctx, geo[i], descs[i].cdef_stride, &ins[i], &outs[i]);
        if (st != VACO_AV1PF_OK)
            return st;
        if (!params_clean(descs[i], geo[i])) {
            // Parameter errors need the samples to order them exactly.
            std::vector<uint16_t> host[3];
            vaco_av1pf_picture hpic = ins[i];
            HIPCHECK(hipStreamSynchronize(s));
            for (int p = 0; p < geo[i].nplanes; p++) {
                host[p].resize((size_t)geo[i].in_w[p] * geo[i].in_h[p]);
                HIPCHECK(hipMemcpy2D(host[p].data(), geo[i].in_w[p] * 2, ins[i].plane[p].data,
                                     ins[i].plane[p].stride * 2, geo[i].in_w[p] * 2, geo[i].in_h[p],
                                     hipMemcpyDeviceToHost));
                hpic.plane[p].data = host[p].data();
                hpic.plane[p].stride = geo[i].in_w[p];
            }
            HostPlane hin[3];
            host_planes_of(&hpic, geo[i].nplanes, hin);
            const char *msg = "";
            vaco_av1pf_status r = replay_checks(descs[i], geo[i], hin, nullptr, &msg);
            if (r != VACO_AV1PF_OK)
                return fail(ctx, r, msg);
        }
        ws_need += scratch_need(descs[i], geo[i]);
        arena += arena_need(descs[i], geo[i]);
    }
    Scratch ws;
    if (own_ws) {
        // The caller's workspace and staging capacity were sized for this
        // batch (the server forms batches to fit them): nothing grows.
        if (ws_need > own_ws->cap)
            return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "batch exceeds its workspace");
        ws.base = own_ws->base;
        ws.cap = own_ws->cap;
    } else if (ws_need > ctx->ws_cap) {
        // Intermediates of every frame of the batch; grown only when a batch
        // needs more than any batch before it.
        HIPCHECK(hipStreamSynchronize(s));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->ws_use.done)
            HIPCHECK(hipEventSynchronize(ctx->ws_use.done));
        if (ctx->ws)
            HIPCHECK(hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_cap = 0;
        HIPCHECK(hipMalloc((void **)&ctx->ws, ws_need * sizeof(uint16_t)));
        ctx->ws_cap = ws_need;
    }
    if (!own_ws) {
        ws.base = ctx->ws;
        ws.cap = ctx->ws_cap;
    }
    if (slot_index)
        *slot_index = ctx->next_slot;
    Arena *a;
    vaco_av1pf_status st = take_slot(ctx, &a);
    if (st != VACO_AV1PF_OK)
        return st;
    if (own_ws && arena > a->cap)
        return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "batch exceeds its staging slot");
    if ((st = grow_slot(ctx, a, arena)) != VACO_AV1PF_OK)
        return st;
    if (!own_ws && (st = acquire(ctx, ctx->ws_use, s)) != VACO_AV1PF_OK)
        return st;
    Plan plan;
    plan.tune = ctx->tune;
    plan.host_units = true;
    for (uint32_t i = 0; i < n; i++) {
        DevPlane din[3], dout[3];
        for (int p = 0; p < geo[i].nplanes; p++) {
            din[p] = DevPlane{ins[i].plane[p].data, ins[i].plane[p].stride};
            dout[p] = DevPlane{outs[i].plane[p].data, outs[i].plane[p].stride};
        }
        st = plan_frame(ctx, descs[i], geo[i], din, dout, a, a->status + i, ws, plan);
        if (st != VACO_AV1PF_OK)
            return st;
    }
    if ((st = launch_plan(ctx, plan, a, s)) != VACO_AV1PF_OK)
        return st;
    if ((st = close_slot(ctx, a, s, done != nullptr, own_ws ? nullptr : &ctx->ws_use)) != VACO_AV1PF_OK)
        return st;
    if (statuses) {
        *statuses = a->status;
        if (done)
            *done = a->done_ev;
        return VACO_AV1PF_OK;
    }
    a->descs.assign(descs, descs + n);
    a->ins.assign(ins, ins + n);
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_finish(vaco_av1pf_ctx *ctx)
{
    if (!ctx)
        return VACO_AV1PF_INVALID_ARGUMENT;
    HIPCHECK(hipSetDevice(ctx->device));
    // The streams the pending frames ran on are synchronised (which also
    // makes their outputs visible to the host); every pending slot's work
    // has then completed, and the slots' deferred checks are settled.
    hipStream_t used[kSlots];
    int n_used = 0;
    for (int i = 0; i < kSlots; i++) {
        const Arena &a = ctx->slot[i];
        if (a.pending && a.stream && a.stream != ctx->stream &&
            std::find(used, used + n_used, a.stream) == used + n_used)
            used[n_used++] = a.stream;
    }
    for (int i = 0; i < n_used; i++)
        HIPCHECK(hipStreamSynchronize(used[i]));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < kSlots; i++) {
        Arena &a = ctx->slot[i];
        if (a.pending && !a.stream)  // settled by its owner; its stream may be gone
            HIPCHECK(hipEventSynchronize(a.done_ev));
        a.pending = false;
        vaco_av1pf_status st = resolve_slot(ctx, a);
        if (st != VACO_AV1PF_OK)
            return st;
    }
    // Nothing is in flight: no later call need wait for an earlier one (and
    // the callers' streams may now go away).
    for (Use &u : ctx->set_use)
        u = Use{};
    ctx->ws_use = ctx->chain_use = Use{};
    const vaco_av1pf_status r = ctx->deferred;
    if (r != VACO_AV1PF_OK)
        fail(ctx, r, ctx->deferred_msg);
    ctx->deferred = VACO_AV1PF_OK;
    ctx->deferred_msg = "";
    return r;
}

// ------------------------------------------------------------ graphs --

extern "C" vaco_av1pf_status vaco_av1pf_use_graphs(vaco_av1pf_ctx *ctx, int enable)
{
    if (!ctx)
        return VACO_AV1PF_INVALID_ARGUMENT;
    ctx->use_graphs = enable != 0;
    return VACO_AV1PF_OK;
}

// ------------------------------------------------------------ tuning --

static bool tuning_ok(const vaco_av1pf_tuning &t)
{
    return (t.superres_rows == 4 || t.superres_rows == 8 || t.superres_rows == 16) &&
           t.cdef_xcd_order <= 1 && t.lr_order <= 1 &&
           (t.lr_tile_rows == 64 || t.lr_tile_rows == 32 || t.lr_tile_rows == 16) &&
           (t.cdef_waves == 16 || t.cdef_waves == 8) && t.chain <= 1;
}

extern "C" vaco_av1pf_status vaco_av1pf_get_tuning(const vaco_av1pf_ctx *ctx, vaco_av1pf_tuning *out)
{
    if (!ctx || !out)
        return VACO_AV1PF_INVALID_ARGUMENT;
    *out = ctx->tune;
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_set_tuning(vaco_av1pf_ctx *ctx, const vaco_av1pf_tuning *t)
{
    if (!ctx || !t || !tuning_ok(*t))
        return VACO_AV1PF_INVALID_ARGUMENT;
    ctx->tune = *t;
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_autotune(vaco_av1pf_ctx *ctx, uint32_t n,
                                                 const vaco_av1pf_frame_desc *descs,
                                                 const vaco_av1pf_picture *ins,
                                                 vaco_av1pf_picture *outs, uint32_t reps,
                                                 vaco_av1pf_tuning *out)
{
    if (!ctx || !out || reps == 0 || n == 0)
        return VACO_AV1PF_INVALID_ARGUMENT;
    vaco_av1pf_status st = vaco_av1pf_finish(ctx);
    if (st != VACO_AV1PF_OK)
        return st;
    const bool was_profiling = ctx->prof;
    if ((st = vaco_av1pf_profile(ctx, 1)) != VACO_AV1PF_OK)
        return st;
    double ms[VACO_AV1PF_KERNEL_COUNT];
    uint64_t launches[VACO_AV1PF_KERNEL_COUNT];
    // Time of `kernel` per call with tuning t: the median of `reps` calls
    // (robust both to a call slowed by something else on the device and to
    // the wide spread of one frame's short launches, where the fastest of a
    // few calls picks noise), after one untimed call.
    // One frame is timed through the per-frame path (whose CDEF launch reads
    // its side information in place and copies the staged bytes), several
    // through one batched call: each path's own kernels.
    auto run = [&]() {
        if (n == 1)
            return vaco_av1pf_frame_device(ctx, descs, ins, outs, nullptr);
        return av1pf_enqueue_batch(ctx, n, descs, ins, outs, nullptr, nullptr, nullptr, nullptr, nullptr);
    };
    std::vector<double> samples;
    auto measure = [&](const vaco_av1pf_tuning &t, int kernel, double *res) {
        ctx->tune = t;
        *res = 0.0;
        samples.clear();
        vaco_av1pf_status r = run();
        if (r == VACO_AV1PF_OK)
            r = vaco_av1pf_finish(ctx);
        if (r == VACO_AV1PF_OK)
            r = vaco_av1pf_profile_read(ctx, nullptr, nullptr, nullptr);
        for (uint32_t i = 0; i < reps && r == VACO_AV1PF_OK; i++) {
            r = run();
            if (r == VACO_AV1PF_OK)
                r = vaco_av1pf_finish(ctx);
            if (r == VACO_AV1PF_OK)
                r = vaco_av1pf_profile_read(ctx, ms, launches, nullptr);
            // kernel < 0: every kernel of the call, summed
            double v = 0;
            uint64_t nl = 0;
            for (int k = 0; k < VACO_AV1PF_KERNEL_COUNT; k++)
                if (kernel < 0 || k == kernel) {
                    v += ms[k];
                    nl += launches[k];
                }
            if (r == VACO_AV1PF_OK && nl)
                samples.push_back(v);
        }
        if (r == VACO_AV1PF_OK && !samples.empty()) {
            std::sort(samples.begin(), samples.end());
            const size_t h = samples.size() / 2;
            *res = samples.size() % 2 ? samples[h] : 0.5 * (samples[h - 1] + samples[h]);
        }
        return r;
    };
    vaco_av1pf_tuning best = ctx->tune, t;
    // The separate kernels are tuned first; for one frame the chain is
    // decided after them, batches (which never chain) keep the value.
    const uint8_t chain_in = best.chain;
    best.chain = 0;
    // One field at a time: the current value is timed first, and another
    // replaces it only when at least 2% faster, so that timing noise between
    // equivalent variants does not flip the choice. A value listed in
    // `costly` costs more outside the kernel's time than the others (the
    // plain CDEF order reads each superblock's halo lines into a different L2
    // than its neighbours: 2.3x the algorithmic bytes against 1.1x; the
    // by-kind restoration order makes the host build a tile list), so it
    // must be at least 5% faster. A kernel absent from the batch times 0 and keeps its
    // value.
    auto tune_field = [&](uint8_t vaco_av1pf_tuning::*field, std::initializer_list<uint8_t> values,
                          int kernel, int costly = -1) {
        double keep_ms = 0, tm = 0;
        if (st != VACO_AV1PF_OK || (st = measure(best, kernel, &keep_ms)) != VACO_AV1PF_OK || keep_ms <= 0)
            return;
        uint8_t keep = best.*field;
        for (uint8_t v : values) {
            if (v == best.*field)
                continue;
            t = best;
            t.*field = v;
            if ((st = measure(t, kernel, &tm)) != VACO_AV1PF_OK)
                return;
            if (tm > 0 && tm < keep_ms * (v == costly ? 0.95 : 0.98)) {
                keep_ms = tm;
                keep = v;
            }
        }
        best.*field = keep;
    };
    tune_field(&vaco_av1pf_tuning::superres_rows, {4, 8, 16}, VACO_AV1PF_KERNEL_SUPERRES);
    tune_field(&vaco_av1pf_tuning::cdef_xcd_order, {0, 1}, VACO_AV1PF_KERNEL_CDEF, 0);
    tune_field(&vaco_av1pf_tuning::cdef_waves, {16, 8}, VACO_AV1PF_KERNEL_CDEF);
    // By kind, the host builds the tile list (a pass over every tile); for a
    // batch that host time rivals the kernels', so plane order (no list)
    // yields only to a clearly faster kernel.
    tune_field(&vaco_av1pf_tuning::lr_order, {0, 1}, VACO_AV1PF_KERNEL_RESTORATION, n > 1 ? 0 : -1);
    tune_field(&vaco_av1pf_tuning::lr_tile_rows, {64, 32, 16}, VACO_AV1PF_KERNEL_RESTORATION);
    // One frame: the chained launch against the separate ones (every kernel
    // of the call, summed; the chain also saves the gap between launches,
    // which no kernel time holds, so it needs only to be no slower). If it
    // is kept, the orders and the tile height are tuned again on it.
    best.chain = chain_in;
    if (n == 1 && st == VACO_AV1PF_OK) {
        double sep_ms = 0, chain_ms = 0;
        t = best;
        t.chain = 0;
        if ((st = measure(t, -1, &sep_ms)) == VACO_AV1PF_OK) {
            t.chain = 1;
            // 0 when the frame does not chain (no CDEF, no restoration, or superres)
            st = measure(t, VACO_AV1PF_KERNEL_CHAIN, &chain_ms);
        }
        if (st == VACO_AV1PF_OK) {
            best.chain = chain_ms > 0 && chain_ms <= sep_ms ? 1 : 0;
            if (best.chain) {
                tune_field(&vaco_av1pf_tuning::cdef_xcd_order, {0, 1}, VACO_AV1PF_KERNEL_CHAIN, 0);
                tune_field(&vaco_av1pf_tuning::lr_order, {0, 1}, VACO_AV1PF_KERNEL_CHAIN);
                tune_field(&vaco_av1pf_tuning::lr_tile_rows, {64, 32, 16}, VACO_AV1PF_KERNEL_CHAIN);
            }
        }
    }
    ctx->tune = best;
    const vaco_av1pf_status ps = vaco_av1pf_profile(ctx, was_profiling ? 1 : 0);
    if (st != VACO_AV1PF_OK)
        return st;
    *out = best;
    return ps;
}

extern "C" vaco_av1pf_status vaco_av1pf_set_chain_spin_limit(vaco_av1pf_ctx *ctx, uint32_t polls)
{
    if (!ctx)
        return VACO_AV1PF_INVALID_ARGUMENT;
    ctx->chain_spin_limit = polls;
    return VACO_AV1PF_OK;
}

// ------------------------------------------------ static-scratch entry --

// The per-frame stage with every device buffer it needs inside `scratch`
// (`bytes` of 256-byte aligned device memory): the status word and the
// staged side information and work descriptors in the first
// kStaticStageBytes, the intermediate images after them. Runs on stream s
// and waits for it; allocates no device or pinned memory (the staged bytes
// are built in a per-thread host buffer and copied up by the runtime).
vaco_av1pf_status av1pf_frame_static(const vaco_av1pf_frame_desc *desc, const vaco_av1pf_picture *in,
                                     vaco_av1pf_picture *out, hipStream_t s, uint8_t *scratch, size_t bytes,
                                     const char **msg)
{
    *msg = "";
    FrameGeo g;
    frame_geometry(*desc, &g);
    vaco_av1pf_status st = check_geometry(g, desc->cdef_stride, in, out, msg);
    if (st != VACO_AV1PF_OK)
        return st;
    const size_t inter = scratch_need(*desc, g) * sizeof(uint16_t);
    if (arena_need(*desc, g) > kStaticStageBytes || bytes < kStaticStageBytes ||
        inter > bytes - kStaticStageBytes) {
        *msg = "frame needs more device memory than the static scratch holds";
        return VACO_AV1PF_LIMIT_EXCEEDED;
    }
    if (!params_clean(*desc, g)) {
        // Parameter errors need the samples to order them exactly.
        if (hipStreamSynchronize(s) != hipSuccess) {
            *msg = "stream synchronisation failed";
            return VACO_AV1PF_DEVICE_ERROR;
        }
        const vaco_av1pf_status r = av1pf_replay_device_frame(nullptr, *desc, *in, msg);
        if (r != VACO_AV1PF_OK)
            return r;
    }
    thread_local std::vector<uint8_t> host;
    host.resize(kStaticStageBytes);
    Arena a;
    a.h = host.data();
    a.d = scratch;
    a.cap = kStaticStageBytes;
    a.pageable = true;
    uint32_t *dstatus = nullptr;
    uint32_t *hstatus = a.take<uint32_t>(1, &dstatus);
    *hstatus = 0;  // cleared by the upload itself
    Scratch ws;
    ws.base = reinterpret_cast<uint16_t *>(scratch + kStaticStageBytes);
    ws.cap = (bytes - kStaticStageBytes) / sizeof(uint16_t);
    Plan plan;
    DevPlane din[3], dout[3];
    for (int p = 0; p < g.nplanes; p++) {
        din[p] = DevPlane{in->plane[p].data, in->plane[p].stride};
        dout[p] = DevPlane{out->plane[p].data, out->plane[p].stride};
    }
    st = plan_frame(nullptr, *desc, g, din, dout, &a, dstatus, ws, plan);
    if (st == VACO_AV1PF_OK)
        st = launch_plan(nullptr, plan, &a, s);
    if (st != VACO_AV1PF_OK) {
        *msg = st == VACO_AV1PF_DEVICE_ERROR ? "kernel launch failed" : "static scratch exhausted";
        (void)hipStreamSynchronize(s);
        return st;
    }
    uint32_t flag = 0;
    hipError_t e = hipMemcpyAsync(&flag, dstatus, sizeof flag, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        *msg = hipGetErrorString(e);
        return VACO_AV1PF_DEVICE_ERROR;
    }
    if (flag & kFlagSampleRange)
        return av1pf_replay_device_frame(nullptr, *desc, *in, msg);
    return VACO_AV1PF_OK;
}

// ------------------------------------------------- kernel-level entries --

extern "C" vaco_av1pf_status vaco_av1pf_cdef_blocks(vaco_av1pf_ctx *ctx, uint32_t n,
                                                    const uint16_t *inputs,
                                                    const vaco_av1pf_filter_params *params,
                                                    uint8_t *dirs, uint32_t *vars, uint16_t *out)
{
    if (!ctx || (n && (!inputs || !params || !dirs || !vars || !out)))
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    ctx->err = "";
    // cdef.rs:79-87 and 142-154, per case in order.
    for (uint32_t i = 0; i < n; i++) {
        const char *msg = nullptr;
        if (!check_find_direction(inputs + (size_t)i * 144, params[i].bit_depth, &msg) ||
            !check_filter_block(inputs + (size_t)i * 144, params[i], &msg))
            return fail(ctx, VACO_AV1PF_INVALID_DATA, msg);
    }
    if (n == 0)
        return VACO_AV1PF_OK;
    HIPCHECK(hipSetDevice(ctx->device));
    static_assert(sizeof(vaco_av1pf_filter_params) == 16, "filter params layout");
    const size_t bin = (size_t)n * 144 * 2, bpar = (size_t)n * 16, bout = (size_t)n * 64 * 2;
    uint8_t *buf = nullptr;
    HIPCHECK(hipMalloc(&buf, bin + bpar + bout + n * 4 + n));
    uint16_t *din = (uint16_t *)buf;
    void *dpar = buf + bin;
    uint16_t *dout = (uint16_t *)(buf + bin + bpar);
    uint32_t *dvar = (uint32_t *)(buf + bin + bpar + bout);
    uint8_t *ddir = buf + bin + bpar + bout + n * 4;
    hipError_t e = hipMemcpyAsync(din, inputs, bin, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dpar, params, bpar, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_cdef_cases(din, dpar, ddir, dvar, dout, n, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, bout, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(vars, dvar, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dirs, ddir, n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return hip_fail(ctx, e);
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_upscale_plane(vaco_av1pf_ctx *ctx, const vaco_av1pf_plane *in,
                                                      const vaco_av1pf_sr_plane_config *cfg,
                                                      vaco_av1pf_plane *out)
{
    if (!ctx || !in || !cfg || !out || !in->data || !out->data)
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    ctx->err = "";
    const char *msg = nullptr;
    if (!check_superres_plane(*cfg, in->width, in->height, &msg))  // superres.rs:128-143
        return fail(ctx, VACO_AV1PF_INVALID_DATA, msg);
    if (out->width != cfg->output_width || out->height != cfg->height || out->stride < out->width ||
        in->stride < in->width)
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "plane geometry");
    // k_superres keeps source positions in 32 bits: x * step < 2^16 * 2^14.
    if (cfg->output_width > kMaxSuperresWidth)
        return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "superres output wider than 65536 samples");
    HIPCHECK(hipSetDevice(ctx->device));
    const int rows_per_block = ctx->tune.superres_rows;
    const size_t nin = (size_t)in->width * cfg->height, nout = (size_t)cfg->output_width * cfg->height;
    uint8_t *buf = nullptr;
    HIPCHECK(hipMalloc(&buf, (nin + nout) * 2 + 512));
    uint16_t *din = (uint16_t *)buf, *dout = din + nin;
    uint8_t *tail = (uint8_t *)(dout + nout);
    SrJob *dj = (SrJob *)(((uintptr_t)tail + 255) & ~(uintptr_t)255);
    uint32_t *dst = (uint32_t *)(dj + 1);
    SrJob j;
    memset(&j, 0, sizeof j);
    int32_t step, start;
    superres_phase(cfg->visible_width, cfg->output_width, &step, &start);
    j.in = din;
    j.out = dout;
    j.status = dst;
    j.rows = nullptr;
    j.in_width = in->width;
    j.in_stride = in->width;
    j.out_stride = cfg->output_width;
    j.out_width = cfg->output_width;
    j.n_rows = cfg->height;
    j.segs = (cfg->output_width + kSrSeg - 1) / kSrSeg;
    j.step = step;
    j.start = start;
    j.block_begin = 0;
    j.max_sample = (uint16_t)((1u << cfg->bit_depth) - 1);
    uint32_t zero = 0, flag = 0;
    hipError_t e = hipMemcpy2DAsync(din, in->width * 2, in->data, in->stride * 2, in->width * 2,
                                    cfg->height, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dj, &j, sizeof j, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst, &zero, 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_superres(dj, 1, sr_job_blocks(j.n_rows, j.segs, rows_per_block), rows_per_block,
                            ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(out->data, out->stride * 2, dout, cfg->output_width * 2,
                             cfg->output_width * 2, cfg->height, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&flag, dst, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return hip_fail(ctx, e);
    if (flag & kFlagSampleRange)  // superres.rs:201-203
        return fail(ctx, VACO_AV1PF_INVALID_DATA, "AV1 superres sample exceeds bit depth");
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_restore_plane(vaco_av1pf_ctx *ctx,
                                                      const vaco_av1pf_plane *before,
                                                      const vaco_av1pf_plane *after,
                                                      const vaco_av1pf_lr_plane_config *cfg,
                                                      const vaco_av1pf_lr_unit *units,
                                                      uint32_t n_units, vaco_av1pf_plane *out)
{
    if (!ctx || !before || !after || !cfg || !out || !before->data || !after->data || !out->data ||
        (n_units && !units))
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "null argument");
    ctx->err = "";
    HostPlane hb{before->data, before->width, before->height, before->stride};
    HostPlane ha{after->data, after->width, after->height, after->stride};
    const char *msg = nullptr;
    if (!check_restore_plane(*cfg, hb, ha, units, n_units, &msg))  // restoration.rs:135-168
        return fail(ctx, VACO_AV1PF_INVALID_DATA, msg);
    if (out->width != cfg->width || out->height != cfg->height || out->stride < out->width)
        return fail(ctx, VACO_AV1PF_INVALID_ARGUMENT, "plane geometry");
    if (cfg->width > 65536 || cfg->height > 65536)  // tile list entries hold 16-bit segment indices
        return fail(ctx, VACO_AV1PF_LIMIT_EXCEEDED, "restoration plane larger than 65536 samples a side");
    HIPCHECK(hipSetDevice(ctx->device));
    std::vector<LrRowSeg> segs;
    std::vector<LrColSeg> cols;
    lr_row_segments(cfg->height, cfg->unit_size, cfg->subsampling_y != 0, segs);
    lr_col_segments(cfg->width, cfg->unit_size, cols);
    // tiles cut to the tuned height, as the frame path cuts them
    const vaco_av1pf_tuning tune = ctx->tune;
    std::vector<LrRowSeg> rows;
    lr_cut_rows(segs, tune.lr_tile_rows, rows);
    const size_t nb = (size_t)before->height * before->width, na = (size_t)after->height * after->width;
    const size_t no = (size_t)cfg->width * cfg->height;
    uint32_t ucols, urows;
    lr_unit_counts(cfg->width, cfg->height, cfg->unit_size, &ucols, &urows);
    // tile lists by kernel kind (0 copy, 1 Wiener, 2 self-guided)
    std::vector<uint64_t> tiles[3];
    for (uint32_t ri = 0; ri < rows.size(); ri++)
        for (uint32_t ci = 0; ci < cols.size(); ci++) {
            const uint8_t k = units[rows[ri].unit_row * ucols + cols[ci].unit_col].kind;
            const int kind = k == VACO_AV1PF_UNIT_WIENER ? 1 : k == VACO_AV1PF_UNIT_SELF_GUIDED ? 2 : 0;
            tiles[tune.lr_order ? 0 : kind].push_back(lr_entry(ri, ci, 0, kind));
        }
    const size_t ntiles = rows.size() * cols.size();
    const size_t bytes_tail = rows.size() * sizeof(LrRowSeg) + cols.size() * sizeof(LrColSeg) +
                              n_units * sizeof(LrUnit) + sizeof(LrPlane) + ntiles * 8 + 10 * 256;
    uint8_t *buf = nullptr;
    HIPCHECK(hipMalloc(&buf, (nb + na + no) * 2 + bytes_tail));
    uint16_t *db = (uint16_t *)buf, *da = db + nb, *dout = da + na;
    uint8_t *p8 = (uint8_t *)(((uintptr_t)(dout + no) + 255) & ~(uintptr_t)255);
    LrPlane *dl = (LrPlane *)p8;
    p8 += (sizeof(LrPlane) + 255) & ~size_t(255);
    LrRowSeg *drs = (LrRowSeg *)p8;
    p8 += (rows.size() * sizeof(LrRowSeg) + 255) & ~size_t(255);
    LrColSeg *dcs = (LrColSeg *)p8;
    p8 += (cols.size() * sizeof(LrColSeg) + 255) & ~size_t(255);
    LrUnit *du = (LrUnit *)p8;
    p8 += (n_units * sizeof(LrUnit) + 255) & ~size_t(255);
    // one list for k_lr (lr_tile_list)
    uint64_t *dlist = (uint64_t *)p8;
    uint32_t nlist[3];
    for (int k = 0; k < 3; k++)
        nlist[k] = (uint32_t)tiles[k].size();
    std::vector<uint64_t> list(ntiles);
    lr_tile_list(tiles, tune.lr_order == 0, list.data());
    p8 += (ntiles * 8 + 255) & ~size_t(255);
    uint32_t *dst = (uint32_t *)p8;
    LrPlane L;
    memset(&L, 0, sizeof L);
    L.before = db;
    L.after = da;
    L.out = dout;
    L.rows = drs;
    L.cols = dcs;
    L.units = du;
    L.status = dst;
    L.before_stride = before->width;
    L.after_stride = after->width;
    L.out_stride = cfg->width;
    L.width = cfg->width;
    L.height = cfg->height;
    L.n_rows = (uint32_t)rows.size();
    L.n_cols = (uint32_t)cols.size();
    L.unit_cols = ucols;
    L.block_begin = 0;
    L.bit_depth = cfg->bit_depth;
    L.ss_y = cfg->subsampling_y ? 1 : 0;
    uint32_t zero = 0;
    hipError_t e = hipMemcpy2DAsync(db, before->width * 2, before->data, before->stride * 2,
                                    before->width * 2, before->height, hipMemcpyHostToDevice,
                                    ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(da, after->width * 2, after->data, after->stride * 2, after->width * 2,
                             after->height, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dl, &L, sizeof L, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(drs, rows.data(), rows.size() * sizeof(LrRowSeg), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dcs, cols.data(), cols.size() * sizeof(LrColSeg), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess && n_units)
        e = hipMemcpyAsync(du, units, n_units * sizeof(LrUnit), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !list.empty() && tune.lr_order == 0)
        e = hipMemcpyAsync(dlist, list.data(), list.size() * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst, &zero, 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_lr(dl, 1, tune.lr_order == 0 ? dlist : nullptr, nlist[0] + nlist[1] + nlist[2],
                      tune.lr_order != 0, tune.lr_tile_rows,
                      ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(out->data, out->stride * 2, dout, cfg->width * 2, cfg->width * 2,
                             cfg->height, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return hip_fail(ctx, e);
    return VACO_AV1PF_OK;
}
