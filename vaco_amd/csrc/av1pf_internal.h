// av1pf_internal.h - runtime entry points shared by the library's host
// sources (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/vaco_av1pf.h"

// Device memory for the intermediate images of one batch, owned by the
// caller (cap in samples).
struct av1pf_workspace {
    uint16_t *base;
    size_t cap;
};

// The batched call with optional ownership of the per-frame status checks
// (see av1pf_host.cpp). own_ws: the batch's intermediates go there, and the
// batch must fit it and one staging slot (nothing is allocated); such calls
// do not order themselves after calls on other streams. slot_index, if not
// NULL, receives the staging slot the batch uses.
vaco_av1pf_status av1pf_enqueue_batch(vaco_av1pf_ctx *ctx, uint32_t n,
                                      const vaco_av1pf_frame_desc *descs,
                                      const vaco_av1pf_picture *ins, vaco_av1pf_picture *outs,
                                      hipStream_t s, const uint32_t **statuses, hipEvent_t *done,
                                      const av1pf_workspace *own_ws, int *slot_index);

// Staging bytes (one slot) and intermediate samples one frame needs in a
// batch, and the same for the largest frame the context accepts.
size_t av1pf_frame_stage_bytes(const vaco_av1pf_frame_desc &d);
size_t av1pf_frame_scratch_samples(const vaco_av1pf_frame_desc &d);
size_t av1pf_worst_frame_stage_bytes(const vaco_av1pf_ctx *ctx);
size_t av1pf_worst_frame_scratch_samples(const vaco_av1pf_ctx *ctx);
// Grow every staging slot to at least `bytes` (waiting for their work);
// the smallest slot's capacity; the slot the next call takes.
vaco_av1pf_status av1pf_reserve_staging(vaco_av1pf_ctx *ctx, size_t bytes);
size_t av1pf_staging_capacity(const vaco_av1pf_ctx *ctx);
int av1pf_next_slot(const vaco_av1pf_ctx *ctx);
// The batch in `slot` has completed and its owner has read its status
// words: the slot no longer refers to the batch's stream or event, so the
// stream may be destroyed.
void av1pf_settle_slot(vaco_av1pf_ctx *ctx, int slot);
// Staging bytes of a batch beyond its frames' own (descriptors, rounding).
constexpr size_t kBatchStageBase = 1 << 16;

// The reference's outcome for a frame whose kernels raised the sample-range
// flag, from a host copy of its device input.
vaco_av1pf_status av1pf_replay_device_frame(vaco_av1pf_ctx *ctx, const vaco_av1pf_frame_desc &d,
                                            const vaco_av1pf_picture &in, const char **msg);

// Device bytes a frame holds while in flight: input and output planes plus
// the intermediates the stage allocates for it.
size_t av1pf_frame_device_bytes(const vaco_av1pf_frame_desc &d);

int av1pf_ctx_device(const vaco_av1pf_ctx *ctx);

// The per-frame stage inside caller-provided device memory `scratch`
// (`bytes` long, 256-byte aligned); synchronous on stream s. *msg receives
// the message of a failure.
vaco_av1pf_status av1pf_frame_static(const vaco_av1pf_frame_desc *desc, const vaco_av1pf_picture *in,
                                     vaco_av1pf_picture *out, hipStream_t s, uint8_t *scratch, size_t bytes,
                                     const char **msg);
