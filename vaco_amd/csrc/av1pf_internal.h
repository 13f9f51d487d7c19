// av1pf_internal.h - runtime entry points shared by the library's host
// sources (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/vaco_av1pf.h"

// The batched call with optional ownership of the per-frame status checks
// (see av1pf_host.cpp).
vaco_av1pf_status av1pf_enqueue_batch(vaco_av1pf_ctx *ctx, uint32_t n,
                                      const vaco_av1pf_frame_desc *descs,
                                      const vaco_av1pf_picture *ins, vaco_av1pf_picture *outs,
                                      hipStream_t s, const uint32_t **statuses, hipEvent_t *done);

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
