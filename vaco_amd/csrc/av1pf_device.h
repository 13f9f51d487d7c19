// av1pf_device.h - work descriptors shared by the host runtime and the
// gfx950 kernels of the AV1 post-reconstruction filter stage.
//
// Every launch receives an array of work items (frames, planes or plane
// jobs). Each item owns a contiguous range of workgroups starting at
// `block_begin`; a workgroup finds its item by binary search. The same
// kernels therefore run one frame (per-frame path) or many frames of mixed
// geometry and bit depth (batched path) per launch.
#pragma once

#include <stdint.h>

namespace av1pf {

// Sentinel stored in LDS for positions outside a plane (CDEF taps there are
// "unavailable", cdef.rs:161-166 / decode.rs:828-831). Valid samples are at
// most 4095, so the value never collides with in-range data.
constexpr uint16_t kUnavailable = 0xFFFF;

// RestorationUnit discriminants (restoration.rs:46-65), as in the C ABI.
constexpr uint8_t kUnitNone = 0, kUnitWiener = 1, kUnitSelfGuided = 2;

// Bit set in a frame's status word when a kernel read a sample above
// (1 << bit_depth) - 1. The host then replays the reference's checks in
// order to report the exact error (or none).
constexpr uint32_t kFlagSampleRange = 1u;

// ----------------------------------------------------------------- CDEF --

// One frame's CDEF work: one workgroup per 64x64 luma superblock, all planes
// (decode.rs:739-861).
struct CdefFrame {
    const uint16_t *in[3];
    uint16_t *out[3];
    uint32_t width[3], height[3];      // padded plane geometry
    uint32_t in_stride[3], out_stride[3];
    const int8_t *cdef_idx;            // per 64x64, -1 = none
    const uint8_t *skip8;              // per 8x8 luma block, nonzero = skipped
    uint32_t idx_stride, skip_stride;
    uint32_t sb_cols, sb_rows;
    uint32_t *status;                  // kFlag* bits
    uint32_t block_begin;
    uint8_t bit_depth, ss_x, ss_y, nplanes;
    uint8_t damping;                   // CdefDamping (includes +3)
    uint8_t vec_ok;                    // inputs 8-byte aligned, strides multiples of 4
    uint8_t out_u8;                    // out[] holds one byte per sample (8-bit frame output)
    uint8_t out_nt;                    // store out[] non-temporally (no superres reads it next)
    uint32_t crop_w[3], crop_h[3];     // stores limited to this region of each output plane
    uint8_t y_pri[8], y_sec[8], uv_pri[8], uv_sec[8];
};

// ------------------------------------------------------------- superres --

// One plane of one image: superres.rs:163-216. Workgroups cover blocks of
// R rows (a launch parameter) x segments of kSrSeg output samples.
constexpr int kSrSeg = 1024;

// Workgroups of a superres job of n_rows rows and segs segments.
inline uint32_t sr_job_blocks(uint32_t n_rows, uint32_t segs, uint32_t rows_per_block)
{
    return (n_rows + rows_per_block - 1) / rows_per_block * segs;
}

struct SrJob {
    const uint16_t *in;
    uint16_t *out;
    uint32_t *status;
    const uint32_t *rows;              // NULL = every row; else row list
    uint32_t in_width;                 // padded width (edge clamp bound)
    uint32_t in_stride, out_stride;
    uint32_t out_width;
    uint32_t n_rows;                   // rows to produce (list length or height)
    uint32_t segs;                     // segments per row
    int32_t step;                      // 14-bit fixed-point step
    int32_t start;                     // initial subpel position
    uint32_t block_begin;
    uint16_t max_sample;
    uint8_t out_u8;                    // out holds one byte per sample (8-bit frame output)
    uint8_t vec_ok;                    // in 8-byte aligned, in_stride a multiple of 4
};

// -------------------------------------------------------- restoration --

// A band of plane rows that lies inside one restoration stripe
// (restoration.rs:218-231) and one unit row (restoration.rs:124-133).
struct LrRowSeg {
    int32_t y0, h;
    int32_t stripe_start, stripe_end;
    uint32_t unit_row;
};

// A band of at most 64 columns inside one unit column.
struct LrColSeg {
    int32_t x0, w;
    uint32_t unit_col;
};

// restoration::RestorationUnit, same layout as vaco_av1pf_lr_unit.
struct LrUnit {
    uint8_t kind, set;
    int16_t xqd[2];
    int16_t vertical[3];
    int16_t horizontal[3];
};

// One plane's restoration work: row segments x column segments.
struct LrPlane {
    const uint16_t *before;            // pre-CDEF (upscaled when superres, unless sr_step)
    const uint16_t *after;             // CDEF output (upscaled when superres, unless sr_step)
    uint16_t *out;
    const LrRowSeg *rows;
    const LrColSeg *cols;
    const LrUnit *units;               // NULL: every unit is None (plain copy)
    uint32_t *status;
    uint32_t before_stride, after_stride, out_stride;
    uint32_t width, height;            // visible geometry
    uint32_t n_rows, n_cols, unit_cols;
    uint32_t block_begin;
    // Superres folded into the staging (superres.rs:163-216): sr_step != 0
    // means `before` and `after` are the images before upscaling, sr_in_width
    // samples wide (padded), and each staged sample is upscaled where it is
    // staged, with this plane's step and initial position.
    int32_t sr_step, sr_start;
    uint32_t sr_in_width;
    uint8_t bit_depth, ss_y;
    uint8_t vec_ok;                    // before/after 8-byte aligned, strides multiples of 4
    uint8_t out_u8;                    // out holds one byte per sample (8-bit frame output)
    uint8_t out_vec;                   // out 8-byte aligned, out_stride a multiple of 4
};

// ------------------------------------------------------ chained launch --

// Ordering state of a chained launch (CDEF, then restoration in the same
// grid; restoration tiles wait for the superblocks they read). Both arrays
// live in device memory owned by the context and are never reset: flags[sb]
// holds the epoch of the last chained launch that finished superblock sb,
// and *copied counts the workgroups, over every chained launch of the
// context, that have finished their share of the staged-bytes copy.
struct ChainSync {
    uint32_t *flags;
    unsigned long long *copied;
    unsigned long long copied_target;  // *copied once this launch's copy is complete
    uint32_t epoch;                    // nonzero, distinct from every earlier launch's
    uint32_t spin_limit;               // polls before a waiting workgroup gives up
};

// Bit set in a frame's status word when a chained launch's restoration tile
// gave up waiting for its superblocks (never expected: workgroups are
// dispatched in order, so every superblock's workgroup is running or done).
constexpr uint32_t kFlagChainStall = 2u;

}  // namespace av1pf
