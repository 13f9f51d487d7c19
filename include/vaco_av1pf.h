/*
 * vaco_av1pf.h - C ABI of the MI355X (gfx950) AV1 post-reconstruction filter
 * stage: CDEF -> superres -> loop restoration, on 8/10/12-bit planes.
 *
 * This is the drop-in boundary for the stage that matteopolak/vaco runs as
 * three private per-frame calls in
 *   crates/codec/vaco-codec-av1/src/decode.rs:554-559
 *     (copy_picture, apply_cdef, apply_superres, apply_restoration),
 * and for the public kernel functions of the same crate
 *   src/cdef.rs:79-230, src/superres.rs:163-274, src/restoration.rs:108-207.
 *
 * Conventions
 *  - Plain C types and pointers only; no torch or HIP types in signatures
 *    (streams are passed as `void*` holding a hipStream_t, NULL = the
 *    context's own stream).
 *  - Planes are u16 row-major, exactly like framebuf::Plane
 *    (framebuf.rs:14-20), with an explicit stride in samples (>= width).
 *  - Every entry point returns a vaco_av1pf_status; the human-readable
 *    message of the last failure on a context is vaco_av1pf_last_error().
 *    Messages equal the reference's `Error::InvalidData(&'static str)` text
 *    where the reference has one.
 *  - Nothing aborts or throws across this boundary.
 *  - A context is not internally synchronised: use one per host thread
 *    (the reference threads `&mut Budget` through a single thread).
 */
#ifndef VACO_AV1PF_H
#define VACO_AV1PF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VACO_AV1PF_ABI_VERSION 3

/* vaco_core::Error classes (crates/core/vaco-core/src/error.rs:14-59). */
typedef enum vaco_av1pf_status {
    VACO_AV1PF_OK = 0,
    VACO_AV1PF_INVALID_DATA = 1,    /* Error::InvalidData(msg)                 */
    VACO_AV1PF_UNSUPPORTED = 2,     /* Error::Unsupported(msg)                 */
    VACO_AV1PF_LIMIT_EXCEEDED = 3,  /* Error::LimitExceeded (fuel, capacity)   */
    VACO_AV1PF_DEVICE_ERROR = 4,    /* HIP runtime failure (no reference peer) */
    VACO_AV1PF_INVALID_ARGUMENT = 5 /* NULL pointer / bad handle (Rust
                                       references cannot be null, so the
                                       reference has no peer for this)        */
} vaco_av1pf_status;

typedef struct vaco_av1pf_ctx vaco_av1pf_ctx;

/* ---------------------------------------------------------------- CDEF --- */

/* cdef::Strength, cdef.rs:14-20 (8-bit units; secondary in {0,1,2,4}). */
typedef struct vaco_av1pf_strength {
    uint8_t primary;
    uint8_t secondary;
} vaco_av1pf_strength;

/* cdef::CdefParams, cdef.rs:23-33. `damping` is CdefDamping (includes the
 * syntax's +3). `enabled` = 0 disables CDEF for the frame. */
typedef struct vaco_av1pf_cdef_params {
    uint8_t enabled;
    uint8_t damping;
    vaco_av1pf_strength y[8];
    vaco_av1pf_strength uv[8];
} vaco_av1pf_cdef_params;

/* cdef::FilterParams, cdef.rs:35-54. */
typedef struct vaco_av1pf_filter_params {
    uint32_t width;      /* 4 or 8 */
    uint32_t height;     /* 4 or 8 */
    uint8_t edges;       /* left=1 right=2 top=4 bottom=8 */
    uint8_t bit_depth;   /* 8, 10, 12 */
    uint8_t direction;   /* 0..7 */
    uint8_t damping;     /* plane damping incl. bit-depth and chroma adjust */
    uint16_t primary;    /* after variance adjustment and <<(bd-8) */
    uint16_t secondary;  /* after <<(bd-8) */
} vaco_av1pf_filter_params;

/* ---------------------------------------------------------- superres --- */

/* superres::PlaneConfig, superres.rs:93-103. */
typedef struct vaco_av1pf_sr_plane_config {
    uint32_t visible_width;
    uint32_t output_width;
    uint32_t height;
    uint8_t bit_depth;
} vaco_av1pf_sr_plane_config;

/* ------------------------------------------------------- restoration --- */

/* restoration::RestorationType, restoration.rs:15-27 (frame-level mode). */
enum {
    VACO_AV1PF_RESTORE_NONE = 0,
    VACO_AV1PF_RESTORE_SWITCHABLE = 1,
    VACO_AV1PF_RESTORE_WIENER = 2,
    VACO_AV1PF_RESTORE_SELF_GUIDED = 3
};

/* restoration::RestorationUnit discriminant, restoration.rs:46-65. */
enum {
    VACO_AV1PF_UNIT_NONE = 0,
    VACO_AV1PF_UNIT_WIENER = 1,
    VACO_AV1PF_UNIT_SELF_GUIDED = 2
};

/* restoration::RestorationUnit, restoration.rs:46-65. Fields not used by
 * `kind` are ignored. Coefficient ranges are validated exactly as
 * restoration.rs:67-91 does. */
typedef struct vaco_av1pf_lr_unit {
    uint8_t kind;           /* VACO_AV1PF_UNIT_* */
    uint8_t set;            /* self-guided: lr_sgr_set, 0..15 */
    int16_t xqd[2];         /* self-guided: LrSgrXqd */
    int16_t vertical[3];    /* Wiener: outermost first */
    int16_t horizontal[3];  /* Wiener: outermost first */
} vaco_av1pf_lr_unit;

/* restoration::PlaneConfig, restoration.rs:93-106. */
typedef struct vaco_av1pf_lr_plane_config {
    uint32_t width;
    uint32_t height;
    uint32_t unit_size;     /* 32, 64, 128, 256 */
    uint8_t bit_depth;
    uint8_t subsampling_y;
} vaco_av1pf_lr_plane_config;

/* ------------------------------------------------------------ planes --- */

/* framebuf::Plane, framebuf.rs:14-20, with an explicit stride. `data` is a
 * host pointer for the host entry points and a device pointer for the
 * *_device entry points. */
typedef struct vaco_av1pf_plane {
    uint16_t *data;
    uint32_t width;
    uint32_t height;
    uint32_t stride;        /* samples between rows, >= width */
} vaco_av1pf_plane;

/* framebuf::Picture, framebuf.rs:78-83. plane[1], plane[2] unused (data may
 * be NULL) when the frame is monochrome. */
typedef struct vaco_av1pf_picture {
    vaco_av1pf_plane plane[3];
} vaco_av1pf_picture;

/* Everything decode_frame() holds for the filter stage (decode.rs:506-561,
 * FrameCtx decode.rs:185-214, FrameSize frame_header.rs:71-85).
 *
 * Input picture geometry is the Mi-padded reconstruction (decode.rs:518-530):
 *   mi_cols = 2*((coded_width+7)>>3), mi_rows = 2*((coded_height+7)>>3),
 *   luma mi_cols*4 x mi_rows*4, chroma (luma_w>>ss_x) x (luma_h>>ss_y).
 * Output picture geometry is the reference's final ctx.pic
 * (vaco_av1pf_frame_output_geometry computes it). */
typedef struct vaco_av1pf_frame_desc {
    uint32_t coded_width;
    uint32_t upscaled_width;    /* == coded_width unless use_superres */
    uint32_t coded_height;
    uint8_t use_superres;
    uint8_t bit_depth;          /* 8, 10, 12 */
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    uint8_t monochrome;
    vaco_av1pf_cdef_params cdef;
    /* cdef_idx per 64x64 luma unit, ceil(mi_rows/16) rows of cdef_stride
     * entries (cdef_stride = ceil(mi_cols/16)); -1 = not read (decode.rs:543,
     * 774-776). May be NULL when cdef.enabled == 0. */
    const int8_t *cdef_idx;
    uint32_t cdef_stride;
    /* One byte per 8x8 luma block, (mi_rows/2) rows of (mi_cols/2): nonzero
     * when all four 4x4 MiCells of the block are skip (decode.rs:892-896;
     * vaco_av1pf_skip8_from_mi derives it from a MiCell skip grid). */
    const uint8_t *skip8;
    /* FrameRestoration, restoration.rs:29-44. */
    uint8_t lr_type[3];         /* VACO_AV1PF_RESTORE_* */
    uint32_t lr_unit_size[3];
    /* Unit maps per plane, row-major, unit_counts() long (decode.rs:612-638) */
    const vaco_av1pf_lr_unit *lr_units[3];
    uint32_t lr_unit_count[3];
} vaco_av1pf_frame_desc;

/* ------------------------------------------------------ context/errors --- */

/* Create a context on HIP device `device`. Device buffers for frames up to
 * max_width x max_height luma samples (padded or upscaled, whichever is
 * larger) are allocated here, so the per-frame calls allocate nothing.
 * No reference peer: the reference allocates per stage through Budget. */
vaco_av1pf_status vaco_av1pf_create(int device, uint32_t max_width,
                                    uint32_t max_height, vaco_av1pf_ctx **out);
void vaco_av1pf_destroy(vaco_av1pf_ctx *ctx);
/* Static string describing the last failure on ctx ("" if none). */
const char *vaco_av1pf_last_error(const vaco_av1pf_ctx *ctx);
/* ABI version of the loaded library (VACO_AV1PF_ABI_VERSION). */
uint32_t vaco_av1pf_abi_version(void);

/* ------------------------------------------------- whole-stage entries --- */

/* Output plane geometry of the stage for `desc`: the reference's final
 * ctx.pic after apply_cdef/apply_superres/apply_restoration
 * (decode.rs:734, 857, 887). Host-only, no GPU. */
vaco_av1pf_status vaco_av1pf_frame_output_geometry(
    const vaco_av1pf_frame_desc *desc, uint32_t width[3], uint32_t height[3]);

/* Input (Mi-padded) plane geometry for `desc` (decode.rs:518-530). */
vaco_av1pf_status vaco_av1pf_frame_input_geometry(
    const vaco_av1pf_frame_desc *desc, uint32_t width[3], uint32_t height[3]);

/* The whole stage on host planes: replaces decode.rs:554-559 as a unit
 * (copy_picture + apply_cdef + apply_superres + apply_restoration).
 * `in` is read only; `out` receives the final picture. `fuel`, if not NULL,
 * is the Budget fuel still available (budget.rs:229-237): on return it holds
 * the remaining fuel; exhaustion returns VACO_AV1PF_LIMIT_EXCEEDED at the
 * same point the reference would. Synchronous. */
vaco_av1pf_status vaco_av1pf_frame(vaco_av1pf_ctx *ctx,
                                   const vaco_av1pf_frame_desc *desc,
                                   const vaco_av1pf_picture *in,
                                   vaco_av1pf_picture *out, uint64_t *fuel);

/* The same stage on device-resident planes, enqueued on `stream`
 * (hipStream_t; NULL = the context's stream). Host-side validation errors
 * return at once; sample-range errors found by the kernels are reported by
 * the next vaco_av1pf_finish() on this context. The side-information arrays
 * `desc` points to and the input planes must stay unchanged until
 * vaco_av1pf_finish() returns: an error is replayed from them. Up to 8
 * frames are in flight per context; the 9th call waits for the 1st. Calls
 * share the context's intermediate images, so a call on another stream than
 * the previous call's waits, on the device, for that call's work: calls on
 * different streams are correct but do not overlap. */
vaco_av1pf_status vaco_av1pf_frame_device(vaco_av1pf_ctx *ctx,
                                          const vaco_av1pf_frame_desc *desc,
                                          const vaco_av1pf_picture *in,
                                          vaco_av1pf_picture *out,
                                          void *stream);

/* Batched form of vaco_av1pf_frame_device: n (<= 256) frames of any
 * geometry, bit depth and chroma layout, each with its own desc and
 * device planes. All frames' work descriptors and side information go up
 * in one copy, and each stage kernel is launched once over the tiles of
 * every frame. Output is identical to n vaco_av1pf_frame_device calls.
 * Same lifetime rule, deferred errors (the first failing frame's error is
 * reported) and cross-stream ordering as vaco_av1pf_frame_device. The
 * intermediates of a batch live in a context workspace that grows (one
 * device allocation, after the context's work has finished) when a batch
 * needs more than any before it; the server below sizes its own at creation
 * and never allocates afterwards. No reference peer: the reference filters
 * one frame per decode_frame call. */
vaco_av1pf_status vaco_av1pf_batch_device(vaco_av1pf_ctx *ctx, uint32_t n,
                                          const vaco_av1pf_frame_desc *descs,
                                          const vaco_av1pf_picture *ins,
                                          vaco_av1pf_picture *outs, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vaco_av1pf_frame_device without a context. It allocates
 * no device or pinned memory: every device buffer it needs (status word,
 * staged side information, intermediate images) lives in one file-scope
 * device array of VACO_AV1PF_STATIC_SCRATCH_BYTES that each GPU holds from
 * the library's first use on it until the process ends. It runs on the
 * calling thread's current device and is synchronous: it returns once the
 * output planes are written, with the status and, in *msg (may be NULL),
 * the message vaco_av1pf_frame would report. Frames whose staging exceeds
 * 2 MiB or whose intermediates exceed the rest of the array (a 4:2:0 frame
 * with restoration up to 3840x2176, with or without superres when every
 * plane restores; not a 4:4:4 frame of that size) return
 * VACO_AV1PF_LIMIT_EXCEEDED before any device work. Every call on a
 * GPU shares that GPU's array and leaves the last call's contents in it; the
 * library serialises this process's calls on one GPU, and calls from other
 * processes use their own copy. */
#define VACO_AV1PF_STATIC_SCRATCH_BYTES (32u << 20)
vaco_av1pf_status vaco_av1pf_frame_device_static(const vaco_av1pf_frame_desc *desc,
                                                 const vaco_av1pf_picture *in,
                                                 vaco_av1pf_picture *out, void *stream,
                                                 const char **msg);

/* ------------------------------------------------------ frame output --- */

/* Output planes in the layout of the decoder's Frame (pic_to_frame,
 * decode.rs:3795-3852): one byte per sample at 8-bit, two (little-endian)
 * above; rows stride_bytes apart. */
typedef struct vaco_av1pf_frame_planes {
    void *data[3];
    uint32_t stride_bytes[3];
} vaco_av1pf_frame_planes;

/* pic_to_frame's plane sizes (decode.rs:3821-3829): luma upscaled_width x
 * coded_height, chroma rounded up; and the bytes per sample. */
vaco_av1pf_status vaco_av1pf_frame_layout(const vaco_av1pf_frame_desc *desc, uint32_t width[3],
                                          uint32_t height[3], uint32_t *bytes_per_sample);

/* decode.rs:554-560 as one call: the stage on device planes, with the final
 * stage storing straight into frame planes as pic_to_frame and its blit
 * (decode.rs:3863-3891) would write them: cropped to the frame, converted to
 * bytes at 8-bit. This removes the conversion pass and halves the output
 * bytes of 8-bit frames. Frame samples blit would not write (a chroma column
 * or row that restoration's rounded-down planes do not reach) are left as
 * they are. Asynchronous like vaco_av1pf_frame_device. */
vaco_av1pf_status vaco_av1pf_frame_device_to_frame(vaco_av1pf_ctx *ctx,
                                                   const vaco_av1pf_frame_desc *desc,
                                                   const vaco_av1pf_picture *in,
                                                   const vaco_av1pf_frame_planes *out, void *stream);

/* ------------------------------------------------ parameter producers --- */

/* The uncompressed-header syntax that produces the stage's frame-level
 * parameters, for decoders without their own header parser. Bits are read
 * MSB first; reads past the end return zeros and set `overrun`
 * (vaco-bitstream/src/reader.rs:72-210). */
typedef struct vaco_av1pf_bit_reader {
    const uint8_t *data;
    uint64_t size_bits;
    uint64_t pos;      /* bits consumed, including any past the end */
    uint32_t overrun;  /* 1 once a read went past the end */
} vaco_av1pf_bit_reader;

/* The sequence-header fields the syntax depends on. */
typedef struct vaco_av1pf_seq_info {
    uint8_t enable_cdef, enable_restoration, enable_superres;
    uint8_t use_128x128_superblock;
    uint8_t subsampling_x, subsampling_y;
} vaco_av1pf_seq_info;

void vaco_av1pf_bits_init(vaco_av1pf_bit_reader *r, const uint8_t *data, size_t len);
uint32_t vaco_av1pf_bits_get(vaco_av1pf_bit_reader *r, uint32_t n);
/* frame_header.rs:1034-1072 (cdef_params): fills *out (zeroed, enabled = 0
 * when CDEF is off, which reads no bits) and returns cdef_bits. */
uint32_t vaco_av1pf_parse_cdef_params(vaco_av1pf_bit_reader *r, const vaco_av1pf_seq_info *seq,
                                      int coded_lossless, int allow_intrabc, uint32_t num_planes,
                                      vaco_av1pf_cdef_params *out);
/* frame_header.rs:1074-1123 (lr_params): types as VACO_AV1PF_RESTORE_*
 * values, unit sizes in samples (0 when restoration is unused). */
void vaco_av1pf_parse_lr_params(vaco_av1pf_bit_reader *r, const vaco_av1pf_seq_info *seq,
                                int all_lossless, int allow_intrabc, uint32_t num_planes,
                                uint8_t types[3], uint32_t unit_size[3]);
/* frame_header.rs:687-704 (superres_params): returns the coded width. */
uint32_t vaco_av1pf_parse_superres_params(vaco_av1pf_bit_reader *r, const vaco_av1pf_seq_info *seq,
                                          uint32_t upscaled_width, uint8_t *use_superres);

/* The AV1 symbol decoder (spec 8.2; crates/signal/vaco-codec-msac/src/
 * av1.rs:27-190), one per tile: the entropy-coded reader of the per-block
 * side info below (cdef_idx, restoration units). */
typedef struct vaco_av1pf_symbol_decoder {
    vaco_av1pf_bit_reader bits;  /* raw bits of the tile */
    uint32_t value;              /* SymbolValue */
    uint32_t range;              /* SymbolRange */
    int64_t max_bits;            /* SymbolMaxBits; negative once padding is read */
    uint32_t disable_cdf_update;
} vaco_av1pf_symbol_decoder;

/* init_symbol(size), av1.rs:48-65. */
void vaco_av1pf_sd_init(vaco_av1pf_symbol_decoder *sd, const uint8_t *data, size_t size,
                        int disable_cdf_update);
/* read_symbol, av1.rs:80-123: cdf holds the n ascending thresholds
 * (cdf[n-1] == 32768) and the adaptation counter cdf[n]; n >= 2. The cdf
 * adapts in place unless disable_cdf_update (av1.rs:195-216). */
uint32_t vaco_av1pf_sd_read_symbol(vaco_av1pf_symbol_decoder *sd, uint16_t *cdf, uint32_t n);
/* read_bool / read_literal(n) (MSB first, low 32 bits kept), av1.rs:129-143. */
uint32_t vaco_av1pf_sd_read_bool(vaco_av1pf_symbol_decoder *sd);
uint32_t vaco_av1pf_sd_read_literal(vaco_av1pf_symbol_decoder *sd, uint32_t n);
/* exit_symbol, av1.rs:152-157: skip the tile's unread trailing bits. */
void vaco_av1pf_sd_exit(vaco_av1pf_symbol_decoder *sd);
/* More than the 14 padding bits consumed, or a raw read past the end
 * (av1.rs:161-165). */
int vaco_av1pf_sd_overrun(const vaco_av1pf_symbol_decoder *sd);

/* Per-tile state of the restoration-unit syntax: its three adaptive CDFs
 * (tables/default_cdf.rs:4725-4731, cdf.rs:128-130) and the reference
 * coefficients the next unit is coded against (decode.rs:1039-1040). */
typedef struct vaco_av1pf_lr_tile_state {
    uint16_t use_wiener[3];
    uint16_t use_sgrproj[3];
    uint16_t restoration_type[4];
    int16_t ref_sgr_xqd[3][2];
    int16_t ref_lr_wiener[3][2][3];
} vaco_av1pf_lr_tile_state;

/* The state at the start of a tile. */
void vaco_av1pf_lr_tile_state_init(vaco_av1pf_lr_tile_state *st);

/* read_restoration_units (decode.rs:1097-1164, with read_restoration_unit,
 * read_wiener_unit, read_sgrproj_unit and the sub-exponential codes,
 * 1166-1386) for the superblock at Mi position (mi_row, mi_col) of block
 * size b_size (BLOCK_64X64 = 12, BLOCK_128X128 = 15): every unit whose
 * top-left corner the superblock covers is read into units[plane][row *
 * unit_cols + col] for each plane whose desc->lr_type is not NONE.
 * desc supplies the geometry (upscaled_width, coded_height), subsampling,
 * monochrome, lr_type, lr_unit_size and, as each units[] length,
 * lr_unit_count. Like the reference it assumes one tile and no superres.
 * Returns VACO_AV1PF_INVALID_DATA with the reference's message in *msg (if
 * msg is not NULL) where the reference returns InvalidData. */
vaco_av1pf_status vaco_av1pf_read_lr_units(vaco_av1pf_symbol_decoder *sd,
                                           vaco_av1pf_lr_tile_state *st,
                                           const vaco_av1pf_frame_desc *desc, uint32_t mi_row,
                                           uint32_t mi_col, uint8_t b_size,
                                           vaco_av1pf_lr_unit *units[3], const char **msg);

/* read_cdef (decode.rs:2087-2145) for a block of size mi_size at Mi
 * position (mi_row, mi_col): when CDEF is enabled, the block is not skip and
 * its 64x64 unit's cdef_idx is still -1, reads cdef_bits bits and stores
 * them in every 64x64 unit the block covers. cdef_idx has cdef_rows rows of
 * cdef_stride entries (-1 = not read yet); entries outside it are ignored,
 * as the reference's checked accesses ignore them. */
void vaco_av1pf_read_cdef(vaco_av1pf_symbol_decoder *sd, int cdef_enabled, uint32_t cdef_bits,
                          int8_t *cdef_idx, uint32_t cdef_stride, uint32_t cdef_rows,
                          uint32_t mi_row, uint32_t mi_col, uint8_t mi_size, int skip);

/* ------------------------------------------------------------ graphs --- */

/* No reference counterpart. With graphs enabled, the kernel chain of a
 * frame-device or batched call is captured once per launch shape (frame
 * geometry and staging slot) as a hipGraph and replayed by later calls of
 * the same shape: one graph launch instead of a launch per kernel. Output
 * is identical either way. Not used while profiling. */
vaco_av1pf_status vaco_av1pf_use_graphs(vaco_av1pf_ctx *ctx, int enable);

/* ------------------------------------------------------------ tuning --- */

/* No reference counterpart: the launch variants an autotuner chooses
 * between. Every variant computes the same output bit for bit; they differ
 * only in work order and workgroup shape. */
typedef struct vaco_av1pf_tuning {
    uint8_t superres_rows;   /* rows per superres workgroup: 4, 8 or 16 */
    uint8_t cdef_xcd_order;  /* 1: superblocks numbered XCD-contiguously */
    uint8_t lr_order;        /* 0: restoration tiles by unit type, self-guided first
                                (a tile list the host builds);
                                1: in plane order, numbered XCD-contiguously (no
                                list: each workgroup finds its tile) */
    uint8_t lr_tile_rows;    /* restoration tile height: 64, 32 or 16 rows */
    uint8_t cdef_waves;      /* waves per CDEF workgroup (one superblock): 16 or 8 */
    uint8_t chain;           /* 1: a single frame with CDEF and restoration but no
                                superres runs both in one launch (restoration tiles
                                wait for the superblocks they read; 8-wave CDEF
                                workgroups, cdef_waves unused); 0: one launch each */
} vaco_av1pf_tuning;

vaco_av1pf_status vaco_av1pf_get_tuning(const vaco_av1pf_ctx *ctx, vaco_av1pf_tuning *out);
/* INVALID_ARGUMENT for values outside the ranges above. Applies to calls
 * made after it returns. */
vaco_av1pf_status vaco_av1pf_set_tuning(vaco_av1pf_ctx *ctx, const vaco_av1pf_tuning *tuning);
/* Runs `n` frames (device planes; one frame as vaco_av1pf_frame_device,
 * several as one vaco_av1pf_batch_device call, so that each path is tuned on
 * its own kernels) `reps` times
 * with each variant of each kernel (after one untimed run), times the
 * kernels by their dispatch events (the median of the `reps` runs counts),
 * keeps the fastest variant per kernel as the context's tuning and returns
 * it in *out. A field leaves its current value only for a variant at least
 * 2% faster, and for the CDEF order that moves more HBM bytes
 * (cdef_xcd_order 0) at least 5% faster. `chain` is decided on one frame by
 * the chained launch's time against the sum of the separate launches', and
 * the orders and tile height are then tuned on the chained launch.
 * Synchronous; the outputs hold the frames' results. */
vaco_av1pf_status vaco_av1pf_autotune(vaco_av1pf_ctx *ctx, uint32_t n,
                                      const vaco_av1pf_frame_desc *descs,
                                      const vaco_av1pf_picture *ins, vaco_av1pf_picture *outs,
                                      uint32_t reps, vaco_av1pf_tuning *out);

/* No reference counterpart (test hook). Sets how many times a restoration
 * tile of a chained launch (tuning.chain) polls for the superblocks it reads
 * before it gives up; a tile that gives up lets the frame fail with
 * DEVICE_ERROR ("... gave up waiting for its CDEF superblocks") instead of
 * waiting on. The default (2^22 polls, about a second) is never reached:
 * workgroups are dispatched in order, so every superblock a tile waits for
 * is being filtered. 0 makes every tile that finds a superblock unfinished
 * give up at once, so the error path can be exercised without a stall. */
vaco_av1pf_status vaco_av1pf_set_chain_spin_limit(vaco_av1pf_ctx *ctx, uint32_t polls);

/* ----------------------------------------------------------- serving --- */

/* No reference peer: a request loop over many streams on one device. The
 * server owns `ctx` while it exists (do not call other entries on it).
 * Requests (device-resident planes) are admitted against an HBM budget,
 * batched dynamically and run with vaco_av1pf_batch_device's kernels:
 *  - a batch launches when fewer than max_inflight batches are on the GPU
 *    and either the GPU is idle, max_batch requests wait, or the oldest
 *    waiting request has waited window_us (continuous batching: requests
 *    arriving while batches run join the next one);
 *  - admission is FIFO; a request is admitted when the device bytes of the
 *    frames in flight plus its own stay within hbm_budget_bytes;
 *  - each batch in flight runs on a stream of its own (max_inflight of
 *    them) with an intermediate-image workspace of its own, so batches
 *    overlap on the device; a batch completes as soon as its own work has.
 *    The HIP runtime maps streams onto a pool of GPU_MAX_HW_QUEUES hardware
 *    queues per process (4 by default) and streams sharing a queue run in
 *    order, so overlap needs a queue per stream: set GPU_MAX_HW_QUEUES to
 *    at least the process's stream count (each context has 1, the server
 *    max_inflight more, plus the caller's own) before the runtime starts.
 * Every device and pinned buffer the server uses is allocated by
 * vaco_av1pf_server_create (one workspace per stream, up to 2 GiB each but
 * at least the largest frame's; staging slots of up to 16 MiB); a batch is
 * closed early rather than outgrow them, and nothing is allocated after
 * creation. Results are identical to the same frame run alone. */
typedef struct vaco_av1pf_server vaco_av1pf_server;

typedef struct vaco_av1pf_server_config {
    uint64_t hbm_budget_bytes;  /* inputs + outputs + intermediates in flight */
    uint32_t window_us;         /* batching window while the GPU is busy */
    uint32_t max_batch;         /* frames per launch, 1..256 */
    uint32_t max_inflight;      /* batches on the GPU at once, 1..4 */
} vaco_av1pf_server_config;

typedef struct vaco_av1pf_completion {
    uint64_t tag;
    int32_t status;       /* vaco_av1pf_status of this frame */
    uint32_t batch_size;  /* frames in the launch that ran it */
    uint64_t t_submit_ns; /* vaco_av1pf_now_ns() clock */
    uint64_t t_launch_ns;
    uint64_t t_done_ns;   /* when the server saw the batch complete */
    /* Device clock, ns since the server was created: when the batch's stream
     * reached its first command, and when its last command finished. Two
     * batches whose [start, end] intervals intersect ran on the device at
     * the same time. */
    uint64_t t_gpu_start_ns;
    uint64_t t_gpu_end_ns;
} vaco_av1pf_completion;

vaco_av1pf_status vaco_av1pf_server_create(vaco_av1pf_ctx *ctx,
                                           const vaco_av1pf_server_config *cfg,
                                           vaco_av1pf_server **out);
/* Queue one frame (device planes; `desc`'s arrays and the planes must stay
 * valid until its completion is polled). Thread-safe. */
vaco_av1pf_status vaco_av1pf_server_submit(vaco_av1pf_server *srv,
                                           const vaco_av1pf_frame_desc *desc,
                                           const vaco_av1pf_picture *in,
                                           vaco_av1pf_picture *out, uint64_t tag);
/* Copy up to `max` completions out; returns how many. Non-blocking. */
uint32_t vaco_av1pf_server_poll(vaco_av1pf_server *srv, vaco_av1pf_completion *out,
                                uint32_t max);
/* Device bytes one frame holds while in flight (what admission counts). */
uint64_t vaco_av1pf_server_frame_bytes(const vaco_av1pf_frame_desc *desc);
/* Block until every submitted request has completed. */
vaco_av1pf_status vaco_av1pf_server_drain(vaco_av1pf_server *srv);
void vaco_av1pf_server_destroy(vaco_av1pf_server *srv);
/* Monotonic clock of the completion timestamps, in nanoseconds. */
uint64_t vaco_av1pf_now_ns(void);

/* Wait for the context's enqueued work and report deferred errors. */
vaco_av1pf_status vaco_av1pf_finish(vaco_av1pf_ctx *ctx);

/* ------------------------------------------- kernel-level GPU entries --- */

/* cdef::find_direction (cdef.rs:79-124) and cdef::filter_block
 * (cdef.rs:142-205) for `n` independent cases on the GPU. inputs: n*144
 * samples (12x12, block at (2,2)); the direction search reads the 8x8 at
 * (2,2). outputs: dirs[n], vars[n], out[n*64] (stride 8, zero outside
 * width x height). Host buffers. */
vaco_av1pf_status vaco_av1pf_cdef_blocks(vaco_av1pf_ctx *ctx, uint32_t n,
                                         const uint16_t *inputs,
                                         const vaco_av1pf_filter_params *params,
                                         uint8_t *dirs, uint32_t *vars,
                                         uint16_t *out);

/* superres::upscale_plane, superres.rs:163-216, on the GPU. Host planes;
 * in->width is the padded width used for edge clamping. out must be
 * output_width x height. */
vaco_av1pf_status vaco_av1pf_upscale_plane(vaco_av1pf_ctx *ctx,
                                           const vaco_av1pf_plane *in,
                                           const vaco_av1pf_sr_plane_config *cfg,
                                           vaco_av1pf_plane *out);

/* restoration::restore_plane, restoration.rs:178-207, on the GPU. Host
 * planes; out must be cfg->width x cfg->height. */
vaco_av1pf_status vaco_av1pf_restore_plane(vaco_av1pf_ctx *ctx,
                                           const vaco_av1pf_plane *before_cdef,
                                           const vaco_av1pf_plane *after_cdef,
                                           const vaco_av1pf_lr_plane_config *cfg,
                                           const vaco_av1pf_lr_unit *units,
                                           uint32_t n_units,
                                           vaco_av1pf_plane *out);

/* --------------------------------------------------------- profiling --- */

/* Kernel identifiers of the profiling counters. */
enum {
    VACO_AV1PF_KERNEL_CDEF = 0,
    VACO_AV1PF_KERNEL_SUPERRES = 1,
    VACO_AV1PF_KERNEL_RESTORATION = 2,
    VACO_AV1PF_KERNEL_UPLOAD = 3,
    VACO_AV1PF_KERNEL_CHAIN = 4, /* CDEF and restoration in one launch (tuning.chain) */
    VACO_AV1PF_KERNEL_COUNT = 5
};

/* No reference counterpart (measurement hook). While enabled, every kernel
 * launch of the frame entries records HIP events at the start and end of
 * its dispatch (the interval a kernel trace reports);
 * vaco_av1pf_profile_read() synchronises the context and returns,
 * per kernel id, the summed launch durations in milliseconds, the launch
 * counts and the algorithmic bytes of those launches since the last read.
 * Algorithmic bytes count every sample a kernel must read once and every
 * sample it writes once, 2 bytes each (CDEF: its planes in and out;
 * superres: visible source rows in, output rows out; restoration: its
 * tiles in and out; upload: the staged side information, read from host
 * memory and written to the device; chain: the input planes in and the
 * restored planes out). Any argument may be NULL. Enabling
 * allocates the event pool. */
vaco_av1pf_status vaco_av1pf_profile(vaco_av1pf_ctx *ctx, int enable);
vaco_av1pf_status vaco_av1pf_profile_read(vaco_av1pf_ctx *ctx,
                                          double ms[VACO_AV1PF_KERNEL_COUNT],
                                          uint64_t launches[VACO_AV1PF_KERNEL_COUNT],
                                          uint64_t bytes[VACO_AV1PF_KERNEL_COUNT]);

/* ------------------------------------------------ host-only helpers --- */

/* cdef::adjust_strength, cdef.rs:208-220. */
uint16_t vaco_av1pf_adjust_strength(uint16_t primary, uint32_t variance);
/* cdef::chroma_direction, cdef.rs:222-230. */
uint8_t vaco_av1pf_chroma_direction(uint8_t direction, uint8_t sub_x,
                                    uint8_t sub_y);
/* restoration::PlaneConfig::unit_counts, restoration.rs:111-120. */
void vaco_av1pf_lr_unit_counts(const vaco_av1pf_lr_plane_config *cfg,
                               uint32_t *cols, uint32_t *rows);
/* restoration::PlaneConfig::unit_index, restoration.rs:124-133. */
uint32_t vaco_av1pf_lr_unit_index(const vaco_av1pf_lr_plane_config *cfg,
                                  uint32_t x, uint32_t y);
/* decode.rs:892-896 cdef_block_skipped over a MiCell skip grid
 * (mi_rows x mi_cols, nonzero = skip) into the 8x8 map of the frame desc. */
void vaco_av1pf_skip8_from_mi(const uint8_t *mi_skip, uint32_t mi_rows,
                              uint32_t mi_cols, uint8_t *skip8);

#ifdef __cplusplus
}
#endif

#endif /* VACO_AV1PF_H */
