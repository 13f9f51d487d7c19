/*
 * av1pf_oracle.h - CPU oracle for the AV1 post-reconstruction filter stage.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of
 * matteopolak/vaco's scalar Rust (crates/codec/vaco-codec-av1/src/cdef.rs,
 * superres.rs, restoration.rs, framebuf.rs and the frame glue in
 * decode.rs:506-896). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline; the product library never links or calls it.
 *
 * Every function returns 0 on success or a vaco_av1pf_status code, and on
 * failure stores the reference's error message in *err.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_find_direction, orc_filter_block   dav1d CDEF corpus, 7,680 cases
 *   orc_upscale_plane                      dav1d superres corpus, 18 cases
 *   orc_restore_plane                      dav1d LR corpus, 300 + 48 cases
 *   orc_adjust_strength                    closed form, exhaustive log2 buckets
 *   orc_chroma_direction                   spec table, exhaustive
 *   orc_lr_unit_counts / orc_lr_unit_index reference unit-geometry test vectors
 *   orc_pic_to_frame                       closed form of blit (crop, byte order)
 *                                          on hand-checked planes
 *   orc_frame                              invariants + brute force against the
 *                                          pinned block/plane functions;
 *                                          "parity unpinned" for the frame glue
 *                                          beyond those (no reference vector
 *                                          exercises apply_cdef/superres->LR).
 */
#ifndef AV1PF_ORACLE_H
#define AV1PF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/vaco_av1pf.h"

#ifdef __cplusplus
extern "C" {
#endif

int orc_find_direction(const uint16_t block[64], int bit_depth, uint8_t *dir,
                       uint32_t *var, const char **err);
int orc_filter_block(const uint16_t input[144],
                     const vaco_av1pf_filter_params *p, uint16_t out[64],
                     const char **err);
uint16_t orc_adjust_strength(uint16_t primary, uint32_t variance);
uint8_t orc_chroma_direction(uint8_t direction, int sub_x, int sub_y);

/* fuel: NULL = unlimited, else remaining fuel (decremented). */
int orc_upscale_plane(const vaco_av1pf_plane *in,
                      const vaco_av1pf_sr_plane_config *cfg,
                      vaco_av1pf_plane *out, uint64_t *fuel, const char **err);

void orc_lr_unit_counts(const vaco_av1pf_lr_plane_config *cfg, size_t *cols,
                        size_t *rows);
size_t orc_lr_unit_index(const vaco_av1pf_lr_plane_config *cfg, size_t x,
                         size_t y);
int orc_restore_plane(const vaco_av1pf_plane *before,
                      const vaco_av1pf_plane *after,
                      const vaco_av1pf_lr_plane_config *cfg,
                      const vaco_av1pf_lr_unit *units, size_t n_units,
                      vaco_av1pf_plane *out, const char **err);

/* Output geometry of orc_frame for desc. */
int orc_frame_output_geometry(const vaco_av1pf_frame_desc *d, uint32_t w[3],
                              uint32_t h[3]);
/* decode.rs:554-559 as a unit. `out` planes must have the geometry
 * orc_frame_output_geometry reports; they are written in full. */
int orc_frame(const vaco_av1pf_frame_desc *d, const vaco_av1pf_picture *in,
              vaco_av1pf_picture *out, uint64_t *fuel, const char **err);

/* pic_to_frame with blit, decode.rs:3795-3891: the final picture `pic`
 * (orc_frame's output) written into frame planes of upscaled_width x
 * coded_height (chroma rounded up), one byte per sample at 8-bit and two
 * little-endian bytes above, rows stride_bytes apart. Each plane writes its
 * min(frame, picture) region; other frame bytes are left as they are. */
void orc_pic_to_frame(const vaco_av1pf_frame_desc *d, const vaco_av1pf_picture *pic,
                      uint8_t *const out[3], const uint32_t stride_bytes[3]);

#ifdef __cplusplus
}
#endif

#endif
