/*
 * vaco_av1pf_gen.h - seeded synthetic workloads for the AV1 filter stage.
 *
 * No reference counterpart: the reference has no workload generator for
 * this stage. Frames and side information are a pure function of
 * (config, frame index), computed with integer arithmetic only
 * (splitmix64 counter hashing; Gaussian noise as a 12-term Irwin-Hall sum
 * in 16-bit fixed point), so every host regenerates them bit-identically.
 *
 * Content: luma = smooth 2-D gradient + N(0, sigma) noise (+ optional hard
 * half-plane edges at angles from a fixed integer table), clipped to the
 * bit depth; chroma built the same way at half resolution (4:2:0).
 * Side information: CDEF enabled, damping U[3,6], eight presets with
 * primary U[0,15] and secondary U{0,1,2,4} for luma and chroma, cdef_idx
 * U[0,7] per 64x64, each 8x8 skipped with probability skip_permille/1000;
 * loop restoration Switchable on all planes with every unit U{None, Wiener,
 * self-guided}, Wiener taps uniform in [-5,10], [-23,8], [-17,46] (chroma
 * outer tap 0), sgrproj set U[0,15], xqd0 U[-96,31], xqd1 U[-32,95].
 */
#ifndef VACO_AV1PF_GEN_H
#define VACO_AV1PF_GEN_H

#include "vaco_av1pf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vaco_av1pf_gen_config {
    uint32_t width;           /* output (upscaled) luma width */
    uint32_t height;          /* luma height */
    uint32_t superres_denom;  /* 8 = off, 9..16 = fixed, 0 = U[9,16] per frame */
    uint8_t bit_depth;        /* 8, 10 */
    uint8_t hard_edges;       /* random half-plane edges per frame */
    uint8_t monochrome;       /* 0 = 4:2:0 */
    uint8_t lr_enabled;       /* 0 = no loop restoration */
    uint16_t noise_sigma;     /* in sample units at bit_depth */
    uint16_t lr_unit_size;    /* luma and chroma restoration unit size */
    uint16_t skip_permille;   /* 8x8 skip probability in 1/1000 */
    uint16_t pad;
    uint64_t seed;
} vaco_av1pf_gen_config;

typedef struct vaco_av1pf_gen_layout {
    uint32_t coded_width, upscaled_width, coded_height;
    uint32_t superres_denom;          /* 8 when superres is off */
    uint32_t plane_width[3], plane_height[3];  /* Mi-padded input planes */
    uint32_t cdef_stride, cdef_idx_count, skip8_count;
    uint32_t lr_unit_count[3];
} vaco_av1pf_gen_layout;

/* Geometry and side-information sizes of frame `frame_index`. */
vaco_av1pf_status vaco_av1pf_gen_get_layout(const vaco_av1pf_gen_config *cfg,
                                        uint32_t frame_index,
                                        vaco_av1pf_gen_layout *out);

/* Fill frame `frame_index`: planes (caller buffers of the layout's plane
 * geometry), cdef_idx[cdef_idx_count], skip8[skip8_count],
 * units[p][lr_unit_count[p]], and `desc` (which points at those arrays). */
vaco_av1pf_status vaco_av1pf_gen_frame(const vaco_av1pf_gen_config *cfg,
                                       uint32_t frame_index,
                                       vaco_av1pf_frame_desc *desc,
                                       vaco_av1pf_picture *planes,
                                       int8_t *cdef_idx, uint8_t *skip8,
                                       vaco_av1pf_lr_unit *units0,
                                       vaco_av1pf_lr_unit *units1,
                                       vaco_av1pf_lr_unit *units2);

#ifdef __cplusplus
}
#endif

#endif /* VACO_AV1PF_GEN_H */
