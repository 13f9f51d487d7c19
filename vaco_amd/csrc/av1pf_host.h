// av1pf_host.h - host-side geometry, validation and error-order replay of
// the AV1 post-reconstruction filter stage (no GPU code).
#pragma once

#include <stdint.h>

#include "../../include/vaco_av1pf.h"

namespace av1pf {

// Frame geometry implied by a desc (decode.rs:518-530, superres.rs:226-252,
// decode.rs:587-610) and which stages run.
struct FrameGeo {
    int nplanes;
    uint32_t mi_cols, mi_rows, idx_rows;
    uint32_t in_w[3], in_h[3];          // Mi-padded input planes
    bool cdef, superres, lr;
    uint32_t sr_vis[3], sr_out_w[3], sr_h[3];
    int32_t sr_step[3], sr_start[3];
    uint32_t lr_w[3], lr_h[3];          // restoration visible geometry
    uint32_t out_w[3], out_h[3];        // stage output
    bool has_idx, has_skip;
};

struct HostPlane {
    const uint16_t *data;
    uint32_t width, height, stride;
};

void frame_geometry(const vaco_av1pf_frame_desc &d, FrameGeo *g);

// True when no parameter of `d` can make the reference fail; sample range
// and fuel are not considered here.
bool params_clean(const vaco_av1pf_frame_desc &d, const FrameGeo &g);

// Fuel the reference charges for this frame (decode.rs:788, superres.rs:174).
uint64_t required_fuel(const vaco_av1pf_frame_desc &d, const FrameGeo &g);

// Replays the reference's checks in order over the host input planes and
// returns the first failure (status + message), or OK. Charges fuel.
vaco_av1pf_status replay_checks(const vaco_av1pf_frame_desc &d, const FrameGeo &g,
                                const HostPlane in[3], uint64_t *fuel, const char **msg);

// Single-function checks (cdef.rs:68-87, 142-154; superres.rs:128-143;
// restoration.rs:135-168).
bool check_find_direction(const uint16_t *input144, uint8_t bit_depth, const char **msg);
bool check_filter_block(const uint16_t *input144, const vaco_av1pf_filter_params &p,
                        const char **msg);
bool check_superres_plane(const vaco_av1pf_sr_plane_config &c, uint32_t in_w, uint32_t in_h,
                          const char **msg);
bool check_restore_plane(const vaco_av1pf_lr_plane_config &c, const HostPlane &before,
                         const HostPlane &after, const vaco_av1pf_lr_unit *units, uint32_t n,
                         const char **msg);

void superres_phase(uint32_t visible, uint32_t output, int32_t *step, int32_t *start);
void lr_unit_counts(uint32_t width, uint32_t height, uint32_t unit, uint32_t *cols, uint32_t *rows);
uint16_t host_adjust_strength(uint16_t primary, uint32_t variance);
uint8_t host_chroma_direction(uint8_t direction, uint8_t sub_x, uint8_t sub_y);

}  // namespace av1pf
