// av1pf_replay.cpp - geometry, parameter validation and the error-order
// replay of the AV1 post-reconstruction filter stage.
//
// The GPU kernels compute results and flag out-of-range samples; they do
// not decide which error the reference would report first. That decision
// is made here, on the host, only when something can fail: the checks of
// matteopolak/vaco's stage (crates/codec/vaco-codec-av1/src/) are walked
// in the reference's order without computing any filter output, except the
// luma direction search whose variance feeds a parameter check.
#include <algorithm>
#include <cstring>
#include <vector>

#include "av1pf_host.h"

namespace av1pf {

namespace {

const char *const kCdefDepth = "CDEF bit depth must be 8, 10, or 12";
const char *const kCdefSample = "CDEF sample exceeds bit depth";
const char *const kCdefBlock = "invalid CDEF block parameters";
const char *const kCdefIndex = "CDEF strength index exceeds table";
const char *const kSrDepth = "AV1 superres bit depth must be 8, 10, or 12";
const char *const kSrGeometry = "invalid AV1 superres plane geometry";
const char *const kSrSample = "AV1 superres sample exceeds bit depth";
const char *const kLrGeometry = "AV1 restoration plane geometry or bit depth";
const char *const kLrCount = "AV1 restoration unit count";
const char *const kLrCoeff = "AV1 restoration coefficient out of range";
const char *const kLrSample = "AV1 restoration sample exceeds bit depth";
const char *const kLrNoUnit = "AV1 active restoration has no unit size";
const char *const kFuel = "fuel";

bool depth_ok(unsigned bd) { return bd == 8 || bd == 10 || bd == 12; }

uint32_t div_ceil_shift(uint32_t v, unsigned s) { return (v + (1u << s) - 1) >> s; }

// restoration.rs:67-91
bool unit_ok(const vaco_av1pf_lr_unit &u)
{
    static const int lo[3] = {-5, -23, -17}, hi[3] = {10, 8, 46};
    switch (u.kind) {
    case VACO_AV1PF_UNIT_NONE:
        return true;
    case VACO_AV1PF_UNIT_WIENER:
        for (int k = 0; k < 3; k++)
            if (u.vertical[k] < lo[k] || u.vertical[k] > hi[k] || u.horizontal[k] < lo[k] ||
                u.horizontal[k] > hi[k])
                return false;
        return true;
    case VACO_AV1PF_UNIT_SELF_GUIDED:
        return u.set < 16 && u.xqd[0] >= -96 && u.xqd[0] <= 31 && u.xqd[1] >= -32 &&
               u.xqd[1] <= 95;
    default:
        return false;
    }
}

int ilog2(uint32_t v)
{
    int r = 0;
    while (v >>= 1)
        r++;
    return r;
}

// cdef.rs:79-124 without the checks (callers validated the samples).
void direction_of(const uint16_t *block64, int bd, uint8_t *dir, uint32_t *var)
{
    int32_t partial[8][15] = {};
    int32_t counts[8][15] = {};
    const int shift = bd - 8;
    for (int idx = 0; idx < 64; idx++) {
        const int i = idx / 8, j = idx % 8;
        const int bins[8] = {i + j, i + j / 2, i, 3 + i - j / 2, 7 + i - j, 3 + j - i / 2, j, i / 2 + j};
        const int32_t v = (int32_t)(block64[idx] >> shift) - 128;
        for (int d = 0; d < 8; d++) {
            partial[d][bins[d]] += v;
            counts[d][bins[d]]++;
        }
    }
    int64_t cost[8];
    for (int d = 0; d < 8; d++) {
        cost[d] = 0;
        for (int b = 0; b < 15; b++)
            if (counts[d][b])
                cost[d] += (int64_t)partial[d][b] * partial[d][b] * (840 / counts[d][b]);
    }
    int best = 0;
    for (int d = 1; d < 8; d++)
        if (cost[d] > cost[best])
            best = d;
    *dir = (uint8_t)best;
    const int64_t v = (cost[best] - cost[(best + 4) & 7]) >> 10;
    *var = (v < 0 || v > 0xffffffffLL) ? 0 : (uint32_t)v;
}

uint16_t get_clamped(const HostPlane &p, int64_t x, int64_t y)
{
    if (p.width == 0 || p.height == 0)
        return 0;
    x = std::min<int64_t>(std::max<int64_t>(x, 0), p.width - 1);
    y = std::min<int64_t>(std::max<int64_t>(y, 0), p.height - 1);
    return p.data[(size_t)y * p.stride + x];
}

}  // namespace

uint16_t host_adjust_strength(uint16_t primary, uint32_t variance)
{
    if (variance == 0)
        return 0;
    const uint32_t scaled = variance >> 6;
    uint32_t lg = scaled == 0 ? 0 : (uint32_t)ilog2(scaled);
    lg = std::min<uint32_t>(lg, 12);
    return (uint16_t)(((uint32_t)primary * (4 + lg) + 8) >> 4);
}

uint8_t host_chroma_direction(uint8_t direction, uint8_t sub_x, uint8_t sub_y)
{
    static const uint8_t xo[8] = {7, 0, 2, 4, 5, 6, 6, 6};
    static const uint8_t yo[8] = {1, 2, 2, 2, 3, 4, 6, 0};
    if (sub_x && !sub_y)
        return xo[direction & 7];
    if (!sub_x && sub_y)
        return yo[direction & 7];
    return direction & 7;
}

void lr_unit_counts(uint32_t width, uint32_t height, uint32_t unit, uint32_t *cols, uint32_t *rows)
{
    // restoration.rs:111-120
    const uint64_t c = unit ? ((uint64_t)width + unit / 2) / unit : 0;
    const uint64_t r = unit ? ((uint64_t)height + unit / 2) / unit : 0;
    *cols = (uint32_t)std::max<uint64_t>(1, c);
    *rows = (uint32_t)std::max<uint64_t>(1, r);
}

void superres_phase(uint32_t visible, uint32_t output, int32_t *step, int32_t *start)
{
    // superres.rs:145-152, 168
    if (output == 0) {
        *step = 0;
        *start = 0;
        return;
    }
    const int64_t vis = visible, out = output;
    const int64_t st = ((vis << 14) + out / 2) / out;
    const int64_t error = out * st - (vis << 14);
    const int64_t s0 = (-((out - vis) << 13) + out / 2) / out + (1 << 7) - error / 2;
    *step = (int32_t)st;
    *start = (int32_t)(s0 & 0x3fff);
}

void frame_geometry(const vaco_av1pf_frame_desc &d, FrameGeo *g)
{
    memset(g, 0, sizeof *g);
    g->nplanes = d.monochrome ? 1 : 3;
    g->mi_cols = 2 * ((d.coded_width + 7) >> 3);
    g->mi_rows = 2 * ((d.coded_height + 7) >> 3);
    g->idx_rows = (g->mi_rows + 15) / 16;
    const unsigned sx = d.subsampling_x ? 1 : 0, sy = d.subsampling_y ? 1 : 0;
    g->in_w[0] = g->mi_cols * 4;
    g->in_h[0] = g->mi_rows * 4;
    for (int p = 1; p < g->nplanes; p++) {
        g->in_w[p] = g->in_w[0] >> sx;
        g->in_h[p] = g->in_h[0] >> sy;
    }
    bool any = false;
    for (int i = 0; i < 8; i++)
        any |= d.cdef.y[i].primary || d.cdef.y[i].secondary || d.cdef.uv[i].primary ||
               d.cdef.uv[i].secondary;
    g->cdef = d.cdef.enabled && any;  // decode.rs:741-748
    g->superres = d.use_superres != 0;
    g->lr = d.lr_type[0] || d.lr_type[1] || d.lr_type[2];  // restoration.rs:41-43
    for (int p = 0; p < g->nplanes; p++) {
        const unsigned px = p ? sx : 0, py = p ? sy : 0;
        g->sr_vis[p] = div_ceil_shift(d.coded_width, px);
        g->sr_out_w[p] = div_ceil_shift(d.upscaled_width, px);
        g->sr_h[p] = div_ceil_shift(d.coded_height, py);
        superres_phase(g->sr_vis[p], g->sr_out_w[p], &g->sr_step[p], &g->sr_start[p]);
        g->lr_w[p] = d.upscaled_width >> px;  // decode.rs:604-605 (floor)
        g->lr_h[p] = d.coded_height >> py;
        if (g->lr) {
            g->out_w[p] = g->lr_w[p];
            g->out_h[p] = g->lr_h[p];
        } else if (g->superres) {
            g->out_w[p] = g->sr_out_w[p];
            g->out_h[p] = g->sr_h[p];
        } else {
            g->out_w[p] = g->in_w[p];
            g->out_h[p] = g->in_h[p];
        }
    }
    g->has_idx = d.cdef_idx != nullptr;
    g->has_skip = d.skip8 != nullptr;
}

bool params_clean(const vaco_av1pf_frame_desc &d, const FrameGeo &g)
{
    if (!depth_ok(d.bit_depth))
        return false;
    const int shift = d.bit_depth - 8;
    if (g.cdef) {
        if (d.cdef.damping + shift > 12 || (g.nplanes > 1 && d.cdef.damping + shift < 1))
            return false;
        for (int i = 0; i < 8; i++)
            if (d.cdef.y[i].primary > 15 || d.cdef.y[i].secondary > 4 ||
                d.cdef.uv[i].primary > 15 || d.cdef.uv[i].secondary > 4)
                return false;
        const size_t n = (size_t)g.idx_rows * d.cdef_stride;
        for (size_t i = 0; i < n; i++)
            if (d.cdef_idx[i] > 7)
                return false;
    }
    const uint32_t img_w[3] = {g.in_w[0], g.in_w[1], g.in_w[2]};
    if (g.superres) {
        for (int p = 0; p < g.nplanes; p++)
            if (g.sr_vis[p] == 0 || g.sr_out_w[p] <= g.sr_vis[p] || g.sr_h[p] == 0 ||
                g.sr_vis[p] > img_w[p] || g.sr_h[p] > g.in_h[p])
                return false;
    }
    if (g.lr) {
        for (int p = 0; p < g.nplanes; p++) {
            if (d.lr_unit_size[p] == 0)
                return false;
            if (d.lr_type[p] == VACO_AV1PF_RESTORE_NONE)
                continue;
            const uint32_t u = d.lr_unit_size[p];
            const uint32_t src_w = g.superres ? g.sr_out_w[p] : g.in_w[p];
            const uint32_t src_h = g.superres ? g.sr_h[p] : g.in_h[p];
            if (!(u == 32 || u == 64 || u == 128 || u == 256) || g.lr_w[p] == 0 || g.lr_h[p] == 0 ||
                g.lr_w[p] > src_w || g.lr_h[p] > src_h)
                return false;
            uint32_t cols, rows;
            lr_unit_counts(g.lr_w[p], g.lr_h[p], u, &cols, &rows);
            if ((uint64_t)cols * rows != d.lr_unit_count[p] || !d.lr_units[p])
                return false;
            for (uint32_t i = 0; i < d.lr_unit_count[p]; i++)
                if (!unit_ok(d.lr_units[p][i]))
                    return false;
        }
    }
    return true;
}

uint64_t required_fuel(const vaco_av1pf_frame_desc &d, const FrameGeo &g)
{
    uint64_t fuel = 0;
    if (g.cdef) {
        for (uint32_t r = 0; r < g.mi_rows; r += 2)
            for (uint32_t c = 0; c < g.mi_cols; c += 2) {
                const int idx = d.cdef_idx[(size_t)(r >> 4) * d.cdef_stride + (c >> 4)];
                if (idx >= 0 && !d.skip8[(size_t)(r / 2) * (g.mi_cols / 2) + c / 2])
                    fuel++;
            }
    }
    if (g.superres)
        for (int p = 0; p < g.nplanes; p++)
            fuel += (uint64_t)g.sr_out_w[p] * g.sr_h[p];
    return fuel;
}

bool check_find_direction(const uint16_t *input144, uint8_t bit_depth, const char **msg)
{
    if (!depth_ok(bit_depth)) {
        *msg = kCdefDepth;
        return false;
    }
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++)
            if (input144[(i + 2) * 12 + j + 2] >= (1u << bit_depth)) {
                *msg = kCdefSample;
                return false;
            }
    return true;
}

bool check_filter_block(const uint16_t *input144, const vaco_av1pf_filter_params &p,
                        const char **msg)
{
    if (!depth_ok(p.bit_depth)) {
        *msg = kCdefDepth;
        return false;
    }
    const int shift = p.bit_depth - 8;
    bool bad = !(p.width == 4 || p.width == 8) || !(p.height == 4 || p.height == 8) ||
               p.direction > 7 || p.edges > 15 || p.damping > 12 || p.primary > (15 << shift) ||
               p.secondary > (4 << shift);
    for (int i = 0; i < 144 && !bad; i++)
        bad = input144[i] >= (1u << p.bit_depth);
    if (bad) {
        *msg = kCdefBlock;
        return false;
    }
    return true;
}

bool check_superres_plane(const vaco_av1pf_sr_plane_config &c, uint32_t in_w, uint32_t in_h,
                          const char **msg)
{
    if (!depth_ok(c.bit_depth)) {
        *msg = kSrDepth;
        return false;
    }
    if (c.visible_width == 0 || c.output_width <= c.visible_width || c.height == 0 ||
        c.visible_width > in_w || c.height > in_h) {
        *msg = kSrGeometry;
        return false;
    }
    return true;
}

bool check_restore_plane(const vaco_av1pf_lr_plane_config &c, const HostPlane &before,
                         const HostPlane &after, const vaco_av1pf_lr_unit *units, uint32_t n,
                         const char **msg)
{
    if (!depth_ok(c.bit_depth) ||
        !(c.unit_size == 32 || c.unit_size == 64 || c.unit_size == 128 || c.unit_size == 256) ||
        c.width == 0 || c.height == 0 || c.width > 0x7fffffffu - 8 || c.height > 0x7fffffffu - 64 ||
        c.width > before.width || c.width > after.width || c.height > before.height ||
        c.height > after.height) {
        *msg = kLrGeometry;
        return false;
    }
    uint32_t cols, rows;
    lr_unit_counts(c.width, c.height, c.unit_size, &cols, &rows);
    if ((uint64_t)cols * rows != n) {
        *msg = kLrCount;
        return false;
    }
    for (uint32_t i = 0; i < n; i++)
        if (!unit_ok(units[i])) {
            *msg = kLrCoeff;
            return false;
        }
    const uint16_t mx = (uint16_t)((1u << c.bit_depth) - 1);
    for (const HostPlane *pl : {&before, &after})
        for (uint32_t y = 0; y < c.height; y++)
            for (uint32_t x = 0; x < c.width; x++)
                if (pl->data[(size_t)y * pl->stride + x] > mx) {
                    *msg = kLrSample;
                    return false;
                }
    return true;
}

vaco_av1pf_status replay_checks(const vaco_av1pf_frame_desc &d, const FrameGeo &g,
                                const HostPlane in[3], uint64_t *fuel, const char **msg)
{
    const int np = g.nplanes;
    const int bd = d.bit_depth;
    const uint32_t maxv = depth_ok(bd) ? (1u << bd) - 1 : 0xffffu;
    // Samples CDEF replaced (they are in range by construction).
    std::vector<uint8_t> filtered[3];
    auto fail = [&](vaco_av1pf_status st, const char *m) {
        *msg = m;
        return st;
    };

    // ---- apply_cdef, decode.rs:739-861
    if (g.cdef) {
        for (int p = 0; p < np; p++)
            filtered[p].assign((size_t)g.in_w[p] * g.in_h[p], 0);
        const int shift = bd - 8;
        for (uint32_t r = 0; r < g.mi_rows; r += 2)
            for (uint32_t c = 0; c < g.mi_cols; c += 2) {
                const int idx = d.cdef_idx[(size_t)(r >> 4) * d.cdef_stride + (c >> 4)];
                if (idx < 0 || d.skip8[(size_t)(r / 2) * (g.mi_cols / 2) + c / 2])
                    continue;
                uint16_t luma[64];
                for (int i = 0; i < 64; i++)
                    luma[i] = get_clamped(in[0], c * 4 + i % 8, r * 4 + (i >> 3));
                if (!depth_ok(bd))
                    return fail(VACO_AV1PF_INVALID_DATA, kCdefDepth);
                for (int i = 0; i < 64; i++)
                    if (luma[i] >= (1u << bd))
                        return fail(VACO_AV1PF_INVALID_DATA, kCdefSample);
                uint8_t dir;
                uint32_t var;
                direction_of(luma, bd, &dir, &var);
                if (fuel) {  // decode.rs:788
                    if (*fuel < 1) {
                        *fuel = 0;
                        return fail(VACO_AV1PF_LIMIT_EXCEEDED, kFuel);
                    }
                    *fuel -= 1;
                }
                for (int p = 0; p < np; p++) {
                    const bool chroma = p != 0;
                    const unsigned sx = chroma && d.subsampling_x, sy = chroma && d.subsampling_y;
                    const int64_t x = (int64_t)(c * 4) >> sx, y = (int64_t)(r * 4) >> sy;
                    const int w = 8 >> sx, h = 8 >> sy;
                    if (idx >= 8)
                        return fail(VACO_AV1PF_INVALID_DATA, kCdefIndex);
                    const vaco_av1pf_strength st = chroma ? d.cdef.uv[idx] : d.cdef.y[idx];
                    const uint16_t pri = (uint16_t)(st.primary << shift);
                    const uint16_t sec = (uint16_t)(st.secondary << shift);
                    vaco_av1pf_filter_params fp;
                    fp.width = (uint32_t)w;
                    fp.height = (uint32_t)h;
                    fp.bit_depth = (uint8_t)bd;
                    fp.direction = pri == 0 ? 0
                                   : chroma ? host_chroma_direction(dir, d.subsampling_x, d.subsampling_y)
                                            : dir;
                    fp.primary = chroma ? pri : host_adjust_strength(pri, var);
                    fp.secondary = sec;
                    fp.damping = (uint8_t)(d.cdef.damping + shift - (chroma ? 1 : 0));
                    fp.edges = (uint8_t)((x > 0) | ((x + w < (int64_t)in[p].width) << 1) |
                                         ((y > 0) << 2) | ((y + h < (int64_t)in[p].height) << 3));
                    uint16_t nb[144];
                    memset(nb, 0, sizeof nb);
                    for (int ny = 0; ny < h + 4; ny++)
                        for (int nx = 0; nx < w + 4; nx++)
                            nb[ny * 12 + nx] = get_clamped(in[p], x + nx - 2, y + ny - 2);
                    const char *m = nullptr;
                    if (!check_filter_block(nb, fp, &m))
                        return fail(VACO_AV1PF_INVALID_DATA, m);
                    for (int oy = 0; oy < h; oy++)
                        for (int ox = 0; ox < w; ox++)
                            filtered[p][(size_t)(y + oy) * g.in_w[p] + x + ox] = 1;
                }
            }
    }
    // Sample of the CDEF image (input where CDEF did not write).
    auto cdef_bad = [&](int p, uint32_t x, uint32_t y) {
        if (g.cdef && filtered[p][(size_t)y * g.in_w[p] + x])
            return false;
        return in[p].data[(size_t)y * in[p].stride + x] > maxv;
    };

    // ---- apply_superres, decode.rs:865-890 / superres.rs:163-216
    if (g.superres) {
        for (int p = 0; p < np; p++) {
            vaco_av1pf_sr_plane_config c;
            c.visible_width = g.sr_vis[p];
            c.output_width = g.sr_out_w[p];
            c.height = g.sr_h[p];
            c.bit_depth = (uint8_t)bd;
            const char *m = nullptr;
            if (!check_superres_plane(c, g.in_w[p], g.in_h[p], &m))
                return fail(VACO_AV1PF_INVALID_DATA, m);
            for (uint32_t y = 0; y < c.height; y++) {
                if (fuel) {
                    if (*fuel < c.output_width) {
                        *fuel = 0;
                        return fail(VACO_AV1PF_LIMIT_EXCEEDED, kFuel);
                    }
                    *fuel -= c.output_width;
                }
                for (uint32_t x = 0; x < c.output_width; x++) {
                    const int32_t src = -(1 << 14) + g.sr_start[p] + (int32_t)x * g.sr_step[p];
                    const int32_t px = src >> 14;
                    for (int k = 0; k < 8; k++) {
                        const int32_t sxx = std::min<int32_t>(std::max<int32_t>(px + k - 3, 0),
                                                              (int32_t)g.in_w[p] - 1);
                        if (cdef_bad(p, (uint32_t)sxx, y))
                            return fail(VACO_AV1PF_INVALID_DATA, kSrSample);
                    }
                }
            }
        }
    }

    // ---- apply_restoration, decode.rs:683-737 / restoration.rs:135-168
    if (g.lr) {
        for (int p = 0; p < np; p++) {
            if (d.lr_unit_size[p] == 0)
                return fail(VACO_AV1PF_INVALID_DATA, kLrNoUnit);
            if (d.lr_type[p] == VACO_AV1PF_RESTORE_NONE)
                continue;
            const uint32_t u = d.lr_unit_size[p];
            const uint32_t src_w = g.superres ? g.sr_out_w[p] : g.in_w[p];
            const uint32_t src_h = g.superres ? g.sr_h[p] : g.in_h[p];
            if (!depth_ok(bd) || !(u == 32 || u == 64 || u == 128 || u == 256) || g.lr_w[p] == 0 ||
                g.lr_h[p] == 0 || g.lr_w[p] > src_w || g.lr_h[p] > src_h)
                return fail(VACO_AV1PF_INVALID_DATA, kLrGeometry);
            uint32_t cols, rows;
            lr_unit_counts(g.lr_w[p], g.lr_h[p], u, &cols, &rows);
            if ((uint64_t)cols * rows != d.lr_unit_count[p] || !d.lr_units[p])
                return fail(VACO_AV1PF_INVALID_DATA, kLrCount);
            for (uint32_t i = 0; i < d.lr_unit_count[p]; i++)
                if (!unit_ok(d.lr_units[p][i]))
                    return fail(VACO_AV1PF_INVALID_DATA, kLrCoeff);
            if (!g.superres) {  // upscaled images are clipped to the depth
                for (uint32_t y = 0; y < g.lr_h[p]; y++)
                    for (uint32_t x = 0; x < g.lr_w[p]; x++)
                        if (in[p].data[(size_t)y * in[p].stride + x] > maxv || cdef_bad(p, x, y))
                            return fail(VACO_AV1PF_INVALID_DATA, kLrSample);
            }
        }
    }
    return VACO_AV1PF_OK;
}

}  // namespace av1pf
