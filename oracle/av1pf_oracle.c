/*
 * av1pf_oracle.c - CPU oracle (TEST INFRASTRUCTURE ONLY) for the AV1
 * post-reconstruction filter stage of matteopolak/vaco.
 *
 * A line-by-line restatement in C of the reference's scalar Rust. It keeps
 * the reference's structure on purpose (12x12 gathers with full validation
 * per CDEF block, per-pixel unit_index, per-pixel box-sum recomputation in
 * self-guided restoration, 64-bit intermediates, a whole-picture copy per
 * stage), because it doubles as the timed "reference CPU path". Citations
 * are crates/codec/vaco-codec-av1/src/<file>:<line> of the reference.
 *
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; the product library never links it.
 */
#include "av1pf_oracle.h"

#include <stdlib.h>
#include <string.h>

#define ERR(msg)              \
    do {                      \
        *err = (msg);         \
        return VACO_AV1PF_INVALID_DATA; \
    } while (0)

/* ------------------------------------------------------------------------
 * framebuf::Plane (framebuf.rs:14-68)
 * ---------------------------------------------------------------------- */
typedef struct {
    uint16_t *data;
    int64_t w, h, stride;
} oplane;

static oplane op_from(const vaco_av1pf_plane *p)
{
    oplane o;
    o.data = p->data;
    o.w = p->width;
    o.h = p->height;
    o.stride = p->stride;
    return o;
}

static int op_alloc(oplane *o, int64_t w, int64_t h)
{
    o->w = w;
    o->h = h;
    o->stride = w;
    o->data = (uint16_t *)calloc((size_t)(w * h > 0 ? w * h : 1), 2);
    return o->data != NULL;
}

/* Plane::get_clamped, framebuf.rs:49-59. */
static uint16_t op_get(const oplane *o, int64_t x, int64_t y)
{
    int64_t mx = o->w > 0 ? o->w - 1 : 0, my = o->h > 0 ? o->h - 1 : 0;
    if (o->w == 0 || o->h == 0)
        return 0;
    if (x < 0) x = 0;
    if (x > mx) x = mx;
    if (y < 0) y = 0;
    if (y > my) y = my;
    return o->data[y * o->stride + x];
}

/* Plane::set, framebuf.rs:61-68. */
static void op_set(oplane *o, int64_t x, int64_t y, uint16_t v)
{
    if (x >= 0 && y >= 0 && x < o->w && y < o->h)
        o->data[y * o->stride + x] = v;
}

static int fuel_take(uint64_t *fuel, uint64_t n, const char **err)
{
    /* Budget::consume_fuel, vaco-limits/src/budget.rs:229-237. */
    if (fuel == NULL)
        return 0;
    if (n > *fuel) {
        *fuel = 0;
        *err = "fuel"; /* LimitError::FuelExhausted -> LimitExceeded { limit: "fuel" }, vaco-limits/src/error.rs:88-92 */
        return VACO_AV1PF_LIMIT_EXCEEDED;
    }
    *fuel -= n;
    return 0;
}

/* ------------------------------------------------------------------------
 * CDEF (cdef.rs)
 * ---------------------------------------------------------------------- */

/* DIRECTIONS, cdef.rs:56-66, stored as (row, column). */
static const int DIRS[8][2][2] = {
    {{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}},
    {{0, 1}, {1, 2}},   {{1, 1}, {2, 2}},  {{1, 0}, {2, 1}},
    {{1, 0}, {2, 0}},   {{1, 0}, {2, -1}},
};

static int ilog2_u32(uint32_t v) /* u32::ilog2 for v > 0 */
{
    int r = 0;
    while (v >>= 1)
        r++;
    return r;
}

/* find_direction, cdef.rs:79-124. */
int orc_find_direction(const uint16_t block[64], int bit_depth, uint8_t *dir,
                       uint32_t *var, const char **err)
{
    int32_t partial[8][15];
    uint32_t counts[8][15];
    int64_t costs[8];
    int shift, best, d, b, index;
    int64_t v;

    if (bit_depth != 8 && bit_depth != 10 && bit_depth != 12)
        ERR("CDEF bit depth must be 8, 10, or 12"); /* cdef.rs:68-73 */
    shift = bit_depth - 8;
    memset(partial, 0, sizeof partial);
    memset(counts, 0, sizeof counts);
    for (index = 0; index < 64; index++) {
        uint16_t sample = block[index];
        int i = index / 8, j = index % 8;
        int bins[8];
        int32_t value;
        if (sample >= (1u << bit_depth))
            ERR("CDEF sample exceeds bit depth");
        bins[0] = i + j;
        bins[1] = i + j / 2;
        bins[2] = i;
        bins[3] = 3 + i - j / 2;
        bins[4] = 7 + i - j;
        bins[5] = 3 + j - i / 2;
        bins[6] = j;
        bins[7] = i / 2 + j;
        value = (int32_t)(sample >> shift) - 128;
        for (d = 0; d < 8; d++) {
            partial[d][bins[d]] += value;
            counts[d][bins[d]] += 1;
        }
    }
    for (d = 0; d < 8; d++) {
        costs[d] = 0;
        for (b = 0; b < 15; b++) {
            uint32_t count = counts[d][b];
            if (count != 0) { /* 840u32.checked_div(count) */
                int64_t weight = 840 / count;
                int64_t sum = partial[d][b];
                costs[d] += sum * sum * weight;
            }
        }
    }
    best = 0;
    for (d = 1; d < 8; d++)
        if (costs[d] > costs[best]) /* first maximum wins, cdef.rs:116 */
            best = d;
    *dir = (uint8_t)best;
    v = (costs[best] - costs[(best + 4) & 7]) >> 10;
    *var = (v < 0 || v > 0xffffffffLL) ? 0 : (uint32_t)v; /* try_from.unwrap_or(0) */
    return 0;
}

/* constrain, cdef.rs:126-133. */
static int32_t constrain(int32_t difference, uint16_t threshold, uint8_t damping)
{
    uint32_t lg, shift;
    int32_t magnitude, lim, m;
    if (threshold == 0)
        return 0;
    lg = (uint32_t)ilog2_u32(threshold);
    shift = damping > lg ? damping - lg : 0; /* saturating_sub */
    magnitude = difference < 0 ? -difference : difference;
    lim = (int32_t)threshold - (magnitude >> shift);
    if (lim < 0)
        lim = 0;
    m = magnitude < lim ? magnitude : lim;
    return difference < 0 ? -m : (difference > 0 ? m : 0); /* signum * m */
}

/* filter_block, cdef.rs:142-205. */
int orc_filter_block(const uint16_t input[144], const vaco_av1pf_filter_params *p,
                     uint16_t out[64], const char **err)
{
    int shift, i, x, y, distance, s, t;
    int primary_taps[2];
    int x_min, x_max, y_min, y_max;
    if (p->bit_depth != 8 && p->bit_depth != 10 && p->bit_depth != 12)
        ERR("CDEF bit depth must be 8, 10, or 12");
    shift = p->bit_depth - 8;
    {
        int bad = !(p->width == 4 || p->width == 8) ||
                  !(p->height == 4 || p->height == 8) || p->direction > 7 ||
                  p->edges > 15 || p->damping > 12 ||
                  p->primary > (15 << shift) || p->secondary > (4 << shift);
        for (i = 0; i < 144 && !bad; i++)
            if (input[i] >= (1u << p->bit_depth))
                bad = 1;
        if (bad)
            ERR("invalid CDEF block parameters");
    }
    memset(out, 0, 64 * sizeof(uint16_t));
    if (((p->primary >> shift) & 1) == 0) {
        primary_taps[0] = 4;
        primary_taps[1] = 2;
    } else {
        primary_taps[0] = 3;
        primary_taps[1] = 3;
    }
    x_min = (p->edges & 1) ? -2 : 0;
    x_max = (int)p->width + ((p->edges & 2) ? 2 : 0);
    y_min = (p->edges & 4) ? -2 : 0;
    y_max = (int)p->height + ((p->edges & 8) ? 2 : 0);
    for (y = 0; y < (int)p->height; y++) {
        for (x = 0; x < (int)p->width; x++) {
            int32_t current = input[(y + 2) * 12 + x + 2];
            int32_t minimum = current, maximum = current, sum = 0, filtered;
            for (distance = 0; distance < 2; distance++) {
                for (s = 0; s < 2; s++) {
                    int sign = s == 0 ? -1 : 1;
                    int dirs3[3], taps3[3];
                    uint16_t str3[3];
                    dirs3[0] = p->direction;
                    str3[0] = p->primary;
                    taps3[0] = primary_taps[distance];
                    dirs3[1] = (p->direction + 6) & 7;
                    str3[1] = p->secondary;
                    taps3[1] = 2 - distance;
                    dirs3[2] = (p->direction + 2) & 7;
                    str3[2] = p->secondary;
                    taps3[2] = 2 - distance;
                    for (t = 0; t < 3; t++) {
                        int dy = DIRS[dirs3[t]][distance][0];
                        int dx = DIRS[dirs3[t]][distance][1];
                        int nx = x + sign * dx, ny = y + sign * dy;
                        int32_t neighbor;
                        if (nx < x_min || nx >= x_max || ny < y_min || ny >= y_max)
                            continue;
                        neighbor = input[(ny + 2) * 12 + nx + 2];
                        if (neighbor < minimum) minimum = neighbor;
                        if (neighbor > maximum) maximum = neighbor;
                        sum += taps3[t] * constrain(neighbor - current, str3[t], p->damping);
                    }
                }
            }
            filtered = current + ((8 + sum - (sum < 0)) >> 4);
            if (filtered < minimum) filtered = minimum;
            if (filtered > maximum) filtered = maximum;
            out[y * 8 + x] = (uint16_t)filtered;
        }
    }
    return 0;
}

/* adjust_strength, cdef.rs:208-220. */
uint16_t orc_adjust_strength(uint16_t primary, uint32_t variance)
{
    uint32_t scaled, lg;
    if (variance == 0)
        return 0;
    scaled = variance >> 6;
    lg = scaled == 0 ? 0 : (uint32_t)ilog2_u32(scaled);
    if (lg > 12)
        lg = 12;
    return (uint16_t)((((uint32_t)primary) * (4 + lg) + 8) >> 4);
}

/* chroma_direction, cdef.rs:222-230. */
uint8_t orc_chroma_direction(uint8_t direction, int sub_x, int sub_y)
{
    static const uint8_t x_only[8] = {7, 0, 2, 4, 5, 6, 6, 6};
    static const uint8_t y_only[8] = {1, 2, 2, 2, 3, 4, 6, 0};
    if (sub_x && !sub_y)
        return x_only[direction & 7];
    if (!sub_x && sub_y)
        return y_only[direction & 7];
    return direction & 7;
}

/* ------------------------------------------------------------------------
 * Superres (superres.rs)
 * ---------------------------------------------------------------------- */

/* UPSCALE_FILTER, superres.rs:25-90 (AV1 spec Upscale_Filter[64][8]). */
static const int16_t UPSCALE[64][8] = {
    {0, 0, 0, 128, 0, 0, 0, 0},       {0, 0, -1, 128, 2, -1, 0, 0},
    {0, 1, -3, 127, 4, -2, 1, 0},     {0, 1, -4, 127, 6, -3, 1, 0},
    {0, 2, -6, 126, 8, -3, 1, 0},     {0, 2, -7, 125, 11, -4, 1, 0},
    {-1, 2, -8, 125, 13, -5, 2, 0},   {-1, 3, -9, 124, 15, -6, 2, 0},
    {-1, 3, -10, 123, 18, -6, 2, -1}, {-1, 3, -11, 122, 20, -7, 3, -1},
    {-1, 4, -12, 121, 22, -8, 3, -1}, {-1, 4, -13, 120, 25, -9, 3, -1},
    {-1, 4, -14, 118, 28, -9, 3, -1}, {-1, 4, -15, 117, 30, -10, 4, -1},
    {-1, 5, -16, 116, 32, -11, 4, -1}, {-1, 5, -16, 114, 35, -12, 4, -1},
    {-1, 5, -17, 112, 38, -12, 4, -1}, {-1, 5, -18, 111, 40, -13, 5, -1},
    {-1, 5, -18, 109, 43, -14, 5, -1}, {-1, 6, -19, 107, 45, -14, 5, -1},
    {-1, 6, -19, 105, 48, -15, 5, -1}, {-1, 6, -19, 103, 51, -16, 5, -1},
    {-1, 6, -20, 101, 53, -16, 6, -1}, {-1, 6, -20, 99, 56, -17, 6, -1},
    {-1, 6, -20, 97, 58, -17, 6, -1}, {-1, 6, -20, 95, 61, -18, 6, -1},
    {-2, 7, -20, 93, 64, -18, 6, -2}, {-2, 7, -20, 91, 66, -19, 6, -1},
    {-2, 7, -20, 88, 69, -19, 6, -1}, {-2, 7, -20, 86, 71, -19, 6, -1},
    {-2, 7, -20, 84, 74, -20, 7, -2}, {-2, 7, -20, 81, 76, -20, 7, -1},
    {-2, 7, -20, 79, 79, -20, 7, -2}, {-1, 7, -20, 76, 81, -20, 7, -2},
    {-2, 7, -20, 74, 84, -20, 7, -2}, {-1, 6, -19, 71, 86, -20, 7, -2},
    {-1, 6, -19, 69, 88, -20, 7, -2}, {-1, 6, -19, 66, 91, -20, 7, -2},
    {-2, 6, -18, 64, 93, -20, 7, -2}, {-1, 6, -18, 61, 95, -20, 6, -1},
    {-1, 6, -17, 58, 97, -20, 6, -1}, {-1, 6, -17, 56, 99, -20, 6, -1},
    {-1, 6, -16, 53, 101, -20, 6, -1}, {-1, 5, -16, 51, 103, -19, 6, -1},
    {-1, 5, -15, 48, 105, -19, 6, -1}, {-1, 5, -14, 45, 107, -19, 6, -1},
    {-1, 5, -14, 43, 109, -18, 5, -1}, {-1, 5, -13, 40, 111, -18, 5, -1},
    {-1, 4, -12, 38, 112, -17, 5, -1}, {-1, 4, -12, 35, 114, -16, 5, -1},
    {-1, 4, -11, 32, 116, -16, 5, -1}, {-1, 4, -10, 30, 117, -15, 4, -1},
    {-1, 3, -9, 28, 118, -14, 4, -1}, {-1, 3, -9, 25, 120, -13, 4, -1},
    {-1, 3, -8, 22, 121, -12, 4, -1}, {-1, 3, -7, 20, 122, -11, 3, -1},
    {-1, 2, -6, 18, 123, -10, 3, -1}, {0, 2, -6, 15, 124, -9, 3, -1},
    {0, 2, -5, 13, 125, -8, 2, -1},   {0, 1, -4, 11, 125, -7, 2, 0},
    {0, 1, -3, 8, 126, -6, 2, 0},     {0, 1, -3, 6, 127, -4, 1, 0},
    {0, 1, -2, 4, 127, -3, 1, 0},     {0, 0, -1, 2, 128, -1, 0, 0},
};

/* subpel_start, superres.rs:145-152 (C `/` truncates like Rust's). */
static int64_t subpel_start(int64_t vis, int64_t outw, int64_t step)
{
    int64_t error = outw * step - (vis << 14);
    int64_t start = (-((outw - vis) << 13) + outw / 2) / outw + (1 << 7) - error / 2;
    return start & 0x3fff;
}

static int upscale_op(const oplane *input, const vaco_av1pf_sr_plane_config *c,
                      oplane *output, uint64_t *fuel, const char **err)
{
    int64_t vis, outw, padded, step, start, max_sample, x, y;
    int k, rc;
    /* validate, superres.rs:128-143 */
    if (c->bit_depth != 8 && c->bit_depth != 10 && c->bit_depth != 12)
        ERR("AV1 superres bit depth must be 8, 10, or 12");
    if (c->visible_width == 0 || c->output_width <= c->visible_width ||
        c->height == 0 || c->visible_width > input->w || c->height > input->h)
        ERR("invalid AV1 superres plane geometry");
    vis = c->visible_width;
    outw = c->output_width;
    padded = input->w;
    step = ((vis << 14) + outw / 2) / outw; /* superres.rs:168 */
    start = subpel_start(vis, outw, step);
    max_sample = (1LL << c->bit_depth) - 1;
    if (!op_alloc(output, outw, c->height))
        ERR("allocation failed");
    for (y = 0; y < (int64_t)c->height; y++) {
        if ((rc = fuel_take(fuel, (uint64_t)outw, err)) != 0) /* superres.rs:174 */
            return rc;
        for (x = 0; x < outw; x++) {
            int64_t src_x = -(1LL << 14) + start + x * step;
            int64_t src_x_px = src_x >> 14;
            int64_t phase = (src_x & 0x3fff) >> 8;
            int64_t sum = 0, rounded;
            for (k = 0; k < 8; k++) {
                int64_t sx = src_x_px + k - 3;
                uint16_t sample;
                if (sx < 0) sx = 0;
                if (sx > padded - 1) sx = padded - 1;
                sample = input->data[y * input->stride + sx];
                if ((int64_t)sample > max_sample)
                    ERR("AV1 superres sample exceeds bit depth");
                sum += (int64_t)sample * UPSCALE[phase][k];
            }
            rounded = (sum + 64) >> 7;
            if (rounded < 0) rounded = 0;
            if (rounded > max_sample) rounded = max_sample;
            op_set(output, x, y, (uint16_t)rounded);
        }
    }
    return 0;
}

/* upscale_plane, superres.rs:163-216. */
int orc_upscale_plane(const vaco_av1pf_plane *in, const vaco_av1pf_sr_plane_config *cfg,
                      vaco_av1pf_plane *out, uint64_t *fuel, const char **err)
{
    oplane i = op_from(in), o;
    int64_t y;
    int rc;
    o.data = NULL;
    rc = upscale_op(&i, cfg, &o, fuel, err);
    if (rc != 0) {
        free(o.data);
        return rc;
    }
    for (y = 0; y < o.h; y++)
        memcpy(out->data + y * out->stride, o.data + y * o.stride, (size_t)o.w * 2);
    free(o.data);
    return 0;
}

/* ------------------------------------------------------------------------
 * Loop restoration (restoration.rs)
 * ---------------------------------------------------------------------- */

/* PlaneConfig::unit_counts, restoration.rs:111-120. */
void orc_lr_unit_counts(const vaco_av1pf_lr_plane_config *c, size_t *cols, size_t *rows)
{
    size_t u = c->unit_size;
    size_t cc = u ? (c->width + u / 2) / u : 0;
    size_t rr = u ? (c->height + u / 2) / u : 0;
    *cols = cc < 1 ? 1 : cc;
    *rows = rr < 1 ? 1 : rr;
}

/* PlaneConfig::unit_index, restoration.rs:124-133. */
size_t orc_lr_unit_index(const vaco_av1pf_lr_plane_config *c, size_t x, size_t y)
{
    size_t cols, rows, row, col, u = c->unit_size;
    orc_lr_unit_counts(c, &cols, &rows);
    row = u ? (y + (8u >> (c->subsampling_y ? 1 : 0))) / u : 0;
    if (row > rows - 1) row = rows - 1;
    col = u ? x / u : 0;
    if (col > cols - 1) col = cols - 1;
    return row * cols + col;
}

/* RestorationUnit::validate, restoration.rs:67-91. */
static int unit_valid(const vaco_av1pf_lr_unit *u)
{
    static const int lo[3] = {-5, -23, -17}, hi[3] = {10, 8, 46};
    int k;
    switch (u->kind) {
    case VACO_AV1PF_UNIT_NONE:
        return 1;
    case VACO_AV1PF_UNIT_WIENER:
        for (k = 0; k < 3; k++) {
            if (u->vertical[k] < lo[k] || u->vertical[k] > hi[k]) return 0;
            if (u->horizontal[k] < lo[k] || u->horizontal[k] > hi[k]) return 0;
        }
        return 1;
    case VACO_AV1PF_UNIT_SELF_GUIDED:
        return u->set < 16 && u->xqd[0] >= -96 && u->xqd[0] <= 31 &&
               u->xqd[1] >= -32 && u->xqd[1] <= 95;
    default:
        return 0; /* not a RestorationUnit variant */
    }
}

/* Source, restoration.rs:209-245. */
typedef struct {
    const oplane *before, *after;
    const vaco_av1pf_lr_plane_config *c;
    int32_t stripe_start, stripe_end;
} osource;

static osource source_new(const oplane *before, const oplane *after,
                          const vaco_av1pf_lr_plane_config *c, int64_t y)
{
    osource s;
    int sub_y = c->subsampling_y ? 1 : 0;
    int64_t stripe_height = 64 >> sub_y, offset = 8 >> sub_y;
    int64_t stripe = (y + offset) / stripe_height;
    s.before = before;
    s.after = after;
    s.c = c;
    s.stripe_start = (int32_t)(stripe * stripe_height - offset);
    s.stripe_end = s.stripe_start + (int32_t)stripe_height - 1;
    return s;
}

static int64_t source_get(const osource *s, int32_t x, int32_t y)
{
    int32_t w = (int32_t)s->c->width, h = (int32_t)s->c->height;
    uint16_t sample;
    if (x < 0) x = 0;
    if (x > w - 1) x = w - 1;
    if (y < 0) y = 0;
    if (y > h - 1) y = h - 1;
    if (y < s->stripe_start)
        sample = op_get(s->before, x, y > s->stripe_start - 2 ? y : s->stripe_start - 2);
    else if (y > s->stripe_end)
        sample = op_get(s->before, x, y < s->stripe_end + 2 ? y : s->stripe_end + 2);
    else
        sample = op_get(s->after, x, y);
    return sample;
}

/* round2, restoration.rs:251-257. */
static int64_t round2(int64_t v, int shift)
{
    return shift == 0 ? v : (v + (1LL << (shift - 1))) >> shift;
}

/* wiener + taps, restoration.rs:259-283. */
static int64_t wiener(const osource *s, int32_t x, int32_t y, const int16_t vertical[3],
                      const int16_t horizontal[3])
{
    int depth = s->c->bit_depth;
    int round_h = depth == 12 ? 5 : 3;
    int64_t offset = 1LL << (depth + 7 - round_h - 1);
    int64_t limit = (1LL << (depth + 1 + 7 - round_h)) - 1;
    int64_t ht[7], vt[7], sum = 0;
    int row, col;
    ht[0] = ht[6] = horizontal[0];
    ht[1] = ht[5] = horizontal[1];
    ht[2] = ht[4] = horizontal[2];
    ht[3] = 128 - 2 * ((int64_t)horizontal[0] + horizontal[1] + horizontal[2]);
    vt[0] = vt[6] = vertical[0];
    vt[1] = vt[5] = vertical[1];
    vt[2] = vt[4] = vertical[2];
    vt[3] = 128 - 2 * ((int64_t)vertical[0] + vertical[1] + vertical[2]);
    for (row = 0; row < 7; row++) {
        int64_t h = 0, r;
        for (col = 0; col < 7; col++)
            h += ht[col] * source_get(s, x + col - 3, y + row - 3);
        r = round2(h, round_h);
        if (r < -offset) r = -offset;
        if (r > limit - offset) r = limit - offset;
        sum += vt[row] * r;
    }
    return round2(sum, 14 - round_h);
}

/* SGR_PARAMS, restoration.rs:286-303: {r0, e0, r1, e1}. */
static const int SGR[16][4] = {
    {2, 12, 1, 4},  {2, 15, 1, 6},  {2, 18, 1, 8},  {2, 21, 1, 9},
    {2, 24, 1, 10}, {2, 29, 1, 11}, {2, 36, 1, 12}, {2, 45, 1, 13},
    {2, 56, 1, 14}, {2, 68, 1, 15}, {0, 0, 1, 5},   {0, 0, 1, 8},
    {0, 0, 1, 11},  {0, 0, 1, 14},  {2, 30, 0, 0},  {2, 75, 0, 0},
};

/* box_ab, restoration.rs:314-339. */
static void box_ab(const osource *s, int32_t x, int32_t y, int radius, int64_t epsilon,
                   int64_t *a_out, int64_t *b_out)
{
    int64_t n = (int64_t)(2 * radius + 1) * (2 * radius + 1);
    int64_t n2e = n * n * epsilon;
    int64_t scale = ((1LL << 20) + n2e / 2) / n2e;
    int64_t sum_sq = 0, sum = 0, a, d, variance, z, a2, reciprocal;
    int dx, dy, bd = s->c->bit_depth;
    for (dy = -radius; dy <= radius; dy++)
        for (dx = -radius; dx <= radius; dx++) {
            int64_t sample = source_get(s, x + dx, y + dy);
            sum_sq += sample * sample;
            sum += sample;
        }
    a = round2(sum_sq, 2 * (bd - 8));
    d = round2(sum, bd - 8);
    variance = a * n - d * d;
    if (variance < 0)
        variance = 0;
    z = round2(variance * scale, 20);
    if (z >= 255)
        a2 = 256;
    else if (z == 0)
        a2 = 1;
    else
        a2 = ((z << 8) + z / 2) / (z + 1);
    reciprocal = ((1LL << 12) + n / 2) / n;
    *a_out = a2;
    *b_out = round2((256 - a2) * sum * reciprocal, 12);
}

/* guided_pass, restoration.rs:341-371. */
static int64_t guided_pass(const osource *s, int32_t x, int32_t y, int pass, int radius,
                           int64_t epsilon)
{
    int64_t original = op_get(s->after, x, y);
    int64_t a = 0, b = 0;
    int dx, dy, shift;
    if (radius == 0)
        return original << 4;
    for (dy = -1; dy <= 1; dy++)
        for (dx = -1; dx <= 1; dx++) {
            int64_t weight;
            if (pass == 0)
                weight = ((y + dy) & 1) == 0 ? 0 : (dx == 0 ? 6 : 5);
            else
                weight = (dx == 0 || dy == 0) ? 4 : 3;
            if (weight != 0) {
                int64_t aa, bb;
                box_ab(s, x + dx, y + dy, radius, epsilon, &aa, &bb);
                a += weight * aa;
                b += weight * bb;
            }
        }
    shift = (pass == 0 && (y & 1) != 0) ? 4 : 5;
    return round2(a * original + b, 8 + shift - 4);
}

/* self_guided, restoration.rs:373-384. */
static int64_t self_guided(const osource *s, int32_t x, int32_t y, int set, const int16_t xqd[2])
{
    const int *p = SGR[set < 16 ? set : 0];
    int64_t w0 = xqd[0], w1 = xqd[1];
    int64_t original = (int64_t)op_get(s->after, x, y) << 4;
    int64_t first = guided_pass(s, x, y, 0, p[0], p[1]);
    int64_t second = guided_pass(s, x, y, 1, p[2], p[3]);
    return round2(w0 * first + w1 * original + (128 - w0 - w1) * second, 11);
}

/* PlaneConfig::validate, restoration.rs:135-168. */
static int lr_validate(const vaco_av1pf_lr_plane_config *c, const oplane *before,
                       const oplane *after, const vaco_av1pf_lr_unit *units, size_t n,
                       const char **err)
{
    size_t cols, rows, i;
    int64_t x, y;
    uint16_t max;
    if (!(c->bit_depth == 8 || c->bit_depth == 10 || c->bit_depth == 12) ||
        !(c->unit_size == 32 || c->unit_size == 64 || c->unit_size == 128 ||
          c->unit_size == 256) ||
        c->width == 0 || c->height == 0 || c->width > 0x7fffffffu - 8 ||
        c->height > 0x7fffffffu - 64 || c->width > before->w || c->width > after->w ||
        c->height > before->h || c->height > after->h)
        ERR("AV1 restoration plane geometry or bit depth");
    orc_lr_unit_counts(c, &cols, &rows);
    if (cols * rows != n)
        ERR("AV1 restoration unit count");
    for (i = 0; i < n; i++)
        if (!unit_valid(&units[i]))
            ERR("AV1 restoration coefficient out of range");
    max = (uint16_t)((1u << c->bit_depth) - 1);
    for (y = 0; y < (int64_t)c->height; y++)
        for (x = 0; x < (int64_t)c->width; x++)
            if (op_get(before, x, y) > max || op_get(after, x, y) > max)
                ERR("AV1 restoration sample exceeds bit depth");
    return 0;
}

static int restore_op(const oplane *before, const oplane *after,
                      const vaco_av1pf_lr_plane_config *c, const vaco_av1pf_lr_unit *units,
                      size_t n, oplane *output, const char **err)
{
    int64_t x, y;
    int rc = lr_validate(c, before, after, units, n, err);
    if (rc != 0)
        return rc;
    if (!op_alloc(output, c->width, c->height))
        ERR("allocation failed");
    for (y = 0; y < (int64_t)c->height; y++) {
        osource s = source_new(before, after, c, y);
        for (x = 0; x < (int64_t)c->width; x++) {
            size_t ui = orc_lr_unit_index(c, (size_t)x, (size_t)y);
            const vaco_av1pf_lr_unit *u;
            int64_t value;
            if (ui >= n) {
                free(output->data);
                output->data = NULL;
                ERR("AV1 restoration unit index");
            }
            u = &units[ui];
            if (u->kind == VACO_AV1PF_UNIT_WIENER)
                value = wiener(&s, (int32_t)x, (int32_t)y, u->vertical, u->horizontal);
            else if (u->kind == VACO_AV1PF_UNIT_SELF_GUIDED)
                value = self_guided(&s, (int32_t)x, (int32_t)y, u->set, u->xqd);
            else
                value = op_get(after, x, y);
            if (value < 0) value = 0;
            if (value > (1LL << c->bit_depth) - 1) value = (1LL << c->bit_depth) - 1;
            op_set(output, x, y, (uint16_t)value);
        }
    }
    return 0;
}

/* restore_plane, restoration.rs:178-207. */
int orc_restore_plane(const vaco_av1pf_plane *before, const vaco_av1pf_plane *after,
                      const vaco_av1pf_lr_plane_config *cfg,
                      const vaco_av1pf_lr_unit *units, size_t n_units,
                      vaco_av1pf_plane *out, const char **err)
{
    oplane b = op_from(before), a = op_from(after), o;
    int64_t y;
    int rc;
    o.data = NULL;
    rc = restore_op(&b, &a, cfg, units, n_units, &o, err);
    if (rc != 0)
        return rc;
    for (y = 0; y < o.h; y++)
        memcpy(out->data + y * out->stride, o.data + y * o.stride, (size_t)o.w * 2);
    free(o.data);
    return 0;
}

/* ------------------------------------------------------------------------
 * Frame glue (decode.rs:506-896)
 * ---------------------------------------------------------------------- */

typedef struct {
    oplane p[3];
    int nplanes;
} opic;

static void opic_free(opic *p)
{
    int i;
    for (i = 0; i < p->nplanes; i++) {
        free(p->p[i].data);
        p->p[i].data = NULL;
    }
}

/* copy_picture, decode.rs:652-671 (also the CDEF output seed, 750-770). */
static int opic_copy(const opic *src, opic *dst)
{
    int i;
    int64_t y;
    dst->nplanes = src->nplanes;
    for (i = 0; i < src->nplanes; i++) {
        if (!op_alloc(&dst->p[i], src->p[i].w, src->p[i].h))
            return 0;
        for (y = 0; y < src->p[i].h; y++)
            memcpy(dst->p[i].data + y * dst->p[i].stride,
                   src->p[i].data + y * src->p[i].stride, (size_t)src->p[i].w * 2);
    }
    return 1;
}

static int lr_active(const vaco_av1pf_frame_desc *d)
{
    /* FrameRestoration::is_active, restoration.rs:41-43 */
    return d->lr_type[0] != 0 || d->lr_type[1] != 0 || d->lr_type[2] != 0;
}

/* apply_cdef, decode.rs:739-861 (with cdef_block_skipped 892-896). */
static int apply_cdef(const vaco_av1pf_frame_desc *d, opic *pic, uint64_t *fuel,
                      const char **err)
{
    const vaco_av1pf_cdef_params *cp = &d->cdef;
    int64_t mi_cols = 2 * (((int64_t)d->coded_width + 7) >> 3);
    int64_t mi_rows = 2 * (((int64_t)d->coded_height + 7) >> 3);
    int all_zero = 1, i, rc, plane;
    int shift = d->bit_depth - 8;
    int64_t r, c;
    opic filtered;
    for (i = 0; i < 8; i++)
        if (cp->y[i].primary || cp->y[i].secondary || cp->uv[i].primary ||
            cp->uv[i].secondary)
            all_zero = 0;
    if (!cp->enabled || all_zero)
        return 0;
    if (!opic_copy(pic, &filtered))
        ERR("allocation failed");
    for (r = 0; r < mi_rows; r += 2) {
        for (c = 0; c < mi_cols; c += 2) {
            int64_t slot = (r >> 4) * d->cdef_stride + (c >> 4);
            int index = d->cdef_idx ? d->cdef_idx[slot] : -1;
            uint16_t luma[64];
            uint8_t direction;
            uint32_t variance;
            if (index < 0 || d->skip8[(r / 2) * (mi_cols / 2) + c / 2])
                continue;
            for (i = 0; i < 64; i++)
                luma[i] = op_get(&pic->p[0], c * 4 + i % 8, r * 4 + (i >> 3));
            rc = orc_find_direction(luma, d->bit_depth, &direction, &variance, err);
            if (rc == 0)
                rc = fuel_take(fuel, 1, err); /* decode.rs:788 */
            if (rc != 0) {
                opic_free(&filtered);
                return rc;
            }
            for (plane = 0; plane < pic->nplanes; plane++) {
                const oplane *source = &pic->p[plane];
                oplane *dest = &filtered.p[plane];
                int chroma = plane != 0;
                int sx = chroma && d->subsampling_x, sy = chroma && d->subsampling_y;
                int64_t x = (c * 4) >> sx, y = (r * 4) >> sy;
                int width = 8 >> sx, height = 8 >> sy, nx, ny, ox, oy;
                vaco_av1pf_strength st;
                uint16_t primary, secondary, neigh[144], out[64];
                vaco_av1pf_filter_params fp;
                if (index >= 8) {
                    opic_free(&filtered);
                    ERR("CDEF strength index exceeds table");
                }
                st = chroma ? cp->uv[index] : cp->y[index];
                primary = (uint16_t)(st.primary << shift);
                secondary = (uint16_t)(st.secondary << shift);
                fp.direction = primary == 0 ? 0
                               : chroma ? orc_chroma_direction(direction, d->subsampling_x,
                                                               d->subsampling_y)
                                        : direction;
                fp.primary = chroma ? primary : orc_adjust_strength(primary, variance);
                fp.secondary = secondary;
                memset(neigh, 0, sizeof neigh);
                for (ny = 0; ny < height + 4; ny++)
                    for (nx = 0; nx < width + 4; nx++)
                        neigh[ny * 12 + nx] = op_get(source, x + nx - 2, y + ny - 2);
                fp.edges = (uint8_t)((x > 0) | ((x + width < source->w) << 1) |
                                     ((y > 0) << 2) | ((y + height < source->h) << 3));
                fp.width = (uint32_t)width;
                fp.height = (uint32_t)height;
                fp.bit_depth = d->bit_depth;
                /* u8 arithmetic, decode.rs:842 */
                fp.damping = (uint8_t)(cp->damping + shift - (chroma ? 1 : 0));
                rc = orc_filter_block(neigh, &fp, out, err);
                if (rc != 0) {
                    opic_free(&filtered);
                    return rc;
                }
                for (oy = 0; oy < height; oy++)
                    for (ox = 0; ox < width; ox++)
                        op_set(dest, x + ox, y + oy, out[oy * 8 + ox]);
            }
        }
    }
    opic_free(pic);
    *pic = filtered;
    return 0;
}

/* upscale_picture, superres.rs:226-274. */
static int upscale_picture(const vaco_av1pf_frame_desc *d, opic *pic, uint64_t *fuel,
                           const char **err)
{
    opic out;
    int i, rc;
    vaco_av1pf_sr_plane_config luma, chroma;
    luma.visible_width = d->coded_width;
    luma.output_width = d->upscaled_width;
    luma.height = d->coded_height;
    luma.bit_depth = d->bit_depth;
    chroma.visible_width = (d->coded_width + (1u << d->subsampling_x) - 1) >> d->subsampling_x;
    chroma.output_width = (d->upscaled_width + (1u << d->subsampling_x) - 1) >> d->subsampling_x;
    chroma.height = (d->coded_height + (1u << d->subsampling_y) - 1) >> d->subsampling_y;
    chroma.bit_depth = d->bit_depth;
    out.nplanes = pic->nplanes;
    for (i = 0; i < 3; i++)
        out.p[i].data = NULL;
    for (i = 0; i < pic->nplanes; i++) {
        rc = upscale_op(&pic->p[i], i == 0 ? &luma : &chroma, &out.p[i], fuel, err);
        if (rc != 0) {
            opic_free(&out);
            return rc;
        }
    }
    opic_free(pic);
    *pic = out;
    return 0;
}

/* restoration_plane_config, decode.rs:587-610. */
static int lr_plane_config(const vaco_av1pf_frame_desc *d, int plane,
                           vaco_av1pf_lr_plane_config *c, const char **err)
{
    int chroma = plane != 0;
    int sx = chroma && d->subsampling_x, sy = chroma && d->subsampling_y;
    if (d->lr_unit_size[plane] == 0)
        ERR("AV1 active restoration has no unit size");
    c->width = d->upscaled_width >> sx;
    c->height = d->coded_height >> sy;
    c->bit_depth = d->bit_depth;
    c->unit_size = d->lr_unit_size[plane];
    c->subsampling_y = (uint8_t)(chroma && d->subsampling_y);
    return 0;
}

/* apply_restoration + restored_plane + copy_visible_plane, decode.rs:673-737. */
static int apply_restoration(const vaco_av1pf_frame_desc *d, const opic *before, opic *pic,
                             const char **err)
{
    opic out;
    int plane, rc;
    int64_t x, y;
    out.nplanes = pic->nplanes;
    for (plane = 0; plane < 3; plane++)
        out.p[plane].data = NULL;
    for (plane = 0; plane < pic->nplanes; plane++) {
        vaco_av1pf_lr_plane_config c;
        rc = lr_plane_config(d, plane, &c, err);
        if (rc == 0) {
            if (d->lr_type[plane] == VACO_AV1PF_RESTORE_NONE) {
                if (!op_alloc(&out.p[plane], c.width, c.height)) {
                    *err = "allocation failed";
                    rc = VACO_AV1PF_INVALID_DATA;
                } else {
                    for (y = 0; y < c.height; y++)
                        for (x = 0; x < c.width; x++)
                            op_set(&out.p[plane], x, y, op_get(&pic->p[plane], x, y));
                }
            } else {
                rc = restore_op(&before->p[plane], &pic->p[plane], &c, d->lr_units[plane],
                                d->lr_unit_count[plane], &out.p[plane], err);
            }
        }
        if (rc != 0) {
            opic_free(&out);
            return rc;
        }
    }
    opic_free(pic);
    *pic = out;
    return 0;
}

static void input_geometry(const vaco_av1pf_frame_desc *d, uint32_t w[3], uint32_t h[3])
{
    uint32_t mi_cols = 2 * ((d->coded_width + 7) >> 3);
    uint32_t mi_rows = 2 * ((d->coded_height + 7) >> 3);
    w[0] = mi_cols * 4;
    h[0] = mi_rows * 4;
    w[1] = w[2] = d->monochrome ? 0 : w[0] >> d->subsampling_x;
    h[1] = h[2] = d->monochrome ? 0 : h[0] >> d->subsampling_y;
}

int orc_frame_output_geometry(const vaco_av1pf_frame_desc *d, uint32_t w[3], uint32_t h[3])
{
    int i;
    if (lr_active(d)) {
        w[0] = d->upscaled_width;
        h[0] = d->coded_height;
        w[1] = w[2] = d->upscaled_width >> d->subsampling_x;
        h[1] = h[2] = d->coded_height >> d->subsampling_y;
    } else if (d->use_superres) {
        w[0] = d->upscaled_width;
        h[0] = d->coded_height;
        w[1] = w[2] = (d->upscaled_width + (1u << d->subsampling_x) - 1) >> d->subsampling_x;
        h[1] = h[2] = (d->coded_height + (1u << d->subsampling_y) - 1) >> d->subsampling_y;
    } else {
        input_geometry(d, w, h);
    }
    if (d->monochrome)
        for (i = 1; i < 3; i++)
            w[i] = h[i] = 0;
    return 0;
}

/* decode.rs:554-559: pre-CDEF copy, CDEF, superres, restoration.
 * With superres and restoration both active, the pre-CDEF image is upscaled
 * too (AV1 spec 7.17: restoration reads UpscaledCdefFrame and, at stripe
 * borders, the upscaled pre-CDEF frame). The reference decoder refuses that
 * combination before the stage runs (decode.rs:574-577). */
int orc_frame(const vaco_av1pf_frame_desc *d, const vaco_av1pf_picture *in,
              vaco_av1pf_picture *out, uint64_t *fuel, const char **err)
{
    opic pic, pre;
    int i, rc = 0, has_pre = 0;
    int64_t y;
    uint32_t ow[3], oh[3];
    pic.nplanes = d->monochrome ? 1 : 3;
    pre.nplanes = 0;
    for (i = 0; i < pic.nplanes; i++)
        pic.p[i] = op_from(&in->plane[i]);
    {
        /* own a copy so stages can free their inputs */
        opic tmp;
        if (!opic_copy(&pic, &tmp))
            ERR("allocation failed");
        pic = tmp;
    }
    if (lr_active(d)) {
        if (!opic_copy(&pic, &pre)) {
            opic_free(&pic);
            ERR("allocation failed");
        }
        has_pre = 1;
    }
    rc = apply_cdef(d, &pic, fuel, err);
    if (rc == 0 && d->use_superres) {
        rc = upscale_picture(d, &pic, fuel, err);
        if (rc == 0 && has_pre)
            rc = upscale_picture(d, &pre, NULL, err);
    }
    if (rc == 0 && has_pre)
        rc = apply_restoration(d, &pre, &pic, err);
    if (rc == 0) {
        orc_frame_output_geometry(d, ow, oh);
        for (i = 0; i < pic.nplanes; i++) {
            if (pic.p[i].w != ow[i] || pic.p[i].h != oh[i]) {
                *err = "internal geometry mismatch";
                rc = VACO_AV1PF_INVALID_DATA;
                break;
            }
            for (y = 0; y < pic.p[i].h; y++)
                memcpy(out->plane[i].data + y * out->plane[i].stride,
                       pic.p[i].data + y * pic.p[i].stride, (size_t)pic.p[i].w * 2);
        }
    }
    opic_free(&pic);
    if (has_pre)
        opic_free(&pre);
    return rc;
}

/* ------------------------------------------------------------------------
 * pic_to_frame (decode.rs:3795-3852) and blit (decode.rs:3863-3891)
 * ---------------------------------------------------------------------- */
void orc_pic_to_frame(const vaco_av1pf_frame_desc *d, const vaco_av1pf_picture *pic,
                      uint8_t *const out[3], const uint32_t stride_bytes[3])
{
    const int64_t w = d->upscaled_width, h = d->coded_height;
    const int64_t sx = d->subsampling_x, sy = d->subsampling_y;
    const int two_bytes = d->bit_depth > 8;
    const int np = d->monochrome ? 1 : 3;
    int p;
    for (p = 0; p < np; p++) {
        /* decode.rs:3827-3829: chroma plane size by div_ceil */
        const int64_t fw = p ? (w + (1 << sx) - 1) >> sx : w;
        const int64_t fh = p ? (h + (1 << sy) - 1) >> sy : h;
        const oplane src = op_from(&pic->plane[p]);
        /* blit over min(frame, plane) (decode.rs:3831-3850) */
        const int64_t bw = fw < src.w ? fw : src.w, bh = fh < src.h ? fh : src.h;
        int64_t x, y;
        for (y = 0; y < bh; y++) {
            uint8_t *row = out[p] + (size_t)y * stride_bytes[p];
            for (x = 0; x < bw; x++) {
                const uint16_t v = op_get(&src, x, y);
                if (two_bytes) {
                    row[2 * x] = (uint8_t)(v & 0xff);
                    row[2 * x + 1] = (uint8_t)(v >> 8);
                } else {
                    row[x] = v <= 255 ? (uint8_t)v : 0; /* u8::try_from(v).unwrap_or(0) */
                }
            }
        }
    }
}
