// av1pf_gen.cpp - seeded synthetic frames and side information for the AV1
// filter stage (include/vaco_av1pf_gen.h). Integer arithmetic only.
#include <algorithm>
#include <cstring>

#include "../../include/vaco_av1pf_gen.h"

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t s;
    uint64_t next() { return splitmix64(s++ * 0x2545F4914F6CDD1Dull); }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
    int32_t range(int32_t lo, int32_t hi) { return lo + (int32_t)below((uint32_t)(hi - lo + 1)); }
};

// cos/sin of k*pi/32 in Q12.
const int16_t kTrig[32][2] = {
    {4096, 0},     {4076, 401},   {4017, 799},   {3920, 1189},  {3784, 1567},  {3612, 1931},
    {3406, 2276},  {3166, 2598},  {2896, 2896},  {2598, 3166},  {2276, 3406},  {1931, 3612},
    {1567, 3784},  {1189, 3920},  {799, 4017},   {401, 4076},   {0, 4096},     {-401, 4076},
    {-799, 4017},  {-1189, 3920}, {-1567, 3784}, {-1931, 3612}, {-2276, 3406}, {-2598, 3166},
    {-2896, 2896}, {-3166, 2598}, {-3406, 2276}, {-3612, 1931}, {-3784, 1567}, {-3920, 1189},
    {-4017, 799},  {-4076, 401},
};

uint64_t frame_key(const vaco_av1pf_gen_config &c, uint32_t frame)
{
    return splitmix64(splitmix64(c.seed) ^ ((uint64_t)frame * 0xD1B54A32D192ED03ull));
}

uint32_t denom_of(const vaco_av1pf_gen_config &c, uint32_t frame)
{
    if (c.superres_denom == 0) {
        Rng r{frame_key(c, frame) ^ 0x5D5D5D5Dull};
        return 9 + r.below(8);
    }
    return c.superres_denom;
}

struct Edge {
    int32_t c, s, px, py, delta;
};

}  // namespace

extern "C" vaco_av1pf_status vaco_av1pf_gen_get_layout(const vaco_av1pf_gen_config *cfg, uint32_t frame,
                                                   vaco_av1pf_gen_layout *out)
{
    if (!cfg || !out || cfg->width == 0 || cfg->height == 0 ||
        !(cfg->bit_depth == 8 || cfg->bit_depth == 10 || cfg->bit_depth == 12))
        return VACO_AV1PF_INVALID_ARGUMENT;
    memset(out, 0, sizeof *out);
    const uint32_t d = denom_of(*cfg, frame);
    if (d < 8 || d > 16)
        return VACO_AV1PF_INVALID_ARGUMENT;
    out->superres_denom = d;
    out->upscaled_width = cfg->width;
    // frame_header.rs:687-704 (coded width from the superres denominator)
    out->coded_width = d == 8 ? cfg->width : (cfg->width * 8 + d / 2) / d;
    out->coded_height = cfg->height;
    const uint32_t mi_cols = 2 * ((out->coded_width + 7) >> 3);
    const uint32_t mi_rows = 2 * ((out->coded_height + 7) >> 3);
    out->plane_width[0] = mi_cols * 4;
    out->plane_height[0] = mi_rows * 4;
    const int np = cfg->monochrome ? 1 : 3;
    for (int p = 1; p < np; p++) {
        out->plane_width[p] = out->plane_width[0] >> 1;
        out->plane_height[p] = out->plane_height[0] >> 1;
    }
    out->cdef_stride = (mi_cols + 15) / 16;
    out->cdef_idx_count = out->cdef_stride * ((mi_rows + 15) / 16);
    out->skip8_count = (mi_cols / 2) * (mi_rows / 2);
    if (cfg->lr_enabled) {
        for (int p = 0; p < np; p++) {
            const uint32_t w = cfg->width >> (p ? 1 : 0), h = cfg->height >> (p ? 1 : 0);
            const uint32_t u = cfg->lr_unit_size;
            const uint32_t cols = std::max<uint32_t>(1, (w + u / 2) / u);
            const uint32_t rows = std::max<uint32_t>(1, (h + u / 2) / u);
            out->lr_unit_count[p] = cols * rows;
        }
    }
    return VACO_AV1PF_OK;
}

extern "C" vaco_av1pf_status vaco_av1pf_gen_frame(const vaco_av1pf_gen_config *cfg, uint32_t frame,
                                                  vaco_av1pf_frame_desc *desc,
                                                  vaco_av1pf_picture *planes, int8_t *cdef_idx,
                                                  uint8_t *skip8, vaco_av1pf_lr_unit *units0,
                                                  vaco_av1pf_lr_unit *units1,
                                                  vaco_av1pf_lr_unit *units2)
{
    vaco_av1pf_gen_layout L;
    vaco_av1pf_status st = vaco_av1pf_gen_get_layout(cfg, frame, &L);
    if (st != VACO_AV1PF_OK)
        return st;
    if (!desc || !planes || !cdef_idx || !skip8)
        return VACO_AV1PF_INVALID_ARGUMENT;
    const int np = cfg->monochrome ? 1 : 3;
    vaco_av1pf_lr_unit *units[3] = {units0, units1, units2};
    for (int p = 0; p < np; p++) {
        const vaco_av1pf_plane &pl = planes->plane[p];
        if (!pl.data || pl.width != L.plane_width[p] || pl.height != L.plane_height[p] ||
            pl.stride < pl.width)
            return VACO_AV1PF_INVALID_ARGUMENT;
        if (cfg->lr_enabled && !units[p])
            return VACO_AV1PF_INVALID_ARGUMENT;
    }
    const uint64_t fk = frame_key(*cfg, frame);
    const int bd = cfg->bit_depth;
    const int32_t maxv = (1 << bd) - 1;
    Rng rng{fk ^ 0xC0FFEEull};

    // Hard edges (luma coordinates), shared by all planes.
    Edge edges[255];
    const int ne = cfg->hard_edges;
    for (int k = 0; k < ne; k++) {
        const uint32_t a = rng.below(32);
        edges[k].c = kTrig[a][0];
        edges[k].s = kTrig[a][1];
        edges[k].px = (int32_t)rng.below(L.plane_width[0]);
        edges[k].py = (int32_t)rng.below(L.plane_height[0]);
        const int32_t mag = (24 + (int32_t)rng.below(48)) << (bd - 8);
        edges[k].delta = (rng.below(2) ? mag : -mag);
    }

    // Planes: gradient + Irwin-Hall noise + edges.
    for (int p = 0; p < np; p++) {
        const vaco_av1pf_plane &pl = planes->plane[p];
        const uint32_t W = pl.width, H = pl.height;
        const int sub = p ? 1 : 0;
        const uint64_t pk = splitmix64(fk ^ (0xA0761D6478BD642Full * (uint64_t)(p + 1)));
        const int64_t sigma = cfg->noise_sigma;
        for (uint32_t y = 0; y < H; y++) {
            uint16_t *row = pl.data + (size_t)y * pl.stride;
            for (uint32_t x = 0; x < W; x++) {
                const int32_t g8 = 48 + (int32_t)((x * 128u) / W) + (int32_t)((y * 64u) / H) +
                                   (p == 1 ? -16 : p == 2 ? 16 : 0);
                int64_t v = (int64_t)g8 << (bd - 8);
                uint64_t h = splitmix64(pk ^ (((uint64_t)y << 32) | x));
                int64_t s = 0;
                for (int k = 0; k < 3; k++) {
                    s += (int64_t)(h & 0xFFFF) + (int64_t)((h >> 16) & 0xFFFF) +
                         (int64_t)((h >> 32) & 0xFFFF) + (int64_t)(h >> 48);
                    h = splitmix64(h);
                }
                // 12 uniforms on [0, 65535]: mean 393210, sd ~65536.
                v += ((s - 393210) * sigma + (1 << 15)) >> 16;
                const int32_t lx = (int32_t)(x << sub), ly = (int32_t)(y << sub);
                for (int k = 0; k < ne; k++) {
                    const int64_t side = (int64_t)(lx - edges[k].px) * edges[k].c +
                                         (int64_t)(ly - edges[k].py) * edges[k].s;
                    if (side > 0)
                        v += edges[k].delta;
                }
                row[x] = (uint16_t)std::min<int64_t>(std::max<int64_t>(v, 0), maxv);
            }
        }
    }

    // CDEF side information (frame_header.rs:1034-1072 semantics).
    memset(desc, 0, sizeof *desc);
    desc->coded_width = L.coded_width;
    desc->upscaled_width = L.upscaled_width;
    desc->coded_height = L.coded_height;
    desc->use_superres = L.superres_denom != 8;
    desc->bit_depth = (uint8_t)bd;
    desc->subsampling_x = cfg->monochrome ? 0 : 1;
    desc->subsampling_y = cfg->monochrome ? 0 : 1;
    desc->monochrome = cfg->monochrome;
    desc->cdef.enabled = 1;
    desc->cdef.damping = (uint8_t)(3 + rng.below(4));
    static const uint8_t sec_of[4] = {0, 1, 2, 4};
    for (int i = 0; i < 8; i++) {
        desc->cdef.y[i].primary = (uint8_t)rng.below(16);
        desc->cdef.y[i].secondary = sec_of[rng.below(4)];
        desc->cdef.uv[i].primary = (uint8_t)rng.below(16);
        desc->cdef.uv[i].secondary = sec_of[rng.below(4)];
    }
    for (uint32_t i = 0; i < L.cdef_idx_count; i++)
        cdef_idx[i] = (int8_t)rng.below(8);
    for (uint32_t i = 0; i < L.skip8_count; i++)
        skip8[i] = rng.below(1000) < cfg->skip_permille ? 1 : 0;
    desc->cdef_idx = cdef_idx;
    desc->cdef_stride = L.cdef_stride;
    desc->skip8 = skip8;

    // Loop restoration (restoration.rs:46-91 ranges).
    if (cfg->lr_enabled) {
        for (int p = 0; p < np; p++) {
            desc->lr_type[p] = VACO_AV1PF_RESTORE_SWITCHABLE;
            desc->lr_unit_size[p] = cfg->lr_unit_size;
            desc->lr_units[p] = units[p];
            desc->lr_unit_count[p] = L.lr_unit_count[p];
            for (uint32_t i = 0; i < L.lr_unit_count[p]; i++) {
                vaco_av1pf_lr_unit &u = units[p][i];
                memset(&u, 0, sizeof u);
                u.kind = (uint8_t)rng.below(3);
                if (u.kind == VACO_AV1PF_UNIT_WIENER) {
                    for (int t = 0; t < 2; t++) {
                        int16_t *taps = t ? u.horizontal : u.vertical;
                        taps[0] = (int16_t)(p ? 0 : rng.range(-5, 10));
                        taps[1] = (int16_t)rng.range(-23, 8);
                        taps[2] = (int16_t)rng.range(-17, 46);
                    }
                } else if (u.kind == VACO_AV1PF_UNIT_SELF_GUIDED) {
                    u.set = (uint8_t)rng.below(16);
                    u.xqd[0] = (int16_t)rng.range(-96, 31);
                    u.xqd[1] = (int16_t)rng.range(-32, 95);
                }
            }
        }
    }
    return VACO_AV1PF_OK;
}
