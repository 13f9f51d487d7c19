// av1pf_syntax.cpp - the producers of the stage's parameters: the
// uncompressed-header syntax of CDEF, loop restoration and superres (AV1
// 5.9.8, 5.9.19, 5.9.20), and the entropy-coded per-block side info (the
// symbol decoder, cdef_idx, restoration units; AV1 8.2, 5.11.56-5.11.58).
// Host code; it fills the fields of vaco_av1pf_frame_desc that a decoder
// would otherwise copy from its own parsers.
//
// Reference (matteopolak/vaco, crates/codec/vaco-codec-av1/src/
// frame_header.rs): parse_superres_params 687-704, parse_cdef_params
// 1034-1063 with read_cdef_strength 1065-1072, parse_lr_params 1074-1123;
// the bit reader's contract (MSB first, zeros past the end, overrun
// recorded) is crates/core/vaco-bitstream/src/reader.rs:72-210.
#include <string.h>

#include <algorithm>

#include "../../include/vaco_av1pf.h"
#include "av1pf_host.h"

extern "C" void vaco_av1pf_bits_init(vaco_av1pf_bit_reader *r, const uint8_t *data, size_t len)
{
    r->data = data;
    r->size_bits = data ? (uint64_t)len * 8 : 0;
    r->pos = 0;
    r->overrun = 0;
}

extern "C" uint32_t vaco_av1pf_bits_get(vaco_av1pf_bit_reader *r, uint32_t n)
{
    if (n > 32)
        n = 32;
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; i++, r->pos++) {
        uint32_t bit = 0;
        if (r->pos < r->size_bits)
            bit = (r->data[r->pos >> 3] >> (7 - (r->pos & 7))) & 1u;
        else
            r->overrun = 1;
        v = (v << 1) | bit;
    }
    return v;
}

// frame_header.rs:1065-1072: 4-bit primary, 2-bit secondary with 3 -> 4.
static vaco_av1pf_strength read_strength(vaco_av1pf_bit_reader *r)
{
    vaco_av1pf_strength s;
    s.primary = (uint8_t)vaco_av1pf_bits_get(r, 4);
    const uint32_t sec = vaco_av1pf_bits_get(r, 2);
    s.secondary = (uint8_t)(sec == 3 ? 4 : sec);
    return s;
}

extern "C" uint32_t vaco_av1pf_parse_cdef_params(vaco_av1pf_bit_reader *r,
                                                 const vaco_av1pf_seq_info *seq, int coded_lossless,
                                                 int allow_intrabc, uint32_t num_planes,
                                                 vaco_av1pf_cdef_params *out)
{
    memset(out, 0, sizeof *out);
    if (coded_lossless || allow_intrabc || !seq->enable_cdef)
        return 0;  // CDEF off, no bits read
    out->enabled = 1;
    out->damping = (uint8_t)(vaco_av1pf_bits_get(r, 2) + 3);
    const uint32_t bits = vaco_av1pf_bits_get(r, 2);
    for (uint32_t i = 0; i < (1u << bits); i++) {
        out->y[i] = read_strength(r);
        if (num_planes > 1)
            out->uv[i] = read_strength(r);
    }
    return bits;
}

extern "C" void vaco_av1pf_parse_lr_params(vaco_av1pf_bit_reader *r, const vaco_av1pf_seq_info *seq,
                                           int all_lossless, int allow_intrabc, uint32_t num_planes,
                                           uint8_t types[3], uint32_t unit_size[3])
{
    for (int p = 0; p < 3; p++) {
        types[p] = VACO_AV1PF_RESTORE_NONE;
        unit_size[p] = 0;
    }
    if (all_lossless || allow_intrabc || !seq->enable_restoration)
        return;
    bool uses_lr = false, uses_chroma_lr = false;
    for (uint32_t p = 0; p < num_planes && p < 3; p++) {
        // lr_type codes 0..3 are None, Switchable, Wiener, SelfGuided: the
        // VACO_AV1PF_RESTORE_* values.
        const uint32_t t = vaco_av1pf_bits_get(r, 2);
        types[p] = (uint8_t)t;
        if (t != 0) {
            uses_lr = true;
            if (p > 0)
                uses_chroma_lr = true;
        }
    }
    if (!uses_lr)
        return;
    uint32_t shift = vaco_av1pf_bits_get(r, 1);
    if (seq->use_128x128_superblock)
        shift += 1;
    else if (shift)
        shift += vaco_av1pf_bits_get(r, 1);
    const uint32_t size = 64u << shift;
    const uint32_t uv_shift =
        (seq->subsampling_x && seq->subsampling_y && uses_chroma_lr) ? vaco_av1pf_bits_get(r, 1) : 0;
    unit_size[0] = size;
    unit_size[1] = unit_size[2] = size >> uv_shift;
}

extern "C" uint32_t vaco_av1pf_parse_superres_params(vaco_av1pf_bit_reader *r,
                                                     const vaco_av1pf_seq_info *seq,
                                                     uint32_t upscaled_width, uint8_t *use_superres)
{
    const bool use = seq->enable_superres && vaco_av1pf_bits_get(r, 1) != 0;
    *use_superres = use ? 1 : 0;
    if (!use)
        return upscaled_width;
    const uint32_t denom = vaco_av1pf_bits_get(r, 3) + 9;  // SUPERRES_DENOM_MIN
    return (upscaled_width * 8 + denom / 2) / denom;      // SUPERRES_NUM = 8
}

// ------------------------------------------------------ symbol decoder --
//
// AV1 8.2 as the reference restates it (crates/signal/vaco-codec-msac/src/
// av1.rs): init_symbol 48-65, read_symbol 80-123 (its adaptation 195-216),
// read_bool 129-132, read_literal 135-142, exit_symbol 152-157, overrun
// 161-165. Raw bits come from the header bit reader above (zeros past the
// end, overrun = a read past the end).

namespace {

constexpr uint32_t kProbShift = 6;  // EC_PROB_SHIFT
constexpr uint32_t kMinProb = 4;    // EC_MIN_PROB

uint32_t floor_log2(uint32_t v) { return v ? 31u - (uint32_t)__builtin_clz(v) : 0u; }

void adapt(uint16_t *cdf, uint32_t symbol, uint32_t n)
{
    const uint16_t count = cdf[n];
    const uint32_t rate = 3 + (count > 15) + (count > 31) + std::min(floor_log2(n), 2u);
    uint32_t tmp = 0;
    for (uint32_t i = 0; i + 1 < n; i++) {
        if (i == symbol)
            tmp = 1u << 15;
        const uint32_t cur = cdf[i];
        cdf[i] = (uint16_t)(tmp < cur ? cur - ((cur - tmp) >> rate) : cur + ((tmp - cur) >> rate));
    }
    if (cdf[n] < 32)
        cdf[n]++;
}

}  // namespace

extern "C" void vaco_av1pf_sd_init(vaco_av1pf_symbol_decoder *sd, const uint8_t *data, size_t size,
                                   int disable_cdf_update)
{
    vaco_av1pf_bits_init(&sd->bits, data, size);
    const uint64_t total = data ? (uint64_t)size * 8 : 0;
    const uint32_t num_bits = (uint32_t)std::min<uint64_t>(total, 15);
    const uint32_t buf = vaco_av1pf_bits_get(&sd->bits, num_bits);
    sd->value = ((1u << 15) - 1) ^ (buf << (15 - num_bits));
    sd->range = 1u << 15;
    sd->max_bits = (int64_t)total - 15;
    sd->disable_cdf_update = disable_cdf_update ? 1 : 0;
}

extern "C" uint32_t vaco_av1pf_sd_read_symbol(vaco_av1pf_symbol_decoder *sd, uint16_t *cdf,
                                              uint32_t n)
{
    // The first symbol whose scaled threshold the value reaches; the last
    // threshold is 0, so the search always ends by symbol n - 1.
    uint32_t cur = sd->range, prev, symbol = 0;
    for (;; symbol++) {
        prev = cur;
        const uint32_t f = (1u << 15) - (symbol < n ? cdf[symbol] : (1u << 15));
        cur = ((sd->range >> 8) * (f >> kProbShift)) >> (7 - kProbShift);
        cur += kMinProb * (n > symbol + 1 ? n - symbol - 1 : 0);
        if (sd->value >= cur)
            break;
    }
    sd->range = prev - cur;
    sd->value -= cur;
    // Renormalise: the range back to 15 bits, new bits (zeros once
    // SymbolMaxBits runs out) shifted into the inverted value.
    const uint32_t bits = 15 - floor_log2(sd->range);
    sd->range <<= bits;
    const uint32_t num_bits =
        (uint32_t)std::min<int64_t>(bits, std::max<int64_t>(sd->max_bits, 0));
    const uint32_t fresh = vaco_av1pf_bits_get(&sd->bits, num_bits) << (bits - num_bits);
    sd->value = fresh ^ (((sd->value + 1) << bits) - 1);
    sd->max_bits -= bits;
    if (!sd->disable_cdf_update)
        adapt(cdf, symbol, n);
    return symbol;
}

extern "C" uint32_t vaco_av1pf_sd_read_bool(vaco_av1pf_symbol_decoder *sd)
{
    uint16_t cdf[3] = {1u << 14, 1u << 15, 0};
    return vaco_av1pf_sd_read_symbol(sd, cdf, 2);
}

extern "C" uint32_t vaco_av1pf_sd_read_literal(vaco_av1pf_symbol_decoder *sd, uint32_t n)
{
    uint32_t x = 0;
    for (uint32_t i = 0; i < n; i++)
        x = (x << 1) | vaco_av1pf_sd_read_bool(sd);
    return x;
}

extern "C" void vaco_av1pf_sd_exit(vaco_av1pf_symbol_decoder *sd)
{
    if (sd->max_bits > 0) {
        sd->bits.pos += (uint64_t)sd->max_bits;
        if (sd->bits.pos > sd->bits.size_bits)
            sd->bits.overrun = 1;
    }
}

extern "C" int vaco_av1pf_sd_overrun(const vaco_av1pf_symbol_decoder *sd)
{
    return sd->max_bits < -14 || sd->bits.overrun;
}

// ------------------------------------------- per-block side-info syntax --
//
// read_restoration_units and the unit syntax (decode.rs:1097-1386) and
// read_cdef (decode.rs:2087-2145). Block dimensions in 4x4 units are
// Num_4x4_Blocks_Wide / _High (tables/conversion.rs:37-44).

namespace {

constexpr uint8_t kNum4x4Wide[22] = {1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 1, 4, 2, 8, 4, 16};
constexpr uint8_t kNum4x4High[22] = {1, 2, 1, 2, 4, 2, 4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 4, 1, 8, 2, 16, 4};

// Whether self-guided set `set` runs pass `pass` (its radius is nonzero,
// restoration.rs:286-312): pass 0 is off for sets 10..13, pass 1 for 14, 15.
bool sgr_pass_used(uint32_t set, int pass)
{
    if (set > 15)
        return false;
    return pass == 0 ? (set < 10 || set > 13) : set < 14;
}

struct Syntax {
    vaco_av1pf_symbol_decoder *sd;
    const char *err = nullptr;

    bool fail(const char *m)
    {
        err = m;
        return false;
    }

    // decode.rs:1368-1386
    bool ns(uint32_t symbols, uint32_t *out)
    {
        if (symbols == 0)
            return fail("AV1 non-power-of-two symbols");
        if (symbols == 1) {
            *out = 0;
            return true;
        }
        const uint32_t width = 32 - (uint32_t)__builtin_clz(symbols - 1);
        const uint32_t m = (1u << width) - symbols;
        const uint32_t v = vaco_av1pf_sd_read_literal(sd, width - 1);
        *out = v < m ? v : v * 2 - m + vaco_av1pf_sd_read_bool(sd);
        return true;
    }

    // decode.rs:1340-1366
    bool subexp(uint32_t symbols, uint32_t k, uint32_t *out)
    {
        if (symbols == 0 || k >= 31)
            return fail("AV1 subexp parameters");
        uint32_t i = 0, mk = 0;
        for (;;) {
            const uint32_t b2 = i == 0 ? k : k + i - 1;
            if (b2 >= 31)
                return fail("AV1 subexp bit width");
            const uint32_t a = 1u << b2;
            if ((uint64_t)symbols <= (uint64_t)mk + 3ull * a) {
                uint32_t v;
                if (!ns(symbols - mk, &v))
                    return false;
                *out = v + mk;
                return true;
            }
            if (vaco_av1pf_sd_read_bool(sd) == 0) {
                *out = vaco_av1pf_sd_read_literal(sd, b2) + mk;
                return true;
            }
            i++;
            mk += a;
        }
    }

    static int32_t inverse_recenter(int32_t r, int32_t v)
    {
        if (v > 2 * r)
            return v;
        return (v & 1) ? r - ((v + 1) >> 1) : r + (v >> 1);
    }

    // decode.rs:1293-1334: a value in [low, high] coded against `ref`.
    bool signed_subexp(int32_t low, int32_t high, uint32_t k, int32_t ref, int16_t *out)
    {
        const int32_t count = high - low + 1, r = ref - low;
        if (count <= 0 || r < 0 || r >= count)
            return fail("AV1 subexp reference range");
        uint32_t v;
        if (!subexp((uint32_t)count, k, &v))
            return false;
        const int32_t max = count - 1;
        const int32_t x = 2 * r <= count ? inverse_recenter(r, (int32_t)v)
                                         : max - inverse_recenter(max - r, (int32_t)v);
        *out = (int16_t)(x + low);
        return true;
    }

    // decode.rs:1221-1264
    bool wiener(vaco_av1pf_lr_tile_state *st, int plane, vaco_av1pf_lr_unit *u)
    {
        static const int32_t lo[3] = {-5, -23, -17}, hi[3] = {10, 8, 46};
        int16_t f[2][3];
        for (int pass = 0; pass < 2; pass++)
            for (int c = 0; c < 3; c++) {
                if (plane != 0 && c == 0) {
                    f[pass][c] = 0;
                    continue;
                }
                if (!signed_subexp(lo[c], hi[c], (uint32_t)c + 1, st->ref_lr_wiener[plane][pass][c],
                                   &f[pass][c]))
                    return false;
            }
        memcpy(st->ref_lr_wiener[plane], f, sizeof f);
        u->kind = VACO_AV1PF_UNIT_WIENER;
        memcpy(u->vertical, f[0], sizeof f[0]);
        memcpy(u->horizontal, f[1], sizeof f[1]);
        return true;
    }

    // decode.rs:1266-1291. The coefficients are coded with k = 3, as the
    // reference codes them.
    bool sgrproj(vaco_av1pf_lr_tile_state *st, int plane, vaco_av1pf_lr_unit *u)
    {
        const uint32_t set = vaco_av1pf_sd_read_literal(sd, 4);
        int16_t xqd[2] = {0, 0};
        for (int pass = 0; pass < 2; pass++) {
            if (sgr_pass_used(set, pass)) {
                if (!signed_subexp(pass == 0 ? -96 : -32, pass == 0 ? 31 : 95, 3,
                                   st->ref_sgr_xqd[plane][pass], &xqd[pass]))
                    return false;
            } else if (pass == 1) {
                xqd[1] = (int16_t)std::min(std::max(128 - (int32_t)xqd[0], -32), 95);
            }
        }
        memcpy(st->ref_sgr_xqd[plane], xqd, sizeof xqd);
        u->kind = VACO_AV1PF_UNIT_SELF_GUIDED;
        u->set = (uint8_t)set;
        u->xqd[0] = xqd[0];
        u->xqd[1] = xqd[1];
        return true;
    }

    // decode.rs:1166-1219
    bool unit(vaco_av1pf_lr_tile_state *st, int plane, uint8_t mode, vaco_av1pf_lr_unit *u)
    {
        memset(u, 0, sizeof *u);
        uint8_t sel = VACO_AV1PF_RESTORE_NONE;
        switch (mode) {
        case VACO_AV1PF_RESTORE_WIENER:
            if (vaco_av1pf_sd_read_symbol(sd, st->use_wiener, 2))
                sel = VACO_AV1PF_RESTORE_WIENER;
            break;
        case VACO_AV1PF_RESTORE_SELF_GUIDED:
            if (vaco_av1pf_sd_read_symbol(sd, st->use_sgrproj, 2))
                sel = VACO_AV1PF_RESTORE_SELF_GUIDED;
            break;
        case VACO_AV1PF_RESTORE_SWITCHABLE:
            switch (vaco_av1pf_sd_read_symbol(sd, st->restoration_type, 3)) {
            case 0: break;
            case 1: sel = VACO_AV1PF_RESTORE_WIENER; break;
            case 2: sel = VACO_AV1PF_RESTORE_SELF_GUIDED; break;
            default: return fail("AV1 restoration type symbol");
            }
            break;
        default:
            break;
        }
        if (sel == VACO_AV1PF_RESTORE_WIENER)
            return wiener(st, plane, u);
        if (sel == VACO_AV1PF_RESTORE_SELF_GUIDED)
            return sgrproj(st, plane, u);
        u->kind = VACO_AV1PF_UNIT_NONE;
        return true;
    }
};

uint32_t ceil_div(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

extern "C" void vaco_av1pf_lr_tile_state_init(vaco_av1pf_lr_tile_state *st)
{
    memset(st, 0, sizeof *st);
    const uint16_t w[3] = {11570, 32768, 0}, s[3] = {16855, 32768, 0}, t[4] = {9413, 22581, 32768, 0};
    memcpy(st->use_wiener, w, sizeof w);
    memcpy(st->use_sgrproj, s, sizeof s);
    memcpy(st->restoration_type, t, sizeof t);
    for (int p = 0; p < 3; p++) {
        st->ref_sgr_xqd[p][0] = -32;
        st->ref_sgr_xqd[p][1] = 31;
        for (int pass = 0; pass < 2; pass++) {
            st->ref_lr_wiener[p][pass][0] = 3;
            st->ref_lr_wiener[p][pass][1] = -7;
            st->ref_lr_wiener[p][pass][2] = 15;
        }
    }
}

extern "C" vaco_av1pf_status vaco_av1pf_read_lr_units(vaco_av1pf_symbol_decoder *sd,
                                                      vaco_av1pf_lr_tile_state *st,
                                                      const vaco_av1pf_frame_desc *d, uint32_t mi_row,
                                                      uint32_t mi_col, uint8_t b_size,
                                                      vaco_av1pf_lr_unit *units[3], const char **msg)
{
    if (msg)
        *msg = "";
    if (!sd || !st || !d || !units) {
        if (msg)
            *msg = "null argument";
        return VACO_AV1PF_INVALID_ARGUMENT;
    }
    auto invalid = [&](const char *m) {
        if (msg)
            *msg = m;
        return VACO_AV1PF_INVALID_DATA;
    };
    if (d->lr_type[0] == VACO_AV1PF_RESTORE_NONE && d->lr_type[1] == VACO_AV1PF_RESTORE_NONE &&
        d->lr_type[2] == VACO_AV1PF_RESTORE_NONE)
        return VACO_AV1PF_OK;
    const uint32_t w4 = b_size < 22 ? kNum4x4Wide[b_size] : 0;
    const uint32_t h4 = b_size < 22 ? kNum4x4High[b_size] : 0;
    if (w4 == 0 || h4 == 0)
        return invalid("AV1 restoration superblock geometry");
    Syntax syn{sd};
    const int num_planes = d->monochrome ? 1 : 3;
    for (int plane = 0; plane < num_planes; plane++) {
        const uint8_t mode = d->lr_type[plane];
        if (mode == VACO_AV1PF_RESTORE_NONE)
            continue;
        const uint32_t unit = d->lr_unit_size[plane];
        if (unit == 0)
            return invalid("AV1 active restoration has no unit size");
        const uint32_t sx = plane && d->subsampling_x, sy = plane && d->subsampling_y;
        uint32_t cols, rows;
        av1pf::lr_unit_counts(d->upscaled_width >> sx, d->coded_height >> sy, unit, &cols, &rows);
        const uint64_t sr = ((uint64_t)mi_row * 4) >> sy, sc = ((uint64_t)mi_col * 4) >> sx;
        const uint64_t sh = ((uint64_t)h4 * 4) >> sy, sw = ((uint64_t)w4 * 4) >> sx;
        const uint32_t r0 = std::min(ceil_div(sr, unit), rows), r1 = std::min(ceil_div(sr + sh, unit), rows);
        const uint32_t c0 = std::min(ceil_div(sc, unit), cols), c1 = std::min(ceil_div(sc + sw, unit), cols);
        for (uint32_t ur = r0; ur < r1; ur++)
            for (uint32_t uc = c0; uc < c1; uc++) {
                vaco_av1pf_lr_unit u;
                if (!syn.unit(st, plane, mode, &u))
                    return invalid(syn.err);
                if (!units[plane])
                    return invalid("AV1 restoration plane unit map");
                const uint64_t slot = (uint64_t)ur * cols + uc;
                if (slot >= d->lr_unit_count[plane])
                    return invalid("AV1 restoration unit map length");
                units[plane][slot] = u;
            }
    }
    return VACO_AV1PF_OK;
}

extern "C" void vaco_av1pf_read_cdef(vaco_av1pf_symbol_decoder *sd, int cdef_enabled,
                                     uint32_t cdef_bits, int8_t *cdef_idx, uint32_t cdef_stride,
                                     uint32_t cdef_rows, uint32_t mi_row, uint32_t mi_col,
                                     uint8_t mi_size, int skip)
{
    if (skip || !cdef_enabled)
        return;
    const uint64_t len = cdef_idx ? (uint64_t)cdef_stride * cdef_rows : 0;
    const uint32_t ur = mi_row & ~15u, uc = mi_col & ~15u;
    const uint64_t idx = (uint64_t)(ur / 16) * cdef_stride + uc / 16;
    // An index outside the map reads as -1 (not read yet), as the reference's
    // checked read does, so the bits are still consumed.
    if (idx < len && cdef_idx[idx] != -1)
        return;
    const uint32_t value = vaco_av1pf_sd_read_literal(sd, cdef_bits);
    const uint32_t w4 = mi_size < 22 ? kNum4x4Wide[mi_size] : 1;
    const uint32_t h4 = mi_size < 22 ? kNum4x4High[mi_size] : 1;
    for (uint32_t i = ur; i < ur + h4; i += 16)
        for (uint32_t j = uc; j < uc + w4; j += 16) {
            const uint64_t slot = (uint64_t)(i / 16) * cdef_stride + j / 16;
            if (slot < len)
                cdef_idx[slot] = (int8_t)value;
        }
}
