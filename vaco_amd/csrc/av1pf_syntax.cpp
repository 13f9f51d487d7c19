// av1pf_syntax.cpp - the producers of the stage's frame-level parameters:
// the uncompressed-header syntax of CDEF, loop restoration and superres
// (AV1 5.9.8, 5.9.19, 5.9.20). Host code; it fills the fields of
// vaco_av1pf_frame_desc that a decoder would otherwise copy from its own
// header parser.
//
// Reference (matteopolak/vaco, crates/codec/vaco-codec-av1/src/
// frame_header.rs): parse_superres_params 687-704, parse_cdef_params
// 1034-1063 with read_cdef_strength 1065-1072, parse_lr_params 1074-1123;
// the bit reader's contract (MSB first, zeros past the end, overrun
// recorded) is crates/core/vaco-bitstream/src/reader.rs:72-210.
#include <string.h>

#include "../../include/vaco_av1pf.h"

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
