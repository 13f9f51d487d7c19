"""The stage driven by side information the library's own parameter
producers decoded, as a decoder would drive it: uncompressed-header bits for
cdef_params and lr_params (frame_header.rs:1034-1123) through
vaco_av1pf_parse_cdef_params / _parse_lr_params, and one tile of
entropy-coded restoration units and cdef_idx (decode.rs:1097-1164,
2087-2145) through vaco_av1pf_read_lr_units / vaco_av1pf_read_cdef, in the
reference's read order: per 64x64 superblock its restoration units, then its
32x32 blocks, each reading cdef_idx when it is not skip and its 64x64 unit
has none yet. The skip map comes from the MiCell grid through
vaco_av1pf_skip8_from_mi. The GPU output must equal the oracle's.

The header bits are written here MSB first from the syntax tables; the tile
bytes come from the arithmetic encoder of test_cpu_symbols, which is derived
from the decoder's interval rule and shares no code with the decoder.
"""
import random
import sys

import numpy as np
import pytest

import frames as F
from oracle import oracle as O
from test_cpu_symbols import Encoder, as_tuple, encode_units, lr_desc
from vaco_amd import abi, av1pf

pytestmark = pytest.mark.gpu

BLOCK_32X32, BLOCK_64X64 = 9, 12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class BitWriter:
    def __init__(self):
        self.bits = []

    def put(self, v, n):
        self.bits += [(v >> i) & 1 for i in range(n - 1, -1, -1)]

    def bytes(self):
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(int("".join(map(str, b[i:i + 8])), 2) for i in range(0, len(b), 8))


def write_headers(rng, w, cdef_bits, lr_syntax, shift, extra, uv_shift, chroma420):
    """cdef_params then lr_params for 64x64 superblocks; returns the bytes
    and the strengths written (secondary as coded, 0..3)."""
    damping_m3 = rng.randrange(4)
    w.put(damping_m3, 2)
    w.put(cdef_bits, 2)
    strengths = []
    for _ in range(1 << cdef_bits):
        s = [rng.randrange(16), rng.randrange(4), rng.randrange(16), rng.randrange(4)]
        w.put(s[0], 4)
        w.put(s[1], 2)
        w.put(s[2], 4)
        w.put(s[3], 2)
        strengths.append(s)
    for t in lr_syntax:
        w.put(t, 2)
    w.put(shift, 1)
    if shift:
        w.put(extra, 1)
    if chroma420 and any(lr_syntax[1:]):
        w.put(uv_shift, 1)
    return damping_m3, strengths


CASES = [
    # coded w, h, bit depth, cdef_bits, lr type syntax per plane, unit shift, extra, uv shift
    (200, 136, 8, 3, (1, 1, 1), 0, 0, 1),
    (331, 203, 10, 2, (2, 3, 1), 1, 1, 0),
    (96, 72, 8, 1, (3, 0, 2), 1, 0, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}-bd{c[2]}-bits{c[3]}")
def test_gpu_stage_on_side_info_decoded_by_the_parameter_producers(case):
    cw, ch, bd, cdef_bits, lr_syntax, shift, extra, uv_shift = case
    rng = random.Random(hash(case))
    say(f"side info: {case}")

    # ---- uncompressed header: cdef_params, lr_params
    bw = BitWriter()
    damping_m3, strengths = write_headers(rng, bw, cdef_bits, lr_syntax, shift, extra, uv_shift, True)
    bits = av1pf.Bits(bw.bytes())
    seq = av1pf.seq_info(cdef=True, restoration=True)
    got_bits, cdef = av1pf.parse_cdef_params(bits, seq)
    types, units = av1pf.parse_lr_params(bits, seq)
    assert bits.pos == len(bw.bits) and not bits.overrun
    assert got_bits == cdef_bits and cdef.enabled and cdef.damping == damping_m3 + 3
    for i, s in enumerate(strengths):
        sec = lambda v: 4 if v == 3 else v
        assert (cdef.y[i].primary, cdef.y[i].secondary, cdef.uv[i].primary, cdef.uv[i].secondary) == \
            (s[0], sec(s[1]), s[2], sec(s[3]))
    assert list(types) == list(lr_syntax)
    luma_unit = 64 << (shift + extra)
    assert units[0] == luma_unit

    # ---- frame desc from the parsed fields
    d, counts = lr_desc(cw, ch, types, units)
    d.bit_depth = bd
    d.cdef = cdef

    # ---- one tile: restoration units and cdef_idx per superblock
    mi_cols, mi_rows = 2 * ((cw + 7) >> 3), 2 * ((ch + 7) >> 3)
    cdef_stride, cdef_rows = (mi_cols + 15) // 16, (mi_rows + 15) // 16
    block_skip = {}
    want_idx = np.full((cdef_rows, cdef_stride), -1, np.int8)

    def blocks_of(r, c):
        return [(r + dr, c + dc) for dr in (0, 8) for dc in (0, 8) if r + dr < mi_rows and c + dc < mi_cols]

    def per_sb(enc, r, c):
        for (br, bc) in blocks_of(r, c):
            skip = rng.random() < 0.3
            block_skip[(br, bc)] = skip
            if not skip and want_idx[br >> 4, bc >> 4] < 0:
                v = rng.randrange(1 << cdef_bits)
                enc.literal(v, cdef_bits)
                want_idx[br >> 4, bc >> 4] = v

    data, want_units, order = encode_units(rng, d, counts, False, per_sb)

    maps = [(abi.LrUnit * (cc * rr))() if d.lr_type[p] else None for p, (cc, rr) in enumerate(counts)]
    idx = np.full((cdef_rows, cdef_stride), -1, np.int8)
    sd = av1pf.SymbolDecoder(data, False)
    st = av1pf.lr_tile_state()
    for (r, c) in order:
        av1pf.read_lr_units(sd, st, d, r, c, BLOCK_64X64, maps)
        for (br, bc) in blocks_of(r, c):
            av1pf.read_cdef(sd, True, cdef_bits, idx, br, bc, BLOCK_32X32, block_skip[(br, bc)])
    assert not sd.overrun
    assert np.array_equal(idx, want_idx)
    for p in range(3):
        if d.lr_type[p]:
            cols = counts[p][0]
            for (ur, uc), wu in want_units[p].items():
                assert as_tuple(maps[p][ur * cols + uc]) == wu

    # ---- skip map from the MiCell grid, planes, and the stage
    mi_skip = np.zeros((mi_rows, mi_cols), np.uint8)
    for (br, bc), skip in block_skip.items():
        mi_skip[br:br + 8, bc:bc + 8] = skip
    skip8 = av1pf.skip8_from_mi(mi_skip)
    nprng = np.random.default_rng(hash(case) & 0xFFFF)
    _, _, geo = F.padded_geometry(cw, ch, 1, 1, False)
    planes = [F.textured_plane(nprng, hh, ww, bd) for (hh, ww) in geo]
    d.cdef_idx = idx.ctypes.data_as(abi.C.POINTER(abi.C.c_int8))
    d.cdef_stride = cdef_stride
    d.skip8 = skip8.ctypes.data_as(abi.C.POINTER(abi.C.c_uint8))
    for p in range(3):
        if maps[p] is not None:
            d.lr_units[p] = abi.C.cast(maps[p], abi.C.POINTER(abi.LrUnit))
    frame = av1pf.Frame(planes, d, keep=(idx, skip8, maps))
    want = O.frame(d, planes)
    with av1pf.Context(512, 256) as ctx:
        got = ctx.frame(frame)
    for p, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape and np.array_equal(g, w), f"plane {p}"
