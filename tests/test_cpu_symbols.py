"""The entropy-coded side-info syntax (host code, no GPU): the AV1 symbol
decoder and the readers of cdef_idx and restoration units built on it.

Pins:
* the reference's libaom 3.14.1 trace of 640 symbols
  (crates/signal/vaco-codec-msac/tests/av1_oracle.rs:12-32) and its adaptive
  trace and unit tests (crates/signal/vaco-codec-msac/src/av1.rs:218-315);
* read_cdef's own tests (crates/codec/vaco-codec-av1/src/
  cdef_frame_tests.rs:107-131) and its fill rule;
* for the adaptive decoder, the sub-exponential codes and the
  restoration-unit syntax, round trips through an arithmetic encoder written
  here from the decoder's interval rule (AV1 8.2.6: a symbol k leaves the
  value in [cur_k, cur_{k-1}) of the range), with exact big-integer
  arithmetic, so it shares no renormalisation or bit-feeding code with the
  decoder under test.
"""
import ctypes as C
import random

import numpy as np
import pytest

import golden_inputs as G
from vaco_amd import abi, av1pf


# ------------------------------------------------------------ encoder ---

def adapt(cdf, k, n):
    """The adaptation rule of AV1 8.2.6 (cdf: n thresholds + counter)."""
    count = cdf[n]
    rate = 3 + (count > 15) + (count > 31) + min(n.bit_length() - 1, 2)
    tmp = 0
    for i in range(n - 1):
        if i == k:
            tmp = 1 << 15
        cdf[i] = cdf[i] - ((cdf[i] - tmp) >> rate) if tmp < cdf[i] else cdf[i] + ((tmp - cdf[i]) >> rate)
    if cdf[n] < 32:
        cdf[n] += 1


class Encoder:
    """Inverse of the AV1 symbol decoder: tracks the lower end L and size R
    of the interval the decoder's value must land in, scaled to the T bits
    the decoder has consumed; the stream is the complement of L over T bits
    (the decoder keeps its value inverted)."""

    def __init__(self, adaptive=True):
        self.L, self.R, self.T, self.adaptive = 0, 1 << 15, 15, adaptive

    def symbol(self, k, cdf):
        n = len(cdf) - 1
        R = self.R

        def cur(i):
            if i < 0:
                return R
            return (((R >> 8) * ((32768 - cdf[i]) >> 6)) >> 1) + 4 * (n - i - 1)

        lo, hi = cur(k), cur(k - 1)
        assert lo < hi
        self.L += lo
        self.R = hi - lo
        d = 15 - (self.R.bit_length() - 1)
        self.L <<= d
        self.R <<= d
        self.T += d
        if self.adaptive:
            adapt(cdf, k, n)

    def bool(self, b):
        self.symbol(b, [1 << 14, 1 << 15, 0])  # a fresh cdf each time, as read_bool

    def literal(self, v, n):
        for i in range(n - 1, -1, -1):
            self.bool((v >> i) & 1)

    def ns(self, v, n):
        if n == 1:
            return
        w = (n - 1).bit_length()
        m = (1 << w) - n
        if v < m:
            self.literal(v, w - 1)
        else:
            t = v + m
            self.literal(t >> 1, w - 1)
            self.bool(t & 1)

    def subexp(self, v, n, k):
        i = mk = 0
        while True:
            b2 = k if i == 0 else k + i - 1
            a = 1 << b2
            if n <= mk + 3 * a:
                self.ns(v - mk, n - mk)
                return
            more = v >= mk + a
            self.bool(int(more))
            if not more:
                self.literal(v - mk, b2)
                return
            i += 1
            mk += a

    def signed_subexp(self, x, low, high, k, ref):
        n, x, r = high - low + 1, x - low, ref - low

        def recenter(r, x):
            if x > 2 * r:
                return x
            return (x - r) << 1 if x >= r else ((r - x) << 1) - 1

        v = recenter(r, x) if 2 * r <= n else recenter(n - 1 - r, n - 1 - x)
        self.subexp(v, n, k)

    def finish(self) -> bytes:
        nbytes = (self.T + 7) // 8
        x = (1 << self.T) - 1 - self.L
        return (x << (8 * nbytes - self.T)).to_bytes(nbytes, "big") + b"\x00\x00"


# ------------------------------------------------------- symbol decoder ---

def test_libaom_trace_without_adaptation():
    # av1_oracle.rs:12-32
    data = bytes(((i * 73 + (i >> 2) * 19 + 0xB4) & 255) for i in range(512))
    expected = G.read("av1_libaom_symbols.bin")
    got = []
    for n in (2, 3, 4, 7, 16):
        cdf = av1pf.cdf_array([i * i * 32768 // (n * n) for i in range(1, n + 1)] + [0])
        sd = av1pf.SymbolDecoder(data, True)
        got += [sd.read_symbol(cdf) for _ in range(128)]
        assert not sd.overrun
    assert bytes(got) == expected


def test_adaptive_trace_and_disabled_update():
    # av1.rs:302-315: six symbols of an adapting two-symbol cdf
    sd = av1pf.SymbolDecoder(bytes([0xB4, 0x2F, 0x91, 0x0C]), False)
    cdf = av1pf.cdf_array([1 << 14, 1 << 15, 0])
    assert [sd.read_symbol(cdf) for _ in range(6)] == [1, 0, 1, 1, 1, 0]
    # av1.rs:283-293
    sd = av1pf.SymbolDecoder(bytes([0x12, 0x34, 0x56, 0x78]), True)
    cdf = av1pf.cdf_array([8192, 1 << 15, 0])
    sd.read_symbol(cdf)
    assert list(cdf) == [8192, 1 << 15, 0]


def test_padding_bits_and_overrun():
    # av1.rs:224-236: fourteen padding bits past the data, then overrun (sticky)
    sd = av1pf.SymbolDecoder(bytes([0]), True)
    assert not sd.overrun
    while sd.max_bits >= -14:
        assert not sd.overrun
        sd.read_bool()
    assert sd.overrun
    sd.exit()
    assert sd.overrun
    assert av1pf.SymbolDecoder(b"", True).overrun
    for n in range(8):  # av1.rs:265-275
        sd = av1pf.SymbolDecoder(bytes([0xA5] * n), False)
        for _ in range(32):
            sd.read_bool()
        sd.exit()


def test_exit_symbol_skips_the_unread_bits():
    enc = Encoder()
    for b in (1, 0, 1):
        enc.bool(b)
    data = enc.finish()
    sd = av1pf.SymbolDecoder(data, False)
    assert [sd.read_bool() for _ in range(3)] == [1, 0, 1]
    left = sd.max_bits
    pos = sd.sd.bits.pos
    sd.exit()
    assert sd.sd.bits.pos == pos + max(left, 0)


def test_literals_fold_bools_msb_first_and_keep_low_32_bits():
    # av1.rs:238-246 and 277-281
    data = bytes([0xA5] * 32)
    a, b = av1pf.SymbolDecoder(data, True), av1pf.SymbolDecoder(data, True)
    want = 0
    for _ in range(65):
        want = ((want << 1) | b.read_bool()) & 0xFFFFFFFF
    assert a.read_literal(65) == want
    data = bytes([0x3C, 0x91, 0x77, 0x02, 0xF0])
    a, b = av1pf.SymbolDecoder(data, False), av1pf.SymbolDecoder(data, False)
    v = a.read_literal(6)
    w = 0
    for _ in range(6):
        w = 2 * w + b.read_bool()
    assert v == w


def test_adapted_cdf_stays_sorted_and_terminated():
    # av1.rs:248-263
    for seed in range(64):
        data = bytes(((i * seed + seed) & 255) for i in range(64))
        sd = av1pf.SymbolDecoder(data, False)
        cdf = av1pf.cdf_array([8192, 16384, 24576, 1 << 15, 0])
        for _ in range(200):
            assert sd.read_symbol(cdf) < 4
            assert cdf[3] == 1 << 15 and cdf[0] <= cdf[1] <= cdf[2] and cdf[4] <= 32


@pytest.mark.parametrize("adaptive", [True, False])
def test_random_symbol_streams_round_trip(adaptive):
    rng = random.Random(5 + adaptive)
    for trial in range(40):
        n = rng.choice((2, 3, 4, 5, 8, 13, 16))
        # a random valid cdf: strictly increasing thresholds, last 32768
        cuts = sorted(rng.sample(range(1, 32768), n - 1))
        init = cuts + [32768, 0]
        syms = [rng.choice(range(n)) if rng.random() < 0.5 else rng.choice((0, n - 1)) for _ in range(300)]
        enc = Encoder(adaptive)
        ecdf = list(init)
        for s in syms:
            enc.symbol(s, ecdf)
        sd = av1pf.SymbolDecoder(enc.finish(), not adaptive)
        dcdf = av1pf.cdf_array(init)
        assert [sd.read_symbol(dcdf) for _ in syms] == syms, (trial, n)
        assert list(dcdf) == ecdf
        assert not sd.overrun


# ------------------------------------------------------------ read_cdef ---

BLOCK_8X8, BLOCK_64X64, BLOCK_128X128 = 3, 12, 15


def test_zero_index_bits_assign_entry_zero_on_the_first_non_skip_block():
    # cdef_frame_tests.rs:107-131
    idx = np.full((1, 1), -1, np.int8)
    sd = av1pf.SymbolDecoder(bytes(8), True)
    av1pf.read_cdef(sd, True, 0, idx, 0, 0, BLOCK_8X8, True)
    assert idx[0, 0] == -1
    av1pf.read_cdef(sd, True, 0, idx, 0, 0, BLOCK_8X8, False)
    assert idx[0, 0] == 0


def test_cdef_index_read_once_per_unit_and_spread_over_large_blocks():
    enc = Encoder()
    values = [5, 2, 7, 1]
    for v in values:
        enc.literal(v, 3)
    data = enc.finish()
    idx = np.full((3, 3), -1, np.int8)
    sd = av1pf.SymbolDecoder(data, False)
    # disabled CDEF reads nothing
    av1pf.read_cdef(sd, False, 3, idx, 0, 0, BLOCK_8X8, False)
    assert (idx == -1).all()
    # a 128x128 block at the origin fills its four 64x64 units
    av1pf.read_cdef(sd, True, 3, idx, 0, 0, BLOCK_128X128, False)
    assert (idx[:2, :2] == 5).all() and (idx[2, :] == -1).all() and (idx[:, 2] == -1).all()
    # a later block in a unit already read reads nothing
    av1pf.read_cdef(sd, True, 3, idx, 4, 20, BLOCK_8X8, False)
    # an 8x8 block in unit (0, 2)
    av1pf.read_cdef(sd, True, 3, idx, 12, 36, BLOCK_8X8, False)
    assert idx[0, 2] == 2 and (idx[2, :] == -1).all()
    # a unit outside the map still consumes its bits (the reference's checked read)
    av1pf.read_cdef(sd, True, 3, idx, 64, 0, BLOCK_8X8, False)
    av1pf.read_cdef(sd, True, 3, idx, 32, 32, BLOCK_64X64, False)
    assert idx[2, 2] == 1


# ------------------------------------------------- restoration units ---

WIENER_RANGE = [(-5, 10, 1), (-23, 8, 2), (-17, 46, 3)]


def sgr_uses(set_, pas):
    return (set_ < 10 or set_ > 13) if pas == 0 else set_ < 14


def lr_desc(w, h, types, units, ssx=1, ssy=1, mono=False):
    d = abi.FrameDesc()
    d.coded_width = d.upscaled_width = w
    d.coded_height = h
    d.bit_depth = 8
    d.subsampling_x, d.subsampling_y, d.monochrome = ssx, ssy, int(mono)
    counts = []
    for p in range(3):
        d.lr_type[p] = types[p]
        d.lr_unit_size[p] = units[p]
        sx, sy = (ssx, ssy) if p else (0, 0)
        pw, ph = w >> sx, h >> sy
        u = max(units[p], 1)
        cols, rows = max(1, (pw + u // 2) // u), max(1, (ph + u // 2) // u)
        counts.append((cols, rows))
        d.lr_unit_count[p] = cols * rows
    return d, counts


def encode_units(rng, d, counts, sb128, per_sb=None):
    """Random units per plane, and the tile bytes coding them in the order
    the decoder reads them: superblocks in raster order, planes in order, then
    the units whose top-left sample lies inside the superblock. per_sb(enc,
    r, c), if given, codes the rest of each superblock after its units."""
    mi_cols = 2 * ((d.coded_width + 7) >> 3)
    mi_rows = 2 * ((d.coded_height + 7) >> 3)
    sb4 = 32 if sb128 else 16
    nplanes = 1 if d.monochrome else 3
    want = [dict() for _ in range(3)]
    enc = Encoder()
    cdf_w, cdf_s, cdf_t = [11570, 32768, 0], [16855, 32768, 0], [9413, 22581, 32768, 0]
    ref_w = [[[3, -7, 15], [3, -7, 15]] for _ in range(3)]
    ref_s = [[-32, 31] for _ in range(3)]
    order = []
    for r in range(0, mi_rows, sb4):
        for c in range(0, mi_cols, sb4):
            order.append((r, c))
            for p in range(nplanes):
                t = d.lr_type[p]
                if t == abi.RESTORE_NONE:
                    continue
                sx, sy = (d.subsampling_x, d.subsampling_y) if p else (0, 0)
                u = d.lr_unit_size[p]
                cols, rows = counts[p]
                y0, x0 = (r * 4) >> sy, (c * 4) >> sx
                hh, ww = (sb4 * 4) >> sy, (sb4 * 4) >> sx
                for ur in range(rows):
                    for uc in range(cols):
                        if not (y0 <= ur * u < y0 + hh and x0 <= uc * u < x0 + ww):
                            continue
                        if t == abi.RESTORE_SWITCHABLE:
                            kind = rng.choice((0, 1, 2))
                            enc.symbol(kind, cdf_t)
                        elif t == abi.RESTORE_WIENER:
                            kind = rng.choice((0, 1))
                            enc.symbol(kind, cdf_w)
                        else:
                            kind = rng.choice((0, 2))
                            enc.symbol(kind // 2, cdf_s)
                        if kind == 1:
                            f = [[0, 0, 0], [0, 0, 0]]
                            for pas in range(2):
                                for k, (lo, hi, kk) in enumerate(WIENER_RANGE):
                                    if p and k == 0:
                                        continue
                                    v = rng.randint(lo, hi) if rng.random() < 0.7 else rng.choice((lo, hi))
                                    enc.signed_subexp(v, lo, hi, kk, ref_w[p][pas][k])
                                    f[pas][k] = v
                            ref_w[p] = [list(f[0]), list(f[1])]
                            want[p][(ur, uc)] = ("W", tuple(f[0]), tuple(f[1]))
                        elif kind == 2:
                            s = rng.randrange(16)
                            enc.literal(s, 4)
                            x = [0, 0]
                            for pas, (lo, hi) in enumerate(((-96, 31), (-32, 95))):
                                if sgr_uses(s, pas):
                                    x[pas] = rng.randint(lo, hi)
                                    enc.signed_subexp(x[pas], lo, hi, 3, ref_s[p][pas])
                                elif pas == 1:
                                    x[1] = min(max(128 - x[0], -32), 95)
                            ref_s[p] = list(x)
                            want[p][(ur, uc)] = ("S", s, tuple(x))
                        else:
                            want[p][(ur, uc)] = ("N",)
            if per_sb:
                per_sb(enc, r, c)
    return enc.finish(), want, order


def as_tuple(u):
    if u.kind == abi.UNIT_WIENER:
        return ("W", tuple(u.vertical), tuple(u.horizontal))
    if u.kind == abi.UNIT_SELF_GUIDED:
        return ("S", u.set, tuple(u.xqd))
    return ("N",)


LR_CASES = [
    # w, h, types, unit sizes, ssx, ssy, mono, 128x128 superblocks
    (200, 136, (1, 1, 1), (64, 32, 32), 1, 1, False, False),
    (200, 136, (2, 3, 1), (64, 64, 64), 1, 1, False, False),
    (260, 190, (1, 0, 3), (128, 64, 64), 1, 1, False, True),
    (130, 100, (3, 2, 2), (32, 32, 32), 0, 0, False, False),
    (150, 150, (1, 1, 1), (64, 64, 64), 1, 0, False, False),
    (96, 70, (1, 0, 0), (32, 0, 0), 1, 1, True, False),
    (300, 72, (2, 2, 2), (256, 128, 128), 1, 1, False, True),
]


@pytest.mark.parametrize("case", LR_CASES, ids=lambda c: f"{c[0]}x{c[1]}-{c[2]}-{c[3]}")
def test_restoration_units_round_trip(case):
    w, h, types, units, ssx, ssy, mono, sb128 = case
    d, counts = lr_desc(w, h, types, units, ssx, ssy, mono)
    for seed in range(3):
        rng = random.Random(hash(case) ^ seed)
        data, want, order = encode_units(rng, d, counts, sb128)
        maps = [(abi.LrUnit * (c * r))() if d.lr_type[p] else None for p, (c, r) in enumerate(counts)]
        for p, m in enumerate(maps):
            if m is not None:
                for u in m:
                    u.kind = 99  # every unit must be written
        sd = av1pf.SymbolDecoder(data, False)
        st = av1pf.lr_tile_state()
        for (r, c) in order:
            av1pf.read_lr_units(sd, st, d, r, c, BLOCK_128X128 if sb128 else BLOCK_64X64, maps)
        assert not sd.overrun
        for p in range(1 if mono else 3):
            if not d.lr_type[p]:
                continue
            cols, rows = counts[p]
            assert len(want[p]) == cols * rows, (p, len(want[p]), cols, rows)
            for (ur, uc), wu in want[p].items():
                assert as_tuple(maps[p][ur * cols + uc]) == wu, (case, seed, p, ur, uc)


def test_restoration_tile_state_defaults():
    # default_cdf.rs:4725-4731; decode.rs:1039-1040
    st = av1pf.lr_tile_state()
    assert list(st.use_wiener) == [11570, 32768, 0] and list(st.use_sgrproj) == [16855, 32768, 0]
    assert list(st.restoration_type) == [9413, 22581, 32768, 0]
    assert all(list(st.ref_sgr_xqd[p]) == [-32, 31] for p in range(3))
    assert all(list(st.ref_lr_wiener[p][k]) == [3, -7, 15] for p in range(3) for k in range(2))


def test_restoration_unit_reader_errors():
    d, counts = lr_desc(128, 128, (1, 0, 0), (64, 0, 0))
    sd = av1pf.SymbolDecoder(bytes(16), False)
    st = av1pf.lr_tile_state()
    m = (abi.LrUnit * 4)()
    with pytest.raises(av1pf.Av1Error) as e:
        av1pf.read_lr_units(sd, st, d, 0, 0, 22, [m, None, None])
    assert e.value.message == "AV1 restoration superblock geometry"
    d.lr_unit_size[0] = 0
    with pytest.raises(av1pf.Av1Error) as e:
        av1pf.read_lr_units(sd, st, d, 0, 0, BLOCK_64X64, [m, None, None])
    assert e.value.message == "AV1 active restoration has no unit size"
    d.lr_unit_size[0] = 64
    d.lr_unit_count[0] = 0
    with pytest.raises(av1pf.Av1Error) as e:
        av1pf.read_lr_units(sd, st, d, 0, 0, BLOCK_64X64, [m, None, None])
    assert e.value.message == "AV1 restoration unit map length"
    # no active plane: nothing read, whatever the block size
    d.lr_type[0] = 0
    pos = sd.sd.bits.pos
    av1pf.read_lr_units(sd, st, d, 0, 0, 22, [m, None, None])
    assert sd.sd.bits.pos == pos
