"""The header-syntax producers of the stage's parameters (host code, no GPU),
pinned to the reference's own test vectors
(crates/codec/vaco-codec-av1/src/frame_header.rs:1235-1372) and to the
closed forms of AV1 5.9.8 / 5.9.19 / 5.9.20.
"""
import random

from vaco_amd import abi, av1pf


def test_restoration_modes_are_remapped_and_chroma_units_subsampled():
    # frame_header.rs:1235-1262: Y=Wiener, U=SGR, V=switchable; unit shift 2, UV shift 1.
    seq = av1pf.seq_info(restoration=True, sb128=False, ssx=True, ssy=True)
    bits = av1pf.Bits(bytes([0b1011_0111, 0b1000_0000]))
    types, units = av1pf.parse_lr_params(bits, seq, num_planes=3)
    assert types == [abi.RESTORE_WIENER, abi.RESTORE_SELF_GUIDED, abi.RESTORE_SWITCHABLE]
    assert units == [256, 128, 128]
    assert bits.pos == 9 and not bits.overrun
    # the 128x128-superblock branch reads one shift bit
    seq = av1pf.seq_info(restoration=True, sb128=True)
    bits = av1pf.Bits(bytes([0b1000_0000]))
    types, units = av1pf.parse_lr_params(bits, seq, num_planes=1)
    assert units == [128, 128, 128] and bits.pos == 3


def test_disabled_restoration_reads_nothing_and_none_modes_read_no_size():
    # frame_header.rs:1265-1284
    for lossless, intrabc, enabled in ((True, False, True), (False, True, True), (False, False, False)):
        seq = av1pf.seq_info(restoration=enabled)
        bits = av1pf.Bits(bytes([255]))
        types, units = av1pf.parse_lr_params(bits, seq, all_lossless=lossless, allow_intrabc=intrabc)
        assert types == [0, 0, 0] and units == [0, 0, 0] and bits.pos == 0
    bits = av1pf.Bits(bytes([0]))
    types, units = av1pf.parse_lr_params(bits, av1pf.seq_info(restoration=True))
    assert types == [0, 0, 0] and bits.pos == 6


def test_restoration_unit_size_closed_form():
    # lr_unit_shift (+ lr_unit_extra_shift) and lr_uv_shift, every combination
    for sb128 in (False, True):
        for shift_bits in ((0,), (1, 0), (1, 1)) if not sb128 else ((0,), (1,)):
            for uv in (0, 1):
                # Y, U, V switchable, then the shift bits and the UV shift
                stream = [0, 1, 0, 1, 0, 1] + list(shift_bits) + [uv]
                data = int("".join(map(str, stream)).ljust(16, "0"), 2).to_bytes(2, "big")
                bits = av1pf.Bits(data)
                types, units = av1pf.parse_lr_params(bits, av1pf.seq_info(restoration=True, sb128=sb128))
                shift = (shift_bits[0] + 1) if sb128 else (shift_bits[0] + (shift_bits[1] if shift_bits[0] else 0))
                assert units == [64 << shift, (64 << shift) >> uv, (64 << shift) >> uv], (sb128, shift_bits, uv)
                assert bits.pos == 6 + len(shift_bits) + 1
    # 4:4:4 and 4:2:2 never read the UV shift
    for ssx, ssy in ((False, False), (True, False)):
        bits = av1pf.Bits(bytes([0b0101_0100, 0]))
        _, units = av1pf.parse_lr_params(bits, av1pf.seq_info(restoration=True, ssx=ssx, ssy=ssy))
        assert units == [64, 64, 64] and bits.pos == 7


def test_cdef_single_entry_keeps_strengths_and_damping():
    # frame_header.rs:1327-1352: damping=2(+3), bits=0, y=(9,3->4), uv=(6,2)
    bits = av1pf.Bits(bytes([0b1000_1001, 0b1101_1010]))
    n, p = av1pf.parse_cdef_params(bits, av1pf.seq_info(cdef=True), num_planes=3)
    assert n == 0 and p.enabled == 1 and p.damping == 5
    assert (p.y[0].primary, p.y[0].secondary) == (9, 4)
    assert (p.uv[0].primary, p.uv[0].secondary) == (6, 2)
    assert bits.pos == 16 and not bits.overrun


def test_cdef_disabled_syntax_reads_nothing():
    # frame_header.rs:1355-1371
    for enable, lossless, intrabc in ((False, False, False), (True, True, False), (True, False, True)):
        bits = av1pf.Bits(b"")
        n, p = av1pf.parse_cdef_params(bits, av1pf.seq_info(cdef=enable), coded_lossless=lossless,
                                       allow_intrabc=intrabc)
        assert n == 0 and p.enabled == 0 and bits.pos == 0 and not bits.overrun


def test_cdef_table_size_and_monochrome():
    rng = random.Random(7)
    for cdef_bits in range(4):
        for planes in (1, 3):
            entries = [(rng.randrange(16), rng.randrange(4), rng.randrange(16), rng.randrange(4))
                       for _ in range(1 << cdef_bits)]
            s = f"{1:02b}{cdef_bits:02b}"
            for yp, ys, up, us in entries:
                s += f"{yp:04b}{ys:02b}" + (f"{up:04b}{us:02b}" if planes > 1 else "")
            data = int(s.ljust((len(s) + 7) // 8 * 8, "0"), 2).to_bytes((len(s) + 7) // 8, "big")
            bits = av1pf.Bits(data)
            n, p = av1pf.parse_cdef_params(bits, av1pf.seq_info(cdef=True), num_planes=planes)
            assert n == cdef_bits and p.damping == 4 and bits.pos == len(s)
            for i, (yp, ys, up, us) in enumerate(entries):
                assert (p.y[i].primary, p.y[i].secondary) == (yp, 4 if ys == 3 else ys)
                want_uv = (up, 4 if us == 3 else us) if planes > 1 else (0, 0)
                assert (p.uv[i].primary, p.uv[i].secondary) == want_uv
            for i in range(1 << cdef_bits, 8):
                assert (p.y[i].primary, p.y[i].secondary) == (0, 0)


def test_superres_coded_width_closed_form():
    # frame_header.rs:687-704: coded = (upscaled * 8 + d / 2) / d, d = 9 + coded_denom
    for d in range(9, 17):
        bits = av1pf.Bits(bytes([0x80 | ((d - 9) << 4)]))
        w, use = av1pf.parse_superres_params(bits, av1pf.seq_info(superres=True), 3840)
        assert use and w == (30720 + d // 2) // d and bits.pos == 4
    assert av1pf.parse_superres_params(av1pf.Bits(bytes([0x80])), av1pf.seq_info(superres=True), 3840)[0] == 3413
    bits = av1pf.Bits(bytes([0x7f]))
    assert av1pf.parse_superres_params(bits, av1pf.seq_info(superres=True), 1920) == (1920, False)
    assert bits.pos == 1
    bits = av1pf.Bits(bytes([0xff]))
    assert av1pf.parse_superres_params(bits, av1pf.seq_info(superres=False), 1920) == (1920, False)
    assert bits.pos == 0


def test_reads_past_the_end_return_zero_and_flag_overrun():
    rng = random.Random(11)
    for n in range(40):
        data = bytes(rng.randrange(256) for _ in range(rng.randrange(3)))
        bits = av1pf.Bits(data)
        av1pf.parse_lr_params(bits, av1pf.seq_info(restoration=True))
        av1pf.parse_cdef_params(bits, av1pf.seq_info(cdef=True))
        assert bits.overrun == (bits.pos > 8 * len(data))
    bits = av1pf.Bits(bytes([0xff]))
    assert bits.get(12) == 0xff0 and bits.overrun and bits.pos == 12
