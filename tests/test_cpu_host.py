"""CPU-only checks: the C ABI library loads and exports what include/*.h
declares, the host helpers agree with the oracle, the workload generator is
deterministic, and the oracle's frame glue satisfies the invariants the
reference's own frame tests state (cdef_frame_tests.rs:71-131,
tests/restoration.rs:229-264) plus brute force against the pinned block
functions. No GPU work is started here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import frames as F
import golden_inputs as G
from oracle import oracle as O
from vaco_amd import abi, av1pf, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("vaco_av1pf.h", "vaco_av1pf_gen.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(vaco_av1pf_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(av1pf.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    header = open(os.path.join(ROOT, "include", "vaco_av1pf.h")).read()
    declared = int(re.search(r"#define VACO_AV1PF_ABI_VERSION (\d+)", header).group(1))
    assert abi.ABI_VERSION == declared
    assert av1pf.lib().vaco_av1pf_abi_version() == declared


def test_struct_sizes_match_the_header_layout():
    assert C.sizeof(abi.FilterParams) == 16
    assert C.sizeof(abi.LrUnit) == 18
    assert C.sizeof(abi.Plane) == 24


def test_server_config_and_frame_bytes_host_checks():
    L = av1pf.lib()
    assert C.sizeof(abi.ServerConfig) == 24 and C.sizeof(abi.Completion) == 56
    h = C.c_void_p()
    bad = [abi.ServerConfig(1 << 30, 2000, 0, 2), abi.ServerConfig(1 << 30, 2000, 257, 2),
           abi.ServerConfig(1 << 30, 2000, 64, 0), abi.ServerConfig(0, 2000, 64, 2)]
    for cfg in bad:  # rejected before any device work
        assert L.vaco_av1pf_server_create(C.c_void_p(1), C.byref(cfg), C.byref(h)) == abi.INVALID_ARGUMENT
    good = abi.ServerConfig(1 << 30, 2000, 64, 2)
    assert L.vaco_av1pf_server_create(None, C.byref(good), C.byref(h)) == abi.INVALID_ARGUMENT
    # admission bytes: input + output planes plus intermediates, at least the planes themselves
    f = av1pf.generate(workloads.cfg1(), 0)
    planes = sum(p.size for p in f.planes) + sum(h * w for h, w in f.output_geometry())
    assert 2 * planes <= av1pf.frame_bytes(f) <= 2 * 4 * planes
    g = av1pf.generate(workloads.cfg0(), 0)
    assert av1pf.frame_bytes(g) < av1pf.frame_bytes(f)


def test_adjust_strength_closed_form_buckets():
    # AV1 7.15.1: priStr * (4 + min(FloorLog2(var >> 6), 12)) + 8 >> 4, 0 when var == 0.
    cases = {(15, 0): 0, (15, 1): (15 * 4 + 8) >> 4, (15, 127): (15 * 4 + 8) >> 4, (15, 128): (15 * 5 + 8) >> 4,
             (60, 1 << 18): 60, (60, 0xFFFFFFFF): 60, (7, 64 << 5): (7 * 9 + 8) >> 4}
    for (p, v), want in cases.items():
        assert av1pf.adjust_strength(p, v) == want == O.adjust_strength(p, v)
    for p in range(0, 61):
        for lg in range(0, 27):
            for v in {(64 << lg), (64 << lg) + 1, (128 << lg) - 1}:
                v &= 0xFFFFFFFF
                assert av1pf.adjust_strength(p, v) == O.adjust_strength(p, v)
                assert O.adjust_strength(p, v) <= p


def test_chroma_direction_table_is_the_spec_map():
    # AV1 7.15.1 Cdef_Uv_Dir for 4:2:2 (sub_x only) and 4:4:0 (sub_y only)
    x_only = [7, 0, 2, 4, 5, 6, 6, 6]
    y_only = [1, 2, 2, 2, 3, 4, 6, 0]
    for d in range(8):
        assert O.chroma_direction(d, True, False) == x_only[d] == av1pf.chroma_direction(d, True, False)
        assert O.chroma_direction(d, False, True) == y_only[d] == av1pf.chroma_direction(d, False, True)
        assert O.chroma_direction(d, True, True) == d == av1pf.chroma_direction(d, True, True)
        assert O.chroma_direction(d, False, False) == d


def test_unit_geometry_host_helpers_match_reference_vectors():
    cfg = abi.LrPlaneConfig(95, 161, 64, 8, 0)
    assert av1pf.lr_unit_counts(cfg) == (1, 3)
    assert [av1pf.lr_unit_index(cfg, 94, y) for y in (55, 56, 119, 120)] == [0, 1, 1, 2]
    chroma = abi.LrPlaneConfig(48, 81, 32, 8, 1)
    assert av1pf.lr_unit_counts(chroma) == (2, 3)
    assert av1pf.lr_unit_index(chroma, 32, 27) == 1 and av1pf.lr_unit_index(chroma, 32, 28) == 3
    rng = np.random.default_rng(1)
    for _ in range(300):
        c = abi.LrPlaneConfig(int(rng.integers(1, 900)), int(rng.integers(1, 900)),
                              int(rng.choice([32, 64, 128, 256])), 8, int(rng.integers(0, 2)))
        assert av1pf.lr_unit_counts(c) == O.lr_unit_counts(c)
        for _ in range(5):
            x, y = int(rng.integers(0, c.width)), int(rng.integers(0, c.height))
            assert av1pf.lr_unit_index(c, x, y) == O.lr_unit_index(c, x, y)


def test_skip8_requires_all_four_cells():
    mi = np.ones((4, 6), np.uint8)
    mi[0, 1] = 0
    mi[3, 5] = 0
    s = av1pf.skip8_from_mi(mi)
    assert s.tolist() == [[0, 1, 1], [1, 1, 0]]


def test_generator_is_deterministic_and_in_range():
    a = av1pf.generate(workloads.cfg0(), 3)
    b = av1pf.generate(workloads.cfg0(), 3)
    c = av1pf.generate(workloads.cfg0(), 4)
    for p in range(3):
        assert np.array_equal(a.planes[p], b.planes[p])
        assert a.planes[p].max() <= 255
    assert not np.array_equal(a.planes[0], c.planes[0])
    assert [a.desc.cdef.y[i].primary for i in range(8)] == [b.desc.cdef.y[i].primary for i in range(8)]
    assert 3 <= a.desc.cdef.damping <= 6
    f2 = av1pf.generate(workloads.cfg2(), 0)
    d = f2.desc
    assert d.use_superres == 1 and d.bit_depth == 10 and d.upscaled_width == 3840
    denom = None
    for k in range(9, 17):
        if (3840 * 8 + k // 2) // k == d.coded_width:
            denom = k
    assert denom is not None
    assert f2.planes[0].shape == (2160, 2 * ((d.coded_width + 7) >> 3) * 4)
    assert f2.planes[0].max() <= 1023


def test_generated_noise_has_the_configured_sigma():
    cfg = workloads.make(256, 256, 8, sigma=6, edges=0, lr=False, seed=7)
    f = av1pf.generate(cfg, 0)
    y = f.planes[0].astype(np.float64)
    resid = y[:, 1:] - y[:, :-1]  # gradient is nearly flat between neighbours
    sd = resid.std() / np.sqrt(2)
    assert 5.0 < sd < 7.0


# ---------------------------------------------------- oracle frame glue ---

def test_zero_strengths_and_disabled_cdef_are_identity():
    for kwargs in ({"zero_strengths": True}, {"cdef": False}):
        f = F.random_frame(11, 80, 72, lr=False, **kwargs)
        out = O.frame(f.desc, f.planes)
        for p in range(3):
            assert np.array_equal(out[p], f.planes[p])


def test_negative_index_and_skipped_blocks_are_untouched():
    f = F.random_frame(12, 136, 72, lr=False, idx_neg_p=0.5, skip_p=0.5)
    out = O.frame(f.desc, f.planes)
    d = f.desc
    mi_cols = 2 * ((d.coded_width + 7) >> 3)
    idx = f._keep[0]
    skip = f._keep[1].reshape(-1, mi_cols // 2)
    for by in range(skip.shape[0]):
        for bx in range(skip.shape[1]):
            i = idx[(by * 2 >> 4) * d.cdef_stride + (bx * 2 >> 4)]
            if i < 0 or skip[by, bx]:
                assert np.array_equal(out[0][by * 8:by * 8 + 8, bx * 8:bx * 8 + 8],
                                      f.planes[0][by * 8:by * 8 + 8, bx * 8:bx * 8 + 8])
                assert np.array_equal(out[1][by * 4:by * 4 + 4, bx * 4:bx * 4 + 4],
                                      f.planes[1][by * 4:by * 4 + 4, bx * 4:bx * 4 + 4])


@pytest.mark.parametrize("ss", [(1, 1), (0, 0), (1, 0), (0, 1)])
def test_frame_cdef_equals_block_functions_on_gathered_neighbourhoods(ss):
    f = F.random_frame(20 + ss[0] * 2 + ss[1], 72, 56, lr=False, ssx=ss[0], ssy=ss[1])
    out = O.frame(f.desc, f.planes)
    want = F.brute_cdef(f)
    for p in range(3):
        assert np.array_equal(out[p], want[p]), p


def test_superres_of_constant_plane_is_constant():
    for bd in (8, 10, 12):
        v = (1 << bd) - 3
        plane = np.full((9, 40), v, np.uint16)
        out = O.upscale_plane(plane, 37, 70, 9, bd)
        assert (out == v).all()


def test_restoration_none_copies_cdef_output_visible_region():
    f = F.random_frame(30, 100, 70, lr=True, unit_kinds=(0,))
    out = O.frame(f.desc, f.planes)
    nolr = F.random_frame(30, 100, 70, lr=False, unit_kinds=(0,))
    ref = O.frame(nolr.desc, nolr.planes)
    assert np.array_equal(out[0], ref[0][:70, :100])
    assert np.array_equal(out[1], ref[1][:35, :50])


def test_fuel_is_charged_per_block_and_per_superres_row():
    f = F.random_frame(31, 64, 32, lr=False, skip_p=0.0, idx_neg_p=0.0, upscaled_w=80)
    # 8 x 4 luma blocks = 32 CDEF fuel; superres 80*32 + 2 * 40*16 per plane rows
    need = 32 + 80 * 32 + 2 * 40 * 16
    O.frame(f.desc, f.planes, fuel=need)
    with pytest.raises(O.OracleError) as e:
        O.frame(f.desc, f.planes, fuel=need - 1)
    assert e.value.status == abi.LIMIT_EXCEEDED and e.value.message == "fuel"


def test_out_of_range_sample_is_reported_by_the_first_stage_reading_it():
    f = F.random_frame(32, 64, 32, lr=False, skip_p=0.0, idx_neg_p=0.0)
    f.planes[0][3, 3] = 300
    with pytest.raises(O.OracleError) as e:
        O.frame(f.desc, f.planes)
    assert e.value.message == "CDEF sample exceeds bit depth"


def test_pic_to_frame_oracle_crops_and_packs_like_blit():
    # decode.rs:3795-3891 on hand-checkable planes: a padded 8-bit picture is
    # cropped to coded size, an out-of-range 8-bit sample becomes 0
    # (u8::try_from(..).unwrap_or(0)), 10-bit samples are two little-endian
    # bytes, and a frame plane larger than the picture plane keeps its extra
    # row untouched.
    d = abi.FrameDesc()
    d.coded_width, d.upscaled_width, d.coded_height = 5, 5, 3
    d.bit_depth, d.subsampling_x, d.subsampling_y = 8, 1, 1
    y = np.arange(8 * 4, dtype=np.uint16).reshape(4, 8) * 9  # padded 8x4
    y[1, 2] = 300
    u = np.array([[1, 2, 3, 4]], dtype=np.uint16)                # 1 row: frame chroma has 2
    v = np.array([[5, 6, 7, 8]], dtype=np.uint16)
    fr = O.pic_to_frame(d, [y, u, v], [(3, 5), (2, 3), (2, 3)], 1)
    want_y = (np.arange(8 * 4).reshape(4, 8) * 9)[:3, :5].astype(np.uint8)
    want_y[1, 2] = 0
    assert np.array_equal(fr[0], want_y)
    assert fr[1].tolist() == [[1, 2, 3], [0, 0, 0]] and fr[2].tolist() == [[5, 6, 7], [0, 0, 0]]
    d.bit_depth = 10
    y10 = np.full((4, 8), 0x3A5, np.uint16)
    fr = O.pic_to_frame(d, [y10, u, v], [(3, 5), (2, 3), (2, 3)], 2)
    assert fr[0].dtype == np.uint16 and (fr[0] == 0x3A5).all()
    assert fr[0].view(np.uint8)[0, :2].tolist() == [0xA5, 0x03]


def test_frame_layout_matches_pic_to_frame_sizes():
    f = F.random_frame(seed=3, coded_w=93, coded_h=75, upscaled_w=161)
    geo, bps = av1pf.frame_layout(f)
    assert geo == [(75, 161), (38, 81), (38, 81)] and bps == 1
    g = F.random_frame(seed=4, coded_w=77, coded_h=45, mono=True, bd=10)
    geo, bps = av1pf.frame_layout(g)
    assert geo == [(45, 77)] and bps == 2
