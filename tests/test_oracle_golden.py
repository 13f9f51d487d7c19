"""The oracle against every golden vector the reference holds for this path.

Mirrors the reference's own tests:
  crates/tool/vaco-checkasm/tests/av1_cdef.rs:106-124
  crates/codec/vaco-codec-av1/tests/superres.rs:37-84
  crates/codec/vaco-codec-av1/tests/restoration.rs:22-264
"""
import numpy as np
import pytest

import golden_inputs as G
from oracle import oracle as O
from vaco_amd import abi


@pytest.fixture(scope="module")
def cdef():
    return G.cdef_cases()


def test_cdef_every_direction_strength_and_damping_matches_dav1d(cdef):
    inputs, params, exp_dir, exp_var, exp_out = cdef
    bad = []
    for i in range(len(params)):
        d, v = O.find_direction(G.luma_of(inputs[i]), params[i].bit_depth)
        out = O.filter_block(inputs[i], params[i]).reshape(64)
        if d != exp_dir[i] or v != exp_var[i] or not np.array_equal(out, exp_out[i]):
            bad.append(i)
    assert not bad, f"{len(bad)} mismatching cases, first {bad[:10]}"


def test_cdef_corpus_exercises_all_directions_and_edge_masks(cdef):
    _, params, exp_dir, _, _ = cdef
    assert len(params) == 7680
    assert set(int(d) for d in exp_dir) == set(range(8))
    assert set(p.edges for p in params) == set(range(16))


def test_superres_matches_every_pinned_dav1d_case():
    for case_id, (bd, vis, padded, out_w, h, inp, exp) in enumerate(G.superres_cases()):
        got = O.upscale_plane(inp, vis, out_w, h, bd)
        assert np.array_equal(got, exp), f"superres case {case_id}"


def test_restoration_matches_all_pixels_across_depths_units_and_stripes():
    pixels = 0
    for (bd, w, h, ssy, unit, mode, before, after, units, cfg, exp) in G.restoration_cases():
        got = O.restore_plane(before, after, cfg, units)
        assert np.array_equal(got, exp), f"depth={bd} {w}x{h} mode={mode}"
        pixels += w * h
    assert pixels == 2_089_260


def test_constant_plane_rounding_matches_dav1d_at_each_bit_depth():
    for bd, s, plane, exp in G.constant_cases():
        cfg = abi.LrPlaneConfig(7, 9, 64, bd, 0)
        wiener = abi.make_unit(abi.UNIT_WIENER, vertical=(-5, -23, -17), horizontal=(10, 8, 46))
        assert np.array_equal(O.restore_plane(plane, plane, cfg, [wiener]), plane)
        sgr = abi.make_unit(abi.UNIT_SELF_GUIDED, set=s, xqd=(-32, 31))
        assert np.array_equal(O.restore_plane(plane, plane, cfg, [sgr]), exp), f"bd={bd} set={s}"


def test_unit_geometry_merges_short_tails_and_shifts_vertical_boundaries():
    cfg = abi.LrPlaneConfig(95, 161, 64, 8, 0)
    assert O.lr_unit_counts(cfg) == (1, 3)
    assert O.lr_unit_index(cfg, 94, 55) == 0
    assert O.lr_unit_index(cfg, 94, 56) == 1
    assert O.lr_unit_index(cfg, 94, 119) == 1
    assert O.lr_unit_index(cfg, 94, 120) == 2
    chroma = abi.LrPlaneConfig(48, 81, 32, 8, 1)
    assert O.lr_unit_counts(chroma) == (2, 3)
    assert O.lr_unit_index(chroma, 32, 27) == 1
    assert O.lr_unit_index(chroma, 32, 28) == 3


def test_invalid_configuration_and_coefficients_are_rejected():
    plane = np.full((4, 4), 100, np.uint16)
    none = [abi.make_unit()]
    for bad in [abi.LrPlaneConfig(4, 4, 0, 8, 0), abi.LrPlaneConfig(4, 4, 64, 16, 0),
                abi.LrPlaneConfig(5, 4, 64, 8, 0), abi.LrPlaneConfig(4, 0, 64, 8, 0)]:
        with pytest.raises(O.OracleError) as e:
            O.restore_plane(plane, plane, bad, none)
        assert e.value.message == "AV1 restoration plane geometry or bit depth"
    cfg = abi.LrPlaneConfig(4, 4, 64, 8, 0)
    for unit in [abi.make_unit(abi.UNIT_SELF_GUIDED, set=16), abi.make_unit(abi.UNIT_SELF_GUIDED, xqd=(32, 0)),
                 abi.make_unit(abi.UNIT_WIENER, vertical=(11, 0, 0))]:
        with pytest.raises(O.OracleError) as e:
            O.restore_plane(plane, plane, cfg, [unit])
        assert e.value.message == "AV1 restoration coefficient out of range"
    with pytest.raises(O.OracleError) as e:
        O.restore_plane(plane, plane, cfg, [])
    assert e.value.message == "AV1 restoration unit count"


def test_none_preserves_visible_cdef_pixels_and_out_of_depth_input_is_rejected():
    before = np.full((12, 12), 16, np.uint16)
    after = np.full((12, 12), 235, np.uint16)
    cfg = abi.LrPlaneConfig(7, 9, 64, 8, 0)
    out = O.restore_plane(before, after, cfg, [abi.make_unit()])
    assert out.shape == (9, 7) and (out == 235).all()
    after[8, 6] = 256
    with pytest.raises(O.OracleError) as e:
        O.restore_plane(before, after, cfg, [abi.make_unit()])
    assert e.value.message == "AV1 restoration sample exceeds bit depth"
