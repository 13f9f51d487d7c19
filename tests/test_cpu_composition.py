"""Pins the oracle's superres -> loop restoration composition, which no
reference vector covers: the reference decoder refuses restoration after
superres (decode.rs:574-577), so no reference output exists for it.

The composition is rebuilt here only from pieces that golden vectors pin:

* CDEF: ``frames.brute_cdef``, the frame walk spelled out over the block
  functions (cdef.rs), which the dav1d CDEF corpus pins;
* superres: ``oracle.upscale_plane`` per plane (dav1d superres corpus),
  with upscale_picture's plane geometry (superres.rs:226-274): luma
  coded_width -> upscaled_width over coded_height rows; chroma widths and
  heights by div_ceil; the input is the Mi-padded plane, whose width bounds
  the edge clamp;
* restoration: ``oracle.restore_plane`` per plane (dav1d restoration
  corpus), reading the upscaled pre-CDEF image at stripe borders and the
  upscaled CDEF image elsewhere (AV1 7.17), with
  restoration_plane_config's geometry (decode.rs:587-610): width
  upscaled_width >> ss_x, height coded_height >> ss_y (floor); a plane of
  type None copies the visible region (decode.rs:673-681).

Frames have odd coded and upscaled sizes, so div_ceil and floor differ, in
every chroma layout and at 8, 10 and 12 bits. Plausible mistakes in the
composition (floor chroma widths for superres, the upscaled CDEF image in
place of the pre-CDEF one at stripe borders) are shown to change the result,
so this check would catch them.
"""
import numpy as np
import pytest

import frames as F
from oracle import oracle as O
from vaco_amd import abi


def _sr_geometry(d, plane, chroma_round):
    """(visible, output, height) of upscale_picture for `plane`;
    chroma_round(v, shift) rounds chroma sizes (div_ceil in the reference)."""
    if plane == 0:
        return d.coded_width, d.upscaled_width, d.coded_height
    sx, sy = d.subsampling_x, d.subsampling_y
    return (chroma_round(d.coded_width, sx), chroma_round(d.upscaled_width, sx),
            chroma_round(d.coded_height, sy))


def div_ceil(v, s):
    return (v + (1 << s) - 1) >> s


def floor_shift(v, s):
    return v >> s


def composed(f, chroma_round=div_ceil, before_is_cdef=False):
    """CDEF -> superres of both images -> restoration, from pinned pieces."""
    d = f.desc
    cdef = F.brute_cdef(f)
    out = []
    for p in range(len(f.planes)):
        vis, ow, h = _sr_geometry(d, p, chroma_round)
        up_cdef = O.upscale_plane(cdef[p], vis, ow, h, d.bit_depth)
        up_pre = up_cdef if before_is_cdef else O.upscale_plane(f.planes[p], vis, ow, h, d.bit_depth)
        sx = d.subsampling_x if p else 0
        sy = d.subsampling_y if p else 0
        cfg = abi.LrPlaneConfig(d.upscaled_width >> sx, d.coded_height >> sy, d.lr_unit_size[p], d.bit_depth,
                                1 if (p and sy) else 0)
        if d.lr_type[p] == abi.RESTORE_NONE:
            out.append(up_cdef[:cfg.height, :cfg.width].copy())
            continue
        units = [d.lr_units[p][i] for i in range(d.lr_unit_count[p])]
        out.append(O.restore_plane(up_pre, up_cdef, cfg, units))
    return out


CASES = [
    # (seed, coded_w, coded_h, upscaled_w, bd, ssx, ssy, mono, unit)
    (101, 93, 75, 161, 8, 1, 1, False, 64),
    (102, 77, 69, 131, 10, 1, 1, False, 32),
    (103, 61, 47, 113, 12, 1, 1, False, 64),
    (104, 69, 45, 125, 8, 0, 0, False, 32),
    (105, 85, 39, 147, 12, 0, 0, False, 64),
    (106, 59, 71, 101, 10, 1, 0, False, 32),
    (107, 83, 53, 149, 8, 1, 0, False, 64),
    (108, 73, 67, 127, 12, 0, 1, False, 32),
    (109, 91, 41, 155, 10, 0, 1, False, 64),
    (110, 71, 57, 123, 8, 0, 0, True, 32),
    (111, 97, 131, 171, 10, 1, 1, False, 64),
]


def _frame(case, lr_types=(1, 1, 1)):
    seed, cw, ch, uw, bd, ssx, ssy, mono, unit = case
    return F.random_frame(seed, cw, ch, bd=bd, ssx=ssx, ssy=ssy, mono=mono, upscaled_w=uw, unit=unit,
                          lr_types=lr_types)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"s{c[0]}-{c[1]}x{c[2]}to{c[3]}-bd{c[4]}-ss{c[5]}{c[6]}"
                         + ("-mono" if c[7] else ""))
def test_oracle_superres_then_restoration_equals_pinned_composition(case):
    f = _frame(case)
    want = composed(f)
    got = O.frame(f.desc, f.planes)
    assert len(got) == len(want)
    for p, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, (p, g.shape, w.shape)
        assert np.array_equal(g, w), f"plane {p}: {int((g != w).sum())} samples differ"


def test_composition_with_per_plane_types_and_a_plane_without_restoration():
    f = _frame(CASES[0], lr_types=(0, 1, 0))
    for g, w in zip(O.frame(f.desc, f.planes), composed(f)):
        assert np.array_equal(g, w)


def test_the_composition_check_detects_floor_chroma_widths():
    """At odd widths a superres chroma plane sized by floor instead of
    div_ceil (superres.rs:248-250) uses another step and start phase; the
    check above must see the difference."""
    f = _frame(CASES[0])
    assert f.desc.coded_width % 2 == 1
    got = O.frame(f.desc, f.planes)
    wrong = composed(f, chroma_round=floor_shift)
    assert np.array_equal(got[0], wrong[0])  # luma is unaffected
    assert not np.array_equal(got[1], wrong[1])


def test_the_composition_check_detects_cdef_rows_at_stripe_borders():
    """Restoration reads rows across a stripe border from the pre-CDEF image
    (restoration.rs:236-241); taking them from the CDEF image instead must
    change the result."""
    f = _frame(CASES[10])
    got = O.frame(f.desc, f.planes)
    wrong = composed(f, before_is_cdef=True)
    assert any(not np.array_equal(g, w) for g, w in zip(got, wrong))
