"""CDEF keeps each pair's clamp range (cdef.rs:200-202: the smallest and the
largest tap) with packed f16 maximum3 / minimum3 instructions on the samples'
u16 bit patterns (``max3_u16`` / ``min3_u16`` in
vaco_amd/csrc/av1pf_cdef_body.h). That is exact when f16 orders every sample
value 0 .. 4095 (12-bit at most) as the integers: those bit patterns are +0,
subnormals and small normals. These checks state it for every value and for
the three-way forms; the GPU parity tests pin the instructions themselves.
"""
import numpy as np

MAXV = 4095


def test_sample_bit_patterns_order_as_f16_like_the_integers():
    v = np.arange(MAXV + 1, dtype=np.uint16)
    h = v.view(np.float16)
    assert np.all(np.isfinite(h)) and not np.any(np.signbit(h))
    assert np.all(np.diff(h.astype(np.float64)) > 0)  # strictly increasing, subnormals included
    assert h[1] > 0 and h[0x3FF] < h[0x400]            # smallest and largest subnormal


def test_three_way_f16_max_and_min_equal_the_u16_ones():
    rng = np.random.default_rng(7)
    a, b, c = (rng.integers(0, MAXV + 1, 200_000, dtype=np.uint16) for _ in range(3))
    ha, hb, hc = (x.view(np.float16) for x in (a, b, c))
    mx = np.maximum(np.maximum(ha, hb), hc).view(np.uint16)
    mn = np.minimum(np.minimum(ha, hb), hc).view(np.uint16)
    assert np.array_equal(mx, np.maximum(np.maximum(a, b), c))
    assert np.array_equal(mn, np.minimum(np.minimum(a, b), c))


def test_values_past_the_sample_range_break_the_order():
    # Why the kernels flag any sample above the bit depth: from 0x7C00 the
    # patterns are infinities and NaNs, from 0x8000 negative.
    h = np.array([0x7BFF, 0x7C00, 0x7E00, 0x8001], dtype=np.uint16).view(np.float16)
    assert np.isfinite(h[0]) and np.isinf(h[1]) and np.isnan(h[2]) and h[3] < 0
