"""The self-guided filter's box arithmetic runs in 32 bits on the GPU
(``ab_of`` in vaco_amd/csrc/av1pf_lr.hip); the reference computes it in i64
(box_ab, restoration.rs:314-339). These checks prove, for every radius /
epsilon pair of SGR_PARAMS (restoration.rs:285-302) and every bit depth, that
the 32-bit form gives the reference's (a2, b2) for every reachable box.

The kernel replaces ``z = round2(p * scale, 20)`` by
``(min(p, thr) * scale + 2^19) >> 20`` with
``thr = ceil((255 * 2^20 - 2^19) / scale)``. z is non-decreasing in p, so the
two agree on a2 for every p >= 0 exactly when z(thr - 1) < 255 <= z(thr) (then
p < thr gives the same z below 255, computed without overflow, and p >= thr
gives a2 = 256 both ways) and thr * scale + 2^19 fits in 32 bits.
"""
import numpy as np
import pytest

# SGR_PARAMS, restoration.rs:285-302: [(r0, e0), (r1, e1)] per set.
SGR_PARAMS = [
    [(2, 12), (1, 4)], [(2, 15), (1, 6)], [(2, 18), (1, 8)], [(2, 21), (1, 9)],
    [(2, 24), (1, 10)], [(2, 29), (1, 11)], [(2, 36), (1, 12)], [(2, 45), (1, 13)],
    [(2, 56), (1, 14)], [(2, 68), (1, 15)], [(0, 0), (1, 5)], [(0, 0), (1, 8)],
    [(0, 0), (1, 11)], [(0, 0), (1, 14)], [(2, 30), (0, 0)], [(2, 75), (0, 0)],
]
PAIRS = sorted({pr for s in SGR_PARAMS for pr in s if pr[0] != 0})
U32 = 1 << 32


def round2(v, s):
    return v if s == 0 else (v + (1 << (s - 1))) >> s


def scale_of(r, eps):
    """restoration.rs:315-317."""
    n = (2 * r + 1) ** 2
    n2e = n * n * eps
    return ((1 << 20) + n2e // 2) // n2e


def thr_of(scale, ceil=True):
    num = (255 << 20) - (1 << 19)
    return -(-num // scale) if ceil else num // scale


def ref_ab(r, eps, s, q, bd):
    """box_ab (restoration.rs:314-339) from the box's sum s and sum of squares q."""
    n = (2 * r + 1) ** 2
    a = round2(q, 2 * (bd - 8))
    d = round2(s, bd - 8)
    p = max(a * n - d * d, 0)
    z = round2(p * scale_of(r, eps), 20)
    a2 = 256 if z >= 255 else 1 if z == 0 else ((z << 8) + z // 2) // (z + 1)
    recip = ((1 << 12) + n // 2) // n
    return a2, round2((256 - a2) * s * recip, 12)


def kernel_ab(r, eps, s, q, bd):
    """ab_of's arithmetic, every intermediate checked against 32 bits."""
    n = (2 * r + 1) ** 2
    scale = scale_of(r, eps)
    thr = thr_of(scale)
    assert q < 1 << 31 and s < 1 << 24
    a = round2(q, 2 * (bd - 8))
    d = round2(s, bd - 8)
    assert a < 1 << 24 and d < 1 << 24 and a * n < 1 << 31 and d * d < 1 << 31  # __umul24, int
    p = max(a * n - d * d, 0)
    zin = min(p, thr) * scale + (1 << 19)
    assert zin < U32
    z = zin >> 20
    a2 = 256 if z >= 255 else 1 if z == 0 else ((z << 8) + z // 2) // (z + 1)
    recip = ((1 << 12) + n // 2) // n
    m = (256 - a2) * recip
    assert m < 1 << 24
    bin_ = m * s + (1 << 11)
    assert bin_ < U32
    b2 = bin_ >> 12
    assert b2 < 1 << 21  # packed as b2 << 9 | a2 in one word
    return a2, b2


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"r{p[0]}e{p[1]}")
def test_clamped_product_gives_the_reference_z_for_every_variance(pair):
    r, eps = pair
    scale = scale_of(r, eps)
    thr = thr_of(scale)
    assert round2((thr - 1) * scale, 20) < 255 <= round2(thr * scale, 20)
    assert thr * scale + (1 << 19) < U32


def test_a_floor_threshold_would_be_wrong_for_some_parameter_set():
    """The check above distinguishes the ceiling from a floor division."""
    bad = [pr for pr in PAIRS if round2(thr_of(scale_of(*pr), ceil=False) * scale_of(*pr), 20) < 255]
    assert bad


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_kernel_box_arithmetic_equals_the_reference_on_extreme_and_random_boxes(bd):
    rng = np.random.default_rng(bd)
    mx = (1 << bd) - 1
    for r, eps in PAIRS:
        n = (2 * r + 1) ** 2
        boxes = [np.zeros(n, np.int64), np.full(n, mx, np.int64), np.full(n, mx // 2, np.int64)]
        for k in range(1, n):  # k samples at the maximum, the rest at 0: the widest variances
            b = np.zeros(n, np.int64)
            b[:k] = mx
            boxes.append(b)
        for _ in range(300):
            lo, hi = sorted(rng.integers(0, mx + 1, size=2))
            boxes.append(rng.integers(lo, hi + 1, size=n).astype(np.int64))
        for b in boxes:
            s, q = int(b.sum()), int((b * b).sum())
            assert kernel_ab(r, eps, s, q, bd) == ref_ab(r, eps, s, q, bd), (r, eps, bd, s, q)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_kernel_box_arithmetic_equals_the_reference_around_the_threshold(bd):
    """Sums of squares whose variance lands next to each threshold."""
    for r, eps in PAIRS:
        n = (2 * r + 1) ** 2
        thr = thr_of(scale_of(r, eps))
        # a box of value v everywhere but one sample v + t: p = (n - 1) * t^2
        # (up to rounding); scan t so that p sweeps across thr
        t0 = int(np.sqrt(thr / (n - 1)))
        for t in range(max(0, t0 - 3), t0 + 4):
            for v in (0, 1 << (bd - 8)):
                b = np.full(n, v, np.int64)
                b[0] = min(v + (t << (bd - 8)), (1 << bd) - 1)
                s, q = int(b.sum()), int((b * b).sum())
                assert kernel_ab(r, eps, s, q, bd) == ref_ab(r, eps, s, q, bd)
