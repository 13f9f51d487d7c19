"""vaco_av1pf_frame_device_static, the entry point that allocates nothing,
against vaco_av1pf_frame_device (bit-exact) through the C ABI: two calls in a
row, the second with a smaller frame; an error replayed with the reference's
message; and the size limit refused before any device work.
"""
import sys

import numpy as np
import pytest

import frames as F
from oracle import oracle as O
from vaco_amd import abi, av1pf, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_static(f):
    from vaco_amd import hip
    dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
    dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
    av1pf.frame_device_static(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out])
    return [t.download() for t in dev_out]


def run_device(f):
    from vaco_amd import hip
    dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
    dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
    with av1pf.Context(3840, 2176) as ctx:
        ctx.frame_device(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out])
        ctx.finish()
        return [t.download() for t in dev_out]


def same(a, b, label):
    assert len(a) == len(b), label
    for p, (x, y) in enumerate(zip(a, b)):
        assert x.shape == y.shape and np.array_equal(x, y), f"{label} plane {p}"


def test_static_entry_equals_the_device_entry_on_two_calls_the_second_smaller():
    big = av1pf.generate(workloads.make(3840, 2160, 10, sigma=24, unit=64, seed=77), 0)
    small = av1pf.generate(workloads.cfg0(), 1)
    say("static entry: 4K 10-bit frame, then 352x288 8-bit")
    got_big = run_static(big)
    got_small = run_static(small)
    same(got_big, run_device(big), "4K")
    same(got_small, run_device(small), "352x288")
    same(got_small, O.frame(small.desc, small.planes), "352x288 vs oracle")


def test_static_entry_on_superres_and_odd_geometry_frames():
    for case in (dict(seed=12, coded_w=93, coded_h=75, upscaled_w=160),
                 dict(seed=26, coded_w=120, coded_h=64, bd=12, ssx=0, ssy=0, upscaled_w=200),
                 dict(seed=16, coded_w=80, coded_h=72, cdef=False, lr=False)):
        say(f"static entry: {case}")
        f = F.random_frame(**case)
        same(run_static(f), run_device(f), str(case))


def test_static_entry_reports_the_reference_error_for_an_out_of_range_sample():
    f = F.random_frame(50, 64, 64, skip_p=0.0, idx_neg_p=0.0)
    f.planes[0][9, 9] = 999
    with pytest.raises(O.OracleError) as want:
        O.frame(f.desc, f.planes)
    with pytest.raises(av1pf.Av1Error) as got:
        run_static(f)
    assert (got.value.status, got.value.message) == (want.value.status, want.value.message)


def test_static_entry_runs_a_4k_superres_frame():
    """4K 10-bit with superres (denominator 9) and restoration: restoration
    upscales inside its staging, so the only intermediate is the CDEF image
    (~22 MB), which fits the static scratch."""
    f = av1pf.generate(workloads.cfg2(), 1)
    say("static: 4K superres frame")
    got = run_static(f)
    want = run_device(f)
    for p, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), p


def test_static_entry_refuses_a_frame_larger_than_its_scratch():
    from vaco_amd import hip
    # 4K 4:4:4 with CDEF and restoration: a ~50 MB CDEF image
    f = F.random_frame(501, 3840, 2160, ssx=0, ssy=0, unit=256)
    dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
    dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
    with pytest.raises(av1pf.Av1Error) as e:
        av1pf.frame_device_static(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out])
    assert e.value.status == abi.LIMIT_EXCEEDED and "static scratch" in e.value.message
