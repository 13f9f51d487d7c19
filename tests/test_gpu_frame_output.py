"""The fused frame output on the GPU (vaco_av1pf_frame_device_to_frame):
the stage written straight into pic_to_frame's frame planes equals the
oracle's final picture passed through the oracle's pic_to_frame, byte for
byte, whichever stage runs last (restoration, superres, CDEF or none).
"""
import sys

import numpy as np
import pytest

import frames as F
from oracle import oracle as O
from vaco_amd import av1pf, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    dict(seed=401, coded_w=64, coded_h=64),                                    # LR last, 8-bit
    dict(seed=402, coded_w=131, coded_h=97, bd=10),                            # LR last, 10-bit, odd sizes
    dict(seed=403, coded_w=93, coded_h=75, upscaled_w=160),                    # superres then LR
    dict(seed=404, coded_w=93, coded_h=75, upscaled_w=161, lr=False),          # superres last, odd width
    dict(seed=405, coded_w=77, coded_h=45, lr=False),                          # CDEF last, cropped
    dict(seed=406, coded_w=77, coded_h=45, lr=False, bd=10),
    dict(seed=407, coded_w=70, coded_h=50, cdef=False, lr=False),              # nothing runs
    dict(seed=408, coded_w=70, coded_h=50, cdef=False, lr=False, bd=12),
    dict(seed=409, coded_w=77, coded_h=45, mono=True),
    dict(seed=410, coded_w=72, coded_h=80, ssx=0, ssy=0, lr=False),
    dict(seed=411, coded_w=72, coded_h=80, ssx=1, ssy=0),
]


@pytest.fixture(scope="module")
def ctx():
    c = av1pf.Context(1024, 576)
    yield c
    c.close()


@pytest.mark.parametrize("case", CASES, ids=lambda c: "s%d" % c["seed"])
def test_gpu_frame_output_matches_pic_to_frame(ctx, case):
    from vaco_amd import hip
    say("frame output", case)
    f = F.random_frame(**case)
    geo, bps = av1pf.frame_layout(f)
    want = O.pic_to_frame(f.desc, O.frame(f.desc, f.planes), geo, bps)
    dt = np.uint8 if bps == 1 else np.uint16
    dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
    dev_out = [hip.DeviceArray.from_host(np.zeros(g, dt)) for g in geo]
    ctx.frame_device_to_frame(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out],
                              [g[1] * bps for g in geo])
    ctx.finish()
    for p, w in enumerate(want):
        got = dev_out[p].download()
        assert np.array_equal(got, w), (case, p, int((got != w).sum()))


@pytest.mark.parametrize("workload,frames", [("cfg0", 2), ("cfg1", 1)],
                         ids=["352x288-8bit", "1920x1080-8bit"])
def test_gpu_frame_output_of_the_8bit_workloads_is_the_u8_picture(workload, frames):
    from vaco_amd import hip
    say(f"frame output: {workload} frames")
    cfg = getattr(workloads, workload)()
    with av1pf.Context(int(cfg.width), (int(cfg.height) + 63) // 64 * 64) as ctx:
        check_u8_frames(ctx, workload, frames, hip)


def check_u8_frames(ctx, workload, frames, hip):
    for k in range(frames):
        f = av1pf.generate(getattr(workloads, workload)(), k)
        geo, bps = av1pf.frame_layout(f)
        assert bps == 1
        want = [p.astype(np.uint8) for p in O.frame(f.desc, f.planes)]
        dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
        dev_out = [hip.DeviceArray(g, np.uint8) for g in geo]
        ctx.frame_device_to_frame(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out], [g[1] for g in geo])
        ctx.finish()
        for p, w in enumerate(want):
            assert np.array_equal(dev_out[p].download(), w), (k, p)
