"""GPU checks of the batched workload at its real shapes: 64 frames of
mixed geometry (640x360 / 1280x720 / 1920x1080 / 3840x2160), bit depth
(8 / 10) and superres (half the frames, denominator 9..16) in one
vaco_av1pf_batch_device call. Every frame must equal the per-frame path's
output, and one frame of each shape x depth x superres class must equal the
oracle. Bit-exact (tolerance 0).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from vaco_amd import av1pf, workloads

pytestmark = pytest.mark.gpu

MIXED = workloads.mixed_batch(64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _class(cfg):
    return (int(cfg.width), int(cfg.height), int(cfg.bit_depth), int(cfg.superres_denom) != 8)


CLASSES = {}
for _i, (_c, _k) in enumerate(MIXED):
    CLASSES.setdefault(_class(_c), _i)


@pytest.fixture(scope="module")
def mixed_run():
    from vaco_amd import hip
    say("mixed batch: generating 64 frames")
    frames = [av1pf.generate(c, i) for c, i in MIXED]
    ctx = av1pf.Context(3840, 2176)
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    out_b = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    out_pf = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    ptr = lambda arrs: [[t.ptr for t in ts] for ts in arrs]
    say("mixed batch: one batched call over 64 frames")
    ctx.batch_device(frames, ptr(dev_in), ptr(out_b))
    ctx.finish()
    say("mixed batch: 64 per-frame calls")
    for k, f in enumerate(frames):
        ctx.frame_device(f, ptr(dev_in)[k], ptr(out_pf)[k])
    ctx.finish()
    got_b = [[t.download() for t in ts] for ts in out_b]
    got_pf = [[t.download() for t in ts] for ts in out_pf]
    ctx.close()
    yield frames, got_b, got_pf


def test_gpu_mixed_batch_of_64_frames_equals_the_per_frame_path(mixed_run):
    frames, got_b, got_pf = mixed_run
    assert len(frames) == 64
    assert sum(f.desc.use_superres for f in frames) > 16
    for k in range(64):
        for p, (b, f) in enumerate(zip(got_b[k], got_pf[k])):
            assert np.array_equal(b, f), f"frame {k} plane {p}"


@pytest.mark.parametrize("cls", sorted(CLASSES), ids=lambda c: f"{c[0]}x{c[1]}-bd{c[2]}" + ("-sr" if c[3] else ""))
def test_gpu_mixed_batch_class_matches_oracle(mixed_run, cls):
    frames, got_b, _ = mixed_run
    k = CLASSES[cls]
    f = frames[k]
    say(f"mixed batch frame {k}: {cls}, coded width {f.desc.coded_width}")
    want = O.frame(f.desc, f.planes)
    for p, (g, w) in enumerate(zip(got_b[k], want)):
        assert g.shape == w.shape and np.array_equal(g, w), f"frame {k} plane {p}"
