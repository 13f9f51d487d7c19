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
    # the batch with the variants the tuner picks for it, the per-frame calls
    # with the defaults
    ctx.tuning = {"superres_rows": 4, "lr_tile_rows": 32, "cdef_waves": 8}
    say("mixed batch: one batched call over 64 frames")
    ctx.batch_device(frames, ptr(dev_in), ptr(out_b))
    ctx.finish()
    ctx.tuning = {"superres_rows": 8, "lr_tile_rows": 64, "cdef_waves": 16}
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


# ---------------------------------------------------------------- serving ---
# The serving workload at its real shapes: 32 streams of fixed 720p / 1080p /
# 2160p resolution and 8 / 10-bit depth, Poisson arrivals at 3,000 frames/s,
# for 0.5 s (1,500 requests), 32 GiB budget, 2 ms window.

STREAM_CFGS = workloads.serve_streams(32)
STREAM_CLASSES = {}
for _s, _c in enumerate(STREAM_CFGS):
    STREAM_CLASSES.setdefault((int(_c.width), int(_c.height), int(_c.bit_depth)), _s)


@pytest.fixture(scope="module")
def served_run():
    from vaco_amd import streams
    say("serving: 32 Poisson streams at 3000 frames/s for 0.5 s")
    run = streams.serve_poisson(0.5, fps=3000.0)
    say(f"serving: {run.stats}")
    yield run


def test_gpu_served_streams_equal_the_per_frame_path(served_run):
    st = served_run.stats
    assert st["frames"] == 1500 and st["spot_checks"] >= 15
    for i, (s, k, planes) in served_run.checked.items():
        for p, (a, b) in enumerate(zip(planes, served_run.reference[(s, k)])):
            assert np.array_equal(a, b), f"request {i} (stream {s}, frame {k}) plane {p}"


@pytest.mark.parametrize("cls", sorted(STREAM_CLASSES), ids=lambda c: f"{c[0]}x{c[1]}-bd{c[2]}")
def test_gpu_served_stream_class_matches_oracle(served_run, cls):
    s = STREAM_CLASSES[cls]
    f = served_run.frames[s][0]
    say(f"serving: stream {s} {cls} frame 0 vs oracle")
    want = O.frame(f.desc, f.planes)
    for p, (g, w) in enumerate(zip(served_run.reference[(s, 0)], want)):
        assert g.shape == w.shape and np.array_equal(g, w), f"stream {s} plane {p}"
