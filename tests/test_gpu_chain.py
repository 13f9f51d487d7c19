"""The chained launch (tuning.chain): one frame's CDEF and loop restoration
in one launch, each restoration tile waiting for the superblocks whose CDEF
output it reads. Bit-exact against the oracle under every order and tile
height, in every chroma layout and bit depth; the chained kernel is the one
that ran; many chained launches in a row (the completion flags carry each
launch's epoch and are never cleared); the give-up path of a tile's bounded
wait reports an error instead of waiting on.
"""
import itertools
import sys

import numpy as np
import pytest

import frames as F
from oracle import oracle as O
from vaco_amd import abi, av1pf, hip, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# Frames with CDEF and restoration and no superres: the chained path.
CASES = [
    dict(seed=301, coded_w=131, coded_h=97),
    dict(seed=302, coded_w=200, coded_h=130, bd=10, unit=128, uv_unit=64),
    dict(seed=303, coded_w=150, coded_h=90, bd=12, ssx=0, ssy=0, unit=32),
    dict(seed=304, coded_w=136, coded_h=72, ssx=1, ssy=0, bd=10),
    dict(seed=305, coded_w=72, coded_h=136, ssx=0, ssy=1),
    dict(seed=306, coded_w=77, coded_h=45, mono=True),
    dict(seed=307, coded_w=260, coded_h=200, unit=256, uv_unit=128, lr_types=(1, 0, 1)),
    dict(seed=308, coded_w=8, coded_h=300, unit=64),
    dict(seed=309, coded_w=333, coded_h=8, bd=10, unit=32),
    dict(seed=310, coded_w=190, coded_h=190, idx_neg_p=0.6, skip_p=0.5, unit_kinds=(2,)),
    dict(seed=311, coded_w=128, coded_h=128, zero_strengths=True, unit_kinds=(1,)),  # CDEF off: not chained
    dict(seed=312, coded_w=129, coded_h=65, lr_types=(0, 1, 1), unit_kinds=(0,)),
    dict(seed=313, coded_w=171, coded_h=130, upscaled_w=300, bd=10),  # superres: not chained
]


# Zero strengths everywhere turn CDEF off (decode.rs:741-748), so that frame
# runs restoration alone; a superres frame keeps its launches apart.
CHAINED = [not c.get("zero_strengths") and not c.get("upscaled_w") for c in CASES] + [True]


@pytest.fixture(scope="module")
def ctx():
    c = av1pf.Context(1920, 1088)
    yield c
    c.close()


@pytest.fixture(scope="module")
def frames():
    fs = [F.random_frame(**c) for c in CASES] + [av1pf.generate(workloads.cfg0(), 1)]
    want = [O.frame(f.desc, f.planes) for f in fs]
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in fs]
    dev_out = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in fs]
    return fs, want, dev_in, dev_out


def ptrs(arrs):
    return [t.ptr for t in arrs]


def assert_equal(dev_out, want, label):
    for p, w in enumerate(want):
        got = dev_out[p].download()
        if not np.array_equal(got, w):
            ys, xs = np.nonzero(got != w)
            raise AssertionError(f"{label} plane {p}: {len(ys)} samples differ, first at y={ys[0]} x={xs[0]}")


def run_frames(ctx, frames, label, expect_chain=True):
    fs, want, dev_in, dev_out = frames
    for k, f in enumerate(fs):
        dev_out[k][0].zero()
        ctx.profile(True)
        ctx.profile_read()
        ctx.frame_device(f, ptrs(dev_in[k]), ptrs(dev_out[k]))
        ctx.finish()
        prof = ctx.profile_read()
        ctx.profile(False)
        if expect_chain and CHAINED[k]:
            assert prof["chain"][1] == 1 and prof["cdef"][1] == 0 and prof["restoration"][1] == 0, (label, k, prof)
        else:
            assert prof["chain"][1] == 0, (label, k, prof)
        assert_equal(dev_out[k], want[k], f"{label} frame {k} ({CASES[k] if k < len(CASES) else 'cfg0'})")


@pytest.mark.parametrize("cdef_xcd,lr_order,lr_rows", list(itertools.product((0, 1), (0, 1), (64, 32, 16))))
def test_gpu_chain_matches_oracle_under_every_variant(ctx, frames, cdef_xcd, lr_order, lr_rows):
    label = f"chain cdef_xcd={cdef_xcd} lr_order={lr_order} lr_rows={lr_rows}"
    say(label)
    ctx.tuning = {"chain": 1, "cdef_xcd_order": cdef_xcd, "lr_order": lr_order, "lr_tile_rows": lr_rows}
    try:
        run_frames(ctx, frames, label)
    finally:
        ctx.tuning = {"chain": 0, "cdef_xcd_order": 1, "lr_order": 0, "lr_tile_rows": 64}


def test_gpu_chain_off_runs_separate_launches(ctx, frames):
    say("chain off")
    ctx.tuning = {"chain": 0}
    run_frames(ctx, frames, "chain off", expect_chain=False)


def test_gpu_chain_many_launches_in_a_row(ctx, frames):
    """Launches back to back without waiting in between: each launch's flags
    carry its own epoch, so a tile never takes the previous frame's flags
    for its own superblocks."""
    fs, want, dev_in, dev_out = frames
    say("chain: 120 launches back to back")
    ctx.tuning = {"chain": 1}
    order = [k % len(fs) for k in range(120)]
    for k in order:
        ctx.frame_device(fs[k], ptrs(dev_in[k]), ptrs(dev_out[k]))
    ctx.finish()
    for k in range(len(fs)):
        assert_equal(dev_out[k], want[k], f"back to back frame {k}")


def test_gpu_chain_1080p_workload(ctx):
    """The 1080p 8-bit workload (128x128 units) through the chained launch."""
    for i in range(2):
        f = av1pf.generate(workloads.cfg1(), i)
        say(f"chain: 1080p workload frame {i}")
        ctx.tuning = {"chain": 1}
        want = O.frame(f.desc, f.planes)
        dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
        dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
        ctx.profile(True)
        ctx.profile_read()
        ctx.frame_device(f, ptrs(dev_in), ptrs(dev_out))
        ctx.finish()
        prof = ctx.profile_read()
        ctx.profile(False)
        assert prof["chain"][1] == 1, prof
        assert_equal(dev_out, want, f"1080p frame {i}")


def test_gpu_chain_gives_up_and_reports_instead_of_waiting(ctx):
    """With the tiles' poll limit at 0, a tile that finds a superblock
    unfinished gives up at once: the call fails with DEVICE_ERROR and the
    give-up message (nothing waits), and the context then runs the same frame
    exactly with the default limit."""
    f = av1pf.generate(workloads.cfg1(), 0)
    want = O.frame(f.desc, f.planes)
    dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
    dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
    say("chain: poll limit 0")
    ctx.tuning = {"chain": 1}
    ctx.set_chain_spin_limit(0)
    try:
        gave_up = False
        for _ in range(3):
            ctx.frame_device(f, ptrs(dev_in), ptrs(dev_out))
            try:
                ctx.finish()
            except av1pf.Av1Error as e:
                assert e.status == abi.DEVICE_ERROR
                assert "gave up waiting" in str(e)
                gave_up = True
        assert gave_up
    finally:
        ctx.set_chain_spin_limit(1 << 22)
    say("chain: default poll limit again")
    ctx.frame_device(f, ptrs(dev_in), ptrs(dev_out))
    ctx.finish()
    assert_equal(dev_out, want, "after the give-up test")
