"""Launch variants (tuning) on the GPU: every variant gives the oracle's
output bit for bit, on the per-frame and the batched path; the autotuner
returns a valid choice and leaves the outputs exact.
"""
import itertools
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


DEFAULT = {"superres_rows": 8, "cdef_xcd_order": 1, "lr_order": 0, "lr_tile_rows": 64, "cdef_waves": 16, "chain": 0}


@pytest.fixture(scope="module")
def ctx():
    c = av1pf.Context(3840, 2176)
    yield c
    c.close()


CASES = [
    dict(seed=201, coded_w=131, coded_h=97, bd=10),
    dict(seed=202, coded_w=93, coded_h=75, upscaled_w=160),
    dict(seed=203, coded_w=171, coded_h=130, upscaled_w=300, bd=10, unit=32),
    dict(seed=204, coded_w=200, coded_h=130, unit=128, uv_unit=64),
    dict(seed=205, coded_w=72, coded_h=80, ssx=1, ssy=0),
    dict(seed=206, coded_w=77, coded_h=45, mono=True),
]


@pytest.fixture(scope="module")
def batch():
    from vaco_amd import hip
    frames = [F.random_frame(**c) for c in CASES] + [av1pf.generate(workloads.cfg0(), 0)]
    want = [O.frame(f.desc, f.planes) for f in frames]
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    dev_out = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    return frames, want, dev_in, dev_out


def ptrs(arrs):
    return [[t.ptr for t in ts] for ts in arrs]


def check(dev_out, want, label):
    for k, w in enumerate(want):
        for p, plane in enumerate(w):
            got = dev_out[k][p].download()
            assert np.array_equal(got, plane), (label, k, p)


def test_gpu_tuning_defaults_and_validation(ctx):
    say("tuning: defaults and rejection")
    assert ctx.tuning == DEFAULT
    for bad in ({"superres_rows": 5}, {"cdef_xcd_order": 2}, {"lr_order": 3}, {"lr_tile_rows": 48},
                {"cdef_waves": 4}, {"cdef_waves": 12}, {"chain": 2}):
        with pytest.raises(av1pf.Av1Error) as e:
            ctx.tuning = bad
        assert e.value.status == abi.INVALID_ARGUMENT
    assert ctx.tuning["superres_rows"] == 8


@pytest.mark.parametrize("rows,cdef_xcd,lr_order,lr_rows,waves",
                         list(itertools.product((4, 8, 16), (0, 1), (0, 1), (64, 32, 16), (16, 8))))
def test_gpu_every_launch_variant_matches_oracle(ctx, batch, rows, cdef_xcd, lr_order, lr_rows, waves):
    frames, want, dev_in, dev_out = batch
    label = f"rows={rows} cdef_xcd={cdef_xcd} lr_order={lr_order} lr_rows={lr_rows} cdef_waves={waves}"
    say("variant", label)
    # chain off: the per-frame calls below run the separate kernels of this
    # variant (the chained launch has tests of its own)
    ctx.tuning = {"superres_rows": rows, "cdef_xcd_order": cdef_xcd, "lr_order": lr_order, "lr_tile_rows": lr_rows,
                  "cdef_waves": waves, "chain": 0}
    try:
        ctx.batch_device(frames, ptrs(dev_in), ptrs(dev_out))
        ctx.finish()
        check(dev_out, want, "batch " + label)
        for k, f in enumerate(frames[:3]):
            ctx.frame_device(f, ptrs(dev_in)[k], ptrs(dev_out)[k])
        ctx.finish()
        check(dev_out[:3], want[:3], "per-frame " + label)
    finally:
        ctx.tuning = DEFAULT


def test_gpu_autotune_returns_a_valid_choice_and_exact_outputs(ctx, batch):
    frames, want, dev_in, dev_out = batch
    say("autotune")
    t = ctx.autotune(frames, ptrs(dev_in), ptrs(dev_out), reps=2)
    assert t["superres_rows"] in (4, 8, 16) and t["cdef_xcd_order"] in (0, 1) and t["lr_order"] in (0, 1)
    assert t["lr_tile_rows"] in (64, 32, 16) and t["cdef_waves"] in (16, 8) and t["chain"] in (0, 1)
    assert ctx.tuning == t
    check(dev_out, want, "after autotune")
    ctx.batch_device(frames, ptrs(dev_in), ptrs(dev_out))
    ctx.finish()
    check(dev_out, want, f"tuned {t}")
    ctx.tuning = DEFAULT
