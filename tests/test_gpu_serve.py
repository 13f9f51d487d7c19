"""The serving loop on the GPU: a frame's result is the same whether it ran
alone or batched with other streams' frames, and equals the oracle's;
errors stay with their frame; admission never exceeds the HBM budget.
Bit-exact (tolerance 0): the stage is integer arithmetic.
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


@pytest.fixture(scope="module")
def ctx():
    c = av1pf.Context(3840, 2176)
    yield c
    c.close()


def device_io(frames):
    from vaco_amd import hip
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    dev_out = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    return dev_in, dev_out


def ptrs(arrs):
    return [t.ptr for t in arrs]


def mixed_frames():
    cases = [
        dict(seed=101, coded_w=64, coded_h=64),
        dict(seed=102, coded_w=131, coded_h=97, bd=10),
        dict(seed=103, coded_w=93, coded_h=75, upscaled_w=160),
        dict(seed=104, coded_w=77, coded_h=45, mono=True),
        dict(seed=105, coded_w=200, coded_h=130, unit=128, uv_unit=64),
        dict(seed=106, coded_w=72, coded_h=80, ssx=0, ssy=0),
        dict(seed=107, coded_w=171, coded_h=130, upscaled_w=300, bd=10, unit=32),
        dict(seed=108, coded_w=80, coded_h=72, cdef=False),
    ]
    frames = [F.random_frame(**c) for c in cases]
    frames += [av1pf.generate(workloads.cfg0(), k) for k in range(4)]
    return frames


def test_gpu_served_frames_alone_and_in_a_mix_are_identical(ctx):
    frames = mixed_frames()
    n = len(frames)
    dev_in, out_alone = device_io(frames)
    _, out_mix = device_io(frames)
    say(f"serve: {n} frames alone")
    with av1pf.Server(ctx, window_us=2000, max_batch=64) as srv:
        for k, f in enumerate(frames):
            srv.submit(f, ptrs(dev_in[k]), ptrs(out_alone[k]), tag=k)
            srv.drain()
            (c,) = srv.poll()
            assert c[0] == k and c[1] == abi.OK and c[2] == 1, c
    say(f"serve: {n} frames in a mix (reverse order, one burst)")
    with av1pf.Server(ctx, window_us=2000, max_batch=64) as srv:
        for k in reversed(range(n)):
            srv.submit(frames[k], ptrs(dev_in[k]), ptrs(out_mix[k]), tag=1000 + k)
        srv.drain()
        done = srv.poll()
    assert sorted(c[0] for c in done) == [1000 + k for k in range(n)]
    assert all(c[1] == abi.OK for c in done), done
    assert max(c[2] for c in done) > 1, "a burst should be batched"
    for c in done:
        assert c[3] <= c[4] <= c[5], c
    for k, f in enumerate(frames):
        alone = [t.download() for t in out_alone[k]]
        mix = [t.download() for t in out_mix[k]]
        want = O.frame(f.desc, f.planes)
        for p in range(len(want)):
            assert np.array_equal(alone[p], want[p]), (k, p, "alone vs oracle")
            assert np.array_equal(mix[p], alone[p]), (k, p, "mix vs alone")


def test_gpu_frames_admitted_while_a_batch_runs_overlap_it_and_stay_exact(ctx):
    """Continuous batching with work overlapping on the device: four 4K
    frames then nine small frames submitted in one burst, no batching
    window, at most 4 frames per batch and two batches in flight. Some batch
    is launched while another is still on the GPU, and on the device clock
    the two run at the same time (on one shared stream the later batch could
    only start after the earlier one ended). Every frame's result equals its
    result alone; the small frames equal the oracle."""
    order = [av1pf.generate(workloads.cfg2(), k) for k in (0, 1, 2, 3)]
    smalls = [F.random_frame(200 + k, 64, 64) for k in range(9)]
    order += smalls
    dev_in, out_alone = device_io(order)
    _, out_mix = device_io(order)
    say("serve: per-frame results alone")
    for k, f in enumerate(order):
        ctx.frame_device(f, ptrs(dev_in[k]), ptrs(out_alone[k]))
    ctx.finish()
    say("serve: four 4K frames then nine small ones in one burst, window 0, 2 batches in flight")
    with av1pf.Server(ctx, window_us=0, max_batch=4, max_inflight=2) as srv:
        calls = [srv.submit_call(f, ptrs(dev_in[k]), ptrs(out_mix[k]), tag=k) for k, f in enumerate(order)]
        for c in calls:
            c()
        srv.drain()
        done = srv.poll()
    assert sorted(c[0] for c in done) == list(range(len(order)))
    assert all(c[1] == abi.OK for c in done), done
    # per batch: (host launch, host seen done, device start, device end)
    batches = sorted({(c[4], c[5], c[6], c[7]) for c in done})
    assert len(batches) >= 2 and all(0 < b[2] < b[3] for b in batches), batches
    assert any(l2 > l1 and l2 < d1 and s2 < e1 and s1 < e2
               for (l1, d1, s1, e1) in batches for (l2, d2, s2, e2) in batches), batches
    for k, f in enumerate(order):
        mix = [t.download() for t in out_mix[k]]
        alone = [t.download() for t in out_alone[k]]
        for p in range(len(alone)):
            assert np.array_equal(mix[p], alone[p]), (k, p, "mix vs alone")
    for f in smalls:
        idx = order.index(f)
        for p, w in enumerate(O.frame(f.desc, f.planes)):
            assert np.array_equal(out_mix[idx][p].download(), w), (idx, p, "vs oracle")


def test_gpu_served_errors_stay_with_their_frame(ctx):
    frames = [F.random_frame(60 + k, 64, 64, skip_p=0.0, idx_neg_p=0.0) for k in range(5)]
    frames[1].planes[0][9, 9] = 999          # flagged by the kernels, settled by the host replay
    frames[3].desc.cdef.damping = 13         # caught before launch
    want = []
    for f in frames:
        try:
            want.append(("ok", O.frame(f.desc, f.planes)))
        except O.OracleError as e:
            want.append(("err", e.status))
    assert want[1][0] == "err" and want[3][0] == "err"
    dev_in, dev_out = device_io(frames)
    say("serve: errors in a mix")
    with av1pf.Server(ctx, window_us=5000, max_batch=8) as srv:
        for k, f in enumerate(frames):
            srv.submit(f, ptrs(dev_in[k]), ptrs(dev_out[k]), tag=k)
        srv.drain()
        got = {c[0]: c[1] for c in srv.poll()}
    for k, w in enumerate(want):
        if w[0] == "err":
            assert got[k] == w[1], (k, got[k], w)
        else:
            assert got[k] == abi.OK, (k, got[k])
            for p, plane in enumerate(w[1]):
                assert np.array_equal(dev_out[k][p].download(), plane), (k, p)


def test_gpu_server_admission_respects_the_hbm_budget(ctx):
    frames = [av1pf.generate(workloads.cfg0(), k) for k in range(6)]
    per = av1pf.frame_bytes(frames[0])
    assert all(av1pf.frame_bytes(f) == per for f in frames)
    dev_in, dev_out = device_io(frames)
    say(f"serve: budget of two frames ({2 * per} bytes)")
    with av1pf.Server(ctx, hbm_budget_bytes=2 * per + per // 2, window_us=500, max_batch=64,
                      max_inflight=2) as srv:
        big = av1pf.generate(workloads.cfg1(), 0)
        with pytest.raises(av1pf.Av1Error) as e:
            srv.submit_call(big, [0, 0, 0], [0, 0, 0], tag=99)()
        assert e.value.status == abi.LIMIT_EXCEEDED
        for k, f in enumerate(frames):
            srv.submit(f, ptrs(dev_in[k]), ptrs(dev_out[k]), tag=k)
        srv.drain()
        done = srv.poll()
    assert len(done) == len(frames) and all(c[1] == abi.OK for c in done)
    # at most two frames on the device at once: a launch holds <= 2 frames,
    # and two launches overlap only when each holds one
    assert max(c[2] for c in done) <= 2, done
    for k, f in enumerate(frames):
        want = O.frame(f.desc, f.planes)
        for p, plane in enumerate(want):
            assert np.array_equal(dev_out[k][p].download(), plane), (k, p)
